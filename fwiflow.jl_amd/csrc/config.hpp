// Host-side configuration of one operator call: the parameter file
// (deps/CustomOps/FWI/Src/Parameter.cpp:16-178), the survey file
// (Src_Rec.cu:19-245), input validation that the reference does with exit(1),
// and the small host tables the kernels read (C-PML profiles, time windows).
#pragma once
#include <string>
#include <vector>

#include "../../include/fwiflow.h"

namespace fwiflow {

struct Status {
  int code = FWIFLOW_OK;
  std::string msg;
  bool ok() const { return code == FWIFLOW_OK; }
  static Status Ok() { return Status(); }
  static Status Err(int c, std::string m) {
    Status s;
    s.code = c;
    s.msg = std::move(m);
    return s;
  }
};

// Parameter file.  Field names as in the JSON (Utils.jl:21-63).
struct Para {
  int nz = 0, nx = 0, nSteps = 0, nPml = 0, nPad = 0;
  // kept as float, as Parameter.h:191-195 does
  float dz = 0, dx = 0, dt = 0, f0 = 0;
  std::string survey_fname, data_dir, scratch_dir;
  bool save_scratch = false;
  // optional residual stages (Parameter.cpp:146-174), off by default
  bool if_win = false, if_filter = false, if_src_update = false;
  float filter[4] = {0, 0, 0, 0};  // Hz: taper-in start/end, taper-out start/end
  int nzA() const { return nz - nPad; }  // rows the stencils touch
  bool conditioned() const { return if_win || if_filter || if_src_update; }
};

// Longest trace transform of the band-pass filter and the source update:
// 2*nSteps samples, held in the 160 KB of LDS of one workgroup.
constexpr int kMaxTransform = 20480;
// Largest prime factor of 2*nSteps those transforms handle.
constexpr int kMaxRadix = 61;

// One shot of the group, indices already offset by nPml (Src_Rec.cu:86,90,105,113).
struct Shot {
  int id = 0;
  int z_src = 0, x_src = 0;
  std::vector<int> z_rec, x_rec;
  // per-trace windows (s) and weights, "if_win" only (Src_Rec.cu:156-199)
  std::vector<float> win_start, win_end, weights;
};

struct Survey {
  int nShots = 0;
  std::vector<Shot> shots;  // group order
  int total_rec() const {
    int n = 0;
    for (const auto& s : shots) n += (int)s.z_rec.size();
    return n;
  }
};

Status read_first_line(const std::string& fname, std::string& line);
Status parse_para(const std::string& fname, Para& p);
Status parse_survey(const Para& p, int group_size, const int* shot_ids, Survey& sv);
Status check_calc_id(int calc_id);
// Geometry bounds that the kernels rely on (ring rows inside the update region etc.).
Status check_geometry(const Para& p);

// Host model preparation: transpose TF row-major [NZ][NX] to z-fastest and
// MPa -> Pa (libCUFD.cu:72-78); also returns max Vp for the Courant check.
void model_to_zfast(const Para& p, const double* Lambda, const double* Mu, const double* Den,
                    bool input_zfast, std::vector<float>& lam, std::vector<float>& mu,
                    std::vector<float>& den, float& cp_max);
// utilities.cu:225-240
Status check_courant(const Para& p, float cp_max);

// C-PML profiles: 6 arrays K, a, b, K_half, a_half, b_half of length N
// (utilities.cu:242-358, CpAve forced to 3000).
void cpml_profiles(int N, int nPml, float dh, float f0, float dt, float* out6);
// w(t)^2 of the default sin^2/cos^2 window (utilities.cu:707-747); returns
// all-ones when the window would be rejected ("Window error 2").
void window_squared(int nt, float dt, float ratio, float* out);

}  // namespace fwiflow
