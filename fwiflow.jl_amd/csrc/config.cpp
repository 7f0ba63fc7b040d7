#include "config.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>

#include "json.hpp"

namespace fwiflow {

namespace {
constexpr double kPI = 3.141592653589793238462643383279502884197169;

Status need_int(const json::Value& o, const char* key, int& out, const std::string& file) {
  const json::Value* v = o.get(key);
  if (!v) return Status::Err(FWIFLOW_ERR_JSON, file + ": missing key \"" + key + "\"");
  if (v->kind != json::Value::Number || !v->is_int)
    return Status::Err(FWIFLOW_ERR_JSON, file + ": key \"" + key + "\" is not an integer");
  out = (int)v->num;
  return Status::Ok();
}

Status need_num(const json::Value& o, const char* key, double& out, const std::string& file) {
  const json::Value* v = o.get(key);
  if (!v) return Status::Err(FWIFLOW_ERR_JSON, file + ": missing key \"" + key + "\"");
  if (v->kind != json::Value::Number)
    return Status::Err(FWIFLOW_ERR_JSON, file + ": key \"" + key + "\" is not a number");
  out = v->num;
  return Status::Ok();
}

Status need_str(const json::Value& o, const char* key, std::string& out, const std::string& file) {
  const json::Value* v = o.get(key);
  if (!v) return Status::Err(FWIFLOW_ERR_JSON, file + ": missing key \"" + key + "\"");
  if (v->kind != json::Value::String)
    return Status::Err(FWIFLOW_ERR_JSON, file + ": key \"" + key + "\" is not a string");
  out = v->str;
  return Status::Ok();
}

Status int_array(const json::Value& o, const char* key, std::vector<int>& out,
                 const std::string& where) {
  const json::Value* v = o.get(key);
  if (!v) return Status::Err(FWIFLOW_ERR_JSON, where + ": missing key \"" + key + "\"");
  if (v->kind != json::Value::Array)
    return Status::Err(FWIFLOW_ERR_JSON, where + ": key \"" + key + "\" is not an array");
  out.clear();
  for (const auto& e : v->arr) {
    if (e.kind != json::Value::Number || !e.is_int)
      return Status::Err(FWIFLOW_ERR_JSON, where + ": \"" + key + "\" holds a non-integer");
    out.push_back((int)e.num);
  }
  return Status::Ok();
}

Status num_array(const json::Value& o, const char* key, std::vector<float>& out,
                 const std::string& where) {
  const json::Value* v = o.get(key);
  if (!v) return Status::Err(FWIFLOW_ERR_JSON, where + ": missing key \"" + key + "\"");
  if (v->kind != json::Value::Array)
    return Status::Err(FWIFLOW_ERR_JSON, where + ": key \"" + key + "\" is not an array");
  out.clear();
  for (const auto& e : v->arr) {
    if (e.kind != json::Value::Number)
      return Status::Err(FWIFLOW_ERR_JSON, where + ": \"" + key + "\" holds a non-number");
    out.push_back((float)e.num);  // Src_Rec.cu:164,172,190: GetDouble() into float
  }
  return Status::Ok();
}

Status need_bool(const json::Value& o, const char* key, bool& out, const std::string& file) {
  const json::Value* v = o.get(key);
  if (!v) return Status::Ok();
  if (v->kind != json::Value::Bool)
    return Status::Err(FWIFLOW_ERR_JSON, file + ": \"" + key + "\" is not a bool");
  out = v->b;
  return Status::Ok();
}

int largest_prime_factor(int n) {
  int best = 1;
  for (int p = 2; (long long)p * p <= n; p++)
    while (n % p == 0) {
      best = p;
      n /= p;
    }
  return n > 1 ? std::max(best, n) : best;
}

#define TRY(expr)                  \
  do {                             \
    Status _s = (expr);            \
    if (!_s.ok()) return _s;       \
  } while (0)
}  // namespace

Status read_first_line(const std::string& fname, std::string& line) {
  std::ifstream f(fname);
  if (!f.is_open()) return Status::Err(FWIFLOW_ERR_IO, "cannot open \"" + fname + "\"");
  // the reference reads one line with getline (Parameter.cpp:28, Src_Rec.cu:31)
  std::getline(f, line);
  return Status::Ok();
}

Status check_calc_id(int calc_id) {
  if (calc_id < 0 || calc_id > 2)
    return Status::Err(FWIFLOW_ERR_CALC_ID, "invalid calc_id " + std::to_string(calc_id));
  return Status::Ok();
}

Status parse_para(const std::string& fname, Para& p) {
  std::string line, err;
  TRY(read_first_line(fname, line));
  json::Value j;
  if (!json::parse(line, j, err)) return Status::Err(FWIFLOW_ERR_JSON, fname + ": " + err);
  if (j.kind != json::Value::Object)
    return Status::Err(FWIFLOW_ERR_JSON, fname + ": not a JSON object");
  double dz, dx, dt, f0;
  TRY(need_int(j, "nz", p.nz, fname));
  TRY(need_int(j, "nx", p.nx, fname));
  TRY(need_num(j, "dz", dz, fname));
  TRY(need_num(j, "dx", dx, fname));
  TRY(need_int(j, "nSteps", p.nSteps, fname));
  TRY(need_int(j, "nPoints_pml", p.nPml, fname));
  TRY(need_int(j, "nPad", p.nPad, fname));
  TRY(need_num(j, "dt", dt, fname));
  TRY(need_num(j, "f0", f0, fname));
  TRY(need_str(j, "survey_fname", p.survey_fname, fname));
  TRY(need_str(j, "data_dir_name", p.data_dir, fname));
  p.dz = (float)dz;
  p.dx = (float)dx;
  p.dt = (float)dt;
  p.f0 = (float)f0;
  if (const json::Value* s = j.get("scratch_dir_name")) {
    if (s->kind != json::Value::String)
      return Status::Err(FWIFLOW_ERR_JSON, fname + ": \"scratch_dir_name\" is not a string");
    p.scratch_dir = s->str;
    p.save_scratch = true;
  }
  TRY(need_bool(j, "if_win", p.if_win, fname));
  TRY(need_bool(j, "if_src_update", p.if_src_update, fname));
  if (const json::Value* f = j.get("filter")) {
    // Parameter.cpp:155-167: any "filter" key switches the band-pass on
    std::vector<float> v;
    TRY(num_array(j, "filter", v, fname));
    if (v.size() != 4)
      return Status::Err(FWIFLOW_ERR_JSON, fname + ": \"filter\" needs 4 frequencies");
    for (int q = 0; q < 4; q++) p.filter[q] = v[q];
    p.if_filter = true;
    (void)f;
  }
  TRY(check_geometry(p));
  if (p.if_filter || p.if_src_update) {
    const int L = 2 * p.nSteps;
    if (L > kMaxTransform || largest_prime_factor(L) > kMaxRadix)
      return Status::Err(FWIFLOW_ERR_UNSUPPORTED,
                         "band-pass filter / source update: 2*nSteps = " + std::to_string(L) +
                             " must be <= " + std::to_string(kMaxTransform) +
                             " with prime factors <= " + std::to_string(kMaxRadix));
  }
  return Status::Ok();
}

Status check_geometry(const Para& p) {
  if (p.nz <= 0 || p.nx <= 0 || p.nPad < 0 || p.nPad >= p.nz)
    return Status::Err(FWIFLOW_ERR_SHAPE, "bad grid size nz/nx/nPad");
  if (p.nSteps < 2) return Status::Err(FWIFLOW_ERR_SHAPE, "nSteps must be >= 2");
  if (!(p.dz > 0) || !(p.dx > 0) || !(p.dt > 0))
    return Status::Err(FWIFLOW_ERR_SHAPE, "dz, dx and dt must be positive");
  // the boundary ring spans rows/columns nPml-2 .. nPml+2 on each side and
  // must sit inside the update region [2, n-3]
  if (p.nPml < 4) return Status::Err(FWIFLOW_ERR_SHAPE, "nPoints_pml must be >= 4");
  if (p.nzA() - 2 * p.nPml < 6 || p.nx - 2 * p.nPml < 6)
    return Status::Err(FWIFLOW_ERR_SHAPE, "interior (grid minus C-PML) must be >= 6 cells");
  return Status::Ok();
}

Status parse_survey(const Para& p, int group_size, const int* shot_ids, Survey& sv) {
  if (group_size <= 0 || !shot_ids)
    return Status::Err(FWIFLOW_ERR_ARG, "group_size must be > 0 with a shot_ids array");
  const std::string& fname = p.survey_fname;
  std::string line, err;
  TRY(read_first_line(fname, line));
  json::Value j;
  if (!json::parse(line, j, err)) return Status::Err(FWIFLOW_ERR_JSON, fname + ": " + err);
  if (j.kind != json::Value::Object)
    return Status::Err(FWIFLOW_ERR_JSON, fname + ": not a JSON object");
  TRY(need_int(j, "nShots", sv.nShots, fname));
  sv.shots.clear();
  const int zlo = 2, zhi = p.nzA() - 3, xlo = 2, xhi = p.nx - 3;
  for (int i = 0; i < group_size; i++) {
    const int id = shot_ids[i];
    if (id < 0 || id >= sv.nShots)
      return Status::Err(FWIFLOW_ERR_SHAPE, "shot id " + std::to_string(id) +
                                                " outside [0, nShots=" +
                                                std::to_string(sv.nShots) + ")");
    const std::string key = "shot" + std::to_string(id);
    const json::Value* o = j.get(key);
    if (!o || o->kind != json::Value::Object)
      return Status::Err(FWIFLOW_ERR_JSON, fname + ": missing object \"" + key + "\"");
    Shot s;
    s.id = id;
    int nrec = 0;
    TRY(need_int(*o, "z_src", s.z_src, key));
    TRY(need_int(*o, "x_src", s.x_src, key));
    TRY(need_int(*o, "nrec", nrec, key));
    TRY(int_array(*o, "z_rec", s.z_rec, key));
    TRY(int_array(*o, "x_rec", s.x_rec, key));
    if (nrec <= 0 || (int)s.z_rec.size() != nrec || (int)s.x_rec.size() != nrec)
      return Status::Err(FWIFLOW_ERR_SHAPE, key + ": nrec does not match z_rec/x_rec sizes");
    if (p.if_win) {
      TRY(num_array(*o, "win_start", s.win_start, key));
      TRY(num_array(*o, "win_end", s.win_end, key));
      TRY(num_array(*o, "weights", s.weights, key));
      if ((int)s.win_start.size() != nrec || (int)s.win_end.size() != nrec ||
          (int)s.weights.size() != nrec)
        return Status::Err(FWIFLOW_ERR_SHAPE,
                           key + ": win_start/win_end/weights need nrec values each");
    }
    s.z_src += p.nPml;
    s.x_src += p.nPml;
    if (s.z_src < zlo || s.z_src > zhi || s.x_src < xlo || s.x_src > xhi)
      return Status::Err(FWIFLOW_ERR_SHAPE, key + ": source outside the grid");
    for (int r = 0; r < nrec; r++) {
      s.z_rec[r] += p.nPml;
      s.x_rec[r] += p.nPml;
      if (s.z_rec[r] < zlo || s.z_rec[r] > zhi || s.x_rec[r] < xlo || s.x_rec[r] > xhi)
        return Status::Err(FWIFLOW_ERR_SHAPE, key + ": receiver " + std::to_string(r) +
                                                  " outside the grid");
    }
    sv.shots.push_back(std::move(s));
  }
  return Status::Ok();
}

void model_to_zfast(const Para& p, const double* Lambda, const double* Mu, const double* Den,
                    bool input_zfast, std::vector<float>& lam, std::vector<float>& mu,
                    std::vector<float>& den, float& cp_max) {
  const int nz = p.nz, nx = p.nx;
  const size_t n = (size_t)nz * nx;
  lam.resize(n);
  mu.resize(n);
  den.resize(n);
  for (int j = 0; j < nx; j++)
    for (int i = 0; i < nz; i++) {
      const size_t src = input_zfast ? (size_t)j * nz + i : (size_t)i * nx + j;
      const size_t dst = (size_t)j * nz + i;
      lam[dst] = (float)(Lambda[src] * 1e6);
      mu[dst] = (float)(Mu[src] * 1e6);
      den[dst] = (float)Den[src];
    }
  // Vp as velInit computes it (utilities.cu:119-121), max for the Courant number
  float m = 0.f;
  bool first = true;
  for (size_t k = 0; k < n; k++) {
    const float cp = (float)std::sqrt((lam[k] + 2.0 * mu[k]) / den[k]);
    if (first || cp > m) m = cp;
    first = false;
  }
  cp_max = m;
}

Status check_courant(const Para& p, float cp_max) {
  const float dh_min = p.dz < p.dx ? p.dz : p.dx;
  const float courant =
      (float)(cp_max * p.dt * std::sqrt(2.0f) * (1.0 / 24.0 + 9.0 / 8.0) / dh_min);
  if (!(courant <= 1.0f))
    return Status::Err(FWIFLOW_ERR_COURANT,
                       "Courant number " + std::to_string(courant) + " > 1 (unstable)");
  return Status::Ok();
}

void cpml_profiles(int N, int nPml, float dh, float f0, float dt, float* out6) {
  float* K = out6;
  float* a = out6 + N;
  float* b = out6 + 2 * N;
  float* Kh = out6 + 3 * N;
  float* ah = out6 + 4 * N;
  float* bh = out6 + 5 * N;
  // Komatitsch & Martin (2007) C-PML with the reference's constants:
  // N = 8, R = 8e-4, K_max = 2, alpha_max = pi f0, damping profile
  // d0 (0.25 x + 0.75 x^8) with d0 = -(N+1) c ln(R) / (2 L), c = 3000 m/s.
  const float npow = 8.f, kmax = 2.f;
  const float alpha_max = (float)(2.0 * kPI * (f0 / 2.0));
  const float L = nPml * dh;
  const float d0 = (float)(-(npow + 1.f) * 3000.f * std::log(0.0008f) / (2.0 * L));
  auto profile = [&](float depth, float& d, float& k, float& al) {
    const float xn = depth / L;
    const float xp = std::pow(xn, npow);
    d = d0 * (0.25f * xn + 0.75f * xp + 0.f * std::pow(xn, 2.f * npow));
    k = (float)(1.0 + (kmax - 1.0) * xp);
    al = (float)(alpha_max * (1.0 - xn));
  };
  for (int i = 0; i < N; i++) {
    float d = 0.f, k = 1.f, al = 0.f, dh_ = 0.f, kh = 1.f, alh = 0.f;
    float depth = (nPml - i) * dh;
    if (depth >= 0.f) profile(depth, d, k, al);
    depth = (float)((nPml - i - 0.5) * dh);
    if (depth >= 0.f) profile(depth, dh_, kh, alh);
    depth = (nPml - N + i) * dh;
    if (depth >= 0.f) profile(depth, d, k, al);
    depth = (float)((nPml - N + i + 0.5) * dh);
    if (depth >= 0.f) profile(depth, dh_, kh, alh);
    if (al < 0.f) al = 0.f;
    if (alh < 0.f) alh = 0.f;
    K[i] = k;
    Kh[i] = kh;
    b[i] = std::exp((float)(-(d / k + al) * dt));
    bh[i] = std::exp((float)(-(dh_ / kh + alh) * dt));
    a[i] = std::fabs(d) > 1.0e-6 ? (float)(d * (b[i] - 1.0) / (k * (d + k * al))) : 0.f;
    ah[i] = std::fabs(dh_) > 1.0e-6 ? (float)(dh_ * (bh[i] - 1.0) / (kh * (dh_ + kh * alh)))
                                    : 0.f;
  }
}

void window_squared(int nt, float dt, float ratio, float* out) {
  const float t0 = 0.f;
  const float t3 = nt * dt;
  const float offset = nt * dt * ratio;
  for (int it = 0; it < nt; it++) out[it] = 1.f;
  if (2.0 * offset >= t3 - t0) return;
  const float t1 = t0 + offset, t2 = t3 - offset;
  for (int it = 0; it < nt; it++) {
    const float t = it * dt;
    float w;
    if (t >= t0 && t < t1)
      w = (float)std::sin(kPI / 2.0 * (t - t0) / (t1 - t0));
    else if (t >= t1 && t < t2)
      w = 1.f;
    else if (t >= t2 && t < t3)
      w = (float)std::cos(kPI / 2.0 * (t - t2) / (t3 - t2));
    else
      w = 0.f;
    out[it] = w * w;
  }
}

}  // namespace fwiflow
