// Data conditioning of the residual: per-trace windows and weights, band-pass
// filter, source-signature update and the adjoint applications that follow
// the misfit (libCUFD.cu:254-309; utilities.cu:596-1112).
//
// Traces stay in the engine's time-major seismogram layout: sample t of
// trace r at base[t * stride + r].  Every transform is one 2*nSteps-point
// complex FFT per PAIR of real traces (x_r + i x_{r+1}), held whole in the
// LDS of one workgroup (mixed radix 4/2/3/5/7/..., decimation in time,
// in place after a digit-reversed load).  The filters that follow are real
// (band-pass) or Hermitian (source update), so the packed spectrum is
// filtered as it stands and the two traces come back as the real and
// imaginary parts.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "config.hpp"

namespace fwiflow {

struct FftDesc {
  int L;       // transform length
  int m;       // number of radix passes
  int p[32];   // radix of each level, level 0 outermost
};

// Radix plan, digit-reversal load permutation and twiddles exp(-2 pi i j / L)
// (computed in double, stored as float) for length L.
Status fft_plan(int L, FftDesc& f, std::vector<int>& perm, std::vector<float2>& tw);

// Bins 0..kcov-1 of the half spectrum that the reference's frequency-domain
// kernels touch (grid sized for nt columns of 32 threads: the Nyquist bin nt
// is skipped when nt is a multiple of 32).
int touched_bins(int nt);

// amp(f)^2 of the band-pass per bin 0..nt of the 2*nt-point grid, in the
// reference's float arithmetic (cuda_bp_filter1d, utilities.cu:596-626).
void bandpass_table(int nt, float dt, const float* f, std::vector<float2>& H);

// ntr traces read from in[t * in_stride + r] and written to
// out[t * out_stride + r] (out may equal in, with the same stride).
struct TraceSet {
  const float* in;
  int in_stride;
  float* out;
  int out_stride;
  int ntr;
};

// out = crop(IFFT(H * FFT(pad(in * padwin)))) * scale; bins kk >= kcov of the
// half spectrum are left as they are; H is a half spectrum (nt+1 bins),
// applied with its Hermitian extension.  padwin (2*nt values) may be null.
hipError_t trace_filter(hipStream_t st, const FftDesc& f, const int* perm, const float2* tw,
                        int nt, TraceSet ts, const float* padwin, const float2* H, int kcov,
                        float scale);

// Half spectra (bins 0..nt) of pad(in * padwin), [trace][bin]; ts.out unused.
hipError_t trace_spectra(hipStream_t st, const FftDesc& f, const int* perm, const float2* tw,
                         int nt, TraceSet ts, const float* padwin, float2* spec);

// coef[k] = sum_r conj(C_r[k]) O_r[k] / (sum_r |C_r[k]|^2 + 1e-6), sums in
// double (cuda_spectrum_update, utilities.cu:768-841); DC and Nyquist real.
hipError_t spectrum_ratio(hipStream_t st, const float2* spec_obs, const float2* spec_cal,
                          int ntr, int nf, float2* coef);

// Time window: default table w2[t] when wstart is null, else per-trace
// windows (ratio, start/end seconds) and weights (utilities.cu:653-704).
hipError_t window_traces(hipStream_t st, int nt, TraceSet ts, const float* w2,
                         const float* wstart, const float* wend, const float* weight, float ratio,
                         float dt);

// res = obs - syn (sample 0 set to 0, gpuMinus utilities.cu:154-167) into
// res[t * res_stride + r]; per-block sums of res^2 in double into partial[0..blocks).
hipError_t shot_residual(hipStream_t st, int nt, int ntr, const float* obs, const float* syn,
                         int in_stride, float* res, int res_stride, double* partial, int blocks);

}  // namespace fwiflow
