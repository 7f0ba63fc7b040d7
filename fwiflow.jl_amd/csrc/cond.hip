// Data conditioning kernels (declarations and layout notes in cond.hpp).
#include "cond.hpp"

#include <cmath>

namespace fwiflow {

namespace {

constexpr double kPI = 3.141592653589793238462643383279502884197169;
constexpr int kThreads = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Radix-P butterfly on a[base + q*M], q < P.  Decimation in time (DIF =
// false): the inputs Y_q[k] are twiddled by W_{PM}^{qk} = tw[q*k*tws], then
// combined by the P-point DFT (W_P^x = tw[x * L/P]) into X[k + j*M] at
// a[base + j*M].  Decimation in frequency (DIF = true) is its transpose: the
// P-point DFT first, then output j twiddled by W_{PM}^{jk}.
template <int P, bool DIF>
__device__ __forceinline__ void bfly(float2* a, int base, int M, int k, int tws, int L,
                                     const float2* __restrict__ tw) {
  float2 z[P], o[P];
#pragma unroll
  for (int q = 0; q < P; q++) {
    const float2 v = a[base + q * M];
    z[q] = (!DIF && q) ? cmul(v, tw[q * k * tws]) : v;
  }
  if (P == 2) {
    o[0] = cadd(z[0], z[1]);
    o[1] = csub(z[0], z[1]);
  } else if (P == 4) {
    const float2 t0 = cadd(z[0], z[2]), t1 = csub(z[0], z[2]);
    const float2 t2 = cadd(z[1], z[3]), t3 = csub(z[1], z[3]);
    o[0] = cadd(t0, t2);
    o[2] = csub(t0, t2);
    o[1] = make_float2(t1.x + t3.y, t1.y - t3.x);  // t1 - i t3
    o[3] = make_float2(t1.x - t3.y, t1.y + t3.x);  // t1 + i t3
  } else {
    const int wp = L / P;
#pragma unroll
    for (int j = 0; j < P; j++) {
      float2 acc = z[0];
#pragma unroll
      for (int q = 1; q < P; q++) acc = cadd(acc, cmul(z[q], tw[((q * j) % P) * wp]));
      o[j] = acc;
    }
  }
#pragma unroll
  for (int j = 0; j < P; j++) a[base + j * M] = (DIF && j) ? cmul(o[j], tw[j * k * tws]) : o[j];
}

// Any radix up to 64 (prime factors 11..61), p^2 complex products.
__device__ void bfly_any(float2* a, int p, bool dif, int base, int M, int k, int tws, int L,
                         const float2* __restrict__ tw) {
  float2 z[64], o[64];
  for (int q = 0; q < p; q++) {
    const float2 v = a[base + q * M];
    z[q] = (!dif && q) ? cmul(v, tw[q * k * tws]) : v;
  }
  const int wp = L / p;
  for (int j = 0; j < p; j++) {
    float2 acc = z[0];
    for (int q = 1; q < p; q++) acc = cadd(acc, cmul(z[q], tw[((q * j) % p) * wp]));
    o[j] = acc;
  }
  for (int j = 0; j < p; j++) a[base + j * M] = (dif && j) ? cmul(o[j], tw[j * k * tws]) : o[j];
}

template <bool DIF>
__device__ __forceinline__ void pass(float2* a, int p, int M, const FftDesc& f,
                                     const float2* __restrict__ tw) {
  const int pM = p * M, nb = f.L / p, tws = f.L / pM;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const int k = b % M, base = (b / M) * pM + k;
    switch (p) {
      case 2: bfly<2, DIF>(a, base, M, k, tws, f.L, tw); break;
      case 3: bfly<3, DIF>(a, base, M, k, tws, f.L, tw); break;
      case 4: bfly<4, DIF>(a, base, M, k, tws, f.L, tw); break;
      case 5: bfly<5, DIF>(a, base, M, k, tws, f.L, tw); break;
      case 7: bfly<7, DIF>(a, base, M, k, tws, f.L, tw); break;
      default: bfly_any(a, p, DIF, base, M, k, tws, f.L, tw); break;
    }
  }
  __syncthreads();
}

// Forward transform of the L values in LDS, digit-reversed on entry (the
// load permutation), natural order on exit.  All threads take part.
__device__ void fft_dit(float2* a, const FftDesc& f, const float2* __restrict__ tw) {
  int M = 1;
  for (int l = f.m - 1; l >= 0; --l) {
    pass<false>(a, f.p[l], M, f, tw);
    M *= f.p[l];
  }
}

// The same transform from natural order: bin n ends at a[perm[n]] (the DFT
// matrix is symmetric, so the transposed DIT stages in reverse order give it).
__device__ void fft_dif(float2* a, const FftDesc& f, const float2* __restrict__ tw) {
  int M = f.L;
  for (int l = 0; l < f.m; ++l) {
    M /= f.p[l];
    pass<true>(a, f.p[l], M, f, tw);
  }
}

// Traces r, r+1 (r = 2*blockIdx.x) zero-padded to L, times padwin, packed as
// x_r + i x_{r+1}, into LDS at their digit-reversed positions.
__device__ void load_pair(float2* a, const FftDesc& f, const int* __restrict__ perm, int nt,
                          const float* in, int stride, int ntr, const float* __restrict__ padwin) {
  const int r = 2 * blockIdx.x;
  const bool two = r + 1 < ntr;
  for (int n = threadIdx.x; n < f.L; n += blockDim.x) {
    float x1 = 0.f, x2 = 0.f;
    if (n < nt) {
      const size_t o = (size_t)n * stride + r;
      x1 = in[o];
      if (two) x2 = in[o + 1];
      if (padwin) {
        const float w = padwin[n];
        x1 *= w;
        x2 *= w;
      }
    }
    a[perm[n]] = make_float2(x1, x2);
  }
  __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_trace_filter(FftDesc f, const int* __restrict__ perm,
                                                           const float2* __restrict__ tw, int nt,
                                                           TraceSet ts,
                                                           const float* __restrict__ padwin,
                                                           const float2* __restrict__ H, int kcov,
                                                           float scale) {
  extern __shared__ float2 lds[];
  load_pair(lds, f, perm, nt, ts.in, ts.in_stride, ts.ntr, padwin);
  fft_dit(lds, f, tw);
  // filter with the Hermitian extension of H, and conjugate: the second
  // forward transform then yields the conjugate of the inverse transform
  const int L = f.L, half = L / 2;
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    float2 v = lds[k];
    const int kk = k <= half ? k : L - k;
    if (kk < kcov) {
      float2 h = H[kk];
      if (kk == 0 || kk == half)
        h.y = 0.f;  // DC and Nyquist: C2R uses the real part only
      else if (k > half)
        h.y = -h.y;
      v = cmul(v, h);
    }
    lds[k] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  fft_dif(lds, f, tw);
  const int r = 2 * blockIdx.x;
  const bool two = r + 1 < ts.ntr;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    const float2 v = lds[perm[t]];
    const size_t o = (size_t)t * ts.out_stride + r;
    ts.out[o] = v.x * scale;
    if (two) ts.out[o + 1] = -v.y * scale;
  }
}

__global__ __launch_bounds__(kThreads) void k_trace_spectra(FftDesc f, const int* __restrict__ perm,
                                                            const float2* __restrict__ tw, int nt,
                                                            const float* __restrict__ in,
                                                            int stride, int ntr,
                                                            const float* __restrict__ padwin,
                                                            float2* __restrict__ spec) {
  extern __shared__ float2 lds[];
  load_pair(lds, f, perm, nt, in, stride, ntr, padwin);  // one read per element, no aliasing
  fft_dit(lds, f, tw);
  const int L = f.L, nf = L / 2 + 1;
  const int r = 2 * blockIdx.x;
  const bool two = r + 1 < ntr;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    const float2 z = lds[k], w = lds[(L - k) % L];
    // X1 = (Z[k] + conj Z[L-k]) / 2, X2 = (Z[k] - conj Z[L-k]) / (2i)
    spec[(size_t)r * nf + k] = make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
    if (two) spec[(size_t)(r + 1) * nf + k] = make_float2(0.5f * (z.y + w.y), -0.5f * (z.x - w.x));
  }
}

constexpr int kRatioRows = 16;  // receiver slices of one k_spectrum_ratio block

__global__ __launch_bounds__(1024) void k_spectrum_ratio(const float2* __restrict__ so,
                                                         const float2* __restrict__ sc, int ntr,
                                                         int nf, float2* __restrict__ coef) {
  __shared__ double red[3][kRatioRows][64];
  const int k = blockIdx.x * 64 + threadIdx.x;
  double nr = 0.0, ni = 0.0, dd = 0.0;
  if (k < nf)
    for (int r = threadIdx.y; r < ntr; r += kRatioRows) {
      const float2 c = sc[(size_t)r * nf + k], o = so[(size_t)r * nf + k];
      const double cx = c.x, cy = c.y, ox = o.x, oy = o.y;
      nr += cx * ox + cy * oy;  // conj(c) * o
      ni += cx * oy - cy * ox;
      dd += cx * cx + cy * cy;
    }
  red[0][threadIdx.y][threadIdx.x] = nr;
  red[1][threadIdx.y][threadIdx.x] = ni;
  red[2][threadIdx.y][threadIdx.x] = dd;
  __syncthreads();
  if (threadIdx.y == 0 && k < nf) {
    for (int y = 1; y < kRatioRows; y++) {
      nr += red[0][y][threadIdx.x];
      ni += red[1][y][threadIdx.x];
      dd += red[2][y][threadIdx.x];
    }
    const double den = dd + 1e-6;
    float2 c = make_float2((float)(nr / den), (float)(ni / den));
    if (k == 0 || k == nf - 1) c.y = 0.f;
    coef[k] = c;
  }
}

__global__ __launch_bounds__(256) void k_window_traces(int nt, TraceSet ts, const float* __restrict__ w2,
                                                       const float* __restrict__ ws,
                                                       const float* __restrict__ we,
                                                       const float* __restrict__ wt, float ratio,
                                                       float dt) {
  const int ntr = ts.ntr;
  const size_t total = (size_t)nt * ntr;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(i / (size_t)ntr), r = (int)(i % (size_t)ntr);
    float v = ts.in[(size_t)t * ts.in_stride + r];
    if (!ws) {
      v = v * w2[t];
    } else if (!(ratio > 0.5f)) {
      // cuda_window with windows (utilities.cu:653-704), float as there,
      // the ramps evaluated in double
      float t0 = ws[r], t3 = we[r];
      const float t_max = (float)nt * dt;
      if (t0 < 0.f) t0 = 0.f;
      if (t0 > t_max) t0 = t_max;
      if (t3 < 0.f) t3 = 0.f;
      if (t3 > t_max) t3 = t_max;
      const float offset = (t3 - t0) * ratio;
      if (offset > 0.f) {
        const float t1 = t0 + offset, t2 = t3 - offset;
        const float tt = (float)t * dt;
        float amp;
        if (tt >= t0 && tt < t1)
          amp = (float)sin(kPI / 2.0 * (double)(tt - t0) / (double)(t1 - t0));
        else if (tt >= t1 && tt < t2)
          amp = 1.f;
        else if (tt >= t2 && tt < t3)
          amp = (float)cos(kPI / 2.0 * (double)(tt - t2) / (double)(t3 - t2));
        else
          amp = 0.f;
        v = v * (amp * amp * wt[r]);
      }
    }
    ts.out[(size_t)t * ts.out_stride + r] = v;
  }
}

__global__ __launch_bounds__(256) void k_shot_residual(int nt, int ntr, const float* __restrict__ obs,
                                                       const float* __restrict__ syn, int in_stride,
                                                       float* __restrict__ res, int res_stride,
                                                       double* __restrict__ partial) {
  __shared__ double red[256];
  const size_t total = (size_t)nt * ntr;
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(i / (size_t)ntr), r = (int)(i % (size_t)ntr);
    const size_t o = (size_t)t * in_stride + r;
    const float rv = t > 0 ? obs[o] - syn[o] : 0.f;
    acc += (double)rv * (double)rv;
    res[(size_t)t * res_stride + r] = rv;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

int grid_for(size_t total) {
  const size_t b = (total + 255) / 256;
  return (int)(b < 4096 ? (b ? b : 1) : 4096);
}

hipError_t set_lds(const void* fn, size_t bytes) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace

Status fft_plan(int L, FftDesc& f, std::vector<int>& perm, std::vector<float2>& tw) {
  if (L < 2 || L > kMaxTransform)
    return Status::Err(FWIFLOW_ERR_UNSUPPORTED, "transform length out of range");
  f.L = L;
  f.m = 0;
  int n = L;
  auto take = [&](int p) {
    while (n % p == 0 && f.m < 32) {
      f.p[f.m++] = p;
      n /= p;
    }
  };
  take(4);
  take(2);
  for (int p = 3; p <= kMaxRadix && n > 1; p += 2) take(p);
  if (n != 1) return Status::Err(FWIFLOW_ERR_UNSUPPORTED, "transform length has a large prime factor");
  // input index n = q0 + p0 (q1 + p1 (q2 + ...)) lands at
  // sum_l q_l * L / (p0 ... p_l), the slot of its leaf in the recursion
  perm.assign(L, 0);
  for (int i = 0; i < L; i++) {
    int v = i, span = L, pos = 0;
    for (int l = 0; l < f.m; l++) {
      const int q = v % f.p[l];
      v /= f.p[l];
      span /= f.p[l];
      pos += q * span;
    }
    perm[i] = pos;
  }
  tw.resize(L);
  for (int j = 0; j < L; j++) {
    const double a = 2.0 * kPI * (double)j / (double)L;
    tw[j] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  return Status::Ok();
}

int touched_bins(int nt) {
  const int grid = (nt + 31) / 32 * 32;
  return grid < nt + 1 ? grid : nt + 1;
}

void bandpass_table(int nt, float dt, const float* f, std::vector<float2>& H) {
  const int L = 2 * nt, nf = nt + 1, kc = touched_bins(nt);
  H.assign(nf, make_float2(1.f, 0.f));
  const float df = (float)(1.0 / dt / L);
  for (int k = 0; k < kc; k++) {
    const float freq = (float)k * df;
    float a;
    if (freq >= f[0] && freq < f[1])
      a = (float)std::sin(kPI / 2.0 * (double)(freq - f[0]) / (double)(f[1] - f[0]));
    else if (freq >= f[1] && freq < f[2])
      a = 1.f;
    else if (freq >= f[2] && freq < f[3])
      a = (float)std::cos(kPI / 2.0 * (double)(freq - f[2]) / (double)(f[3] - f[2]));
    else
      a = 0.f;
    H[k] = make_float2(a * a, 0.f);
  }
}

hipError_t trace_filter(hipStream_t st, const FftDesc& f, const int* perm, const float2* tw, int nt,
                        TraceSet ts, const float* padwin, const float2* H, int kcov, float scale) {
  if (ts.ntr <= 0) return hipSuccess;
  const size_t lds = (size_t)f.L * sizeof(float2);
  hipError_t e = set_lds((const void*)k_trace_filter, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_trace_filter, dim3((ts.ntr + 1) / 2), dim3(kThreads), lds, st, f, perm, tw, nt,
                     ts, padwin, H, kcov, scale);
  return hipGetLastError();
}

hipError_t trace_spectra(hipStream_t st, const FftDesc& f, const int* perm, const float2* tw, int nt,
                         TraceSet ts, const float* padwin, float2* spec) {
  if (ts.ntr <= 0) return hipSuccess;
  const size_t lds = (size_t)f.L * sizeof(float2);
  hipError_t e = set_lds((const void*)k_trace_spectra, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_trace_spectra, dim3((ts.ntr + 1) / 2), dim3(kThreads), lds, st, f, perm, tw,
                     nt, ts.in, ts.in_stride, ts.ntr, padwin, spec);
  return hipGetLastError();
}

hipError_t spectrum_ratio(hipStream_t st, const float2* spec_obs, const float2* spec_cal, int ntr,
                          int nf, float2* coef) {
  hipLaunchKernelGGL(k_spectrum_ratio, dim3((nf + 63) / 64), dim3(64, kRatioRows), 0, st, spec_obs,
                     spec_cal, ntr, nf, coef);
  return hipGetLastError();
}

hipError_t window_traces(hipStream_t st, int nt, TraceSet ts, const float* w2, const float* wstart,
                         const float* wend, const float* weight, float ratio, float dt) {
  if (ts.ntr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_window_traces, dim3(grid_for((size_t)nt * ts.ntr)), dim3(256), 0, st, nt, ts,
                     w2, wstart, wend, weight, ratio, dt);
  return hipGetLastError();
}

hipError_t shot_residual(hipStream_t st, int nt, int ntr, const float* obs, const float* syn,
                         int in_stride, float* res, int res_stride, double* partial, int blocks) {
  hipLaunchKernelGGL(k_shot_residual, dim3(blocks), dim3(256), 0, st, nt, ntr, obs, syn, in_stride,
                     res, res_stride, partial);
  return hipGetLastError();
}

}  // namespace fwiflow
