// CPU oracle for FwiFlow's elastic FWI operator (libCUFD `cufd`).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; only tests/, the smoke check and the CPU-baseline leg of
// bench.py use it, and only as the checker.
//
// It restates, on the CPU, the algorithm of
//   deps/CustomOps/FWI/Src/libCUFD.cu:34-580   (driver, shot loop, time loops)
//   deps/CustomOps/FWI/Src/el_stress.cu:24-131 (stress update fwd / reverse + lambda,mu grads)
//   deps/CustomOps/FWI/Src/el_velocity.cu:21-117 (velocity update fwd / reverse + rho grad)
//   deps/CustomOps/FWI/Src/el_velocity_adj.cu:22-108, el_stress_adj.cu:22-104 (adjoint)
//   deps/CustomOps/FWI/Src/utilities.cu:58-747 (init, averages, misfit, Courant,
//                                               cpmlInit, ring store, source/receiver, window)
//   deps/CustomOps/FWI/Src/utilities.cu:596-1112, libCUFD.cu:254-309 (optional data
//                                               conditioning: per-trace windows, band-pass,
//                                               source-signature update)
//   deps/CustomOps/FWI/Src/Model.cu:15-93, Cpml.cu:7-117, Src_Rec.cu:19-245, Boundary.cu:5-102
//
// Two precisions:
//   T = float : the reference's arithmetic.  Expressions keep the reference's
//               literal types (2.0, 1e6, 3.0 are double), so C++'s usual
//               conversions promote exactly where the CUDA source promotes.
//               Compiled with -ffp-contract=off (no FMA contraction).
//   T = double: an FP64 run of the same scheme.
// Backward schemes:
//   scheme 0  : the reference's: a 5-layer boundary ring of all 5 wavefield
//               components is stored for every step and the forward wavefield
//               is re-materialised by running the update backwards in time.
//   scheme 1  : exact: full forward state checkpoints every `ckpt_every` steps,
//               segments recomputed forward and replayed backwards.  In exact
//               arithmetic both schemes give the same gradient.
//   scheme 3  : no reference counterpart: the exact adjoint (transpose of the
//               forward step) on the forward states of scheme 1.
//   scheme 2  : scheme 0 plus interior resets: the 5 components on the interior
//               rectangle [nPml, NZ-nPad-1-nPml] x [nPml, NX-1-nPml] are saved
//               before forward step it and written back after reverse step it,
//               for it a positive multiple of `ckpt_every` (bounds the FP32
//               drift of the reversal; no reference counterpart).
// Illumination (no reference counterpart; the preconditioner input of the
// inversion driver): sum over shots and steps it = 0..nSteps-2 of
// vz^2 + vx^2 after the velocity update, over the update region
// [2, NZ-nPad-3] x [2, NX-3], accumulated in double.
//
// Layout inside: z fastest, index = x*NZ + z (el_stress.cu:1-19).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPI = 3.141592653589793238462643383279502884197169;  // utilities.h:15
constexpr double kMEGA = 1e6;                                         // utilities.h:16
constexpr double kRSXXZZ = 3.0;                                       // utilities.h:21

struct Geom {
  int nz, nx, nPml, nPad, nSteps;
  double dz, dx, dt, f0;
};

template <typename T>
struct Prm {
  // the reference keeps dz, dx, dt, f0 as float (Parameter.h:191-195)
  T dz, dx, dt, f0;
  explicit Prm(const Geom& g) : dz((T)g.dz), dx((T)g.dx), dt((T)g.dt), f0((T)g.f0) {}
};

// ---------------------------------------------------------------- planted mistakes
// Mistakes a restatement of the reference could plausibly make, planted on
// request so the oracle's own tests can show that they would catch them
// (0: none, the reference as written).  Set by oracle_set_mutation.
enum Mutation {
  kMutNone = 0,
  kMutVelGate = 1,         // el_velocity: right x-PML gate of el_stress (x > nx-nPml-1)
  kMutAdjVelASwap = 2,     // el_velocity_adj vx: a_x_half in place of a_x
  kMutAdjVelKSwap = 3,     // el_velocity_adj vx: K_x_half in place of K_x
  kMutAdjVelDropPsi = 4,   // el_velocity_adj vx: a_x * d(psi_xx)/dx dropped
  kMutAdjStressDropPsi = 5,  // el_stress_adj sxz: a_x * d(phi_xz_x)/dx dropped
  kMutAdjStressASwap = 6,  // el_stress_adj sxz: a_x_half in place of a_x
};
// Not a mistake, a variant (a flag combined with the above): el_stress_adj
// with the PML gates that the reference comments out (el_stress_adj.cu:66-72,
// 89-96) restored, so its memory variables stay zero in the interior.
constexpr int kVarAdjStressGated = 0x100;
static int g_mutation = kMutNone;

// ---------------------------------------------------------------- C-PML
// utilities.cu:242-358 (cpmlInit).  CpAve is forced to 3000 (utilities.cu:259).
template <typename T>
void cpml_init(T* K, T* a, T* b, T* K_half, T* a_half, T* b_half, int N, int nPml, T dh, T f0,
               T dt) {
  std::vector<T> damp(N, 0), damp_half(N, 0), alpha(N, 0), alpha_half(N, 0);
  T d0_h = 0.0;
  T Rcoef = 0.0008;
  T depth_in_pml = 0.0;
  T depth_normalized = 0.0;
  T thickness_PML = 0.0;
  const T K_MAX_PML = 2.0;
  const T ALPHA_MAX_PML = 2.0 * kPI * (f0 / 2.0);
  const T NPOWER = 8.0;
  const T c1 = 0.25, c2 = 0.75, c3 = 0.0;
  thickness_PML = nPml * dh;
  T CpAve = 3000.0;
  d0_h = -(NPOWER + 1) * CpAve * std::log(Rcoef) / (2.0 * thickness_PML);
  for (int i = 0; i < N; i++) {
    K[i] = 1.0;
    K_half[i] = 1.0;
    a[i] = 0.0;
    a_half[i] = 0.0;
    b[i] = 0.0;
    b_half[i] = 0.0;
  }
  for (int i = 0; i < N; i++) {
    // left edge
    depth_in_pml = (nPml - i) * dh;
    if (depth_in_pml >= 0.0) {
      depth_normalized = depth_in_pml / thickness_PML;
      damp[i] = d0_h * (c1 * depth_normalized + c2 * std::pow(depth_normalized, NPOWER) +
                        c3 * std::pow(depth_normalized, 2 * NPOWER));
      K[i] = 1.0 + (K_MAX_PML - 1.0) * std::pow(depth_normalized, NPOWER);
      alpha[i] = ALPHA_MAX_PML * (1.0 - depth_normalized);
    }
    depth_in_pml = (nPml - i - 0.5) * dh;
    if (depth_in_pml >= 0.0) {
      depth_normalized = depth_in_pml / thickness_PML;
      damp_half[i] = d0_h * (c1 * depth_normalized + c2 * std::pow(depth_normalized, NPOWER) +
                             c3 * std::pow(depth_normalized, 2 * NPOWER));
      K_half[i] = 1.0 + (K_MAX_PML - 1.0) * std::pow(depth_normalized, NPOWER);
      alpha_half[i] = ALPHA_MAX_PML * (1.0 - depth_normalized);
    }
    // right edge
    depth_in_pml = (nPml - N + i) * dh;
    if (depth_in_pml >= 0.0) {
      depth_normalized = depth_in_pml / thickness_PML;
      damp[i] = d0_h * (c1 * depth_normalized + c2 * std::pow(depth_normalized, NPOWER) +
                        c3 * std::pow(depth_normalized, 2 * NPOWER));
      K[i] = 1.0 + (K_MAX_PML - 1.0) * std::pow(depth_normalized, NPOWER);
      alpha[i] = ALPHA_MAX_PML * (1.0 - depth_normalized);
    }
    depth_in_pml = (nPml - N + i + 0.5) * dh;
    if (depth_in_pml >= 0.0) {
      depth_normalized = depth_in_pml / thickness_PML;
      damp_half[i] = d0_h * (c1 * depth_normalized + c2 * std::pow(depth_normalized, NPOWER) +
                             c3 * std::pow(depth_normalized, 2 * NPOWER));
      K_half[i] = 1.0 + (K_MAX_PML - 1.0) * std::pow(depth_normalized, NPOWER);
      alpha_half[i] = ALPHA_MAX_PML * (1.0 - depth_normalized);
    }
    if (alpha[i] < 0.0) alpha[i] = 0.0;
    if (alpha_half[i] < 0.0) alpha_half[i] = 0.0;
    b[i] = std::exp((T)(-(damp[i] / K[i] + alpha[i]) * dt));
    b_half[i] = std::exp((T)(-(damp_half[i] / K_half[i] + alpha_half[i]) * dt));
    if (std::fabs(damp[i]) > 1.0e-6) {
      a[i] = damp[i] * (b[i] - 1.0) / (K[i] * (damp[i] + K[i] * alpha[i]));
    }
    if (std::fabs(damp_half[i]) > 1.0e-6) {
      a_half[i] = damp_half[i] * (b_half[i] - 1.0) /
                  (K_half[i] * (damp_half[i] + K_half[i] * alpha_half[i]));
    }
  }
}

template <typename T>
struct Cpml {
  // z profiles have length NZ - nPad (Cpml.cu:46-48), x profiles length NX
  std::vector<T> K_z, a_z, b_z, K_z_half, a_z_half, b_z_half;
  std::vector<T> K_x, a_x, b_x, K_x_half, a_x_half, b_x_half;
  Cpml(const Geom& g, const Prm<T>& p) {
    const int Nz = g.nz - g.nPad, Nx = g.nx;
    for (auto* v : {&K_z, &a_z, &b_z, &K_z_half, &a_z_half, &b_z_half}) v->assign(Nz, 0);
    for (auto* v : {&K_x, &a_x, &b_x, &K_x_half, &a_x_half, &b_x_half}) v->assign(Nx, 0);
    cpml_init<T>(K_z.data(), a_z.data(), b_z.data(), K_z_half.data(), a_z_half.data(),
                 b_z_half.data(), Nz, g.nPml, p.dz, p.f0, p.dt);
    cpml_init<T>(K_x.data(), a_x.data(), b_x.data(), K_x_half.data(), a_x_half.data(),
                 b_x_half.data(), Nx, g.nPml, p.dx, p.f0, p.dt);
  }
};

// ---------------------------------------------------------------- model
// Model.cu:15-93 with velInit / aveMuInit / aveBycInit (utilities.cu:109-152)
template <typename T>
struct Model {
  int nz, nx;
  std::vector<T> lam, mu, den, ave_mu, byc_a, byc_b, cp;
  Model(const Geom& g, const double* Lambda, const double* Mu, const double* Den) {
    nz = g.nz;
    nx = g.nx;
    const size_t n = (size_t)nz * nx;
    lam.resize(n);
    mu.resize(n);
    den.resize(n);
    // libCUFD.cu:72-78: transpose TF row-major [NZ][NX] to z-fastest, MPa -> Pa
    for (int i = 0; i < nz; i++)
      for (int j = 0; j < nx; j++) {
        lam[(size_t)j * nz + i] = (T)(Lambda[(size_t)i * nx + j] * kMEGA);
        mu[(size_t)j * nz + i] = (T)(Mu[(size_t)i * nx + j] * kMEGA);
        den[(size_t)j * nz + i] = (T)Den[(size_t)i * nx + j];
      }
    ave_mu.assign(n, (T)0.0);
    byc_a.assign(n, (T)(1.0 / 1000.0));
    byc_b.assign(n, (T)(1.0 / 1000.0));
    cp.assign(n, (T)0.0);
    for (size_t k = 0; k < n; k++) cp[k] = std::sqrt((lam[k] + 2.0 * mu[k]) / den[k]);
    for (int x = 2; x <= nx - 3; x++)
      for (int z = 2; z <= nz - 3; z++) {
        const size_t k = (size_t)x * nz + z;
        T a = mu[k], b = mu[k + 1], c = mu[k + nz], d = mu[k + nz + 1];
        if (a == 0.0 || b == 0.0 || c == 0.0 || d == 0.0)
          ave_mu[k] = 0.0;
        else
          ave_mu[k] = 4.0 / (1.0 / a + 1.0 / b + 1.0 / c + 1.0 / d);
        byc_a[k] = 2.0 / (den[k + 1] + den[k]);
        byc_b[k] = 2.0 / (den[k + nz] + den[k]);
      }
  }
};

// ---------------------------------------------------------------- window
// utilities.cu:707-747 (cuda_window without per-trace windows), applied per
// trace to an [ntr][nt] array (time fastest).
template <typename T>
void window_default(int nt, int ntr, T dt, T ratio, T* data) {
  for (int idt = 0; idt < nt; idt++) {
    T window_amp = 1.0;
    T t = idt * dt;
    T t0 = 0;
    T t3 = nt * dt;
    T offset = nt * dt * ratio;
    if (2.0 * offset >= t3 - t0) return;  // "Window error 2": data left as is
    T t1 = t0 + offset;
    T t2 = t3 - offset;
    if (t >= t0 && t < t1)
      window_amp = std::sin(kPI / 2.0 * (t - t0) / (t1 - t0));
    else if (t >= t1 && t < t2)
      window_amp = 1.0;
    else if (t >= t2 && t < t3)
      window_amp = std::cos(kPI / 2.0 * (t - t2) / (t3 - t2));
    else
      window_amp = 0.0;
    for (int r = 0; r < ntr; r++) data[(size_t)r * nt + idt] *= window_amp * window_amp;
  }
}

// ---------------------------------------------------------------- data conditioning
// The optional residual stages of libCUFD.cu:254-309 (all off by default,
// Parameter.cpp:146-174): per-trace windows and weights, band-pass filter,
// source-signature update.  Data are [ntr][nt], time fastest.
//
// The reference transforms with cuFFT in FP32; here every transform is an
// exact DFT evaluated in double by a mixed-radix decimation-in-time
// recursion, and values are rounded to T where the reference stores floats.
// The reference zero-pads a trace to 2*nt samples with intialArrayGPU
// launched over only ceil(nt/32)*32 columns of each 2*nt row
// (utilities.cu:886-887, 946-947, 1051-1052), so it reads uninitialised
// device memory beyond that; the restatement pads with zeros (what a fresh
// allocation holds).
using cplx = std::complex<double>;

// y[k] = sum_n x[n] exp(sign * 2 pi i n k / N); tw[j] = exp(sign * 2 pi i j / N0)
// of the top-level length N0, `twm` = N0 / N.
void dft_rec(const cplx* x, size_t stride, cplx* y, int N, const cplx* tw, size_t twm, int N0) {
  if (N == 1) {
    y[0] = x[0];
    return;
  }
  int p = 2;
  while (N % p) p++;
  const int m = N / p;
  std::vector<cplx> sub(N);
  for (int q = 0; q < p; q++) dft_rec(x + q * stride, stride * p, &sub[(size_t)q * m], m, tw, twm * p, N0);
  for (int k = 0; k < m; k++)
    for (int j = 0; j < p; j++) {
      cplx acc = 0.0;
      for (int q = 0; q < p; q++)
        acc += sub[(size_t)q * m + k] * tw[((size_t)q * (k + (size_t)m * j) % N) * twm];
      y[k + (size_t)m * j] = acc;
    }
}

std::vector<cplx> dft(const std::vector<cplx>& x, int sign) {
  const int N = (int)x.size();
  std::vector<cplx> tw(N), y(N);
  for (int j = 0; j < N; j++) tw[j] = std::polar(1.0, sign * 2.0 * kPI * j / N);
  dft_rec(x.data(), 1, y.data(), N, tw.data(), 1, N);
  return y;
}

// R2C of a length-nt trace zero-padded to L: bins 0..L/2.
template <typename T>
std::vector<cplx> rfft_pad(const T* x, int nt, int L) {
  std::vector<cplx> a(L, 0.0);
  for (int t = 0; t < nt && t < L; t++) a[t] = (double)x[t];
  std::vector<cplx> X = dft(a, -1);
  X.resize(L / 2 + 1);
  return X;
}

// C2R (unnormalised) of half spectrum X (bins 0..L/2): the Hermitian
// extension, imaginary parts of the DC and Nyquist bins ignored.
std::vector<double> irfft(const std::vector<cplx>& X, int L) {
  std::vector<cplx> a(L);
  for (int k = 0; k <= L / 2; k++) a[k] = X[k];
  a[0] = a[0].real();
  if (L % 2 == 0) a[L / 2] = a[L / 2].real();
  for (int k = L / 2 + 1; k < L; k++) a[k] = std::conj(X[L - k]);
  std::vector<cplx> y = dft(a, +1);
  std::vector<double> out(L);
  for (int t = 0; t < L; t++) out[t] = y[t].real();
  return out;
}

// cuda_window with per-trace windows and weights (utilities.cu:653-704).
// An element is left as it is when ratio > 0.5 or the window is empty
// (the kernel's error paths return before scaling).
template <typename T>
void window_trace(int nt, int ntr, T dt, const T* wstart, const T* wend, const T* weight, T ratio,
                  T* data) {
  if (ratio > 0.5) return;
  for (int r = 0; r < ntr; r++) {
    T t0 = wstart[r], t3 = wend[r];
    const T t_max = nt * dt;
    if (t0 < 0.0) t0 = 0.0;
    if (t0 > t_max) t0 = t_max;
    if (t3 < 0.0) t3 = 0.0;
    if (t3 > t_max) t3 = t_max;
    const T offset = (t3 - t0) * ratio;
    if (offset <= 0.0) continue;
    const T t1 = t0 + offset, t2 = t3 - offset;
    for (int idt = 0; idt < nt; idt++) {
      T window_amp = 1.0;
      const T t = idt * dt;
      if (t >= t0 && t < t1)
        window_amp = std::sin(kPI / 2.0 * (t - t0) / (t1 - t0));
      else if (t >= t1 && t < t2)
        window_amp = 1.0;
      else if (t >= t2 && t < t3)
        window_amp = std::cos(kPI / 2.0 * (t - t2) / (t3 - t2));
      else
        window_amp = 0.0;
      data[(size_t)r * nt + idt] *= window_amp * window_amp * weight[r];
    }
  }
}

// Bins 0..kcov-1 of the half spectrum that the reference's frequency-domain
// kernels touch: they run on a grid sized for nt (not nt+1) columns of TX = 32
// threads (utilities.cu:879-880, 1042, utilities.h:18), so the Nyquist bin
// k = nt is skipped when nt is a multiple of 32.
inline int touched_bins(int nt) { return std::min(nt + 1, (nt + 31) / 32 * 32); }

// Squared band-pass response amp(f)^2 per bin of the 2*nt-point grid
// (cuda_bp_filter1d, utilities.cu:596-626); 1 on bins the kernel skips.
template <typename T>
std::vector<T> bandpass_response(int nt, T dt, const T* f) {
  const int L = 2 * nt, nf = nt + 1, kc = touched_bins(nt);
  std::vector<T> h(nf, (T)1);
  const T df = 1.0 / dt / L;
  for (int k = 0; k < kc; k++) {
    const T freq = k * df;
    T a;
    if (freq >= f[0] && freq < f[1])
      a = std::sin(kPI / 2.0 * (freq - f[0]) / (f[1] - f[0]));
    else if (freq >= f[1] && freq < f[2])
      a = 1.0;
    else if (freq >= f[2] && freq < f[3])
      a = std::cos(kPI / 2.0 * (freq - f[2]) / (f[3] - f[2]));
    else
      a = 0.0;
    h[k] = a * a;
  }
  return h;
}

// bp_filter1d (utilities.cu:872-925): pad to 2*nt, R2C, times amp^2, C2R,
// crop, times 1/(2*nt).
template <typename T>
void bp_filter(int nt, int ntr, T dt, const T* f, T* data) {
  const int L = 2 * nt;
  const std::vector<T> h = bandpass_response<T>(nt, dt, f);
  const T factor = (T)1 / (T)L;
  for (int r = 0; r < ntr; r++) {
    T* x = data + (size_t)r * nt;
    std::vector<cplx> X = rfft_pad(x, nt, L);
    for (size_t k = 0; k < X.size(); k++) X[k] *= (double)h[k];
    const std::vector<double> y = irfft(X, L);
    for (int t = 0; t < nt; t++) x[t] = (T)y[t] * factor;
  }
}

// source_update (utilities.cu:927-1035): obs and cal padded to 2*nt and
// windowed (ratio 0.01 over the padded length), transformed; per bin
// coef = sum_r conj(C_r) O_r / (sum_r |C_r|^2 + 1e-6); cal and the source
// (padded, not windowed) are multiplied by coef and transformed back
// (crop, times 1/(2*nt)).  Returns coef (bins 0..nt).
template <typename T>
std::vector<cplx> source_update(int nt, int ntr, T dt, const T* obs, T* cal, T* src) {
  const int L = 2 * nt, nf = nt + 1;
  std::vector<std::vector<cplx>> O(ntr), Cc(ntr);
  for (int r = 0; r < ntr; r++) {
    std::vector<T> po(L, (T)0), pc(L, (T)0);
    std::copy(obs + (size_t)r * nt, obs + (size_t)(r + 1) * nt, po.begin());
    std::copy(cal + (size_t)r * nt, cal + (size_t)(r + 1) * nt, pc.begin());
    window_default<T>(L, 1, dt, (T)0.01, po.data());
    window_default<T>(L, 1, dt, (T)0.01, pc.data());
    O[r] = rfft_pad(po.data(), L, L);
    Cc[r] = rfft_pad(pc.data(), L, L);
  }
  std::vector<cplx> S = rfft_pad(src, nt, L), coef(nf);
  for (int k = 0; k < nf; k++) {
    cplx num = 0.0;
    double den = 0.0;
    for (int r = 0; r < ntr; r++) {
      num += std::conj(Cc[r][k]) * O[r][k];
      den += std::norm(Cc[r][k]);
    }
    cplx c = num / (den + 1e-6);
    coef[k] = cplx((T)c.real(), (T)c.imag());  // the reference keeps coef in FP32 complex
    S[k] *= coef[k];
    for (int r = 0; r < ntr; r++) Cc[r][k] *= coef[k];
  }
  const T factor = (T)1 / (T)L;
  for (int r = 0; r < ntr; r++) {
    const std::vector<double> y = irfft(Cc[r], L);
    for (int t = 0; t < nt; t++) cal[(size_t)r * nt + t] = (T)y[t] * factor;
  }
  const std::vector<double> ys = irfft(S, L);
  for (int t = 0; t < nt; t++) src[t] = (T)ys[t] * factor;
  return coef;
}

// source_update_adj (utilities.cu:1037-1082) with amp_ratio = 1
// (libCUFD.cu:287): pad, window (0.01 over the padded length), R2C, times coef
// (not its conjugate) on the touched bins, C2R, crop, times 1/(2*nt).
template <typename T>
void source_update_adj(int nt, int ntr, T dt, const std::vector<cplx>& coef, T* data) {
  const int L = 2 * nt, kc = touched_bins(nt);
  const T factor = (T)1 / (T)L;
  for (int r = 0; r < ntr; r++) {
    T* x = data + (size_t)r * nt;
    std::vector<T> p(L, (T)0);
    std::copy(x, x + nt, p.begin());
    window_default<T>(L, 1, dt, (T)0.01, p.data());
    std::vector<cplx> X = rfft_pad(p.data(), L, L);
    for (int k = 0; k < kc; k++) X[k] *= coef[k];
    const std::vector<double> y = irfft(X, L);
    for (int t = 0; t < nt; t++) x[t] = (T)y[t] * factor;
  }
}

// ---------------------------------------------------------------- state
template <typename T>
struct State {
  size_t n;
  std::vector<T> vz, vx, szz, sxx, sxz;
  // memory variables, named as in libCUFD.cu:105-106
  std::vector<T> m_dvz_dz, m_dvz_dx, m_dvx_dz, m_dvx_dx;
  std::vector<T> m_dszz_dz, m_dsxx_dx, m_dsxz_dz, m_dsxz_dx;
  explicit State(size_t n_) : n(n_) { zero(); }
  std::vector<std::vector<T>*> all() {
    return {&vz, &vx, &szz, &sxx, &sxz, &m_dvz_dz, &m_dvz_dx, &m_dvx_dz,
            &m_dvx_dx, &m_dszz_dz, &m_dsxx_dx, &m_dsxz_dz, &m_dsxz_dx};
  }
  void zero() {
    for (auto* v : all()) v->assign(n, (T)0);
  }
};

// Optional residual stages (Parameter.cpp:146-174, Src_Rec.cu:156-199).
struct Cond {
  bool if_win = false, if_filter = false, if_src_update = false;
  double filter[4] = {0, 0, 0, 0};
};

struct Ctx {
  Geom g;
  int nthreads;
  Cond cond;
};

#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(nthreads)")

// el_stress.cu:50-88, isFor = true
template <typename T>
void el_stress_fwd(const Ctx& c, const Prm<T>& p, const Model<T>& m, const Cpml<T>& pm,
                   State<T>& s) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  T *vz = s.vz.data(), *vx = s.vx.data(), *szz = s.szz.data(), *sxx = s.sxx.data(),
    *sxz = s.sxz.data();
  T *mvzz = s.m_dvz_dz.data(), *mvxx = s.m_dvx_dx.data(), *mvxz = s.m_dvx_dz.data(),
    *mvzx = s.m_dvz_dx.data();
  const T *lam = m.lam.data(), *mu = m.mu.data(), *amu = m.ave_mu.data();
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nz - nPad - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      T dvz_dz = (c1 * (vz[i] - vz[i - 1]) - c2 * (vz[i + 1] - vz[i - 2])) / dz;
      T dvx_dx = (c1 * (vx[i] - vx[i - nz]) - c2 * (vx[i + nz] - vx[i - 2 * nz])) / dx;
      if (z < nPml || (z > nz - nPml - nPad - 1)) {
        mvzz[i] = pm.b_z[z] * mvzz[i] + pm.a_z[z] * dvz_dz;
        dvz_dz = dvz_dz / pm.K_z[z] + mvzz[i];
      }
      if (x < nPml || x > nx - nPml - 1) {
        mvxx[i] = pm.b_x[x] * mvxx[i] + pm.a_x[x] * dvx_dx;
        dvx_dx = dvx_dx / pm.K_x[x] + mvxx[i];
      }
      szz[i] += ((lam[i] + 2.0 * mu[i]) * dvz_dz + lam[i] * dvx_dx) * dt;
      sxx[i] += (lam[i] * dvz_dz + (lam[i] + 2.0 * mu[i]) * dvx_dx) * dt;
      T dvx_dz = (c1 * (vx[i + 1] - vx[i]) - c2 * (vx[i + 2] - vx[i - 1])) / dz;
      T dvz_dx = (c1 * (vz[i + nz] - vz[i]) - c2 * (vz[i + 2 * nz] - vz[i - nz])) / dx;
      if (z < nPml || (z > nz - nPml - nPad - 1)) {
        mvxz[i] = pm.b_z_half[z] * mvxz[i] + pm.a_z_half[z] * dvx_dz;
        dvx_dz = dvx_dz / pm.K_z_half[z] + mvxz[i];
      }
      if (x < nPml || x > nx - nPml - 1) {
        mvzx[i] = pm.b_x_half[x] * mvzx[i] + pm.a_x_half[x] * dvz_dx;
        dvz_dx = dvz_dx / pm.K_x_half[x] + mvzx[i];
      }
      sxz[i] = sxz[i] + amu[i] * (dvx_dz + dvz_dx) * dt;
    }
  }
}

// el_velocity.cu:45-82, isFor = true.  Note the x-PML test `x > nx - nPml`
// (el_velocity.cu:56,71), one column narrower on the right than el_stress's.
template <typename T>
void el_velocity_fwd(const Ctx& c, const Prm<T>& p, const Model<T>& m, const Cpml<T>& pm,
                     State<T>& s) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  T *vz = s.vz.data(), *vx = s.vx.data(), *szz = s.szz.data(), *sxx = s.sxx.data(),
    *sxz = s.sxz.data();
  T *mszz = s.m_dszz_dz.data(), *msxzx = s.m_dsxz_dx.data(), *msxzz = s.m_dsxz_dz.data(),
    *msxx = s.m_dsxx_dx.data();
  const T *ba = m.byc_a.data(), *bb = m.byc_b.data();
  const int xr = (g_mutation & 0xff) == kMutVelGate ? nx - nPml - 1 : nx - nPml;
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nz - nPad - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      T dszz_dz = (c1 * (szz[i + 1] - szz[i]) - c2 * (szz[i + 2] - szz[i - 1])) / dz;
      T dsxz_dx = (c1 * (sxz[i] - sxz[i - nz]) - c2 * (sxz[i + nz] - sxz[i - 2 * nz])) / dx;
      if (z < nPml || (z > nz - nPml - nPad - 1)) {
        mszz[i] = pm.b_z_half[z] * mszz[i] + pm.a_z_half[z] * dszz_dz;
        dszz_dz = dszz_dz / pm.K_z_half[z] + mszz[i];
      }
      if (x < nPml || x > xr) {
        msxzx[i] = pm.b_x[x] * msxzx[i] + pm.a_x[x] * dsxz_dx;
        dsxz_dx = dsxz_dx / pm.K_x[x] + msxzx[i];
      }
      vz[i] += (dszz_dz + dsxz_dx) * ba[i] * dt;
      T dsxz_dz = (c1 * (sxz[i] - sxz[i - 1]) - c2 * (sxz[i + 1] - sxz[i - 2])) / dz;
      T dsxx_dx = (c1 * (sxx[i + nz] - sxx[i]) - c2 * (sxx[i + 2 * nz] - sxx[i - nz])) / dx;
      if (z < nPml || (z > nz - nPml - nPad - 1)) {
        msxzz[i] = pm.b_z[z] * msxzz[i] + pm.a_z[z] * dsxz_dz;
        dsxz_dz = dsxz_dz / pm.K_z[z] + msxzz[i];
      }
      if (x < nPml || x > xr) {
        msxx[i] = pm.b_x_half[x] * msxx[i] + pm.a_x_half[x] * dsxx_dx;
        dsxx_dx = dsxx_dx / pm.K_x_half[x] + msxx[i];
      }
      vx[i] += (dsxz_dz + dsxx_dx) * bb[i] * dt;
    }
  }
}

// Forward wavefield as seen by the gradient: only the 5 components.
template <typename T>
struct Wave {
  T *vz, *vx, *szz, *sxx, *sxz;
};

// el_velocity.cu:84-117 (isFor = false): reverse velocity step over the
// interior plus the density gradient.  The reference sprays the buoyancy
// derivative with atomicAdd; here the 4 contributions per cell are computed in
// a parallel pass and sprayed serially in the same target order.  The test
// `gidx+1<=gidx<=nx-1-nPml` (el_velocity.cu:109) is always true.
template <typename T>
void el_velocity_rev(const Ctx& c, const Prm<T>& p, const Model<T>& m, Wave<T> w, bool update,
                     const T* vz_adj, const T* vx_adj, T* den_grad, std::vector<T>& ga,
                     std::vector<T>& gb) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  const T *szz = w.szz, *sxx = w.sxx, *sxz = w.sxz;
  const T *ba = m.byc_a.data(), *bb = m.byc_b.data();
  const int z0 = nPml, z1 = nz - nPad - 1 - nPml, x0 = nPml, x1 = nx - 1 - nPml;
  OMP_FOR
  for (int x = x0; x <= x1; x++) {
    for (int z = z0; z <= z1; z++) {
      const size_t i = (size_t)x * nz + z;
      T dszz_dz = (c1 * (szz[i + 1] - szz[i]) - c2 * (szz[i + 2] - szz[i - 1])) / dz;
      T dsxz_dx = (c1 * (sxz[i] - sxz[i - nz]) - c2 * (sxz[i + nz] - sxz[i - 2 * nz])) / dx;
      if (update) w.vz[i] -= (dszz_dz + dsxz_dx) * ba[i] * dt;
      T dsxz_dz = (c1 * (sxz[i] - sxz[i - 1]) - c2 * (sxz[i + 1] - sxz[i - 2])) / dz;
      T dsxx_dx = (c1 * (sxx[i + nz] - sxx[i]) - c2 * (sxx[i + 2 * nz] - sxx[i - nz])) / dx;
      if (update) w.vx[i] -= (dsxz_dz + dsxx_dx) * bb[i] * dt;
      ga[i] = -vz_adj[i] * (dszz_dz + dsxz_dx) * dt * (-std::pow(ba[i], 2) / 2.0);
      gb[i] = -vx_adj[i] * (dsxz_dz + dsxx_dx) * dt * (-std::pow(bb[i], 2) / 2.0);
    }
  }
  for (int x = x0; x <= x1; x++) {
    for (int z = z0; z <= z1; z++) {
      const size_t i = (size_t)x * nz + z;
      den_grad[i] += ga[i];
      den_grad[i] += gb[i];
      if (z + 1 <= nz - nPad - 1 - nPml) den_grad[i + 1] += ga[i];
      den_grad[i + nz] += gb[i];
    }
  }
}

// el_stress.cu:90-131 (isFor = false): reverse stress step over the interior
// plus lambda and mu gradients (x1e6: the caller's moduli are in MPa).  The
// shear term is sprayed onto the 4 mu cells of the harmonic average with
// d(avg)/d(mu_k) = avg^2/(4 mu_k^2).
template <typename T>
void el_stress_rev(const Ctx& c, const Prm<T>& p, const Model<T>& m, Wave<T> w, bool update,
                   const T* szz_adj, const T* sxx_adj, const T* sxz_adj, T* lam_grad, T* mu_grad,
                   std::vector<T>& tsp) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  const T *vz = w.vz, *vx = w.vx;
  const T *lam = m.lam.data(), *mu = m.mu.data(), *amu = m.ave_mu.data();
  const int z0 = nPml, z1 = nz - nPad - 1 - nPml, x0 = nPml, x1 = nx - 1 - nPml;
  OMP_FOR
  for (int x = x0; x <= x1; x++) {
    for (int z = z0; z <= z1; z++) {
      const size_t i = (size_t)x * nz + z;
      T dvz_dz = (c1 * (vz[i] - vz[i - 1]) - c2 * (vz[i + 1] - vz[i - 2])) / dz;
      T dvx_dx = (c1 * (vx[i] - vx[i - nz]) - c2 * (vx[i + nz] - vx[i - 2 * nz])) / dx;
      if (update) {
        w.szz[i] -= ((lam[i] + 2.0 * mu[i]) * dvz_dz + lam[i] * dvx_dx) * dt;
        w.sxx[i] -= (lam[i] * dvz_dz + (lam[i] + 2.0 * mu[i]) * dvx_dx) * dt;
      }
      T dvx_dz = (c1 * (vx[i + 1] - vx[i]) - c2 * (vx[i + 2] - vx[i - 1])) / dz;
      T dvz_dx = (c1 * (vz[i + nz] - vz[i]) - c2 * (vz[i + 2 * nz] - vz[i - nz])) / dx;
      if (update) w.sxz[i] -= amu[i] * (dvx_dz + dvz_dx) * dt;
      lam_grad[i] += -(szz_adj[i] + sxx_adj[i]) * (dvz_dz + dvx_dx) * dt * kMEGA;
      mu_grad[i] += (-2.0 * szz_adj[i] * dvz_dz * dt - 2.0 * sxx_adj[i] * dvx_dx * dt) * kMEGA;
      tsp[i] = 0;
      if (amu[i] != 0.0) {
        T scale = -sxz_adj[i] * (dvx_dz + dvz_dx) * dt * amu[i] /
                  (1.0 / mu[i] + 1.0 / mu[i + 1] + 1.0 / mu[i + nz] + 1.0 / mu[i + nz + 1]) * kMEGA;
        tsp[i] = scale;
      }
    }
  }
  for (int x = x0; x <= x1; x++) {
    for (int z = z0; z <= z1; z++) {
      const size_t i = (size_t)x * nz + z;
      if (amu[i] == 0.0) continue;
      const T scale = tsp[i];
      auto inv2 = [&](size_t k) { return 1.0 / ((double)mu[k] * (double)mu[k]); };
      mu_grad[i] += (T)(inv2(i) * scale);
      if (z + 1 <= nz - nPad - 1 - nPml) mu_grad[i + 1] += (T)(inv2(i + 1) * scale);
      mu_grad[i + nz] += (T)(inv2(i + nz) * scale);
      if (z + 1 <= nz - nPad - 1 - nPml && x + 1 <= nx - 1 - nPml)
        mu_grad[i + nz + 1] += (T)(inv2(i + nz + 1) * scale);
    }
  }
}

// el_velocity_adj.cu:22-108
template <typename T>
void el_velocity_adj(const Ctx& c, const Prm<T>& p, const Model<T>& m, const Cpml<T>& pm,
                     State<T>& a) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  T *vz = a.vz.data(), *vx = a.vx.data();
  const T *szz = a.szz.data(), *sxx = a.sxx.data(), *sxz = a.sxz.data();
  T *mszz = a.m_dszz_dz.data(), *msxzx = a.m_dsxz_dx.data(), *msxzz = a.m_dsxz_dz.data(),
    *msxx = a.m_dsxx_dx.data();
  const T *mvzz = a.m_dvz_dz.data(), *mvzx = a.m_dvz_dx.data(), *mvxz = a.m_dvx_dz.data(),
          *mvxx = a.m_dvx_dx.data();
  const T *lam = m.lam.data(), *mu = m.mu.data(), *amu = m.ave_mu.data();
  const T *ba = m.byc_a.data(), *bb = m.byc_b.data();
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nz - nPad - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      const T lambda = lam[i], muv = mu[i];
      T dpsixx_dx = (-c1 * (mvxx[i + nz] - mvxx[i]) + c2 * (mvxx[i + 2 * nz] - mvxx[i - nz])) / dx;
      T dszz_dx = (-c1 * (szz[i + nz] - szz[i]) + c2 * (szz[i + 2 * nz] - szz[i - nz])) / dx;
      T dsxx_dx = (-c1 * (sxx[i + nz] - sxx[i]) + c2 * (sxx[i + 2 * nz] - sxx[i - nz])) / dx;
      T dpsixz_dz = (-c1 * (mvxz[i] - mvxz[i - 1]) + c2 * (mvxz[i + 1] - mvxz[i - 2])) / dz;
      T dsxz_dz = (-c1 * (sxz[i] - sxz[i - 1]) + c2 * (sxz[i + 1] - sxz[i - 2])) / dz;
      const T ax = (g_mutation & 0xff) == kMutAdjVelASwap ? pm.a_x_half[x] : pm.a_x[x];
      const T kx = (g_mutation & 0xff) == kMutAdjVelKSwap ? pm.K_x_half[x] : pm.K_x[x];
      const T psi_xx = (g_mutation & 0xff) == kMutAdjVelDropPsi ? (T)0 : ax * dpsixx_dx;
      vx[i] += (psi_xx + lambda * dszz_dx / kx * dt + (lambda + 2.0 * muv) * dsxx_dx / kx * dt +
                pm.a_z_half[z] * dpsixz_dz + amu[i] / pm.K_z_half[z] * dsxz_dz * dt);
      if (x < nPml || x > nx - nPml - 1) msxx[i] = pm.b_x_half[x] * msxx[i] + bb[i] * vx[i] * dt;
      if (z < nPml || (z > nz - nPml - nPad - 1))
        msxzz[i] = pm.b_z[z] * msxzz[i] + bb[i] * vx[i] * dt;
      T dpsizz_dz = (-c1 * (mvzz[i + 1] - mvzz[i]) + c2 * (mvzz[i + 2] - mvzz[i - 1])) / dz;
      T dszz_dz = (-c1 * (szz[i + 1] - szz[i]) + c2 * (szz[i + 2] - szz[i - 1])) / dz;
      T dsxx_dz = (-c1 * (sxx[i + 1] - sxx[i]) + c2 * (sxx[i + 2] - sxx[i - 1])) / dz;
      T dpsizx_dx = (-c1 * (mvzx[i] - mvzx[i - nz]) + c2 * (mvzx[i + nz] - mvzx[i - 2 * nz])) / dx;
      T dsxz_dx = (-c1 * (sxz[i] - sxz[i - nz]) + c2 * (sxz[i + nz] - sxz[i - 2 * nz])) / dx;
      vz[i] += (pm.a_z[z] * dpsizz_dz + (lambda + 2.0 * muv) * dszz_dz / pm.K_z[z] * dt +
                lambda * dsxx_dz / pm.K_z[z] * dt + pm.a_x_half[x] * dpsizx_dx +
                amu[i] / pm.K_x_half[x] * dsxz_dx * dt);
      if (x < nPml || x > nx - nPml - 1) msxzx[i] = pm.b_x[x] * msxzx[i] + ba[i] * vz[i] * dt;
      if (z < nPml || (z > nz - nPml - nPad - 1))
        mszz[i] = pm.b_z_half[z] * mszz[i] + ba[i] * vz[i] * dt;
    }
  }
}

// el_stress_adj.cu:22-104.  The 4 stress-side memory variables are updated
// over the whole update region (their PML gates are commented out in the
// reference), so they accumulate in the interior where b = 1.
template <typename T>
void el_stress_adj(const Ctx& c, const Prm<T>& p, const Model<T>& m, const Cpml<T>& pm,
                   State<T>& a) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  const T *vz = a.vz.data(), *vx = a.vx.data();
  T *szz = a.szz.data(), *sxx = a.sxx.data(), *sxz = a.sxz.data();
  const T *mszz = a.m_dszz_dz.data(), *msxzx = a.m_dsxz_dx.data(), *msxzz = a.m_dsxz_dz.data(),
          *msxx = a.m_dsxx_dx.data();
  T *mvzz = a.m_dvz_dz.data(), *mvzx = a.m_dvz_dx.data(), *mvxz = a.m_dvx_dz.data(),
    *mvxx = a.m_dvx_dx.data();
  const T *lam = m.lam.data(), *mu = m.mu.data(), *amu = m.ave_mu.data();
  const T *ba = m.byc_a.data(), *bb = m.byc_b.data();
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nz - nPad - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      const T lambda = lam[i], muv = mu[i];
      T dphi_xz_x_dx =
          (-c1 * (msxzx[i + nz] - msxzx[i]) + c2 * (msxzx[i + 2 * nz] - msxzx[i - nz])) / dx;
      T dvz_dx = (-c1 * (vz[i + nz] - vz[i]) + c2 * (vz[i + 2 * nz] - vz[i - nz])) / dx;
      T dphi_xz_z_dz = (-c1 * (msxzz[i + 1] - msxzz[i]) + c2 * (msxzz[i + 2] - msxzz[i - 1])) / dz;
      T dvx_dz = (-c1 * (vx[i + 1] - vx[i]) + c2 * (vx[i + 2] - vx[i - 1])) / dz;
      const T ax = (g_mutation & 0xff) == kMutAdjStressASwap ? pm.a_x_half[x] : pm.a_x[x];
      const T phi_xz_x = (g_mutation & 0xff) == kMutAdjStressDropPsi ? (T)0 : ax * dphi_xz_x_dx;
      sxz[i] += phi_xz_x + dvz_dx / pm.K_x[x] * ba[i] * dt + pm.a_z[z] * dphi_xz_z_dz +
                dvx_dz / pm.K_z[z] * bb[i] * dt;
      const bool gated = (g_mutation & kVarAdjStressGated) != 0;
      const bool xin = x < nPml || x > nx - nPml - 1, zin = z < nPml || z > nz - nPml - nPad - 1;
      if (!gated || xin) mvzx[i] = pm.b_x_half[x] * mvzx[i] + sxz[i] * amu[i] * dt;
      if (!gated || zin) mvxz[i] = pm.b_z_half[z] * mvxz[i] + sxz[i] * amu[i] * dt;
      T dphi_xx_x_dx =
          (-c1 * (msxx[i] - msxx[i - nz]) + c2 * (msxx[i + nz] - msxx[i - 2 * nz])) / dx;
      T dvx_dx = (-c1 * (vx[i] - vx[i - nz]) + c2 * (vx[i + nz] - vx[i - 2 * nz])) / dx;
      T dphi_zz_z_dz = (-c1 * (mszz[i] - mszz[i - 1]) + c2 * (mszz[i + 1] - mszz[i - 2])) / dz;
      T dvz_dz = (-c1 * (vz[i] - vz[i - 1]) + c2 * (vz[i + 1] - vz[i - 2])) / dz;
      sxx[i] += pm.a_x_half[x] * dphi_xx_x_dx + bb[i] * dvx_dx / pm.K_x_half[x] * dt;
      szz[i] += pm.a_z_half[z] * dphi_zz_z_dz + ba[i] * dvz_dz / pm.K_z_half[z] * dt;
      if (!gated || xin)
        mvxx[i] = pm.b_x[x] * mvxx[i] + lambda * szz[i] * dt + (lambda + 2.0 * muv) * sxx[i] * dt;
      if (!gated || zin)
        mvzz[i] = pm.b_z[z] * mvzz[i] + (lambda + 2.0 * muv) * szz[i] * dt + lambda * sxx[i] * dt;
    }
  }
}

// ---------------------------------------------------------------- exact adjoint
// No reference counterpart: the transpose of the forward time step above
// (el_stress_fwd, add_source, el_velocity_fwd, recording), taken operation
// by operation, so that its gradient is the derivative of the discrete
// misfit in every setting (C-PML, heterogeneous media, Vs = 0 layers).  The
// adjoint variables are dJ/d(state): a stencil difference f -> D f becomes
// its transpose acting on the coefficient-weighted adjoint (coefficients
// inside the derivative), and a C-PML term d~ = d/K + psi', psi' = b psi +
// a d becomes d_adj = g/K + a (g + P), P <- b (g + P) for the adjoint P of
// psi.  Forward gates (including el_velocity's x > nx-nPml) are kept.

// f_adj[j] += (D^T B)[j] for D = D+ (f[k+1]-f[k], c2: f[k+2]-f[k-1]) along a
// stride (1: z, nz: x) over the update region's rows k (B zero elsewhere)
template <typename T>
inline T dplus_T(const T* B, size_t j, size_t st, T c1, T c2, T h) {
  return (c1 * (B[j - st] - B[j]) - c2 * (B[j - 2 * st] - B[j + st])) / h;
}
// D = D- (f[k]-f[k-1], c2: f[k+1]-f[k-2])
template <typename T>
inline T dminus_T(const T* B, size_t j, size_t st, T c1, T c2, T h) {
  return (c1 * (B[j] - B[j + st]) - c2 * (B[j - st] - B[j + 2 * st])) / h;
}

// one C-PML adjoint: g = dJ/d(d~) at a point with a gated memory variable
template <typename T>
inline T pml_T(bool gated, T g, T K, T a, T b, T& P) {
  if (!gated) return g;
  const T tot = g + P;
  P = b * tot;
  return g / K + a * tot;
}

// Transpose of el_velocity_fwd at step it: adds dJ/d(sigma) to a's
// stresses from a's velocities; rho terms of the gradient (interior, as the
// reference's) from the forward stresses sigma^{it+1} in w.
template <typename T>
void exact_velocity_T(const Ctx& c, const Prm<T>& p, const Model<T>& m, const Cpml<T>& pm,
                      State<T>& a, Wave<T> w, T* den_grad) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const int nzA = nz - nPad;
  const size_t n = (size_t)nz * nx;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  const T *ba = m.byc_a.data(), *bb = m.byc_b.data();
  std::vector<T> Bz(n, 0), Bzx(n, 0), Bxz(n, 0), Bx(n, 0), ga(n, 0), gb(n, 0);
  const int xr = (g_mutation & 0xff) == kMutVelGate ? nx - nPml - 1 : nx - nPml;
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nzA - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      const bool pz = z < nPml || z > nz - nPml - nPad - 1, px = x < nPml || x > xr;
      const T gz = dt * ba[i] * a.vz[i], gx = dt * bb[i] * a.vx[i];
      Bz[i] = pml_T(pz, gz, pm.K_z_half[z], pm.a_z_half[z], pm.b_z_half[z], a.m_dszz_dz[i]);
      Bzx[i] = pml_T(px, gz, pm.K_x[x], pm.a_x[x], pm.b_x[x], a.m_dsxz_dx[i]);
      Bxz[i] = pml_T(pz, gx, pm.K_z[z], pm.a_z[z], pm.b_z[z], a.m_dsxz_dz[i]);
      Bx[i] = pml_T(px, gx, pm.K_x_half[x], pm.a_x_half[x], pm.b_x_half[x], a.m_dsxx_dx[i]);
      if (z >= nPml && z <= nzA - 1 - nPml && x >= nPml && x <= nx - 1 - nPml) {
        const T dszz_dz = (c1 * (w.szz[i + 1] - w.szz[i]) - c2 * (w.szz[i + 2] - w.szz[i - 1])) / dz;
        const T dsxz_dx = (c1 * (w.sxz[i] - w.sxz[i - nz]) - c2 * (w.sxz[i + nz] - w.sxz[i - 2 * nz])) / dx;
        const T dsxz_dz = (c1 * (w.sxz[i] - w.sxz[i - 1]) - c2 * (w.sxz[i + 1] - w.sxz[i - 2])) / dz;
        const T dsxx_dx = (c1 * (w.sxx[i + nz] - w.sxx[i]) - c2 * (w.sxx[i + 2 * nz] - w.sxx[i - nz])) / dx;
        // d(ba)/d(rho) = -ba^2/2 for both rho cells of the average
        ga[i] = a.vz[i] * dt * (dszz_dz + dsxz_dx) * (-ba[i] * ba[i] / 2);
        gb[i] = a.vx[i] * dt * (dsxz_dz + dsxx_dx) * (-bb[i] * bb[i] / 2);
      }
    }
  }
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nzA - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      a.szz[i] += dplus_T(Bz.data(), i, 1, c1, c2, dz);
      a.sxz[i] += dminus_T(Bzx.data(), i, (size_t)nz, c1, c2, dx) +
                  dminus_T(Bxz.data(), i, 1, c1, c2, dz);
      a.sxx[i] += dplus_T(Bx.data(), i, (size_t)nz, c1, c2, dx);
    }
  }
  for (int x = 2; x <= nx - 3; x++)
    for (int z = 2; z <= nzA - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      den_grad[i] += ga[i] + gb[i];
      den_grad[i + 1] += ga[i];
      den_grad[i + nz] += gb[i];
    }
}

// Transpose of el_stress_fwd at step it: adds dJ/d(v) to a's velocities from
// a's stresses; lambda / mu terms of the gradient (interior; x1e6 for MPa)
// from the forward velocities v^it in w, the mu spray onto the 4 cells of
// the harmonic average unconditionally.
template <typename T>
void exact_stress_T(const Ctx& c, const Prm<T>& p, const Model<T>& m, const Cpml<T>& pm,
                    State<T>& a, Wave<T> w, T* lam_grad, T* mu_grad) {
  const int nz = c.g.nz, nx = c.g.nx, nPml = c.g.nPml, nPad = c.g.nPad, nthreads = c.nthreads;
  const int nzA = nz - nPad;
  const size_t n = (size_t)nz * nx;
  const T c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const T dz = p.dz, dx = p.dx, dt = p.dt;
  const T *lam = m.lam.data(), *mu = m.mu.data(), *amu = m.ave_mu.data();
  std::vector<T> Azz(n, 0), Axx(n, 0), Axz(n, 0), Azx(n, 0), gs(n, 0);
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nzA - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      const bool pz = z < nPml || z > nz - nPml - nPad - 1, px = x < nPml || x > nx - nPml - 1;
      const T l2m = lam[i] + 2.0 * mu[i];
      const T Gzz = dt * (l2m * a.szz[i] + lam[i] * a.sxx[i]);
      const T Gxx = dt * (lam[i] * a.szz[i] + l2m * a.sxx[i]);
      const T Gs = dt * amu[i] * a.sxz[i];
      Azz[i] = pml_T(pz, Gzz, pm.K_z[z], pm.a_z[z], pm.b_z[z], a.m_dvz_dz[i]);
      Axx[i] = pml_T(px, Gxx, pm.K_x[x], pm.a_x[x], pm.b_x[x], a.m_dvx_dx[i]);
      Axz[i] = pml_T(pz, Gs, pm.K_z_half[z], pm.a_z_half[z], pm.b_z_half[z], a.m_dvx_dz[i]);
      Azx[i] = pml_T(px, Gs, pm.K_x_half[x], pm.a_x_half[x], pm.b_x_half[x], a.m_dvz_dx[i]);
      if (z >= nPml && z <= nzA - 1 - nPml && x >= nPml && x <= nx - 1 - nPml) {
        const T dvz_dz = (c1 * (w.vz[i] - w.vz[i - 1]) - c2 * (w.vz[i + 1] - w.vz[i - 2])) / dz;
        const T dvx_dx = (c1 * (w.vx[i] - w.vx[i - nz]) - c2 * (w.vx[i + nz] - w.vx[i - 2 * nz])) / dx;
        const T dvx_dz = (c1 * (w.vx[i + 1] - w.vx[i]) - c2 * (w.vx[i + 2] - w.vx[i - 1])) / dz;
        const T dvz_dx = (c1 * (w.vz[i + nz] - w.vz[i]) - c2 * (w.vz[i + 2 * nz] - w.vz[i - nz])) / dx;
        lam_grad[i] += (a.szz[i] + a.sxx[i]) * (dvz_dz + dvx_dx) * dt * kMEGA;
        mu_grad[i] += 2.0 * (a.szz[i] * dvz_dz + a.sxx[i] * dvx_dx) * dt * kMEGA;
        gs[i] = a.sxz[i] * (dvx_dz + dvz_dx) * dt * kMEGA;  // dJ/d(amu), per MPa
      }
    }
  }
  OMP_FOR
  for (int x = 2; x <= nx - 3; x++) {
    for (int z = 2; z <= nzA - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      a.vz[i] += dminus_T(Azz.data(), i, 1, c1, c2, dz) +
                 dplus_T(Azx.data(), i, (size_t)nz, c1, c2, dx);
      a.vx[i] += dminus_T(Axx.data(), i, (size_t)nz, c1, c2, dx) +
                 dplus_T(Axz.data(), i, 1, c1, c2, dz);
    }
  }
  // harmonic average: d(amu)/d(mu_k) = amu^2 / (4 mu_k^2)
  for (int x = 2; x <= nx - 3; x++)
    for (int z = 2; z <= nzA - 3; z++) {
      const size_t i = (size_t)x * nz + z;
      if (gs[i] == 0 || amu[i] == 0) continue;
      const double q = (double)amu[i] * (double)amu[i] / 4.0;
      for (size_t k : {i, i + 1, i + (size_t)nz, i + (size_t)nz + 1})
        mu_grad[k] += (T)((double)gs[i] * q / ((double)mu[k] * (double)mu[k]));
    }
}

// ---------------------------------------------------------------- ring
// Boundary.cu:17-41 and from_bnd/to_bnd (utilities.cu:361-424): for every step
// the 5 layers at rows/columns nPml-2 .. nPml+2 of each of the 4 sides.
struct Ring {
  std::vector<size_t> idx;
  Ring(const Geom& g) {
    const int nz = g.nz, nx = g.nx, nPml = g.nPml, nPad = g.nPad;
    const int nzBnd = nz - 2 * nPml - nPad + 4, nxBnd = nx - 2 * nPml + 4, L = 5;
    auto at = [&](int z, int x) { return (size_t)x * nz + z; };
    for (int j = 0; j < L; j++)
      for (int r = 0; r < nzBnd; r++) idx.push_back(at(r + nPml - 2, j + nPml - 2));
    for (int j = 0; j < L; j++)
      for (int r = 0; r < nzBnd; r++) idx.push_back(at(r + nPml - 2, nx - nPml - j - 1 + 2));
    for (int r = 0; r < L; r++)
      for (int j = 0; j < nxBnd; j++) idx.push_back(at(r + nPml - 2, j + nPml - 2));
    for (int r = 0; r < L; r++)
      for (int j = 0; j < nxBnd; j++)
        idx.push_back(at(nz - nPml - nPad - r - 1 + 2, j + nPml - 2));
  }
};

// ---------------------------------------------------------------- misfit
// cuda_cal_objective (utilities.cu:169-205): one block of 512 threads,
// strided partial sums of powf(a, 2) then a tree reduction.
template <typename T>
T cal_objective(const T* err, size_t ng) {
  const int B = 512;
  std::vector<T> sdata(B, (T)0);
  for (int tid = 0; tid < B; tid++) {
    for (size_t s = 0; s < (ng + B - 1) / B; s++) {
      size_t id = s * B + tid;
      T a = (id < ng) ? err[id] : (T)0;
      sdata[tid] += std::pow(a, (T)2);
    }
  }
  for (int s = B / 2; s >= 1; s /= 2)
    for (int tid = 0; tid < s; tid++) sdata[tid] += sdata[tid + s];
  return sdata[0];
}

struct ShotIn {
  int z_src, x_src, nrec;
  const int *z_rec, *x_rec;  // grid indices (JSON value + nPml)
  const double* stf;         // stf row of this shot, nSteps samples
  const float* obs;          // [nrec][nSteps] (calc 0/1)
  float* syn;                // out [nrec][nSteps]
  float* res;                // out [nrec][nSteps] (calc 0/1)
  // per-trace windows and weights of the shot (if_win), nrec each
  const double *win_start, *win_end, *weights;
  // optional outputs (calc 0/1): conditioned obs and syn [nrec][nSteps] and
  // the source after the update [nSteps] (libCUFD.cu:315-329)
  float *cond_obs, *cond_syn, *src_out;
};

template <typename T>
struct Runner {
  Ctx c;
  Prm<T> p;
  Model<T> m;
  Cpml<T> pm;
  int scheme, ckpt_every;
  size_t n;
  std::vector<T> lam_grad, mu_grad, den_grad;
  std::vector<T> ga, gb, tsp;
  std::vector<double> illum;  // sized only when requested

  Runner(const Ctx& c_, const double* L, const double* M, const double* D, int scheme_, int k)
      : c(c_), p(c_.g), m(c_.g, L, M, D), pm(c_.g, p), scheme(scheme_), ckpt_every(k) {
    n = (size_t)c.g.nz * c.g.nx;
    lam_grad.assign(n, 0);
    mu_grad.assign(n, 0);
    den_grad.assign(n, 0);
    ga.assign(n, 0);
    gb.assign(n, 0);
    tsp.assign(n, 0);
  }

  // add_source (utilities.cu:521-555): the 9x9 Gaussian exp(-1000 r^2) is 1 at
  // the centre and underflows to 0 elsewhere, so only the centre cell changes.
  // `float scale = pow(1500.0, 2)` (utilities.cu:528): the szz increment is a
  // float product, the sxx one (RSXXZZ * scale * ...) a double product.
  void add_source(State<T>& s, T amp, int zs, int xs, bool isFor) {
    const T dt = p.dt;
    const size_t i = (size_t)xs * c.g.nz + zs;
    const T scale = (T)std::pow(1500.0, 2);
    const T gauss = std::exp((T)(-1000.0 * 0));
    if (isFor) {
      s.szz[i] += scale * amp * dt * gauss;
      s.sxx[i] += kRSXXZZ * scale * amp * dt * gauss;
    } else {
      s.szz[i] -= scale * amp * dt * gauss;
      s.sxx[i] -= kRSXXZZ * scale * amp * dt * gauss;
    }
  }

  void forward_step(State<T>& s, int it, const std::vector<T>& src, const ShotIn& sh,
                    std::vector<T>* data) {
    el_stress_fwd(c, p, m, pm, s);
    add_source(s, src[it], sh.z_src, sh.x_src, true);
    el_velocity_fwd(c, p, m, pm, s);
    if (data) {
      // recording (utilities.cu:557-567): sample it+1 of each trace
      const int nz = c.g.nz, nSteps = c.g.nSteps;
      for (int r = 0; r < sh.nrec; r++) {
        const size_t i = (size_t)sh.x_rec[r] * nz + sh.z_rec[r];
        (*data)[(size_t)r * nSteps + it + 1] = s.szz[i] + kRSXXZZ * s.sxx[i];
      }
    }
  }

  // returns the shot's objective contribution (float sum, FP32 mode)
  T run_shot(int calc_id, const ShotIn& sh, double* stf_grad_row) {
    const Geom& g = c.g;
    const int nSteps = g.nSteps, nz = g.nz;
    const bool if_res = calc_id != 2, withAdj = calc_id == 1;
    // Src_Rec.cu:132-143: stf row -> float -> window(ratio 0.001)
    std::vector<T> src(nSteps);
    for (int it = 0; it < nSteps; it++) src[it] = (T)sh.stf[it];
    window_default<T>(nSteps, 1, p.dt, (T)0.001, src.data());
    std::vector<T> src_window(nSteps, (T)1);  // d(src)/d(stf), for the exact adjoint
    window_default<T>(nSteps, 1, p.dt, (T)0.001, src_window.data());
    const std::vector<T> src_fwd = src;  // the source update replaces src for the reversal

    State<T> s(n);
    std::vector<T> data((size_t)sh.nrec * nSteps, (T)0);
    Ring ring(g);
    const size_t rl = ring.idx.size();
    std::vector<T> rbuf;  // scheme 0: [it][field][ring]
    std::vector<std::vector<T>> ckpt;  // scheme 1
    const int k = std::max(1, ckpt_every);
    if (withAdj && (scheme == 0 || scheme == 2)) rbuf.assign((size_t)nSteps * 5 * rl, (T)0);
    std::vector<std::vector<T>> interior;  // scheme 2: [checkpoint][5 * NZ*NX]
    for (int it = 0; it <= nSteps - 2; it++) {
      if (withAdj && scheme == 2 && it > 0 && it % k == 0) {
        std::vector<T> cp;
        cp.reserve(5 * n);
        for (auto* v : {&s.vz, &s.vx, &s.szz, &s.sxx, &s.sxz}) cp.insert(cp.end(), v->begin(), v->end());
        interior.push_back(std::move(cp));
      }
      if (withAdj && (scheme == 0 || scheme == 2)) {
        T* dst = rbuf.data() + (size_t)it * 5 * rl;
        const T* f[5] = {s.szz.data(), s.sxz.data(), s.sxx.data(), s.vz.data(), s.vx.data()};
        for (int q = 0; q < 5; q++)
          for (size_t j = 0; j < rl; j++) dst[q * rl + j] = f[q][ring.idx[j]];
      }
      if (withAdj && (scheme == 1 || scheme == 3) && it % k == 0) {
        std::vector<T> cp;
        cp.reserve(13 * n);
        for (auto* v : s.all()) cp.insert(cp.end(), v->begin(), v->end());
        ckpt.push_back(std::move(cp));
      }
      forward_step(s, it, src, sh, &data);
      if (!illum.empty()) {
        const int nzA = g.nz - g.nPad;
        for (int x = 2; x <= g.nx - 3; x++)
          for (int z = 2; z <= nzA - 3; z++) {
            const size_t i = (size_t)x * nz + z;
            illum[i] += (double)s.vz[i] * (double)s.vz[i] + (double)s.vx[i] * (double)s.vx[i];
          }
      }
    }
    for (size_t j = 0; j < data.size(); j++) sh.syn[j] = (float)data[j];
    if (!if_res) return (T)0;

    // residual (libCUFD.cu:254-313)
    const Cond& cd = c.cond;
    std::vector<T> obs((size_t)sh.nrec * nSteps), res(obs.size());
    for (size_t j = 0; j < obs.size(); j++) obs[j] = (T)sh.obs[j];
    const T win_ratio = 0.005;
    std::vector<T> ws, we, wt;
    if (cd.if_win) {
      for (int r = 0; r < sh.nrec; r++) {
        ws.push_back((T)sh.win_start[r]);
        we.push_back((T)sh.win_end[r]);
        wt.push_back((T)sh.weights[r]);
      }
      window_trace<T>(nSteps, sh.nrec, p.dt, ws.data(), we.data(), wt.data(), win_ratio, obs.data());
      window_trace<T>(nSteps, sh.nrec, p.dt, ws.data(), we.data(), wt.data(), win_ratio, data.data());
    } else {
      window_default<T>(nSteps, sh.nrec, p.dt, win_ratio, obs.data());
      window_default<T>(nSteps, sh.nrec, p.dt, win_ratio, data.data());
    }
    T fl[4];
    for (int q = 0; q < 4; q++) fl[q] = (T)cd.filter[q];
    if (cd.if_filter) {
      bp_filter<T>(nSteps, sh.nrec, p.dt, fl, obs.data());
      bp_filter<T>(nSteps, sh.nrec, p.dt, fl, data.data());
    }
    std::vector<cplx> coef;
    if (cd.if_src_update) coef = source_update<T>(nSteps, sh.nrec, p.dt, obs.data(), data.data(), src.data());
    for (int r = 0; r < sh.nrec; r++)
      for (int t = 0; t < nSteps; t++) {
        const size_t j = (size_t)r * nSteps + t;
        res[j] = t > 0 ? obs[j] - data[j] : (T)0;
      }
    T obj = cal_objective(res.data(), res.size());
    if (cd.if_src_update) source_update_adj<T>(nSteps, sh.nrec, p.dt, coef, res.data());
    if (cd.if_filter) bp_filter<T>(nSteps, sh.nrec, p.dt, fl, res.data());
    if (cd.if_win)
      window_trace<T>(nSteps, sh.nrec, p.dt, ws.data(), we.data(), wt.data(), (T)0.1, res.data());
    else
      window_default<T>(nSteps, sh.nrec, p.dt, win_ratio, res.data());
    for (size_t j = 0; j < res.size(); j++) sh.res[j] = (float)res[j];
    if (sh.cond_obs)
      for (size_t j = 0; j < obs.size(); j++) sh.cond_obs[j] = (float)obs[j];
    if (sh.cond_syn)
      for (size_t j = 0; j < data.size(); j++) sh.cond_syn[j] = (float)data[j];
    if (sh.src_out)
      for (int it = 0; it < nSteps; it++) sh.src_out[it] = (float)src[it];
    if (!withAdj) return obj;

    // ------------------------------------------------ backward (libCUFD.cu:334-457)
    State<T> a(n);
    std::vector<T> stf_grad(nSteps, (T)0);
    auto res_injection = [&](int it) {
      for (int r = 0; r < sh.nrec; r++) {
        const size_t i = (size_t)sh.x_rec[r] * nz + sh.z_rec[r];
        a.szz[i] += res[(size_t)r * nSteps + it];
        a.sxx[i] += kRSXXZZ * res[(size_t)r * nSteps + it];
      }
    };
    auto source_grad = [&](int it) {
      const size_t i = (size_t)sh.x_src * nz + sh.z_src;
      stf_grad[it] = -(a.szz[i] + kRSXXZZ * a.sxx[i]) * p.dt;
    };
    if (scheme != 3) {
      el_velocity_adj(c, p, m, pm, a);
      res_injection(nSteps - 1);
      el_stress_adj(c, p, m, pm, a);
    }

    if (scheme == 0 || scheme == 2) {
      auto reset = [&](int it) {  // scheme 2: interior state before forward step it
        const std::vector<T>& cp = interior[(size_t)(it / k) - 1];
        const int nzA = g.nz - g.nPad, P0 = g.nPml;
        T* f[5] = {s.vz.data(), s.vx.data(), s.szz.data(), s.sxx.data(), s.sxz.data()};
        for (int q = 0; q < 5; q++)
          for (int x = P0; x <= g.nx - 1 - P0; x++)
            for (int z = P0; z <= nzA - 1 - P0; z++) {
              const size_t i = (size_t)x * nz + z;
              f[q][i] = cp[(size_t)q * n + i];
            }
      };
      auto restore = [&](int it, bool stress) {
        const T* src_ = rbuf.data() + (size_t)it * 5 * rl;
        if (stress) {
          T* f[3] = {s.szz.data(), s.sxz.data(), s.sxx.data()};
          for (int q = 0; q < 3; q++)
            for (size_t j = 0; j < rl; j++) f[q][ring.idx[j]] = src_[q * rl + j];
        } else {
          T* f[2] = {s.vz.data(), s.vx.data()};
          for (int q = 0; q < 2; q++)
            for (size_t j = 0; j < rl; j++) f[q][ring.idx[j]] = src_[(3 + q) * rl + j];
        }
      };
      Wave<T> w{s.vz.data(), s.vx.data(), s.szz.data(), s.sxx.data(), s.sxz.data()};
      for (int it = nSteps - 2; it >= 0; it--) {
        source_grad(it);
        el_velocity_rev(c, p, m, w, true, a.vz.data(), a.vx.data(), den_grad.data(), ga, gb);
        restore(it, false);
        add_source(s, src[it], sh.z_src, sh.x_src, false);
        el_stress_rev(c, p, m, w, true, a.szz.data(), a.sxx.data(), a.sxz.data(),
                      lam_grad.data(), mu_grad.data(), tsp);
        restore(it, true);
        if (scheme == 2 && it > 0 && it % k == 0) reset(it);
        el_velocity_adj(c, p, m, pm, a);
        res_injection(it);
        el_stress_adj(c, p, m, pm, a);
      }
    } else {
      // exact scheme: for each segment [s0, s1) recompute forward states from
      // the checkpoint, keeping sigma^{it+1} and v^{it} for every step.
      const int nseg = (int)ckpt.size();
      std::vector<T> sig((size_t)k * 3 * n), vel((size_t)k * 2 * n);
      for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * k, s1 = std::min(s0 + k, nSteps - 1);
        {
          size_t off = 0;
          for (auto* v : s.all()) {
            std::copy(ckpt[sg].begin() + off, ckpt[sg].begin() + off + n, v->begin());
            off += n;
          }
        }
        for (int it = s0; it < s1; it++) {
          const size_t q = (size_t)(it - s0);
          std::copy(s.vz.begin(), s.vz.end(), vel.begin() + (q * 2 + 0) * n);
          std::copy(s.vx.begin(), s.vx.end(), vel.begin() + (q * 2 + 1) * n);
          forward_step(s, it, src_fwd, sh, nullptr);
          std::copy(s.szz.begin(), s.szz.end(), sig.begin() + (q * 3 + 0) * n);
          std::copy(s.sxx.begin(), s.sxx.end(), sig.begin() + (q * 3 + 1) * n);
          std::copy(s.sxz.begin(), s.sxz.end(), sig.begin() + (q * 3 + 2) * n);
        }
        for (int it = s1 - 1; it >= s0; it--) {
          const size_t q = (size_t)(it - s0);
          T* vz = vel.data() + (q * 2 + 0) * n;
          T* vx = vel.data() + (q * 2 + 1) * n;
          T* szz = sig.data() + (q * 3 + 0) * n;
          T* sxx = sig.data() + (q * 3 + 1) * n;
          T* sxz = sig.data() + (q * 3 + 2) * n;
          if (scheme == 3) {
            // transpose of forward step it: recording of sample it+1 (the
            // residual is -dJ/d(data)), velocity update, source, stress update
            for (int r = 0; r < sh.nrec; r++) {
              const size_t i = (size_t)sh.x_rec[r] * nz + sh.z_rec[r];
              const T f = -res[(size_t)r * nSteps + it + 1];
              a.szz[i] += f;
              a.sxx[i] += kRSXXZZ * f;
            }
            exact_velocity_T(c, p, m, pm, a, Wave<T>{vz, vx, szz, sxx, sxz}, den_grad.data());
            const size_t is = (size_t)sh.x_src * nz + sh.z_src;
            stf_grad[it] = (T)std::pow(1500.0, 2) * p.dt * (a.szz[is] + kRSXXZZ * a.sxx[is]) *
                           src_window[it];
            exact_stress_T(c, p, m, pm, a, Wave<T>{vz, vx, szz, sxx, sxz}, lam_grad.data(),
                           mu_grad.data());
            continue;
          }
          source_grad(it);
          el_velocity_rev(c, p, m, Wave<T>{vz, vx, szz, sxx, sxz}, false, a.vz.data(),
                          a.vx.data(), den_grad.data(), ga, gb);
          el_stress_rev(c, p, m, Wave<T>{vz, vx, szz, sxx, sxz}, false, a.szz.data(),
                        a.sxx.data(), a.sxz.data(), lam_grad.data(), mu_grad.data(), tsp);
          el_velocity_adj(c, p, m, pm, a);
          res_injection(it);
          el_stress_adj(c, p, m, pm, a);
        }
      }
    }
    for (int it = 0; it < nSteps; it++) stf_grad_row[it] = (double)stf_grad[it];
    return obj;
  }
};

template <typename T>
int run_all(const Ctx& c, int scheme, int ckpt_every, int calc_id, int group_size,
            const ShotIn* shots, const double* L, const double* M, const double* D,
            double* misfit, double* gL, double* gM, double* gD, double* gS,
            double* illum) {
  Runner<T> R(c, L, M, D, scheme, ckpt_every);
  if (illum) R.illum.assign(R.n, 0.0);
  T h_l2Obj = 0.0;
  std::vector<double> dummy(c.g.nSteps);
  for (int s = 0; s < group_size; s++) {
    double* row = (calc_id == 1 && gS) ? gS + (size_t)s * c.g.nSteps : dummy.data();
    h_l2Obj += R.run_shot(calc_id, shots[s], row);
  }
  const int nz = c.g.nz, nx = c.g.nx;
  if (calc_id == 1) {
    for (int i = 0; i < nz; i++)
      for (int j = 0; j < nx; j++) {
        if (gL) gL[(size_t)i * nx + j] = R.lam_grad[(size_t)j * nz + i];
        if (gM) gM[(size_t)i * nx + j] = R.mu_grad[(size_t)j * nz + i];
        if (gD) gD[(size_t)i * nx + j] = R.den_grad[(size_t)j * nz + i];
      }
  }
  if (calc_id == 0 && misfit) *misfit = 0.5 * h_l2Obj;
  if (illum)
    for (int i = 0; i < nz; i++)
      for (int j = 0; j < nx; j++) illum[(size_t)i * nx + j] = R.illum[(size_t)j * nz + i];
  return 0;
}

}  // namespace

extern "C" {

// Plant one of the Mutation mistakes, optionally with the kVarAdjStressGated
// flag (0 removes both), for the oracle's tests.
void oracle_set_mutation(int m) { g_mutation = m; }

typedef struct {
  int nz, nx, nPml, nPad, nSteps;
  double dz, dx, dt, f0;
} oracle_geom;

// Optional residual stages ("if_win", "filter", "if_src_update" of the
// parameter file) and their outputs.  win_start/win_end/weights are
// concatenated over the group's shots (one value per receiver, seconds and
// weight); cond_obs/cond_syn are laid out like obs, src_out is
// [group_size][nSteps].  Output pointers may be NULL.
typedef struct {
  int if_win, if_filter, if_src_update;
  double filter[4];
  const double *win_start, *win_end, *weights;
  float *cond_obs, *cond_syn, *src_out;
} oracle_cond;

// Full cufd restatement.  Inputs follow the reference's conventions:
// Lambda/Mu in MPa and Den, row-major [NZ][NX]; stf rows already selected per
// shot ([group_size][nSteps]); receiver/source indices already offset by nPml
// (Src_Rec.cu:86,90,105,113).  obs/syn/res are concatenated per shot,
// [nrec][nSteps] each.  illum (optional, [NZ][NX]) receives the forward
// illumination; cond (optional) switches on the conditioning stages.
// Returns 0, or -1 when if_win is set without window arrays.
int oracle_cufd(int precision, int scheme, int ckpt_every, int calc_id, const oracle_geom* g,
                int group_size, const int* z_src, const int* x_src, const int* nrec,
                const int* z_rec, const int* x_rec, const double* Lambda, const double* Mu,
                const double* Den, const double* stf_rows, const float* obs, float* syn,
                float* res, double* misfit, double* grad_lambda, double* grad_mu,
                double* grad_den, double* grad_stf, int nthreads, double* illum,
                const oracle_cond* cond) {
  Ctx c;
  c.g = Geom{g->nz, g->nx, g->nPml, g->nPad, g->nSteps, g->dz, g->dx, g->dt, g->f0};
  c.nthreads = nthreads > 0 ? nthreads : 1;
  if (cond) {
    c.cond.if_win = cond->if_win != 0;
    c.cond.if_filter = cond->if_filter != 0;
    c.cond.if_src_update = cond->if_src_update != 0;
    for (int q = 0; q < 4; q++) c.cond.filter[q] = cond->filter[q];
    if (c.cond.if_win && (!cond->win_start || !cond->win_end || !cond->weights)) return -1;
  }
  std::vector<ShotIn> shots(group_size);
  size_t roff = 0, doff = 0;
  for (int s = 0; s < group_size; s++) {
    ShotIn& sh = shots[s];
    sh.z_src = z_src[s];
    sh.x_src = x_src[s];
    sh.nrec = nrec[s];
    sh.z_rec = z_rec + roff;
    sh.x_rec = x_rec + roff;
    sh.stf = stf_rows + (size_t)s * g->nSteps;
    sh.obs = obs ? obs + doff : nullptr;
    sh.syn = syn + doff;
    sh.res = res ? res + doff : nullptr;
    const bool w = cond && cond->if_win;
    sh.win_start = w ? cond->win_start + roff : nullptr;
    sh.win_end = w ? cond->win_end + roff : nullptr;
    sh.weights = w ? cond->weights + roff : nullptr;
    sh.cond_obs = cond && cond->cond_obs ? cond->cond_obs + doff : nullptr;
    sh.cond_syn = cond && cond->cond_syn ? cond->cond_syn + doff : nullptr;
    sh.src_out = cond && cond->src_out ? cond->src_out + (size_t)s * g->nSteps : nullptr;
    roff += nrec[s];
    doff += (size_t)nrec[s] * g->nSteps;
  }
  if (precision == 64)
    return run_all<double>(c, scheme, ckpt_every, calc_id, group_size, shots.data(), Lambda, Mu,
                           Den, misfit, grad_lambda, grad_mu, grad_den, grad_stf, illum);
  return run_all<float>(c, scheme, ckpt_every, calc_id, group_size, shots.data(), Lambda, Mu,
                        Den, misfit, grad_lambda, grad_mu, grad_den, grad_stf, illum);
}

// C-PML profiles of one axis, out = 6 x N: K, a, b, K_half, a_half, b_half.
void oracle_cpml(int precision, int N, int nPml, double dh, double f0, double dt, double* out) {
  if (precision == 64) {
    std::vector<double> v(6 * (size_t)N);
    cpml_init<double>(&v[0], &v[N], &v[2 * N], &v[3 * N], &v[4 * N], &v[5 * N], N, nPml, dh, f0,
                      dt);
    std::copy(v.begin(), v.end(), out);
  } else {
    std::vector<float> v(6 * (size_t)N);
    cpml_init<float>(&v[0], &v[N], &v[2 * N], &v[3 * N], &v[4 * N], &v[5 * N], N, nPml,
                     (float)dh, (float)f0, (float)dt);
    for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
  }
}

// Derived model arrays, z-fastest [x][z]: ave_mu, byc_a, byc_b, cp (4 x NZ*NX).
void oracle_model(int precision, const oracle_geom* g, const double* Lambda, const double* Mu,
                  const double* Den, double* out) {
  Geom G{g->nz, g->nx, g->nPml, g->nPad, g->nSteps, g->dz, g->dx, g->dt, g->f0};
  const size_t n = (size_t)g->nz * g->nx;
  if (precision == 64) {
    Model<double> m(G, Lambda, Mu, Den);
    for (size_t i = 0; i < n; i++) {
      out[i] = m.ave_mu[i];
      out[n + i] = m.byc_a[i];
      out[2 * n + i] = m.byc_b[i];
      out[3 * n + i] = m.cp[i];
    }
  } else {
    Model<float> m(G, Lambda, Mu, Den);
    for (size_t i = 0; i < n; i++) {
      out[i] = m.ave_mu[i];
      out[n + i] = m.byc_a[i];
      out[2 * n + i] = m.byc_b[i];
      out[3 * n + i] = m.cp[i];
    }
  }
}

// Time window applied in place to [ntr][nt] data.
void oracle_window(int precision, int nt, int ntr, double dt, double ratio, double* data) {
  if (precision == 64) {
    window_default<double>(nt, ntr, dt, ratio, data);
  } else {
    std::vector<float> v(data, data + (size_t)nt * ntr);
    window_default<float>(nt, ntr, (float)dt, (float)ratio, v.data());
    for (size_t i = 0; i < v.size(); i++) data[i] = v[i];
  }
}

// Exact DFT of n complex values (interleaved re, im), sign -1 forward, +1
// inverse (unnormalised), in place.
void oracle_dft(int n, int sign, double* data) {
  std::vector<cplx> x(n);
  for (int i = 0; i < n; i++) x[i] = cplx(data[2 * i], data[2 * i + 1]);
  const std::vector<cplx> y = dft(x, sign);
  for (int i = 0; i < n; i++) {
    data[2 * i] = y[i].real();
    data[2 * i + 1] = y[i].imag();
  }
}

// Conditioning stages on [ntr][nt] data in place (restated reference
// functions, for pinning).  op 0: per-trace window (params: ratio; arrays
// wstart, wend, weight); op 1: band-pass (params: f0..f3); op 2: source
// update (data = cal, aux = obs [ntr][nt], src = [nt] in/out, coef_out =
// 2*(nt+1) doubles); op 3: its adjoint (coef_in = 2*(nt+1) doubles).
int oracle_condition(int precision, int op, int nt, int ntr, double dt, const double* params,
                     const double* wstart, const double* wend, const double* weight,
                     double* data, const double* aux, double* src, double* coef) {
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    const size_t n = (size_t)nt * ntr;
    std::vector<T> d(data, data + n);
    if (op == 0) {
      std::vector<T> a(wstart, wstart + ntr), b(wend, wend + ntr), w(weight, weight + ntr);
      window_trace<T>(nt, ntr, (T)dt, a.data(), b.data(), w.data(), (T)params[0], d.data());
    } else if (op == 1) {
      T f[4] = {(T)params[0], (T)params[1], (T)params[2], (T)params[3]};
      bp_filter<T>(nt, ntr, (T)dt, f, d.data());
    } else if (op == 2) {
      std::vector<T> o(aux, aux + n), s(src, src + nt);
      const std::vector<cplx> cf = source_update<T>(nt, ntr, (T)dt, o.data(), d.data(), s.data());
      for (int t = 0; t < nt; t++) src[t] = s[t];
      for (int k = 0; k <= nt; k++) {
        coef[2 * k] = cf[k].real();
        coef[2 * k + 1] = cf[k].imag();
      }
    } else if (op == 3) {
      std::vector<cplx> cf(nt + 1);
      for (int k = 0; k <= nt; k++) cf[k] = cplx(coef[2 * k], coef[2 * k + 1]);
      source_update_adj<T>(nt, ntr, (T)dt, cf, d.data());
    } else {
      return -1;
    }
    for (size_t i = 0; i < n; i++) data[i] = d[i];
    return 0;
  };
  return precision == 64 ? run(0.0) : run(0.0f);
}

// Ring cell indices (z-fastest linear index), count returned; out may be null.
int oracle_ring(const oracle_geom* g, long long* out) {
  Geom G{g->nz, g->nx, g->nPml, g->nPad, g->nSteps, g->dz, g->dx, g->dt, g->f0};
  Ring r(G);
  if (out)
    for (size_t i = 0; i < r.idx.size(); i++) out[i] = (long long)r.idx[i];
  return (int)r.idx.size();
}

}  // extern "C"
