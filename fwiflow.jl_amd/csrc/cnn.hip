// gfx950 kernels of the learned preconditioner: a 4-layer 3x3 convolutional
// network (4 -> 32 -> 32 -> 32 -> 3 channels, GELU after the first three
// layers, zero padding), FP32 weights and accumulation, BF16 activations.
// No reference counterpart: FwiFlow's plain method is L-BFGS on the raw
// gradient; this network maps (gradient of cp, cs, rho; illumination) to a
// search direction and is trained on the plain inversion's accepted updates.
//
// Layouts: feature maps are planar [channel][row][column] (rows = depth z,
// columns = x fastest), i.e. the row-major [NZ][NX] of the operator's
// gradients, one plane per channel.  Parameters are one flat FP32 vector:
// per layer W [Cout][Cin][3][3] then b [Cout].
//
// Kernels:
//   k_conv       3x3 convolution over a 16x16 pixel tile per workgroup, the
//                input tile (+1 halo) staged in LDS 8 channels at a time,
//                weights read as wave-uniform scalars; epilogues: hidden
//                layer (store z and GELU(z) in BF16), output layer (FP32),
//                input-gradient (multiply by GELU'(z) of the layer below)
//   k_flip       W [Cout][Cin][ky][kx] -> W' [Cin][Cout][2-ky][2-kx], so the
//                input gradient is the forward convolution with W'
//   k_wgrad      weight / bias gradient partial sums of one 16x16 tile, both
//                operands staged in LDS; one thread per (Cin, tap) pair
//   k_sum_rows   deterministic reduction of the per-tile partials
//   k_loss_grad  masked MSE and its gradient
//   k_adam       Adam update
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "cnn.hpp"

namespace fwiflow {
namespace {

#define TRY(expr)            \
  do {                       \
    Status s_ = (expr);      \
    if (!s_.ok()) return s_; \
  } while (0)

#define CNN_HIP(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return Status::Err(FWIFLOW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kTile = 16;       // output tile edge
constexpr int kHalo = kTile + 2;
constexpr int kCinChunk = 8;    // input channels staged in LDS per pass
constexpr int kMaxC = 32;

typedef unsigned short bf16_t;

// FP32 -> BF16, round to nearest even (finite inputs; NaN kept as NaN)
__device__ inline bf16_t to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ inline float from_bf16(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

template <typename T>
__device__ inline float load_act(const T* p, size_t i);
template <>
__device__ inline float load_act<float>(const float* p, size_t i) {
  return p[i];
}
template <>
__device__ inline float load_act<bf16_t>(const bf16_t* p, size_t i) {
  return from_bf16(p[i]);
}

enum Epilogue { kHidden = 0, kOutput = 1, kInputGrad = 2 };

// out[co][y][x] = b[co] + sum_{ci,ky,kx} w[co][ci][ky][kx] in[ci][y+ky-1][x+kx-1]
template <typename TIn, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv(const TIn* __restrict__ in, int Cin, int H, int W,
                                              const float* __restrict__ w,
                                              const float* __restrict__ b,
                                              float* __restrict__ out_f32,
                                              bf16_t* __restrict__ z_bf16,
                                              bf16_t* __restrict__ a_bf16,
                                              const bf16_t* __restrict__ zin) {
  __shared__ float tile[kCinChunk][kHalo * kHalo];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const int x = x0 + tx, y = y0 + ty;
  const size_t plane = (size_t)H * W;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = b ? b[co] : 0.f;
  for (int c0 = 0; c0 < Cin; c0 += kCinChunk) {
    const int nc = min(kCinChunk, Cin - c0);
    __syncthreads();
    for (int k = threadIdx.x; k < nc * kHalo * kHalo; k += 256) {
      const int c = k / (kHalo * kHalo), r = k % (kHalo * kHalo);
      const int yy = y0 - 1 + r / kHalo, xx = x0 - 1 + r % kHalo;
      float v = 0.f;
      if (yy >= 0 && yy < H && xx >= 0 && xx < W)
        v = load_act<TIn>(in, (size_t)(c0 + c) * plane + (size_t)yy * W + xx);
      tile[c][r] = v;
    }
    __syncthreads();
    for (int c = 0; c < nc; ++c) {
      float v[9];
#pragma unroll
      for (int ky = 0; ky < 3; ++ky)
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) v[ky * 3 + kx] = tile[c][(ty + ky) * kHalo + tx + kx];
      const float* wc = w + (size_t)(c0 + c) * 9;
#pragma unroll
      for (int co = 0; co < COUT; ++co) {
        const float* wk = wc + (size_t)co * Cin * 9;
        float s = acc[co];
#pragma unroll
        for (int t = 0; t < 9; ++t) s += wk[t] * v[t];
        acc[co] = s;
      }
    }
  }
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
#pragma unroll
  for (int co = 0; co < COUT; ++co) {
    const size_t o = (size_t)co * plane + p;
    if (EPI == kHidden) {
      z_bf16[o] = to_bf16(acc[co]);
      a_bf16[o] = to_bf16(gelu(acc[co]));
    } else if (EPI == kOutput) {
      out_f32[o] = acc[co];
    } else {  // input gradient of a hidden layer's activation -> of its z
      out_f32[o] = acc[co] * gelu_grad(from_bf16(zin[o]));
    }
  }
}

// W [Cout][Cin][3][3] -> W' [Cin][Cout][3][3] with the taps reversed
__global__ void k_flip(const float* __restrict__ w, int Cout, int Cin, float* __restrict__ wf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Cout * Cin * 9) return;
  const int t = k % 9, ci = (k / 9) % Cin, co = k / (9 * Cin);
  wf[((size_t)ci * Cout + co) * 9 + (8 - t)] = w[k];
}

// Partial dW[co][ci][t] and db[co] of one 16x16 tile: thread (ci, t) sums over
// the tile's pixels in[ci][p + tap t] * dz[co][p] for every co.
template <typename TIn, int COUT>
__global__ __launch_bounds__(320) void k_wgrad(const TIn* __restrict__ in, int Cin, int H, int W,
                                               const float* __restrict__ dz,
                                               float* __restrict__ part) {
  extern __shared__ float sm[];
  float* sin = sm;                              // [Cin][18*18]
  float* sdz = sm + (size_t)Cin * kHalo * kHalo;  // [COUT][16*16]
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const size_t plane = (size_t)H * W;
  for (int k = threadIdx.x; k < Cin * kHalo * kHalo; k += blockDim.x) {
    const int c = k / (kHalo * kHalo), r = k % (kHalo * kHalo);
    const int yy = y0 - 1 + r / kHalo, xx = x0 - 1 + r % kHalo;
    float v = 0.f;
    if (yy >= 0 && yy < H && xx >= 0 && xx < W)
      v = load_act<TIn>(in, (size_t)c * plane + (size_t)yy * W + xx);
    sin[k] = v;
  }
  for (int k = threadIdx.x; k < COUT * kTile * kTile; k += blockDim.x) {
    const int c = k / (kTile * kTile), r = k % (kTile * kTile);
    const int yy = y0 + r / kTile, xx = x0 + r % kTile;
    sdz[k] = (yy < H && xx < W) ? dz[(size_t)c * plane + (size_t)yy * W + xx] : 0.f;
  }
  __syncthreads();
  const size_t nw = (size_t)COUT * Cin * 9;
  float* pw = part + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * (nw + COUT);
  const int j = threadIdx.x;
  if (j < Cin * 9) {
    const int ci = j / 9, t = j % 9, ky = t / 3, kx = t % 3;
    float acc[COUT];
#pragma unroll
    for (int co = 0; co < COUT; ++co) acc[co] = 0.f;
    const float* si = sin + (size_t)ci * kHalo * kHalo;
    for (int py = 0; py < kTile; ++py)
      for (int px = 0; px < kTile; ++px) {
        const float a = si[(py + ky) * kHalo + px + kx];
        const int q = py * kTile + px;
#pragma unroll
        for (int co = 0; co < COUT; ++co) acc[co] += a * sdz[co * kTile * kTile + q];
      }
#pragma unroll
    for (int co = 0; co < COUT; ++co) pw[((size_t)co * Cin + ci) * 9 + t] = acc[co];
  } else if (j < Cin * 9 + COUT) {
    const int co = j - Cin * 9;
    float s = 0.f;
    for (int q = 0; q < kTile * kTile; ++q) s += sdz[co * kTile * kTile + q];
    pw[nw + co] = s;
  }
}

// out[k] = sum_r rows[r][k] (fixed order: deterministic)
__global__ void k_sum_rows(const float* __restrict__ rows, int nrows, int n,
                           float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float s = 0.f;
  for (int r = 0; r < nrows; ++r) s += rows[(size_t)r * n + k];
  out[k] = s;
}

// loss = 0.5 / count * sum mask (out - target)^2 over C planes; d out
__global__ void k_loss_grad(const float* __restrict__ out, const float* __restrict__ target,
                            const float* __restrict__ mask, int C, size_t plane, float inv_count,
                            float* __restrict__ dout, double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < C * plane;
       k += (size_t)gridDim.x * blockDim.x) {
    const float m = mask[k % plane];
    const float r = (out[k] - target[k]) * m;
    dout[k] = r * m * inv_count;
    acc += 0.5 * (double)r * (double)r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] * (double)inv_count;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

constexpr int kLayers = 4;
constexpr int kCin[kLayers] = {4, 32, 32, 32};
constexpr int kCout[kLayers] = {32, 32, 32, 3};

}  // namespace

struct CnnState {
  int dev = 0;
  hipStream_t stream = nullptr;
  int H = 0, W = 0;
  std::vector<void*> allocs;
  size_t off[kLayers] = {};  // start of each layer's W in the flat vector
  int nparams = 0;
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  float *wflip = nullptr;     // flipped weights of one layer
  float *x0 = nullptr;        // network input [4][H][W]
  bf16_t *z[3] = {}, *a[3] = {};  // hidden pre-activations and activations
  float *out = nullptr, *target = nullptr, *mask = nullptr;
  float *d0 = nullptr, *d1 = nullptr;  // gradients w.r.t. z of the current / lower layer
  float* part = nullptr;      // per-tile weight-gradient partials
  double* lpart = nullptr;
  int lblocks = 512;
  long step = 0;

  ~CnnState() {
    if (stream) (void)hipSetDevice(dev);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <typename T>
  Status alloc(T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 4);
    if (e != hipSuccess) return Status::Err(FWIFLOW_ERR_MEMORY, "hipMalloc failed in the network");
    allocs.push_back(q);
    *p = (T*)q;
    return Status::Ok();
  }
  dim3 tiles() const { return dim3((W + kTile - 1) / kTile, (H + kTile - 1) / kTile); }
};

namespace {

template <typename TIn>
void launch_conv(CnnState& s, int layer, const TIn* in, int epi, float* out_f32, bf16_t* zb,
                 bf16_t* ab, const bf16_t* zin, const float* w, const float* bias, int Cin,
                 int Cout) {
  const dim3 grid = s.tiles();
  if (Cout == 32) {
    if (epi == kHidden)
      hipLaunchKernelGGL((k_conv<TIn, 32, kHidden>), grid, dim3(256), 0, s.stream, in, Cin, s.H,
                         s.W, w, bias, out_f32, zb, ab, zin);
    else if (epi == kOutput)
      hipLaunchKernelGGL((k_conv<TIn, 32, kOutput>), grid, dim3(256), 0, s.stream, in, Cin, s.H,
                         s.W, w, bias, out_f32, zb, ab, zin);
    else
      hipLaunchKernelGGL((k_conv<TIn, 32, kInputGrad>), grid, dim3(256), 0, s.stream, in, Cin,
                         s.H, s.W, w, bias, out_f32, zb, ab, zin);
  } else if (Cout == 3) {
    hipLaunchKernelGGL((k_conv<TIn, 3, kOutput>), grid, dim3(256), 0, s.stream, in, Cin, s.H, s.W,
                       w, bias, out_f32, zb, ab, zin);
  } else {  // Cout == 4: input gradient of the first layer is never needed
  }
  (void)layer;
}

template <typename TIn>
void launch_wgrad(CnnState& s, const TIn* in, int Cin, int Cout, const float* dz, float* gw) {
  const dim3 grid = s.tiles();
  const size_t lds = ((size_t)Cin * kHalo * kHalo + (size_t)Cout * kTile * kTile) * sizeof(float);
  if (Cout == 32)
    hipLaunchKernelGGL((k_wgrad<TIn, 32>), grid, dim3(320), lds, s.stream, in, Cin, s.H, s.W, dz,
                       s.part);
  else
    hipLaunchKernelGGL((k_wgrad<TIn, 3>), grid, dim3(320), lds, s.stream, in, Cin, s.H, s.W, dz,
                       s.part);
  const int n = Cout * Cin * 9 + Cout;
  hipLaunchKernelGGL(k_sum_rows, dim3((n + 255) / 256), dim3(256), 0, s.stream, s.part,
                     (int)(grid.x * grid.y), n, gw);
}

Status forward(CnnState& s) {
  const float* P = s.params;
  // layer 0: FP32 input
  launch_conv<float>(s, 0, s.x0, kHidden, nullptr, s.z[0], s.a[0], nullptr, P + s.off[0],
                     P + s.off[0] + kCout[0] * kCin[0] * 9, kCin[0], kCout[0]);
  for (int l = 1; l < 3; ++l)
    launch_conv<bf16_t>(s, l, s.a[l - 1], kHidden, nullptr, s.z[l], s.a[l], nullptr,
                        P + s.off[l], P + s.off[l] + kCout[l] * kCin[l] * 9, kCin[l], kCout[l]);
  launch_conv<bf16_t>(s, 3, s.a[2], kOutput, s.out, nullptr, nullptr, nullptr, P + s.off[3],
                      P + s.off[3] + kCout[3] * kCin[3] * 9, kCin[3], kCout[3]);
  CNN_HIP(hipGetLastError());
  return Status::Ok();
}

// loss and parameter gradients for the staged input / target / mask
Status backward(CnnState& s, double* loss) {
  const size_t plane = (size_t)s.H * s.W;
  // mask population for the mean
  std::vector<float> hm(plane);
  CNN_HIP(hipMemcpyAsync(hm.data(), s.mask, plane * sizeof(float), hipMemcpyDeviceToHost,
                         s.stream));
  CNN_HIP(hipStreamSynchronize(s.stream));
  double cnt = 0;
  for (float v : hm) cnt += v != 0.f;
  const float inv = cnt > 0 ? (float)(1.0 / (cnt * kCout[3])) : 0.f;
  hipLaunchKernelGGL(k_loss_grad, dim3(s.lblocks), dim3(256), 0, s.stream, s.out, s.target,
                     s.mask, kCout[3], plane, inv, s.d0, s.lpart);
  float* G = s.grads;
  const float* P = s.params;
  // layer 3 (output): dz = d out
  launch_wgrad<bf16_t>(s, s.a[2], kCin[3], kCout[3], s.d0, G + s.off[3]);
  float* dcur = s.d0;
  float* dnext = s.d1;
  for (int l = 3; l >= 1; --l) {
    // gradient w.r.t. z of layer l-1: conv of dz_l with flipped W_l, times GELU'(z_{l-1})
    const int n = kCout[l] * kCin[l] * 9;
    hipLaunchKernelGGL(k_flip, dim3((n + 255) / 256), dim3(256), 0, s.stream, P + s.off[l],
                       kCout[l], kCin[l], s.wflip);
    const int cin_t = kCout[l], cout_t = kCin[l];  // transposed roles
    const dim3 grid = s.tiles();
    if (cout_t == 32) {
      hipLaunchKernelGGL((k_conv<float, 32, kInputGrad>), grid, dim3(256), 0, s.stream, dcur,
                         cin_t, s.H, s.W, s.wflip, nullptr, dnext, nullptr, nullptr, s.z[l - 1]);
    }
    // weight gradient of layer l-1
    if (l - 1 == 0)
      launch_wgrad<float>(s, s.x0, kCin[0], kCout[0], dnext, G + s.off[0]);
    else
      launch_wgrad<bf16_t>(s, s.a[l - 2], kCin[l - 1], kCout[l - 1], dnext, G + s.off[l - 1]);
    std::swap(dcur, dnext);
  }
  CNN_HIP(hipGetLastError());
  std::vector<double> lp(s.lblocks);
  CNN_HIP(hipMemcpyAsync(lp.data(), s.lpart, lp.size() * sizeof(double), hipMemcpyDeviceToHost,
                         s.stream));
  CNN_HIP(hipStreamSynchronize(s.stream));
  double L = 0;
  for (double v : lp) L += v;
  if (loss) *loss = L;
  return Status::Ok();
}

}  // namespace

Cnn::Cnn() = default;
Cnn::~Cnn() = default;

Status Cnn::create(int gpu, int height, int width, unsigned seed) {
  if (height < 1 || width < 1) return Status::Err(FWIFLOW_ERR_ARG, "network grid must be >= 1x1");
  int ndev = 0;
  CNN_HIP(hipGetDeviceCount(&ndev));
  if (gpu < 0 || gpu >= ndev) return Status::Err(FWIFLOW_ERR_HIP, "gpu_id not present");
  s_.reset(new CnnState());
  CnnState& s = *s_;
  s.dev = gpu;
  s.H = height;
  s.W = width;
  CNN_HIP(hipSetDevice(gpu));
  CNN_HIP(hipStreamCreate(&s.stream));
  // the weight-gradient tiles stage up to 74 KB in LDS (gfx950: 160 KB per CU)
  const int lds_max = (int)(((size_t)kMaxC * kHalo * kHalo + (size_t)kMaxC * kTile * kTile) *
                            sizeof(float));
  CNN_HIP(hipFuncSetAttribute((const void*)k_wgrad<bf16_t, 32>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  CNN_HIP(hipFuncSetAttribute((const void*)k_wgrad<bf16_t, 3>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  CNN_HIP(hipFuncSetAttribute((const void*)k_wgrad<float, 32>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  size_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    s.off[l] = o;
    o += (size_t)kCout[l] * kCin[l] * 9 + kCout[l];
  }
  s.nparams = (int)o;
  // He-normal weights for the GELU layers, smaller for the linear output
  // layer, zero biases; deterministic in `seed`
  std::vector<float> h(o, 0.f);
  std::mt19937 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    const double fan_in = 9.0 * kCin[l];
    const double sd = (l < kLayers - 1 ? std::sqrt(2.0 / fan_in) : 0.1 * std::sqrt(1.0 / fan_in));
    std::normal_distribution<double> nd(0.0, sd);
    for (int k = 0; k < kCout[l] * kCin[l] * 9; ++k) h[s.off[l] + k] = (float)nd(rng);
  }
  const size_t plane = (size_t)height * width;
  const dim3 t = s.tiles();
  TRY(s.alloc(&s.params, o));
  TRY(s.alloc(&s.grads, o));
  TRY(s.alloc(&s.m1, o));
  TRY(s.alloc(&s.m2, o));
  TRY(s.alloc(&s.wflip, (size_t)kMaxC * kMaxC * 9));
  TRY(s.alloc(&s.x0, 4 * plane));
  for (int l = 0; l < 3; ++l) {
    TRY(s.alloc(&s.z[l], (size_t)kMaxC * plane));
    TRY(s.alloc(&s.a[l], (size_t)kMaxC * plane));
  }
  TRY(s.alloc(&s.out, 3 * plane));
  TRY(s.alloc(&s.target, 3 * plane));
  TRY(s.alloc(&s.mask, plane));
  TRY(s.alloc(&s.d0, (size_t)kMaxC * plane));
  TRY(s.alloc(&s.d1, (size_t)kMaxC * plane));
  TRY(s.alloc(&s.part, (size_t)t.x * t.y * (kMaxC * kMaxC * 9 + kMaxC)));
  TRY(s.alloc(&s.lpart, (size_t)s.lblocks));
  CNN_HIP(hipMemcpy(s.params, h.data(), o * sizeof(float), hipMemcpyHostToDevice));
  CNN_HIP(hipMemset(s.m1, 0, o * sizeof(float)));
  CNN_HIP(hipMemset(s.m2, 0, o * sizeof(float)));
  return Status::Ok();
}

int Cnn::num_params() const { return s_ ? s_->nparams : 0; }

Status Cnn::get_params(float* out) const {
  if (!s_ || !out) return Status::Err(FWIFLOW_ERR_ARG, "null network or output");
  CNN_HIP(hipSetDevice(s_->dev));
  CNN_HIP(hipMemcpy(out, s_->params, s_->nparams * sizeof(float), hipMemcpyDeviceToHost));
  return Status::Ok();
}

Status Cnn::set_params(const float* in) {
  if (!s_ || !in) return Status::Err(FWIFLOW_ERR_ARG, "null network or input");
  CNN_HIP(hipSetDevice(s_->dev));
  CNN_HIP(hipMemcpy(s_->params, in, s_->nparams * sizeof(float), hipMemcpyHostToDevice));
  CNN_HIP(hipMemset(s_->m1, 0, s_->nparams * sizeof(float)));
  CNN_HIP(hipMemset(s_->m2, 0, s_->nparams * sizeof(float)));
  s_->step = 0;
  return Status::Ok();
}

Status Cnn::stage(const float* input, const float* target, const float* mask) {
  CnnState& s = *s_;
  const size_t plane = (size_t)s.H * s.W;
  CNN_HIP(hipMemcpyAsync(s.x0, input, 4 * plane * sizeof(float), hipMemcpyHostToDevice,
                         s.stream));
  if (target)
    CNN_HIP(hipMemcpyAsync(s.target, target, 3 * plane * sizeof(float), hipMemcpyHostToDevice,
                           s.stream));
  if (mask)
    CNN_HIP(hipMemcpyAsync(s.mask, mask, plane * sizeof(float), hipMemcpyHostToDevice, s.stream));
  return Status::Ok();
}

Status Cnn::forward(const float* input, float* output) {
  if (!s_ || !input || !output) return Status::Err(FWIFLOW_ERR_ARG, "null network or buffer");
  CnnState& s = *s_;
  CNN_HIP(hipSetDevice(s.dev));
  TRY(stage(input, nullptr, nullptr));
  TRY(fwiflow::forward(s));
  CNN_HIP(hipMemcpyAsync(output, s.out, 3 * (size_t)s.H * s.W * sizeof(float),
                         hipMemcpyDeviceToHost, s.stream));
  CNN_HIP(hipStreamSynchronize(s.stream));
  return Status::Ok();
}

Status Cnn::gradients(const float* input, const float* target, const float* mask, double* loss,
                      float* grad) {
  if (!s_ || !input || !target || !mask)
    return Status::Err(FWIFLOW_ERR_ARG, "null network or buffer");
  CnnState& s = *s_;
  CNN_HIP(hipSetDevice(s.dev));
  TRY(stage(input, target, mask));
  TRY(fwiflow::forward(s));
  TRY(backward(s, loss));
  if (grad)
    CNN_HIP(hipMemcpy(grad, s.grads, s.nparams * sizeof(float), hipMemcpyDeviceToHost));
  return Status::Ok();
}

Status Cnn::train_step(const float* input, const float* target, const float* mask, float lr,
                       double* loss) {
  TRY(gradients(input, target, mask, loss, nullptr));
  CnnState& s = *s_;
  s.step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)s.step), c2 = 1.f - std::pow(b2, (float)s.step);
  hipLaunchKernelGGL(k_adam, dim3((s.nparams + 255) / 256), dim3(256), 0, s.stream, s.params,
                     s.grads, s.m1, s.m2, s.nparams, lr, b1, b2, 1e-8f, c1, c2);
  CNN_HIP(hipGetLastError());
  CNN_HIP(hipStreamSynchronize(s.stream));
  return Status::Ok();
}

}  // namespace fwiflow
