// gfx950 engine of the elastic FWI operator: the time-step kernels, the plan
// (device-resident model, data and wavefields of one parameter file) and the
// job runner that drives one plan, or several plans on different grids,
// through one set of launches per time step.
//
// Device data layout (all FP32, z fastest so a wave64 row is 64 consecutive
// depth cells, one 256-byte coalesced access):
//   model      lam, mu, ave_mu, byc_a, byc_b          [NX][NZ]
//   wavefield  vz, vx, szz, sxx, sxz                  [shot][component][NX][NZ]
//   adjoint    same 5 components                      [shot][component][NX][NZ]
//   C-PML memory variables (8)                        [shot][variable][NX][NZ]
//   boundary ring, 5 layers x 4 sides, 5 components   [shot][step][component][ring slot]
//   seismograms syn/obs/res, time-major               [nSteps][all receivers]
//   gradient accumulators (5, summed over shots)      [NX][NZ]
//   exact adjoint only: transposed-update operands    [launch shot][slot][NX][NZ]
//
// Two adjoints: the reference's (el_velocity_adj / el_stress_adj, the default)
// and the exact one, the transpose of the forward step (a plan option, no
// reference counterpart; k_exact_a / k_exact_b).
//
// Arithmetic.  Every wavefield update evaluates the reference's expression in
// the reference's order and types, so that the fields follow its FP32
// arithmetic rather than a re-association of it:
//   * a stencil difference is DIVIDED by dz or dx, correctly rounded:
//     q = x*y + (x - d*(x*y))*y with y = RN(1/d), one Newton correction on a
//     correctly rounded reciprocal, which equals IEEE x/d for every normal x
//     when d passes the exhaustive check at plan creation (check_spacing);
//     C-PML K divisions take two corrections;
//   * the reference's double literals promote exactly where its source says
//     (2.0*mu, RSXXZZ*..., MEGA), so those terms are evaluated in FP64 and
//     rounded once to FP32 on assignment, as C++ does;
//   * no multiply-add contraction anywhere in this file.
// Gradient accumulators are the one place that sums in another order: the
// contributions of a launch's shots are added in registers, and the chain
// rule to (lambda, mu, rho) is applied once at the end.
//
// Reference kernels restated here: el_stress.cu:50-131, el_velocity.cu:45-117,
// el_velocity_adj.cu:22-108, el_stress_adj.cu:22-104, utilities.cu:125-205,
// 361-424, 506-593, 653-747; driver sequence libCUFD.cu:157-463.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cond.hpp"
#include "engine.hpp"

#pragma clang fp contract(off)

// Device memory of the allocation-free entry point (fwiflow_cufd_noalloc):
// every buffer of such a call is carved out of this array, which each GPU
// the library runs on holds from module load to process end.  Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returned for the current device (scratch_device).
static __device__ __align__(256) unsigned char scratch[FWIFLOW_SCRATCH_BYTES];

namespace fwiflow {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return Status::Err(FWIFLOW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY(expr)            \
  do {                       \
    Status s_ = (expr);      \
    if (!s_.ok()) return s_; \
  } while (0)

// 16-byte raw buffer load (the LLVM AMDGPU intrinsic, bound by name)
typedef float f4 __attribute__((ext_vector_type(4)));
__device__ f4 raw_load4(__amdgpu_buffer_rsrc_t r, int vo, int so, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v4f32");

namespace {

constexpr int kBZ = 64;            // threads along z (one wave64 row)
constexpr int kCondBlocks = 512;   // misfit partial sums per shot of the conditioned residual
constexpr int kPartials = 1024;    // misfit partial sums of one residual launch
// el_stress.cu:42-43: float c1 = 9.0/8.0, c2 = 1.0/24.0
constexpr float kC1 = (float)(9.0 / 8.0);
constexpr float kC2 = (float)(1.0 / 24.0);
// add_source (utilities.cu:528): float scale = pow(1500.0, 2)
constexpr float kSrcScale = 2250000.0f;
// RSXXZZ (utilities.h:21) is the double literal 3.0
constexpr double kRS = 3.0;

// ------------------------------------------------------------ arithmetic
// x / d for d with the correctly rounded reciprocal y: one Newton correction
// (exact for every normal x when check_spacing(d) holds).
__device__ __forceinline__ float div1(float x, float d, float y) {
  const float q = x * y;
  return __builtin_fmaf(__builtin_fmaf(-d, q, x), y, q);
}
// two corrections: correctly rounded for any normal divisor (the first makes
// q faithful, the second rounds it)
__device__ __forceinline__ float div2(float x, float d, float y) {
  float q = x * y;
  q = __builtin_fmaf(__builtin_fmaf(-d, q, x), y, q);
  return __builtin_fmaf(__builtin_fmaf(-d, q, x), y, q);
}
__device__ __forceinline__ double ddiv2(double x, double d, double y) {
  double q = x * y;
  q = __builtin_fma(__builtin_fma(-d, q, x), y, q);
  return __builtin_fma(__builtin_fma(-d, q, x), y, q);
}

// A C-PML stretching coefficient K of one cell and its reciprocals; x / K is
// x itself where K == 1 (outside the C-PML).
struct Kdiv {
  float K, r;
  double r64;
  bool one;
};
__device__ __forceinline__ Kdiv kdiv_of(float K) {
  Kdiv k;
  k.K = K;
  k.one = K == 1.f;
  k.r = k.one ? 1.f : 1.f / K;
  k.r64 = k.one ? 1.0 : 1.0 / (double)K;
  return k;
}
__device__ __forceinline__ float kdiv(float x, const Kdiv& k) { return k.one ? x : div2(x, k.K, k.r); }
__device__ __forceinline__ double kdiv(double x, const Kdiv& k) {
  return k.one ? x : ddiv2(x, (double)k.K, k.r64);
}

// ------------------------------------------------------------ descriptors
struct Geo {
  int nz, nx, nzA, nPml, nSteps;
  int ns;          // shots of this part in the launch
  int s0;          // plan index of its first shot
  long long N;     // cells per shot
  float dt, dz, dx, rdz, rdx;
  int ringW, ringH, ringlen;
  int R;           // receivers of the plan: row length of the time-major seismograms
  int rs0;         // ring: index of the part's first shot in the ring buffer
  int rnt;         // ring: steps held
};

struct Mdl {
  const float *lam, *mu, *amu, *ba, *bb;
};

struct Pml {
  const float *Kz, *az, *bz, *Kzh, *azh, *bzh;  // length nzA
  const float *Kx, *ax, *bx, *Kxh, *axh, *bxh;  // length nx
};

// Wavefield arrays are [shot][component][cell]: the 5 components of a state
// (vz, vx, szz, sxx, sxz) and the 8 C-PML memory variables, named after the
// reference's d_mem_* arrays.
enum WaveComp { kVz = 0, kVx = 1, kSzz = 2, kSxx = 3, kSxz = 4, kWaveComps = 5 };
enum MemComp {
  kMvzz = 0,   // mem_dvz_dz
  kMvzx = 1,   // mem_dvz_dx
  kMvxz = 2,   // mem_dvx_dz
  kMvxx = 3,   // mem_dvx_dx
  kMszz = 4,   // mem_dszz_dz
  kMsxx = 5,   // mem_dsxx_dx
  kMsxzz = 6,  // mem_dsxz_dz
  kMsxzx = 7,  // mem_dsxz_dx
  kMemComps = 8
};

struct Src {
  const int* zx;     // [shot][2] grid indices of the source
  const float* amp;  // [shot][nSteps] windowed source rows (Src_Rec.cu:132-143)
};

// Residual injection folded into the adjoint stress update (one-row surveys
// with one receiver per cell): at[plan shot][x] is the receiver of column x
// in row z, or -1.
struct Inj {
  const float* res;  // time-major residual, row t at res + t * R; null: no injection
  const int* at;
  int z;
};

// One plan's shots in one launch.  Every per-shot pointer already points at
// the part's first shot.
struct Part {
  Geo g;
  Mdl m;
  Pml c;
  float* w;         // forward state [shot][5][cell] from the part's first shot
  float* a;         // adjoint state, same layout
  float* p;         // C-PML memory variables [shot][8][cell] (forward ones in
                    // forward launches, adjoint ones in backward launches)
  Src src;          // source rows: injected (forward) or removed (reversal)
  Inj inj;
  float* ring;      // boundary ring buffer (null: not saved)
  float* acc;       // 5 gradient accumulators
  float* stf_grad;  // [shot][nSteps]
  float* illum;     // forward illumination (null: off)
  const int* rec;   // [receiver][3] = (z, x, plan shot), from the part's first receiver
  float* syn;       // seismograms + the part's first receiver (row t at + t * R)
  int nrec;         // receivers of the part's shots
  int gzU;          // tiles along z of the update-region tiling
  int gzR;          // tiles along z of the ring-rectangle tiling
};

// Kernel argument: one part by value, or a table of parts in device memory.
template <int NP>
struct Parts;
template <>
struct Parts<1> {
  Part p;
  __device__ __forceinline__ const Part& get(int) const { return p; }
};
template <>
struct Parts<0> {
  const Part* p;
  __device__ __forceinline__ const Part& get(int i) const { return p[i]; }
};

// Workgroup map entries: kind (3 bits), part (5 bits), tile or receiver
// block (24 bits).  A launch's grid is 8 * L workgroups; workgroup b runs on
// XCD b % 8, so the host lays out each XCD's list of work at b = 8 j + xcd.
enum Kind : unsigned {
  kNop = 0,
  kStressFwd = 1,
  kVelFwd = 2,
  kRecord = 3,
  kRevStress = 4,
  kAdjVel = 5,
  kAdjStress = 6,
  kRevVel = 7,
};
constexpr int kMaxPartsTable = 31;
__host__ __device__ inline unsigned wg_entry(unsigned kind, unsigned part, unsigned tile) {
  return kind << 29 | part << 24 | tile;
}

// Boundary ring (Boundary.cu:17-41, utilities.cu:361-424): rows and columns
// nPml-2 .. nPml+2 on each side of the interior, i.e. the band of width 5 of
// the rectangle [nPml-2, nzA-nPml+1] x [nPml-2, nx-nPml+1].  Slot numbering is
// this engine's own: top band [W][5], bottom band [W][5], left band [5][H-10],
// right band [5][H-10], z fastest inside each so a wave writes contiguously.
__host__ __device__ inline int ring_slot(int z, int x, const Geo& g) {
  const int zr = z - (g.nPml - 2), xr = x - (g.nPml - 2);
  const int H = g.ringH, W = g.ringW;
  if (zr < 0 || zr >= H || xr < 0 || xr >= W) return -1;
  if (zr < 5) return xr * 5 + zr;
  if (zr >= H - 5) return 5 * W + xr * 5 + (zr - (H - 5));
  const int hz = H - 10, zc = zr - 5;
  if (xr < 5) return 10 * W + xr * hz + zc;
  if (xr >= W - 5) return 10 * W + 5 * hz + (xr - (W - 5)) * hz + zc;
  return -1;
}

// Ring layout [ring shot][held step][component][slot]; `s` counts from the
// part's first shot.  The ring holds rnt consecutive steps: step it is row
// it % rnt (all steps of the sweep when rnt = nSteps - 1, one checkpoint
// segment otherwise).
__device__ inline float* ring_ptr(float* ring, const Geo& g, int it, int comp, int s, int slot) {
  const size_t row = ((size_t)g.rs0 + s) * g.rnt + (unsigned)it % (unsigned)g.rnt;
  return ring + (row * 5 + comp) * (size_t)g.ringlen + slot;
}

// Stencil loads and stores go through buffer resources: one scalar resource
// per field of a part, the shot's byte offset in an SGPR (soffset) and the
// cell's per-lane byte offset in a VGPR, with the z neighbours at immediate
// offsets.  A wave's addresses then cost no per-lane 64-bit arithmetic.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rsrc(const float* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, (int)0xffffffff,
                                           0x00020000);
}
// Per-lane byte offsets of a cell's stencil: z-2 (the z neighbours follow
// 4 B apart) and the x neighbours x-2, x-1, x+1, x+2.
struct Cell {
  int zm2, xm2, xm1, xp1, xp2;
};
__device__ __forceinline__ Cell cell_of(unsigned bo, int nz) {
  const int b = (int)bo, r = 4 * nz;
  return Cell{b - 8, b - 2 * r, b - r, b + r, b + 2 * r};
}
__device__ __forceinline__ float ldb(Rsrc r, int so, int vo) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
// A [shot][component][cell] array of one part: its buffer resource and the
// bytes of one component.  Accessors take the shot's byte offset `so` and a
// component.
struct Arr {
  Rsrc r;
  int nb;
};
__device__ __forceinline__ Arr arr(const float* p, int nb) { return Arr{rsrc(p), nb}; }
// z neighbour k (-2..2) and x neighbour k (-2, -1, 1, 2) of the cell
__device__ __forceinline__ float lz(const Arr& A, int so, int comp, const Cell& c, int k) {
  return ldb(A.r, so + comp * A.nb, c.zm2 + 4 * (k + 2));
}
__device__ __forceinline__ float lx(const Arr& A, int so, int comp, const Cell& c, int k) {
  return ldb(A.r, so + comp * A.nb, k == -2 ? c.xm2 : k == -1 ? c.xm1 : k == 1 ? c.xp1 : c.xp2);
}
// z neighbours k0 .. k0+3 of the cell in one 16-byte load
__device__ __forceinline__ f4 lz4(const Arr& A, int so, int comp, const Cell& c, int k0) {
  return raw_load4(A.r, c.zm2 + 4 * (k0 + 2), so + comp * A.nb, 0);
}
__device__ __forceinline__ void st(const Arr& A, int so, int comp, const Cell& c, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), A.r, c.zm2 + 8,
                                        so + comp * A.nb, 0);
}

// Column of the calling thread in a tile whose first column is x0.  A
// workgroup is 64 x tile_x threads, so each wave64 is one column: the column
// is wave-uniform, and the x-direction C-PML coefficients and x branches it
// selects are scalar.
__device__ __forceinline__ int wave_col(int x0) {
  return __builtin_amdgcn_readfirstlane(x0 + (int)threadIdx.y);
}

// Staggered 4th-order differences as the reference writes them:
// forward / reverse (c1 (a - b) - c2 (c - d)) / h, adjoint (-c1 (a - b) + c2 (c - d)) / h.
__device__ __forceinline__ float dfw(float a, float b, float c, float d, float h, float rh) {
  return div1(kC1 * (a - b) - kC2 * (c - d), h, rh);
}
__device__ __forceinline__ float dad(float a, float b, float c, float d, float h, float rh) {
  return div1(-kC1 * (a - b) + kC2 * (c - d), h, rh);
}

// ------------------------------------------------------------------ forward
// The kernels walk all shots of a part inside each thread, so model
// parameters and C-PML coefficients are read once per cell.  Shots are taken
// U at a time: every load of the U shots is issued before the first store,
// so a wave keeps U shots' worth of reads in flight (the stores of shot s may
// alias the loads of shot s+1 as far as the compiler knows, so it cannot
// reorder them itself).

// el_stress.cu:50-88 + add_source (utilities.cu:521-555) + from_bnd
// (utilities.cu:361-391): the ring cells save sigma^it and v^it before the
// update of step it.
struct StressIn {
  float vz0, vzm1, vzp1, vzm2, vzpx, vzp2x, vzmx;
  float vx0, vxmx, vxpx, vxm2x, vxp1, vxp2, vxm1;
  float szz, sxx, sxz;
  float mvzz, mvxz, mvxx, mvzx;
  int sz, sx;
};

template <int U>
__device__ __forceinline__ void stress_fwd(const Part& P, int tile, int it) {
  const Geo& g = P.g;
  const int z = (tile % P.gzU) * kBZ + (int)threadIdx.x;
  const int x = wave_col((tile / P.gzU) * (int)blockDim.y);
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const Cell C = cell_of(bo, nz);
  const int Nb = (int)(g.N * 4);  // bytes per shot
  const Arr W = arr(P.w, Nb);
  const Arr M = arr(P.p, Nb);
  const float lam = P.m.lam[i], mu = P.m.mu[i], amu = P.m.amu[i];
  const double l2m = (double)lam + 2.0 * (double)mu;  // (d_Lambda + 2.0*d_Mu)
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml - 1;
  float bz = 0.f, az = 0.f, bzh = 0.f, azh = 0.f, bx = 0.f, ax = 0.f, bxh = 0.f, axh = 0.f;
  Kdiv kz{}, kzh{}, kx{}, kxh{};
  if (pz) {
    bz = P.c.bz[z], az = P.c.az[z], kz = kdiv_of(P.c.Kz[z]);
    bzh = P.c.bzh[z], azh = P.c.azh[z], kzh = kdiv_of(P.c.Kzh[z]);
  }
  if (px) {
    bx = P.c.bx[x], ax = P.c.ax[x], kx = kdiv_of(P.c.Kx[x]);
    bxh = P.c.bxh[x], axh = P.c.axh[x], kxh = kdiv_of(P.c.Kxh[x]);
  }
  const int slot = P.ring ? ring_slot(z, x, g) : -1;
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const double dt64 = (double)dt;
  const int ns = g.ns;
  for (int s0 = 0; s0 < ns; s0 += U) {
    StressIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      StressIn& q = r[u];
      const f4 vzc = lz4(W, sw, kVz, C, -2), vxc = lz4(W, sw, kVx, C, -1);
      q.vzm2 = vzc[0], q.vzm1 = vzc[1], q.vz0 = vzc[2], q.vzp1 = vzc[3];
      q.vzpx = lx(W, sw, kVz, C, 1), q.vzp2x = lx(W, sw, kVz, C, 2), q.vzmx = lx(W, sw, kVz, C, -1);
      q.vxm1 = vxc[0], q.vx0 = vxc[1], q.vxp1 = vxc[2], q.vxp2 = vxc[3];
      q.vxmx = lx(W, sw, kVx, C, -1), q.vxpx = lx(W, sw, kVx, C, 1), q.vxm2x = lx(W, sw, kVx, C, -2);
      q.szz = lz(W, sw, kSzz, C, 0), q.sxx = lz(W, sw, kSxx, C, 0), q.sxz = lz(W, sw, kSxz, C, 0);
      if (pz) q.mvzz = lz(M, sm, kMvzz, C, 0), q.mvxz = lz(M, sm, kMvxz, C, 0);
      if (px) q.mvxx = lz(M, sm, kMvxx, C, 0), q.mvzx = lz(M, sm, kMvzx, C, 0);
      q.sz = P.src.zx[2 * s], q.sx = P.src.zx[2 * s + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      const StressIn& q = r[u];
      float dvz_dz = dfw(q.vz0, q.vzm1, q.vzp1, q.vzm2, dz, rdz);
      float dvx_dx = dfw(q.vx0, q.vxmx, q.vxpx, q.vxm2x, dx, rdx);
      if (pz) {
        const float mm = bz * q.mvzz + az * dvz_dz;
        st(M, sm, kMvzz, C, mm);
        dvz_dz = kdiv(dvz_dz, kz) + mm;
      }
      if (px) {
        const float mm = bx * q.mvxx + ax * dvx_dx;
        st(M, sm, kMvxx, C, mm);
        dvx_dx = kdiv(dvx_dx, kx) + mm;
      }
      float szz = q.szz, sxx = q.sxx, sxz = q.sxz;
      if (slot >= 0) {
        *ring_ptr(P.ring, g, it, 0, s, slot) = szz;
        *ring_ptr(P.ring, g, it, 1, s, slot) = sxz;
        *ring_ptr(P.ring, g, it, 2, s, slot) = sxx;
        *ring_ptr(P.ring, g, it, 3, s, slot) = q.vz0;
        *ring_ptr(P.ring, g, it, 4, s, slot) = q.vx0;
      }
      szz = (float)((double)szz + (l2m * (double)dvz_dz + (double)(lam * dvx_dx)) * dt64);
      sxx = (float)((double)sxx + ((double)(lam * dvz_dz) + l2m * (double)dvx_dx) * dt64);
      float dvx_dz = dfw(q.vxp1, q.vx0, q.vxp2, q.vxm1, dz, rdz);
      float dvz_dx = dfw(q.vzpx, q.vz0, q.vzp2x, q.vzmx, dx, rdx);
      if (pz) {
        const float mm = bzh * q.mvxz + azh * dvx_dz;
        st(M, sm, kMvxz, C, mm);
        dvx_dz = kdiv(dvx_dz, kzh) + mm;
      }
      if (px) {
        const float mm = bxh * q.mvzx + axh * dvz_dx;
        st(M, sm, kMvzx, C, mm);
        dvz_dx = kdiv(dvz_dx, kxh) + mm;
      }
      sxz = sxz + amu * (dvx_dz + dvz_dx) * dt;
      if (z == q.sz && x == q.sx) {
        // add_source: szz += scale*amp*dt*gauss (float), sxx += RSXXZZ*scale*amp*dt*gauss
        // (double); the 9x9 Gaussian is 1 at the centre and 0 elsewhere in FP32
        const float amp = P.src.amp[(size_t)s * g.nSteps + it];
        szz = szz + kSrcScale * amp * dt;
        sxx = (float)((double)sxx + kRS * (double)kSrcScale * (double)amp * dt64);
      }
      st(W, sw, kSzz, C, szz);
      st(W, sw, kSxx, C, sxx);
      st(W, sw, kSxz, C, sxz);
    }
  }
}

// el_velocity.cu:45-82.  x-PML test is `x > nx - nPml` (el_velocity.cu:56,71).
// With an illumination buffer, the part's sum over shots of vz^2 + vx^2 is
// added to it (forward illumination; no reference counterpart).
struct VelIn {
  float szz0, szzp1, szzp2, szzm1;
  float sxz0, sxzmx, sxzpx, sxzm2x, sxzm1, sxzp1, sxzm2;
  float sxxpx, sxx0, sxxp2x, sxxmx;
  float mszz, msxzz, msxzx, msxx;
  float vz, vx;
};

template <int U>
__device__ __forceinline__ void vel_fwd(const Part& P, int tile) {
  const Geo& g = P.g;
  const int z = (tile % P.gzU) * kBZ + (int)threadIdx.x;
  const int x = wave_col((tile / P.gzU) * (int)blockDim.y);
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const Cell C = cell_of(bo, nz);
  const int Nb = (int)(g.N * 4);  // bytes per shot
  const Arr W = arr(P.w, Nb);
  const Arr M = arr(P.p, Nb);
  const float ba = P.m.ba[i], bb = P.m.bb[i];
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml;
  float bz = 0.f, az = 0.f, bzh = 0.f, azh = 0.f, bx = 0.f, ax = 0.f, bxh = 0.f, axh = 0.f;
  Kdiv kz{}, kzh{}, kx{}, kxh{};
  if (pz) {
    bz = P.c.bz[z], az = P.c.az[z], kz = kdiv_of(P.c.Kz[z]);
    bzh = P.c.bzh[z], azh = P.c.azh[z], kzh = kdiv_of(P.c.Kzh[z]);
  }
  if (px) {
    bx = P.c.bx[x], ax = P.c.ax[x], kx = kdiv_of(P.c.Kx[x]);
    bxh = P.c.bxh[x], axh = P.c.axh[x], kxh = kdiv_of(P.c.Kxh[x]);
  }
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const int ns = g.ns;
  float il = 0.f;  // illumination of this part's shots
  for (int s0 = 0; s0 < ns; s0 += U) {
    VelIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      VelIn& q = r[u];
      const f4 szc = lz4(W, sw, kSzz, C, -1), sxc = lz4(W, sw, kSxz, C, -2);
      q.szzm1 = szc[0], q.szz0 = szc[1], q.szzp1 = szc[2], q.szzp2 = szc[3];
      q.sxzm2 = sxc[0], q.sxzm1 = sxc[1], q.sxz0 = sxc[2], q.sxzp1 = sxc[3];
      q.sxzmx = lx(W, sw, kSxz, C, -1), q.sxzpx = lx(W, sw, kSxz, C, 1), q.sxzm2x = lx(W, sw, kSxz, C, -2);
      q.sxxpx = lx(W, sw, kSxx, C, 1), q.sxx0 = lz(W, sw, kSxx, C, 0), q.sxxp2x = lx(W, sw, kSxx, C, 2), q.sxxmx = lx(W, sw, kSxx, C, -1);
      if (pz) q.mszz = lz(M, sm, kMszz, C, 0), q.msxzz = lz(M, sm, kMsxzz, C, 0);
      if (px) q.msxzx = lz(M, sm, kMsxzx, C, 0), q.msxx = lz(M, sm, kMsxx, C, 0);
      q.vz = lz(W, sw, kVz, C, 0), q.vx = lz(W, sw, kVx, C, 0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      const VelIn& q = r[u];
      float dszz_dz = dfw(q.szzp1, q.szz0, q.szzp2, q.szzm1, dz, rdz);
      float dsxz_dx = dfw(q.sxz0, q.sxzmx, q.sxzpx, q.sxzm2x, dx, rdx);
      if (pz) {
        const float mm = bzh * q.mszz + azh * dszz_dz;
        st(M, sm, kMszz, C, mm);
        dszz_dz = kdiv(dszz_dz, kzh) + mm;
      }
      if (px) {
        const float mm = bx * q.msxzx + ax * dsxz_dx;
        st(M, sm, kMsxzx, C, mm);
        dsxz_dx = kdiv(dsxz_dx, kx) + mm;
      }
      const float vz = q.vz + (dszz_dz + dsxz_dx) * ba * dt;
      float dsxz_dz = dfw(q.sxz0, q.sxzm1, q.sxzp1, q.sxzm2, dz, rdz);
      float dsxx_dx = dfw(q.sxxpx, q.sxx0, q.sxxp2x, q.sxxmx, dx, rdx);
      if (pz) {
        const float mm = bz * q.msxzz + az * dsxz_dz;
        st(M, sm, kMsxzz, C, mm);
        dsxz_dz = kdiv(dsxz_dz, kz) + mm;
      }
      if (px) {
        const float mm = bxh * q.msxx + axh * dsxx_dx;
        st(M, sm, kMsxx, C, mm);
        dsxx_dx = kdiv(dsxx_dx, kxh) + mm;
      }
      const float vx = q.vx + (dsxz_dz + dsxx_dx) * bb * dt;
      st(W, sw, kVz, C, vz);
      st(W, sw, kVx, C, vx);
      il += vz * vz + vx * vx;
    }
  }
  if (P.illum) P.illum[i] += il;
}

// recording (utilities.cu:557-567): szz + RSXXZZ sxx (a double sum rounded to
// FP32) at each receiver of the part, written to row it+1 of the seismograms;
// one receiver per thread of the workgroup.
__device__ __forceinline__ void record(const Part& P, int blk, int it) {
  const int per = (int)(blockDim.x * blockDim.y);
  const int r = blk * per + (int)(threadIdx.y * blockDim.x + threadIdx.x);
  if (r >= P.nrec) return;
  const Geo& g = P.g;
  const int z = P.rec[3 * r], x = P.rec[3 * r + 1], s = P.rec[3 * r + 2] - g.s0;
  const float* w = P.w + (size_t)s * kWaveComps * g.N + (long long)x * g.nz + z;
  P.syn[(size_t)(it + 1) * g.R + r] = (float)((double)w[kSzz * g.N] + kRS * (double)w[kSxx * g.N]);
}

// ------------------------------------------------------------------ backward
// Reverse-time re-materialisation (el_velocity.cu:84-117 with isFor = false)
// fused with the ring restore of v (to_bnd, utilities.cu:394-424) and the
// density-gradient accumulation.  The thread domain is the ring rectangle;
// interior cells update, ring cells take the stored values.
// acc_a += vz_adj (dszz/dz + dsxz/dx), acc_b += vx_adj (dsxz/dz + dsxx/dx).
struct RevVelIn {
  float szzp1, szz0, szzp2, szzm1;
  float sxz0, sxzmx, sxzpx, sxzm2x, sxzm1, sxzp1, sxzm2;
  float sxxpx, sxx0, sxxp2x, sxxmx;
  float avz, avx, vz, vx;
};

template <int U>
__device__ __forceinline__ void rev_vel(const Part& P, int tile, int it) {
  const Geo& g = P.g;
  const int z = (tile % P.gzR) * kBZ + (int)threadIdx.x + (g.nPml - 2);
  const int x = wave_col((tile / P.gzR) * (int)blockDim.y) + (g.nPml - 2);
  if (z > g.nzA - g.nPml + 1 || x > g.nx - g.nPml + 1) return;
  const bool in = z >= g.nPml && z <= g.nzA - 1 - g.nPml && x >= g.nPml && x <= g.nx - 1 - g.nPml;
  const int slot = ring_slot(z, x, g);
  if (!in && slot < 0) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const Cell C = cell_of(bo, nz);
  const int Nb = (int)(g.N * 4);  // bytes per shot
  const Arr W = arr(P.w, Nb);
  const Arr A = arr(P.a, Nb);
  const float ba = P.m.ba[i], bb = P.m.bb[i];
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const int ns = g.ns;
  float ga = 0.f, gb = 0.f;
  for (int s0 = 0; s0 < ns; s0 += U) {
    RevVelIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb;
      RevVelIn& q = r[u];
      if (in) {
        const f4 szc = lz4(W, sw, kSzz, C, -1), sxc = lz4(W, sw, kSxz, C, -2);
        q.szzm1 = szc[0], q.szz0 = szc[1], q.szzp1 = szc[2], q.szzp2 = szc[3];
        q.sxzm2 = sxc[0], q.sxzm1 = sxc[1], q.sxz0 = sxc[2], q.sxzp1 = sxc[3];
        q.sxzmx = lx(W, sw, kSxz, C, -1), q.sxzpx = lx(W, sw, kSxz, C, 1), q.sxzm2x = lx(W, sw, kSxz, C, -2);
        q.sxxpx = lx(W, sw, kSxx, C, 1), q.sxx0 = lz(W, sw, kSxx, C, 0), q.sxxp2x = lx(W, sw, kSxx, C, 2), q.sxxmx = lx(W, sw, kSxx, C, -1);
        q.avz = lz(A, sw, kVz, C, 0), q.avx = lz(A, sw, kVx, C, 0);
      }
      if (slot < 0) {
        q.vz = lz(W, sw, kVz, C, 0), q.vx = lz(W, sw, kVx, C, 0);
      } else {
        q.vz = *ring_ptr(P.ring, g, it, 3, s, slot);
        q.vx = *ring_ptr(P.ring, g, it, 4, s, slot);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb;
      const RevVelIn& q = r[u];
      float vz = q.vz, vx = q.vx;
      if (in) {
        const float dszz_dz = dfw(q.szzp1, q.szz0, q.szzp2, q.szzm1, dz, rdz);
        const float dsxz_dx = dfw(q.sxz0, q.sxzmx, q.sxzpx, q.sxzm2x, dx, rdx);
        const float dsxz_dz = dfw(q.sxz0, q.sxzm1, q.sxzp1, q.sxzm2, dz, rdz);
        const float dsxx_dx = dfw(q.sxxpx, q.sxx0, q.sxxp2x, q.sxxmx, dx, rdx);
        const float fz = dszz_dz + dsxz_dx, fx = dsxz_dz + dsxx_dx;
        ga += q.avz * fz;
        gb += q.avx * fx;
        if (slot < 0) {
          vz = vz - fz * ba * dt;
          vx = vx - fx * bb * dt;
        }
      }
      st(W, sw, kVz, C, vz);
      st(W, sw, kVx, C, vx);
    }
  }
  if (in) {
    P.acc[3 * g.N + i] += ga;
    P.acc[4 * g.N + i] += gb;
  }
}

// el_stress.cu:90-131 (isFor = false) preceded by add_source(-) and followed
// by the ring restore of sigma.  acc_l += (szz_adj + sxx_adj)(dvz/dz + dvx/dx),
// acc_m += szz_adj dvz/dz + sxx_adj dvx/dx, acc_s += sxz_adj (dvx/dz + dvz/dx).
struct RevStressIn {
  float vz0, vzm1, vzp1, vzm2, vzpx, vzp2x, vzmx;
  float vx0, vxmx, vxpx, vxm2x, vxp1, vxp2, vxm1;
  float aszz, asxx, asxz;
  float szz, sxx, sxz;
  int sz, sx;
};

template <int U>
__device__ __forceinline__ void rev_stress(const Part& P, int tile, int it) {
  const Geo& g = P.g;
  const int z = (tile % P.gzR) * kBZ + (int)threadIdx.x + (g.nPml - 2);
  const int x = wave_col((tile / P.gzR) * (int)blockDim.y) + (g.nPml - 2);
  if (z > g.nzA - g.nPml + 1 || x > g.nx - g.nPml + 1) return;
  const bool in = z >= g.nPml && z <= g.nzA - 1 - g.nPml && x >= g.nPml && x <= g.nx - 1 - g.nPml;
  const int slot = ring_slot(z, x, g);
  if (!in && slot < 0) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const Cell C = cell_of(bo, nz);
  const int Nb = (int)(g.N * 4);  // bytes per shot
  const Arr W = arr(P.w, Nb);
  const Arr A = arr(P.a, Nb);
  const float lam = P.m.lam[i], mu = P.m.mu[i], amu = P.m.amu[i];
  const double l2m = (double)lam + 2.0 * (double)mu;
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const double dt64 = (double)dt;
  const int ns = g.ns;
  float gl = 0.f, gm = 0.f, gs = 0.f;
  for (int s0 = 0; s0 < ns; s0 += U) {
    RevStressIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb;
      RevStressIn& q = r[u];
      if (in) {
        const f4 vzc = lz4(W, sw, kVz, C, -2), vxc = lz4(W, sw, kVx, C, -1);
        q.vzm2 = vzc[0], q.vzm1 = vzc[1], q.vz0 = vzc[2], q.vzp1 = vzc[3];
        q.vzpx = lx(W, sw, kVz, C, 1), q.vzp2x = lx(W, sw, kVz, C, 2), q.vzmx = lx(W, sw, kVz, C, -1);
        q.vxm1 = vxc[0], q.vx0 = vxc[1], q.vxp1 = vxc[2], q.vxp2 = vxc[3];
        q.vxmx = lx(W, sw, kVx, C, -1), q.vxpx = lx(W, sw, kVx, C, 1), q.vxm2x = lx(W, sw, kVx, C, -2);
        q.aszz = lz(A, sw, kSzz, C, 0), q.asxx = lz(A, sw, kSxx, C, 0), q.asxz = lz(A, sw, kSxz, C, 0);
      }
      if (slot < 0) {
        q.szz = lz(W, sw, kSzz, C, 0), q.sxx = lz(W, sw, kSxx, C, 0), q.sxz = lz(W, sw, kSxz, C, 0);
        q.sz = P.src.zx[2 * s], q.sx = P.src.zx[2 * s + 1];
      } else {
        q.szz = *ring_ptr(P.ring, g, it, 0, s, slot);
        q.sxz = *ring_ptr(P.ring, g, it, 1, s, slot);
        q.sxx = *ring_ptr(P.ring, g, it, 2, s, slot);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb;
      const RevStressIn& q = r[u];
      float szz = q.szz, sxx = q.sxx, sxz = q.sxz;
      if (in) {
        const float dvz_dz = dfw(q.vz0, q.vzm1, q.vzp1, q.vzm2, dz, rdz);
        const float dvx_dx = dfw(q.vx0, q.vxmx, q.vxpx, q.vxm2x, dx, rdx);
        const float dvx_dz = dfw(q.vxp1, q.vx0, q.vxp2, q.vxm1, dz, rdz);
        const float dvz_dx = dfw(q.vzpx, q.vz0, q.vzp2x, q.vzmx, dx, rdx);
        gl += (q.aszz + q.asxx) * (dvz_dz + dvx_dx);
        gm += q.aszz * dvz_dz + q.asxx * dvx_dx;
        gs += q.asxz * (dvx_dz + dvz_dx);
        if (slot < 0) {
          if (z == q.sz && x == q.sx) {
            const float amp = P.src.amp[(size_t)s * g.nSteps + it];
            szz = szz - kSrcScale * amp * dt;
            sxx = (float)((double)sxx - kRS * (double)kSrcScale * (double)amp * dt64);
          }
          szz = (float)((double)szz - (l2m * (double)dvz_dz + (double)(lam * dvx_dx)) * dt64);
          sxx = (float)((double)sxx - ((double)(lam * dvz_dz) + l2m * (double)dvx_dx) * dt64);
          sxz = sxz - amu * (dvx_dz + dvz_dx) * dt;
        }
      }
      st(W, sw, kSzz, C, szz);
      st(W, sw, kSxx, C, sxx);
      st(W, sw, kSxz, C, sxz);
    }
  }
  if (in) {
    P.acc[i] += gl;
    P.acc[g.N + i] += gm;
    P.acc[2 * g.N + i] += gs;
  }
}

// el_velocity_adj.cu:22-108, plus source_grad (utilities.cu:582-593) for step
// `it` < nSteps-1 (the adjoint stress it reads is not touched by this update).
// The C-PML memory terms a*dpsi are evaluated only where a != 0; elsewhere
// they are exactly zero in the reference as well, and adding a zero leaves
// the sum unchanged.
struct AdjVelIn {
  float szz0, szzpx, szzp2x, szzmx, szzp1, szzp2, szzm1;
  float sxx0, sxxpx, sxxp2x, sxxmx, sxxp1, sxxp2, sxxm1;
  float sxz0, sxzm1, sxzp1, sxzm2, sxzmx, sxzpx, sxzm2x;
  float q1[4], q2[4], q3[4], q4[4];  // vxx (x), vxz (z), vzz (z), vzx (x) stencils
  float m_sxx, m_sxzz, m_sxzx, m_szz;
  float vx, vz;
  int sz, sx;
};

template <int U>
__device__ __forceinline__ void adj_vel(const Part& P, int tile, int it) {
  const Geo& g = P.g;
  const int z = (tile % P.gzU) * kBZ + (int)threadIdx.x;
  const int x = wave_col((tile / P.gzU) * (int)blockDim.y);
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const Cell C = cell_of(bo, nz);
  const int Nb = (int)(g.N * 4);  // bytes per shot
  const Arr A = arr(P.a, Nb);
  const Arr M = arr(P.p, Nb);
  const float lam = P.m.lam[i], mu = P.m.mu[i], amu = P.m.amu[i], ba = P.m.ba[i], bb = P.m.bb[i];
  const double l2m = (double)lam + 2.0 * (double)mu;
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml - 1;
  const float ax = P.c.ax[x], bx = P.c.bx[x], axh = P.c.axh[x], bxh = P.c.bxh[x];
  const float az = P.c.az[z], bz = P.c.bz[z], azh = P.c.azh[z], bzh = P.c.bzh[z];
  const Kdiv kx = kdiv_of(P.c.Kx[x]), kz = kdiv_of(P.c.Kz[z]);
  const float amu_kzh = kdiv(amu, kdiv_of(P.c.Kzh[z]));  // d_ave_Mu / d_K_z_half
  const float amu_kxh = kdiv(amu, kdiv_of(P.c.Kxh[x]));  // d_ave_Mu / d_K_x_half
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const double dt64 = (double)dt;
  const int ns = g.ns;
  const bool stf = it < g.nSteps - 1 && P.stf_grad;
  for (int s0 = 0; s0 < ns; s0 += U) {
    AdjVelIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      AdjVelIn& q = r[u];
      const f4 szc = lz4(A, sw, kSzz, C, -1), sxc = lz4(A, sw, kSxx, C, -1), sxzc = lz4(A, sw, kSxz, C, -2);
      q.szzm1 = szc[0], q.szz0 = szc[1], q.szzp1 = szc[2], q.szzp2 = szc[3];
      q.szzpx = lx(A, sw, kSzz, C, 1), q.szzp2x = lx(A, sw, kSzz, C, 2), q.szzmx = lx(A, sw, kSzz, C, -1);
      q.sxxm1 = sxc[0], q.sxx0 = sxc[1], q.sxxp1 = sxc[2], q.sxxp2 = sxc[3];
      q.sxxpx = lx(A, sw, kSxx, C, 1), q.sxxp2x = lx(A, sw, kSxx, C, 2), q.sxxmx = lx(A, sw, kSxx, C, -1);
      q.sxzm2 = sxzc[0], q.sxzm1 = sxzc[1], q.sxz0 = sxzc[2], q.sxzp1 = sxzc[3];
      q.sxzmx = lx(A, sw, kSxz, C, -1), q.sxzpx = lx(A, sw, kSxz, C, 1), q.sxzm2x = lx(A, sw, kSxz, C, -2);
      if (ax != 0.f) {  // psi_vxx at x+1, x, x+2, x-1
        q.q1[0] = lx(M, sm, kMvxx, C, 1), q.q1[1] = lz(M, sm, kMvxx, C, 0), q.q1[2] = lx(M, sm, kMvxx, C, 2), q.q1[3] = lx(M, sm, kMvxx, C, -1);
      }
      if (azh != 0.f) {  // psi_vxz at z, z-1, z+1, z-2
        const f4 t = lz4(M, sm, kMvxz, C, -2);
        q.q2[0] = t[2], q.q2[1] = t[1], q.q2[2] = t[3], q.q2[3] = t[0];
      }
      if (az != 0.f) {  // psi_vzz at z+1, z, z+2, z-1
        const f4 t = lz4(M, sm, kMvzz, C, -1);
        q.q3[0] = t[2], q.q3[1] = t[1], q.q3[2] = t[3], q.q3[3] = t[0];
      }
      if (axh != 0.f) {  // psi_vzx at x, x-1, x+1, x-2
        q.q4[0] = lz(M, sm, kMvzx, C, 0), q.q4[1] = lx(M, sm, kMvzx, C, -1), q.q4[2] = lx(M, sm, kMvzx, C, 1), q.q4[3] = lx(M, sm, kMvzx, C, -2);
      }
      if (px) q.m_sxx = lz(M, sm, kMsxx, C, 0), q.m_sxzx = lz(M, sm, kMsxzx, C, 0);
      if (pz) q.m_sxzz = lz(M, sm, kMsxzz, C, 0), q.m_szz = lz(M, sm, kMszz, C, 0);
      q.vx = lz(A, sw, kVx, C, 0), q.vz = lz(A, sw, kVz, C, 0);
      q.sz = P.src.zx[2 * s], q.sx = P.src.zx[2 * s + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      const AdjVelIn& q = r[u];
      if (stf && z == q.sz && x == q.sx)  // -(szz_adj + RSXXZZ sxx_adj) dt, double, stored as float
        P.stf_grad[(size_t)s * g.nSteps + it] = (float)(-((double)q.szz0 + kRS * (double)q.sxx0) * dt64);
      // update vx: a_x dpsi_xx/dx + lambda dszz/dx / K_x dt + (lambda + 2.0 mu) dsxx/dx / K_x dt
      //            + a_z_half dpsi_xz/dz + ave_mu / K_z_half dsxz/dz dt
      const float dszz_dx = dad(q.szzpx, q.szz0, q.szzp2x, q.szzmx, dx, rdx);
      const float dsxx_dx = dad(q.sxxpx, q.sxx0, q.sxxp2x, q.sxxmx, dx, rdx);
      const float dsxz_dz = dad(q.sxz0, q.sxzm1, q.sxzp1, q.sxzm2, dz, rdz);
      float t = kdiv(lam * dszz_dx, kx) * dt;
      if (ax != 0.f) t = ax * dad(q.q1[0], q.q1[1], q.q1[2], q.q1[3], dx, rdx) + t;
      double S = (double)t + kdiv(l2m * (double)dsxx_dx, kx) * dt64;
      if (azh != 0.f) S = S + (double)(azh * dad(q.q2[0], q.q2[1], q.q2[2], q.q2[3], dz, rdz));
      S = S + (double)(amu_kzh * dsxz_dz * dt);
      const float vx = (float)((double)q.vx + S);
      st(A, sw, kVx, C, vx);
      if (px) st(M, sm, kMsxx, C, bxh * q.m_sxx + bb * vx * dt);
      if (pz) st(M, sm, kMsxzz, C, bz * q.m_sxzz + bb * vx * dt);
      // update vz: a_z dpsi_zz/dz + (lambda + 2.0 mu) dszz/dz / K_z dt + lambda dsxx/dz / K_z dt
      //            + a_x_half dpsi_zx/dx + ave_mu / K_x_half dsxz/dx dt
      const float dszz_dz = dad(q.szzp1, q.szz0, q.szzp2, q.szzm1, dz, rdz);
      const float dsxx_dz = dad(q.sxxp1, q.sxx0, q.sxxp2, q.sxxm1, dz, rdz);
      const float dsxz_dx = dad(q.sxz0, q.sxzmx, q.sxzpx, q.sxzm2x, dx, rdx);
      S = kdiv(l2m * (double)dszz_dz, kz) * dt64;
      if (az != 0.f) S = (double)(az * dad(q.q3[0], q.q3[1], q.q3[2], q.q3[3], dz, rdz)) + S;
      S = S + (double)(kdiv(lam * dsxx_dz, kz) * dt);
      if (axh != 0.f) S = S + (double)(axh * dad(q.q4[0], q.q4[1], q.q4[2], q.q4[3], dx, rdx));
      S = S + (double)(amu_kxh * dsxz_dx * dt);
      const float vz = (float)((double)q.vz + S);
      st(A, sw, kVz, C, vz);
      if (px) st(M, sm, kMsxzx, C, bx * q.m_sxzx + ba * vz * dt);
      if (pz) st(M, sm, kMszz, C, bzh * q.m_szz + ba * vz * dt);
    }
  }
}

// el_stress_adj.cu:22-104.  The 4 stress-side memory variables are updated
// without PML gates in the reference; they are read only through a*dpsi
// terms, i.e. where a != 0 (the C-PML layers) plus the 2-cell reach of the
// stencil, so they are kept in that band (`zlive`/`xlive`) and nowhere else.
// Residual injection folded into the update (res_injection,
// utilities.cu:569-580): at a receiver's cell szz_adj += res and sxx_adj +=
// RSXXZZ res before the update reads them (the update reads only its own
// cell's stresses, so the sums are those of a separate injection pass).
struct AdjStressIn {
  float vz0, vzpx, vzp2x, vzmx, vzm1, vzp1, vzm2;
  float vx0, vxp1, vxp2, vxm1, vxmx, vxpx, vxm2x;
  float q1[4], q2[4], q3[4], q4[4];  // sxzx (x), sxzz (z), sxx (x), szz (z) stencils
  float m_vzx, m_vxz, m_vxx, m_vzz;
  float sxz, sxx, szz;
};

template <int U>
__device__ __forceinline__ void adj_stress(const Part& P, int tile, int it) {
  const Geo& g = P.g;
  const int z = (tile % P.gzU) * kBZ + (int)threadIdx.x;
  const int x = wave_col((tile / P.gzU) * (int)blockDim.y);
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const Cell C = cell_of(bo, nz);
  const int Nb = (int)(g.N * 4);  // bytes per shot
  const Arr A = arr(P.a, Nb);
  const Arr M = arr(P.p, Nb);
  const float lam = P.m.lam[i], mu = P.m.mu[i], amu = P.m.amu[i], ba = P.m.ba[i], bb = P.m.bb[i];
  const double l2m = (double)lam + 2.0 * (double)mu;
  const bool zlive = z <= g.nPml + 1 || z >= g.nzA - g.nPml - 2;
  const bool xlive = x <= g.nPml + 1 || x >= g.nx - g.nPml - 2;
  const float ax = P.c.ax[x], bx = P.c.bx[x], axh = P.c.axh[x], bxh = P.c.bxh[x];
  const float az = P.c.az[z], bz = P.c.bz[z], azh = P.c.azh[z], bzh = P.c.bzh[z];
  const Kdiv kx = kdiv_of(P.c.Kx[x]), kz = kdiv_of(P.c.Kz[z]);
  const Kdiv kxh = kdiv_of(P.c.Kxh[x]), kzh = kdiv_of(P.c.Kzh[z]);
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const double dt64 = (double)dt;
  const int ns = g.ns;
  const bool inj_row = P.inj.res && z == P.inj.z;
  const float* res_row = P.inj.res ? P.inj.res + (size_t)it * g.R : nullptr;
  for (int s0 = 0; s0 < ns; s0 += U) {
    AdjStressIn r[U];
    float resv[U];
    bool has[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      resv[u] = 0.f;
      has[u] = false;
      if (s >= ns) break;
      if (inj_row) {
        const int k = P.inj.at[(size_t)(g.s0 + s) * g.nx + x];
        if (k >= 0) {
          resv[u] = res_row[k];
          has[u] = true;
        }
      }
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      AdjStressIn& q = r[u];
      const f4 vzc = lz4(A, sw, kVz, C, -2), vxc = lz4(A, sw, kVx, C, -1);
      q.vzm2 = vzc[0], q.vzm1 = vzc[1], q.vz0 = vzc[2], q.vzp1 = vzc[3];
      q.vzpx = lx(A, sw, kVz, C, 1), q.vzp2x = lx(A, sw, kVz, C, 2), q.vzmx = lx(A, sw, kVz, C, -1);
      q.vxm1 = vxc[0], q.vx0 = vxc[1], q.vxp1 = vxc[2], q.vxp2 = vxc[3];
      q.vxmx = lx(A, sw, kVx, C, -1), q.vxpx = lx(A, sw, kVx, C, 1), q.vxm2x = lx(A, sw, kVx, C, -2);
      if (ax != 0.f) {  // psi_sxzx at x+1, x, x+2, x-1
        q.q1[0] = lx(M, sm, kMsxzx, C, 1), q.q1[1] = lz(M, sm, kMsxzx, C, 0), q.q1[2] = lx(M, sm, kMsxzx, C, 2), q.q1[3] = lx(M, sm, kMsxzx, C, -1);
      }
      if (az != 0.f) {  // psi_sxzz at z+1, z, z+2, z-1
        const f4 t = lz4(M, sm, kMsxzz, C, -1);
        q.q2[0] = t[2], q.q2[1] = t[1], q.q2[2] = t[3], q.q2[3] = t[0];
      }
      if (axh != 0.f) {  // psi_sxx at x, x-1, x+1, x-2
        q.q3[0] = lz(M, sm, kMsxx, C, 0), q.q3[1] = lx(M, sm, kMsxx, C, -1), q.q3[2] = lx(M, sm, kMsxx, C, 1), q.q3[3] = lx(M, sm, kMsxx, C, -2);
      }
      if (azh != 0.f) {  // psi_szz at z, z-1, z+1, z-2
        const f4 t = lz4(M, sm, kMszz, C, -2);
        q.q4[0] = t[2], q.q4[1] = t[1], q.q4[2] = t[3], q.q4[3] = t[0];
      }
      if (xlive) q.m_vzx = lz(M, sm, kMvzx, C, 0), q.m_vxx = lz(M, sm, kMvxx, C, 0);
      if (zlive) q.m_vxz = lz(M, sm, kMvxz, C, 0), q.m_vzz = lz(M, sm, kMvzz, C, 0);
      q.sxz = lz(A, sw, kSxz, C, 0), q.sxx = lz(A, sw, kSxx, C, 0), q.szz = lz(A, sw, kSzz, C, 0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= ns) break;
      const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb;
      const AdjStressIn& q = r[u];
      // sxz += a_x dphi_xz_x/dx + dvz/dx / K_x ba dt + a_z dphi_xz_z/dz + dvx/dz / K_z bb dt
      const float dvz_dx = dad(q.vzpx, q.vz0, q.vzp2x, q.vzmx, dx, rdx);
      const float dvx_dz = dad(q.vxp1, q.vx0, q.vxp2, q.vxm1, dz, rdz);
      float e = kdiv(dvz_dx, kx) * ba * dt;
      if (ax != 0.f) e = ax * dad(q.q1[0], q.q1[1], q.q1[2], q.q1[3], dx, rdx) + e;
      if (az != 0.f) e = e + az * dad(q.q2[0], q.q2[1], q.q2[2], q.q2[3], dz, rdz);
      e = e + kdiv(dvx_dz, kz) * bb * dt;
      const float sxz = q.sxz + e;
      st(A, sw, kSxz, C, sxz);
      if (xlive) st(M, sm, kMvzx, C, bxh * q.m_vzx + sxz * amu * dt);
      if (zlive) st(M, sm, kMvxz, C, bzh * q.m_vxz + sxz * amu * dt);
      const float dvx_dx = dad(q.vx0, q.vxmx, q.vxpx, q.vxm2x, dx, rdx);
      const float dvz_dz = dad(q.vz0, q.vzm1, q.vzp1, q.vzm2, dz, rdz);
      float sxx0 = q.sxx, szz0 = q.szz;
      if (has[u]) {
        szz0 = szz0 + resv[u];
        sxx0 = (float)((double)sxx0 + kRS * (double)resv[u]);
      }
      // sxx += a_x_half dphi_xx/dx + bb dvx/dx / K_x_half dt; szz likewise in z
      float t = kdiv(bb * dvx_dx, kxh) * dt;
      if (axh != 0.f) t = axh * dad(q.q3[0], q.q3[1], q.q3[2], q.q3[3], dx, rdx) + t;
      const float sxx = sxx0 + t;
      t = kdiv(ba * dvz_dz, kzh) * dt;
      if (azh != 0.f) t = azh * dad(q.q4[0], q.q4[1], q.q4[2], q.q4[3], dz, rdz) + t;
      const float szz = szz0 + t;
      st(A, sw, kSxx, C, sxx);
      st(A, sw, kSzz, C, szz);
      // psi_xx = b_x psi + lambda szz dt + (lambda + 2.0 mu) sxx dt;
      // psi_zz = b_z psi + (lambda + 2.0 mu) szz dt + lambda sxx dt
      if (xlive)
        st(M, sm, kMvxx, C, (float)((double)(bx * q.m_vxx + lam * szz * dt) + l2m * (double)sxx * dt64));
      if (zlive)
        st(M, sm, kMvzz, C, (float)((double)(bz * q.m_vzz) + l2m * (double)szz * dt64 + (double)(lam * sxx * dt)));
    }
  }
}

// ------------------------------------------------------------ exact adjoint
// No reference counterpart: the transpose of the forward time step
// (el_stress_fwd, add_source, el_velocity_fwd, recording), taken operation by
// operation, so that the gradient is the derivative of the discrete misfit
// also where waves enter the C-PML and in zero-shear layers.  The adjoint
// state keeps the reference's sign (the residual is injected with +, so it is
// -dJ/d(state)) and the gradient sums keep the meaning of the reference
// mode's, so k_finalize turns both into (lambda, mu, rho) gradients.
//
// A stencil difference f -> D f of the forward step becomes its transpose
// acting on the coefficient-weighted adjoint, and a C-PML term
// d~ = d / K + psi', psi' = b psi + a d becomes, for the adjoint P of psi,
// d_adj = g / K + a (g + P), P <- b (g + P) (only where the forward updates
// psi, with the forward's gates, including el_velocity.cu:56,71's
// x > nx - nPml).  Each transposed update runs as two passes over the
// update region: the first forms the weighted operands of every cell
// (xt, zero outside the update region), the second applies the transposed
// stencils to them.  The second pass of one update and the first of the next
// touch only their own cell's adjoint, so they share a launch, and so does
// the reverse re-materialisation step that reads the same forward field:
//   k_exact_a(it): reverse velocity step it (v^{it+1} -> v^it from
//     sigma^{it+1}), transposed stress update of step it+1 into v_adj,
//     velocity operands of step it and the density sums from sigma^{it+1};
//   k_exact_b(it): residual of sample it+1 (folded injection), transposed
//     velocity update of step it into sigma_adj, stf gradient, stress
//     operands of step it and the lambda / mu sums from v^it, reverse stress
//     step it (sigma^{it+1} -> sigma^it from v^it).
// Each reversal writes only its cell's own forward field, which the adjoint
// half of its launch does not read.  xt holds two operand sets per shot
// (slots 0-3: Azz, Axx, Axz, Azx of the stress update; 4-7: Bz, Bzx, Bxz, Bx
// of the velocity update).  Outside the C-PML frame (cells whose row or
// column is gated) Azx = Axz, Bzx = Bz and Bx = Bxz, so slots 3, 5 and 7 are
// written and read only inside the frame, and a reader takes slots 2, 4 and 6
// for neighbours outside it.
__device__ __forceinline__ float pml_t(float g, const Kdiv& K, float a, float b, float& P) {
  const float tot = g + P;
  P = b * tot;
  return kdiv(g, K) + a * tot;
}
// The transposed differences are written through dfw: for a stride st,
// (D+)^T B at j = (c1 (B[j-st] - B[j]) - c2 (B[j-2st] - B[j+st])) / h and
// (D-)^T B at j = (c1 (B[j] - B[j+st]) - c2 (B[j-st] - B[j+2st])) / h, for
// D+ f = (c1 (f[k+1] - f[k]) - c2 (f[k+2] - f[k-1])) / h and
// D- f = (c1 (f[k] - f[k-1]) - c2 (f[k+1] - f[k-2])) / h.

// x neighbour k (-2..2) of the cell from operand slot `own` when the
// neighbour lies in the C-PML frame (`f`), else from slot `base`; the slot
// goes into the per-lane offset, so `f` may differ between lanes.
__device__ __forceinline__ float lxf(const Arr& X, int sx, int own, int base, const Cell& C, int k,
                                     bool f) {
  const int v = k == 0 ? C.zm2 + 8 : k == -2 ? C.xm2 : k == -1 ? C.xm1 : k == 1 ? C.xp1 : C.xp2;
  return ldb(X.r, sx, v + (f ? own : base) * X.nb);
}

// The cell of the calling thread: 64 x 4 tiles over the update region,
// z fastest, in XCD bands (workgroup b runs on XCD b % 8, and XCD k takes
// the k-th eighth of the tiles, so the columns a tile shares with its
// neighbours stay in that XCD's L2).  The grid is 8 * ceil(tiles / 8)
// workgroups of 64 x 4; false for threads outside the update region.
__device__ __forceinline__ bool exact_cell(const Geo& g, int& z, int& x) {
  const int gz = (g.nzA + kBZ - 1) / kBZ, gx = (g.nx + 3) / 4;
  const int tiles = gz * gx, per = (tiles + 7) / 8;
  const int b = (int)blockIdx.x, t = (b & 7) * per + (b >> 3);
  if (t >= tiles) return false;
  z = (t % gz) * kBZ + (int)threadIdx.x;
  x = wave_col((t / gz) * 4);
  return z >= 2 && z <= g.nzA - 3 && x >= 2 && x <= g.nx - 3;
}

__global__ __launch_bounds__(256) void k_exact_a(Part P, float* __restrict__ xt, int it, int first) {
  const Geo& g = P.g;
  int z, x;
  if (!exact_cell(g, z, x)) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const bool in = z >= g.nPml && z <= g.nzA - 1 - g.nPml && x >= g.nPml && x <= g.nx - 1 - g.nPml;
  const int slot = ring_slot(z, x, g);
  const bool re = in || slot >= 0;  // the cell takes part in the reversal
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml;  // el_velocity's gate
  const Cell C = cell_of((unsigned)i * 4u, nz);
  const int Nb = (int)(g.N * 4);
  const Arr W = arr(P.w, Nb), A = arr(P.a, Nb), M = arr(P.p, Nb), X = arr(xt, Nb);
  const float ba = P.m.ba[i], bb = P.m.bb[i];
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  Kdiv kzh{}, kz{}, kx{}, kxh{};
  float azh = 0.f, bzh = 0.f, az = 0.f, bz = 0.f, ax = 0.f, bx = 0.f, axh = 0.f, bxh = 0.f;
  if (pz) {
    kzh = kdiv_of(P.c.Kzh[z]), azh = P.c.azh[z], bzh = P.c.bzh[z];
    kz = kdiv_of(P.c.Kz[z]), az = P.c.az[z], bz = P.c.bz[z];
  }
  if (px) {
    kx = kdiv_of(P.c.Kx[x]), ax = P.c.ax[x], bx = P.c.bx[x];
    kxh = kdiv_of(P.c.Kxh[x]), axh = P.c.axh[x], bxh = P.c.bxh[x];
  }
  float ga = 0.f, gb = 0.f;
  for (int s = 0; s < g.ns; ++s) {
    const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb, sx = s * 8 * Nb;
    // forward stresses sigma^{it+1} around the cell (interior cells)
    float fz = 0.f, fx = 0.f;
    if (in) {
      const f4 szc = lz4(W, sw, kSzz, C, -1), sxc = lz4(W, sw, kSxz, C, -2);
      const float sxzmx = lx(W, sw, kSxz, C, -1), sxzpx = lx(W, sw, kSxz, C, 1),
                  sxzm2x = lx(W, sw, kSxz, C, -2);
      const float sxxpx = lx(W, sw, kSxx, C, 1), sxx0 = lz(W, sw, kSxx, C, 0),
                  sxxp2x = lx(W, sw, kSxx, C, 2), sxxmx = lx(W, sw, kSxx, C, -1);
      const float dszz_dz = dfw(szc[2], szc[1], szc[3], szc[0], dz, rdz);
      const float dsxz_dx = dfw(sxc[2], sxzmx, sxzpx, sxzm2x, dx, rdx);
      const float dsxz_dz = dfw(sxc[2], sxc[1], sxc[3], sxc[0], dz, rdz);
      const float dsxx_dx = dfw(sxxpx, sxx0, sxxp2x, sxxmx, dx, rdx);
      fz = dszz_dz + dsxz_dx;
      fx = dsxz_dz + dsxx_dx;
    }
    float wvz = 0.f, wvx = 0.f;
    if (re) {
      if (slot < 0) {
        wvz = lz(W, sw, kVz, C, 0), wvx = lz(W, sw, kVx, C, 0);
      } else {
        wvz = *ring_ptr(P.ring, g, it, 3, s, slot);
        wvx = *ring_ptr(P.ring, g, it, 4, s, slot);
      }
    }
    float vz = lz(A, sw, kVz, C, 0), vx = lz(A, sw, kVx, C, 0);
    if (!first) {
      const f4 azz = lz4(X, sx, 0, C, -1);  // Azz at z-1 .. z+2
      const f4 axz = lz4(X, sx, 2, C, -2);  // Axz at z-2 .. z+1
      // Azx (slot 3 in the stress update's frame, else Axz)
      auto fs = [&](int k) { return pz || x + k < g.nPml || x + k > g.nx - g.nPml - 1; };
      const float azxm = lxf(X, sx, 3, 2, C, -1, fs(-1)), azx0 = lxf(X, sx, 3, 2, C, 0, fs(0)),
                  azxm2 = lxf(X, sx, 3, 2, C, -2, fs(-2)), azxp = lxf(X, sx, 3, 2, C, 1, fs(1));
      const float axx0 = lz(X, sx, 1, C, 0), axxp = lx(X, sx, 1, C, 1),
                  axxm = lx(X, sx, 1, C, -1), axxp2 = lx(X, sx, 1, C, 2);
      // (D-)^T Azz along z + (D+)^T Azx along x; (D-)^T Axx along x + (D+)^T Axz along z
      vz = vz + (dfw(azz[1], azz[2], azz[0], azz[3], dz, rdz) + dfw(azxm, azx0, azxm2, azxp, dx, rdx));
      vx = vx + (dfw(axx0, axxp, axxm, axxp2, dx, rdx) + dfw(axz[1], axz[2], axz[0], axz[3], dz, rdz));
      st(A, sw, kVz, C, vz);
      st(A, sw, kVx, C, vx);
    }
    const float gz = dt * ba * vz, gx = dt * bb * vx;
    float oz = gz, ozx = gz, oxz = gx, ox = gx;
    if (pz) {
      float q = lz(M, sm, kMszz, C, 0);
      oz = pml_t(gz, kzh, azh, bzh, q);
      st(M, sm, kMszz, C, q);
      q = lz(M, sm, kMsxzz, C, 0);
      oxz = pml_t(gx, kz, az, bz, q);
      st(M, sm, kMsxzz, C, q);
    }
    if (px) {
      float q = lz(M, sm, kMsxzx, C, 0);
      ozx = pml_t(gz, kx, ax, bx, q);
      st(M, sm, kMsxzx, C, q);
      q = lz(M, sm, kMsxx, C, 0);
      ox = pml_t(gx, kxh, axh, bxh, q);
      st(M, sm, kMsxx, C, q);
    }
    st(X, sx, 4, C, oz);
    st(X, sx, 6, C, oxz);
    if (pz || px) {  // the velocity update's frame
      st(X, sx, 5, C, ozx);
      st(X, sx, 7, C, ox);
    }
    if (in) {
      ga += vz * fz;
      gb += vx * fx;
    }
    if (re) {
      if (slot < 0) {  // el_velocity.cu:84-117, isFor = false
        wvz = wvz - fz * ba * dt;
        wvx = wvx - fx * bb * dt;
      }
      st(W, sw, kVz, C, wvz);
      st(W, sw, kVx, C, wvx);
    }
  }
  if (in) {
    P.acc[3 * g.N + i] += ga;
    P.acc[4 * g.N + i] += gb;
  }
}

__global__ __launch_bounds__(256) void k_exact_b(Part P, float* __restrict__ xt,
                                                 const float* __restrict__ srcwin, int it) {
  const Geo& g = P.g;
  int z, x;
  if (!exact_cell(g, z, x)) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const bool in = z >= g.nPml && z <= g.nzA - 1 - g.nPml && x >= g.nPml && x <= g.nx - 1 - g.nPml;
  const int slot = ring_slot(z, x, g);
  const bool re = in || slot >= 0;  // the cell takes part in the reversal
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml - 1;  // el_stress's gate
  const Cell C = cell_of((unsigned)i * 4u, nz);
  const int Nb = (int)(g.N * 4);
  const Arr W = arr(P.w, Nb), A = arr(P.a, Nb), M = arr(P.p, Nb), X = arr(xt, Nb);
  const float lam = P.m.lam[i], mu = P.m.mu[i], amu = P.m.amu[i];
  const double l2m64 = (double)lam + 2.0 * (double)mu;
  const float l2m = (float)l2m64;
  const float dt = g.dt, dz = g.dz, dx = g.dx, rdz = g.rdz, rdx = g.rdx;
  const double dt64 = (double)dt;
  Kdiv kz{}, kzh{}, kx{}, kxh{};
  float az = 0.f, bz = 0.f, azh = 0.f, bzh = 0.f, ax = 0.f, bx = 0.f, axh = 0.f, bxh = 0.f;
  if (pz) {
    kz = kdiv_of(P.c.Kz[z]), az = P.c.az[z], bz = P.c.bz[z];
    kzh = kdiv_of(P.c.Kzh[z]), azh = P.c.azh[z], bzh = P.c.bzh[z];
  }
  if (px) {
    kx = kdiv_of(P.c.Kx[x]), ax = P.c.ax[x], bx = P.c.bx[x];
    kxh = kdiv_of(P.c.Kxh[x]), axh = P.c.axh[x], bxh = P.c.bxh[x];
  }
  const bool inj_row = P.inj.res && z == P.inj.z;
  const float* res_row = P.inj.res ? P.inj.res + (size_t)(it + 1) * g.R : nullptr;
  float gl = 0.f, gm = 0.f, gs = 0.f;
  for (int s = 0; s < g.ns; ++s) {
    const int sw = s * kWaveComps * Nb, sm = s * kMemComps * Nb, sx = s * 8 * Nb;
    const int zs = P.src.zx[2 * s], xs = P.src.zx[2 * s + 1];
    // forward velocities v^it around the cell (interior cells)
    float dvz_dz = 0.f, dvx_dx = 0.f, dvx_dz = 0.f, dvz_dx = 0.f;
    if (in) {
      const f4 vzc = lz4(W, sw, kVz, C, -2), vxc = lz4(W, sw, kVx, C, -1);
      const float vzpx = lx(W, sw, kVz, C, 1), vzp2x = lx(W, sw, kVz, C, 2),
                  vzmx = lx(W, sw, kVz, C, -1);
      const float vxmx = lx(W, sw, kVx, C, -1), vxpx = lx(W, sw, kVx, C, 1),
                  vxm2x = lx(W, sw, kVx, C, -2);
      dvz_dz = dfw(vzc[2], vzc[1], vzc[3], vzc[0], dz, rdz);
      dvx_dx = dfw(vxc[1], vxmx, vxpx, vxm2x, dx, rdx);
      dvx_dz = dfw(vxc[2], vxc[1], vxc[3], vxc[0], dz, rdz);
      dvz_dx = dfw(vzpx, vzc[2], vzp2x, vzmx, dx, rdx);
    }
    float wszz = 0.f, wsxx = 0.f, wsxz = 0.f;
    if (re) {
      if (slot < 0) {
        wszz = lz(W, sw, kSzz, C, 0), wsxx = lz(W, sw, kSxx, C, 0), wsxz = lz(W, sw, kSxz, C, 0);
      } else {
        wszz = *ring_ptr(P.ring, g, it, 0, s, slot);
        wsxz = *ring_ptr(P.ring, g, it, 1, s, slot);
        wsxx = *ring_ptr(P.ring, g, it, 2, s, slot);
      }
    }
    float szz = lz(A, sw, kSzz, C, 0), sxx = lz(A, sw, kSxx, C, 0), sxz = lz(A, sw, kSxz, C, 0);
    if (inj_row) {
      const int k = P.inj.at[(size_t)(g.s0 + s) * g.nx + x];
      if (k >= 0) {
        const float v = res_row[k];
        szz = szz + v;
        sxx = (float)((double)sxx + kRS * (double)v);
      }
    }
    {
      const f4 bzc = lz4(X, sx, 4, C, -2);   // Bz at z-2 .. z+1
      const f4 bxzc = lz4(X, sx, 6, C, -1);  // Bxz at z-1 .. z+2
      // Bzx and Bx (slots 5 and 7 in the velocity update's frame, else Bz and Bxz)
      auto fv = [&](int k) { return pz || x + k < g.nPml || x + k > g.nx - g.nPml; };
      const float bzx0 = lxf(X, sx, 5, 4, C, 0, fv(0)), bzxp = lxf(X, sx, 5, 4, C, 1, fv(1)),
                  bzxm = lxf(X, sx, 5, 4, C, -1, fv(-1)), bzxp2 = lxf(X, sx, 5, 4, C, 2, fv(2));
      const float bxm = lxf(X, sx, 7, 6, C, -1, fv(-1)), bx0 = lxf(X, sx, 7, 6, C, 0, fv(0)),
                  bxm2 = lxf(X, sx, 7, 6, C, -2, fv(-2)), bxp = lxf(X, sx, 7, 6, C, 1, fv(1));
      // (D+)^T Bz along z; (D-)^T Bzx along x + (D-)^T Bxz along z; (D+)^T Bx along x
      szz = szz + dfw(bzc[1], bzc[2], bzc[0], bzc[3], dz, rdz);
      sxz = sxz + (dfw(bzx0, bzxp, bzxm, bzxp2, dx, rdx) + dfw(bxzc[1], bxzc[2], bxzc[0], bxzc[3], dz, rdz));
      sxx = sxx + dfw(bxm, bx0, bxm2, bxp, dx, rdx);
    }
    st(A, sw, kSzz, C, szz);
    st(A, sw, kSxx, C, sxx);
    st(A, sw, kSxz, C, sxz);
    // d(szz, sxx)/d(stf) at the source: 1500^2 dt and the window (add_source)
    if (P.stf_grad && z == zs && x == xs)
      P.stf_grad[(size_t)s * g.nSteps + it] =
          (float)(-(double)kSrcScale * dt64 * ((double)szz + kRS * (double)sxx) *
                  (double)srcwin[it]);
    const float Gzz = dt * (l2m * szz + lam * sxx);
    const float Gxx = dt * (lam * szz + l2m * sxx);
    const float Gs = dt * amu * sxz;
    float ozz = Gzz, oxx = Gxx, oxz = Gs, ozx = Gs;
    if (pz) {
      float q = lz(M, sm, kMvzz, C, 0);
      ozz = pml_t(Gzz, kz, az, bz, q);
      st(M, sm, kMvzz, C, q);
      q = lz(M, sm, kMvxz, C, 0);
      oxz = pml_t(Gs, kzh, azh, bzh, q);
      st(M, sm, kMvxz, C, q);
    }
    if (px) {
      float q = lz(M, sm, kMvxx, C, 0);
      oxx = pml_t(Gxx, kx, ax, bx, q);
      st(M, sm, kMvxx, C, q);
      q = lz(M, sm, kMvzx, C, 0);
      ozx = pml_t(Gs, kxh, axh, bxh, q);
      st(M, sm, kMvzx, C, q);
    }
    st(X, sx, 0, C, ozz);
    st(X, sx, 1, C, oxx);
    st(X, sx, 2, C, oxz);
    if (pz || px) st(X, sx, 3, C, ozx);  // the stress update's frame
    if (in) {
      gl += (szz + sxx) * (dvz_dz + dvx_dx);
      gm += szz * dvz_dz + sxx * dvx_dx;
      gs += sxz * (dvx_dz + dvz_dx);
    }
    if (re) {
      if (slot < 0) {  // add_source(-) then el_stress.cu:90-131, isFor = false
        if (z == zs && x == xs) {
          const float amp = P.src.amp[(size_t)s * g.nSteps + it];
          wszz = wszz - kSrcScale * amp * dt;
          wsxx = (float)((double)wsxx - kRS * (double)kSrcScale * (double)amp * dt64);
        }
        wszz = (float)((double)wszz - (l2m64 * (double)dvz_dz + (double)(lam * dvx_dx)) * dt64);
        wsxx = (float)((double)wsxx - ((double)(lam * dvz_dz) + l2m64 * (double)dvx_dx) * dt64);
        wsxz = wsxz - amu * (dvx_dz + dvz_dx) * dt;
      }
      st(W, sw, kSzz, C, wszz);
      st(W, sw, kSxx, C, wsxx);
      st(W, sw, kSxz, C, wsxz);
    }
  }
  if (in) {
    P.acc[i] += gl;
    P.acc[g.N + i] += gm;
    P.acc[2 * g.N + i] += gs;
  }
}

// ------------------------------------------------------------ step kernels
template <int NP>
__device__ __forceinline__ const Part& part_of(const Parts<NP>& ps, unsigned e) {
  return ps.get((int)((e >> 24) & 31u));
}

template <int NP, int U>
__global__ __launch_bounds__(256) void k_fwd_stress(Parts<NP> ps, const unsigned* __restrict__ map,
                                                    int it) {
  const unsigned e = map[blockIdx.x];
  if ((e >> 29) == kNop) return;
  stress_fwd<U>(part_of(ps, e), (int)(e & 0xffffffu), it);
}

// the velocity update of step it and the recording of the stresses it read
// (recording reads only stresses, which the velocity update does not touch)
template <int NP, int U>
__global__ __launch_bounds__(256) void k_fwd_vel(Parts<NP> ps, const unsigned* __restrict__ map,
                                                 int it) {
  const unsigned e = map[blockIdx.x];
  const unsigned kind = e >> 29;
  if (kind == kVelFwd)
    vel_fwd<U>(part_of(ps, e), (int)(e & 0xffffffu));
  else if (kind == kRecord)
    record(part_of(ps, e), (int)(e & 0xffffffu), it);
}

// Backward pairs with disjoint reads and writes share a launch:
//   a: rev_stress(it) + adj_vel(it)   (rev_stress reads the adjoint stress that
//      adj_vel also only reads; adj_vel writes the adjoint velocity that
//      rev_stress does not touch)
//   b: adj_stress(it) + rev_vel(it-1) (rev_vel reads the adjoint velocity that
//      adj_stress also only reads; adj_stress writes the adjoint stress that
//      rev_vel does not touch)
template <int NP, int U>
__global__ __launch_bounds__(256) void k_bwd_a(Parts<NP> ps, const unsigned* __restrict__ map,
                                               int it) {
  const unsigned e = map[blockIdx.x];
  const unsigned kind = e >> 29;
  if (kind == kRevStress)
    rev_stress<U>(part_of(ps, e), (int)(e & 0xffffffu), it);
  else if (kind == kAdjVel)
    adj_vel<U>(part_of(ps, e), (int)(e & 0xffffffu), it);
}

template <int NP, int U>
__global__ __launch_bounds__(256) void k_bwd_b(Parts<NP> ps, const unsigned* __restrict__ map,
                                               int it) {
  const unsigned e = map[blockIdx.x];
  const unsigned kind = e >> 29;
  if (kind == kAdjStress)
    adj_stress<U>(part_of(ps, e), (int)(e & 0xffffffu), it);
  else if (kind == kRevVel)
    rev_vel<U>(part_of(ps, e), (int)(e & 0xffffffu), it - 1);
}

// res_injection (utilities.cu:569-580) for surveys whose receivers are not one
// per cell in one row: one thread per injection cell adds its receivers'
// residuals in receiver order (szz_adj += res, sxx_adj += RSXXZZ res), the
// order in which the reference's per-receiver += lands on a shared cell when
// its threads run in index order.  cells[c] = (plan shot, z, x); receivers
// of cell c are rid[ptr[c] .. ptr[c+1]).
__global__ void k_res_inject(Geo g, float* __restrict__ adj, const int* __restrict__ cells,
                             const int* __restrict__ ptr, const int* __restrict__ rid, int c0,
                             int ncells, const float* __restrict__ row) {
  const int c = c0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= c0 + ncells) return;
  float* a = adj + (size_t)cells[3 * c] * kWaveComps * g.N + (long long)cells[3 * c + 2] * g.nz +
             cells[3 * c + 1];
  float szz = a[kSzz * g.N], sxx = a[kSxx * g.N];
  for (int j = ptr[c]; j < ptr[c + 1]; ++j) {
    const float v = row[rid[j]];
    szz = szz + v;
    sxx = (float)((double)sxx + kRS * (double)v);
  }
  a[kSzz * g.N] = szz;
  a[kSxx * g.N] = sxx;
}

// Residual (libCUFD.cu:258-313 without the optional filters) over receivers
// [r0, r0 + nr) of the plan: window obs and syn with w^2, res = obs_w - syn_w
// (first sample 0, gpuMinus utilities.cu:154-167), sum res^2
// (cuda_cal_objective, here in FP64), then window res again.
__global__ __launch_bounds__(256) void k_residual(int nSteps, int R, int r0, int nr,
                                                  const float* __restrict__ w2,
                                                  const float* __restrict__ obs,
                                                  const float* __restrict__ syn,
                                                  float* __restrict__ res,
                                                  double* __restrict__ partial) {
  __shared__ double red[256];
  const size_t total = (size_t)nSteps * nr;
  double acc = 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(k / (size_t)nr);
    const size_t j = (size_t)t * R + r0 + (k - (size_t)t * nr);
    const float ww = w2[t];
    // obs and syn are windowed separately (two roundings, as the reference's
    // separate cuda_window launches do)
    const float ow = obs[j] * ww;
    const float sw = syn[j] * ww;
    const float rv = t > 0 ? ow - sw : 0.f;
    acc += (double)rv * (double)rv;
    res[j] = rv * ww;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) acc += partial[k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Chain rule from the accumulated derivatives to (lambda, mu, rho), once per
// run.  Deterministic gather form of the reference's atomicAdd sprays
// (el_stress.cu:113-124, el_velocity.cu:104-112), with the same target
// conditions (including the always-true `gidx+1<=gidx<=nx-1-nPml`).  With
// `exact` (the exact adjoint) every cell of a harmonic or arithmetic average
// receives its share: the conditions that drop the last interior row and
// column's neighbours are not applied.
__global__ __launch_bounds__(256) void k_finalize(Geo g, Mdl m, const float* __restrict__ acc,
                                                  float* __restrict__ grad, int exact) {
  const int z = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int x = (int)blockIdx.y;
  if (z >= g.nz) return;
  const int nz = g.nz;
  const long long N = g.N;
  const float* acc_l = acc;
  const float* acc_m = acc + N;
  const float* acc_s = acc + 2 * N;
  const float* acc_a = acc + 3 * N;
  const float* acc_b = acc + 4 * N;
  const int z0 = g.nPml, z1 = g.nzA - 1 - g.nPml, x0 = g.nPml, x1 = g.nx - 1 - g.nPml;
  auto inI = [&](int zz, int xx) { return zz >= z0 && zz <= z1 && xx >= x0 && xx <= x1; };
  const long long i = (long long)x * nz + z;
  const float dt = g.dt;
  float gl = 0.f, gm = 0.f, gd = 0.f;
  if (inI(z, x)) {
    gl = -acc_l[i] * dt * 1.0e6f;
    gm = -2.f * acc_m[i] * dt * 1.0e6f;
  }
  auto shear = [&](int zz, int xx) -> float {
    const long long j = (long long)xx * nz + zz;
    const float am = m.amu[j];
    if (am == 0.f) return 0.f;
    const float S = 1.f / m.mu[j] + 1.f / m.mu[j + 1] + 1.f / m.mu[j + nz] + 1.f / m.mu[j + nz + 1];
    return -acc_s[j] * dt * am / S * 1.0e6f;
  };
  float t = 0.f;
  bool any = false;
  if (inI(z, x) && m.amu[i] != 0.f) t += shear(z, x), any = true;
  const bool zc = exact || z <= z1, xc = exact || x <= x1;
  if (inI(z - 1, x) && zc && m.amu[i - 1] != 0.f) t += shear(z - 1, x), any = true;
  if (inI(z, x - 1) && m.amu[i - nz] != 0.f) t += shear(z, x - 1), any = true;
  if (inI(z - 1, x - 1) && zc && xc && m.amu[i - nz - 1] != 0.f)
    t += shear(z - 1, x - 1), any = true;
  if (any) {
    const float mu = m.mu[i];
    gm += t / (mu * mu);
  }
  auto ga = [&](long long j) { const float b = m.ba[j]; return acc_a[j] * dt * (0.5f * b * b); };
  auto gb = [&](long long j) { const float b = m.bb[j]; return acc_b[j] * dt * (0.5f * b * b); };
  if (inI(z, x)) gd += ga(i) + gb(i);
  if (inI(z - 1, x) && zc) gd += ga(i - 1);
  if (inI(z, x - 1)) gd += gb(i - nz);
  grad[i] = gl;
  grad[N + i] = gm;
  grad[2 * N + i] = gd;
}

// aveMuInit / aveBycInit (utilities.cu:125-152) over [2, nz-3] x [2, nx-3];
// defaults 0 and 1/1000 elsewhere (Model.cu:67-73).
__global__ void k_model_init(Geo g, const float* __restrict__ mu, const float* __restrict__ den,
                             float* __restrict__ amu, float* __restrict__ ba,
                             float* __restrict__ bb) {
  const int z = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int x = (int)blockIdx.y;
  if (z >= g.nz) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  float am = 0.f, fa = 1.f / 1000.f, fb = 1.f / 1000.f;
  if (z >= 2 && z <= g.nz - 3 && x >= 2 && x <= g.nx - 3) {
    const float A = mu[i], B = mu[i + 1], C = mu[i + nz], D = mu[i + nz + 1];
    if (A == 0.f || B == 0.f || C == 0.f || D == 0.f)
      am = 0.f;
    else
      am = (float)(4.0 / (1.0 / A + 1.0 / B + 1.0 / C + 1.0 / D));
    fa = (float)(2.0 / (double)(den[i + 1] + den[i]));
    fb = (float)(2.0 / (double)(den[i + nz] + den[i]));
  }
  amu[i] = am;
  ba[i] = fa;
  bb[i] = fb;
}

// [rows][cols] -> [cols][rows] through LDS tiles.
__global__ void k_transpose(const float* __restrict__ in, float* __restrict__ out, int rows,
                            int cols) {
  __shared__ float tile[32][33];
  const int c = blockIdx.x * 32 + threadIdx.x;
  const int r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int r = r0 + k;
    if (r < rows && c < cols) tile[k][threadIdx.x] = in[(size_t)r * cols + c];
  }
  __syncthreads();
  const int r = r0 + threadIdx.x;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int cc = blockIdx.x * 32 + k;
    if (r < rows && cc < cols) out[(size_t)cc * rows + r] = tile[threadIdx.x][k];
  }
}

}  // namespace

// ===================================================================== host
namespace {

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

// div1(x, d, RN(1/d)) == x / d for every normal x: checked over every
// mantissa of the binades [1, 2) and [2, 4) (the computation scales exactly
// with the exponent of x).  Cached per spacing; about 30 ms per new value.
bool spacing_ok(float d) {
  static std::mutex mu;
  static std::map<float, bool> seen;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = seen.find(d);
    if (it != seen.end()) return it->second;
  }
  const float y = 1.0f / d;
  bool ok = std::isnormal(d) && std::isnormal(y);
  for (unsigned e = 127; e <= 128 && ok; ++e)
    for (unsigned m = 0; m < (1u << 23); ++m) {
      const unsigned u = e << 23 | m;
      float x;
      std::memcpy(&x, &u, 4);
      const float q = x * y;
      const float v = std::fma(std::fma(-d, q, x), y, q);
      if (v != x / d) {
        ok = false;
        break;
      }
    }
  std::lock_guard<std::mutex> lk(mu);
  seen[d] = ok;
  return ok;
}

// ---------------------------------------------------- workgroup maps
// Each XCD's ordered list of work; the launch grid interleaves them so that
// workgroup b = 8 j + k is entry j of XCD k's list.
struct Lists {
  std::vector<unsigned> l[8];
  size_t longest() const {
    size_t n = 0;
    for (const auto& v : l) n = std::max(n, v.size());
    return n;
  }
  bool empty() const { return longest() == 0; }
};

// tiles [0, n) in 8 contiguous bands (z fastest, then x): neighbouring tiles,
// which share halo columns, share an XCD's L2
void band(Lists& L, unsigned kind, int part, int n) {
  for (int k = 0; k < 8; k++)
    for (int t = (int)((long long)n * k / 8); t < (int)((long long)n * (k + 1) / 8); t++)
      L.l[k].push_back(wg_entry(kind, (unsigned)part, (unsigned)t));
}

// The update-region tiling, whose first `hl` and last `hr` tile columns are
// heavier (the x-direction C-PML): each XCD gets one contiguous run of the
// heavy tiles, first, then one contiguous run of the others.
void balanced(Lists& L, unsigned kind, int part, int gz, int gx, int hl, int hr) {
  std::vector<int> heavy, light;
  for (int bx = 0; bx < gx; bx++) {
    const bool h = bx < hl || bx >= gx - hr;
    for (int bz = 0; bz < gz; bz++) (h ? heavy : light).push_back(bx * gz + bz);
  }
  const size_t H = heavy.size(), Ln = light.size();
  for (int k = 0; k < 8; k++) {
    for (size_t q = H * k / 8; q < H * (k + 1) / 8; q++)
      L.l[k].push_back(wg_entry(kind, (unsigned)part, (unsigned)heavy[q]));
    for (size_t q = Ln * k / 8; q < Ln * (k + 1) / 8; q++)
      L.l[k].push_back(wg_entry(kind, (unsigned)part, (unsigned)light[q]));
  }
}

// items [0, n) dealt round robin
void dealt(Lists& L, unsigned kind, int part, int n) {
  for (int t = 0; t < n; t++) L.l[t % 8].push_back(wg_entry(kind, (unsigned)part, (unsigned)t));
}

void append(Lists& A, const Lists& B) {
  for (int k = 0; k < 8; k++) A.l[k].insert(A.l[k].end(), B.l[k].begin(), B.l[k].end());
}

// per XCD, the two lists interleaved in proportion to their lengths, so both
// halves of a paired launch sweep their tiles at the same pace
Lists interleave(const Lists& A, const Lists& B) {
  Lists M;
  for (int k = 0; k < 8; k++) {
    const auto &a = A.l[k], &b = B.l[k];
    const size_t na = a.size(), n = na + b.size();
    size_t ia = 0, ib = 0;
    for (size_t j = 0; j < n; j++) {
      const bool takeA = ia < na && (ib >= b.size() || (j + 1) * na / n > ia);
      M.l[k].push_back(takeA ? a[ia++] : b[ib++]);
    }
  }
  return M;
}

std::vector<unsigned> flatten(const Lists& L) {
  const size_t n = L.longest();
  std::vector<unsigned> m(8 * n, wg_entry(kNop, 0, 0));
  for (int k = 0; k < 8; k++)
    for (size_t j = 0; j < L.l[k].size(); j++) m[8 * j + k] = L.l[k][j];
  return m;
}

}  // namespace

// A launch bracketed by timing events.
struct TimedLaunch {
  std::string name;
  hipEvent_t a, b;
  bool pooled = true;  // events go back to the pool (not those of a captured graph)
};

// A shot group's time loop captured as a graph, and its timed launches.
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  std::vector<TimedLaunch> recs;
};

// Device-resident workgroup maps, part tables and captured time loops of a
// job, kept across runs.
struct DevMap {
  unsigned* ptr = nullptr;
  int n = 0;  // workgroups of the launch
};

// Bump allocator over a fixed device region (the scratch array of the
// allocation-free entry point): 256-byte aligned pieces, never returned
// before the region is dropped as a whole.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  Status take(void** p, size_t nb) {
    const size_t at = (used + 255) & ~(size_t)255;
    const size_t n = nb ? nb : 4;
    if (at > cap || n > cap - at)
      return Status::Err(FWIFLOW_ERR_MEMORY,
                         "device scratch exhausted: " + std::to_string(n) + " more bytes needed, " +
                             std::to_string(at) + " of the " + std::to_string(cap) +
                             " scratch bytes (FWIFLOW_SCRATCH_BYTES) in use");
    *p = base + at;
    used = at + n;
    return Status::Ok();
  }
  size_t free_bytes() const {
    const size_t at = (used + 255) & ~(size_t)255;
    return at < cap ? cap - at : 0;
  }
};

struct JobCache {
  int dev = 0;
  Arena* arena = nullptr;  // device memory comes from here instead of hipMalloc when set
  std::map<std::string, DevMap> maps;
  Part* table = nullptr;  // part tables of the multi-part launches of the last run
  size_t table_cap = 0;
  std::map<std::string, GraphEntry> graphs;
  ~JobCache() {
    (void)hipSetDevice(dev);
    clear_graphs();
    if (arena) return;
    for (auto& kv : maps) (void)hipFree(kv.second.ptr);
    if (table) (void)hipFree(table);
  }
  Status device_alloc(void** p, size_t nb) {
    if (arena) return arena->take(p, nb);
    HIP_TRY(hipMalloc(p, nb));
    return Status::Ok();
  }
  void clear_graphs() {
    for (auto& kv : graphs) {
      (void)hipGraphExecDestroy(kv.second.exec);
      for (auto& r : kv.second.recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
      }
    }
    graphs.clear();
  }
  Status map(const std::string& key, const Lists& L, DevMap& out) {
    auto it = maps.find(key);
    if (it != maps.end()) {
      out = it->second;
      return Status::Ok();
    }
    const std::vector<unsigned> m = flatten(L);
    DevMap dm;
    dm.n = (int)m.size();
    TRY(device_alloc((void**)&dm.ptr, std::max<size_t>(1, m.size()) * sizeof(unsigned)));
    if (!m.empty())
      HIP_TRY(hipMemcpy(dm.ptr, m.data(), m.size() * sizeof(unsigned), hipMemcpyHostToDevice));
    maps[key] = dm;
    out = dm;
    return Status::Ok();
  }
  Status upload_table(const std::vector<Part>& parts) {
    if (parts.size() > table_cap) {
      if (table && !arena) HIP_TRY(hipFree(table));
      table = nullptr;
      table_cap = 0;
      TRY(device_alloc((void**)&table, parts.size() * sizeof(Part)));
      table_cap = parts.size();
    }
    if (!parts.empty())
      HIP_TRY(hipMemcpy(table, parts.data(), parts.size() * sizeof(Part), hipMemcpyHostToDevice));
    return Status::Ok();
  }
};

// ===================================================================== device state
struct DeviceState {
  int dev = 0;
  hipStream_t stream = nullptr;
  Geo g{};  // whole plan: ns = plan shots, s0 = 0
  int R = 0;
  std::vector<int> roff;  // first receiver of each shot (+ total)
  // engine settings (environment variables read at plan creation)
  int tile_x = 4;          // x extent of a workgroup tile (blockDim.y)
  bool xcd = true;         // XCD-banded workgroup order
  bool balance = true;     // C-PML columns shared evenly by the XCDs (update-region launches)
  bool mix = true;         // backward pair halves interleaved (pair a)
  bool mix_b = true;       // ... (pair b)
  bool pair_a = true;      // reverse stress + adjoint velocity in one launch
  bool pair_b = true;      // adjoint stress + reverse velocity in one launch
  int unroll = 1;          // shots whose loads a backward thread issues together
  int unroll_fwd = 2;      // ... a forward thread
  int shot_group = -1;     // shots per forward launch (0: all, -1: automatic)
  int shot_group_bwd = -1; // shots per backward launch (0: all, -1: automatic)
  bool fold_inject = true; // residual injection inside the adjoint stress update
  bool part_table = false; // single-part launches also read their part from device memory
  bool graphs = false;     // shot groups' time loops captured and replayed as graphs
                           // (measured no faster at 608x1780; timed runs stay eager)
  int ckpt_every = 0;      // 0: boundary ring every step; k > 0: checkpoints every k steps
  int adjoint = 0;         // 0: the reference's adjoint; 1: the exact adjoint
  double ring_budget = 32e9;  // device bytes a gradient run may hold for ring + checkpoints
  // model + tables
  float *lam = nullptr, *mu = nullptr, *den = nullptr, *amu = nullptr, *ba = nullptr,
        *bb = nullptr;
  float* pml = nullptr;   // 6*nzA + 6*nx
  float* win2 = nullptr;  // nSteps
  int* src = nullptr;     // [ns][2]
  float* srcamp = nullptr;  // [ns][nSteps] windowed source rows
  float* srcwin = nullptr;  // [nSteps] the source rows' window (exact adjoint's stf gradient)
  int* rec = nullptr;     // [R][3] (z, x, plan shot)
  int* rec_at = nullptr;  // [ns][nx] receiver of each column (one-row surveys)
  int rec_row = -1;
  // residual injection cells (surveys the folded injection does not cover)
  int *icell = nullptr, *iptr = nullptr, *irid = nullptr;
  std::vector<int> coff;  // first injection cell of each shot (+ total)
  // wavefields
  float *fwd = nullptr, *mem = nullptr, *adj = nullptr, *madj = nullptr;
  float* ring = nullptr;
  size_t ring_floats = 0;
  float* xt = nullptr;  // exact adjoint: operands [backward group shot][8][cell]
  size_t xt_floats = 0;
  float* snap = nullptr;
  size_t snap_floats = 0;
  float *syn = nullptr, *obs = nullptr, *res = nullptr, *stage = nullptr;
  float *acc = nullptr, *grad = nullptr, *stf_grad = nullptr, *illum = nullptr;
  double* partial = nullptr;
  int npartial = 0;
  double* misfit = nullptr;
  bool model_set = false, stf_set = false, obs_set = false;
  bool illum_on = false, illum_valid = false;
  double misfit_host = 0.0;
  // residual conditioning, shot by shot
  bool cond = false;
  FftDesc fft{};
  int* fperm = nullptr;
  float2* ftw = nullptr;
  float2* bpH = nullptr;
  float* wpad = nullptr;
  float *wst = nullptr, *wen = nullptr, *wwt = nullptr;
  float *cobs = nullptr, *csyn = nullptr;
  float2 *spo = nullptr, *spc = nullptr;
  float2* coef = nullptr;
  float* srcout = nullptr;
  double* cpart = nullptr;
  int nrec_max = 0;
  bool src_updated = false;
  std::vector<float> keep_obs, keep_syn;
  // last gradient run's memory plan and shot groups
  int last_k = 0, last_rows_sg = 0, last_groups_f = 0, last_groups_b = 0;
  long long last_ring_bytes = 0, last_snap_bytes = 0;
  std::unique_ptr<JobCache> cache;  // maps of this plan's own runs
  std::vector<void*> allocs;
  long long bytes = 0;
  // allocation-free plans: buffers carved from the device scratch array, the
  // stream shared by all such plans of the device
  Arena arena_store;
  Arena* arena = nullptr;
  bool own_stream = true;
  // timing
  bool timing = false;
  using Rec = TimedLaunch;
  std::vector<Rec>* sink = nullptr;  // while a graph is captured: its sampled launches
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long long>> totals;

  ~DeviceState() {
    if (stream) (void)hipSetDevice(dev);
    cache.reset();
    for (auto& r : pending) {
      if (!r.pooled) continue;  // owned by a captured graph
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  Status alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t nb = count * sizeof(T);
    if (arena) {
      TRY(arena->take(&q, nb));
      bytes += (long long)nb;
      *p = (T*)q;
      return Status::Ok();
    }
    hipError_t e = hipMalloc(&q, nb ? nb : 4);
    if (e != hipSuccess)
      return Status::Err(e == hipErrorOutOfMemory ? FWIFLOW_ERR_MEMORY : FWIFLOW_ERR_HIP,
                         "hipMalloc(" + std::to_string(nb) + " bytes): " + hipGetErrorString(e));
    allocs.push_back(q);
    bytes += (long long)nb;
    *p = (T*)q;
    return Status::Ok();
  }
  template <typename T>
  void release(T*& p, size_t count) {
    if (!p) return;
    if (arena) {  // arena pieces are dropped with the plan
      bytes -= (long long)(count * sizeof(T));
      p = nullptr;
      return;
    }
    auto it = std::find(allocs.begin(), allocs.end(), (void*)p);
    if (it != allocs.end()) allocs.erase(it);
    (void)hipFree(p);
    bytes -= (long long)(count * sizeof(T));
    p = nullptr;
  }

  hipEvent_t take_event() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  // a sampled step launch passes these to hipExtLaunchKernel, which stamps
  // them when the kernel itself starts and ends (no launch gap inside)
  hipEvent_t kernel_start = nullptr, kernel_end = nullptr;
  template <typename F>
  void launch(const char* name, bool sample, F&& f) {
    if (timing && sample && !sink) {
      Rec r{name, take_event(), take_event(), true};
      kernel_start = r.a;
      kernel_end = r.b;
      f();
      kernel_start = kernel_end = nullptr;
      pending.push_back(r);
    } else if (timing && sample) {
      Rec r{name, take_event(), take_event(), false};
      (void)hipEventRecord(r.a, stream);
      f();
      (void)hipEventRecord(r.b, stream);
      sink->push_back(r);
    } else {
      f();
    }
  }
  void collect_timing() {
    for (auto& r : pending) {
      float ms = 0.f;
      (void)hipEventSynchronize(r.b);
      (void)hipEventElapsedTime(&ms, r.a, r.b);
      auto& t = totals[r.name];
      t.first += ms;
      t.second += 1;
      if (r.pooled) {
        pool.push_back(r.a);
        pool.push_back(r.b);
      }
    }
    pending.clear();
  }

  Pml pml_ptrs() const {
    const int nzA = g.nzA, nx = g.nx;
    const float* z = pml;
    const float* x = pml + 6 * nzA;
    return Pml{z, z + nzA, z + 2 * nzA, z + 3 * nzA, z + 4 * nzA, z + 5 * nzA,
               x, x + nx,  x + 2 * nx,  x + 3 * nx,  x + 4 * nx,  x + 5 * nx};
  }
  Mdl mdl() const { return Mdl{lam, mu, amu, ba, bb}; }
  size_t F() const { return (size_t)g.ns * g.N; }
  // first cell of shot s0 in a [shot][comps][cell] array
  float* shot(float* base, int comps, int s0) const { return base + (size_t)s0 * comps * g.N; }
  int gzU() const { return (g.nzA + kBZ - 1) / kBZ; }
  int gxU() const { return (g.nx + tile_x - 1) / tile_x; }
  int gzR() const { return (g.ringH + kBZ - 1) / kBZ; }
  int gxR() const { return (g.ringW + tile_x - 1) / tile_x; }
  // ~40 B per cell and shot of forward state stay cache-resident from step to
  // step; the backward holds forward and adjoint state (~64 B), and the exact
  // adjoint its two operand sets besides (+32 B).  224 MB of the 256 MB
  // Infinity Cache for them (measured at 608x1780: 5 shots per forward
  // launch, 3 per backward launch; 2 for the exact adjoint, 4.87
  // shot-gradients/s against 4.37 with 3).
  double fwd_bytes_per_shot() const { return 40.0 * (double)g.N; }
  double bwd_bytes_per_shot() const { return (adjoint ? 96.0 : 64.0) * (double)g.N; }
  size_t ring_step_floats() const { return (size_t)5 * g.ringlen; }
  size_t snap_floats_per_shot() const { return (size_t)13 * g.N; }
  // zero shots [s0, s0+n) of a [shot][comps][cell] array
  // on stream st (the stream of the run, which may be another plan's)
  Status zero_shots(float* base, int comps, int s0, int n, hipStream_t st) {
    HIP_TRY(hipMemsetAsync(shot(base, comps, s0), 0, (size_t)n * comps * g.N * sizeof(float), st));
    return Status::Ok();
  }
};

// Address of the scratch array on each device and the stream of the
// allocation-free plans there, looked up once per device with that device
// current and kept until the process ends.
namespace {

struct ScratchDevice {
  bool ready = false;
  char* ptr = nullptr;
  hipStream_t stream = nullptr;
};
std::mutex g_scratch_mu;
std::vector<ScratchDevice> g_scratch_dev;

Status scratch_device(int dev, char** base, hipStream_t* stream) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1);
  ScratchDevice& sd = g_scratch_dev[dev];
  if (!sd.ready) {
    HIP_TRY(hipSetDevice(dev));
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    HIP_TRY(hipStreamCreate(&sd.stream));
    sd.ptr = (char*)ptr;
    sd.ready = true;
  }
  *base = sd.ptr;
  *stream = sd.stream;
  return Status::Ok();
}

}  // namespace

Plan::Plan() = default;
Plan::~Plan() = default;

Status Plan::create(const std::string& para_fname, int gpu, int group_size, const int* shot_ids,
                    bool in_scratch) {
  TRY(parse_para(para_fname, para));
  TRY(parse_survey(para, group_size, shot_ids, survey));
  gpu_id = gpu;
  if (!spacing_ok(para.dz) || !spacing_ok(para.dx))
    return Status::Err(FWIFLOW_ERR_UNSUPPORTED,
                       "grid spacing: the reciprocal division of the stencils is not correctly "
                       "rounded for dz = " + std::to_string(para.dz) + " or dx = " +
                           std::to_string(para.dx));
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (gpu < 0 || gpu >= ndev)
    return Status::Err(FWIFLOW_ERR_HIP, "gpu_id " + std::to_string(gpu) + " not present (" +
                                            std::to_string(ndev) + " devices)");
  d_.reset(new DeviceState());
  DeviceState& d = *d_;
  d.dev = gpu;
  d.tile_x = std::max(1, std::min(4, env_int("FWIFLOW_TILE_X", 4)));  // <= 256 threads
  d.xcd = env_int("FWIFLOW_XCD_REMAP", 1) != 0;
  d.balance = env_int("FWIFLOW_XCD_BALANCE", 1) != 0 && d.xcd;
  d.mix = env_int("FWIFLOW_PAIR_MIX", 1) != 0;
  d.mix_b = env_int("FWIFLOW_PAIR_MIX_B", 1) != 0;
  d.pair_a = env_int("FWIFLOW_PAIR_A", 1) != 0;
  d.pair_b = env_int("FWIFLOW_PAIR_B", 1) != 0;
  d.unroll = env_int("FWIFLOW_SHOT_UNROLL", 1) >= 2 ? 2 : 1;
  d.unroll_fwd = env_int("FWIFLOW_SHOT_UNROLL_FWD", 2) >= 2 ? 2 : 1;
  d.shot_group = std::max(-1, env_int("FWIFLOW_SHOT_GROUP", -1));
  d.shot_group_bwd = std::max(-1, env_int("FWIFLOW_SHOT_GROUP_BWD", -1));
  d.fold_inject = env_int("FWIFLOW_FOLD_INJECT", 1) != 0;
  d.part_table = env_int("FWIFLOW_PART_TABLE", 0) != 0;
  d.graphs = env_int("FWIFLOW_GRAPHS", 0) != 0;
  d.adjoint = env_int("FWIFLOW_EXACT_ADJOINT", 0) != 0 ? 1 : 0;
  if (in_scratch) {
    // every buffer from the device's scratch array, no graphs (instantiating
    // one allocates), the device's shared stream
    char* base = nullptr;
    TRY(scratch_device(gpu, &base, &d.stream));
    d.own_stream = false;
    d.arena_store.base = base;
    d.arena_store.cap = FWIFLOW_SCRATCH_BYTES;
    d.arena = &d.arena_store;
    d.graphs = false;
  }
  HIP_TRY(hipSetDevice(gpu));
  if (!d.stream) HIP_TRY(hipStreamCreate(&d.stream));
  Geo& g = d.g;
  g.nz = para.nz;
  g.nx = para.nx;
  g.nzA = para.nzA();
  g.nPml = para.nPml;
  g.ns = group_size;
  g.s0 = 0;
  g.nSteps = para.nSteps;
  g.N = (long long)para.nz * para.nx;
  g.dt = para.dt;
  g.dz = para.dz;
  g.dx = para.dx;
  g.rdz = 1.f / para.dz;
  g.rdx = 1.f / para.dx;
  g.ringH = g.nzA - 2 * g.nPml + 4;
  g.ringW = g.nx - 2 * g.nPml + 4;
  g.ringlen = 10 * g.ringW + 10 * (g.ringH - 10);
  d.R = survey.total_rec();
  g.R = d.R;
  g.rs0 = 0;
  g.rnt = std::max(1, g.nSteps - 1);

  const size_t N = (size_t)g.N;
  TRY(d.alloc(&d.lam, N));
  TRY(d.alloc(&d.mu, N));
  TRY(d.alloc(&d.den, N));
  TRY(d.alloc(&d.amu, N));
  TRY(d.alloc(&d.ba, N));
  TRY(d.alloc(&d.bb, N));
  std::vector<float> tab(6 * (size_t)g.nzA + 6 * (size_t)g.nx);
  cpml_profiles(g.nzA, g.nPml, para.dz, para.f0, para.dt, tab.data());
  cpml_profiles(g.nx, g.nPml, para.dx, para.f0, para.dt, tab.data() + 6 * g.nzA);
  TRY(d.alloc(&d.pml, tab.size()));
  HIP_TRY(hipMemcpy(d.pml, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> w2(g.nSteps);
  window_squared(g.nSteps, para.dt, 0.005f, w2.data());
  TRY(d.alloc(&d.win2, w2.size()));
  HIP_TRY(hipMemcpy(d.win2, w2.data(), w2.size() * sizeof(float), hipMemcpyHostToDevice));
  // the source rows' window (Src_Rec.cu:132-143): d(source amplitude)/d(stf)
  window_squared(g.nSteps, para.dt, 0.001f, w2.data());
  TRY(d.alloc(&d.srcwin, w2.size()));
  HIP_TRY(hipMemcpy(d.srcwin, w2.data(), w2.size() * sizeof(float), hipMemcpyHostToDevice));
  std::vector<int> src(2 * g.ns), rec(3 * (size_t)d.R);
  size_t r = 0;
  d.roff.assign(g.ns + 1, 0);
  for (int s = 0; s < g.ns; s++) {
    const Shot& sh = survey.shots[s];
    d.roff[s] = (int)r;
    src[2 * s] = sh.z_src;
    src[2 * s + 1] = sh.x_src;
    for (size_t k = 0; k < sh.z_rec.size(); k++, r++) {
      rec[3 * r] = sh.z_rec[k];
      rec[3 * r + 1] = sh.x_rec[k];
      rec[3 * r + 2] = s;
    }
  }
  d.roff[g.ns] = (int)r;
  {
    // one-row surveys with at most one receiver per cell get the injection
    // inside the adjoint stress update
    bool one_row = d.R > 0;
    const int row = d.R > 0 ? rec[0] : -1;
    std::vector<int> at((size_t)g.ns * g.nx, -1);
    for (int k = 0; k < d.R && one_row; k++) {
      int& slot = at[(size_t)rec[3 * k + 2] * g.nx + rec[3 * k + 1]];
      if (rec[3 * k] != row || slot >= 0) one_row = false;
      slot = k;
    }
    if (one_row && d.fold_inject) {
      TRY(d.alloc(&d.rec_at, at.size()));
      HIP_TRY(hipMemcpy(d.rec_at, at.data(), at.size() * sizeof(int), hipMemcpyHostToDevice));
      d.rec_row = row;
    }
  }
  {
    // injection cells of each shot, receivers of a cell in receiver order
    std::vector<int> cells, ptr{0}, rid;
    d.coff.assign(g.ns + 1, 0);
    for (int s = 0; s < g.ns; s++) {
      d.coff[s] = (int)cells.size() / 3;
      std::map<std::pair<int, int>, std::vector<int>> by;  // (x, z) -> receivers
      for (int k = d.roff[s]; k < d.roff[s + 1]; k++)
        by[{rec[3 * k + 1], rec[3 * k]}].push_back(k);
      for (auto& kv : by) {
        cells.insert(cells.end(), {s, kv.first.second, kv.first.first});
        rid.insert(rid.end(), kv.second.begin(), kv.second.end());
        ptr.push_back((int)rid.size());
      }
    }
    d.coff[g.ns] = (int)cells.size() / 3;
    TRY(d.alloc(&d.icell, std::max<size_t>(1, cells.size())));
    TRY(d.alloc(&d.iptr, ptr.size()));
    TRY(d.alloc(&d.irid, std::max<size_t>(1, rid.size())));
    if (!cells.empty())
      HIP_TRY(hipMemcpy(d.icell, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.iptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!rid.empty())
      HIP_TRY(hipMemcpy(d.irid, rid.data(), rid.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  TRY(d.alloc(&d.src, src.size()));
  HIP_TRY(hipMemcpy(d.src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
  TRY(d.alloc(&d.rec, std::max<size_t>(3, rec.size())));
  if (!rec.empty())
    HIP_TRY(hipMemcpy(d.rec, rec.data(), rec.size() * sizeof(int), hipMemcpyHostToDevice));
  TRY(d.alloc(&d.srcamp, (size_t)g.ns * g.nSteps));
  const size_t F = (size_t)g.ns * N;
  TRY(d.alloc(&d.fwd, 5 * F));
  TRY(d.alloc(&d.mem, 8 * F));
  d.cond = para.conditioned() || para.save_scratch;
  if (d.cond) {
    const int nt = g.nSteps, L = 2 * nt;
    for (int s = 0; s < g.ns; s++) d.nrec_max = std::max(d.nrec_max, d.roff[s + 1] - d.roff[s]);
    TRY(d.alloc(&d.cobs, (size_t)nt * d.nrec_max));
    TRY(d.alloc(&d.csyn, (size_t)nt * d.nrec_max));
    TRY(d.alloc(&d.cpart, (size_t)g.ns * kCondBlocks));
    if (para.if_win) {
      std::vector<float> a, b, c;
      for (const Shot& sh : survey.shots) {
        a.insert(a.end(), sh.win_start.begin(), sh.win_start.end());
        b.insert(b.end(), sh.win_end.begin(), sh.win_end.end());
        c.insert(c.end(), sh.weights.begin(), sh.weights.end());
      }
      TRY(d.alloc(&d.wst, a.size()));
      TRY(d.alloc(&d.wen, b.size()));
      TRY(d.alloc(&d.wwt, c.size()));
      HIP_TRY(hipMemcpy(d.wst, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d.wen, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d.wwt, c.data(), c.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (para.if_filter || para.if_src_update) {
      std::vector<int> perm;
      std::vector<float2> tw;
      TRY(fft_plan(L, d.fft, perm, tw));
      TRY(d.alloc(&d.fperm, perm.size()));
      TRY(d.alloc(&d.ftw, tw.size()));
      HIP_TRY(hipMemcpy(d.fperm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d.ftw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    if (para.if_filter) {
      std::vector<float2> H;
      bandpass_table(nt, para.dt, para.filter, H);
      TRY(d.alloc(&d.bpH, H.size()));
      HIP_TRY(hipMemcpy(d.bpH, H.data(), H.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    if (para.if_src_update) {
      std::vector<float> wp(L);
      window_squared(L, para.dt, 0.01f, wp.data());
      TRY(d.alloc(&d.wpad, wp.size()));
      HIP_TRY(hipMemcpy(d.wpad, wp.data(), wp.size() * sizeof(float), hipMemcpyHostToDevice));
      TRY(d.alloc(&d.spo, (size_t)d.nrec_max * (nt + 1)));
      TRY(d.alloc(&d.spc, (size_t)d.nrec_max * (nt + 1)));
      TRY(d.alloc(&d.coef, (size_t)nt + 1));
      TRY(d.alloc(&d.srcout, (size_t)g.ns * nt));
    }
  }
  TRY(d.alloc(&d.syn, std::max<size_t>(1, (size_t)g.nSteps * d.R)));
  TRY(d.alloc(&d.misfit, 1));
  return Status::Ok();
}

Status Plan::set_model(const double* Lambda, const double* Mu, const double* Den, bool zfast) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (!Lambda || !Mu || !Den) return Status::Err(FWIFLOW_ERR_ARG, "null model pointer");
  std::vector<float> lam, mu, den;
  model_to_zfast(para, Lambda, Mu, Den, zfast, lam, mu, den, cp_max);
  TRY(check_courant(para, cp_max));
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  const size_t nb = (size_t)d.g.N * sizeof(float);
  HIP_TRY(hipMemcpyAsync(d.lam, lam.data(), nb, hipMemcpyHostToDevice, d.stream));
  HIP_TRY(hipMemcpyAsync(d.mu, mu.data(), nb, hipMemcpyHostToDevice, d.stream));
  HIP_TRY(hipMemcpyAsync(d.den, den.data(), nb, hipMemcpyHostToDevice, d.stream));
  hipLaunchKernelGGL(k_model_init, dim3((d.g.nz + 255) / 256, d.g.nx), dim3(256), 0, d.stream,
                     d.g, d.mu, d.den, d.amu, d.ba, d.bb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(d.stream));
  d.model_set = true;
  return Status::Ok();
}

Status Plan::set_stf(const double* stf) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (!stf) return Status::Err(FWIFLOW_ERR_ARG, "null stf pointer");
  DeviceState& d = *d_;
  const int nSteps = d.g.nSteps;
  // Src_Rec.cu:132-143: row shot_id -> float -> window ratio 0.001
  std::vector<float> w2(nSteps), amps((size_t)d.g.ns * nSteps);
  window_squared(nSteps, para.dt, 0.001f, w2.data());
  for (int s = 0; s < d.g.ns; s++) {
    const double* row = stf + (size_t)survey.shots[s].id * nSteps;
    for (int it = 0; it < nSteps; it++) amps[(size_t)s * nSteps + it] = (float)row[it] * w2[it];
  }
  HIP_TRY(hipSetDevice(d.dev));
  HIP_TRY(hipMemcpy(d.srcamp, amps.data(), amps.size() * sizeof(float), hipMemcpyHostToDevice));
  d.stf_set = true;
  return Status::Ok();
}

static Status ensure_obs_buffers(DeviceState& d) {
  const size_t T = std::max<size_t>(1, (size_t)d.g.nSteps * d.R);
  if (!d.obs) TRY(d.alloc(&d.obs, T));
  if (!d.res) TRY(d.alloc(&d.res, T));
  if (!d.stage) TRY(d.alloc(&d.stage, T));
  return Status::Ok();
}

Status Plan::set_obs(const float* obs) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (!obs) return Status::Err(FWIFLOW_ERR_ARG, "null obs pointer");
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  TRY(ensure_obs_buffers(d));
  const size_t T = (size_t)d.g.nSteps * d.R;
  // host [R][nSteps] (concatenated Shot files) -> device time-major [nSteps][R]
  HIP_TRY(hipMemcpyAsync(d.stage, obs, T * sizeof(float), hipMemcpyHostToDevice, d.stream));
  dim3 blk(32, 8), grd((d.g.nSteps + 31) / 32, (d.R + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grd, blk, 0, d.stream, d.stage, d.obs, d.R, d.g.nSteps);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(d.stream));
  d.obs_set = true;
  return Status::Ok();
}

Status Plan::load_obs() {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  const int nSteps = para.nSteps;
  std::vector<float> host((size_t)nSteps * d_->R);
  size_t off = 0;
  for (const Shot& sh : survey.shots) {
    const std::string fn = para.data_dir + "/Shot" + std::to_string(sh.id) + ".bin";
    FILE* fp = std::fopen(fn.c_str(), "rb");
    if (!fp) return Status::Err(FWIFLOW_ERR_IO, "cannot open \"" + fn + "\"");
    const size_t n = sh.z_rec.size() * (size_t)nSteps;
    const size_t got = std::fread(host.data() + off, sizeof(float), n, fp);
    std::fclose(fp);
    if (got != n)
      return Status::Err(FWIFLOW_ERR_IO, "\"" + fn + "\" holds " + std::to_string(got) +
                                             " floats, expected " + std::to_string(n));
    off += n;
  }
  return set_obs(host.data());
}

Status Plan::set_checkpoint(int k) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (k < 0) return Status::Err(FWIFLOW_ERR_ARG, "checkpoint interval must be >= 0");
  d_->ckpt_every = k;
  return Status::Ok();
}

Status Plan::set_adjoint(int mode) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (mode != 0 && mode != 1) return Status::Err(FWIFLOW_ERR_ARG, "adjoint mode must be 0 or 1");
  d_->adjoint = mode;
  return Status::Ok();
}

int Plan::adjoint() const { return d_ ? d_->adjoint : 0; }

Status Plan::set_memory_budget(double bytes) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (!(bytes > 0)) return Status::Err(FWIFLOW_ERR_ARG, "memory budget must be positive");
  d_->ring_budget = bytes;
  return Status::Ok();
}

Status Plan::set_launch(int tile_x, int shots_fwd, int shots_bwd) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (tile_x < 1 || tile_x > 4) return Status::Err(FWIFLOW_ERR_ARG, "tile width must be 1..4");
  if (shots_fwd < -1 || shots_bwd < -1)
    return Status::Err(FWIFLOW_ERR_ARG, "shots per launch must be >= -1");
  d_->tile_x = tile_x;
  d_->shot_group = shots_fwd;
  d_->shot_group_bwd = shots_bwd;
  if (d_->cache) d_->cache->clear_graphs();  // captured loops bake the old shape in
  return Status::Ok();
}

Status Plan::launch(int* tile_x, int* shots_fwd, int* shots_bwd) const {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (tile_x) *tile_x = d_->tile_x;
  if (shots_fwd) *shots_fwd = d_->shot_group;
  if (shots_bwd) *shots_bwd = d_->shot_group_bwd;
  return Status::Ok();
}

Status Plan::memory_plan(int* k, int* shots_per_sweep, long long* ring_bytes,
                         long long* checkpoint_bytes) const {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if (k) *k = d_->last_k;
  if (shots_per_sweep) *shots_per_sweep = d_->last_rows_sg;
  if (ring_bytes) *ring_bytes = d_->last_ring_bytes;
  if (checkpoint_bytes) *checkpoint_bytes = d_->last_snap_bytes;
  return Status::Ok();
}

Status Plan::set_timing(bool on) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  d_->timing = on;
  d_->totals.clear();
  return Status::Ok();
}

Status Plan::timing(const std::string& name, double* ms, long long* count) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  auto it = d_->totals.find(name);
  if (ms) *ms = it == d_->totals.end() ? 0.0 : it->second.first;
  if (count) *count = it == d_->totals.end() ? 0 : it->second.second;
  return Status::Ok();
}

long long Plan::device_bytes() const { return d_ ? d_->bytes : 0; }

int Plan::shots_per_launch() const {
  if (!d_) return 0;
  return d_->last_groups_f > 0 ? (d_->g.ns + d_->last_groups_f - 1) / d_->last_groups_f : 0;
}
int Plan::shots_per_launch_bwd() const {
  if (!d_) return 0;
  return d_->last_groups_b > 0 ? (d_->g.ns + d_->last_groups_b - 1) / d_->last_groups_b : 0;
}
int Plan::launch_groups(int backward) const {
  if (!d_) return 0;
  return backward ? d_->last_groups_b : d_->last_groups_f;
}

// Residual with the optional conditioning stages for shots [sb, se), one shot
// at a time (libCUFD.cu:254-313): window obs and syn (per-trace windows and
// weights, or the default ratio-0.005 window), band-pass both, source update
// (syn and the source row replaced), res = obs - syn and its misfit, then the
// adjoint stages on res: source-update filter, band-pass, window (ratio 0.1
// with per-trace windows).  `keep` copies the conditioned obs and syn to the
// host for the scratch files.
static Status condition_residual(DeviceState& d, const Para& para, bool keep, int sb, int se,
                                 hipStream_t st) {
  const Geo& g = d.g;
  const int nt = g.nSteps, nf = nt + 1, kc = touched_bins(nt);
  const float scale = 1.0f / (float)(2 * nt);
  const float dt = para.dt;
  if (keep && d.keep_obs.size() != (size_t)d.R * nt) {
    d.keep_obs.assign((size_t)d.R * nt, 0.f);
    d.keep_syn.assign((size_t)d.R * nt, 0.f);
  }
  std::vector<float> h;
  for (int s = sb; s < se; s++) {
    const int r0 = d.roff[s], n = d.roff[s + 1] - r0;
    const float* ws = para.if_win ? d.wst + r0 : nullptr;
    const float* we = para.if_win ? d.wen + r0 : nullptr;
    const float* wt = para.if_win ? d.wwt + r0 : nullptr;
    const TraceSet O{d.cobs, n, d.cobs, n, n}, Y{d.csyn, n, d.csyn, n, n};
    const TraceSet Rs{d.res + r0, d.R, d.res + r0, d.R, n};
    HIP_TRY(window_traces(st, nt, TraceSet{d.obs + r0, d.R, d.cobs, n, n}, d.win2, ws, we, wt,
                          0.005f, dt));
    HIP_TRY(window_traces(st, nt, TraceSet{d.syn + r0, d.R, d.csyn, n, n}, d.win2, ws, we, wt,
                          0.005f, dt));
    if (para.if_filter) {
      HIP_TRY(trace_filter(st, d.fft, d.fperm, d.ftw, nt, O, nullptr, d.bpH, kc, scale));
      HIP_TRY(trace_filter(st, d.fft, d.fperm, d.ftw, nt, Y, nullptr, d.bpH, kc, scale));
    }
    if (para.if_src_update) {
      HIP_TRY(trace_spectra(st, d.fft, d.fperm, d.ftw, nt, O, d.wpad, d.spo));
      HIP_TRY(trace_spectra(st, d.fft, d.fperm, d.ftw, nt, Y, d.wpad, d.spc));
      HIP_TRY(spectrum_ratio(st, d.spo, d.spc, n, nf, d.coef));
      HIP_TRY(trace_filter(st, d.fft, d.fperm, d.ftw, nt, Y, d.wpad, d.coef, nf, scale));
      const size_t so = (size_t)s * nt;
      HIP_TRY(trace_filter(st, d.fft, d.fperm, d.ftw, nt,
                           TraceSet{d.srcamp + so, 1, d.srcout + so, 1, 1}, nullptr, d.coef, nf,
                           scale));
    }
    HIP_TRY(shot_residual(st, nt, n, d.cobs, d.csyn, n, d.res + r0, d.R,
                          d.cpart + (size_t)s * kCondBlocks, kCondBlocks));
    if (para.if_src_update)
      HIP_TRY(trace_filter(st, d.fft, d.fperm, d.ftw, nt, Rs, d.wpad, d.coef, kc, scale));
    if (para.if_filter)
      HIP_TRY(trace_filter(st, d.fft, d.fperm, d.ftw, nt, Rs, nullptr, d.bpH, kc, scale));
    HIP_TRY(window_traces(st, nt, Rs, d.win2, ws, we, wt, 0.1f, dt));
    if (keep) {
      // [nSteps][n] -> [n][nSteps] (the Shot file layout)
      h.resize((size_t)nt * n);
      for (int q = 0; q < 2; q++) {
        HIP_TRY(hipMemcpyAsync(h.data(), q ? d.csyn : d.cobs, h.size() * sizeof(float),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float* dst = (q ? d.keep_syn.data() : d.keep_obs.data()) + (size_t)r0 * nt;
        for (int t = 0; t < nt; t++)
          for (int r = 0; r < n; r++) dst[(size_t)r * nt + t] = h[(size_t)t * n + r];
      }
    }
  }
  return Status::Ok();
}

// ===================================================================== job runner
// One run of calc_id over one or several plans on one device.  The plans'
// shots are lined up in rows (row r = shot r of every plan that has one); a
// launch of a time-step kernel covers a group of rows of every plan, each
// plan one part of the launch's workgroup map.
//
// Gradient runs take the rows in sweeps (super-groups): a sweep runs the
// forward of its rows, their residual and their backward before the next
// sweep starts, so the boundary store holds one sweep's shots, not all.
// The store is either the reference's ring of every step (k = 0) or
// checkpoints every k steps plus the ring of one k-step segment, the segment
// recomputed from its checkpoint in the backward sweep.
namespace {

struct Member {
  Plan* plan;
  DeviceState* d;
};

struct RowRange {
  int r0, r1;
};

std::vector<RowRange> split_even(int a, int b, int per) {
  std::vector<RowRange> v;
  if (b <= a) return v;
  const int n = (b - a + per - 1) / per;
  for (int k = 0; k < n; k++)
    v.push_back({a + (int)((long long)k * (b - a) / n), a + (int)((long long)(k + 1) * (b - a) / n)});
  return v;
}

enum Phase { kPhFwd = 0, kPhRec = 1, kPhBwd = 2 };

// the parts of one group of rows in one phase
struct GroupParts {
  std::vector<Part> parts;
  std::vector<int> member;  // member of each part
  size_t table_off = 0;     // offset of the parts in the job's device table
  std::string key;          // row range, for the map cache
};

struct StepLaunch {
  const Part* table = nullptr;  // device table (multi-part)
  const Part* single = nullptr;
  const unsigned* map = nullptr;
  int nwg = 0;
};

class Runner {
 public:
  Runner(std::vector<Member> m, JobCache& c) : M(std::move(m)), cache(c) {}
  Status run(int calc_id);

 private:
  std::vector<Member> M;
  JobCache& cache;
  DeviceState* d0 = nullptr;
  hipStream_t st = nullptr;
  dim3 blk;
  int nSteps = 0, rows = 0, gf = 1, gb = 1, S = 1, K = 0, nsnap = 0;
  bool if_res = false, withAdj = false, ckpt = false;
  std::vector<Part> table;  // host copy of every multi-part group's parts

  int shots_in(const Member& m, int r) const { return std::min(r, m.d->g.ns); }
  GroupParts parts_of(RowRange rr, int sg0, Phase ph);
  Status maps_of(const GroupParts& gp, const char* kind, DevMap& out);
  StepLaunch step(const GroupParts& gp, const DevMap& m) const {
    StepLaunch s;
    s.map = m.ptr;
    s.nwg = m.n;
    if (gp.parts.size() == 1 && !d0->part_table)
      s.single = &gp.parts[0];
    else
      s.table = cache.table + gp.table_off;
    return s;
  }
  void fwd_stress(const StepLaunch& s, int it, int U);
  void fwd_vel(const StepLaunch& s, int it, int U);
  void bwd_a(const StepLaunch& s, int it, int U);
  void bwd_b(const StepLaunch& s, int it, int U);
  void exact_a(const Part& P, int it, int first);
  void exact_b(const Part& P, int it);
  float* xt_ = nullptr;            // exact adjoint operands of the running group
  const float* srcwin_ = nullptr;  // and the source window
  Status snapshot(const GroupParts& gp, int sg0, int j, bool save);
  // run `body` (launches only, no memsets) as a graph captured on first use
  Status replay(const std::string& key, const std::function<void()>& body);
  std::string graph_key(const char* phase, const GroupParts& gp,
                        std::initializer_list<const DevMap*> maps) const;
};

std::string Runner::graph_key(const char* phase, const GroupParts& gp,
                              std::initializer_list<const DevMap*> maps) const {
  // FNV-1a over everything the captured launches bake in: parts, maps,
  // unroll factors and whether launches are timed
  unsigned long long h = 1469598103934665603ULL;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ULL;
  };
  for (const Part& P : gp.parts) mix(&P, sizeof(Part));
  for (const DevMap* m : maps) mix(m, sizeof(DevMap));
  const int flags[4] = {d0->unroll, d0->unroll_fwd, d0->timing ? 1 : 0, nSteps};
  mix(flags, sizeof(flags));
  if (gp.parts.size() > 1 || d0->part_table) mix(&cache.table, sizeof(cache.table));
  return std::string(phase) + "/" + gp.key + "/" + std::to_string(h);
}

Status Runner::replay(const std::string& key, const std::function<void()>& body) {
  // events recorded inside a captured graph cannot be timed: timed runs are eager
  if (!d0->graphs || d0->timing) {
    body();
    return Status::Ok();
  }
  auto it = cache.graphs.find(key);
  if (it == cache.graphs.end()) {
    if (cache.graphs.size() >= 512) cache.clear_graphs();
    GraphEntry e;
    d0->sink = &e.recs;
    hipError_t err = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (err == hipSuccess) {
      body();
      hipGraph_t g = nullptr;
      err = hipStreamEndCapture(st, &g);
      if (err == hipSuccess) {
        err = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
      }
    }
    d0->sink = nullptr;
    if (err != hipSuccess) {
      for (auto& r : e.recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
      }
      return Status::Err(FWIFLOW_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(err));
    }
    it = cache.graphs.emplace(key, std::move(e)).first;
  }
  HIP_TRY(hipGraphLaunch(it->second.exec, st));
  if (d0->timing)
    for (const TimedLaunch& r : it->second.recs) d0->pending.push_back(r);
  return Status::Ok();
}

// a sampled launch (timing on) goes through hipExtLaunchKernel with the
// plan's kernel_start / kernel_end events
#define STEP_LAUNCH_ONE(K, a)                                                               \
  do {                                                                                      \
    if (d0->kernel_start)                                                                   \
      hipExtLaunchKernelGGL(K, grid, blk, 0, st, d0->kernel_start, d0->kernel_end, 0, a,    \
                            s.map, it);                                                     \
    else                                                                                    \
      hipLaunchKernelGGL(K, grid, blk, 0, st, a, s.map, it);                                \
  } while (0)
#define STEP_LAUNCH(KER)                                                                    \
  do {                                                                                      \
    if (!s.nwg) break;                                                                      \
    const dim3 grid((unsigned)s.nwg);                                                       \
    if (s.single) {                                                                         \
      const Parts<1> a{*s.single};                                                          \
      if (U == 2)                                                                           \
        STEP_LAUNCH_ONE((KER<1, 2>), a);                                                    \
      else                                                                                  \
        STEP_LAUNCH_ONE((KER<1, 1>), a);                                                    \
    } else {                                                                                \
      const Parts<0> a{s.table};                                                            \
      if (U == 2)                                                                           \
        STEP_LAUNCH_ONE((KER<0, 2>), a);                                                    \
      else                                                                                  \
        STEP_LAUNCH_ONE((KER<0, 1>), a);                                                    \
    }                                                                                       \
  } while (0)

void Runner::fwd_stress(const StepLaunch& s, int it, int U) { STEP_LAUNCH(k_fwd_stress); }
void Runner::fwd_vel(const StepLaunch& s, int it, int U) { STEP_LAUNCH(k_fwd_vel); }
void Runner::bwd_a(const StepLaunch& s, int it, int U) { STEP_LAUNCH(k_bwd_a); }
void Runner::bwd_b(const StepLaunch& s, int it, int U) { STEP_LAUNCH(k_bwd_b); }
#undef STEP_LAUNCH
#undef STEP_LAUNCH_ONE

// exact-adjoint launches: 64 x 4 workgroups over the update region
// (with the reversal of the step: the plan's ring is that of a gradient run)
#define EXACT_LAUNCH(K, ...)                                                                \
  do {                                                                                      \
    const int tiles_ = ((P.g.nzA + kBZ - 1) / kBZ) * ((P.g.nx + 3) / 4);                   \
    const dim3 grid((unsigned)(8 * ((tiles_ + 7) / 8)));                                    \
    const dim3 b4(kBZ, 4);                                                                  \
    if (d0->kernel_start)                                                                   \
      hipExtLaunchKernelGGL(K, grid, b4, 0, st, d0->kernel_start, d0->kernel_end, 0, P,     \
                            __VA_ARGS__);                                                   \
    else                                                                                    \
      hipLaunchKernelGGL(K, grid, b4, 0, st, P, __VA_ARGS__);                               \
  } while (0)
void Runner::exact_a(const Part& P, int it, int first) { EXACT_LAUNCH(k_exact_a, xt_, it, first); }
void Runner::exact_b(const Part& P, int it) { EXACT_LAUNCH(k_exact_b, xt_, srcwin_, it); }
#undef EXACT_LAUNCH

GroupParts Runner::parts_of(RowRange rr, int sg0, Phase ph) {
  GroupParts gp;
  gp.key = std::to_string(rr.r0) + ":" + std::to_string(rr.r1);
  for (size_t mi = 0; mi < M.size(); mi++) {
    DeviceState& d = *M[mi].d;
    const int s0 = shots_in(M[mi], rr.r0), s1 = shots_in(M[mi], rr.r1);
    if (s1 <= s0) continue;
    const int n = s1 - s0;
    Part P{};
    P.g = d.g;
    P.g.ns = n;
    P.g.s0 = s0;
    P.g.rs0 = s0 - shots_in(M[mi], sg0);
    P.g.rnt = ckpt ? K : std::max(1, nSteps - 1);
    P.m = d.mdl();
    P.c = d.pml_ptrs();
    P.w = d.shot(d.fwd, kWaveComps, s0);
    P.a = d.adj ? d.shot(d.adj, kWaveComps, s0) : nullptr;
    P.p = d.shot(ph == kPhBwd && ckpt ? d.madj : d.mem, kMemComps, s0);
    const float* amp = (ph == kPhBwd && d.src_updated) ? d.srcout : d.srcamp;
    P.src = Src{d.src + 2 * s0, amp + (size_t)s0 * nSteps};
    P.inj = Inj{d.rec_at && ph == kPhBwd ? d.res : nullptr, d.rec_at, d.rec_row};
    const bool ring = withAdj && ((ph == kPhFwd && !ckpt) || ph == kPhRec || ph == kPhBwd);
    P.ring = ring ? d.ring : nullptr;
    P.acc = d.acc;
    P.stf_grad = d.stf_grad ? d.stf_grad + (size_t)s0 * nSteps : nullptr;
    P.illum = ph == kPhFwd && d.illum_on ? d.illum : nullptr;
    P.rec = d.rec + 3 * (size_t)d.roff[s0];
    P.syn = d.syn + d.roff[s0];
    P.nrec = d.roff[s1] - d.roff[s0];
    P.gzU = d.gzU();
    P.gzR = d.gzR();
    gp.parts.push_back(P);
    gp.member.push_back((int)mi);
    gp.key += "|" + std::to_string(mi) + ":" + std::to_string(s0) + "-" + std::to_string(s1);
  }
  if (gp.parts.size() > 1 || d0->part_table) {
    gp.table_off = table.size();
    table.insert(table.end(), gp.parts.begin(), gp.parts.end());
  }
  return gp;
}

// Workgroup map of one launch kind over the parts of a group:
//   "S" forward stress, "V" forward velocity + recording, "W" forward velocity,
//   "A" reverse stress + adjoint velocity, "B" adjoint stress + reverse velocity,
//   "AV" adjoint velocity, "AS" adjoint stress, "RV" reverse velocity.
Status Runner::maps_of(const GroupParts& gp, const char* kind, DevMap& out) {
  const std::string key = std::string(kind) + "/" + gp.key + "/" + std::to_string(d0->tile_x);
  const std::string k = kind;
  Lists all;
  for (size_t q = 0; q < gp.parts.size(); q++) {
    const DeviceState& d = *M[gp.member[q]].d;
    const int part = (int)q;
    auto upd = [&](unsigned kd) {
      Lists L;
      const int gz = d.gzU(), gx = d.gxU();
      if (!d.xcd)
        dealt(L, kd, part, gz * gx);
      else if (d.balance)
        balanced(L, kd, part, gz, gx, (d.g.nPml + d.tile_x - 1) / d.tile_x,
                 gx - (d.g.nx - d.g.nPml) / d.tile_x);
      else
        band(L, kd, part, gz * gx);
      return L;
    };
    auto rng = [&](unsigned kd) {
      Lists L;
      if (d.xcd)
        band(L, kd, part, d.gzR() * d.gxR());
      else
        dealt(L, kd, part, d.gzR() * d.gxR());
      return L;
    };
    Lists L;
    if (k == "S") {
      L = upd(kStressFwd);
    } else if (k == "V" || k == "W") {
      L = upd(kVelFwd);
      if (k == "V") {
        Lists r;
        const int per = kBZ * d.tile_x;  // receivers per recording workgroup
        dealt(r, kRecord, part, (gp.parts[q].nrec + per - 1) / per);
        append(L, r);
      }
    } else if (k == "A") {
      const Lists a = rng(kRevStress), b = upd(kAdjVel);
      if (d.mix)
        L = interleave(a, b);
      else
        L = a, append(L, b);
    } else if (k == "B") {
      const Lists a = upd(kAdjStress), b = rng(kRevVel);
      if (d.mix_b)
        L = interleave(a, b);
      else
        L = a, append(L, b);
    } else if (k == "R") {
      L = rng(kRevStress);
    } else if (k == "AV") {
      L = upd(kAdjVel);
    } else if (k == "AS") {
      L = upd(kAdjStress);
    } else if (k == "RV") {
      L = rng(kRevVel);
    }
    append(all, L);
  }
  return cache.map(key, all, out);
}

// Checkpoint j >= 1 of the shots of a group: the 5 wavefield components and
// the 8 forward C-PML memory variables before forward step j * K.  Layout
// [j-1] x {wavefields [shot of the sweep][5][cell], memory variables
// [shot of the sweep][8][cell]}: a group's shots are one contiguous block of
// each.
Status Runner::snapshot(const GroupParts& gp, int sg0, int j, bool save) {
  for (size_t q = 0; q < gp.parts.size(); q++) {
    DeviceState& d = *M[gp.member[q]].d;
    const int s0 = gp.parts[q].g.s0, n = gp.parts[q].g.ns;
    const int rows_m = std::min(S, d.g.ns), rs = s0 - shots_in(M[gp.member[q]], sg0);
    const size_t N = (size_t)d.g.N;
    float* cp = d.snap + (size_t)(j - 1) * (kWaveComps + kMemComps) * rows_m * N;
    float* cw = cp + (size_t)rs * kWaveComps * N;
    float* cm = cp + (size_t)rows_m * kWaveComps * N + (size_t)rs * kMemComps * N;
    float* fw = d.shot(d.fwd, kWaveComps, s0);
    float* fm = d.shot(d.mem, kMemComps, s0);
    const size_t bw = (size_t)n * kWaveComps * N * sizeof(float);
    const size_t bm = (size_t)n * kMemComps * N * sizeof(float);
    HIP_TRY(hipMemcpyAsync(save ? cw : fw, save ? fw : cw, bw, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(save ? cm : fm, save ? fm : cm, bm, hipMemcpyDeviceToDevice, st));
  }
  return Status::Ok();
}

Status Runner::run(int calc_id) {
  TRY(check_calc_id(calc_id));
  if (M.empty()) return Status::Err(FWIFLOW_ERR_ARG, "no plans to run");
  d0 = M[0].d;
  for (const Member& m : M) {
    if (!m.d) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
    if (!m.d->model_set) return Status::Err(FWIFLOW_ERR_ARG, "model not set");
    if (!m.d->stf_set) return Status::Err(FWIFLOW_ERR_ARG, "source time functions not set");
    if (calc_id != 2 && !m.d->obs_set) return Status::Err(FWIFLOW_ERR_ARG, "observed data not set");
    if (m.d->dev != d0->dev) return Status::Err(FWIFLOW_ERR_ARG, "plans of a job must share the GPU");
    if (m.d->g.nSteps != d0->g.nSteps)
      return Status::Err(FWIFLOW_ERR_ARG, "plans of a job must share nSteps");
    if (m.d->tile_x != d0->tile_x)
      return Status::Err(FWIFLOW_ERR_ARG, "plans of a job must share the tile width");
    if (m.d->adjoint && M.size() > 1)
      return Status::Err(FWIFLOW_ERR_UNSUPPORTED, "the exact adjoint runs one plan at a time");
    if (m.d->adjoint && calc_id == 1 && m.plan->para.if_src_update)
      return Status::Err(FWIFLOW_ERR_UNSUPPORTED,
                         "the exact adjoint does not take the source update (if_src_update)");
    m.d->illum_valid = false;
    m.plan->last_calc = -1;
  }
  if (M.size() > 1 && M.size() > (size_t)kMaxPartsTable)
    return Status::Err(FWIFLOW_ERR_ARG, "at most " + std::to_string(kMaxPartsTable) + " plans per job");
  HIP_TRY(hipSetDevice(d0->dev));
  st = d0->stream;
  blk = dim3(kBZ, d0->tile_x);
  nSteps = d0->g.nSteps;
  if_res = calc_id != 2;
  withAdj = calc_id == 1;
  rows = 0;
  double rowf = 0, rowb = 0, ring_row = 0, snap_row = 0;
  for (const Member& m : M) {
    rows = std::max(rows, m.d->g.ns);
    rowf += m.d->fwd_bytes_per_shot();
    rowb += m.d->bwd_bytes_per_shot();
    ring_row += 4.0 * m.d->ring_step_floats();
    snap_row += 4.0 * m.d->snap_floats_per_shot();
  }
  // rows per launch: forward and backward state of a group kept in the
  // Infinity Cache from step to step
  const DeviceState& c0 = *d0;
  auto autosize = [&](double per_row) {
    return std::max(1, std::min(rows, (int)std::lround(224e6 / per_row)));
  };
  gf = c0.shot_group == 0 ? rows : c0.shot_group > 0 ? std::min(c0.shot_group, rows) : autosize(rowf);
  if (c0.shot_group_bwd == 0)
    gb = rows;
  else if (c0.shot_group_bwd > 0)
    gb = std::min(c0.shot_group_bwd, rows);
  else if (c0.shot_group >= 0)
    gb = gf;
  else
    gb = autosize(rowb);
  // the kernels address a part's arrays with 32-bit byte offsets: one part's
  // C-PML memory variables (8 arrays per shot) must stay below 2^31 bytes
  int cap = rows;
  for (const Member& m : M)
    cap = std::min(cap, (int)(2147483647LL / ((long long)kMemComps * 4 * m.d->g.N)));
  if (cap < 1) return Status::Err(FWIFLOW_ERR_SHAPE, "grid too large for one shot per launch");
  gf = std::min(gf, cap);
  gb = std::min(gb, cap);
  // memory plan of a gradient run: rows per sweep S and checkpoint interval K
  const int steps = std::max(1, nSteps - 1);  // forward steps 0 .. nSteps-2
  double budget = c0.ring_budget;
  if (withAdj && c0.arena && M.size() == 1) {
    // scratch-backed plan: ring + checkpoints get what the scratch array has
    // left after the gradient run's other buffers (adjoint fields, C-PML
    // adjoint memory variables, accumulators, misfit partials, maps)
    const DeviceState& d = c0;
    const double N = (double)d.g.N, F = (double)d.F();
    const double other = 4.0 * ((d.adj ? 0 : 5 * F) + (d.madj ? 0 : 8 * F) +
                                (d.acc ? 0 : 5 * N) + (d.grad ? 0 : 3 * N) +
                                (d.stf_grad ? 0 : (double)d.g.ns * nSteps) +
                                (d.illum_on && !d.illum ? N : 0)) +
                         8.0 * rows * kPartials + (1 << 20);
    budget = std::min(budget, std::max(1.0, (double)d.arena->free_bytes() - other));
  }
  K = withAdj ? c0.ckpt_every : 0;
  if (K >= steps) K = 0;
  S = rows;
  if (withAdj) {
    const int l = gf / std::__gcd(gf, gb) * gb;
    auto fit = [&](double per_row) {
      // the most rows within the budget, a multiple of both group sizes if
      // possible (no partial groups inside a sweep), else of the backward's
      int n = (int)std::floor(budget / per_row);
      if (n >= l) n = n / l * l;
      else if (n >= gb) n = n / gb * gb;
      return std::max(1, std::min(rows, n));
    };
    const double full = ring_row * steps;
    if (K == 0 && full * std::min(rows, gb) > budget && steps > 1) {
      // the full ring of one backward group does not fit: checkpoints, at the
      // interval that minimises ring + checkpoint memory
      K = (int)std::lround(std::sqrt((double)steps * snap_row / ring_row));
      K = std::max(1, std::min(K, steps - 1));
    }
    ckpt = K > 0;
    if (ckpt) {
      nsnap = (steps + K - 1) / K - 1;
      S = fit(K * ring_row + nsnap * snap_row);
    } else {
      nsnap = 0;
      S = fit(full);
    }
  }
  // sweeps (super-groups) and their launch groups, as row ranges
  std::vector<RowRange> sweeps;
  for (int a = 0; a < rows; a += S) sweeps.push_back({a, std::min(rows, a + S)});
  std::vector<std::vector<RowRange>> FR(sweeps.size()), BR(sweeps.size());
  for (size_t j = 0; j < sweeps.size(); j++) {
    FR[j] = split_even(sweeps[j].r0, sweeps[j].r1, gf);
    if (withAdj) BR[j] = split_even(sweeps[j].r0, sweeps[j].r1, gb);
  }
  // buffers
  for (const Member& m : M) {
    DeviceState& d = *m.d;
    const size_t N = (size_t)d.g.N, F = d.F();
    if (withAdj) {
      if (!d.adj) TRY(d.alloc(&d.adj, 5 * F));
      if (ckpt && !d.madj) TRY(d.alloc(&d.madj, 8 * F));
      if (!d.acc) TRY(d.alloc(&d.acc, 5 * N));
      if (!d.grad) TRY(d.alloc(&d.grad, 3 * N));
      if (!d.stf_grad) TRY(d.alloc(&d.stf_grad, (size_t)d.g.ns * nSteps));
      if (d.adjoint) {
        const size_t xf = (size_t)std::min(gb, d.g.ns) * 8 * N;
        if (d.xt_floats < xf) {
          d.release(d.xt, d.xt_floats);
          d.xt_floats = 0;
          TRY(d.alloc(&d.xt, xf));
          d.xt_floats = xf;
        }
      }
      const int rows_m = std::min(S, d.g.ns);
      const size_t rf = (size_t)rows_m * (ckpt ? K : steps) * d.ring_step_floats();
      if (d.ring_floats < rf || d.ring_floats > 2 * rf + (1u << 20)) {
        d.release(d.ring, d.ring_floats);
        d.ring_floats = 0;
        TRY(d.alloc(&d.ring, rf));
        d.ring_floats = rf;
      }
      const size_t sf = ckpt ? (size_t)nsnap * 13 * rows_m * N : 0;
      if (d.snap_floats < sf || (sf == 0 && d.snap)) {
        d.release(d.snap, d.snap_floats);
        d.snap_floats = 0;
        if (sf) {
          TRY(d.alloc(&d.snap, sf));
          d.snap_floats = sf;
        }
      }
      d.last_k = K;
      d.last_rows_sg = rows_m;
      d.last_ring_bytes = (long long)(rf * sizeof(float));
      d.last_snap_bytes = (long long)(sf * sizeof(float));
    }
    const int npart = (int)sweeps.size() * kPartials;
    if (d.npartial < npart) {
      d.release(d.partial, (size_t)d.npartial);
      d.npartial = 0;
      TRY(d.alloc(&d.partial, (size_t)npart));
      d.npartial = npart;
    }
    if (d.illum_on && !d.illum) TRY(d.alloc(&d.illum, N));
    d.src_updated = if_res && d.cond && m.plan->para.if_src_update;
    d.last_groups_f = 0;
    d.last_groups_b = 0;
    for (size_t j = 0; j < sweeps.size(); j++) {
      for (RowRange rr : FR[j]) d.last_groups_f += shots_in(m, rr.r1) > shots_in(m, rr.r0);
      for (RowRange rr : BR[j]) d.last_groups_b += shots_in(m, rr.r1) > shots_in(m, rr.r0);
    }
  }
  // parts of every group (pointing at the buffers above)
  table.clear();
  std::vector<std::vector<GroupParts>> FG(sweeps.size()), RG(sweeps.size()), BG(sweeps.size());
  for (size_t j = 0; j < sweeps.size(); j++) {
    for (RowRange rr : FR[j]) FG[j].push_back(parts_of(rr, sweeps[j].r0, kPhFwd));
    for (RowRange rr : BR[j]) {
      BG[j].push_back(parts_of(rr, sweeps[j].r0, kPhBwd));
      if (ckpt) RG[j].push_back(parts_of(rr, sweeps[j].r0, kPhRec));
    }
  }
  if (!table.empty()) TRY(cache.upload_table(table));
  // clear the run's state
  for (const Member& m : M) {
    DeviceState& d = *m.d;
    const size_t N = (size_t)d.g.N, F = d.F();
    HIP_TRY(hipMemsetAsync(d.fwd, 0, 5 * F * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(d.mem, 0, 8 * F * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(d.syn, 0, (size_t)nSteps * d.R * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(d.partial, 0, (size_t)d.npartial * sizeof(double), st));
    if (d.illum_on) HIP_TRY(hipMemsetAsync(d.illum, 0, N * sizeof(float), st));
    if (withAdj) {
      HIP_TRY(hipMemsetAsync(d.acc, 0, 5 * N * sizeof(float), st));
      HIP_TRY(hipMemsetAsync(d.stf_grad, 0, (size_t)d.g.ns * nSteps * sizeof(float), st));
    }
  }
  const int Uf = d0->unroll_fwd, Ub = d0->unroll;
  // every 64th step of each launch kind is bracketed by timing events (when
  // timing is on): events between launches cost the stream about 2% at
  // every 8th step
  const int sample_every = 64;
  for (size_t j = 0; j < sweeps.size(); j++) {
    const RowRange sw = sweeps[j];
    // ------------------------------------------------------ forward
    for (const GroupParts& gp : FG[j]) {
      DevMap mS, mV;
      TRY(maps_of(gp, "S", mS));
      TRY(maps_of(gp, "V", mV));
      const StepLaunch sS = step(gp, mS), sV = step(gp, mV);
      if (ckpt) {  // checkpoint copies between the steps: no graph
        for (int it = 0; it <= nSteps - 2; it++) {
          const bool smp = it % sample_every == 0;
          if (it > 0 && it % K == 0) TRY(snapshot(gp, sw.r0, it / K, true));
          d0->launch("stress_fwd", smp, [&] { fwd_stress(sS, it, Uf); });
          d0->launch("vel_fwd", smp, [&] { fwd_vel(sV, it, Uf); });
        }
      } else {
        TRY(replay(graph_key("F", gp, {&mS, &mV}), [&] {
          for (int it = 0; it <= nSteps - 2; it++) {
            const bool smp = it % sample_every == 0;
            d0->launch("stress_fwd", smp, [&] { fwd_stress(sS, it, Uf); });
            d0->launch("vel_fwd", smp, [&] { fwd_vel(sV, it, Uf); });
          }
        }));
      }
    }
    HIP_TRY(hipGetLastError());
    // ------------------------------------------------------ residual
    if (if_res) {
      for (const Member& m : M) {
        DeviceState& d = *m.d;
        const int sa = shots_in(m, sw.r0), sb = shots_in(m, sw.r1);
        if (sb <= sa) continue;
        if (d.cond) {
          TRY(condition_residual(d, m.plan->para, m.plan->para.save_scratch && withAdj, sa, sb,
                                 st));
        } else {
          const int r0 = d.roff[sa], nr = d.roff[sb] - r0;
          hipLaunchKernelGGL(k_residual, dim3(kPartials), dim3(256), 0, st, nSteps, d.R, r0, nr,
                             d.win2, d.obs, d.syn, d.res, d.partial + j * kPartials);
        }
      }
      HIP_TRY(hipGetLastError());
    }
    if (!withAdj) continue;
    // ------------------------------------------------------ backward
    for (size_t q = 0; q < BG[j].size(); q++) {
      const GroupParts& gp = BG[j][q];
      for (size_t p = 0; p < gp.parts.size(); p++) {
        DeviceState& d = *M[gp.member[p]].d;
        const int s0 = gp.parts[p].g.s0, n = gp.parts[p].g.ns;
        TRY(d.zero_shots(d.adj, kWaveComps, s0, n, st));
        TRY(d.zero_shots(ckpt ? d.madj : d.mem, kMemComps, s0, n, st));
      }
      DevMap mA, mB, mAV, mAS, mRV, mRS, mRW, mR;
      TRY(maps_of(gp, "A", mA));
      TRY(maps_of(gp, "R", mR));
      TRY(maps_of(gp, "B", mB));
      TRY(maps_of(gp, "AV", mAV));
      TRY(maps_of(gp, "AS", mAS));
      TRY(maps_of(gp, "RV", mRV));
      const StepLaunch sA = step(gp, mA), sB = step(gp, mB), sAV = step(gp, mAV),
                       sAS = step(gp, mAS), sRV = step(gp, mRV), sR = step(gp, mR);
      StepLaunch sRS, sRW;
      if (ckpt) {
        TRY(maps_of(RG[j][q], "S", mRS));
        TRY(maps_of(RG[j][q], "W", mRW));
        sRS = step(RG[j][q], mRS);
        sRW = step(RG[j][q], mRW);
      }
      // residual injection of step t for plans without the folded injection
      auto inject = [&](int t) {
        for (size_t p = 0; p < gp.parts.size(); p++) {
          DeviceState& d = *M[gp.member[p]].d;
          if (d.rec_at) continue;
          const int s0 = gp.parts[p].g.s0, s1 = s0 + gp.parts[p].g.ns;
          const int c0 = d.coff[s0], nc = d.coff[s1] - c0;
          if (nc > 0)
            hipLaunchKernelGGL(k_res_inject, dim3((nc + 255) / 256), dim3(256), 0, st, d.g,
                               d.adj, d.icell, d.iptr, d.irid, c0, nc,
                               d.res + (size_t)t * d.R);
        }
      };
      // exact adjoint, reverse steps t1-1 .. t0 of a segment: the forward
      // state enters as (v, sigma) after step t1-1 and each reverse step it
      // re-materialises (v^it, sigma^it) in its two launches.  No initial
      // adjoint step: sample 0 is not recorded, and sample it+1 enters at
      // reverse step it.
      const Part* xp = gp.parts.empty() ? nullptr : &gp.parts[0];
      auto xsegment = [&](int t0, int t1) {
        for (int it = t1 - 1; it >= t0; it--) {
          const bool smp = it % sample_every == 0;
          d0->launch("exact_a", smp, [&] { exact_a(*xp, it, it == nSteps - 2 ? 1 : 0); });
          inject(it + 1);
          d0->launch("exact_b", smp, [&] { exact_b(*xp, it); });
        }
      };
      if (d0->adjoint) {
        for (size_t p = 0; p < gp.parts.size(); p++) {
          DeviceState& d = *M[gp.member[p]].d;
          HIP_TRY(hipMemsetAsync(d.xt, 0, (size_t)gp.parts[p].g.ns * 8 * d.g.N * sizeof(float), st));
        }
        xt_ = d0->xt;
        srcwin_ = d0->srcwin;
      }
      if (d0->adjoint && !ckpt) {
        if (nSteps >= 2)
          TRY(replay(graph_key("X", gp, {}) + "/" + std::to_string((uintptr_t)xt_),
                     [&] { xsegment(0, steps); }));
        HIP_TRY(hipGetLastError());
        continue;
      }
      // reverse steps t1-1 .. t0 of one segment; `last` is the segment that
      // ends at nSteps-1
      auto segment = [&](int t0, int t1, bool last) {
        if (last)
          bwd_b(sB, nSteps - 1, Ub);  // adj_stress(nSteps-1) + rev_vel(nSteps-2)
        else
          bwd_b(sRV, t1, Ub);         // rev_vel(t1-1)
        for (int it = t1 - 1; it >= t0; it--) {
          const bool smp = it % sample_every == 0;
          if (d0->pair_a) {
            d0->launch("bwd_pair_a", smp, [&] { bwd_a(sA, it, Ub); });
          } else {  // the halves as launches of their own (same kernel, one kind each)
            d0->launch("rev_stress", smp, [&] { bwd_a(sR, it, Ub); });
            d0->launch("adj_vel", smp, [&] { bwd_a(sAV, it, Ub); });
          }
          inject(it);
          if (it > t0 && d0->pair_b) {
            d0->launch("bwd_pair_b", smp, [&] { bwd_b(sB, it, Ub); });
          } else if (it > t0) {
            d0->launch("adj_stress", smp, [&] { bwd_b(sAS, it, Ub); });
            d0->launch("rev_vel", smp, [&] { bwd_b(sRV, it, Ub); });
          } else {
            bwd_b(sAS, it, Ub);
          }
        }
      };
      if (!ckpt) {
        // the whole backward sweep of the group, from the initial adjoint
        // step at t = nSteps-1 (libCUFD.cu:353-373)
        TRY(replay(graph_key("B", gp, {&mA, &mB, &mAV, &mAS, &mRV, &mR}), [&] {
          bwd_a(sAV, nSteps - 1, Ub);
          inject(nSteps - 1);
          if (nSteps < 2)
            bwd_b(sAS, nSteps - 1, Ub);
          else
            segment(0, steps, true);
        }));
        HIP_TRY(hipGetLastError());
        continue;
      }
      // checkpoints: each segment is recomputed from its checkpoint, then
      // reversed with its ring
      if (d0->adjoint) {
        if (nSteps < 2) continue;
      } else {
        bwd_a(sAV, nSteps - 1, Ub);
        inject(nSteps - 1);
        if (nSteps < 2) {
          bwd_b(sAS, nSteps - 1, Ub);
          continue;
        }
      }
      const int nseg = (steps + K - 1) / K;
      for (int jj = nseg - 1; jj >= 0; jj--) {
        const int t0 = jj * K, t1 = std::min(t0 + K, steps);
        if (jj == 0) {
          for (size_t p = 0; p < gp.parts.size(); p++) {
            DeviceState& d = *M[gp.member[p]].d;
            TRY(d.zero_shots(d.fwd, kWaveComps, gp.parts[p].g.s0, gp.parts[p].g.ns, st));
            TRY(d.zero_shots(d.mem, kMemComps, gp.parts[p].g.s0, gp.parts[p].g.ns, st));
          }
        } else {
          TRY(snapshot(RG[j][q], sw.r0, jj, false));
        }
        for (int it = t0; it < t1; it++) {
          fwd_stress(sRS, it, Uf);
          fwd_vel(sRW, it, Uf);
        }
        if (d0->adjoint)
          xsegment(t0, t1);
        else
          segment(t0, t1, jj == nseg - 1);
      }
      HIP_TRY(hipGetLastError());
    }
  }
  // misfit and gradients
  for (const Member& m : M) {
    DeviceState& d = *m.d;
    if (if_res) {
      if (d.cond)
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, d.cpart,
                           d.g.ns * kCondBlocks, d.misfit);
      else
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, d.partial,
                           (int)sweeps.size() * kPartials, d.misfit);
    }
    if (withAdj)
      hipLaunchKernelGGL(k_finalize, dim3((d.g.nz + 255) / 256, d.g.nx), dim3(256), 0, st, d.g,
                         d.mdl(), d.acc, d.grad, d.adjoint);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  if (d0->timing) d0->collect_timing();
  for (const Member& m : M) {
    DeviceState& d = *m.d;
    if (if_res) HIP_TRY(hipMemcpy(&d.misfit_host, d.misfit, sizeof(double), hipMemcpyDeviceToHost));
    m.plan->last_calc = calc_id;
    d.illum_valid = d.illum_on;
  }
  return Status::Ok();
}

}  // namespace

Status Plan::run(int calc_id) {
  TRY(check_calc_id(calc_id));
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  HIP_TRY(hipSetDevice(d_->dev));
  if (!d_->cache) {
    d_->cache.reset(new JobCache());
    d_->cache->dev = d_->dev;
    d_->cache->arena = d_->arena;
  }
  Runner r({Member{this, d_.get()}}, *d_->cache);
  return r.run(calc_id);
}

Job::Job() = default;
Job::~Job() {
  if (cache_) {
    (void)hipSetDevice(cache_->dev);
    cache_.reset();
  }
}

Status Job::add(Plan* p) {
  if (!p || !p->d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  if ((int)plans_.size() >= kMaxPartsTable)
    return Status::Err(FWIFLOW_ERR_ARG, "at most " + std::to_string(kMaxPartsTable) + " plans per job");
  for (Plan* q : plans_)
    if (q == p) return Status::Err(FWIFLOW_ERR_ARG, "plan already in the job");
  if (p->d_->arena)
    return Status::Err(FWIFLOW_ERR_ARG, "a plan in the device scratch cannot join a job");
  if (!plans_.empty()) {
    const DeviceState& a = *plans_[0]->d_;
    if (p->d_->dev != a.dev) return Status::Err(FWIFLOW_ERR_ARG, "plans of a job must share the GPU");
    if (p->d_->g.nSteps != a.g.nSteps)
      return Status::Err(FWIFLOW_ERR_ARG, "plans of a job must share nSteps");
  }
  plans_.push_back(p);
  cache_.reset();  // maps depend on the members
  return Status::Ok();
}

Status Job::run(int calc_id) {
  if (plans_.empty()) return Status::Err(FWIFLOW_ERR_ARG, "empty job");
  HIP_TRY(hipSetDevice(plans_[0]->d_->dev));
  // Fused launches (every launch covers a group of shots of every plan) or
  // the plans one after another.  The fused groups are sized by the
  // Infinity-Cache rule on the plans' summed bytes per shot, so a fused
  // backward launch holds fewer shots of each plan than that plan's own
  // launches would, and every launch re-reads every plan's model and
  // gradient accumulators for those few shots.  Fusing pays only where the
  // plans are small enough that their own launches would be no larger;
  // FWIFLOW_JOB_FUSE = 1 / 0 forces either.
  int fuse = env_int("FWIFLOW_JOB_FUSE", -1);
  if (fuse < 0 && plans_.size() > 1) {
    auto group = [](double bytes_per_row, int rows) {
      return std::max(1, std::min(rows, (int)std::lround(224e6 / bytes_per_row)));
    };
    double sum = 0;
    int rows = 0, solo = 1 << 30;
    for (Plan* p : plans_) {
      const DeviceState& d = *p->d_;
      sum += d.bwd_bytes_per_shot();
      rows = std::max(rows, d.g.ns);
      solo = std::min(solo, d.shot_group_bwd > 0 ? std::min(d.shot_group_bwd, d.g.ns)
                                                 : group(d.bwd_bytes_per_shot(), d.g.ns));
    }
    const int fused = plans_[0]->d_->shot_group_bwd > 0
                          ? std::min(plans_[0]->d_->shot_group_bwd, rows)
                          : group(sum, rows);
    fuse = fused >= solo ? 1 : 0;
  }
  // the exact adjoint runs one plan at a time
  for (Plan* p : plans_)
    if (p->d_->adjoint) fuse = 0;
  last_fused_ = fuse != 0 || plans_.size() == 1;
  if (!last_fused_) {
    for (Plan* p : plans_) TRY(p->run(calc_id));
    return Status::Ok();
  }
  if (!cache_) {
    cache_.reset(new JobCache());
    cache_->dev = plans_[0]->d_->dev;
  }
  std::vector<Member> m;
  for (Plan* p : plans_) m.push_back(Member{p, p->d_.get()});
  Runner r(std::move(m), *cache_);
  return r.run(calc_id);
}

int Job::size() const { return (int)plans_.size(); }
bool Job::fused() const { return last_fused_; }

Status Plan::misfit(double* out) {
  if (!d_ || (last_calc != 0 && last_calc != 1))
    return Status::Err(FWIFLOW_ERR_ARG, "no calc_id 0 or 1 result");
  if (out) *out = 0.5 * d_->misfit_host;
  return Status::Ok();
}

Status Plan::gradients(double* gL, double* gM, double* gD, double* gS) {
  if (!d_ || last_calc != 1) return Status::Err(FWIFLOW_ERR_ARG, "no calc_id 1 result");
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  const int nz = d.g.nz, nx = d.g.nx;
  const size_t N = (size_t)d.g.N;
  std::vector<float> h(3 * N);
  HIP_TRY(hipMemcpy(h.data(), d.grad, 3 * N * sizeof(float), hipMemcpyDeviceToHost));
  double* outs[3] = {gL, gM, gD};
  for (int k = 0; k < 3; k++) {
    if (!outs[k]) continue;
    const float* src = h.data() + k * N;
    // z-fastest -> row-major [NZ][NX] (libCUFD.cu:480-486)
    for (int i = 0; i < nz; i++)
      for (int j = 0; j < nx; j++) outs[k][(size_t)i * nx + j] = src[(size_t)j * nz + i];
  }
  if (gS) {
    std::vector<float> s((size_t)d.g.ns * d.g.nSteps);
    HIP_TRY(hipMemcpy(s.data(), d.stf_grad, s.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < s.size(); k++) gS[k] = s[k];
  }
  return Status::Ok();
}

Status Plan::set_illumination(bool on) {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  d_->illum_on = on;
  return Status::Ok();
}

Status Plan::illumination(double* out) {
  if (!d_ || !d_->illum_valid)
    return Status::Err(FWIFLOW_ERR_ARG, "no run with the illumination enabled");
  if (!out) return Status::Err(FWIFLOW_ERR_ARG, "null output pointer");
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  const int nz = d.g.nz, nx = d.g.nx;
  std::vector<float> h((size_t)d.g.N);
  HIP_TRY(hipMemcpy(h.data(), d.illum, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < nx; j++) out[(size_t)i * nx + j] = h[(size_t)j * nz + i];
  return Status::Ok();
}

Status Plan::data(float* syn) {
  if (!d_ || last_calc < 0) return Status::Err(FWIFLOW_ERR_ARG, "no run result");
  if (!syn) return Status::Err(FWIFLOW_ERR_ARG, "null output pointer");
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  TRY(ensure_obs_buffers(d));
  const size_t T = (size_t)d.g.nSteps * d.R;
  // time-major [nSteps][R] -> [R][nSteps] (Shot file layout, utilities.cu:557-567)
  dim3 blk(32, 8), grd((d.R + 31) / 32, (d.g.nSteps + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grd, blk, 0, d.stream, d.syn, d.stage, d.g.nSteps, d.R);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(syn, d.stage, T * sizeof(float), hipMemcpyDeviceToHost, d.stream));
  HIP_TRY(hipStreamSynchronize(d.stream));
  return Status::Ok();
}

Status Plan::write_data() {
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  std::vector<float> h((size_t)para.nSteps * d_->R);
  TRY(data(h.data()));
  size_t off = 0;
  for (const Shot& sh : survey.shots) {
    const std::string fn = para.data_dir + "/Shot" + std::to_string(sh.id) + ".bin";
    FILE* fp = std::fopen(fn.c_str(), "wb");
    if (!fp) return Status::Err(FWIFLOW_ERR_IO, "cannot write \"" + fn + "\"");
    const size_t n = sh.z_rec.size() * (size_t)para.nSteps;
    const size_t put = std::fwrite(h.data() + off, sizeof(float), n, fp);
    std::fclose(fp);
    if (put != n) return Status::Err(FWIFLOW_ERR_IO, "short write to \"" + fn + "\"");
    off += n;
  }
  return Status::Ok();
}

Status Plan::residual(float* res) {
  if (!d_ || (last_calc != 0 && last_calc != 1))
    return Status::Err(FWIFLOW_ERR_ARG, "no calc_id 0 or 1 result");
  if (!res) return Status::Err(FWIFLOW_ERR_ARG, "null output pointer");
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  const size_t T = (size_t)d.g.nSteps * d.R;
  dim3 blk(32, 8), grd((d.R + 31) / 32, (d.g.nSteps + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grd, blk, 0, d.stream, d.res, d.stage, d.g.nSteps, d.R);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(res, d.stage, T * sizeof(float), hipMemcpyDeviceToHost, d.stream));
  HIP_TRY(hipStreamSynchronize(d.stream));
  return Status::Ok();
}

Status Plan::source(float* rows) {
  if (!d_ || !d_->stf_set) return Status::Err(FWIFLOW_ERR_ARG, "no source set");
  if (!rows) return Status::Err(FWIFLOW_ERR_ARG, "null output pointer");
  DeviceState& d = *d_;
  HIP_TRY(hipSetDevice(d.dev));
  const float* p = d.src_updated ? d.srcout : d.srcamp;
  HIP_TRY(hipMemcpy(rows, p, (size_t)d.g.ns * d.g.nSteps * sizeof(float), hipMemcpyDeviceToHost));
  return Status::Ok();
}

// libCUFD.cu:493-511: after a gradient run, per shot the residual, the
// conditioned synthetic and observed data and the source row.
Status Plan::write_scratch() {
  if (!d_ || last_calc != 1) return Status::Err(FWIFLOW_ERR_ARG, "no calc_id 1 result");
  DeviceState& d = *d_;
  if (!para.save_scratch) return Status::Ok();
  const int nt = para.nSteps;
  if (d.keep_obs.size() != (size_t)d.R * nt)
    return Status::Err(FWIFLOW_ERR_ARG, "conditioned data of the last run were not kept");
  std::vector<float> res((size_t)d.R * nt), src((size_t)d.g.ns * nt);
  TRY(residual(res.data()));
  TRY(source(src.data()));
  auto put = [&](const std::string& fn, const float* p, size_t n) -> Status {
    FILE* fp = std::fopen(fn.c_str(), "wb");
    if (!fp) return Status::Err(FWIFLOW_ERR_IO, "cannot write \"" + fn + "\"");
    const size_t w = std::fwrite(p, sizeof(float), n, fp);
    std::fclose(fp);
    if (w != n) return Status::Err(FWIFLOW_ERR_IO, "short write to \"" + fn + "\"");
    return Status::Ok();
  };
  for (int s = 0; s < d.g.ns; s++) {
    const std::string id = std::to_string(survey.shots[s].id), dir = para.scratch_dir + "/";
    const size_t off = (size_t)d.roff[s] * nt, n = (size_t)(d.roff[s + 1] - d.roff[s]) * nt;
    TRY(put(dir + "Residual_Shot" + id + ".bin", res.data() + off, n));
    TRY(put(dir + "Syn_Shot" + id + ".bin", d.keep_syn.data() + off, n));
    TRY(put(dir + "CondObs_Shot" + id + ".bin", d.keep_obs.data() + off, n));
    TRY(put(dir + "src_updated" + id + ".bin", src.data() + (size_t)s * nt, (size_t)nt));
  }
  return Status::Ok();
}

}  // namespace fwiflow
