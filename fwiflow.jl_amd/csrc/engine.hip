// gfx950 engine of the elastic FWI operator.
//
// Device data layout (all FP32, z fastest so a wave64 row is 64 consecutive
// depth cells, one 256-byte coalesced access):
//   model      lam, mu, ave_mu, byc_a, byc_b          [NX][NZ]
//   wavefield  vz, vx, szz, sxx, sxz  per shot        [shot][NX][NZ]
//   adjoint    same 5 components                      [shot][NX][NZ]
//   C-PML memory variables (8), per shot              [shot][NX][NZ]
//   boundary ring, 5 layers x 4 sides, 5 components   [shot][step][component][ring]
//   seismograms syn/obs/res, time-major               [nSteps][all receivers]
//   gradient accumulators (5, summed over shots)      [NX][NZ]
// One thread owns one cell and loops over every shot of the group: the model
// parameters and C-PML coefficients are read once per cell per step and the
// gradient contributions of all shots are summed in registers before one
// read-modify-write per step.
//
// Reference kernels restated here: el_stress.cu:24-131, el_velocity.cu:21-117,
// el_velocity_adj.cu:22-108, el_stress_adj.cu:22-104, utilities.cu:125-205,
// 361-424, 521-593, 653-747; driver sequence libCUFD.cu:157-463.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cond.hpp"
#include "engine.hpp"

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

namespace {

constexpr int kBZ = 64;  // threads along z (one wave64 row)
constexpr int kCondBlocks = 512;  // misfit partial sums per shot of the conditioned residual
constexpr float kC1 = 9.0f / 8.0f;
constexpr float kC2 = 1.0f / 24.0f;

struct Geo {
  int nz, nx, nzA, nPml, ns, nSteps;
  long long N;  // cells per shot
  float dt, rdz, rdx;
  int ringW, ringH, ringlen;
  int s0;  // plan index of the launch's first shot
};

struct Mdl {
  const float *lam, *mu, *amu, *ba, *bb;
};

struct Pml {
  const float *Kz, *az, *bz, *Kzh, *azh, *bzh;  // length nzA
  const float *Kx, *ax, *bx, *Kxh, *axh, *bxh;  // length nx
};

struct Wave {
  float *vz, *vx, *szz, *sxx, *sxz;
};

// C-PML memory variables, named after the reference's d_mem_* arrays
struct Mem {
  float* vzz;   // mem_dvz_dz
  float* vzx;   // mem_dvz_dx
  float* vxz;   // mem_dvx_dz
  float* vxx;   // mem_dvx_dx
  float* szz;   // mem_dszz_dz
  float* sxx;   // mem_dsxx_dx
  float* sxzz;  // mem_dsxz_dz
  float* sxzx;  // mem_dsxz_dx
};

struct Src {
  const int* zx;     // [ns][2] grid indices of the source
  const float* inj;  // [ns][nSteps] stress increment 1500^2 * stf_w(t) * dt
};

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

// Ring layout [plan shot][step][component][slot]; `s` counts from the
// launch's first shot.  (A step-major order, one step's slices of a shot
// group contiguous, measured no faster: 5.73 vs 5.81 shot-gradients/s.)
__device__ inline float* ring_ptr(float* ring, const Geo& g, int it, int comp, int s, int slot) {
  const size_t row = ((size_t)g.s0 + s) * g.nSteps + it;
  return ring + (row * 5 + comp) * (size_t)g.ringlen + slot;
}

// 1-D launch over a 2-D tile grid.  Workgroups are dealt to the 8 XCDs
// round-robin by linear id; with `xcd` set, the ids are permuted so that each
// XCD walks one contiguous band of tiles (z fastest, then x), and tiles that
// share x-halo columns share that XCD's L2.
// A launch may also split its shots into `chunks` ranges of `spt` shots, one
// range per copy of the tile grid (workgroup id = chunk * n + tile before
// the XCD permutation); tile_of returns the chunk.
// With `perm` set (one chunk only), the tile of workgroup b is perm[b]: a
// host-built order that also balances the heavy C-PML columns over the XCDs.
struct Tiling {
  int gz, gx, n, xcd;
  int chunks = 1, spt = 1 << 30;
  const int* perm = nullptr;
};

__device__ inline int tile_of(const Tiling& t, int b, int& bz, int& bx) {
  if (t.perm && t.chunks == 1) {
    b = t.perm[b];
  } else if (t.xcd) {
    const int per = t.n * t.chunks / 8;
    if (b < per * 8) b = (b % 8) * per + b / 8;
  }
  const int c = b / t.n;
  b -= c * t.n;
  bz = b % t.gz;
  bx = b / t.gz;
  return c;
}
__device__ inline int tile_of(const Tiling& t, int& bz, int& bx) {
  return tile_of(t, (int)blockIdx.x, bz, bx);
}

// Loads and stores at a byte offset from a wave-uniform base: the per-shot,
// per-neighbour base is uniform across the wave and the per-lane part of
// every address is one 32-bit offset shared by all fields and shots.
__device__ __forceinline__ float ldo(const float* base, unsigned bo) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + bo);
}
__device__ __forceinline__ void sto(float* base, unsigned bo, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + bo) = v;
}

// ------------------------------------------------------------------ forward
// The split kernels below walk all shots of the launch inside each thread, so
// model parameters and C-PML coefficients are read once per cell.  Shots are
// taken U at a time: every load of the U shots is issued before the first
// store, so a wave keeps U shots' worth of reads in flight instead of one
// (the stores of shot s may alias the loads of shot s+1 as far as the
// compiler knows, so it cannot reorder them itself).

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
__global__ __launch_bounds__(256) void k_stress_fwd(Tiling tl, Geo g, Mdl m, Pml c, Wave w, Mem p, Src src,
                                                    int it, float* __restrict__ ring) {
  int tz, tx;
  const int chunk = tile_of(tl, tz, tx);
  const int sb = chunk * tl.spt, se = min(g.ns, sb + tl.spt);  // shots of this workgroup
  const int z = tz * kBZ + threadIdx.x;
  const int x = tx * (int)blockDim.y + threadIdx.y;
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i];
  const float l2m = lam + 2.f * mu;
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml - 1;
  float bz = 0.f, az = 0.f, rkz = 1.f, bzh = 0.f, azh = 0.f, rkzh = 1.f;
  float bx = 0.f, ax = 0.f, rkx = 1.f, bxh = 0.f, axh = 0.f, rkxh = 1.f;
  if (pz) {
    bz = c.bz[z], az = c.az[z], rkz = 1.f / c.Kz[z];
    bzh = c.bzh[z], azh = c.azh[z], rkzh = 1.f / c.Kzh[z];
  }
  if (px) {
    bx = c.bx[x], ax = c.ax[x], rkx = 1.f / c.Kx[x];
    bxh = c.bxh[x], axh = c.axh[x], rkxh = 1.f / c.Kxh[x];
  }
  const int slot = ring ? ring_slot(z, x, g) : -1;
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  for (int s0 = sb; s0 < se; s0 += U) {
    StressIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const float* vz = w.vz + so;
      const float* vx = w.vx + so;
      StressIn& q = r[u];
      q.vz0 = ldo(vz, bo), q.vzm1 = ldo(vz - 1, bo), q.vzp1 = ldo(vz + 1, bo), q.vzm2 = ldo(vz - 2, bo);
      q.vzpx = ldo(vz + nz, bo), q.vzp2x = ldo(vz + 2 * nz, bo), q.vzmx = ldo(vz - nz, bo);
      q.vx0 = ldo(vx, bo), q.vxmx = ldo(vx - nz, bo), q.vxpx = ldo(vx + nz, bo), q.vxm2x = ldo(vx - 2 * nz, bo);
      q.vxp1 = ldo(vx + 1, bo), q.vxp2 = ldo(vx + 2, bo), q.vxm1 = ldo(vx - 1, bo);
      q.szz = ldo(w.szz + so, bo), q.sxx = ldo(w.sxx + so, bo), q.sxz = ldo(w.sxz + so, bo);
      if (pz) q.mvzz = ldo(p.vzz + so, bo), q.mvxz = ldo(p.vxz + so, bo);
      if (px) q.mvxx = ldo(p.vxx + so, bo), q.mvzx = ldo(p.vzx + so, bo);
      q.sz = src.zx[2 * s], q.sx = src.zx[2 * s + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const StressIn& q = r[u];
      float dvz_dz = (kC1 * (q.vz0 - q.vzm1) - kC2 * (q.vzp1 - q.vzm2)) * rdz;
      float dvx_dx = (kC1 * (q.vx0 - q.vxmx) - kC2 * (q.vxpx - q.vxm2x)) * rdx;
      float dvx_dz = (kC1 * (q.vxp1 - q.vx0) - kC2 * (q.vxp2 - q.vxm1)) * rdz;
      float dvz_dx = (kC1 * (q.vzpx - q.vz0) - kC2 * (q.vzp2x - q.vzmx)) * rdx;
      if (pz) {
        const float a = bz * q.mvzz + az * dvz_dz;
        sto(p.vzz + so, bo, a);
        dvz_dz = dvz_dz * rkz + a;
        const float b = bzh * q.mvxz + azh * dvx_dz;
        sto(p.vxz + so, bo, b);
        dvx_dz = dvx_dz * rkzh + b;
      }
      if (px) {
        const float a = bx * q.mvxx + ax * dvx_dx;
        sto(p.vxx + so, bo, a);
        dvx_dx = dvx_dx * rkx + a;
        const float b = bxh * q.mvzx + axh * dvz_dx;
        sto(p.vzx + so, bo, b);
        dvz_dx = dvz_dx * rkxh + b;
      }
      float szz = q.szz, sxx = q.sxx, sxz = q.sxz;
      if (slot >= 0) {
        *ring_ptr(ring, g, it, 0, s, slot) = szz;
        *ring_ptr(ring, g, it, 1, s, slot) = sxz;
        *ring_ptr(ring, g, it, 2, s, slot) = sxx;
        *ring_ptr(ring, g, it, 3, s, slot) = q.vz0;
        *ring_ptr(ring, g, it, 4, s, slot) = q.vx0;
      }
      szz += (l2m * dvz_dz + lam * dvx_dx) * dt;
      sxx += (lam * dvz_dz + l2m * dvx_dx) * dt;
      sxz += amu * (dvx_dz + dvz_dx) * dt;
      if (z == q.sz && x == q.sx) {
        const float inj = src.inj[(size_t)s * g.nSteps + it];
        szz += inj;
        sxx += 3.f * inj;
      }
      sto(w.szz + so, bo, szz);
      sto(w.sxx + so, bo, sxx);
      sto(w.sxz + so, bo, sxz);
    }
  }
}

// el_velocity.cu:45-82.  x-PML test is `x > nx - nPml` (el_velocity.cu:56,71).
// With `illum` set, the launch's sum over shots of vz^2 + vx^2 is added to it
// (forward illumination; no reference counterpart).
struct VelIn {
  float szz0, szzp1, szzp2, szzm1;
  float sxz0, sxzmx, sxzpx, sxzm2x, sxzm1, sxzp1, sxzm2;
  float sxxpx, sxx0, sxxp2x, sxxmx;
  float mszz, msxzz, msxzx, msxx;
  float vz, vx;
};

template <int U>
__device__ __forceinline__ void k_vel_fwd_body(int bid, Tiling tl, Geo g, Mdl m, Pml c, Wave w, Mem p,
                                                float* __restrict__ illum) {
  int tz, tx;
  const int chunk = tile_of(tl, bid, tz, tx);
  const int sb = chunk * tl.spt, se = min(g.ns, sb + tl.spt);  // shots of this workgroup
  const int z = tz * kBZ + threadIdx.x;
  const int x = tx * (int)blockDim.y + threadIdx.y;
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const float ba = m.ba[i], bb = m.bb[i];
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml;
  float bz = 0.f, az = 0.f, rkz = 1.f, bzh = 0.f, azh = 0.f, rkzh = 1.f;
  float bx = 0.f, ax = 0.f, rkx = 1.f, bxh = 0.f, axh = 0.f, rkxh = 1.f;
  if (pz) {
    bz = c.bz[z], az = c.az[z], rkz = 1.f / c.Kz[z];
    bzh = c.bzh[z], azh = c.azh[z], rkzh = 1.f / c.Kzh[z];
  }
  if (px) {
    bx = c.bx[x], ax = c.ax[x], rkx = 1.f / c.Kx[x];
    bxh = c.bxh[x], axh = c.axh[x], rkxh = 1.f / c.Kxh[x];
  }
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  float il = 0.f;  // illumination of this launch's shots
  for (int s0 = sb; s0 < se; s0 += U) {
    VelIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const float* szz = w.szz + so;
      const float* sxx = w.sxx + so;
      const float* sxz = w.sxz + so;
      VelIn& q = r[u];
      q.szz0 = ldo(szz, bo), q.szzp1 = ldo(szz + 1, bo), q.szzp2 = ldo(szz + 2, bo), q.szzm1 = ldo(szz - 1, bo);
      q.sxz0 = ldo(sxz, bo), q.sxzmx = ldo(sxz - nz, bo), q.sxzpx = ldo(sxz + nz, bo), q.sxzm2x = ldo(sxz - 2 * nz, bo);
      q.sxzm1 = ldo(sxz - 1, bo), q.sxzp1 = ldo(sxz + 1, bo), q.sxzm2 = ldo(sxz - 2, bo);
      q.sxxpx = ldo(sxx + nz, bo), q.sxx0 = ldo(sxx, bo), q.sxxp2x = ldo(sxx + 2 * nz, bo), q.sxxmx = ldo(sxx - nz, bo);
      if (pz) q.mszz = ldo(p.szz + so, bo), q.msxzz = ldo(p.sxzz + so, bo);
      if (px) q.msxzx = ldo(p.sxzx + so, bo), q.msxx = ldo(p.sxx + so, bo);
      q.vz = ldo(w.vz + so, bo), q.vx = ldo(w.vx + so, bo);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const VelIn& q = r[u];
      float dszz_dz = (kC1 * (q.szzp1 - q.szz0) - kC2 * (q.szzp2 - q.szzm1)) * rdz;
      float dsxz_dx = (kC1 * (q.sxz0 - q.sxzmx) - kC2 * (q.sxzpx - q.sxzm2x)) * rdx;
      float dsxz_dz = (kC1 * (q.sxz0 - q.sxzm1) - kC2 * (q.sxzp1 - q.sxzm2)) * rdz;
      float dsxx_dx = (kC1 * (q.sxxpx - q.sxx0) - kC2 * (q.sxxp2x - q.sxxmx)) * rdx;
      if (pz) {
        const float a = bzh * q.mszz + azh * dszz_dz;
        sto(p.szz + so, bo, a);
        dszz_dz = dszz_dz * rkzh + a;
        const float b = bz * q.msxzz + az * dsxz_dz;
        sto(p.sxzz + so, bo, b);
        dsxz_dz = dsxz_dz * rkz + b;
      }
      if (px) {
        const float a = bx * q.msxzx + ax * dsxz_dx;
        sto(p.sxzx + so, bo, a);
        dsxz_dx = dsxz_dx * rkx + a;
        const float b = bxh * q.msxx + axh * dsxx_dx;
        sto(p.sxx + so, bo, b);
        dsxx_dx = dsxx_dx * rkxh + b;
      }
      float vz = q.vz, vx = q.vx;
      vz += (dszz_dz + dsxz_dx) * ba * dt;
      vx += (dsxz_dz + dsxx_dx) * bb * dt;
      sto(w.vz + so, bo, vz);
      sto(w.vx + so, bo, vx);
      il += vz * vz + vx * vx;
    }
  }
  if (illum) illum[i] += il;
}

template <int U>
__global__ __launch_bounds__(256) void k_vel_fwd(Tiling tl, Geo g, Mdl m, Pml c, Wave w, Mem p,
                                                float* __restrict__ illum) {
  k_vel_fwd_body<U>(blockIdx.x, tl, g, m, c, w, p, illum);
}

// ------------------------------------------------------------ fused forward
// One time step (el_stress.cu:50-88 + add_source + el_velocity.cu:45-82 and
// the ring save of from_bnd) in one pass.  A workgroup owns every row of one
// shot for the x-segment [x0, x0+tx) and marches along x:
//   column c  : sigma^{n+1}(., c) from v^n(., c-2..c+2)
//   column c-2: v^{n+1}(., c-2) from sigma^{n+1}(., c-4..c)
// x-neighbours are carried in per-thread register rings, z-neighbours are read
// from LDS.  sigma^{n+1} is recomputed on 2 halo columns each side (same code,
// same bits as the owner's value).  State and memory variables are read from
// buffer A and written to buffer B, so halo reads never see a neighbour's new
// values.
struct Seg {
  int tx, nseg, n, xcd;
};

__device__ inline void seg_of(const Seg& q, int ns, int& seg, int& s) {
  int b = blockIdx.x;
  if (q.xcd) {
    const int per = q.n / 8;
    if (b < per * 8) b = (b % 8) * per + b / 8;
  }
  seg = b / ns;
  s = b % ns;
}

__device__ inline int mod3(int c) { return (c + 3) % 3; }

__global__ __launch_bounds__(1024) void k_fwd_step(Seg q, Geo g, Mdl m, Pml c, Wave A, Wave B,
                                                   Mem pA, Mem pB, Src src, int it,
                                                   float* __restrict__ ring) {
  extern __shared__ float lds[];
  const int NT = blockDim.x;
  float* lvz = lds;            // [2][NT] v^n at the current column
  float* lvx = lds + 2 * NT;
  float* lszz = lds + 4 * NT;  // [3][NT] sigma^{n+1} of the last 3 columns
  float* lsxz = lds + 7 * NT;
  const int z = threadIdx.x;
  int seg, s;
  seg_of(q, g.ns, seg, s);
  const int x0 = seg * q.tx;
  const int x1 = min(x0 + q.tx, g.nx);
  const int nz = g.nz;
  const long long so = (long long)s * g.N;
  const bool zval = z < nz;
  const bool zin = z >= 2 && z <= g.nzA - 3;
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  float bz = 0.f, az = 0.f, rkz = 1.f, bzh = 0.f, azh = 0.f, rkzh = 1.f;
  if (zin && pz) {
    bz = c.bz[z], az = c.az[z], rkz = 1.f / c.Kz[z];
    bzh = c.bzh[z], azh = c.azh[z], rkzh = 1.f / c.Kzh[z];
  }
  const int zs = src.zx[2 * s], xs = src.zx[2 * s + 1];
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  auto ld = [&](const float* f, int col) -> float {
    return (zval && col >= 0 && col < g.nx) ? f[so + (long long)col * nz + z] : 0.f;
  };
  // v rings: r*[k] holds column c-2+k
  float vz0 = ld(A.vz, x0 - 4), vz1 = ld(A.vz, x0 - 3), vz2 = ld(A.vz, x0 - 2),
        vz3 = ld(A.vz, x0 - 1), vz4;
  float vx0 = ld(A.vx, x0 - 4), vx1 = ld(A.vx, x0 - 3), vx2 = ld(A.vx, x0 - 2),
        vx3 = ld(A.vx, x0 - 1), vx4;
  // sigma^{n+1} rings: sxz at columns c-4..c, sxx at columns c-3..c
  float sxz0 = 0.f, sxz1 = 0.f, sxz2 = 0.f, sxz3 = 0.f, sxz4 = 0.f;
  float sxx0 = 0.f, sxx1 = 0.f, sxx2 = 0.f, sxx3 = 0.f;
  for (int col = x0 - 2; col <= x1 + 1; ++col) {
    // ---- phase 1: bring v^n(., col+2) in, publish v^n(., col) for z-stencils
    vz4 = ld(A.vz, col + 2);
    vx4 = ld(A.vx, col + 2);
    const int sv = (col & 1) * NT;
    lvz[sv + z] = vz2;
    lvx[sv + z] = vx2;
    __syncthreads();
    // ---- phase 2: sigma^{n+1}(z, col)
    float nszz = 0.f, nsxx = 0.f, nsxz = 0.f;
    if (zval && col >= 0 && col < g.nx) {
      const long long i = (long long)col * nz + z;
      const long long o = so + i;
      float szz = A.szz[o], sxx = A.sxx[o], sxz = A.sxz[o];
      const bool own = col >= x0 && col < x1;
      if (zin && col >= 2 && col <= g.nx - 3) {
        const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i];
        const float l2m = lam + 2.f * mu;
        const float* Vz = lvz + sv + z;
        const float* Vx = lvx + sv + z;
        float dvz_dz = (kC1 * (vz2 - Vz[-1]) - kC2 * (Vz[1] - Vz[-2])) * rdz;
        float dvx_dx = (kC1 * (vx2 - vx1) - kC2 * (vx3 - vx0)) * rdx;
        float dvx_dz = (kC1 * (Vx[1] - vx2) - kC2 * (Vx[2] - Vx[-1])) * rdz;
        float dvz_dx = (kC1 * (vz3 - vz2) - kC2 * (vz4 - vz1)) * rdx;
        if (pz) {
          const float a = bz * pA.vzz[o] + az * dvz_dz;
          dvz_dz = dvz_dz * rkz + a;
          const float b = bzh * pA.vxz[o] + azh * dvx_dz;
          dvx_dz = dvx_dz * rkzh + b;
          if (own) {
            pB.vzz[o] = a;
            pB.vxz[o] = b;
          }
        }
        if (col < g.nPml || col > g.nx - g.nPml - 1) {
          const float a = c.bx[col] * pA.vxx[o] + c.ax[col] * dvx_dx;
          dvx_dx = dvx_dx * (1.f / c.Kx[col]) + a;
          const float b = c.bxh[col] * pA.vzx[o] + c.axh[col] * dvz_dx;
          dvz_dx = dvz_dx * (1.f / c.Kxh[col]) + b;
          if (own) {
            pB.vxx[o] = a;
            pB.vzx[o] = b;
          }
        }
        if (own && ring) {
          const int slot = ring_slot(z, col, g);
          if (slot >= 0) {
            *ring_ptr(ring, g, it, 0, s, slot) = szz;
            *ring_ptr(ring, g, it, 1, s, slot) = sxz;
            *ring_ptr(ring, g, it, 2, s, slot) = sxx;
            *ring_ptr(ring, g, it, 3, s, slot) = vz2;
            *ring_ptr(ring, g, it, 4, s, slot) = vx2;
          }
        }
        szz += (l2m * dvz_dz + lam * dvx_dx) * dt;
        sxx += (lam * dvz_dz + l2m * dvx_dx) * dt;
        sxz += amu * (dvx_dz + dvz_dx) * dt;
        if (z == zs && col == xs) {
          const float inj = src.inj[(size_t)s * g.nSteps + it];
          szz += inj;
          sxx += 3.f * inj;
        }
        if (own) {
          B.szz[o] = szz;
          B.sxx[o] = sxx;
          B.sxz[o] = sxz;
        }
      }
      nszz = szz;
      nsxx = sxx;
      nsxz = sxz;
    }
    const int ss = mod3(col) * NT;
    lszz[ss + z] = nszz;
    lsxz[ss + z] = nsxz;
    sxz0 = sxz1, sxz1 = sxz2, sxz2 = sxz3, sxz3 = sxz4, sxz4 = nsxz;
    sxx0 = sxx1, sxx1 = sxx2, sxx2 = sxx3, sxx3 = nsxx;
    __syncthreads();
    // ---- phase 3: v^{n+1}(z, col-2)
    const int cp = col - 2;
    if (zin && cp >= x0 && cp < x1 && cp >= 2 && cp <= g.nx - 3) {
      const long long i = (long long)cp * nz + z;
      const long long o = so + i;
      const float ba = m.ba[i], bb = m.bb[i];
      const int sp = mod3(cp) * NT + z;
      const float* Szz = lszz + sp;
      const float* Sxz = lsxz + sp;
      float dszz_dz = (kC1 * (Szz[1] - Szz[0]) - kC2 * (Szz[2] - Szz[-1])) * rdz;
      float dsxz_dx = (kC1 * (sxz2 - sxz1) - kC2 * (sxz3 - sxz0)) * rdx;
      float dsxz_dz = (kC1 * (Sxz[0] - Sxz[-1]) - kC2 * (Sxz[1] - Sxz[-2])) * rdz;
      float dsxx_dx = (kC1 * (sxx2 - sxx1) - kC2 * (sxx3 - sxx0)) * rdx;
      if (pz) {
        const float a = bzh * pA.szz[o] + azh * dszz_dz;
        pB.szz[o] = a;
        dszz_dz = dszz_dz * rkzh + a;
        const float b = bz * pA.sxzz[o] + az * dsxz_dz;
        pB.sxzz[o] = b;
        dsxz_dz = dsxz_dz * rkz + b;
      }
      if (cp < g.nPml || cp > g.nx - g.nPml) {
        const float a = c.bx[cp] * pA.sxzx[o] + c.ax[cp] * dsxz_dx;
        pB.sxzx[o] = a;
        dsxz_dx = dsxz_dx * (1.f / c.Kx[cp]) + a;
        const float b = c.bxh[cp] * pA.sxx[o] + c.axh[cp] * dsxx_dx;
        pB.sxx[o] = b;
        dsxx_dx = dsxx_dx * (1.f / c.Kxh[cp]) + b;
      }
      B.vz[o] = vz0 + (dszz_dz + dsxz_dx) * ba * dt;
      B.vx[o] = vx0 + (dsxz_dz + dsxx_dx) * bb * dt;
    }
    vz0 = vz1, vz1 = vz2, vz2 = vz3, vz3 = vz4;
    vx0 = vx1, vx1 = vx2, vx2 = vx3, vx3 = vx4;
  }
}

// ------------------------------------------------------------ fused adjoint
// One adjoint step: el_velocity_adj.cu:22-108, res_injection
// (utilities.cu:569-580), el_stress_adj.cu:22-104 and source_grad
// (utilities.cu:582-593) in one column-marching pass, same scheme as
// k_fwd_step: adjoint velocity at column c, adjoint stress at column c-2,
// state and memory variables read from A and written to B.  The residual is
// injected into the owner's adjoint stress before the stress update, through
// a per-shot, per-column receiver index.
struct RecIdx {
  const int* ptr;  // [ns][nx+1] start of each column's receivers
  const int* zr;   // receiver row
  const int* rg;   // receiver index in the time-major residual row
};

__global__ __launch_bounds__(1024) void k_adj_step(Seg q, Geo g, Mdl m, Pml c, Wave A, Wave B,
                                                   Mem pA, Mem pB, Src src, RecIdx rx,
                                                   const float* __restrict__ res_row, int R,
                                                   float* __restrict__ stf_grad, int stf_it) {
  extern __shared__ float lds[];
  const int NT = blockDim.x;
  float* lszz = lds;             // [2][NT] adjoint stress (old) at column c
  float* lsxx = lds + 2 * NT;
  float* lsxz = lds + 4 * NT;
  float* lvz = lds + 6 * NT;     // [3][NT] adjoint velocity (new), last 3 columns
  float* lvx = lds + 9 * NT;
  float* lpzz = lds + 12 * NT;   // [3][NT] new mem_dsxz_dz
  float* lpsz = lds + 15 * NT;   // [3][NT] new mem_dszz_dz
  const int z = threadIdx.x;
  int seg, s;
  seg_of(q, g.ns, seg, s);
  const int x0 = seg * q.tx;
  const int x1 = min(x0 + q.tx, g.nx);
  const int nz = g.nz;
  const long long so = (long long)s * g.N;
  const bool zval = z < nz;
  const bool zin = z >= 2 && z <= g.nzA - 3;
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool zlive = z <= g.nPml + 1 || z >= g.nzA - g.nPml - 2;
  float az = 0.f, rkz = 1.f, bz = 0.f, azh = 0.f, rkzh = 1.f, bzh = 0.f;
  if (zin) {
    az = c.az[z], rkz = 1.f / c.Kz[z], bz = c.bz[z];
    azh = c.azh[z], rkzh = 1.f / c.Kzh[z], bzh = c.bzh[z];
  }
  const int zs = src.zx[2 * s], xs = src.zx[2 * s + 1];
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  auto ld = [&](const float* f, int col) -> float {
    return (zval && col >= 0 && col < g.nx) ? f[so + (long long)col * nz + z] : 0.f;
  };
  // adjoint stress (old) rings, slot k = column c-2+k
  float zz0 = ld(A.szz, x0 - 4), zz1 = ld(A.szz, x0 - 3), zz2 = ld(A.szz, x0 - 2),
        zz3 = ld(A.szz, x0 - 1), zz4;
  float xx0 = ld(A.sxx, x0 - 4), xx1 = ld(A.sxx, x0 - 3), xx2 = ld(A.sxx, x0 - 2),
        xx3 = ld(A.sxx, x0 - 1), xx4;
  float xz0 = ld(A.sxz, x0 - 4), xz1 = ld(A.sxz, x0 - 3), xz2 = ld(A.sxz, x0 - 2),
        xz3 = ld(A.sxz, x0 - 1), xz4;
  // new adjoint velocity rings: vz at columns c-3..c, vx at c-4..c
  float nvz0 = 0.f, nvz1 = 0.f, nvz2 = 0.f, nvz3 = 0.f;
  float nvx0 = 0.f, nvx1 = 0.f, nvx2 = 0.f, nvx3 = 0.f, nvx4 = 0.f;
  // new velocity-side memory variables: mem_dsxz_dx at c-3..c, mem_dsxx_dx at c-4..c
  float pxz0 = 0.f, pxz1 = 0.f, pxz2 = 0.f, pxz3 = 0.f;
  float pxx0 = 0.f, pxx1 = 0.f, pxx2 = 0.f, pxx3 = 0.f, pxx4 = 0.f;
  for (int col = x0 - 2; col <= x1 + 1; ++col) {
    // ---- phase 1
    zz4 = ld(A.szz, col + 2);
    xx4 = ld(A.sxx, col + 2);
    xz4 = ld(A.sxz, col + 2);
    const int sv = (col & 1) * NT;
    lszz[sv + z] = zz2;
    lsxx[sv + z] = xx2;
    lsxz[sv + z] = xz2;
    __syncthreads();
    // ---- phase 2: adjoint velocity at (z, col)
    float vzn = 0.f, vxn = 0.f, psxz = 0.f, psxx = 0.f, psxzz = 0.f, pszz = 0.f;
    const bool own = col >= x0 && col < x1;
    if (zval && col >= 0 && col < g.nx) {
      const long long i = (long long)col * nz + z;
      const long long o = so + i;
      vzn = A.vz[o];
      vxn = A.vx[o];
      psxz = pA.sxzx[o];
      psxx = pA.sxx[o];
      psxzz = pA.sxzz[o];
      pszz = pA.szz[o];
      if (own && stf_it >= 0 && z == zs && col == xs)
        stf_grad[(size_t)s * g.nSteps + stf_it] = -(zz2 + 3.f * xx2) * dt;
      if (zin && col >= 2 && col <= g.nx - 3) {
        const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i], ba = m.ba[i], bb = m.bb[i];
        const float l2m = lam + 2.f * mu;
        const float ax = c.ax[col], rkx = 1.f / c.Kx[col], bx = c.bx[col];
        const float axh = c.axh[col], rkxh = 1.f / c.Kxh[col], bxh = c.bxh[col];
        const bool px = col < g.nPml || col > g.nx - g.nPml - 1;
        const float* Szz = lszz + sv + z;
        const float* Sxx = lsxx + sv + z;
        const float* Sxz = lsxz + sv + z;
        const float dszz_dx = (-kC1 * (zz3 - zz2) + kC2 * (zz4 - zz1)) * rdx;
        const float dsxx_dx = (-kC1 * (xx3 - xx2) + kC2 * (xx4 - xx1)) * rdx;
        const float dsxz_dz = (-kC1 * (Sxz[0] - Sxz[-1]) + kC2 * (Sxz[1] - Sxz[-2])) * rdz;
        vxn += (lam * dszz_dx * rkx * dt + l2m * dsxx_dx * rkx * dt + amu * rkzh * dsxz_dz * dt);
        if (ax != 0.f) {
          const float* qv = pA.vxx + o;
          vxn += ax * ((-kC1 * (qv[nz] - qv[0]) + kC2 * (qv[2 * nz] - qv[-nz])) * rdx);
        }
        if (azh != 0.f) {
          const float* qv = pA.vxz + o;
          vxn += azh * ((-kC1 * (qv[0] - qv[-1]) + kC2 * (qv[1] - qv[-2])) * rdz);
        }
        if (px) psxx = bxh * psxx + bb * vxn * dt;
        if (pz) psxzz = bz * psxzz + bb * vxn * dt;
        const float dszz_dz = (-kC1 * (Szz[1] - Szz[0]) + kC2 * (Szz[2] - Szz[-1])) * rdz;
        const float dsxx_dz = (-kC1 * (Sxx[1] - Sxx[0]) + kC2 * (Sxx[2] - Sxx[-1])) * rdz;
        const float dsxz_dx = (-kC1 * (xz2 - xz1) + kC2 * (xz3 - xz0)) * rdx;
        vzn += (l2m * dszz_dz * rkz * dt + lam * dsxx_dz * rkz * dt + amu * rkxh * dsxz_dx * dt);
        if (az != 0.f) {
          const float* qv = pA.vzz + o;
          vzn += az * ((-kC1 * (qv[1] - qv[0]) + kC2 * (qv[2] - qv[-1])) * rdz);
        }
        if (axh != 0.f) {
          const float* qv = pA.vzx + o;
          vzn += axh * ((-kC1 * (qv[0] - qv[-nz]) + kC2 * (qv[nz] - qv[-2 * nz])) * rdx);
        }
        if (px) psxz = bx * psxz + ba * vzn * dt;
        if (pz) pszz = bzh * pszz + ba * vzn * dt;
        if (own) {
          B.vx[o] = vxn;
          B.vz[o] = vzn;
          if (px) {
            pB.sxx[o] = psxx;
            pB.sxzx[o] = psxz;
          }
          if (pz) {
            pB.sxzz[o] = psxzz;
            pB.szz[o] = pszz;
          }
        }
      }
    }
    const int sn = mod3(col) * NT;
    lvz[sn + z] = vzn;
    lvx[sn + z] = vxn;
    lpzz[sn + z] = psxzz;
    lpsz[sn + z] = pszz;
    nvz0 = nvz1, nvz1 = nvz2, nvz2 = nvz3, nvz3 = vzn;
    nvx0 = nvx1, nvx1 = nvx2, nvx2 = nvx3, nvx3 = nvx4, nvx4 = vxn;
    pxz0 = pxz1, pxz1 = pxz2, pxz2 = pxz3, pxz3 = psxz;
    pxx0 = pxx1, pxx1 = pxx2, pxx2 = pxx3, pxx3 = pxx4, pxx4 = psxx;
    __syncthreads();
    // ---- phase 3: residual injection + adjoint stress at (z, col-2)
    const int cp = col - 2;
    if (cp >= x0 && cp < x1 && zval) {
      const long long i = (long long)cp * nz + z;
      const long long o = so + i;
      float szz = zz0, sxx = xx0, sxz = xz0;
      const int* rp = rx.ptr + (size_t)s * (g.nx + 1);
      for (int k = rp[cp]; k < rp[cp + 1]; ++k)
        if (rx.zr[k] == z) {
          const float r = res_row[rx.rg[k]];
          szz += r;
          sxx += 3.f * r;
        }
      if (zin && cp >= 2 && cp <= g.nx - 3) {
        const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i], ba = m.ba[i], bb = m.bb[i];
        const float l2m = lam + 2.f * mu;
        const float ax = c.ax[cp], rkx = 1.f / c.Kx[cp], bx = c.bx[cp];
        const float axh = c.axh[cp], rkxh = 1.f / c.Kxh[cp], bxh = c.bxh[cp];
        const bool xlive = cp <= g.nPml + 1 || cp >= g.nx - g.nPml - 2;
        const int sp = mod3(cp) * NT + z;
        const float* Vz = lvz + sp;
        const float* Vx = lvx + sp;
        // nvz slot k = column col-3+k (cp = slot 1); nvx slot k = column col-4+k (cp = slot 2)
        const float dvz_dx = (-kC1 * (nvz2 - nvz1) + kC2 * (nvz3 - nvz0)) * rdx;
        const float dvx_dz = (-kC1 * (Vx[1] - Vx[0]) + kC2 * (Vx[2] - Vx[-1])) * rdz;
        sxz += (dvz_dx * rkx * ba * dt + dvx_dz * rkz * bb * dt);
        if (ax != 0.f)  // mem_dsxz_dx at cp-1..cp+2: pxz slot k = column col-3+k
          sxz += ax * ((-kC1 * (pxz2 - pxz1) + kC2 * (pxz3 - pxz0)) * rdx);
        if (az != 0.f) {
          const float* qv = lpzz + sp;
          sxz += az * ((-kC1 * (qv[1] - qv[0]) + kC2 * (qv[2] - qv[-1])) * rdz);
        }
        const float dvx_dx = (-kC1 * (nvx2 - nvx1) + kC2 * (nvx3 - nvx0)) * rdx;
        const float dvz_dz = (-kC1 * (Vz[0] - Vz[-1]) + kC2 * (Vz[1] - Vz[-2])) * rdz;
        sxx += bb * dvx_dx * rkxh * dt;
        if (axh != 0.f)  // mem_dsxx_dx at cp-2..cp+1: pxx slot k = column col-4+k
          sxx += axh * ((-kC1 * (pxx2 - pxx1) + kC2 * (pxx3 - pxx0)) * rdx);
        szz += ba * dvz_dz * rkzh * dt;
        if (azh != 0.f) {
          const float* qv = lpsz + sp;
          szz += azh * ((-kC1 * (qv[0] - qv[-1]) + kC2 * (qv[1] - qv[-2])) * rdz);
        }
        if (xlive) {
          pB.vzx[o] = bxh * pA.vzx[o] + sxz * amu * dt;
          pB.vxx[o] = bx * pA.vxx[o] + lam * szz * dt + l2m * sxx * dt;
        }
        if (zlive) {
          pB.vxz[o] = bzh * pA.vxz[o] + sxz * amu * dt;
          pB.vzz[o] = bz * pA.vzz[o] + l2m * szz * dt + lam * sxx * dt;
        }
      }
      B.szz[o] = szz;
      B.sxx[o] = sxx;
      B.sxz[o] = sxz;
    }
    zz0 = zz1, zz1 = zz2, zz2 = zz3, zz3 = zz4;
    xx0 = xx1, xx1 = xx2, xx2 = xx3, xx3 = xx4;
    xz0 = xz1, xz1 = xz2, xz2 = xz3, xz3 = xz4;
  }
}

// ------------------------------------------------------------ fused reverse
// One reverse step: el_velocity.cu:84-117 and el_stress.cu:90-131 with
// isFor = false, the ring restores of to_bnd (utilities.cu:394-424) and the
// source removal of add_source, plus all gradient terms, in one column-marching
// pass: v^{it}(c) from sigma^{it+1}(c-2..c+2), then sigma^{it}(c-2) from
// v^{it}(c-4..c).  A workgroup owns one x-segment of TXR columns for a group of
// shots; it walks the group's shots one after another and keeps the gradient
// terms of its cells in LDS (each thread only touches its own row), then adds
// them once to the group's slice of
// the partial accumulators (one slice per group, so no two workgroups write
// the same word).  State is read from A (step it+1) and written to B (step it).
constexpr int TXR = 8;

struct SegG {
  int nseg, ngrp, gsz, n, xcd;
};

__global__ __launch_bounds__(1024) void k_rev_step(SegG q, Geo g, Mdl m, Wave A, Wave B, Wave adj,
                                                   Src src, const float* __restrict__ ring,
                                                   int it, float* __restrict__ accp) {
  extern __shared__ float lds[];
  const int NT = blockDim.x;
  float* lszz = lds;          // [2][NT] sigma^{it+1} at column c
  float* lsxz = lds + 2 * NT;
  float* lvz = lds + 4 * NT;  // [3][NT] v^{it} of the last 3 columns
  float* lvx = lds + 7 * NT;
  float* lacc = lds + 10 * NT;  // [5][TXR][NT] gradient terms of the owned cells
  const int z = threadIdx.x;
  int b = blockIdx.x;
  if (q.xcd) {
    const int per = q.n / 8;
    if (b < per * 8) b = (b % 8) * per + b / 8;
  }
  const int seg = b / q.ngrp, grp = b % q.ngrp;
  const int x0 = seg * TXR;
  const int x1 = min(x0 + TXR, g.nx);
  const int nz = g.nz;
  const bool zval = z < nz;
  const bool zI = z >= g.nPml && z <= g.nzA - 1 - g.nPml;
  const int xI0 = g.nPml, xI1 = g.nx - 1 - g.nPml;
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  for (int k = 0; k < 5 * TXR; ++k) lacc[k * NT + z] = 0.f;
  auto acc = [&](int term, int j) -> float& { return lacc[(term * TXR + j) * NT + z]; };
  const int s0 = grp * q.gsz, s1 = min(s0 + q.gsz, g.ns);
  for (int s = s0; s < s1; ++s) {
    const long long so = (long long)s * g.N;
    auto ld = [&](const float* f, int col) -> float {
      return (zval && col >= 0 && col < g.nx) ? f[so + (long long)col * nz + z] : 0.f;
    };
    const int zs = src.zx[2 * s], xs = src.zx[2 * s + 1];
    // sigma^{it+1} rings, slot k = column c-2+k
    float zz0 = ld(A.szz, x0 - 4), zz1 = ld(A.szz, x0 - 3), zz2 = ld(A.szz, x0 - 2),
          zz3 = ld(A.szz, x0 - 1), zz4;
    float xx0 = ld(A.sxx, x0 - 4), xx1 = ld(A.sxx, x0 - 3), xx2 = ld(A.sxx, x0 - 2),
          xx3 = ld(A.sxx, x0 - 1), xx4;
    float xz0 = ld(A.sxz, x0 - 4), xz1 = ld(A.sxz, x0 - 3), xz2 = ld(A.sxz, x0 - 2),
          xz3 = ld(A.sxz, x0 - 1), xz4;
    // v^{it} rings: vz at columns c-3..c, vx at c-4..c
    float wz0 = 0.f, wz1 = 0.f, wz2 = 0.f, wz3 = 0.f;
    float wx0 = 0.f, wx1 = 0.f, wx2 = 0.f, wx3 = 0.f, wx4 = 0.f;
#pragma unroll 1
    for (int j = 0; j < TXR + 4; ++j) {
      const int col = x0 - 2 + j;
      // ---- phase 1
      zz4 = ld(A.szz, col + 2);
      xx4 = ld(A.sxx, col + 2);
      xz4 = ld(A.sxz, col + 2);
      const int sv = (j & 1) * NT;
      lszz[sv + z] = zz2;
      lsxz[sv + z] = xz2;
      __syncthreads();
      // ---- phase 2: v^{it}(z, col)
      float nvz = 0.f, nvx = 0.f;
      const bool jown = j >= 2 && j < TXR + 2;  // col in [x0, x0+TXR)
      if (zval && col >= 0 && col < g.nx) {
        const long long i = (long long)col * nz + z;
        const long long o = so + i;
        nvz = A.vz[o];
        nvx = A.vx[o];
        const bool own = jown && col < x1;
        const bool inI = zI && col >= xI0 && col <= xI1;
        const int slot = ring_slot(z, col, g);
        if (inI) {
          const float* Szz = lszz + sv + z;
          const float* Sxz = lsxz + sv + z;
          const float dszz_dz = (kC1 * (Szz[1] - zz2) - kC2 * (Szz[2] - Szz[-1])) * rdz;
          const float dsxz_dx = (kC1 * (xz2 - xz1) - kC2 * (xz3 - xz0)) * rdx;
          const float dsxz_dz = (kC1 * (xz2 - Sxz[-1]) - kC2 * (Sxz[1] - Sxz[-2])) * rdz;
          const float dsxx_dx = (kC1 * (xx3 - xx2) - kC2 * (xx4 - xx1)) * rdx;
          const float fz = dszz_dz + dsxz_dx, fx = dsxz_dz + dsxx_dx;
          if (own) {
            acc(3, j - 2) += adj.vz[o] * fz;
            acc(4, j - 2) += adj.vx[o] * fx;
          }
          if (slot < 0) {
            nvz -= fz * m.ba[i] * dt;
            nvx -= fx * m.bb[i] * dt;
          }
        }
        if (slot >= 0) {
          nvz = *ring_ptr(const_cast<float*>(ring), g, it, 3, s, slot);
          nvx = *ring_ptr(const_cast<float*>(ring), g, it, 4, s, slot);
        }
        if (own) {
          B.vz[o] = nvz;
          B.vx[o] = nvx;
        }
      }
      const int sn = (j % 3) * NT;
      lvz[sn + z] = nvz;
      lvx[sn + z] = nvx;
      wz0 = wz1, wz1 = wz2, wz2 = wz3, wz3 = nvz;
      wx0 = wx1, wx1 = wx2, wx2 = wx3, wx3 = wx4, wx4 = nvx;
      __syncthreads();
      // ---- phase 3: sigma^{it}(z, col-2)
      const int cp = col - 2;
      if (j >= 4 && zval && cp < x1) {
        const long long i = (long long)cp * nz + z;
        const long long o = so + i;
        float szz = zz0, sxx = xx0, sxz = xz0;
        const bool inI = zI && cp >= xI0 && cp <= xI1;
        const int slot = ring_slot(z, cp, g);
        if (inI) {
          const int sp = ((j - 2) % 3) * NT + z;
          const float* Vz = lvz + sp;
          const float* Vx = lvx + sp;
          // wz slot k = column col-3+k (cp = slot 1); wx slot k = column col-4+k (cp = slot 2)
          const float dvz_dz = (kC1 * (wz1 - Vz[-1]) - kC2 * (Vz[1] - Vz[-2])) * rdz;
          const float dvx_dx = (kC1 * (wx2 - wx1) - kC2 * (wx3 - wx0)) * rdx;
          const float dvx_dz = (kC1 * (Vx[1] - wx2) - kC2 * (Vx[2] - Vx[-1])) * rdz;
          const float dvz_dx = (kC1 * (wz2 - wz1) - kC2 * (wz3 - wz0)) * rdx;
          const float szza = adj.szz[o], sxxa = adj.sxx[o];
          const int k = j - 4;
          acc(0, k) += (szza + sxxa) * (dvz_dz + dvx_dx);
          acc(1, k) += szza * dvz_dz + sxxa * dvx_dx;
          acc(2, k) += adj.sxz[o] * (dvx_dz + dvz_dx);
          if (slot < 0) {
            const float lam = m.lam[i], mu = m.mu[i];
            const float l2m = lam + 2.f * mu;
            if (z == zs && cp == xs) {
              const float inj = src.inj[(size_t)s * g.nSteps + it];
              szz -= inj;
              sxx -= 3.f * inj;
            }
            szz -= (l2m * dvz_dz + lam * dvx_dx) * dt;
            sxx -= (lam * dvz_dz + l2m * dvx_dx) * dt;
            sxz -= m.amu[i] * (dvx_dz + dvz_dx) * dt;
          }
        }
        if (slot >= 0) {
          szz = *ring_ptr(const_cast<float*>(ring), g, it, 0, s, slot);
          sxz = *ring_ptr(const_cast<float*>(ring), g, it, 1, s, slot);
          sxx = *ring_ptr(const_cast<float*>(ring), g, it, 2, s, slot);
        }
        B.szz[o] = szz;
        B.sxx[o] = sxx;
        B.sxz[o] = sxz;
      }
      zz0 = zz1, zz1 = zz2, zz2 = zz3, zz3 = zz4;
      xx0 = xx1, xx1 = xx2, xx2 = xx3, xx3 = xx4;
      xz0 = xz1, xz1 = xz2, xz2 = xz3, xz3 = xz4;
    }
  }
  if (zI) {
    const long long N = g.N;
    float* P = accp + (size_t)grp * 5 * N;
    for (int j = 0; j < TXR; ++j) {
      const int col = x0 + j;
      if (col < x1 && col >= xI0 && col <= xI1) {
        const long long i = (long long)col * nz + z;
        for (int t = 0; t < 5; ++t) P[t * N + i] += acc(t, j);
      }
    }
  }
}

// Sum the per-group partial accumulators into slice 0.
__global__ void k_sum_groups(float* __restrict__ accp, long long n5, int ngrp) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n5) return;
  float t = accp[k];
  for (int gi = 1; gi < ngrp; ++gi) t += accp[(size_t)gi * n5 + k];
  accp[k] = t;
}

// ------------------------------------------------------------ tiled fused forward
// One time step on a 32x32 output tile, every shot of the group walked inside
// the block.  Per shot: v^n on the tile + 4-cell halo goes to LDS, sigma^{n+1}
// is computed on the tile + 2-cell halo (recomputed by neighbours, same bits)
// into LDS, then v^{n+1} on the tile.  The model parameters of the cells a
// thread handles stay in registers for all shots.  State and memory variables
// are read from A and written to B.
constexpr int TT = 32;
constexpr int TR4 = TT + 8;  // 40
constexpr int TR2 = TT + 4;  // 36
constexpr int NR4 = TR4 * TR4;
constexpr int NR2 = TR2 * TR2;

__global__ __launch_bounds__(1024) void k_fwd_tile(Tiling tl, Geo g, Mdl m, Pml c, Wave A, Wave B,
                                                   Mem pA, Mem pB, Src src, int it,
                                                   float* __restrict__ ring) {
  __shared__ float svz[NR4], svx[NR4];
  __shared__ float sszz[NR2], ssxx[NR2], ssxz[NR2];
  const int t = threadIdx.x;
  int tz, tx;
  tile_of(tl, tz, tx);
  const int z0 = tz * TT, x0 = tx * TT;
  const int nz = g.nz, nx = g.nx;
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  const int oz = t % TT, ox = t / TT;
  const int zo = z0 + oz, xo = x0 + ox;
  const bool updo = zo >= 2 && zo <= g.nzA - 3 && xo >= 2 && xo <= nx - 3;
  const long long io = (long long)xo * nz + zo;
  for (int s = 0; s < g.ns; ++s) {
    const long long so = (long long)s * g.N;
    // ---- v^n on the tile + 4-cell halo, sigma^n on the 2-cell halo region
    float lvz0 = 0.f, lvx0 = 0.f, lvz1 = 0.f, lvx1 = 0.f;
    {
      const int k = t;
      const int zz = z0 - 4 + k % TR4, xx = x0 - 4 + k / TR4;
      if (zz >= 0 && zz < nz && xx >= 0 && xx < nx) {
        const long long o = so + (long long)xx * nz + zz;
        lvz0 = A.vz[o];
        lvx0 = A.vx[o];
      }
    }
    if (t + 1024 < NR4) {
      const int k = t + 1024;
      const int zz = z0 - 4 + k % TR4, xx = x0 - 4 + k / TR4;
      if (zz >= 0 && zz < nz && xx >= 0 && xx < nx) {
        const long long o = so + (long long)xx * nz + zz;
        lvz1 = A.vz[o];
        lvx1 = A.vx[o];
      }
    }
    __syncthreads();  // previous shot's readers are done with the LDS tiles
    svz[t] = lvz0;
    svx[t] = lvx0;
    if (t + 1024 < NR4) {
      svz[t + 1024] = lvz1;
      svx[t + 1024] = lvx1;
    }
    __syncthreads();
    // ---- sigma^{n+1} on the tile + 2-cell halo
    for (int k = t; k < NR2; k += 1024) {
      const int zz = z0 - 2 + k % TR2, xx = x0 - 2 + k / TR2;
      float szz = 0.f, sxx = 0.f, sxz = 0.f;
      if (zz >= 0 && zz < nz && xx >= 0 && xx < nx) {
        const long long i = (long long)xx * nz + zz;
        const long long o = so + i;
        szz = A.szz[o];
        sxx = A.sxx[o];
        sxz = A.sxz[o];
        if (zz >= 2 && zz <= g.nzA - 3 && xx >= 2 && xx <= nx - 3) {
          const bool own = zz >= z0 && zz < z0 + TT && xx >= x0 && xx < x0 + TT;
          const int q = ((xx - x0 + 4) * TR4) + (zz - z0 + 4);
          const float vz0 = svz[q], vx0 = svx[q];
          float dvz_dz = (kC1 * (vz0 - svz[q - 1]) - kC2 * (svz[q + 1] - svz[q - 2])) * rdz;
          float dvx_dx = (kC1 * (vx0 - svx[q - TR4]) - kC2 * (svx[q + TR4] - svx[q - 2 * TR4])) * rdx;
          float dvx_dz = (kC1 * (svx[q + 1] - vx0) - kC2 * (svx[q + 2] - svx[q - 1])) * rdz;
          float dvz_dx = (kC1 * (svz[q + TR4] - vz0) - kC2 * (svz[q + 2 * TR4] - svz[q - TR4])) * rdx;
          if (zz < g.nPml || zz > g.nzA - g.nPml - 1) {
            const float a = c.bz[zz] * pA.vzz[o] + c.az[zz] * dvz_dz;
            dvz_dz = dvz_dz * (1.f / c.Kz[zz]) + a;
            const float b = c.bzh[zz] * pA.vxz[o] + c.azh[zz] * dvx_dz;
            dvx_dz = dvx_dz * (1.f / c.Kzh[zz]) + b;
            if (own) {
              pB.vzz[o] = a;
              pB.vxz[o] = b;
            }
          }
          if (xx < g.nPml || xx > nx - g.nPml - 1) {
            const float a = c.bx[xx] * pA.vxx[o] + c.ax[xx] * dvx_dx;
            dvx_dx = dvx_dx * (1.f / c.Kx[xx]) + a;
            const float b = c.bxh[xx] * pA.vzx[o] + c.axh[xx] * dvz_dx;
            dvz_dx = dvz_dx * (1.f / c.Kxh[xx]) + b;
            if (own) {
              pB.vxx[o] = a;
              pB.vzx[o] = b;
            }
          }
          if (own && ring) {
            const int slot = ring_slot(zz, xx, g);
            if (slot >= 0) {
              *ring_ptr(ring, g, it, 0, s, slot) = szz;
              *ring_ptr(ring, g, it, 1, s, slot) = sxz;
              *ring_ptr(ring, g, it, 2, s, slot) = sxx;
              *ring_ptr(ring, g, it, 3, s, slot) = vz0;
              *ring_ptr(ring, g, it, 4, s, slot) = vx0;
            }
          }
          const float lam = m.lam[i], l2m = lam + 2.f * m.mu[i];
          szz += (l2m * dvz_dz + lam * dvx_dx) * dt;
          sxx += (lam * dvz_dz + l2m * dvx_dx) * dt;
          sxz += m.amu[i] * (dvx_dz + dvz_dx) * dt;
          if (zz == src.zx[2 * s] && xx == src.zx[2 * s + 1]) {
            const float inj = src.inj[(size_t)s * g.nSteps + it];
            szz += inj;
            sxx += 3.f * inj;
          }
          if (own) {
            B.szz[o] = szz;
            B.sxx[o] = sxx;
            B.sxz[o] = sxz;
          }
        }
      }
      sszz[k] = szz;
      ssxx[k] = sxx;
      ssxz[k] = sxz;
    }
    __syncthreads();
    // ---- v^{n+1} on the tile
    if (updo) {
      const long long o = so + io;
      const int q = (ox + 2) * TR2 + (oz + 2);
      const int q4 = (ox + 4) * TR4 + (oz + 4);
      float dszz_dz = (kC1 * (sszz[q + 1] - sszz[q]) - kC2 * (sszz[q + 2] - sszz[q - 1])) * rdz;
      float dsxz_dx = (kC1 * (ssxz[q] - ssxz[q - TR2]) - kC2 * (ssxz[q + TR2] - ssxz[q - 2 * TR2])) * rdx;
      float dsxz_dz = (kC1 * (ssxz[q] - ssxz[q - 1]) - kC2 * (ssxz[q + 1] - ssxz[q - 2])) * rdz;
      float dsxx_dx = (kC1 * (ssxx[q + TR2] - ssxx[q]) - kC2 * (ssxx[q + 2 * TR2] - ssxx[q - TR2])) * rdx;
      if (zo < g.nPml || zo > g.nzA - g.nPml - 1) {
        const float a = c.bzh[zo] * pA.szz[o] + c.azh[zo] * dszz_dz;
        pB.szz[o] = a;
        dszz_dz = dszz_dz * (1.f / c.Kzh[zo]) + a;
        const float b = c.bz[zo] * pA.sxzz[o] + c.az[zo] * dsxz_dz;
        pB.sxzz[o] = b;
        dsxz_dz = dsxz_dz * (1.f / c.Kz[zo]) + b;
      }
      if (xo < g.nPml || xo > nx - g.nPml) {  // el_velocity.cu:56,71
        const float a = c.bx[xo] * pA.sxzx[o] + c.ax[xo] * dsxz_dx;
        pB.sxzx[o] = a;
        dsxz_dx = dsxz_dx * (1.f / c.Kx[xo]) + a;
        const float b = c.bxh[xo] * pA.sxx[o] + c.axh[xo] * dsxx_dx;
        pB.sxx[o] = b;
        dsxx_dx = dsxx_dx * (1.f / c.Kxh[xo]) + b;
      }
      B.vz[o] = svz[q4] + (dszz_dz + dsxz_dx) * m.ba[io] * dt;
      B.vx[o] = svx[q4] + (dsxz_dz + dsxx_dx) * m.bb[io] * dt;
    }
  }
}

// recording (utilities.cu:557-567): szz + 3 sxx at each receiver, written to
// row `t` of the time-major seismogram array.
__device__ inline void record_at(const Geo& g, const float* __restrict__ szz,
                                 const float* __restrict__ sxx, const int* __restrict__ rec,
                                 int R, float* __restrict__ row, int r) {
  if (r >= R) return;
  const int z = rec[3 * r], x = rec[3 * r + 1], s = rec[3 * r + 2];
  const long long o = s * g.N + (long long)x * g.nz + z;
  row[r] = szz[o] + 3.f * sxx[o];
}

__global__ void k_record(Geo g, const float* __restrict__ szz, const float* __restrict__ sxx,
                         const int* __restrict__ rec, int R, float* __restrict__ row) {
  record_at(g, szz, sxx, rec, R, row, blockIdx.x * blockDim.x + threadIdx.x);
}

// The velocity update of one step with the recording of that step appended as
// extra workgroups of the same grid: recording reads the stresses, which the
// velocity update does not touch, so both run in one launch.  Workgroups
// [0, nV) update velocities (nV a multiple of 8), the rest record one
// receiver per thread.
template <int U>
__global__ __launch_bounds__(256) void k_vel_fwd_rec(int nV, Tiling tl, Geo g, Mdl m, Pml c, Wave w,
                                                    Mem p, float* __restrict__ illum, Geo gall,
                                                    const float* __restrict__ szz,
                                                    const float* __restrict__ sxx,
                                                    const int* __restrict__ rec, int R,
                                                    float* __restrict__ row) {
  const int b = blockIdx.x;
  if (b < nV) {
    if (b < tl.n * tl.chunks) k_vel_fwd_body<U>(b, tl, g, m, c, w, p, illum);
  } else {
    const int per = (int)(blockDim.x * blockDim.y);
    record_at(gall, szz, sxx, rec, R, row,
              (b - nV) * per + (int)(threadIdx.y * blockDim.x + threadIdx.x));
  }
}

// Residual (libCUFD.cu:258-313 without the optional filters): window obs and
// syn with w^2, res = obs_w - syn_w (first sample 0, gpuMinus utilities.cu:154-167),
// sum res^2 (cuda_cal_objective, here in FP64), then window res again.
__global__ __launch_bounds__(256) void k_residual(int nSteps, int R, const float* __restrict__ w2,
                                                  const float* __restrict__ obs,
                                                  const float* __restrict__ syn,
                                                  float* __restrict__ res,
                                                  double* __restrict__ partial) {
  __shared__ double red[256];
  const size_t total = (size_t)nSteps * R;
  double acc = 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(k / (size_t)R);
    float rv;
    {
      // obs and syn are windowed separately (two roundings, as the reference's
      // separate cuda_window launches do); an FMA contraction of
      // obs*ww - syn*ww would leave a rounding residue when obs == syn.
#pragma clang fp contract(off)
      const float ww = w2[t];
      const float ow = obs[k] * ww;
      const float sw = syn[k] * ww;
      rv = t > 0 ? ow - sw : 0.f;
    }
    const float ww = w2[t];
    acc += (double)rv * (double)rv;
    res[k] = rv * ww;
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

// ------------------------------------------------------------------ backward
// Reverse-time re-materialisation (el_velocity.cu:84-117 with isFor = false)
// fused with the ring restore of v (to_bnd, utilities.cu:394-424) and the
// density-gradient accumulation.  The thread domain is the ring rectangle;
// interior cells update, ring cells take the stored values.
// acc_a += vz_adj (dszz/dz + dsxz/dx), acc_b += vx_adj (dsxz/dz + dsxx/dx).
// Shots are batched U at a time as in the forward kernels.
struct RevVelIn {
  float szzp1, szz0, szzp2, szzm1;
  float sxz0, sxzmx, sxzpx, sxzm2x, sxzm1, sxzp1, sxzm2;
  float sxxpx, sxx0, sxxp2x, sxxmx;
  float avz, avx, vz, vx;
};

template <int U>
__device__ __forceinline__ void k_rev_vel_body(int bid, Tiling tl, Geo g, Mdl m, Wave w, Wave a,
                                                 float* __restrict__ acc_a,
                                                 float* __restrict__ acc_b,
                                                 const float* __restrict__ ring, int it) {
  int tz, tx;
  tile_of(tl, bid, tz, tx);
  const int z = tz * kBZ + threadIdx.x + (g.nPml - 2);
  const int x = tx * (int)blockDim.y + threadIdx.y + (g.nPml - 2);
  if (z > g.nzA - g.nPml + 1 || x > g.nx - g.nPml + 1) return;
  const bool in = z >= g.nPml && z <= g.nzA - 1 - g.nPml && x >= g.nPml && x <= g.nx - 1 - g.nPml;
  const int slot = ring_slot(z, x, g);
  if (!in && slot < 0) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const float ba = m.ba[i], bb = m.bb[i];
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  float ga = 0.f, gb = 0.f;
  for (int s0 = 0; s0 < g.ns; s0 += U) {
    RevVelIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= g.ns) break;
      const long long so = s * g.N;
      RevVelIn& q = r[u];
      if (in) {
        const float* szz = w.szz + so;
        const float* sxx = w.sxx + so;
        const float* sxz = w.sxz + so;
        q.szzp1 = ldo(szz + 1, bo), q.szz0 = ldo(szz, bo), q.szzp2 = ldo(szz + 2, bo), q.szzm1 = ldo(szz - 1, bo);
        q.sxz0 = ldo(sxz, bo), q.sxzmx = ldo(sxz - nz, bo), q.sxzpx = ldo(sxz + nz, bo), q.sxzm2x = ldo(sxz - 2 * nz, bo);
        q.sxzm1 = ldo(sxz - 1, bo), q.sxzp1 = ldo(sxz + 1, bo), q.sxzm2 = ldo(sxz - 2, bo);
        q.sxxpx = ldo(sxx + nz, bo), q.sxx0 = ldo(sxx, bo), q.sxxp2x = ldo(sxx + 2 * nz, bo), q.sxxmx = ldo(sxx - nz, bo);
        q.avz = ldo(a.vz + so, bo), q.avx = ldo(a.vx + so, bo);
      }
      if (slot < 0) {
        q.vz = ldo(w.vz + so, bo), q.vx = ldo(w.vx + so, bo);
      } else {
        q.vz = *ring_ptr(const_cast<float*>(ring), g, it, 3, s, slot);
        q.vx = *ring_ptr(const_cast<float*>(ring), g, it, 4, s, slot);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= g.ns) break;
      const long long so = s * g.N;
      const RevVelIn& q = r[u];
      float vz = q.vz, vx = q.vx;
      if (in) {
        const float dszz_dz = (kC1 * (q.szzp1 - q.szz0) - kC2 * (q.szzp2 - q.szzm1)) * rdz;
        const float dsxz_dx = (kC1 * (q.sxz0 - q.sxzmx) - kC2 * (q.sxzpx - q.sxzm2x)) * rdx;
        const float dsxz_dz = (kC1 * (q.sxz0 - q.sxzm1) - kC2 * (q.sxzp1 - q.sxzm2)) * rdz;
        const float dsxx_dx = (kC1 * (q.sxxpx - q.sxx0) - kC2 * (q.sxxp2x - q.sxxmx)) * rdx;
        const float fz = dszz_dz + dsxz_dx, fx = dsxz_dz + dsxx_dx;
        ga += q.avz * fz;
        gb += q.avx * fx;
        if (slot < 0) {
          vz -= fz * ba * dt;
          vx -= fx * bb * dt;
        }
      }
      sto(w.vz + so, bo, vz);
      sto(w.vx + so, bo, vx);
    }
  }
  if (in) {
    acc_a[i] += ga;
    acc_b[i] += gb;
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_rev_vel(Tiling tl, Geo g, Mdl m, Wave w, Wave a,
                                                 float* __restrict__ acc_a,
                                                 float* __restrict__ acc_b,
                                                 const float* __restrict__ ring, int it) {
  k_rev_vel_body<U>(blockIdx.x, tl, g, m, w, a, acc_a, acc_b, ring, it);
}

// el_stress.cu:90-131 (isFor = false) fused with add_source(-) and the ring
// restore of sigma.  acc_l += (szz_adj + sxx_adj)(dvz/dz + dvx/dx),
// acc_m += szz_adj dvz/dz + sxx_adj dvx/dx, acc_s += sxz_adj (dvx/dz + dvz/dx).
struct RevStressIn {
  float vz0, vzm1, vzp1, vzm2, vzpx, vzp2x, vzmx;
  float vx0, vxmx, vxpx, vxm2x, vxp1, vxp2, vxm1;
  float aszz, asxx, asxz;
  float szz, sxx, sxz;
  int sz, sx;
};

template <int U>
__device__ __forceinline__ void k_rev_stress_body(int bid, Tiling tl, Geo g, Mdl m, Wave w, Wave a, Src src,
                                                    float* __restrict__ acc_l,
                                                    float* __restrict__ acc_m,
                                                    float* __restrict__ acc_s,
                                                    const float* __restrict__ ring, int it) {
  int tz, tx;
  tile_of(tl, bid, tz, tx);
  const int z = tz * kBZ + threadIdx.x + (g.nPml - 2);
  const int x = tx * (int)blockDim.y + threadIdx.y + (g.nPml - 2);
  if (z > g.nzA - g.nPml + 1 || x > g.nx - g.nPml + 1) return;
  const bool in = z >= g.nPml && z <= g.nzA - 1 - g.nPml && x >= g.nPml && x <= g.nx - 1 - g.nPml;
  const int slot = ring_slot(z, x, g);
  if (!in && slot < 0) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i];
  const float l2m = lam + 2.f * mu;
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  float gl = 0.f, gm = 0.f, gs = 0.f;
  for (int s0 = 0; s0 < g.ns; s0 += U) {
    RevStressIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= g.ns) break;
      const long long so = s * g.N;
      RevStressIn& q = r[u];
      if (in) {
        const float* vz = w.vz + so;
        const float* vx = w.vx + so;
        q.vz0 = ldo(vz, bo), q.vzm1 = ldo(vz - 1, bo), q.vzp1 = ldo(vz + 1, bo), q.vzm2 = ldo(vz - 2, bo);
        q.vzpx = ldo(vz + nz, bo), q.vzp2x = ldo(vz + 2 * nz, bo), q.vzmx = ldo(vz - nz, bo);
        q.vx0 = ldo(vx, bo), q.vxmx = ldo(vx - nz, bo), q.vxpx = ldo(vx + nz, bo), q.vxm2x = ldo(vx - 2 * nz, bo);
        q.vxp1 = ldo(vx + 1, bo), q.vxp2 = ldo(vx + 2, bo), q.vxm1 = ldo(vx - 1, bo);
        q.aszz = ldo(a.szz + so, bo), q.asxx = ldo(a.sxx + so, bo), q.asxz = ldo(a.sxz + so, bo);
      }
      if (slot < 0) {
        q.szz = ldo(w.szz + so, bo), q.sxx = ldo(w.sxx + so, bo), q.sxz = ldo(w.sxz + so, bo);
        q.sz = src.zx[2 * s], q.sx = src.zx[2 * s + 1];
      } else {
        q.szz = *ring_ptr(const_cast<float*>(ring), g, it, 0, s, slot);
        q.sxz = *ring_ptr(const_cast<float*>(ring), g, it, 1, s, slot);
        q.sxx = *ring_ptr(const_cast<float*>(ring), g, it, 2, s, slot);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= g.ns) break;
      const long long so = s * g.N;
      const RevStressIn& q = r[u];
      float szz = q.szz, sxx = q.sxx, sxz = q.sxz;
      if (in) {
        const float dvz_dz = (kC1 * (q.vz0 - q.vzm1) - kC2 * (q.vzp1 - q.vzm2)) * rdz;
        const float dvx_dx = (kC1 * (q.vx0 - q.vxmx) - kC2 * (q.vxpx - q.vxm2x)) * rdx;
        const float dvx_dz = (kC1 * (q.vxp1 - q.vx0) - kC2 * (q.vxp2 - q.vxm1)) * rdz;
        const float dvz_dx = (kC1 * (q.vzpx - q.vz0) - kC2 * (q.vzp2x - q.vzmx)) * rdx;
        gl += (q.aszz + q.asxx) * (dvz_dz + dvx_dx);
        gm += q.aszz * dvz_dz + q.asxx * dvx_dx;
        gs += q.asxz * (dvx_dz + dvz_dx);
        if (slot < 0) {
          if (z == q.sz && x == q.sx) {
            const float inj = src.inj[(size_t)s * g.nSteps + it];
            szz -= inj;
            sxx -= 3.f * inj;
          }
          szz = szz - (l2m * dvz_dz + lam * dvx_dx) * dt;
          sxx = sxx - (lam * dvz_dz + l2m * dvx_dx) * dt;
          sxz -= amu * (dvx_dz + dvz_dx) * dt;
        }
      }
      sto(w.szz + so, bo, szz);
      sto(w.sxx + so, bo, sxx);
      sto(w.sxz + so, bo, sxz);
    }
  }
  if (in) {
    acc_l[i] += gl;
    acc_m[i] += gm;
    acc_s[i] += gs;
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_rev_stress(Tiling tl, Geo g, Mdl m, Wave w, Wave a, Src src,
                                                    float* __restrict__ acc_l,
                                                    float* __restrict__ acc_m,
                                                    float* __restrict__ acc_s,
                                                    const float* __restrict__ ring, int it) {
  k_rev_stress_body<U>(blockIdx.x, tl, g, m, w, a, src, acc_l, acc_m, acc_s, ring, it);
}

// el_velocity_adj.cu:22-108, plus source_grad (utilities.cu:582-593) for step
// `it` (the adjoint stress it reads is not touched by this kernel).  The
// C-PML memory terms a*dpsi are evaluated only where a != 0; elsewhere they
// are exactly zero in the reference as well.
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
__device__ __forceinline__ void k_adj_vel_body(int bid, Tiling tl, Geo g, Mdl m, Pml c, Wave a, Mem p, Src src,
                                                 float* __restrict__ stf_grad, int it) {
  int tz, tx;
  const int chunk = tile_of(tl, bid, tz, tx);
  const int sb = chunk * tl.spt, se = min(g.ns, sb + tl.spt);  // shots of this workgroup
  const int z = tz * kBZ + threadIdx.x;
  const int x = tx * (int)blockDim.y + threadIdx.y;
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i], ba = m.ba[i], bb = m.bb[i];
  const float l2m = lam + 2.f * mu;
  const bool pz = z < g.nPml || z > g.nzA - g.nPml - 1;
  const bool px = x < g.nPml || x > g.nx - g.nPml - 1;
  const float ax = c.ax[x], rkx = 1.f / c.Kx[x], bx = c.bx[x];
  const float axh = c.axh[x], rkxh = 1.f / c.Kxh[x], bxh = c.bxh[x];
  const float az = c.az[z], rkz = 1.f / c.Kz[z], bz = c.bz[z];
  const float azh = c.azh[z], rkzh = 1.f / c.Kzh[z], bzh = c.bzh[z];
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  for (int s0 = sb; s0 < se; s0 += U) {
    AdjVelIn r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const float* szz = a.szz + so;
      const float* sxx = a.sxx + so;
      const float* sxz = a.sxz + so;
      AdjVelIn& q = r[u];
      q.szz0 = ldo(szz, bo), q.szzpx = ldo(szz + nz, bo), q.szzp2x = ldo(szz + 2 * nz, bo), q.szzmx = ldo(szz - nz, bo);
      q.szzp1 = ldo(szz + 1, bo), q.szzp2 = ldo(szz + 2, bo), q.szzm1 = ldo(szz - 1, bo);
      q.sxx0 = ldo(sxx, bo), q.sxxpx = ldo(sxx + nz, bo), q.sxxp2x = ldo(sxx + 2 * nz, bo), q.sxxmx = ldo(sxx - nz, bo);
      q.sxxp1 = ldo(sxx + 1, bo), q.sxxp2 = ldo(sxx + 2, bo), q.sxxm1 = ldo(sxx - 1, bo);
      q.sxz0 = ldo(sxz, bo), q.sxzm1 = ldo(sxz - 1, bo), q.sxzp1 = ldo(sxz + 1, bo), q.sxzm2 = ldo(sxz - 2, bo);
      q.sxzmx = ldo(sxz - nz, bo), q.sxzpx = ldo(sxz + nz, bo), q.sxzm2x = ldo(sxz - 2 * nz, bo);
      if (ax != 0.f) {
        const float* v = p.vxx + so;
        q.q1[0] = ldo(v + nz, bo), q.q1[1] = ldo(v, bo), q.q1[2] = ldo(v + 2 * nz, bo), q.q1[3] = ldo(v - nz, bo);
      }
      if (azh != 0.f) {
        const float* v = p.vxz + so;
        q.q2[0] = ldo(v, bo), q.q2[1] = ldo(v - 1, bo), q.q2[2] = ldo(v + 1, bo), q.q2[3] = ldo(v - 2, bo);
      }
      if (az != 0.f) {
        const float* v = p.vzz + so;
        q.q3[0] = ldo(v + 1, bo), q.q3[1] = ldo(v, bo), q.q3[2] = ldo(v + 2, bo), q.q3[3] = ldo(v - 1, bo);
      }
      if (axh != 0.f) {
        const float* v = p.vzx + so;
        q.q4[0] = ldo(v, bo), q.q4[1] = ldo(v - nz, bo), q.q4[2] = ldo(v + nz, bo), q.q4[3] = ldo(v - 2 * nz, bo);
      }
      if (px) q.m_sxx = ldo(p.sxx + so, bo), q.m_sxzx = ldo(p.sxzx + so, bo);
      if (pz) q.m_sxzz = ldo(p.sxzz + so, bo), q.m_szz = ldo(p.szz + so, bo);
      q.vx = ldo(a.vx + so, bo), q.vz = ldo(a.vz + so, bo);
      q.sz = src.zx[2 * s], q.sx = src.zx[2 * s + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const AdjVelIn& q = r[u];
      if (it >= 0 && z == q.sz && x == q.sx)
        stf_grad[(size_t)s * g.nSteps + it] = -(q.szz0 + 3.f * q.sxx0) * dt;
      // update vx
      const float dszz_dx = (-kC1 * (q.szzpx - q.szz0) + kC2 * (q.szzp2x - q.szzmx)) * rdx;
      const float dsxx_dx = (-kC1 * (q.sxxpx - q.sxx0) + kC2 * (q.sxxp2x - q.sxxmx)) * rdx;
      const float dsxz_dz = (-kC1 * (q.sxz0 - q.sxzm1) + kC2 * (q.sxzp1 - q.sxzm2)) * rdz;
      float vx = q.vx + (lam * dszz_dx * rkx * dt + l2m * dsxx_dx * rkx * dt +
                         amu * rkzh * dsxz_dz * dt);
      if (ax != 0.f) vx += ax * ((-kC1 * (q.q1[0] - q.q1[1]) + kC2 * (q.q1[2] - q.q1[3])) * rdx);
      if (azh != 0.f) vx += azh * ((-kC1 * (q.q2[0] - q.q2[1]) + kC2 * (q.q2[2] - q.q2[3])) * rdz);
      sto(a.vx + so, bo, vx);
      if (px) sto(p.sxx + so, bo, bxh * q.m_sxx + bb * vx * dt);
      if (pz) sto(p.sxzz + so, bo, bz * q.m_sxzz + bb * vx * dt);
      // update vz
      const float dszz_dz = (-kC1 * (q.szzp1 - q.szz0) + kC2 * (q.szzp2 - q.szzm1)) * rdz;
      const float dsxx_dz = (-kC1 * (q.sxxp1 - q.sxx0) + kC2 * (q.sxxp2 - q.sxxm1)) * rdz;
      const float dsxz_dx = (-kC1 * (q.sxz0 - q.sxzmx) + kC2 * (q.sxzpx - q.sxzm2x)) * rdx;
      float vz = q.vz + (l2m * dszz_dz * rkz * dt + lam * dsxx_dz * rkz * dt +
                         amu * rkxh * dsxz_dx * dt);
      if (az != 0.f) vz += az * ((-kC1 * (q.q3[0] - q.q3[1]) + kC2 * (q.q3[2] - q.q3[3])) * rdz);
      if (axh != 0.f) vz += axh * ((-kC1 * (q.q4[0] - q.q4[1]) + kC2 * (q.q4[2] - q.q4[3])) * rdx);
      sto(a.vz + so, bo, vz);
      if (px) sto(p.sxzx + so, bo, bx * q.m_sxzx + ba * vz * dt);
      if (pz) sto(p.szz + so, bo, bzh * q.m_szz + ba * vz * dt);
    }
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_adj_vel(Tiling tl, Geo g, Mdl m, Pml c, Wave a, Mem p, Src src,
                                                 float* __restrict__ stf_grad, int it) {
  k_adj_vel_body<U>(blockIdx.x, tl, g, m, c, a, p, src, stf_grad, it);
}

// res_injection (utilities.cu:569-580): szz_adj += res, sxx_adj += 3 res.
__global__ void k_res_inject(Geo g, float* __restrict__ szz, float* __restrict__ sxx,
                             const int* __restrict__ rec, int R, const float* __restrict__ row) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int z = rec[3 * r], x = rec[3 * r + 1], s = rec[3 * r + 2];
  const long long o = s * g.N + (long long)x * g.nz + z;
  const float v = row[r];
  atomicAdd(szz + o, v);
  atomicAdd(sxx + o, 3.f * v);
}

// el_stress_adj.cu:22-104.  The 4 stress-side memory variables are updated
// without PML gates in the reference; they are read only through a*dpsi terms,
// i.e. where a != 0 (the C-PML layers) plus the 2-cell reach of the stencil,
// so they are kept in that band (`zlive`/`xlive`) and nowhere else.
// Residual injection folded into the adjoint stress update (res_injection,
// utilities.cu:569-580): at a receiver's cell szz_adj += res and sxx_adj +=
// 3 res before the update reads them (the update reads only its own cell's
// stresses, so the sum is the same as a separate injection pass).  Used when
// every receiver lies in one row z and no two receivers of a shot share a
// cell: at[shot][x] is then the receiver of column x (index into the
// residual row) or -1.
struct Inj {
  const float* row;  // residual row of the step, all receivers; null: no injection
  const int* at;     // [ns][nx]
  int z;
};

struct AdjStressIn {
  float vz0, vzpx, vzp2x, vzmx, vzm1, vzp1, vzm2;
  float vx0, vxp1, vxp2, vxm1, vxmx, vxpx, vxm2x;
  float q1[4], q2[4], q3[4], q4[4];  // sxzx (x), sxzz (z), sxx (x), szz (z) stencils
  float m_vzx, m_vxz, m_vxx, m_vzz;
  float sxz, sxx, szz;
};

template <int U>
__device__ __forceinline__ void k_adj_stress_body(int bid, Tiling tl, Geo g, Mdl m, Pml c, Wave a, Mem p,
                                                  Inj inj) {
  int tz, tx;
  const int chunk = tile_of(tl, bid, tz, tx);
  const int sb = chunk * tl.spt, se = min(g.ns, sb + tl.spt);  // shots of this workgroup
  const int z = tz * kBZ + threadIdx.x;
  const int x = tx * (int)blockDim.y + threadIdx.y;
  if (z < 2 || z > g.nzA - 3 || x < 2 || x > g.nx - 3) return;
  const int nz = g.nz;
  const long long i = (long long)x * nz + z;
  const unsigned bo = (unsigned)i * 4u;  // byte offset of the cell in a per-shot array
  const float lam = m.lam[i], mu = m.mu[i], amu = m.amu[i], ba = m.ba[i], bb = m.bb[i];
  const float l2m = lam + 2.f * mu;
  const bool zlive = z <= g.nPml + 1 || z >= g.nzA - g.nPml - 2;
  const bool xlive = x <= g.nPml + 1 || x >= g.nx - g.nPml - 2;
  const float ax = c.ax[x], rkx = 1.f / c.Kx[x], bx = c.bx[x];
  const float axh = c.axh[x], rkxh = 1.f / c.Kxh[x], bxh = c.bxh[x];
  const float az = c.az[z], rkz = 1.f / c.Kz[z], bz = c.bz[z];
  const float azh = c.azh[z], rkzh = 1.f / c.Kzh[z], bzh = c.bzh[z];
  const float dt = g.dt, rdz = g.rdz, rdx = g.rdx;
  const bool inj_row = inj.row && z == inj.z;
  for (int s0 = sb; s0 < se; s0 += U) {
    AdjStressIn r[U];
    float resv[U];
    bool has[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      resv[u] = 0.f;
      has[u] = false;
      if (s >= se) break;
      if (inj_row) {
        const int k = inj.at[(size_t)(g.s0 + s) * g.nx + x];
        if (k >= 0) {
          resv[u] = inj.row[k];
          has[u] = true;
        }
      }
      const long long so = s * g.N;
      const float* vz = a.vz + so;
      const float* vx = a.vx + so;
      AdjStressIn& q = r[u];
      q.vz0 = ldo(vz, bo), q.vzpx = ldo(vz + nz, bo), q.vzp2x = ldo(vz + 2 * nz, bo), q.vzmx = ldo(vz - nz, bo);
      q.vzm1 = ldo(vz - 1, bo), q.vzp1 = ldo(vz + 1, bo), q.vzm2 = ldo(vz - 2, bo);
      q.vx0 = ldo(vx, bo), q.vxp1 = ldo(vx + 1, bo), q.vxp2 = ldo(vx + 2, bo), q.vxm1 = ldo(vx - 1, bo);
      q.vxmx = ldo(vx - nz, bo), q.vxpx = ldo(vx + nz, bo), q.vxm2x = ldo(vx - 2 * nz, bo);
      if (ax != 0.f) {
        const float* v = p.sxzx + so;
        q.q1[0] = ldo(v + nz, bo), q.q1[1] = ldo(v, bo), q.q1[2] = ldo(v + 2 * nz, bo), q.q1[3] = ldo(v - nz, bo);
      }
      if (az != 0.f) {
        const float* v = p.sxzz + so;
        q.q2[0] = ldo(v + 1, bo), q.q2[1] = ldo(v, bo), q.q2[2] = ldo(v + 2, bo), q.q2[3] = ldo(v - 1, bo);
      }
      if (axh != 0.f) {
        const float* v = p.sxx + so;
        q.q3[0] = ldo(v, bo), q.q3[1] = ldo(v - nz, bo), q.q3[2] = ldo(v + nz, bo), q.q3[3] = ldo(v - 2 * nz, bo);
      }
      if (azh != 0.f) {
        const float* v = p.szz + so;
        q.q4[0] = ldo(v, bo), q.q4[1] = ldo(v - 1, bo), q.q4[2] = ldo(v + 1, bo), q.q4[3] = ldo(v - 2, bo);
      }
      if (xlive) q.m_vzx = ldo(p.vzx + so, bo), q.m_vxx = ldo(p.vxx + so, bo);
      if (zlive) q.m_vxz = ldo(p.vxz + so, bo), q.m_vzz = ldo(p.vzz + so, bo);
      q.sxz = ldo(a.sxz + so, bo), q.sxx = ldo(a.sxx + so, bo), q.szz = ldo(a.szz + so, bo);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u;
      if (s >= se) break;
      const long long so = s * g.N;
      const AdjStressIn& q = r[u];
      const float dvz_dx = (-kC1 * (q.vzpx - q.vz0) + kC2 * (q.vzp2x - q.vzmx)) * rdx;
      const float dvx_dz = (-kC1 * (q.vxp1 - q.vx0) + kC2 * (q.vxp2 - q.vxm1)) * rdz;
      float sxz = q.sxz + (dvz_dx * rkx * ba * dt + dvx_dz * rkz * bb * dt);
      if (ax != 0.f) sxz += ax * ((-kC1 * (q.q1[0] - q.q1[1]) + kC2 * (q.q1[2] - q.q1[3])) * rdx);
      if (az != 0.f) sxz += az * ((-kC1 * (q.q2[0] - q.q2[1]) + kC2 * (q.q2[2] - q.q2[3])) * rdz);
      sto(a.sxz + so, bo, sxz);
      if (xlive) sto(p.vzx + so, bo, bxh * q.m_vzx + sxz * amu * dt);
      if (zlive) sto(p.vxz + so, bo, bzh * q.m_vxz + sxz * amu * dt);
      const float dvx_dx = (-kC1 * (q.vx0 - q.vxmx) + kC2 * (q.vxpx - q.vxm2x)) * rdx;
      const float dvz_dz = (-kC1 * (q.vz0 - q.vzm1) + kC2 * (q.vzp1 - q.vzm2)) * rdz;
      float sxx0 = q.sxx, szz0 = q.szz;
      if (has[u]) {
        // 3 res rounded before the add, as the separate injection adds it
#pragma clang fp contract(off)
        const float v3 = 3.f * resv[u];
        szz0 += resv[u];
        sxx0 += v3;
      }
      float sxx = sxx0 + bb * dvx_dx * rkxh * dt;
      if (axh != 0.f) sxx += axh * ((-kC1 * (q.q3[0] - q.q3[1]) + kC2 * (q.q3[2] - q.q3[3])) * rdx);
      float szz = szz0 + ba * dvz_dz * rkzh * dt;
      if (azh != 0.f) szz += azh * ((-kC1 * (q.q4[0] - q.q4[1]) + kC2 * (q.q4[2] - q.q4[3])) * rdz);
      sto(a.sxx + so, bo, sxx);
      sto(a.szz + so, bo, szz);
      if (xlive) sto(p.vxx + so, bo, bx * q.m_vxx + lam * szz * dt + l2m * sxx * dt);
      if (zlive) sto(p.vzz + so, bo, bz * q.m_vzz + l2m * szz * dt + lam * sxx * dt);
    }
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_adj_stress(Tiling tl, Geo g, Mdl m, Pml c, Wave a, Mem p,
                                                   Inj inj) {
  k_adj_stress_body<U>(blockIdx.x, tl, g, m, c, a, p, inj);
}

// Interior checkpoints of the reverse sweep (no reference counterpart): the 5
// components on the interior rectangle [nPml, nzA-1-nPml] x [nPml, nx-1-nPml]
// are saved before forward step it and written back after reverse step it,
// for it a positive multiple of the interval, which bounds the drift of the
// time-reversed re-materialisation.  Layout [shot][checkpoint][component]
// [interior cell, z fastest].  Thread domain: the interior rectangle.
__global__ __launch_bounds__(256) void k_ckpt(Tiling tl, Geo g, Wave w, float* __restrict__ ck,
                                             int slot, int nslots, int save) {
  int tz, tx;
  tile_of(tl, tz, tx);
  const int Ih = g.nzA - 2 * g.nPml, Iw = g.nx - 2 * g.nPml;
  const int zi = tz * kBZ + threadIdx.x;
  const int xi = tx * (int)blockDim.y + threadIdx.y;
  if (zi >= Ih || xi >= Iw) return;
  const long long i = (long long)(xi + g.nPml) * g.nz + (zi + g.nPml);
  const size_t Ic = (size_t)Ih * Iw, j = (size_t)xi * Ih + zi;
  float* f[5] = {w.vz, w.vx, w.szz, w.sxx, w.sxz};
  for (int s = 0; s < g.ns; ++s) {
    float* b = ck + ((size_t)s * nslots + slot) * 5 * Ic + j;
    const long long o = s * g.N + i;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      if (save)
        b[q * Ic] = f[q][o];
      else
        f[q][o] = b[q * Ic];
    }
  }
}

// Backward kernel pairs with disjoint reads and writes, launched as one grid
// so they share a launch and its ramp-up and drain.  The grid holds nA
// workgroups of the first kernel and nB of the second (both multiples of 8,
// so both parts see the same workgroup-to-XCD dealing).  With `mix` set,
// rounds of 8 workgroups (one per XCD) alternate between the parts in
// proportion to their sizes, so both sweep their tiles at the same pace and
// the model parameters and shared adjoint values that both read are fetched
// close together in time, one L2 fill serving both; otherwise the first nA
// workgroups run the first kernel.
//   pair a: rev_stress(it) + adj_vel(it) (rev_stress reads the adjoint stress
//           that adj_vel also only reads; adj_vel writes the adjoint velocity
//           that rev_stress does not touch)
//   pair b: adj_stress(it) + rev_vel(it-1) (rev_vel reads the adjoint velocity
//           that adj_stress also only reads; adj_stress writes the adjoint
//           stress that rev_vel does not touch)
__device__ inline bool pair_part(int b, int nA, int nB, int mix, int& local) {
  if (!mix) {
    local = b < nA ? b : b - nA;
    return b < nA;
  }
  const long long RA = nA / 8, R = RA + nB / 8, r = b / 8;
  const int before = (int)(r * RA / R);  // first-part rounds before round r
  if ((int)((r + 1) * RA / R) > before) {
    local = before * 8 + b % 8;
    return true;
  }
  local = (int)(r - before) * 8 + b % 8;
  return false;
}

template <int U>
__global__ __launch_bounds__(256) void k_bwd_pair_a(int nA, int nB, int mix, Tiling tR, Tiling tU,
                                                   Geo g, Mdl m, Pml c, Wave w, Wave a, Mem p,
                                                   Src src, float* __restrict__ acc,
                                                   const float* __restrict__ ring,
                                                   float* __restrict__ stf_grad, int it) {
  int b;
  if (pair_part(blockIdx.x, nA, nB, mix, b)) {
    if (b < tR.n)
      k_rev_stress_body<U>(b, tR, g, m, w, a, src, acc, acc + g.N, acc + 2 * g.N, ring, it);
  } else if (b < tU.n * tU.chunks) {
    k_adj_vel_body<U>(b, tU, g, m, c, a, p, src, stf_grad, it);
  }
}

template <int U>
__global__ __launch_bounds__(256) void k_bwd_pair_b(int nA, int nB, int mix, Tiling tU, Tiling tR,
                                                   Geo g, Mdl m, Pml c, Wave w, Wave a, Mem p,
                                                   float* __restrict__ acc,
                                                   const float* __restrict__ ring, int it_rev,
                                                   Inj inj) {
  int b;
  if (pair_part(blockIdx.x, nA, nB, mix, b)) {
    if (b < tU.n * tU.chunks) k_adj_stress_body<U>(b, tU, g, m, c, a, p, inj);
  } else if (b < tR.n) {
    k_rev_vel_body<U>(b, tR, g, m, w, a, acc + 3 * g.N, acc + 4 * g.N, ring, it_rev);
  }
}

// Chain rule from the accumulated derivatives to (lambda, mu, rho), once per
// run.  Deterministic gather form of the reference's atomicAdd sprays
// (el_stress.cu:113-124, el_velocity.cu:234-243), with the same target
// conditions (including the always-true `gidx+1<=gidx<=nx-1-nPml`).
__global__ __launch_bounds__(1024) void k_finalize(Tiling tl, Geo g, Mdl m, const float* __restrict__ acc,
                                                  float* __restrict__ grad) {
  int tz, tx;
  tile_of(tl, tz, tx);
  const int z = tz * kBZ + threadIdx.x;
  const int x = tx * (int)blockDim.y + threadIdx.y;
  if (z >= g.nz || x >= g.nx) return;
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
  if (inI(z - 1, x) && z <= z1 && m.amu[i - 1] != 0.f) t += shear(z - 1, x), any = true;
  if (inI(z, x - 1) && m.amu[i - nz] != 0.f) t += shear(z, x - 1), any = true;
  if (inI(z - 1, x - 1) && z <= z1 && x <= x1 && m.amu[i - nz - 1] != 0.f)
    t += shear(z - 1, x - 1), any = true;
  if (any) {
    const float mu = m.mu[i];
    gm += t / (mu * mu);
  }
  auto ga = [&](long long j) { const float b = m.ba[j]; return acc_a[j] * dt * (0.5f * b * b); };
  auto gb = [&](long long j) { const float b = m.bb[j]; return acc_b[j] * dt * (0.5f * b * b); };
  if (inI(z, x)) gd += ga(i) + gb(i);
  if (inI(z - 1, x) && z <= z1) gd += ga(i - 1);
  if (inI(z, x - 1)) gd += gb(i - nz);
  grad[i] = gl;
  grad[N + i] = gm;
  grad[2 * N + i] = gd;
}

// aveMuInit / aveBycInit (utilities.cu:125-152) over [2, nz-3] x [2, nx-3];
// defaults 0 and 1/1000 elsewhere (Model.cu:67-73).
__global__ void k_model_init(Tiling tl, Geo g, const float* __restrict__ mu, const float* __restrict__ den,
                             float* __restrict__ amu, float* __restrict__ ba,
                             float* __restrict__ bb) {
  int tz, tx;
  tile_of(tl, tz, tx);
  const int z = tz * kBZ + threadIdx.x;
  const int x = tx * (int)blockDim.y + threadIdx.y;
  if (z >= g.nz || x >= g.nx) return;
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

// ===================================================================== plan
static Tiling make_tiling(int nz_dom, int nx_dom, int bx, bool xcd) {
  Tiling t;
  t.gz = (nz_dom + kBZ - 1) / kBZ;
  t.gx = (nx_dom + bx - 1) / bx;
  t.n = t.gz * t.gx;
  t.xcd = xcd ? 1 : 0;
  return t;
}

// Workgroup -> tile order for an XCD-banded launch whose first `heavy_left`
// and last `heavy_right` tile columns are heavier (the x-direction C-PML):
// workgroup b runs on XCD b % 8, and each XCD gets one contiguous run of the
// heavy tiles (run first) followed by one contiguous run of the others, so
// every XCD carries the same share of heavy work and neighbouring tiles stay
// on one XCD's L2.  Tile id = column * gz + row (rows fastest).
static std::vector<int> balanced_tile_order(int gz, int gx, int heavy_left, int heavy_right) {
  const int n = gz * gx;
  std::vector<int> heavy, light;
  for (int bx = 0; bx < gx; bx++) {
    const bool h = bx < heavy_left || bx >= gx - heavy_right;
    for (int bz = 0; bz < gz; bz++) (h ? heavy : light).push_back(bx * gz + bz);
  }
  std::vector<int> perm(n);
  size_t ho = 0, lo = 0;
  const size_t H = heavy.size();
  for (int k = 0; k < 8; k++) {
    const int len = n > k ? (n - k + 7) / 8 : 0;  // workgroups b < n with b % 8 == k
    const size_t hk = std::min((size_t)len, H * (k + 1) / 8 - H * k / 8);
    int j = 0;
    for (size_t q = 0; q < hk; q++) perm[(size_t)(j++) * 8 + k] = heavy[ho++];
    while (j < len) {
      if (lo < light.size())
        perm[(size_t)(j++) * 8 + k] = light[lo++];
      else
        perm[(size_t)(j++) * 8 + k] = heavy[ho++];
    }
  }
  return perm;
}

// threads of a fused workgroup: whole wave64s covering the active rows
static int fused_threads(const Geo& g) { return ((g.nzA + 63) / 64) * 64; }

static int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

struct DeviceState {
  int dev = 0;
  hipStream_t stream = nullptr;
  Geo g{};
  int R = 0;
  int ckpt_every = 0;     // interior checkpoint interval of the reverse sweep (0: none)
  float* ck = nullptr;    // interior checkpoints [shot][slot][5][interior]
  size_t ck_floats = 0;
  int tile_x = 4;    // x extent of a workgroup tile (blockDim.y)
  bool xcd = true;   // XCD-banded workgroup order
  bool fused = false; // column-marching fused kernels (FWIFLOW_FUSED=1)
  bool tiled = false;  // tiled fused forward step (FWIFLOW_TILED)
  int seg_x = 32;    // x extent of a fused segment
  int rev_group = 8; // shots per workgroup of the fused reverse step
  bool fused_rev = true;  // fused reverse step (needs its LDS to fit)
  int unroll = 1;         // shots whose loads a split-kernel thread issues together
  int unroll_fwd = 2;     // ... in the forward kernels
  int shot_group = -1;    // shots per launch of the split/tiled kernels (0: all, -1: auto)
  int shot_group_bwd = -1;  // the same for the backward sweep (0: all, -1: auto)
  // Two streams (default priority both): shot groups alternate between them,
  // so one stream's kernel fills the other's launch gap; the backward groups
  // of the second stream accumulate into acc + 5N.
  int streams = 1;
  hipStream_t stream2 = nullptr;
  hipEvent_t join[2] = {nullptr, nullptr};
  int spt = 0;            // shots per thread of the forward/adjoint split kernels (0: all)
  bool pairs = true;      // backward kernel pairs share a launch
  bool mix = true;        // ... with their workgroups interleaved round by round (pair a,
  bool mix_b = true;      // and pair b: slower with 4-shot groups, faster with 3)
  bool balance = true;    // C-PML columns shared evenly by the XCDs (update-region launches)
  int* perm_u = nullptr;  // workgroup -> tile order of the update-region tiling
  bool illum_on = false;  // sum the forward illumination
  bool illum_valid = false;
  // residual conditioning, shot by shot: the optional stages, and the path
  // that keeps the conditioned data for the scratch files
  bool cond = false;
  FftDesc fft{};
  int* fperm = nullptr;                    // [2*nSteps] digit-reversal permutation
  float2* ftw = nullptr;                   // [2*nSteps] twiddles
  float2* bpH = nullptr;                   // [nSteps+1] band-pass amp^2
  float* wpad = nullptr;                   // [2*nSteps] window ratio 0.01 over the padded length
  float *wst = nullptr, *wen = nullptr, *wwt = nullptr;  // [R] per-trace windows and weights
  float *cobs = nullptr, *csyn = nullptr;  // [nSteps][nrec of one shot] conditioned data
  float2 *spo = nullptr, *spc = nullptr;   // [nrec][nSteps+1] spectra (source update)
  float2* coef = nullptr;                  // [nSteps+1] source-update filter of one shot
  float* srcamp = nullptr;                 // [ns][nSteps] source rows as injected (stf * window)
  float* srcout = nullptr;                 // [ns][nSteps] updated source rows
  float* inj_rev = nullptr;                // [ns][nSteps] increments of the reversal
  double* cpart = nullptr;                 // [ns][kCondBlocks] misfit partial sums
  int nrec_max = 0;
  bool src_updated = false;                // the last run updated the source rows
  bool fold_inject = true;                 // residual injection inside the adjoint stress update
  int* rec_at = nullptr;                   // [ns][nx] receiver of each column (one-row surveys)
  int rec_row = -1;                        // the row of those receivers
  std::vector<float> keep_obs, keep_syn;   // conditioned data [R][nSteps] for the scratch files
  float* illum = nullptr; // [NX][NZ]
  std::vector<int> roff;  // first receiver of each shot (+ total)
  std::vector<void*> allocs;
  long long bytes = 0;
  // model + tables
  float *lam = nullptr, *mu = nullptr, *amu = nullptr, *ba = nullptr, *bb = nullptr;
  float* den = nullptr;
  float* pml = nullptr;   // 6*nzA + 6*nx
  float* win2 = nullptr;  // nSteps
  int* src = nullptr;     // [ns][2]
  float* inj = nullptr;   // [ns][nSteps]
  int* rec = nullptr;     // [R][3]
  // wavefields
  float *fwd = nullptr, *adj = nullptr, *mem = nullptr;
  float *fwd2 = nullptr, *mem2 = nullptr, *adj2 = nullptr;  // ping-pong partners
  int *rptr = nullptr, *rzr = nullptr, *rrg = nullptr;       // receivers by column
  float* accp = nullptr;  // per-group partial gradient accumulators
  int accp_groups = 0;
  float* ring = nullptr;
  float *syn = nullptr, *obs = nullptr, *res = nullptr, *stage = nullptr;
  float *acc = nullptr, *grad = nullptr, *stf_grad = nullptr;
  double* partial = nullptr;
  double* misfit = nullptr;
  int npart = 0;
  bool model_set = false, stf_set = false, obs_set = false;
  double misfit_host = 0.0;
  // timing
  bool timing = false;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long long>> totals;

  ~DeviceState() {
    if (stream) (void)hipSetDevice(dev);
    for (auto& r : pending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
    for (auto e : join)
      if (e) (void)hipEventDestroy(e);  // a null handle would leave a sticky error
    for (void* p : allocs) (void)hipFree(p);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  Status alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t nb = count * sizeof(T);
    hipError_t e = hipMalloc(&q, nb ? nb : 4);
    if (e != hipSuccess)
      return Status::Err(e == hipErrorOutOfMemory ? FWIFLOW_ERR_MEMORY : FWIFLOW_ERR_HIP,
                         "hipMalloc(" + std::to_string(nb) + " bytes): " + hipGetErrorString(e));
    allocs.push_back(q);
    bytes += (long long)nb;
    *p = (T*)q;
    return Status::Ok();
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

  // Launch helper: optionally brackets the launch with events.
  template <typename F>
  void launch(const char* name, bool sample, F&& f) {
    if (timing && sample) {
      Rec r{name, take_event(), take_event()};
      (void)hipEventRecord(r.a, stream);
      f();
      (void)hipEventRecord(r.b, stream);
      pending.push_back(r);
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
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }

  Pml pml_ptrs() const {
    const int nzA = g.nzA, nx = g.nx;
    const float* z = pml;
    const float* x = pml + 6 * nzA;
    return Pml{z,     z + nzA,     z + 2 * nzA, z + 3 * nzA, z + 4 * nzA, z + 5 * nzA,
               x,     x + nx,      x + 2 * nx,  x + 3 * nx,  x + 4 * nx,  x + 5 * nx};
  }
  Mdl mdl() const { return Mdl{lam, mu, amu, ba, bb}; }
  Wave wave(float* base) const {
    const size_t F = (size_t)g.ns * g.N;
    return Wave{base, base + F, base + 2 * F, base + 3 * F, base + 4 * F};
  }
  Mem memv(float* b) const {
    const size_t F = (size_t)g.ns * g.N;
    return Mem{b, b + F, b + 2 * F, b + 3 * F, b + 4 * F, b + 5 * F, b + 6 * F, b + 7 * F};
  }
  // Views of shots [s0, s0 + n) of the launch: every per-shot array is
  // offset to shot s0 and the geometry counts n shots.
  Geo group_geo(int s0, int n) const {
    Geo q = g;
    q.s0 = s0;
    q.ns = n;
    return q;
  }
  Wave wave(float* base, int s0) const {
    Wave w = wave(base);
    const size_t o = (size_t)s0 * g.N;
    return Wave{w.vz + o, w.vx + o, w.szz + o, w.sxx + o, w.sxz + o};
  }
  Mem memv(float* b, int s0) const {
    Mem p = memv(b);
    const size_t o = (size_t)s0 * g.N;
    return Mem{p.vzz + o, p.vzx + o, p.vxz + o, p.vxx + o,
               p.szz + o, p.sxx + o, p.sxzz + o, p.sxzx + o};
  }
  // Shots per launch of the time-step kernels.  Automatic: as many shots as
  // keep one group's backward working set (forward and adjoint state, C-PML
  // memory variables, about 48 B per cell and shot) near 224 MB, i.e. inside
  // the 256 MB Infinity Cache next to the model and accumulators.  Measured
  // at 608x1780: 3, 4 and 5 shots give 55.2, 59.4 and 56.4 shot-gradients/s
  // at nt 400; the rule picks 4.
  int group_size() const {
    if (fused || shot_group == 0) return g.ns;
    if (shot_group > 0) return std::min(shot_group, g.ns);
    // ~40 B per cell and shot stay cache-resident from step to step in the
    // forward sweep (5 at 608x1780: 4.9 groups-of-5 measured 7.8 us per
    // shot-step against 8.1 for 4 and 9.3 for 8)
    const double per_shot = 40.0 * (double)g.N;
    const int n = (int)std::lround(224e6 / per_shot);
    return std::max(1, std::min(n, g.ns));
  }
  // The backward sweep holds forward and adjoint state of its group: fewer
  // shots fit the 256 MB Infinity Cache (3 at 608x1780, where 4 shots'
  // 4 x 43 MB of state plus model and accumulators overflow it).
  int group_size_bwd() const {
    if (fused || (shot_group_bwd < 0 && shot_group == 0)) return g.ns;
    if (shot_group_bwd == 0) return g.ns;
    if (shot_group_bwd > 0) return std::min(shot_group_bwd, g.ns);
    if (shot_group > 0) return std::min(shot_group, g.ns);
    const double per_shot = 64.0 * (double)g.N;
    const int n = (int)std::lround(224e6 / per_shot);
    return std::max(1, std::min(n, g.ns));
  }
  Src src_view(int s0) const { return Src{src + 2 * s0, inj + (size_t)s0 * g.nSteps}; }
  Src src_view(int s0, const float* rows) const {
    return Src{src + 2 * s0, rows + (size_t)s0 * g.nSteps};
  }
  float* ring_view(int) const { return ring; }  // ring_ptr adds the group's first shot
};

Plan::Plan() = default;
Plan::~Plan() = default;

Status Plan::create(const std::string& para_fname, int gpu, int group_size, const int* shot_ids) {
  TRY(parse_para(para_fname, para));
  TRY(parse_survey(para, group_size, shot_ids, survey));
  gpu_id = gpu;
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
  d.fused = env_int("FWIFLOW_FUSED", 0) != 0;
  d.tiled = env_int("FWIFLOW_TILED", 0) != 0 && !d.fused;
  d.seg_x = std::max(4, std::min(256, env_int("FWIFLOW_SEG_X", 32)));
  d.rev_group = std::max(1, std::min(64, env_int("FWIFLOW_REV_GROUP", 8)));
  {
    const int u = env_int("FWIFLOW_SHOT_UNROLL", 1);
    d.unroll = u >= 4 ? 4 : (u >= 2 ? 2 : 1);
    // the forward kernels keep occupancy 4-5 with two shots' loads in flight
    // and gained (6.07-6.09 vs 6.05-6.06 shot-gradients/s)
    const int uf = env_int("FWIFLOW_SHOT_UNROLL_FWD", 2);
    d.unroll_fwd = uf >= 4 ? 4 : (uf >= 2 ? 2 : 1);
  }
  d.shot_group = std::max(-1, env_int("FWIFLOW_SHOT_GROUP", -1));
  d.shot_group_bwd = std::max(-1, env_int("FWIFLOW_SHOT_GROUP_BWD", -1));
  d.streams = env_int("FWIFLOW_STREAMS", 1) == 2 ? 2 : 1;
  d.spt = std::max(0, env_int("FWIFLOW_SHOTS_PER_THREAD", 0));
  d.pairs = env_int("FWIFLOW_BWD_PAIRS", 1) != 0;
  d.mix = env_int("FWIFLOW_PAIR_MIX", 1) != 0;
  d.mix_b = env_int("FWIFLOW_PAIR_MIX_B", 1) != 0;
  d.balance = env_int("FWIFLOW_XCD_BALANCE", 1) != 0 && d.xcd;
  HIP_TRY(hipSetDevice(gpu));
  HIP_TRY(hipStreamCreate(&d.stream));
  if (d.streams == 2) {
    HIP_TRY(hipStreamCreate(&d.stream2));
    for (auto& e : d.join) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  Geo& g = d.g;
  g.nz = para.nz;
  g.nx = para.nx;
  g.nzA = para.nzA();
  g.nPml = para.nPml;
  g.ns = group_size;
  g.nSteps = para.nSteps;
  g.N = (long long)para.nz * para.nx;
  g.dt = para.dt;
  g.rdz = 1.f / para.dz;
  g.rdx = 1.f / para.dx;
  g.ringH = g.nzA - 2 * g.nPml + 4;
  g.ringW = g.nx - 2 * g.nPml + 4;
  g.ringlen = 10 * g.ringW + 10 * (g.ringH - 10);
  g.s0 = 0;
  d.R = survey.total_rec();

  const size_t N = (size_t)g.N;
  TRY(d.alloc(&d.lam, N));
  TRY(d.alloc(&d.mu, N));
  TRY(d.alloc(&d.den, N));
  TRY(d.alloc(&d.amu, N));
  TRY(d.alloc(&d.ba, N));
  TRY(d.alloc(&d.bb, N));
  // host tables
  std::vector<float> tab(6 * (size_t)g.nzA + 6 * (size_t)g.nx);
  cpml_profiles(g.nzA, g.nPml, para.dz, para.f0, para.dt, tab.data());
  cpml_profiles(g.nx, g.nPml, para.dx, para.f0, para.dt, tab.data() + 6 * g.nzA);
  TRY(d.alloc(&d.pml, tab.size()));
  HIP_TRY(hipMemcpy(d.pml, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> w2(g.nSteps);
  window_squared(g.nSteps, para.dt, 0.005f, w2.data());
  TRY(d.alloc(&d.win2, w2.size()));
  HIP_TRY(hipMemcpy(d.win2, w2.data(), w2.size() * sizeof(float), hipMemcpyHostToDevice));
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
  d.fold_inject = env_int("FWIFLOW_FOLD_INJECT", 1) != 0;
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
    if (one_row) {
      TRY(d.alloc(&d.rec_at, at.size()));
      HIP_TRY(hipMemcpy(d.rec_at, at.data(), at.size() * sizeof(int), hipMemcpyHostToDevice));
      d.rec_row = row;
    }
  }
  TRY(d.alloc(&d.src, src.size()));
  HIP_TRY(hipMemcpy(d.src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
  TRY(d.alloc(&d.rec, rec.size()));
  HIP_TRY(hipMemcpy(d.rec, rec.data(), rec.size() * sizeof(int), hipMemcpyHostToDevice));
  {
    // receivers of each shot grouped by column (CSR) for the fused adjoint
    std::vector<int> ptr((size_t)g.ns * (g.nx + 1), 0), zr(d.R), rg(d.R);
    size_t base = 0, rglob = 0;
    for (int s = 0; s < g.ns; s++) {
      const Shot& sh = survey.shots[s];
      int* P = ptr.data() + (size_t)s * (g.nx + 1);
      std::vector<int> cnt(g.nx + 1, 0);
      for (size_t k = 0; k < sh.x_rec.size(); k++) cnt[sh.x_rec[k] + 1]++;
      for (int x = 0; x < g.nx; x++) cnt[x + 1] += cnt[x];
      for (int x = 0; x <= g.nx; x++) P[x] = (int)base + cnt[x];
      std::vector<int> fill(cnt.begin(), cnt.end() - 1);
      for (size_t k = 0; k < sh.x_rec.size(); k++) {
        const size_t pos = base + fill[sh.x_rec[k]]++;
        zr[pos] = sh.z_rec[k];
        rg[pos] = (int)(rglob + k);
      }
      base += sh.x_rec.size();
      rglob += sh.x_rec.size();
    }
    TRY(d.alloc(&d.rptr, ptr.size()));
    TRY(d.alloc(&d.rzr, zr.size()));
    TRY(d.alloc(&d.rrg, rg.size()));
    HIP_TRY(hipMemcpy(d.rptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.rzr, zr.data(), zr.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.rrg, rg.data(), rg.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  if (d.balance) {
    const Tiling t = make_tiling(g.nzA, g.nx, d.tile_x, true);
    const int hl = (g.nPml + d.tile_x - 1) / d.tile_x;
    const int hr = t.gx - (g.nx - g.nPml) / d.tile_x;
    const std::vector<int> perm = balanced_tile_order(t.gz, t.gx, hl, hr);
    std::vector<char> seen(perm.size(), 0);
    for (int v : perm) {
      if (v < 0 || v >= t.n || seen[v])
        return Status::Err(FWIFLOW_ERR_HIP, "internal: workgroup order is not a permutation");
      seen[v] = 1;
    }
    TRY(d.alloc(&d.perm_u, perm.size()));
    HIP_TRY(hipMemcpy(d.perm_u, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  TRY(d.alloc(&d.inj, (size_t)g.ns * g.nSteps));
  const size_t F = (size_t)g.ns * N;
  TRY(d.alloc(&d.fwd, 5 * F));
  TRY(d.alloc(&d.mem, 8 * F));
  if (d.tiled) {
    TRY(d.alloc(&d.fwd2, 5 * F));
    TRY(d.alloc(&d.mem2, 8 * F));
    d.fused_rev = false;
  } else if (d.fused && fused_threads(g) <= 1024) {
    TRY(d.alloc(&d.fwd2, 5 * F));
    TRY(d.alloc(&d.mem2, 8 * F));
    const size_t ldsr = (size_t)(10 + 5 * TXR) * fused_threads(g) * sizeof(float);
    d.fused_rev = env_int("FWIFLOW_FUSED_REV", 1) != 0 && ldsr <= 160 * 1024;
    if (d.fused_rev)
      HIP_TRY(hipFuncSetAttribute((const void*)k_rev_step,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsr));
  } else {
    d.fused = false;
    d.fused_rev = false;
  }
  TRY(d.alloc(&d.srcamp, (size_t)g.ns * g.nSteps));
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
      TRY(d.alloc(&d.inj_rev, (size_t)g.ns * nt));
    }
  }
  const size_t T = (size_t)g.nSteps * d.R;
  TRY(d.alloc(&d.syn, T));
  d.npart = 1024;
  TRY(d.alloc(&d.partial, (size_t)d.npart));
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
  const Tiling tA = make_tiling(d.g.nz, d.g.nx, d.tile_x, d.xcd);
  hipLaunchKernelGGL(k_model_init, dim3(tA.n), dim3(kBZ, d.tile_x), 0, d.stream, tA, d.g, d.mu,
                     d.den, d.amu, d.ba, d.bb);
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
  // Src_Rec.cu:132-143: row shot_id -> float -> window ratio 0.001; the
  // increment added to szz is 1500^2 * stf * dt (add_source, utilities.cu:528-538)
  std::vector<float> w2(nSteps), inj((size_t)d.g.ns * nSteps), amps(inj.size());
  window_squared(nSteps, para.dt, 0.001f, w2.data());
  for (int s = 0; s < d.g.ns; s++) {
    const double* row = stf + (size_t)survey.shots[s].id * nSteps;
    for (int it = 0; it < nSteps; it++) {
      const float amp = (float)row[it] * w2[it];
      amps[(size_t)s * nSteps + it] = amp;
      inj[(size_t)s * nSteps + it] = (float)(2250000.0 * (double)amp * (double)para.dt);
    }
  }
  HIP_TRY(hipSetDevice(d.dev));
  HIP_TRY(hipMemcpy(d.inj, inj.data(), inj.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d.srcamp, amps.data(), amps.size() * sizeof(float), hipMemcpyHostToDevice));
  d.stf_set = true;
  return Status::Ok();
}

static Status ensure_obs_buffers(DeviceState& d) {
  const size_t T = (size_t)d.g.nSteps * d.R;
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
  if (k > 0 && d_->fused)
    return Status::Err(FWIFLOW_ERR_UNSUPPORTED,
                       "interior checkpoints need the split or tiled kernels (FWIFLOW_FUSED=0)");
  d_->ckpt_every = k;
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

int Plan::shots_per_launch() const { return d_ ? d_->group_size() : 0; }
int Plan::shots_per_launch_bwd() const { return d_ ? d_->group_size_bwd() : 0; }

// Residual with the optional conditioning stages, one shot at a time
// (libCUFD.cu:254-313): window obs and syn (per-trace windows and weights, or
// the default ratio-0.005 window), band-pass both, source update (syn and the
// source row replaced), res = obs - syn and its misfit, then the adjoint
// stages on res: source-update filter, band-pass, window (ratio 0.1 with
// per-trace windows).  `keep` copies the conditioned obs and syn to the host
// for the scratch files.
static Status condition_residual(DeviceState& d, const Para& para, bool keep) {
  const Geo& g = d.g;
  const int nt = g.nSteps, nf = nt + 1, kc = touched_bins(nt);
  const float scale = 1.0f / (float)(2 * nt);
  const float dt = para.dt;
  hipStream_t st = d.stream;
  if (keep) {
    d.keep_obs.assign((size_t)d.R * nt, 0.f);
    d.keep_syn.assign((size_t)d.R * nt, 0.f);
  }
  std::vector<float> h;
  for (int s = 0; s < g.ns; s++) {
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
      HIP_TRY(source_increment(st, nt, d.srcout + so, dt, d.inj_rev + so));
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
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, d.cpart, g.ns * kCondBlocks,
                     d.misfit);
  HIP_TRY(hipGetLastError());
  return Status::Ok();
}

Status Plan::run(int calc_id) {
  TRY(check_calc_id(calc_id));
  if (!d_) return Status::Err(FWIFLOW_ERR_ARG, "plan not created");
  DeviceState& d = *d_;
  if (!d.model_set) return Status::Err(FWIFLOW_ERR_ARG, "model not set");
  if (!d.stf_set) return Status::Err(FWIFLOW_ERR_ARG, "source time functions not set");
  const bool if_res = calc_id != 2, withAdj = calc_id == 1;
  if (if_res && !d.obs_set) return Status::Err(FWIFLOW_ERR_ARG, "observed data not set");
  d.illum_valid = false;
  last_calc = -1;
  HIP_TRY(hipSetDevice(d.dev));
  const Geo& g = d.g;
  const size_t N = (size_t)g.N, F = (size_t)g.ns * N;
  const size_t T = (size_t)g.nSteps * d.R;
  if (withAdj) {
    if (!d.adj) TRY(d.alloc(&d.adj, 5 * F));
    if (d.fused && !d.adj2) TRY(d.alloc(&d.adj2, 5 * F));
    if (!d.ring) TRY(d.alloc(&d.ring, (size_t)g.nSteps * 5 * g.ns * (size_t)g.ringlen));
    if (!d.acc) TRY(d.alloc(&d.acc, (size_t)(d.streams == 2 ? 10 : 5) * N));
    if (!d.grad) TRY(d.alloc(&d.grad, 3 * N));
    if (!d.stf_grad) TRY(d.alloc(&d.stf_grad, (size_t)g.ns * g.nSteps));
  }
  // interior checkpoints: slot it/k for it = k, 2k, .. <= nSteps-2 (slot 0 unused)
  const int ck_k = withAdj ? d.ckpt_every : 0;
  const int ck_slots = ck_k > 0 ? (g.nSteps - 2) / ck_k + 1 : 0;
  const size_t ck_cells = (size_t)(g.nzA - 2 * g.nPml) * (g.nx - 2 * g.nPml);
  if (ck_k > 0) {
    const size_t need = (size_t)g.ns * ck_slots * 5 * ck_cells;
    if (d.ck_floats < need) {
      TRY(d.alloc(&d.ck, need));  // a smaller earlier buffer stays in the plan's pool
      d.ck_floats = need;
    }
  }
  const Tiling tI = make_tiling(g.nzA - 2 * g.nPml, g.nx - 2 * g.nPml, d.tile_x, d.xcd);
  hipStream_t st = d.stream;
  HIP_TRY(hipMemsetAsync(d.fwd, 0, 5 * F * sizeof(float), st));
  HIP_TRY(hipMemsetAsync(d.mem, 0, 8 * F * sizeof(float), st));
  if (d.fused || d.tiled) {
    HIP_TRY(hipMemsetAsync(d.fwd2, 0, 5 * F * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(d.mem2, 0, 8 * F * sizeof(float), st));
  }
  Tiling tT;
  tT.gz = (g.nzA + TT - 1) / TT;
  tT.gx = (g.nx + TT - 1) / TT;
  tT.n = tT.gz * tT.gx;
  tT.xcd = d.xcd ? 1 : 0;
  HIP_TRY(hipMemsetAsync(d.syn, 0, T * sizeof(float), st));
  float* il_ptr = nullptr;
  if (d.illum_on) {
    if (d.fused || d.tiled)
      return Status::Err(FWIFLOW_ERR_UNSUPPORTED,
                         "illumination is summed by the split forward kernels only");
    if (!d.illum) TRY(d.alloc(&d.illum, N));
    HIP_TRY(hipMemsetAsync(d.illum, 0, N * sizeof(float), st));
    il_ptr = d.illum;
  }

  const Mdl m = d.mdl();
  const Pml c = d.pml_ptrs();
  float* fa = d.fwd;   // state n
  float* fb = d.fwd2;  // state n+1 (fused path)
  float* ma = d.mem;
  float* mb = d.mem2;
  const Src src{d.src, d.inj};
  const dim3 blk(kBZ, d.tile_x);
  Tiling tU = make_tiling(g.nzA, g.nx, d.tile_x, d.xcd);
  tU.perm = d.perm_u;
  const int recBlocks = (d.R + 255) / 256;
  const int sample_every = 8;
  Seg q;
  q.tx = d.seg_x;
  q.nseg = (g.nx + q.tx - 1) / q.tx;
  q.n = q.nseg * g.ns;
  q.xcd = d.xcd ? 1 : 0;
  const int NTf = fused_threads(g);
  const size_t ldsf = (size_t)10 * NTf * sizeof(float);

  // ---------------------------------------------------------- forward
  // The split and tiled paths take the shots in groups of `shot_group`: a
  // group runs every time step before the next group starts, so the state of
  // one group stays resident in the last-level cache from step to step.
  const int G = d.group_size();
  const int U = d.unroll;
  // two streams for the split kernels (the illumination has one writer per cell)
  const bool two = d.streams == 2 && !d.fused && !d.tiled && !d.illum_on;
#define LAUNCH_U(K, grid, block, ...)                                        \
  do {                                                                       \
    if (U == 4)                                                              \
      hipLaunchKernelGGL(K<4>, grid, block, 0, st, __VA_ARGS__);             \
    else if (U == 2)                                                         \
      hipLaunchKernelGGL(K<2>, grid, block, 0, st, __VA_ARGS__);             \
    else                                                                     \
      hipLaunchKernelGGL(K<1>, grid, block, 0, st, __VA_ARGS__);             \
  } while (0)
  if (d.fused) {
    for (int it = 0; it <= g.nSteps - 2; it++) {
      const bool smp = (it % sample_every) == 0;
      float* ring = withAdj ? d.ring : nullptr;
      const Wave wa = d.wave(fa), wb = d.wave(fb);
      const Mem pa = d.memv(ma), pb = d.memv(mb);
      d.launch("fwd_step", smp, [&] {
        hipLaunchKernelGGL(k_fwd_step, dim3(q.n), dim3(NTf), ldsf, st, q, g, m, c, wa, wb, pa, pb,
                           src, it, ring);
      });
      hipLaunchKernelGGL(k_record, dim3(recBlocks), dim3(256), 0, st, g, wb.szz, wb.sxx, d.rec,
                         d.R, d.syn + (size_t)(it + 1) * d.R);
      std::swap(fa, fb);
      std::swap(ma, mb);
    }
  } else {
    float* fa0 = fa;
    float* ma0 = ma;
    // ceil(ns / G) groups of balanced sizes (64 shots by 3: 20 x 3 + 2 x 2)
    const int NGf = (g.ns + G - 1) / G;
    const int U = d.unroll_fwd;  // LAUNCH_U below
    if (two) {
      HIP_TRY(hipEventRecord(d.join[0], st));
      HIP_TRY(hipStreamWaitEvent(d.stream2, d.join[0], 0));
    }
    for (int k = 0; k < NGf; k++) {
      const bool second = two && (k & 1);
      hipStream_t st = second ? d.stream2 : d.stream;  // the launches below use it
      const int s0 = (int)((long long)k * g.ns / NGf);
      const int n = (int)((long long)(k + 1) * g.ns / NGf) - s0;
      const Geo gg = d.group_geo(s0, n);
      Tiling tUg = tU;  // forward/adjoint split kernels: shots spread over chunks
      // (one chunk when the illumination is summed: one writer per cell)
      tUg.spt = d.spt > 0 && !d.illum_on ? std::min(d.spt, n) : n;
      tUg.chunks = (n + tUg.spt - 1) / tUg.spt;
      const dim3 grdUg(tUg.n * tUg.chunks);
      const Src sg = d.src_view(s0);
      float* ring = withAdj ? d.ring_view(s0) : nullptr;
      const int r0 = d.roff[s0], Rg = d.roff[s0 + n] - r0;
      const dim3 rgrid((Rg + 255) / 256);
      fa = fa0, fb = d.fwd2, ma = ma0, mb = d.mem2;
      float* ckg = ck_k > 0 ? d.ck + (size_t)s0 * ck_slots * 5 * ck_cells : nullptr;
      for (int it = 0; it <= g.nSteps - 2; it++) {
        const bool smp = !second && (it % sample_every) == 0;  // events on the first stream
        float* out = fa;  // buffer that holds step it+1 after this step
        if (ckg && it > 0 && it % ck_k == 0)
          hipLaunchKernelGGL(k_ckpt, dim3(tI.n), blk, 0, st, tI, gg, d.wave(fa, s0), ckg,
                             it / ck_k, ck_slots, 1);
        if (d.tiled) {
          const Wave wa = d.wave(fa, s0), wb = d.wave(fb, s0);
          const Mem pa = d.memv(ma, s0), pb = d.memv(mb, s0);
          d.launch("fwd_tile", smp, [&] {
            hipLaunchKernelGGL(k_fwd_tile, dim3(tT.n), dim3(1024), 0, st, tT, gg, m, c, wa, wb,
                               pa, pb, sg, it, ring);
          });
          out = fb;
          std::swap(fa, fb);
          std::swap(ma, mb);
        } else {
          const Wave w = d.wave(fa, s0);
          const Mem p = d.memv(ma, s0);
          d.launch("stress_fwd", smp, [&] {
            LAUNCH_U(k_stress_fwd, grdUg, blk, tUg, gg, m, c, w, p, sg, it, ring);
          });
          // velocity update with this step's recording appended to its grid
          const int nV = (tUg.n * tUg.chunks + 7) / 8 * 8;
          const Wave wo = d.wave(fa);
          float* row = d.syn + (size_t)(it + 1) * d.R + r0;
          d.launch("vel_fwd", smp, [&] {
            const int per = (int)(blk.x * blk.y);
            LAUNCH_U(k_vel_fwd_rec, dim3(nV + (Rg + per - 1) / per), blk, nV, tUg, gg, m, c, w, p,
                     il_ptr, g, wo.szz, wo.sxx, d.rec + 3 * r0, Rg, row);
          });
        }
        if (Rg > 0 && d.tiled) {
          const Wave wo = d.wave(out);
          hipLaunchKernelGGL(k_record, rgrid, dim3(256), 0, st, g, wo.szz, wo.sxx, d.rec + 3 * r0,
                             Rg, d.syn + (size_t)(it + 1) * d.R + r0);
        }
      }
    }
    if (two) {
      HIP_TRY(hipEventRecord(d.join[1], d.stream2));
      HIP_TRY(hipStreamWaitEvent(st, d.join[1], 0));
    }
  }
  HIP_TRY(hipGetLastError());
  const Wave w = d.wave(fa);  // final forward state

  d.src_updated = false;
  if (if_res && d.cond) {
    TRY(condition_residual(d, para, para.save_scratch && withAdj));
    d.src_updated = para.if_src_update;
  } else if (if_res) {
    hipLaunchKernelGGL(k_residual, dim3(d.npart), dim3(256), 0, st, g.nSteps, d.R, d.win2, d.obs,
                       d.syn, d.res, d.partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, d.partial, d.npart, d.misfit);
    HIP_TRY(hipGetLastError());
  }
  // the reversal subtracts the source rows the residual stage left: updated
  // ones after a source update (libCUFD.cu:326-329, 390-394)
  const float* inj_b = d.src_updated ? d.inj_rev : d.inj;
  const Src srcb{d.src, inj_b};

  if (withAdj) {
    // ------------------------------------------------------- backward
    HIP_TRY(hipMemsetAsync(d.adj, 0, 5 * F * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(ma, 0, 8 * F * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(d.acc, 0, (two ? 10 : 5) * N * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(d.stf_grad, 0, (size_t)g.ns * g.nSteps * sizeof(float), st));
    const Tiling tR = make_tiling(g.ringH, g.ringW, d.tile_x, d.xcd);
    const dim3 grdR(tR.n);
    float* acc = d.acc;
    if (d.fused) {
      HIP_TRY(hipMemsetAsync(d.adj2, 0, 5 * F * sizeof(float), st));
      HIP_TRY(hipMemsetAsync(mb, 0, 8 * F * sizeof(float), st));
      float* aa = d.adj;  // adjoint state before the step
      float* ab = d.adj2;
      float* qa = ma;     // adjoint memory variables
      float* qb = mb;
      const RecIdx rx{d.rptr, d.rzr, d.rrg};
      const size_t ldsa = (size_t)18 * NTf * sizeof(float);
      auto adj_step = [&](int res_t, int stf_it) {
        hipLaunchKernelGGL(k_adj_step, dim3(q.n), dim3(NTf), ldsa, st, q, g, m, c, d.wave(aa),
                           d.wave(ab), d.memv(qa), d.memv(qb), srcb, rx,
                           d.res + (size_t)res_t * d.R, d.R, d.stf_grad, stf_it);
        std::swap(aa, ab);
        std::swap(qa, qb);
      };
      SegG qr;
      qr.nseg = (g.nx + TXR - 1) / TXR;
      qr.gsz = d.rev_group;
      qr.ngrp = (g.ns + qr.gsz - 1) / qr.gsz;
      qr.n = qr.nseg * qr.ngrp;
      qr.xcd = d.xcd ? 1 : 0;
      if (d.accp_groups < qr.ngrp) {
        TRY(d.alloc(&d.accp, (size_t)qr.ngrp * 5 * N));
        d.accp_groups = qr.ngrp;
      }
      HIP_TRY(hipMemsetAsync(d.accp, 0, (size_t)qr.ngrp * 5 * N * sizeof(float), st));
      const size_t ldsr = (size_t)(10 + 5 * TXR) * NTf * sizeof(float);
      float* ra = fa;  // forward state at step it+1
      float* rb = fb;  // forward state at step it
      // initial adjoint step at t = nSteps-1 (libCUFD.cu:353-373)
      adj_step(g.nSteps - 1, -1);
      for (int it = g.nSteps - 2; it >= 0; it--) {
        const bool smp = (it % sample_every) == 0;
        const Wave a = d.wave(aa);
        if (d.fused_rev) {
          const Wave wa = d.wave(ra), wb = d.wave(rb);
          d.launch("rev_step", smp, [&] {
            hipLaunchKernelGGL(k_rev_step, dim3(qr.n), dim3(NTf), ldsr, st, qr, g, m, wa, wb, a,
                               srcb, d.ring, it, d.accp);
          });
          std::swap(ra, rb);
        } else {
          d.launch("rev_vel", smp, [&] {
            hipLaunchKernelGGL(k_rev_vel<1>, grdR, blk, 0, st, tR, g, m, w, a, d.accp + 3 * N,
                               d.accp + 4 * N, d.ring, it);
          });
          d.launch("rev_stress", smp, [&] {
            hipLaunchKernelGGL(k_rev_stress<1>, grdR, blk, 0, st, tR, g, m, w, a, src, d.accp,
                               d.accp + N, d.accp + 2 * N, d.ring, it);
          });
        }
        d.launch("adj_step", smp, [&] { adj_step(it, it); });
      }
      const long long n5 = 5 * (long long)N;
      hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)((n5 + 255) / 256)), dim3(256), 0, st,
                         d.accp, n5, qr.ngrp);
      acc = d.accp;
    } else {
      const Wave a_all = d.wave(d.adj);
      const int GB = d.group_size_bwd();
      const int NGb = (g.ns + GB - 1) / GB;
      if (two) {
        HIP_TRY(hipEventRecord(d.join[0], st));
        HIP_TRY(hipStreamWaitEvent(d.stream2, d.join[0], 0));
      }
      float* const acc0 = acc;
      for (int k = 0; k < NGb; k++) {
        const bool second = two && (k & 1);
        hipStream_t st = second ? d.stream2 : d.stream;  // the launches below use it
        float* acc = second ? acc0 + 5 * N : acc0;        // and these accumulators
        const int s0 = (int)((long long)k * g.ns / NGb);
        const int n = (int)((long long)(k + 1) * g.ns / NGb) - s0;
        const Geo gg = d.group_geo(s0, n);
        Tiling tUg = tU;  // forward/adjoint split kernels: shots spread over chunks
        tUg.spt = d.spt > 0 ? std::min(d.spt, n) : n;
        tUg.chunks = (n + tUg.spt - 1) / tUg.spt;
        const dim3 grdUg(tUg.n * tUg.chunks);
        const Src sg = d.src_view(s0, inj_b);
        const float* ring = d.ring_view(s0);
        const int r0 = d.roff[s0], Rg = d.roff[s0 + n] - r0;
        const dim3 rgrid((Rg + 255) / 256);
        const Wave wg = d.wave(fa, s0);  // final forward state of the group
        const Wave a = d.wave(d.adj, s0);
        const Mem pg = d.memv(ma, s0);   // adjoint memory variables
        float* sgrad = d.stf_grad + (size_t)s0 * g.nSteps;
        float* ckg = ck_k > 0 ? d.ck + (size_t)s0 * ck_slots * 5 * ck_cells : nullptr;
        // after reverse step it the group's state is the one before forward step it
        auto reset = [&](int it) {
          if (ckg && it > 0 && it % ck_k == 0)
            hipLaunchKernelGGL(k_ckpt, dim3(tI.n), blk, 0, st, tI, gg, wg, ckg, it / ck_k,
                               ck_slots, 0);
        };
        auto inject = [&](int t) {
          if (Rg > 0)
            hipLaunchKernelGGL(k_res_inject, rgrid, dim3(256), 0, st, g, a_all.szz, a_all.sxx,
                               d.rec + 3 * r0, Rg, d.res + (size_t)t * d.R + r0);
        };
        // with backward pairs the residual of step t is injected by the
        // adjoint stress update of step t itself (k_bwd_pair_b, k_adj_stress)
        const bool fold = d.fold_inject && d.rec_at && d.pairs && g.nSteps >= 2;
        auto inj_at = [&](int t) {
          return Inj{fold ? d.res + (size_t)t * d.R : nullptr, d.rec_at, d.rec_row};
        };
        const Inj none{nullptr, d.rec_at, -1};
        // initial adjoint step at t = nSteps-1 (libCUFD.cu:353-373)
        LAUNCH_U(k_adj_vel, grdUg, blk, tUg, gg, m, c, a, pg, sg, sgrad, -1);
        if (!fold) inject(g.nSteps - 1);
        if (d.pairs && g.nSteps >= 2) {
          const int nU8 = (tUg.n * tUg.chunks + 7) / 8 * 8, nR8 = (tR.n + 7) / 8 * 8;
          const dim3 grdP(nR8 + nU8);
          const int mix = d.mix ? 1 : 0, mix_b = d.mix_b ? 1 : 0;
          // adj_stress(nSteps-1) with rev_vel(nSteps-2)
          LAUNCH_U(k_bwd_pair_b, grdP, blk, nU8, nR8, mix_b, tUg, tR, gg, m, c, wg, a, pg, acc, ring,
                   g.nSteps - 2, inj_at(g.nSteps - 1));
          for (int it = g.nSteps - 2; it >= 0; it--) {
            const bool smp = !second && (it % sample_every) == 0;
            d.launch("bwd_pair_a", smp, [&] {
              LAUNCH_U(k_bwd_pair_a, grdP, blk, nR8, nU8, mix, tR, tUg, gg, m, c, wg, a, pg, sg, acc,
                       ring, sgrad, it);
            });
            if (!fold) inject(it);
            reset(it);
            if (it >= 1) {
              d.launch("bwd_pair_b", smp, [&] {
                LAUNCH_U(k_bwd_pair_b, grdP, blk, nU8, nR8, mix_b, tUg, tR, gg, m, c, wg, a, pg, acc,
                         ring, it - 1, inj_at(it));
              });
            } else {
              LAUNCH_U(k_adj_stress, grdUg, blk, tUg, gg, m, c, a, pg, inj_at(0));
            }
          }
          continue;
        }
        LAUNCH_U(k_adj_stress, grdUg, blk, tUg, gg, m, c, a, pg, none);
        for (int it = g.nSteps - 2; it >= 0; it--) {
          const bool smp = !second && (it % sample_every) == 0;
          d.launch("rev_vel", smp, [&] {
            LAUNCH_U(k_rev_vel, grdR, blk, tR, gg, m, wg, a, acc + 3 * N, acc + 4 * N, ring, it);
          });
          d.launch("rev_stress", smp, [&] {
            LAUNCH_U(k_rev_stress, grdR, blk, tR, gg, m, wg, a, sg, acc, acc + N, acc + 2 * N,
                     ring, it);
          });
          reset(it);
          d.launch("adj_vel", smp, [&] {
            LAUNCH_U(k_adj_vel, grdUg, blk, tUg, gg, m, c, a, pg, sg, sgrad, it);
          });
          inject(it);
          d.launch("adj_stress", smp, [&] {
            LAUNCH_U(k_adj_stress, grdUg, blk, tUg, gg, m, c, a, pg, none);
          });
        }
      }
      if (two) {
        HIP_TRY(hipEventRecord(d.join[1], d.stream2));
        HIP_TRY(hipStreamWaitEvent(st, d.join[1], 0));
        const long long n5 = 5 * (long long)N;
        hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)((n5 + 255) / 256)), dim3(256), 0, st,
                           acc0, n5, 2);
      }
    }
#undef LAUNCH_U
    const Tiling tA = make_tiling(g.nz, g.nx, d.tile_x, d.xcd);
    hipLaunchKernelGGL(k_finalize, dim3(tA.n), blk, 0, st, tA, g, m, acc, d.grad);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(st));
  if (d.timing) d.collect_timing();
  if (if_res) HIP_TRY(hipMemcpy(&d.misfit_host, d.misfit, sizeof(double), hipMemcpyDeviceToHost));
  last_calc = calc_id;
  d.illum_valid = d.illum_on;
  return Status::Ok();
}

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
