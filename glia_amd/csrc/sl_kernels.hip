// sl_kernels.hip -- semi-Lagrangian advection on gfx950 (SemiLagrangianSolver,
// src/pde/AdvectionSolver.cpp:279-704).
//
// A field value at a departure point x + q(x) is the tensor-product cubic Lagrange interpolant of
// the 4 x 4 x 4 grid values around it on the periodic grid (the reference's GPU path evaluates the
// same interpolant with a wrap-mode texture, cubicTex3DFastLagrange,
// src/cuda/InterpolationTexture.cu:270-366).  Offsets q are kept relative to the grid point, in grid
// units, so the fraction that sets the weights is exact in FP32 at any grid size.
//
// One thread per grid point; consecutive threads are consecutive z points, so for a smooth
// velocity the 16 rows of 4 taps a wave gathers are nearly contiguous and are served by L1/L2.
// Each row of 4 z taps is one 16-byte buffer load.
// HBM traffic per point: the offsets (12 B) once, each field read once through the gathers (4 B)
// and written once (4 B).
#include <hip/hip_runtime.h>

#include "sl_launch.h"

namespace glia {
namespace sl {

// lagrange_weights (3rdparty/interp_poly_kernel.cu:56-69): cubic Lagrange basis on the nodes -1, 0,
// 1, 2 at fraction t in [0, 1)
__device__ __forceinline__ void lag_weights(float t, float w[4]) {
  const float tp1 = t + 1.0f, tm1 = t - 1.0f, tm2 = t - 2.0f;
  w[0] = (-1.0f / 6.0f) * t * tm1 * tm2;
  w[1] = 0.5f * tp1 * tm1 * tm2;
  w[2] = -0.5f * tp1 * t * tm2;
  w[3] = (1.0f / 6.0f) * tp1 * t * tm1;
}

__device__ __forceinline__ int wrap(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// the 64 taps of one departure point: row offsets (x, y) and z indices with their weights
struct Stencil {
  int row[4][4];  // (ix * n1 + iy) * n2 (< 2^31: grids up to 1024^3 / 2)
  int iz[4];
  float wx[4], wy[4], wz[4];
};

__device__ __forceinline__ void make_stencil(Stencil& s, const SLDims& d, int i, int j, int k, float qx, float qy,
                                             float qz) {
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  lag_weights(qx - fx, s.wx);
  lag_weights(qy - fy, s.wy);
  lag_weights(qz - fz, s.wz);
  const int bx = i + (int)fx - 1, by = j + (int)fy - 1, bz = k + (int)fz - 1;
  int ix[4], iy[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    ix[a] = wrap(bx + a, d.n0);
    iy[a] = wrap(by + a, d.n1);
    s.iz[a] = wrap(bz + a, d.n2);
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) s.row[a][b] = (ix[a] * d.n1 + iy[b]) * d.n2;
}

// Fields are read through buffer resources: a row of 4 z taps is one buffer_load_dwordx4, which
// needs only dword alignment, and reads outside the field (which the wrapped indices never make)
// return 0 instead of faulting.  Word 3 of the descriptor selects the 32-bit data format.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const float* f, const SLDims& d) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)f, (short)0, (int)(d.ng * 4), 0x00020000);
}

__device__ __forceinline__ float apply_stencil(const Stencil& s, __amdgpu_buffer_rsrc_t f) {
  float acc = 0.0f;
  // the z taps of a row are contiguous unless the stencil wraps around z (3 of n2 columns)
  const bool contiguous = s.iz[3] == s.iz[0] + 3;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    float ya = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int r = s.row[a][b];
      float t0, t1, t2, t3;
      if (contiguous) {
        const u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(f, (r + s.iz[0]) * 4, 0, 0);
        t0 = __uint_as_float(q.x);
        t1 = __uint_as_float(q.y);
        t2 = __uint_as_float(q.z);
        t3 = __uint_as_float(q.w);
      } else {
        t0 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(f, (r + s.iz[0]) * 4, 0, 0));
        t1 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(f, (r + s.iz[1]) * 4, 0, 0));
        t2 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(f, (r + s.iz[2]) * 4, 0, 0));
        t3 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(f, (r + s.iz[3]) * 4, 0, 0));
      }
      const float zb = s.wz[0] * t0 + s.wz[1] * t1 + s.wz[2] * t2 + s.wz[3] * t3;
      ya += s.wy[b] * zb;
    }
    acc += s.wx[a] * ya;
  }
  return acc;
}

__device__ __forceinline__ void point_of(const SLDims& d, int64_t p, int& i, int& j, int& k) {
  const int64_t r = p / d.n2;
  k = (int)(p - r * d.n2);
  j = (int)(r % d.n1);
  i = (int)(r / d.n1);
}

// computeTrajectories (src/pde/AdvectionSolver.cpp:470-590): Euler point X* = x - s v, v* = v(X*),
// departure offset q = -(s/2) (v + v*), with s = dt / h per axis in grid units
__global__ void __launch_bounds__(256) k_sl_trajectories(SLDims d, const float* __restrict__ vx,
                                                         const float* __restrict__ vy, const float* __restrict__ vz,
                                                         float sx, float sy, float sz, float* __restrict__ qx,
                                                         float* __restrict__ qy, float* __restrict__ qz) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.ng) return;
  int i, j, k;
  point_of(d, p, i, j, k);
  const float ux = vx[p], uy = vy[p], uz = vz[p];
  Stencil s;
  make_stencil(s, d, i, j, k, -sx * ux, -sy * uy, -sz * uz);
  // one component at a time (a loop the compiler keeps), so the 64 taps of one field are in
  // registers at once: 3 waves per SIMD instead of 2
  const float* vs[3] = {vx, vy, vz};
  float w[3];
#pragma unroll 1
  for (int a = 0; a < 3; a++) w[a] = apply_stencil(s, field_rsrc(vs[a], d));
  qx[p] = -0.5f * sx * (ux + w[0]);
  qy[p] = -0.5f * sy * (uy + w[1]);
  qz[p] = -0.5f * sz * (uz + w[2]);
}

// SemiLagrangianSolver::solve (src/pde/AdvectionSolver.cpp:592-660) for up to SL_MAX_FIELDS fields
// on one set of departure offsets.  mode 2: out = in(X).  mode 1 (mass conservation):
// ci = in(X), src = -ci div_v(X), c* = ci + dt src, out = ci + dt/2 (src - c* div_v(x)).
__global__ void __launch_bounds__(256) k_sl_apply(SLDims d, SLFields F, const float* __restrict__ qx,
                                                  const float* __restrict__ qy, const float* __restrict__ qz,
                                                  const float* __restrict__ dv, const float* __restrict__ dvi,
                                                  float dt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.ng) return;
  int i, j, k;
  point_of(d, p, i, j, k);
  Stencil s;
  make_stencil(s, d, i, j, k, qx[p], qy[p], qz[p]);
  float dvp = 0.0f, dvip = 0.0f;
  if (F.any_mass) {
    dvp = dv[p];
    dvip = dvi[p];
  }
  for (int f = 0; f < F.n; f++) {
    const float ci = apply_stencil(s, field_rsrc(F.in[f], d));
    float out = ci;
    if (F.mode[f] == 1) {
      const float src = -ci * dvip;
      const float cstar = ci + dt * src;
      out = ci + 0.5f * dt * (src - cstar * dvp);
    }
    F.out[f][p] = out;
  }
}

static unsigned blocks_for(int64_t ng) { return (unsigned)((ng + 255) / 256); }

int launch_sl_trajectories(const SLDims& d, hipStream_t st, const float* vx, const float* vy, const float* vz,
                           float sx, float sy, float sz, float* qx, float* qy, float* qz) {
  k_sl_trajectories<<<blocks_for(d.ng), 256, 0, st>>>(d, vx, vy, vz, sx, sy, sz, qx, qy, qz);
  return (int)hipGetLastError();
}

int launch_sl_apply(const SLDims& d, hipStream_t st, const SLFields& F, const float* qx, const float* qy,
                    const float* qz, const float* dv, const float* dvi, float dt) {
  k_sl_apply<<<blocks_for(d.ng), 256, 0, st>>>(d, F, qx, qy, qz, dv, dvi, dt);
  return (int)hipGetLastError();
}

}  // namespace sl
}  // namespace glia
