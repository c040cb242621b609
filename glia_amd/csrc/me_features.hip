// me_features.hip -- the biophysical features of the model-4 loop
// (PdeOperatorsMassEffect::computeBiophysicalFeatures, src/pde/PdeOperatorsMassEffect.cpp:323-418,
// with writeStats :261-321, computeStress :140-212, computeTumorQuants :214-259 and the statistics of
// src/utils/Utils.cpp:699-880).  Pointwise kernels, FP64 block reductions with a fixed-order finalize
// (deterministic), and order statistics from a device radix sort of masked keys.  Everything stays
// on the stream; the host reads one small array of sums per time step.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "me_features.h"

namespace glia {
namespace mef {

constexpr int RB = 1024;  // blocks of the reduction kernels (partials per sum)
constexpr int TPB = 256;

__device__ __forceinline__ bool in_tc(float s) { return s == 1.0f; }                  // getTCRecon
__device__ __forceinline__ bool in_healthy(float s) { return s != 0.0f && s != 1.0f; }  // getHealthyBrain

// VecField::computeMagnitude (src/utils/Utils.cpp:112-130)
__global__ void __launch_bounds__(TPB) k_mag3(int64_t n, const float* __restrict__ x, const float* __restrict__ y,
                                              const float* __restrict__ z, float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = sqrtf(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
}

// c v, the flux whose divergence is the advection term of dc/dt
__global__ void __launch_bounds__(TPB) k_flux(int64_t n, const float* __restrict__ c, const float* __restrict__ vx,
                                              const float* __restrict__ vy, const float* __restrict__ vz,
                                              float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float ci = c[i];
    ox[i] = ci * vx[i];
    oy[i] = ci * vy[i];
    oz[i] = ci * vz[i];
  }
}

// dc/dt = D c + rho (c - c^2) - div(c v), in the order of computeTumorQuants (:231-246)
__global__ void __launch_bounds__(TPB) k_dcdt(int64_t n, const float* __restrict__ dc, const float* __restrict__ c,
                                              const float* __restrict__ rho, const float* __restrict__ dv,
                                              float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float ci = c[i];
    const float r = (ci - ci * ci) * rho[i];
    out[i] = (dc[i] + r) - dv[i];
  }
}

// computeStressQuants: F = I + grad u (row d = gradient of u_d), det F, the linearized strain
// E = (F + F^T) / 2 - I, the stress S = lam tr(E) I + 2 mu E, tr S, and half the spread of the
// eigenvalues of S (the maximum shear stress).  The eigenvalues of the symmetric 3x3 S come from
// the trigonometric closed form of the characteristic cubic (Smith, CACM 4(4):168, 1961).
__global__ void __launch_bounds__(TPB) k_stress(int64_t n, StressIn g, const float* __restrict__ mu,
                                                const float* __restrict__ lam, float* __restrict__ jac,
                                                float* __restrict__ tr, float* __restrict__ tau) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    float f[9];
#pragma unroll
    for (int k = 0; k < 9; k++) f[k] = g.d[k][i];
    f[0] += 1.0f;
    f[4] += 1.0f;
    f[8] += 1.0f;
    jac[i] = f[0] * (f[4] * f[8] - f[7] * f[5]) - f[1] * (f[3] * f[8] - f[6] * f[5]) + f[2] * (f[3] * f[7] - f[6] * f[4]);
    const float m2 = 2.0f * mu[i], l = lam[i];
    const float exx = f[0] - 1.0f, eyy = f[4] - 1.0f, ezz = f[8] - 1.0f;
    const float exy = 0.5f * (f[1] + f[3]), exz = 0.5f * (f[2] + f[6]), eyz = 0.5f * (f[5] + f[7]);
    const float tre = exx + eyy + ezz;
    const float sxx = l * tre + m2 * exx, syy = l * tre + m2 * eyy, szz = l * tre + m2 * ezz;
    const float sxy = m2 * exy, sxz = m2 * exz, syz = m2 * eyz;
    const float trs = sxx + syy + szz;
    tr[i] = trs;
    const float off = sxy * sxy + sxz * sxz + syz * syz;
    float e0, e1, e2;
    if (off == 0.0f) {
      e0 = sxx;
      e1 = syy;
      e2 = szz;
    } else {
      const float q = trs / 3.0f;
      const float axx = sxx - q, ayy = syy - q, azz = szz - q;
      const float p = sqrtf((axx * axx + ayy * ayy + azz * azz + 2.0f * off) / 6.0f);
      const float ip = 1.0f / p;
      // half the determinant of (S - q I) / p, clamped to [-1, 1]
      const float det = axx * (ayy * azz - syz * syz) - sxy * (sxy * azz - syz * sxz) + sxz * (sxy * syz - ayy * sxz);
      const float r = ip * ip * ip * 0.5f * det;
      const float phi = (r <= -1.0f) ? (float)(M_PI / 3.0) : (r >= 1.0f ? 0.0f : acosf(r) / 3.0f);
      e0 = q + 2.0f * p * cosf(phi);
      e1 = q + 2.0f * p * cosf(phi + (float)(2.0 * M_PI / 3.0));
      e2 = 3.0f * q - e0 - e1;
    }
    const float lo = fminf(e0, fminf(e1, e2)), hi = fmaxf(e0, fmaxf(e1, e2));
    tau[i] = 0.5f * (hi - lo);
  }
}

// fixed-order sum of the per-block partials of nsum sums into acc[slot + s]
__global__ void __launch_bounds__(TPB) k_finalize(const double* __restrict__ part, int nsum, double* __restrict__ acc,
                                                  int slot) {
  __shared__ double red[TPB];
  for (int s = 0; s < nsum; s++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < RB; b += TPB) v += part[(int64_t)s * RB + b];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) acc[slot + s] = red[0];
    __syncthreads();
  }
}

template <int NS>
__device__ __forceinline__ void block_sums(double (&v)[NS], double* __restrict__ part) {
  __shared__ double red[NS][TPB];
#pragma unroll
  for (int s = 0; s < NS; s++) red[s][threadIdx.x] = v[s];
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int s = 0; s < NS; s++) red[s][threadIdx.x] += red[s][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int s = 0; s < NS; s++) part[(int64_t)s * RB + blockIdx.x] = red[s][0];
}

// the location of the largest |grad c| outside the tumor core (computeTumorQuants :248-257): the
// largest value of g (1 - tc), the first index among equal values
__global__ void __launch_bounds__(TPB) k_argmax(int64_t n, const float* __restrict__ g, const float* __restrict__ seg,
                                                float* __restrict__ pv, long long* __restrict__ pi) {
  float best = -INFINITY;
  long long bi = n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float w = g[i] - g[i] * (in_tc(seg[i]) ? 1.0f : 0.0f);
    if (w > best) {
      best = w;
      bi = i;
    }
  }
  __shared__ float rv[TPB];
  __shared__ long long ri[TPB];
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bi;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const float a = rv[threadIdx.x + w];
      const long long ai = ri[threadIdx.x + w];
      if (a > rv[threadIdx.x] || (a == rv[threadIdx.x] && ai < ri[threadIdx.x])) {
        rv[threadIdx.x] = a;
        ri[threadIdx.x] = ai;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = rv[0];
    pi[blockIdx.x] = ri[0];
  }
}

// one block: the argmax over the block partials, then c* = c[index]
__global__ void __launch_bounds__(TPB) k_argmax_fin(const float* __restrict__ pv, const long long* __restrict__ pi,
                                                    const float* __restrict__ c, int64_t n, float* __restrict__ cstar) {
  __shared__ float rv[TPB];
  __shared__ long long ri[TPB];
  float best = -INFINITY;
  long long bi = n;
  for (int b = threadIdx.x; b < RB; b += TPB) {
    if (pv[b] > best || (pv[b] == best && pi[b] < bi)) {
      best = pv[b];
      bi = pi[b];
    }
  }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bi;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const float a = rv[threadIdx.x + w];
      const long long ai = ri[threadIdx.x + w];
      if (a > rv[threadIdx.x] || (a == rv[threadIdx.x] && ai < ri[threadIdx.x])) {
        rv[threadIdx.x] = a;
        ri[threadIdx.x] = ai;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *cstar = (ri[0] < n) ? c[ri[0]] : 0.0f;
}

// sums of writeStats: x, x tc, x healthy, I x (I = |c - c*| < 0.1, computeIndicatorFunction with
// its threshold of 1e-1, src/utils/Utils.cpp:1075-1100), and the voxel counts of tc and healthy
__global__ void __launch_bounds__(TPB) k_sums(int64_t n, const float* __restrict__ x, const float* __restrict__ seg,
                                              const float* __restrict__ c, const float* __restrict__ cstar,
                                              double* __restrict__ part) {
  double v[6] = {0, 0, 0, 0, 0, 0};
  const float cs = *cstar;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float xi = x[i];
    const float s = seg[i];
    // masks as 0/1 factors, as the reference multiplies them (VecPointwiseMult)
    const float t = in_tc(s) ? 1.0f : 0.0f, hb = in_healthy(s) ? 1.0f : 0.0f;
    const float ind = (fabsf(c[i] - cs) < 0.1f) ? 1.0f : 0.0f;
    v[0] += (double)xi;
    v[1] += (double)(t * xi);
    v[2] += (double)(hb * xi);
    v[3] += (double)(ind * xi);
    v[4] += (double)t;
    v[5] += (double)hb;
  }
  block_sums<6>(v, part);
}

// the two-pass variances of computeStd: sums of (x - mean)^2 over the grid, tc and healthy, with the
// means taken from the sums in acc[slot..] (single precision shifts, as VecShift applies them)
__global__ void __launch_bounds__(TPB) k_sqdev(int64_t n, const float* __restrict__ x, const float* __restrict__ seg,
                                               const double* __restrict__ acc, int slot, double* __restrict__ part) {
  const float m = (float)(acc[slot + 0] / (double)n);
  const float mt = (float)(acc[slot + 1] / acc[slot + 4]);
  const float mh = (float)(acc[slot + 2] / acc[slot + 5]);
  double v[3] = {0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float xi = x[i];
    const float s = seg[i];
    const float a = xi - m;
    const float b = (xi - mt) * (in_tc(s) ? 1.0f : 0.0f);
    const float d = (xi - mh) * (in_healthy(s) ? 1.0f : 0.0f);
    v[0] += (double)(a * a);
    v[1] += (double)(b * b);
    v[2] += (double)(d * d);
  }
  block_sums<3>(v, part);
}

// sort keys of a region: x inside (all, tc or healthy), +inf outside, so that the region's values
// come first in ascending order
__global__ void __launch_bounds__(TPB) k_keys(int64_t n, const float* __restrict__ x, const float* __restrict__ seg,
                                              int region, float* __restrict__ keys) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float s = seg[i];
    const bool in = region == 0 || (region == 1 ? in_tc(s) : in_healthy(s));
    keys[i] = in ? x[i] : INFINITY;
  }
}

// computeQuantile: the entry floor(q m) of the m sorted values of the region (q and m in single
// precision, as the reference multiplies them), for q = 0.875, 0.95, 0.99; 0 for an empty region
__global__ void k_pick(const float* __restrict__ sorted, const double* __restrict__ acc, int count_slot, int64_t n,
                       double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const long long m = (count_slot < 0) ? (long long)n : (long long)acc[count_slot];
  const float qs[3] = {0.875f, 0.95f, 0.99f};
  for (int k = 0; k < 3; k++) {
    if (m <= 0) {
      out[k] = 0.0;
      continue;
    }
    long long idx = (long long)floorf(qs[k] * (float)m);
    if (idx >= m) idx = m - 1;
    out[k] = (double)sorted[idx];
  }
}

inline dim3 grid_of(int64_t n) {
  const int64_t b = (n + TPB - 1) / TPB;
  return dim3((unsigned)(b < 4096 ? b : 4096));
}

#define MEF_LAUNCH()                       \
  do {                                     \
    const hipError_t e_ = hipGetLastError(); \
    if (e_ != hipSuccess) return (int)e_;  \
  } while (0)

int launch_mag3(hipStream_t st, int64_t n, const float* x, const float* y, const float* z, float* out) {
  k_mag3<<<grid_of(n), TPB, 0, st>>>(n, x, y, z, out);
  MEF_LAUNCH();
  return 0;
}

int launch_flux(hipStream_t st, int64_t n, const float* c, const float* const v[3], float* const o[3]) {
  k_flux<<<grid_of(n), TPB, 0, st>>>(n, c, v[0], v[1], v[2], o[0], o[1], o[2]);
  MEF_LAUNCH();
  return 0;
}

int launch_dcdt(hipStream_t st, int64_t n, const float* dc, const float* c, const float* rho, const float* dv,
                float* out) {
  k_dcdt<<<grid_of(n), TPB, 0, st>>>(n, dc, c, rho, dv, out);
  MEF_LAUNCH();
  return 0;
}

int launch_stress(hipStream_t st, int64_t n, const StressIn& g, const float* mu, const float* lam, float* jac,
                  float* tr, float* tau) {
  k_stress<<<grid_of(n), TPB, 0, st>>>(n, g, mu, lam, jac, tr, tau);
  MEF_LAUNCH();
  return 0;
}

size_t scratch_bytes(int64_t n) {
  size_t sort_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (const float*)nullptr, (float*)nullptr, (int)n);
  // partials: 6 sums of RB doubles, RB floats + RB indices of the argmax, then the sort's storage
  return sizeof(double) * 6 * RB + sizeof(float) * RB + sizeof(long long) * RB + ((sort_bytes + 255) / 256) * 256;
}

int launch_cstar(hipStream_t st, int64_t n, const float* gradc, const float* seg, const float* c, void* scratch,
                 float* cstar) {
  char* p = (char*)scratch + sizeof(double) * 6 * RB;
  float* pv = (float*)p;
  long long* pi = (long long*)(p + sizeof(float) * RB);
  k_argmax<<<RB, TPB, 0, st>>>(n, gradc, seg, pv, pi);
  MEF_LAUNCH();
  k_argmax_fin<<<1, TPB, 0, st>>>(pv, pi, c, n, cstar);
  MEF_LAUNCH();
  return 0;
}

int launch_stats(hipStream_t st, int64_t n, const float* x, const float* seg, const float* c, const float* cstar,
                 void* scratch, float* keys_in, float* keys_out, double* acc, int slot) {
  double* part = (double*)scratch;
  k_sums<<<RB, TPB, 0, st>>>(n, x, seg, c, cstar, part);
  MEF_LAUNCH();
  k_finalize<<<1, TPB, 0, st>>>(part, 6, acc, slot);
  MEF_LAUNCH();
  k_sqdev<<<RB, TPB, 0, st>>>(n, x, seg, acc, slot, part);
  MEF_LAUNCH();
  k_finalize<<<1, TPB, 0, st>>>(part, 3, acc, slot + 6);
  MEF_LAUNCH();
  char* sort_tmp = (char*)scratch + sizeof(double) * 6 * RB + sizeof(float) * RB + sizeof(long long) * RB;
  size_t sort_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (const float*)keys_in, keys_out, (int)n);
  if (e != hipSuccess) return (int)e;
  for (int r = 0; r < 3; r++) {
    k_keys<<<grid_of(n), TPB, 0, st>>>(n, x, seg, r, keys_in);
    MEF_LAUNCH();
    e = hipcub::DeviceRadixSort::SortKeys(sort_tmp, sort_bytes, (const float*)keys_in, keys_out, (int)n, 0, 32, st);
    if (e != hipSuccess) return (int)e;
    // the quantiles 0.875, 0.95, 0.99 of region r land in acc[slot + 9 + 3 r + k], k = 0, 1, 2
    k_pick<<<1, 64, 0, st>>>(keys_out, acc, r == 0 ? -1 : slot + 3 + r, n, acc + slot + 9 + 3 * r);
    MEF_LAUNCH();
  }
  return 0;
}

}  // namespace mef
}  // namespace glia
