// el_kernels.hip -- pointwise and reduction kernels of the linear-elasticity solver of the
// mass-effect model (VariableLinearElasticitySolver, src/pde/ElasticitySolver.cpp:6-533).  The
// spectral parts (transforms, the strain/force combinations, the preconditioner) run on the RD
// engine's packed spectra; these kernels work on plain FP32 fields of ng values.
#include <hip/hip_runtime.h>

#include "el_launch.h"

namespace glia {
namespace el {

constexpr int EB = 1024;  // blocks of the grid-stride kernels (the partial-sum count)

// computeMaterialProperties (src/pde/ElasticitySolver.cpp:376-455) with computeTumorLame and
// computeScreening (src/cuda/ElasticitySolver.cu:13-34); partial sums of mu, lam and screen per block
__global__ void __launch_bounds__(256) k_el_material(int64_t ng, const float* __restrict__ c, const float* __restrict__ wm,
                                                     const float* __restrict__ gm, const float* __restrict__ csf,
                                                     const float* __restrict__ vt, const float* __restrict__ bg,
                                                     ElMaterial m, float* __restrict__ mu, float* __restrict__ lam,
                                                     float* __restrict__ scr, double* __restrict__ part) {
  double s[3] = {0.0, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride) {
    float a = m.mu_bg * bg[i] + m.mu_h * wm[i] + m.mu_h * gm[i] + m.mu_c * vt[i] + m.mu_c * csf[i];
    float b = m.lam_bg * bg[i] + m.lam_h * wm[i] + m.lam_h * gm[i] + m.lam_c * vt[i] + m.lam_c * csf[i];
    a = a > 0.0f ? a : 0.0f;
    b = b > 0.0f ? b : 0.0f;
    const float ci = c[i];
    if (ci > 0.0f) {
      a += m.mu_t * ci;
      b += m.lam_t * ci;
    }
    float sc = (ci >= 0.005f) ? m.screen_low : m.screen_high;
    if (bg[i] > 0.95f) sc = 1e6f;
    mu[i] = a;
    lam[i] = b;
    scr[i] = sc;
    s[0] += a;
    s[1] += b;
    s[2] += sc;
  }
  __shared__ double red[3][256];
  for (int k = 0; k < 3; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// per-block FP64 partials of <x, y> over the three components
__global__ void __launch_bounds__(256) k_dot3(int64_t ng, Vec3C x, Vec3C y, double* __restrict__ part) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int k = 0; k < 3; k++)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride)
      s += (double)x.p[k][i] * (double)y.p[k][i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// per-block FP64 partials of <z, z> and <z, r> in one pass (part[2 b], part[2 b + 1])
__global__ void __launch_bounds__(256) k_dot3_zz_zr(int64_t ng, Vec3C z, Vec3C r, double* __restrict__ part) {
  double s0 = 0.0, s1 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int k = 0; k < 3; k++)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride) {
      const double zi = z.p[k][i];
      s0 += zi * zi;
      s1 += zi * (double)r.p[k][i];
    }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) part[2 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// fixed-order sums of nsum interleaved partial arrays (part[nsum b + s], b < np) into out[s]
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ part, int np, int nsum,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  for (int s = 0; s < nsum; s++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < np; b += 256) v += part[(int64_t)nsum * b + s];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
    __syncthreads();
  }
}

// the CG update with the step a = beta / dpi taken on the device, dpi = <p, A p> from *dpi; the
// update is skipped when the host will stop on dpi (NaN or infinite, zero, or a sign change from
// dpiold after the first iteration), so x and r stay as the reference leaves them
__global__ void __launch_bounds__(256) k_cg_update_dev(int64_t ng, Vec3 x, Vec3 r, Vec3C p, Vec3C w,
                                                       const double* __restrict__ dpi, double beta, double dpiold,
                                                       int first) {
  const double q = *dpi;
  if (isnan(q) || isinf(q) || q == 0.0 || (!first && q * dpiold <= 0.0)) return;
  const float fa = (float)(beta / q);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int k = 0; k < 3; k++)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride) {
      x.p[k][i] += fa * p.p[k][i];
      r.p[k][i] -= fa * w.p[k][i];
    }
}

// p = z + b p (b = 0: p = z); y = x - y when sub (the initial residual b - A x)
__global__ void __launch_bounds__(256) k_axpby3(int64_t ng, Vec3 p, Vec3C z, double b, int sub) {
  const float fb = (float)b;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int k = 0; k < 3; k++)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride)
      p.p[k][i] = sub ? z.p[k][i] - p.p[k][i] : z.p[k][i] + fb * p.p[k][i];
}

// computeForce's scaling (src/mat/Tumor.cpp:114-158, nonlinearForceScalingCuda): out = ff g, times
// tanh(c) when use_tanh
__global__ void __launch_bounds__(256) k_force_scale(int64_t ng, const float* __restrict__ g, const float* __restrict__ c,
                                                     float ff, int use_tanh, float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride)
    out[i] = use_tanh ? ff * g[i] * tanhf(c[i]) : ff * g[i];
}

int partial_count() { return EB; }

int launch_material(hipStream_t st, int64_t ng, const float* c, const float* wm, const float* gm, const float* csf,
                    const float* vt, const float* bg, const ElMaterial& m, float* mu, float* lam, float* scr,
                    double* part) {
  k_el_material<<<EB, 256, 0, st>>>(ng, c, wm, gm, csf, vt, bg, m, mu, lam, scr, part);
  return (int)hipGetLastError();
}
int launch_force_scale(hipStream_t st, int64_t ng, const float* g, const float* c, float ff, int use_tanh, float* out) {
  k_force_scale<<<EB, 256, 0, st>>>(ng, g, c, ff, use_tanh, out);
  return (int)hipGetLastError();
}
int launch_dot3(hipStream_t st, int64_t ng, const Vec3C& x, const Vec3C& y, double* part) {
  k_dot3<<<EB, 256, 0, st>>>(ng, x, y, part);
  return (int)hipGetLastError();
}
int launch_dot3_zz_zr(hipStream_t st, int64_t ng, const Vec3C& z, const Vec3C& r, double* part) {
  k_dot3_zz_zr<<<EB, 256, 0, st>>>(ng, z, r, part);
  return (int)hipGetLastError();
}
int launch_sum_partials(hipStream_t st, const double* part, int nsum, double* out) {
  k_sum_partials<<<1, 256, 0, st>>>(part, EB, nsum, out);
  return (int)hipGetLastError();
}
int launch_cg_update_dev(hipStream_t st, int64_t ng, const Vec3& x, const Vec3& r, const Vec3C& p, const Vec3C& w,
                         const double* dpi, double beta, double dpiold, int first) {
  k_cg_update_dev<<<EB, 256, 0, st>>>(ng, x, r, p, w, dpi, beta, dpiold, first);
  return (int)hipGetLastError();
}
int launch_axpby3(hipStream_t st, int64_t ng, const Vec3& p, const Vec3C& z, double b, int sub) {
  k_axpby3<<<EB, 256, 0, st>>>(ng, p, z, b, sub);
  return (int)hipGetLastError();
}

}  // namespace el
}  // namespace glia
