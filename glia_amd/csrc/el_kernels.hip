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
                                                     ElMaterial m, Vec3 mu, float* __restrict__ lam, Vec3 scr,
                                                     double* __restrict__ part) {
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
    for (int d = 0; d < 3; d++) {
      mu.p[d][i] = a;
      scr.p[d][i] = sc;
    }
    lam[i] = b;
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
                    const float* vt, const float* bg, const ElMaterial& m, const Vec3& mu, float* lam, const Vec3& scr,
                    double* part) {
  k_el_material<<<EB, 256, 0, st>>>(ng, c, wm, gm, csf, vt, bg, m, mu, lam, scr, part);
  return (int)hipGetLastError();
}
int launch_force_scale(hipStream_t st, int64_t ng, const float* g, const float* c, float ff, int use_tanh, float* out) {
  k_force_scale<<<EB, 256, 0, st>>>(ng, g, c, ff, use_tanh, out);
  return (int)hipGetLastError();
}
int launch_axpby3(hipStream_t st, int64_t ng, const Vec3& p, const Vec3C& z, double b, int sub) {
  k_axpby3<<<EB, 256, 0, st>>>(ng, p, z, b, sub);
  return (int)hipGetLastError();
}

}  // namespace el
}  // namespace glia
