// me_kernels.hip -- pointwise kernels of the mass-effect forward (PdeOperatorsMassEffect::solveState,
// src/pde/PdeOperatorsMassEffect.cpp:421-743) on plain FP32 fields of ng values.  The transforms,
// the elasticity solve and the advection run on their own solvers' kernels.
#include <hip/hip_runtime.h>

#include "me_launch.h"

namespace glia {
namespace me {

constexpr int MB = 1024;  // blocks of the grid-stride kernels (the partial count of launch_vmax)

// MatProp::filterTumor (src/mat/MatProp.cpp:205-224)
__global__ void __launch_bounds__(256) k_filter(int64_t ng, const float* __restrict__ c, float* __restrict__ wm,
                                                float* __restrict__ gm) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride) {
    const float s = 1.0f - c[i];
    wm[i] *= s;
    gm[i] *= s;
  }
}

// VecField::computeMagnitude + VecMax (src/pde/PdeOperatorsMassEffect.cpp:490-491): per-block max |v|
__global__ void __launch_bounds__(256) k_vmax(int64_t ng, const float* __restrict__ vx, const float* __restrict__ vy,
                                              const float* __restrict__ vz, float* __restrict__ part) {
  float m = 0.0f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride)
    m = fmaxf(m, sqrtf(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
  __shared__ float red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// PdeOperatorsRD::reaction(0) (src/pde/PdeOperators.cpp:140-176): c <- a f / (a f + 1) with
// a = c / (1 - c) and f = exp(rho dt); c = 1 stays 1
__global__ void __launch_bounds__(256) k_reaction(int64_t ng, float* __restrict__ c, const float* __restrict__ rho,
                                                  float dt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride) {
    const float f = expf(rho[i] * dt);
    const float a = c[i] / (1.0f - c[i]);
    c[i] = isinf(a) ? 1.0f : a * f / (a * f + 1.0f);
  }
}

struct P3 {
  float* p[3];
};
struct P3C {
  const float* p[3];
};

// velocity from two displacements (src/pde/PdeOperatorsMassEffect.cpp:638-643) and the copy of
// the displacement into u_old (:664)
__global__ void __launch_bounds__(256) k_velocity(int64_t ng, P3 v, P3C u, P3 uold, float inv_dt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int k = 0; k < 3; k++)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride) {
      const float ui = u.p[k][i];
      v.p[k][i] = (ui - uold.p[k][i]) * inv_dt;
      uold.p[k][i] = ui;
    }
}

// the multispecies start: p = c0 clipped at zero (clipTumor), i = 0, ox = 1
__global__ void __launch_bounds__(256) k_species_init(int64_t ng, const float* __restrict__ c0, float* __restrict__ p,
                                                      float* __restrict__ iv, float* __restrict__ ox) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ng; k += stride) {
    const float c = c0[k];
    p[k] = c <= 0.0f ? 0.0f : c;
    iv[k] = 0.0f;
    ox[k] = 1.0f;
  }
}

// out = a + b (+ c): the tumor core p + nec + i and the start's p + i
__global__ void __launch_bounds__(256) k_sum(int64_t ng, const float* __restrict__ a, const float* __restrict__ b,
                                             const float* __restrict__ c, float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ng; k += stride)
    out[k] = c ? a[k] + b[k] + c[k] : a[k] + b[k];
}

int partial_count() { return MB; }
int launch_species_init(hipStream_t st, int64_t ng, const float* c0, float* p, float* i, float* ox) {
  k_species_init<<<MB, 256, 0, st>>>(ng, c0, p, i, ox);
  return (int)hipGetLastError();
}
int launch_sum(hipStream_t st, int64_t ng, const float* a, const float* b, const float* c, float* out) {
  k_sum<<<MB, 256, 0, st>>>(ng, a, b, c, out);
  return (int)hipGetLastError();
}

int launch_filter(hipStream_t st, int64_t ng, const float* c, float* wm, float* gm) {
  k_filter<<<MB, 256, 0, st>>>(ng, c, wm, gm);
  return (int)hipGetLastError();
}
int launch_vmax(hipStream_t st, int64_t ng, const float* vx, const float* vy, const float* vz, float* part) {
  k_vmax<<<MB, 256, 0, st>>>(ng, vx, vy, vz, part);
  return (int)hipGetLastError();
}
int launch_reaction(hipStream_t st, int64_t ng, float* c, const float* rho, float dt) {
  k_reaction<<<MB, 256, 0, st>>>(ng, c, rho, dt);
  return (int)hipGetLastError();
}
int launch_velocity(hipStream_t st, int64_t ng, float* const v[3], const float* const u[3], float* const uold[3],
                    float inv_dt) {
  P3 pv{{v[0], v[1], v[2]}};
  P3C pu{{u[0], u[1], u[2]}};
  P3 po{{uold[0], uold[1], uold[2]}};
  k_velocity<<<MB, 256, 0, st>>>(ng, pv, pu, po, inv_dt);
  return (int)hipGetLastError();
}

}  // namespace me
}  // namespace glia
