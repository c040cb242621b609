// ms_kernels.hip -- pointwise terms of the multispecies model (include/glia_ms.h).  The reference
// computes the reaction rate, the two transition rates and the death switch into four work fields
// and then applies the sources (src/pde/PdeOperatorsMultiSpecies.cpp:10-190); here the rates live
// in registers and one pass reads p, i, n, ox, gm, wm, rho, di and writes p, i, n, ox, gm, wm.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/glia_ms.h"

namespace glia {
namespace ms {

struct Rates {
  float dt, death_rate, ox_source, ox_consumption, alpha_0, beta_0, ox_inv, ox_hypoxia;
};

__global__ void __launch_bounds__(256) k_ms_sources(int64_t n, float* __restrict__ p, float* __restrict__ iv,
                                                    float* __restrict__ nec, float* __restrict__ ox,
                                                    float* __restrict__ gm, float* __restrict__ wm,
                                                    const float* __restrict__ rho, const float* __restrict__ di,
                                                    Rates R) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const float o = ox[k], pk = p[k], ik = iv[k], g = gm[k], w = wm[k];
    // computeReactionRate, computeTransition, computeThesholder (src/cuda/PdeOperators.cu:47-78)
    const float m = rho[k] * (1.0f / (1.0f + expf(-100.0f * (o - R.ox_hypoxia))));
    const float al = R.alpha_0 * (1.0f / (1.0f + expf(100.0f * (o - R.ox_inv))));
    const float be = R.beta_0 * o;
    const float h = 1.0f / (1.0f + expf(100.0f * (o - R.ox_hypoxia)));
    // computeSources (src/cuda/PdeOperators.cu:80-110); ox_heal, reac_ratio and death_ratio are 1
    const float gp = m * pk * (1.0f - pk), gi = m * ik * (1.0f - ik), dh = h * R.death_rate;
    p[k] = pk + R.dt * (gp - al * pk + be * ik - dh * pk);
    iv[k] = ik + R.dt * (gi + al * pk - be * ik - dh * ik);
    nec[k] += R.dt * (dh * (pk + ik + g + w));
    const float on = o + R.dt * (-R.ox_consumption * pk + R.ox_source * (1.0f - o) * (g + w));
    ox[k] = (on <= 0.0f) ? 0.0f : on;
    float f1 = 0.0f, f2 = 0.0f;
    if (g > 0.01f || w > 0.01f) {
      f1 = g / (g + w);
      f2 = w / (g + w);
    }
    f1 = isnan(f1) ? 0.0f : f1;
    f2 = isnan(f2) ? 0.0f : f2;
    const float loss = gp + gi + di[k];
    gm[k] = g - R.dt * (f1 * loss + dh * g);
    wm[k] = w - R.dt * (f2 * loss + dh * w);
  }
}

__global__ void __launch_bounds__(256) k_ms_coefficients(int64_t n, float* __restrict__ rho_f, float* __restrict__ k_f,
                                                         const float* __restrict__ bg, const float* __restrict__ gm,
                                                         const float* __restrict__ vt, const float* __restrict__ csf,
                                                         float rho, float kk, float gm_r_scale, float gm_k_scale) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const float base = bg[k] + vt[k] + csf[k], g = gm[k];
    const float tr = 1.0f - (base + gm_r_scale * g), tk = 1.0f - (base + gm_k_scale * g);
    rho_f[k] = (tr < 0.0f ? 0.0f : tr) * rho;
    k_f[k] = (tk < 0.0f ? 0.0f : tk) * kk;
  }
}

__global__ void __launch_bounds__(256) k_ms_segmentation(int64_t n, const float* __restrict__ bg,
                                                         const float* __restrict__ gm, const float* __restrict__ wm,
                                                         const float* __restrict__ vt, const float* __restrict__ csf,
                                                         const float* __restrict__ c, float* __restrict__ seg) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const float ck = c[k], a = 1.0f - ck;
    float mx = bg[k];
    int ct = 0;
    if (ck > mx) { mx = ck; ct = 1; }
    if (wm[k] * a > mx) { mx = wm[k] * a; ct = 6; }
    if (gm[k] * a > mx) { mx = gm[k] * a; ct = 5; }
    if (vt[k] * a > mx) { mx = vt[k] * a; ct = 7; }
    if (csf[k] * a > mx) { mx = csf[k] * a; ct = 8; }
    seg[k] = (float)ct;
  }
}

__global__ void __launch_bounds__(256) k_ms_clip(int64_t n, float* __restrict__ x) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) x[k] = (x[k] <= 0.0f) ? 0.0f : x[k];
}

__global__ void __launch_bounds__(256) k_ms_edema(int64_t n, const float* __restrict__ iv, float thr, float* __restrict__ ed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) ed[k] = (iv[k] > thr) ? 1.0f : 0.0f;
}

static unsigned blocks(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 4096); }

}  // namespace ms
}  // namespace glia

extern "C" {

void glia_ms_params_init(glia_ms_params* p) {
  if (!p) return;
  p->death_rate = 4;
  p->ox_source = 1;
  p->ox_consumption = 5;
  p->alpha_0 = 0.1;
  p->beta_0 = 0.01;
  p->ox_inv = 0.8;
  p->ox_hypoxia = 0.6;
  p->invasive_threshold = 0.001;
}

int glia_ms_sources(int64_t n, float* p, float* i, float* nec, float* ox, float* gm, float* wm, const float* rho,
                    const float* di, double dt, const glia_ms_params* prm, void* stream) {
  if (n < 0 || !p || !i || !nec || !ox || !gm || !wm || !rho || !di) return GLIA_RD_ERR_ARG;
  if (n == 0) return 0;
  glia_ms_params q;
  glia_ms_params_init(&q);
  if (prm) q = *prm;
  glia::ms::Rates R{(float)dt, (float)q.death_rate, (float)q.ox_source, (float)q.ox_consumption, (float)q.alpha_0,
                    (float)q.beta_0, (float)q.ox_inv, (float)q.ox_hypoxia};
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  glia::ms::k_ms_sources<<<blocks, 256, 0, (hipStream_t)stream>>>(n, p, i, nec, ox, gm, wm, rho, di, R);
  return (int)hipGetLastError();
}

int glia_ms_update_coefficients(int64_t n, float* rho_field, float* k_field, const float* bg, const float* gm,
                                const float* vt, const float* csf, double rho, double k, double r_gm_wm,
                                double k_gm_wm, void* stream) {
  if (n < 0 || !rho_field || !k_field || !bg || !gm || !vt || !csf) return GLIA_RD_ERR_ARG;
  if (n == 0) return 0;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  glia::ms::k_ms_coefficients<<<blocks, 256, 0, (hipStream_t)stream>>>(
      n, rho_field, k_field, bg, gm, vt, csf, (float)rho, (float)k, (float)(1 - r_gm_wm), (float)(1 - k_gm_wm));
  return (int)hipGetLastError();
}

int glia_ms_segmentation(int64_t n, const float* bg, const float* gm, const float* wm, const float* vt,
                         const float* csf, const float* c, float* seg, void* stream) {
  if (n < 0 || !bg || !gm || !wm || !vt || !csf || !c || !seg) return GLIA_RD_ERR_ARG;
  if (n == 0) return 0;
  glia::ms::k_ms_segmentation<<<glia::ms::blocks(n), 256, 0, (hipStream_t)stream>>>(n, bg, gm, wm, vt, csf, c, seg);
  return (int)hipGetLastError();
}

int glia_ms_clip(int64_t n, float* x, void* stream) {
  if (n < 0 || !x) return GLIA_RD_ERR_ARG;
  if (n == 0) return 0;
  glia::ms::k_ms_clip<<<glia::ms::blocks(n), 256, 0, (hipStream_t)stream>>>(n, x);
  return (int)hipGetLastError();
}

int glia_ms_edema_mask(int64_t n, const float* i, double threshold, float* ed, void* stream) {
  if (n < 0 || !i || !ed) return GLIA_RD_ERR_ARG;
  if (n == 0) return 0;
  glia::ms::k_ms_edema<<<glia::ms::blocks(n), 256, 0, (hipStream_t)stream>>>(n, i, (float)threshold, ed);
  return (int)hipGetLastError();
}

}  // extern "C"
