// el_capi.cpp -- extern "C" boundary of the linear-elasticity solver (include/glia_el.h).
//
// The operator runs in spectral space on a one-member RD engine: the three displacement spectra
// U_d are formed once, then
//   div u -> lam div u -> its spectrum L,
//   the six strain terms mu (d_j u_i + d_i u_j) -> their spectra S_ij,
//   f_i = inverse of i (sum_j w_j S_ij + w_i L), minus screen u_i,
// i.e. 20 transforms per application (the reference's sequence of gradient and divergence calls
// makes 48).  The engine's real slots RS_WM (forward input) and RS_GM (inverse output) stage the
// transforms; its complex slots hold U (CS_X, CS_A, CS_H), L (CS_P), S (CS_T0..CS_T5) and one
// combination (CS_W); CS_T7 is the inverse transform's work slot.
#include <cmath>
#include <memory>
#include <utility>
#include <new>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/glia_el.h"
#include "el_launch.h"
#include "rd_solver.h"

using glia::el::Vec3;
using glia::el::Vec3C;

namespace {
enum {
  KSP_CONVERGED_RTOL = 2,
  KSP_CONVERGED_ATOL = 3,
  KSP_DIVERGED_ITS = -3,
  KSP_DIVERGED_DTOL = -4,
  KSP_DIVERGED_INDEFINITE_PC = -8,
  KSP_DIVERGED_NANORINF = -9,
  KSP_DIVERGED_INDEFINITE_MAT = -10
};
constexpr int CU[3] = {glia::CS_X, glia::CS_A, glia::CS_H};
constexpr int CS_L = glia::CS_P, CS_C = glia::CS_W;
constexpr int CS_S[6] = {glia::CS_T0, glia::CS_T1, glia::CS_T2, glia::CS_T3, glia::CS_T4, glia::CS_T5};
// strain pair (i, j), i <= j, -> index into CS_S
int sidx(int i, int j) {
  if (i > j) std::swap(i, j);
  return i == 0 ? j : (i == 1 ? 2 + j : 5);
}
}  // namespace

struct glia_el_s {
  std::unique_ptr<glia::RDEnsemble> e;
  int64_t ng = 0;
  hipStream_t st = nullptr;
  float *mu = nullptr, *lam = nullptr, *scr = nullptr;
  // CG vectors: x staging, b, r, z, p, w.  r, z and p live in real fields of the engine that the
  // solver does not otherwise use, so the transforms read and write them in place
  float* v[6][3] = {};
  double* part = nullptr;
  std::vector<double> hpart;
  double avg[3] = {0, 0, 0};
  bool have_material = false;
  std::string err;
  ~glia_el_s() {
    for (float* p : {mu, lam, scr})
      if (p) (void)hipFree(p);
    for (int k = 0; k < 6; k++)
      for (float* p : v[k])
        if (p && !engine_owned(k)) (void)hipFree(p);
    if (part) (void)hipFree(part);
  }
  static bool engine_owned(int k) { return k == 2 || k == 3 || k == 4; }  // VR, VZ, VP
};

namespace {

int fail(glia_el_s* h, int code, const std::string& msg) {
  h->err = msg;
  return code;
}
int hip_fail(glia_el_s* h, int e, const char* what) {
  h->err = std::string(what) + ": " + hipGetErrorString((hipError_t)e);
  return e;
}
#define EL_HIP(h, x)                                   \
  do {                                                 \
    const int e_ = (int)(x);                           \
    if (e_ != 0) return hip_fail(h, e_, #x);           \
  } while (0)
#define EL_ENG(h, x)                                   \
  do {                                                 \
    const int e_ = (x);                                \
    if (e_ != 0) return fail(h, e_, h->e->error());    \
  } while (0)

float* slot(glia_el_s* h, int rs) { return h->e->field_ptr(0, rs); }

// the engine's real field at p, or -1
int engine_slot(glia_el_s* h, const float* p) {
  for (int s = 0; s < glia::RS_NUM; s++)
    if (slot(h, s) == p) return s;
  return -1;
}

// real field (device) -> packed spectrum in complex slot cs, in place when the field is one of the
// engine's, else through RS_WM
int forward(glia_el_s* h, const float* src, int cs) {
  int s = engine_slot(h, src);
  if (s < 0) {
    EL_HIP(h, hipMemcpyAsync(slot(h, glia::RS_WM), src, sizeof(float) * h->ng, hipMemcpyDeviceToDevice, h->st));
    s = glia::RS_WM;
  }
  EL_ENG(h, h->e->fft_r2c(s, cs));
  return 0;
}

// f = A u (device vectors, f must not alias u)
int apply_op(glia_el_s* h, const float* const u[3], float* const f[3]) {
  const int64_t ng = h->ng;
  float* wm = slot(h, glia::RS_WM);
  float* gmo = slot(h, glia::RS_GM);
  int rc;
  for (int d = 0; d < 3; d++)
    if ((rc = forward(h, u[d], CU[d]))) return rc;
  // lam div u -> L
  {
    const int s[3] = {CU[0], CU[1], CU[2]}, dir[3] = {0, 1, 2};
    EL_ENG(h, h->e->spec_lin(3, s, dir, CS_C));
    EL_ENG(h, h->e->fft_c2r(CS_C, glia::RS_GM));
    EL_HIP(h, glia::el::launch_mul(h->st, ng, h->lam, gmo, wm));
    EL_ENG(h, h->e->fft_r2c(glia::RS_WM, CS_L));
  }
  // mu (d_j u_i + d_i u_j) -> S_ij
  for (int i = 0; i < 3; i++)
    for (int j = i; j < 3; j++) {
      const int s[2] = {CU[i], CU[j]}, dir[2] = {j, i};
      EL_ENG(h, h->e->spec_lin(2, s, dir, CS_C));
      EL_ENG(h, h->e->fft_c2r(CS_C, glia::RS_GM));
      EL_HIP(h, glia::el::launch_mul(h->st, ng, h->mu, gmo, wm));
      EL_ENG(h, h->e->fft_r2c(glia::RS_WM, CS_S[sidx(i, j)]));
    }
  // f_i = div_j(S_ij) + d_i L - screen u_i
  for (int i = 0; i < 3; i++) {
    const int s[4] = {CS_S[sidx(i, 0)], CS_S[sidx(i, 1)], CS_S[sidx(i, 2)], CS_L}, dir[4] = {0, 1, 2, i};
    EL_ENG(h, h->e->spec_lin(4, s, dir, CS_C));
    EL_ENG(h, h->e->fft_c2r(CS_C, glia::RS_GM));
    EL_HIP(h, glia::el::launch_fnms(h->st, ng, gmo, h->scr, u[i], f[i]));
  }
  return 0;
}

// u = B f (device vectors)
int apply_prec(glia_el_s* h, const float* const f[3], float* const u[3]) {
  int rc;
  for (int d = 0; d < 3; d++)
    if ((rc = forward(h, f[d], CU[d]))) return rc;
  const int in[3] = {CU[0], CU[1], CU[2]}, out[3] = {CS_S[0], CS_S[1], CS_S[2]};
  EL_ENG(h, h->e->el_prec(in, out, h->avg[0], h->avg[1], h->avg[2]));
  for (int d = 0; d < 3; d++) {
    const int s = engine_slot(h, u[d]);
    if (s >= 0 && s != glia::RS_WM && s != glia::RS_GM) {
      EL_ENG(h, h->e->fft_c2r(out[d], s));
    } else {
      EL_ENG(h, h->e->fft_c2r(out[d], glia::RS_GM));
      EL_HIP(h, hipMemcpyAsync(u[d], slot(h, glia::RS_GM), sizeof(float) * h->ng, hipMemcpyDeviceToDevice, h->st));
    }
  }
  return 0;
}

int dot(glia_el_s* h, const float* const x[3], const float* const y[3], double* out) {
  Vec3C a{{x[0], x[1], x[2]}}, b{{y[0], y[1], y[2]}};
  EL_HIP(h, glia::el::launch_dot3(h->st, h->ng, a, b, h->part));
  const int np = glia::el::partial_count();
  EL_HIP(h, hipMemcpyAsync(h->hpart.data(), h->part, sizeof(double) * np, hipMemcpyDeviceToHost, h->st));
  EL_HIP(h, hipStreamSynchronize(h->st));
  double s = 0;
  for (int i = 0; i < np; i++) s += h->hpart[i];  // fixed order: deterministic
  *out = s;
  return 0;
}

// <z, z> and <z, r> in one pass and one host read
int dot_zz_zr(glia_el_s* h, const float* const z[3], const float* const r[3], double* zz, double* zr) {
  Vec3C a{{z[0], z[1], z[2]}}, b{{r[0], r[1], r[2]}};
  EL_HIP(h, glia::el::launch_dot3_zz_zr(h->st, h->ng, a, b, h->part));
  const int np = glia::el::partial_count();
  EL_HIP(h, hipMemcpyAsync(h->hpart.data(), h->part, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, h->st));
  EL_HIP(h, hipStreamSynchronize(h->st));
  double s0 = 0, s1 = 0;
  for (int i = 0; i < np; i++) {  // fixed order: deterministic
    s0 += h->hpart[2 * i];
    s1 += h->hpart[2 * i + 1];
  }
  *zz = s0;
  *zr = s1;
  return 0;
}

Vec3 mut(float* const p[3]) { return Vec3{{p[0], p[1], p[2]}}; }
Vec3C cst(float* const p[3]) { return Vec3C{{p[0], p[1], p[2]}}; }

bool valid_kind(int k) { return k == GLIA_RD_HOST || k == GLIA_RD_DEVICE; }

// host or device vector -> device staging v[slot]
int stage_in(glia_el_s* h, const float* const src[3], int kind, int vs) {
  for (int d = 0; d < 3; d++) {
    if (!src[d]) return fail(h, GLIA_RD_ERR_ARG, "null component");
    EL_HIP(h, hipMemcpyAsync(h->v[vs][d], src[d], sizeof(float) * h->ng,
                             kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
  }
  return 0;
}
int stage_out(glia_el_s* h, float* const dst[3], int kind, int vs) {
  for (int d = 0; d < 3; d++) {
    if (!dst[d]) return fail(h, GLIA_RD_ERR_ARG, "null component");
    EL_HIP(h, hipMemcpyAsync(dst[d], h->v[vs][d], sizeof(float) * h->ng,
                             kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->st));
  }
  EL_HIP(h, hipStreamSynchronize(h->st));
  return 0;
}

enum { VX = 0, VB, VR, VZ, VP, VW };

}  // namespace

extern "C" {

void glia_el_params_init(glia_el_params* p) {
  if (!p) return;
  p->E_healthy = 2100;
  p->nu_healthy = 0.4;
  p->E_tumor = 8000;
  p->nu_tumor = 0.45;
  p->E_bg = 15000;
  p->nu_bg = 0.45;
  p->E_csf = 100;
  p->nu_csf = 0.1;
  p->screen_low = 1e-2;
  p->screen_high = 1e2;
}

int glia_el_create(const int n[3], void* stream, glia_el_t* out) {
  if (!n || !out) return GLIA_RD_ERR_ARG;
  *out = nullptr;
  glia::SolverConfig sc;
  for (int i = 0; i < 3; i++) sc.n[i] = n[i];
  sc.nmembers = 1;
  sc.nt = 1;
  sc.stream = (hipStream_t)stream;
  glia_el_s* h = new (std::nothrow) glia_el_s();
  if (!h) return GLIA_RD_ERR_ARG;
  h->e.reset(new (std::nothrow) glia::RDEnsemble(sc));
  int rc = h->e ? h->e->init() : GLIA_RD_ERR_ARG;
  if (rc) {
    delete h;
    return rc;
  }
  h->ng = (int64_t)n[0] * n[1] * n[2];
  h->st = h->e->stream();
  const size_t b = sizeof(float) * h->ng;
  // r, z, p in the engine's real fields that the elasticity solver never stages through (it uses
  // RS_WM and RS_GM)
  const int own[3][3] = {{glia::RS_CSF, glia::RS_VT, glia::RS_FILTER},
                         {glia::RS_OBS, glia::RS_D1, glia::RS_C0},
                         {glia::RS_CT, glia::RS_A0, glia::RS_CH}};
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) h->v[2 + k][d] = slot(h, own[k][d]);
  std::vector<float**> bufs = {&h->mu, &h->lam, &h->scr};
  for (int k = 0; k < 6; k++)
    if (!glia_el_s::engine_owned(k))
      for (float*& p : h->v[k]) bufs.push_back(&p);
  for (float** p : bufs) {
    const hipError_t e = hipMalloc((void**)p, b);
    if (e != hipSuccess) {
      delete h;
      return (int)e;
    }
  }
  const int np = glia::el::partial_count();
  if (hipMalloc((void**)&h->part, sizeof(double) * 3 * np) != hipSuccess) {
    delete h;
    return (int)hipErrorOutOfMemory;
  }
  h->hpart.assign(3 * np, 0.0);
  *out = h;
  return 0;
}

int glia_el_destroy(glia_el_t h) {
  if (!h) return GLIA_RD_ERR_ARG;
  (void)hipStreamSynchronize(h->st);
  delete h;
  return 0;
}

const char* glia_el_last_error(glia_el_t h) { return h ? h->err.c_str() : "null handle"; }

int glia_el_set_material(glia_el_t h, const float* c, const float* wm, const float* gm, const float* csf,
                         const float* vt, const float* bg, const glia_el_params* p, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!c || !wm || !gm || !csf || !vt || !bg || !valid_kind(mem_kind))
    return fail(h, GLIA_RD_ERR_ARG, "null field or bad mem_kind");
  glia_el_params q;
  glia_el_params_init(&q);
  if (p) q = *p;
  auto mu = [](double E, double nu) { return (float)(E / (2 * (1 + nu))); };
  auto lam = [](double E, double nu) { return (float)(nu * E / ((1 + nu) * (1 - 2 * nu))); };
  glia::el::ElMaterial m;
  m.mu_h = mu(q.E_healthy, q.nu_healthy);
  m.mu_t = mu(q.E_tumor, q.nu_tumor);
  m.mu_bg = mu(q.E_bg, q.nu_bg);
  m.mu_c = mu(q.E_csf, q.nu_csf);
  m.lam_h = lam(q.E_healthy, q.nu_healthy);
  m.lam_t = lam(q.E_tumor, q.nu_tumor);
  m.lam_bg = lam(q.E_bg, q.nu_bg);
  m.lam_c = lam(q.E_csf, q.nu_csf);
  m.screen_low = (float)q.screen_low;
  m.screen_high = (float)q.screen_high;
  // the six inputs staged in the CG vectors (free outside a solve)
  const float* src[6] = {c, wm, gm, csf, vt, bg};
  float* dev[6] = {h->v[VX][0], h->v[VX][1], h->v[VX][2], h->v[VB][0], h->v[VB][1], h->v[VB][2]};
  for (int k = 0; k < 6; k++)
    EL_HIP(h, hipMemcpyAsync(dev[k], src[k], sizeof(float) * h->ng,
                             mem_kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
  EL_HIP(h, glia::el::launch_material(h->st, h->ng, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], m, h->mu, h->lam,
                                      h->scr, h->part));
  const int np = glia::el::partial_count();
  EL_HIP(h, hipMemcpyAsync(h->hpart.data(), h->part, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, h->st));
  EL_HIP(h, hipStreamSynchronize(h->st));
  double s[3] = {0, 0, 0};
  for (int i = 0; i < np; i++)
    for (int k = 0; k < 3; k++) s[k] += h->hpart[3 * i + k];
  for (int k = 0; k < 3; k++) h->avg[k] = s[k] / (double)h->ng;
  h->have_material = true;
  return 0;
}

int glia_el_get_material(glia_el_t h, float* mu, float* lam, float* screen, double* avg, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "bad mem_kind");
  if (!h->have_material) return fail(h, GLIA_RD_ERR_STATE, "material not set");
  const float* src[3] = {h->mu, h->lam, h->scr};
  float* dst[3] = {mu, lam, screen};
  for (int k = 0; k < 3; k++)
    if (dst[k])
      EL_HIP(h, hipMemcpyAsync(dst[k], src[k], sizeof(float) * h->ng,
                               mem_kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->st));
  EL_HIP(h, hipStreamSynchronize(h->st));
  if (avg)
    for (int k = 0; k < 3; k++) avg[k] = h->avg[k];
  return 0;
}

int glia_el_apply(glia_el_t h, const float* const u[3], float* const f[3], int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!u || !f || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null vectors or bad mem_kind");
  if (!h->have_material) return fail(h, GLIA_RD_ERR_STATE, "material not set");
  int rc;
  if ((rc = stage_in(h, u, mem_kind, VP))) return rc;
  float* const* out = h->v[VW];
  if ((rc = apply_op(h, h->v[VP], out))) return rc;
  return stage_out(h, f, mem_kind, VW);
}

int glia_el_precondition(glia_el_t h, const float* const f[3], float* const u[3], int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!f || !u || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null vectors or bad mem_kind");
  if (!h->have_material) return fail(h, GLIA_RD_ERR_STATE, "material not set");
  int rc;
  if ((rc = stage_in(h, f, mem_kind, VR))) return rc;
  if ((rc = apply_prec(h, h->v[VR], h->v[VZ]))) return rc;
  return stage_out(h, u, mem_kind, VZ);
}

int glia_el_force(glia_el_t h, const float* c, double forcing_factor, int use_tanh, float* const f[3], int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!c || !f || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null field or bad mem_kind");
  float* wm = slot(h, glia::RS_WM);
  EL_HIP(h, hipMemcpyAsync(wm, c, sizeof(float) * h->ng,
                           mem_kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
  const int n0 = h->e->dims().n0;
  EL_ENG(h, h->e->smooth_field(0, glia::RS_WM, 2.0 * M_PI / n0));  // sigma_smooth = 2 pi / n0
  float* cs = h->v[VB][0];                                           // the smoothed c
  EL_HIP(h, hipMemcpyAsync(cs, wm, sizeof(float) * h->ng, hipMemcpyDeviceToDevice, h->st));
  EL_ENG(h, h->e->fft_r2c(glia::RS_WM, CU[0]));
  for (int d = 0; d < 3; d++) {
    const int s[1] = {CU[0]}, dir[1] = {d};
    EL_ENG(h, h->e->spec_lin(1, s, dir, CS_C));
    EL_ENG(h, h->e->fft_c2r(CS_C, glia::RS_GM));
    EL_HIP(h, glia::el::launch_force_scale(h->st, h->ng, slot(h, glia::RS_GM), cs, (float)forcing_factor,
                                           use_tanh ? 1 : 0, h->v[VW][d]));
  }
  return stage_out(h, f, mem_kind, VW);
}

int glia_el_solve(glia_el_t h, const float* const f[3], float* const u[3], int mem_kind, int* iterations, int* reason,
                  double* rnorm) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!f || !u || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null vectors or bad mem_kind");
  if (!h->have_material) return fail(h, GLIA_RD_ERR_STATE, "material not set");
  const double rtol = 1e-3, atol = 1e-50, dtol = 1e5;
  const int maxit = 100;
  int rc;
  if ((rc = stage_in(h, f, mem_kind, VB)) || (rc = stage_in(h, u, mem_kind, VX))) return rc;
  float* const* x = h->v[VX];
  float* const* b = h->v[VB];
  float* const* r = h->v[VR];
  float* const* z = h->v[VZ];
  float* const* p = h->v[VP];
  float* const* w = h->v[VW];
  // KSPSolve_CG (PETSc, KSP_NORM_PRECONDITIONED, nonzero initial guess) with KSPConvergedDefault
  if ((rc = apply_op(h, x, r))) return rc;
  EL_HIP(h, glia::el::launch_axpby3(h->st, h->ng, mut(r), cst(b), 0.0, 1));  // r = b - A x
  double rnorm0, dp, beta = 0, betaold = 0, dpi = 0, dpiold, a;
  if ((rc = apply_prec(h, b, z)) || (rc = dot(h, z, z, &rnorm0))) return rc;  // ||B b||
  rnorm0 = std::sqrt(rnorm0);
  if ((rc = apply_prec(h, r, z)) || (rc = dot(h, z, z, &dp))) return rc;
  dp = std::sqrt(dp);
  if (rnorm0 == 0) rnorm0 = dp;
  const double ttol = std::fmax(rtol * rnorm0, atol);
  auto converged = [&](double rn) {
    if (std::isnan(rn) || std::isinf(rn)) return (int)KSP_DIVERGED_NANORINF;
    if (rn <= ttol) return (int)(rn < atol ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL);
    if (rn >= dtol * rnorm0) return (int)KSP_DIVERGED_DTOL;
    return 0;
  };
  int its = 0, why = converged(dp), i = 0;
  if (!why) {
    if ((rc = dot(h, z, r, &beta))) return rc;
    do {
      its = i + 1;
      if (beta == 0.0) {
        why = KSP_CONVERGED_ATOL;
        break;
      } else if (i > 0 && beta * betaold < 0.0) {
        why = KSP_DIVERGED_INDEFINITE_PC;
        break;
      }
      EL_HIP(h, glia::el::launch_axpby3(h->st, h->ng, mut(p), cst(z), i ? beta / betaold : 0.0, 0));
      dpiold = dpi;
      if ((rc = apply_op(h, p, w)) || (rc = dot(h, p, w, &dpi))) return rc;
      betaold = beta;
      if (std::isnan(dpi) || std::isinf(dpi)) {
        why = KSP_DIVERGED_NANORINF;
        break;
      }
      if (dpi == 0.0 || (i > 0 && dpi * dpiold <= 0.0)) {
        why = KSP_DIVERGED_INDEFINITE_MAT;
        break;
      }
      a = beta / dpi;
      EL_HIP(h, glia::el::launch_cg_update(h->st, h->ng, mut(x), mut(r), cst(p), cst(w), a));
      // <z, z> for the test and <z, r> for the next beta in one pass (beta is unused when the
      // test stops the loop)
      double zr = 0;
      if ((rc = apply_prec(h, r, z)) || (rc = dot_zz_zr(h, z, r, &dp, &zr))) return rc;
      dp = std::sqrt(dp);
      if ((why = converged(dp))) break;
      beta = zr;
      if (std::isnan(beta) || std::isinf(beta)) {
        why = KSP_DIVERGED_NANORINF;
        break;
      }
      i++;
    } while (i < maxit);
    if (i >= maxit) why = KSP_DIVERGED_ITS;
  }
  if ((rc = stage_out(h, u, mem_kind, VX))) return rc;
  if (iterations) *iterations = its;
  if (reason) *reason = why;
  if (rnorm) *rnorm = dp;
  return 0;
}

}  // extern "C"
