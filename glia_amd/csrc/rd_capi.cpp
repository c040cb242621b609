// rd_capi.cpp -- extern "C" boundary (include/glia_rd.h) over glia::RDEnsemble.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <chrono>
#include <string>
#include <thread>
#include <cmath>
#include <cstdint>
#include <map>
#include <random>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/glia_rd.h"
#include "rd_handle.h"

using glia::RDEnsemble;

static int field_slot(int which) {
  switch (which) {
    case GLIA_RD_FIELD_K: return glia::RS_K;
    case GLIA_RD_FIELD_RHO: return glia::RS_RHO;
    case GLIA_RD_FIELD_WM: return glia::RS_WM;
    case GLIA_RD_FIELD_GM: return glia::RS_GM;
    case GLIA_RD_FIELD_CSF: return glia::RS_CSF;
    case GLIA_RD_FIELD_VT: return glia::RS_VT;
    case GLIA_RD_FIELD_FILTER: return glia::RS_FILTER;
    case GLIA_RD_FIELD_OBS: return glia::RS_OBS;
    case GLIA_RD_FIELD_D1: return glia::RS_D1;
    case GLIA_RD_FIELD_C0: return glia::RS_C0;
    case GLIA_RD_FIELD_CT: return glia::RS_CT;
    case GLIA_RD_FIELD_ALPHA0: return glia::RS_A0;
    case GLIA_RD_FIELD_WORK: return glia::RS_CH;
    default: return -1;
  }
}

static int ret(glia_rd_t h, int rc) {
  if (rc != 0 && h && h->e) h->err = h->e->error();
  return rc;
}

// a rejected argument: the message is what glia_rd_last_error reports
static int arg_err(glia_rd_t h, const char* what) {
  h->err = what;
  return GLIA_RD_ERR_ARG;
}

#define GLIA_H(h)                                      \
  do {                                                 \
    if (!(h) || !(h)->e) return GLIA_RD_ERR_ARG;       \
  } while (0)

extern "C" {

void glia_rd_config_init(glia_rd_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->n[0] = c->n[1] = c->n[2] = 64;
  c->nmembers = 1;
  c->dt = 0.5;
  c->nt = 1;
  c->order = 2;
  c->pcg_rtol = 1e-6;
  c->pcg_maxit = 5000;
  c->checkpoint_slots = 0;
  c->adjoint_store = 0;
  c->phi_store = 0;
  c->k_gm_wm = 0.0;
  c->k_glm_wm = 0.0;
  c->r_gm_wm = 0.0;
  c->r_glm_wm = 0.0;
  c->beta = 0.0;
  c->k_scale = 1.0;
  c->rho_scale = 1.0;
  c->smoothing_factor = 1.0;
  c->stream = nullptr;
}

int glia_rd_create(const glia_rd_config* cfg, glia_rd_t* out) {
  if (!cfg || !out) return GLIA_RD_ERR_ARG;
  *out = nullptr;
  if (cfg->order != 2) return GLIA_RD_ERR_UNSUPPORTED;
  glia::SolverConfig sc;
  for (int i = 0; i < 3; i++) sc.n[i] = cfg->n[i];
  sc.nmembers = cfg->nmembers;
  sc.dt = cfg->dt;
  sc.nt = cfg->nt;
  sc.rtol = cfg->pcg_rtol;
  sc.maxit = cfg->pcg_maxit;
  sc.checkpoint_slots = cfg->checkpoint_slots;
  sc.adjoint_store = cfg->adjoint_store;
  sc.phi_store = cfg->phi_store;
  sc.k_gm_wm = cfg->k_gm_wm;
  sc.k_glm_wm = cfg->k_glm_wm;
  sc.r_gm_wm = cfg->r_gm_wm;
  sc.r_glm_wm = cfg->r_glm_wm;
  sc.beta = cfg->beta;
  sc.k_scale = cfg->k_scale;
  sc.rho_scale = cfg->rho_scale;
  sc.smoothing_factor = cfg->smoothing_factor;
  sc.stream = (hipStream_t)cfg->stream;
  glia_rd_s* h = new (std::nothrow) glia_rd_s();
  if (!h) return GLIA_RD_ERR_ARG;
  h->e.reset(new (std::nothrow) RDEnsemble(sc));
  if (!h->e) {
    delete h;
    return GLIA_RD_ERR_ARG;
  }
  int rc = h->e->init();
  // GLIA_AUTOTUNE=1 tunes every new handle (GLIA_TUNE_CACHE: the choice cache file)
  const char* tune = std::getenv("GLIA_AUTOTUNE");
  if (rc == 0 && tune && std::string(tune) == "1") rc = h->e->autotune(std::getenv("GLIA_TUNE_CACHE"), nullptr);
  if (rc != 0) {
    // keep the handle so the caller can read the message, then destroy it
    h->err = h->e->error();
    *out = h;
    return rc;
  }
  *out = h;
  return 0;
}

int glia_rd_destroy(glia_rd_t h) {
  if (!h) return GLIA_RD_ERR_ARG;
  delete h;
  return 0;
}

const char* glia_rd_last_error(glia_rd_t h) {
  if (!h) return "null handle";
  return h->err.c_str();
}

int glia_rd_sync(glia_rd_t h) {
  GLIA_H(h);
  return ret(h, h->e->sync());
}

int glia_rd_num_members(glia_rd_t h) {
  GLIA_H(h);
  return h->e->nmembers();
}

int glia_rd_get_grid(glia_rd_t h, int* n) {
  GLIA_H(h);
  if (!n) return arg_err(h, "null output pointer");
  for (int d = 0; d < 3; d++) n[d] = h->e->config().n[d];
  return 0;
}

int glia_rd_set_tissue(glia_rd_t h, int member, const float* wm, const float* gm, const float* csf, const float* vt,
                       int mem_kind) {
  GLIA_H(h);
  return ret(h, h->e->set_tissue(member, wm, gm, csf, vt, mem_kind == GLIA_RD_DEVICE));
}

int glia_rd_set_tissue_sinusoidal(glia_rd_t h, int member) {
  GLIA_H(h);
  return ret(h, h->e->set_tissue_sinusoidal(member));
}

int glia_rd_set_coefficients(glia_rd_t h, int member, double kappa, double rho) {
  GLIA_H(h);
  if (member < 0 || member >= h->e->nmembers()) return arg_err(h, "member out of range");
  return ret(h, h->e->set_coefficients(member, kappa, rho));
}

int glia_rd_set_kappa(glia_rd_t h, int member, double kappa) {
  GLIA_H(h);
  if (member < 0 || member >= h->e->nmembers()) return arg_err(h, "member out of range");
  return ret(h, h->e->set_kappa(member, kappa));
}

int glia_rd_set_rho(glia_rd_t h, int member, double rho) {
  GLIA_H(h);
  if (member < 0 || member >= h->e->nmembers()) return arg_err(h, "member out of range");
  return ret(h, h->e->set_rho(member, rho));
}

int glia_rd_set_k_sinusoidal(glia_rd_t h, int member, double scale) {
  GLIA_H(h);
  if (member < 0 || member >= h->e->nmembers()) return arg_err(h, "member out of range");
  return ret(h, h->e->set_k_sinusoidal(member, scale));
}

int glia_rd_prec_factor(glia_rd_t h) {
  GLIA_H(h);
  return ret(h, h->e->prec_factor());
}

int glia_rd_set_ctx_dt(glia_rd_t h, double dt) {
  GLIA_H(h);
  h->e->set_ctx_dt(dt);
  return 0;
}

int glia_rd_get_ctx_dt(glia_rd_t h, double* dt) {
  GLIA_H(h);
  if (!dt) return arg_err(h, "null output pointer");
  *dt = h->e->ctx_dt();
  return 0;
}

int glia_rd_set_beta(glia_rd_t h, double beta) {
  GLIA_H(h);
  if (!(beta >= 0)) return arg_err(h, "beta must be >= 0");
  h->e->set_beta(beta);
  return 0;
}

int glia_rd_set_ratios(glia_rd_t h, double k_gm_wm, double k_glm_wm, double r_gm_wm, double r_glm_wm) {
  GLIA_H(h);
  h->e->set_ratios(k_gm_wm, k_glm_wm, r_gm_wm, r_glm_wm);
  return 0;
}

int glia_rd_get_beta(glia_rd_t h, double* beta) {
  GLIA_H(h);
  if (!beta) return arg_err(h, "null output pointer");
  *beta = h->e->config().beta;
  return 0;
}

int glia_rd_set_basis(glia_rd_t h, int member, int np, const double* centers, double sigma) {
  GLIA_H(h);
  return ret(h, h->e->set_basis(member, np, centers, sigma));
}

int glia_rd_basis_size(glia_rd_t h, int member) {
  GLIA_H(h);
  return h->e->basis_size(member);
}

int glia_rd_set_data(glia_rd_t h, int member, const float* d1, const float* obs_mask, double obs_thr, int mem_kind) {
  GLIA_H(h);
  if (!d1) return arg_err(h, "null data pointer");
  return ret(h, h->e->set_data(member, d1, obs_mask, obs_thr, mem_kind == GLIA_RD_DEVICE));
}

int glia_rd_set_field(glia_rd_t h, int member, int which, const float* src, int mem_kind) {
  GLIA_H(h);
  const int s = field_slot(which);
  if (s < 0) return arg_err(h, "unknown field id");
  return ret(h, h->e->set_field(member, s, src, mem_kind == GLIA_RD_DEVICE));
}

int glia_rd_get_field(glia_rd_t h, int member, int which, float* dst, int mem_kind) {
  GLIA_H(h);
  const int s = field_slot(which);
  if (s < 0 || !dst) return arg_err(h, "unknown field id or null destination");
  return ret(h, h->e->get_field(member, s, dst, mem_kind == GLIA_RD_DEVICE));
}

static bool split_x(RDEnsemble* e, const double* x, int extra, std::vector<std::vector<double>>& out) {
  if (!x) return false;
  const int nm = e->nmembers();
  out.assign(nm, {});
  int64_t off = 0;
  for (int m = 0; m < nm; m++) {
    const int n = e->basis_size(m) + extra;
    out[m].assign(x + off, x + off + n);
    off += n;
  }
  return true;
}

int glia_rd_phi_apply(glia_rd_t h, const double* p) {
  GLIA_H(h);
  std::vector<std::vector<double>> pv;
  if (!split_x(h->e.get(), p, 0, pv)) return arg_err(h, "p does not match the basis sizes");
  int rc = h->e->phi_apply(pv);
  if (rc == 0) rc = h->e->sync();
  return ret(h, rc);
}

int glia_rd_phi_apply_transpose(glia_rd_t h, int which, double* out) {
  GLIA_H(h);
  const int s = field_slot(which);
  if (s < 0 || !out) return arg_err(h, "unknown field id or null output");
  std::vector<std::vector<double>> o;
  const int rc = h->e->phi_apply_transpose(s, o);
  if (rc) return ret(h, rc);
  int64_t off = 0;
  for (auto& v : o) {
    std::memcpy(out + off, v.data(), sizeof(double) * v.size());
    off += (int64_t)v.size();
  }
  return 0;
}

int glia_rd_diffusion_solve(glia_rd_t h, int which, double tau) {
  GLIA_H(h);
  const int s = field_slot(which);
  if (s < 0) return arg_err(h, "unknown field id");
  return ret(h, h->e->diffusion_solve_field(s, tau));
}

int glia_rd_apply_d(glia_rd_t h, int which_in, int which_out) {
  GLIA_H(h);
  const int a = field_slot(which_in), b = field_slot(which_out);
  if (a < 0 || b < 0) return arg_err(h, "unknown field id");
  return ret(h, h->e->apply_d_field(a, b));
}

// packed half spectrum <-> reference layout (kz = 0 and kz = n2/2 planes unpacked by symmetry)
int glia_rd_fft_r2c(glia_rd_t h, int member, const float* f, float* fhat) {
  GLIA_H(h);
  if (!f || !fhat || member < 0 || member >= h->e->nmembers()) return arg_err(h, "null buffer or member out of range");
  const glia::Dims& d = h->e->dims();
  int rc = h->e->set_field(member, glia::RS_CH, f, false);
  if (!rc) rc = h->e->fft_r2c(glia::RS_CH, glia::CS_T7);
  std::vector<float2> P(d.nspec);
  if (!rc) rc = h->e->get_spectrum(member, glia::CS_T7, P.data());
  if (rc) return ret(h, rc);
  const int n0 = d.n0, n1 = d.n1, M = d.m2, n2h = M + 1;
  for (int kx = 0; kx < n0; kx++)
    for (int ky = 0; ky < n1; ky++) {
      const int64_t b = ((int64_t)kx * n1 + ky);
      float* o = fhat + 2 * b * n2h;
      const float2* p = P.data() + b * M;
      for (int kz = 1; kz < M; kz++) {
        o[2 * kz] = p[kz].x;
        o[2 * kz + 1] = p[kz].y;
      }
      const int64_t bm = ((int64_t)((n0 - kx) % n0) * n1 + (n1 - ky) % n1);
      const float2 a = p[0], c = P[bm * M];  // P[k], P[-k]
      // A = (P[k] + conj P[-k]) / 2, B = (P[k] - conj P[-k]) / (2i)
      o[0] = 0.5f * (a.x + c.x);
      o[1] = 0.5f * (a.y - c.y);
      const float dx = a.x - c.x, dy = a.y + c.y;
      o[2 * M] = 0.5f * dy;
      o[2 * M + 1] = -0.5f * dx;
    }
  return 0;
}

int glia_rd_fft_c2r(glia_rd_t h, int member, const float* fhat, float* f) {
  GLIA_H(h);
  if (!f || !fhat || member < 0 || member >= h->e->nmembers()) return arg_err(h, "null buffer or member out of range");
  const glia::Dims& d = h->e->dims();
  const int n0 = d.n0, n1 = d.n1, M = d.m2, n2h = M + 1;
  std::vector<float2> P(d.nspec);
  for (int64_t b = 0; b < (int64_t)n0 * n1; b++) {
    const float* in = fhat + 2 * b * n2h;
    float2* p = P.data() + b * M;
    p[0] = make_float2(in[0] - in[2 * M + 1], in[1] + in[2 * M]);  // X0 + i X_{n2/2}
    for (int kz = 1; kz < M; kz++) p[kz] = make_float2(in[2 * kz], in[2 * kz + 1]);
  }
  int rc = h->e->set_spectrum(member, glia::CS_T6, P.data());
  if (!rc) rc = h->e->fft_c2r(glia::CS_T6, glia::RS_CH);
  if (!rc) rc = h->e->get_field(member, glia::RS_CH, f, false);
  if (rc) return ret(h, rc);
  // the solver's inverse includes 1/N^3; the reference's C2R is unnormalized
  const double s = (double)d.ng;
  for (int64_t i = 0; i < d.ng; i++) f[i] = (float)(f[i] * s);
  return 0;
}

int glia_rd_solve_state(glia_rd_t h, int linearized) {
  GLIA_H(h);
  int rc = h->e->solve_state(linearized, true);
  if (!rc) rc = h->e->sync();
  return ret(h, rc);
}

int glia_rd_solve_adjoint(glia_rd_t h, int linearized) {
  GLIA_H(h);
  int rc = h->e->solve_adjoint(linearized);
  if (!rc) rc = h->e->sync();
  return ret(h, rc);
}

int glia_rd_autotune(glia_rd_t h, const char* cache_path, int* choice) {
  GLIA_H(h);
  return ret(h, h->e->autotune(cache_path, choice));
}

int glia_rd_step_info(glia_rd_t h, long long* launched, long long* reruns, int* cached) {
  GLIA_H(h);
  if (!launched || !reruns || !cached) return arg_err(h, "null output");
  h->e->step_info(launched, reruns, cached);
  return 0;
}

int glia_rd_set_time(glia_rd_t h, double dt, int nt) {
  GLIA_H(h);
  return ret(h, h->e->set_time(dt, nt));
}

int glia_rd_keep_adjoint_history(glia_rd_t h, int enable) {
  GLIA_H(h);
  return ret(h, h->e->keep_adjoint_history(enable != 0));
}

int glia_rd_get_history(glia_rd_t h, int member, int kind, int index, float* dst, int mem_kind) {
  GLIA_H(h);
  if (!dst) return arg_err(h, "null destination");
  return ret(h, h->e->get_history(member, kind, index, dst, mem_kind == GLIA_RD_DEVICE));
}

int glia_rd_set_secondary(glia_rd_t h, int member, double kt) {
  GLIA_H(h);
  int rc = h->e->set_secondary(member, kt);
  if (!rc) rc = h->e->sync();
  return ret(h, rc);
}

int glia_rd_grad_diffusion(glia_rd_t h, double* gk) {
  GLIA_H(h);
  if (!gk) return arg_err(h, "null output");
  for (int m = 0; m < h->e->nmembers(); m++) {
    const int rc = h->e->grad_kappa_rho(m, gk + m, nullptr);
    if (rc) return ret(h, rc);
  }
  return 0;
}

int glia_rd_grad_reaction(glia_rd_t h, double* gr) {
  GLIA_H(h);
  if (!gr) return arg_err(h, "null output");
  for (int m = 0; m < h->e->nmembers(); m++) {
    const int rc = h->e->grad_kappa_rho(m, nullptr, gr + m);
    if (rc) return ret(h, rc);
  }
  return 0;
}

static int eval(glia_rd_t h, int mode, const double* x, double* J, double* dJ, bool grad) {
  GLIA_H(h);
  std::vector<std::vector<double>> xv, g;
  std::vector<double> Jv;
  if (!split_x(h->e.get(), x, 2, xv)) return arg_err(h, "x does not match the basis sizes");
  const int rc = h->e->objective_and_gradient(mode, xv, grad, Jv, g);
  if (rc) return ret(h, rc);
  int64_t off = 0;
  for (int m = 0; m < h->e->nmembers(); m++) {
    if (J) J[m] = Jv[m];
    if (dJ) std::memcpy(dJ + off, g[m].data(), sizeof(double) * g[m].size());
    off += (int64_t)g[m].size();
  }
  return 0;
}

int glia_rd_objective(glia_rd_t h, int mode, const double* x, double* J) {
  if (!J) return arg_err(h, "null output pointer");
  return eval(h, mode, x, J, nullptr, false);
}

int glia_rd_gradient(glia_rd_t h, int mode, const double* x, double* dJ) {
  if (!dJ) return arg_err(h, "null output pointer");
  return eval(h, mode, x, nullptr, dJ, true);
}

int glia_rd_objective_and_gradient(glia_rd_t h, int mode, const double* x, double* J, double* dJ) {
  if (!J || !dJ) return arg_err(h, "null output pointer");
  return eval(h, mode, x, J, dJ, true);
}

int glia_rd_check_gradient(glia_rd_t h, int mode, const double* x, const double* dir, double* err, double* gdiff) {
  GLIA_H(h);
  if (!x) return arg_err(h, "null x");
  const int nm = h->e->nmembers();
  std::vector<std::vector<double>> xv, g;
  if (!split_x(h->e.get(), x, 2, xv)) return arg_err(h, "x does not match the basis sizes");
  std::vector<std::vector<double>> dv(nm);
  if (dir) {
    if (!split_x(h->e.get(), dir, 2, dv)) return arg_err(h, "dir does not match the basis sizes");
  } else {
    std::mt19937_64 gen(0x9e3779b97f4a7c15ull);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int m = 0; m < nm; m++) {
      dv[m].resize(xv[m].size());
      for (double& e : dv[m]) e = u(gen);
    }
  }
  std::vector<double> J0, Ji;
  int rc = h->e->objective_and_gradient(mode, xv, true, J0, g);
  if (rc) return ret(h, rc);
  std::vector<double> gd(nm, 0.0);
  for (int m = 0; m < nm; m++)
    for (size_t k = 0; k < g[m].size(); k++) gd[m] += g[m][k] * dv[m][k];
  std::vector<std::vector<double>> xn(nm), unused;
  for (int i = 0; i < 6; i++) {
    const double hi = std::pow(10.0, -i);
    for (int m = 0; m < nm; m++) {
      xn[m] = xv[m];
      for (size_t k = 0; k < xn[m].size(); k++) xn[m][k] += hi * dv[m][k];
    }
    if ((rc = h->e->objective_and_gradient(mode, xn, false, Ji, unused))) return ret(h, rc);
    for (int m = 0; m < nm; m++) {
      if (err) err[6 * m + i] = std::fabs(Ji[m] - (J0[m] + gd[m] * hi));
      if (gdiff) gdiff[6 * m + i] = (Ji[m] - J0[m]) / hi - gd[m];
    }
  }
  return 0;
}

int glia_rd_hessian(glia_rd_t h, int diffusivity, const double* x, double* y) {
  GLIA_H(h);
  if (!x || !y) return arg_err(h, "null x or y");
  std::vector<std::vector<double>> xv, yv;
  if (!split_x(h->e.get(), x, 2, xv)) return arg_err(h, "x does not match the basis sizes");
  const int rc = h->e->hessian(xv, diffusivity != 0, yv);
  if (rc) return ret(h, rc);
  int64_t off = 0;
  for (int m = 0; m < h->e->nmembers(); m++) {
    std::memcpy(y + off, yv[m].data(), sizeof(double) * yv[m].size());
    off += (int64_t)yv[m].size();
  }
  return 0;
}

int glia_rd_objective_and_gradient_group(int nh, const glia_rd_t* handles, const int* modes, const double* const* x,
                                         double* const* J, double* const* dJ, double* ms) {
  if (nh <= 0 || !handles || !modes || !x || !J || !dJ) return GLIA_RD_ERR_ARG;
  for (int i = 0; i < nh; i++) {
    if (!handles[i] || !handles[i]->e || !x[i] || !J[i] || !dJ[i]) return GLIA_RD_ERR_ARG;
    for (int k = 0; k < i; k++)
      if (handles[k] == handles[i]) return GLIA_RD_ERR_ARG;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GLIA_RD_ERR_STATE;
  std::vector<int> rc(nh, 0);
  std::vector<double> t(nh, 0.0);
  std::vector<std::thread> th;
  th.reserve(nh);
  for (int i = 0; i < nh; i++) {
    th.emplace_back([&, i]() {
      const auto t0 = std::chrono::steady_clock::now();
      const hipError_t e = hipSetDevice(dev);
      rc[i] = (e != hipSuccess) ? (int)e : eval(handles[i], modes[i], x[i], J[i], dJ[i], true);
      t[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
  }
  for (auto& w : th) w.join();
  if (ms)
    for (int i = 0; i < nh; i++) ms[i] = t[i];
  for (int i = 0; i < nh; i++)
    if (rc[i]) return rc[i];
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Lockstep evaluation of several handles: time step i of every handle runs as one graph
// ---------------------------------------------------------------------------------------------
extern "C++" {
namespace {

struct SpanGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
};

// graphs spanning handles, keyed by (handle id, step-cache epoch, cached step address) per handle.
// Per thread, so that evicting never meets a graph that another thread has in flight (every step
// of a lockstep call has completed before the next one is prepared).  Never destroyed at exit:
// the runtime may be gone by then.
std::map<std::vector<uint64_t>, SpanGraph>& span_cache() {
  thread_local auto* cache = new std::map<std::vector<uint64_t>, SpanGraph>();
  return *cache;
}
constexpr size_t kSpanCacheMax = 512;

int span_launch(const std::vector<RDEnsemble*>& e, const std::vector<const RDEnsemble::StepGraph*>& sg,
                const std::vector<hipEvent_t>& ev) {
  std::vector<uint64_t> key;
  key.reserve(3 * e.size());
  for (size_t a = 0; a < e.size(); a++) {
    key.push_back(e[a]->uid());
    key.push_back(e[a]->graph_epoch());
    key.push_back((uint64_t)(uintptr_t)sg[a]);
  }
  auto& cache = span_cache();
  auto it = cache.find(key);
  if (it == cache.end()) {
    if (cache.size() >= kSpanCacheMax) {
      for (auto& kv : cache) {
        (void)hipGraphExecDestroy(kv.second.ex);
        (void)hipGraphDestroy(kv.second.g);
      }
      cache.clear();
    }
    SpanGraph s;
    hipError_t err = hipGraphCreate(&s.g, 0);
    if (err != hipSuccess) return (int)err;
    for (size_t a = 0; a < e.size() && err == hipSuccess; a++) {
      hipGraphNode_t node;
      err = hipGraphAddChildGraphNode(&node, s.g, nullptr, 0, sg[a]->g);  // independent branches
    }
    if (err == hipSuccess) err = hipGraphInstantiate(&s.ex, s.g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
      (void)hipGraphDestroy(s.g);
      return (int)err;
    }
    it = cache.emplace(std::move(key), s).first;
  }
  // the graph runs on the first handle's stream after the work queued on every handle's stream,
  // and every handle's stream continues after it
  hipStream_t s0 = e[0]->stream();
  for (size_t a = 1; a < e.size(); a++) {
    if (e[a]->stream() == s0) continue;
    hipError_t err = hipEventRecord(ev[a], e[a]->stream());
    if (err == hipSuccess) err = hipStreamWaitEvent(s0, ev[a], 0);
    if (err != hipSuccess) return (int)err;
  }
  hipError_t err = hipGraphLaunch(it->second.ex, s0);
  if (err == hipSuccess) err = hipEventRecord(ev[0], s0);
  for (size_t a = 1; a < e.size() && err == hipSuccess; a++)
    if (e[a]->stream() != s0) err = hipStreamWaitEvent(e[a]->stream(), ev[0], 0);
  return (int)err;
}

// one time step of the active handles (step[a] = its index on handle a): every handle's captured
// step in one graph when all have one, else each captured step on its own stream and the eager
// steps in the finish calls
template <class Prepare, class Finish>
int lockstep_step(const std::vector<glia_rd_t>& hs, const std::vector<int>& step, const std::vector<hipEvent_t>& ev,
                  Prepare prepare, Finish finish, long long* spanned) {
  const size_t na = hs.size();
  std::vector<const RDEnsemble::StepGraph*> sg(na, nullptr);
  std::vector<RDEnsemble*> e(na);
  bool all = true;
  for (size_t a = 0; a < na; a++) {
    e[a] = hs[a]->e.get();
    const int rc = prepare(e[a], step[a], &sg[a]);
    if (rc) return ret(hs[a], rc);
    all = all && sg[a] != nullptr;
  }
  if (all && na > 1) {
    const int rc = span_launch(e, sg, ev);
    if (rc) return rc;
    (*spanned)++;
  } else {
    for (size_t a = 0; a < na; a++)
      if (sg[a]) {
        const hipError_t err = hipGraphLaunch(sg[a]->ex, e[a]->stream());
        if (err != hipSuccess) return (int)err;
      }
  }
  for (size_t a = 0; a < na; a++) {
    const int rc = finish(e[a], step[a], sg[a] != nullptr);
    if (rc) return ret(hs[a], rc);
  }
  return 0;
}

}  // namespace
}  // extern "C++"

int glia_rd_objective_and_gradient_lockstep(int nh, const glia_rd_t* handles, const int* modes,
                                            const double* const* x, double* const* J, double* const* dJ,
                                            long long* spanned_steps) {
  if (nh <= 0 || !handles || !modes || !x || !J || !dJ) return GLIA_RD_ERR_ARG;
  bool phased = true;
  for (int i = 0; i < nh; i++) {
    if (!handles[i] || !handles[i]->e || !x[i] || !J[i] || !dJ[i]) return GLIA_RD_ERR_ARG;
    for (int k = 0; k < i; k++)
      if (handles[k] == handles[i]) return GLIA_RD_ERR_ARG;
    phased = phased && handles[i]->e->store_all();
  }
  if (spanned_steps) *spanned_steps = 0;
  if (!phased) {  // checkpointed histories recompute their steps: one handle after another
    for (int i = 0; i < nh; i++) {
      const int rc = eval(handles[i], modes[i], x[i], J[i], dJ[i], true);
      if (rc) return rc;
    }
    return 0;
  }
  std::vector<std::vector<std::vector<double>>> xv(nh);
  int ntmax = 0;
  for (int i = 0; i < nh; i++) {
    if (!split_x(handles[i]->e.get(), x[i], 2, xv[i])) return arg_err(handles[i], "x does not match the basis sizes");
    ntmax = std::max(ntmax, handles[i]->e->nt());
  }
  std::vector<hipEvent_t> ev(nh, nullptr);
  for (int i = 0; i < nh; i++)
    if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
      for (auto v : ev)
        if (v) (void)hipEventDestroy(v);
      return GLIA_RD_ERR_STATE;
    }
  long long spanned = 0;
  auto run = [&]() -> int {
    int rc;
    for (int i = 0; i < nh; i++)
      if ((rc = handles[i]->e->eval_begin(modes[i], xv[i]))) return ret(handles[i], rc);
    std::vector<glia_rd_t> hs;
    std::vector<int> step;
    std::vector<hipEvent_t> eva;
    // forward: step k of every handle with nt > k; adjoint: the k-th step back from each one's end
    for (int dir = 0; dir < 2; dir++) {
      for (int k = 0; k < ntmax; k++) {
        hs.clear();
        step.clear();
        eva.clear();
        for (int i = 0; i < nh; i++) {
          const int nt = handles[i]->e->nt();
          if (k >= nt) continue;
          hs.push_back(handles[i]);
          step.push_back(dir == 0 ? k : nt - 1 - k);
          eva.push_back(ev[i]);
        }
        if (dir == 0)
          rc = lockstep_step(
              hs, step, eva,
              [](RDEnsemble* e, int i, const RDEnsemble::StepGraph** g) { return e->fwd_prepare(i, e->store_half(), g); },
              [](RDEnsemble* e, int i, bool l) { return e->fwd_finish(i, e->store_half(), l); }, &spanned);
        else
          rc = lockstep_step(
              hs, step, eva, [](RDEnsemble* e, int i, const RDEnsemble::StepGraph** g) { return e->adj_prepare(i, g); },
              [](RDEnsemble* e, int i, bool l) { return e->adj_finish(i, l); }, &spanned);
        if (rc) return rc;
      }
      if (dir == 0)
        for (int i = 0; i < nh; i++)
          if ((rc = handles[i]->e->eval_middle())) return ret(handles[i], rc);
    }
    for (int i = 0; i < nh; i++) {
      std::vector<double> Jv;
      std::vector<std::vector<double>> g;
      if ((rc = handles[i]->e->eval_end(Jv, g))) return ret(handles[i], rc);
      int64_t off = 0;
      for (int m = 0; m < handles[i]->e->nmembers(); m++) {
        J[i][m] = Jv[m];
        std::memcpy(dJ[i] + off, g[m].data(), sizeof(double) * g[m].size());
        off += (int64_t)g[m].size();
      }
    }
    return 0;
  };
  const int rc = run();
  for (int i = 0; i < nh; i++) (void)hipStreamSynchronize(handles[i]->e->stream());
  for (auto v : ev) (void)hipEventDestroy(v);
  if (spanned_steps) *spanned_steps = spanned;
  return rc;
}

int glia_rd_stats(glia_rd_t h, int member, glia_rd_stats_t* out) {
  GLIA_H(h);
  if (!out) return arg_err(h, "null output pointer");
  glia::MemberStats s;
  const int rc = h->e->stats(member, &s);
  if (rc) return ret(h, rc);
  out->pcg_iterations = s.its_total;
  out->cn_solves = s.solves;
  out->worst_reason = s.worst_reason;
  out->last_iterations = s.last_its;
  out->last_reason = s.last_reason;
  out->kbar = s.kavg;
  out->prec_coef = s.pc;
  out->filter_sum = s.filter_sum;
  out->misfit = s.m1;
  out->c0_norm2 = s.c0sq;
  return 0;
}

int glia_rd_smooth_field(glia_rd_t h, int member, int which, double sigma) {
  GLIA_H(h);
  const int slot = field_slot(which);
  if (slot < 0) return arg_err(h, "unknown field id");
  return ret(h, h->e->smooth_field(member, slot, sigma));
}

int glia_rd_data_support(glia_rd_t h, int member, int which, double threshold, double sigma, double gvf, int max_np,
                         double* centers, int* np) {
  GLIA_H(h);
  const int slot = field_slot(which);
  if (slot < 0) return arg_err(h, "unknown field id");
  if (!np || max_np < 0 || (max_np > 0 && !centers)) return arg_err(h, "null output pointer or negative max_np");
  std::vector<double> c;
  const int rc = h->e->data_support(member, slot, threshold, sigma, gvf, &c);
  if (rc) return ret(h, rc);
  *np = (int)(c.size() / 3);
  if (*np > max_np) return arg_err(h, "more support centers than max_np (np holds the count)");
  if (*np) std::memcpy(centers, c.data(), sizeof(double) * c.size());
  return 0;
}

int glia_rd_refresh_tissue(glia_rd_t h, int member) {
  GLIA_H(h);
  return ret(h, h->e->refresh_tissue(member));
}

int glia_rd_use_graphs(glia_rd_t h, int enable) {
  GLIA_H(h);
  return ret(h, h->e->set_use_graphs(enable != 0));
}

int glia_rd_profile(glia_rd_t h, int enable) {
  GLIA_H(h);
  return ret(h, h->e->profile_enable(enable != 0));
}

int glia_rd_profile_get(glia_rd_t h, int kind, double* ms, long long* launches) {
  GLIA_H(h);
  if (!ms || !launches) return arg_err(h, "null output pointer");
  return ret(h, h->e->profile_get(kind, ms, launches));
}

int glia_rd_reset_stats(glia_rd_t h) {
  GLIA_H(h);
  return ret(h, h->e->reset_counters());
}

}  // extern "C"
