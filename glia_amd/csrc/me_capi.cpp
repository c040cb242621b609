// me_capi.cpp -- extern "C" boundary (include/glia_me.h): the mass-effect forwards of model 4
// (PdeOperatorsMassEffect::solveState, src/pde/PdeOperatorsMassEffect.cpp:421-743) and model 5
// (PdeOperatorsMultiSpecies::solveState, src/pde/PdeOperatorsMultiSpecies.cpp:235-508) over the
// diffusion engine, the elasticity solver and the advection solver, all on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "../../include/glia_me.h"
#include "../../include/glia_ms.h"
#include "me_features.h"
#include "me_launch.h"
#include "rd_handle.h"

struct glia_me_s {
  glia_rd_t rd = nullptr;
  glia_el_t el = nullptr;
  glia_sl_t sl = nullptr;
  hipStream_t st = nullptr;
  int n[3] = {0, 0, 0};
  int64_t ng = 0;
  // u_old (3), force (3), advection outputs (6: gm, wm, vt, csf and mri or p, nec), the tumor core
  // of the multispecies model (1), its edema and segmentation outputs (2), then the |v| partials
  float* work = nullptr;
  bool have_ms_outputs = false;  // edema and segmentation of a multispecies solve are in work
  std::vector<float> hpart;
  // biophysical features (allocated on the first call that asks for them): kFeatFields fields,
  // the device sums of one step, and the scratch of the reductions and sorts
  float* feat = nullptr;
  double* acc = nullptr;
  void* fscratch = nullptr;
  std::vector<double> hacc;
  std::string err;
};

namespace {

constexpr int kWorkFields = 15;
constexpr int kEdema = 13, kSeg = 14;  // work fields of the multispecies outputs

int fail(glia_me_s* h, int rc, const std::string& msg) {
  h->err = msg;
  return rc;
}

// a handle whose creation failed part way has no engine, solvers or work fields
bool complete(const glia_me_s* h) { return h->rd && h->rd->e && h->el && h->sl && h->work; }

// the CFL limit of a call: p->max_cfl when positive, else the reference's limit of the model
// (14 in PdeOperatorsMassEffect.cpp:472, 8 in PdeOperatorsMultiSpecies.cpp:296); -1 for NaN
double cfl_limit(const glia_me_params* p, double model_default) {
  if (std::isnan(p->max_cfl)) return -1.0;
  return p->max_cfl > 0 ? p->max_cfl : model_default;
}

// feature fields: seg, three components (grad c, then c v), |grad c|, D c, dc/dt, |u|, |v|, grad u
// (9), mu, lam, det F, tr S, max shear, two sort key arrays, c* (one value)
enum { FS_SEG = 0, FS_G = 1, FS_GRADC = 4, FS_DC, FS_DCDT, FS_UMAG, FS_VMAG, FS_GU = 9, FS_MU = 18, FS_LAM, FS_JAC,
       FS_TR, FS_TAU, FS_KEYS, FS_KEYS2, FS_CSTAR, kFeatFields };
constexpr int kFeatures = 8;  // c, dcdt, gradc, u, jac, trT, tau, vel

}  // namespace

#define ME_HIP(h, call)                                                                         \
  do {                                                                                          \
    const hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail(h, (int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// a call into one of the owned solvers: its own message on failure
#define ME_SUB(h, call, lasterr)                                             \
  do {                                                                       \
    const int r_ = (call);                                                   \
    if (r_) return fail(h, r_, std::string(#call) + ": " + (lasterr));       \
  } while (0)

namespace {

// computeGradient of a device field through the engine's transforms (member 0; the engine's
// alpha(0) field and its transform work slots serve as scratch)
int feat_gradient(glia_me_s* h, const float* f, float* const g[3]) {
  glia::RDEnsemble* e = h->rd->e.get();
  float* ra = e->field_ptr(0, glia::RS_A0);
  const size_t fb = sizeof(float) * h->ng;
  ME_HIP(h, hipMemcpyAsync(ra, f, fb, hipMemcpyDeviceToDevice, h->st));
  ME_SUB(h, e->fft_r2c(glia::RS_A0, glia::CS_T0), e->error());
  const int slot = glia::CS_T0;
  for (int d = 0; d < 3; d++) {
    ME_SUB(h, e->spec_lin(1, &slot, &d, glia::CS_T1), e->error());
    ME_SUB(h, e->fft_c2r(glia::CS_T1, glia::RS_A0), e->error());
    ME_HIP(h, hipMemcpyAsync(g[d], ra, fb, hipMemcpyDeviceToDevice, h->st));
  }
  return 0;
}

// computeDivergence of three device fields into out (out may be x[0])
int feat_divergence(glia_me_s* h, const float* const x[3], float* out) {
  glia::RDEnsemble* e = h->rd->e.get();
  float* ra = e->field_ptr(0, glia::RS_A0);
  const size_t fb = sizeof(float) * h->ng;
  const int slots[3] = {glia::CS_T0, glia::CS_T1, glia::CS_T2}, dirs[3] = {0, 1, 2};
  for (int d = 0; d < 3; d++) {
    ME_HIP(h, hipMemcpyAsync(ra, x[d], fb, hipMemcpyDeviceToDevice, h->st));
    ME_SUB(h, e->fft_r2c(glia::RS_A0, slots[d]), e->error());
  }
  ME_SUB(h, e->spec_lin(3, slots, dirs, glia::CS_T3), e->error());
  ME_SUB(h, e->fft_c2r(glia::CS_T3, glia::RS_A0), e->error());
  ME_HIP(h, hipMemcpyAsync(out, ra, fb, hipMemcpyDeviceToDevice, h->st));
  return 0;
}

int feat_alloc(glia_me_s* h) {
  if (h->feat) return 0;
  const int64_t ng = h->ng;
  if (hipMalloc((void**)&h->feat, sizeof(float) * ((int64_t)(kFeatFields - 1) * ng + 64)) != hipSuccess ||
      hipMalloc((void**)&h->acc, sizeof(double) * kFeatures * glia::mef::kPerField) != hipSuccess ||
      hipMalloc(&h->fscratch, glia::mef::scratch_bytes(ng)) != hipSuccess) {
    (void)hipGetLastError();
    if (h->feat) (void)hipFree(h->feat);
    if (h->acc) (void)hipFree(h->acc);
    h->feat = nullptr;
    h->acc = nullptr;
    h->fscratch = nullptr;
    return fail(h, (int)hipErrorOutOfMemory, "out of device memory for the feature fields");
  }
  h->hacc.assign(kFeatures * glia::mef::kPerField, 0.0);
  return 0;
}

// PdeOperatorsMassEffect::computeBiophysicalFeatures (src/pde/PdeOperatorsMassEffect.cpp:323-418) at
// time step i: one row of GLIA_ME_NFEATURES values (the reference's CSV row) and the tumor ratio
// tc / (tc + healthy) in single precision, as the loop's stopping rule reads it
int feat_step(glia_me_s* h, int i, double dt, const float* c, const float* wm, const float* gm, const float* vt,
              const float* csf, const float* bg, float* const u[3], float* const v[3], double* row, float* ratio) {
  int rc = feat_alloc(h);
  if (rc) return rc;
  glia::RDEnsemble* e = h->rd->e.get();
  const int64_t ng = h->ng;
  const hipStream_t st = h->st;
  auto F = [&](int k) { return h->feat + (int64_t)k * ng; };
  float* G[3] = {F(FS_G), F(FS_G + 1), F(FS_G + 2)};
  float* cstar = F(FS_CSTAR);
  // Tumor::computeSegmentation, then tc = [seg == 1] and healthy = [seg not in {0, 1}]
  ME_SUB(h, glia_ms_segmentation(ng, bg, gm, wm, vt, csf, c, F(FS_SEG), (void*)st), "segmentation");
  // computeTumorQuants: |grad c|, dc/dt = D c + rho (c - c^2) - div(c v), c* (:214-259)
  if ((rc = feat_gradient(h, c, G))) return rc;
  ME_HIP(h, (hipError_t)glia::mef::launch_mag3(st, ng, G[0], G[1], G[2], F(FS_GRADC)));
  float* ra = e->field_ptr(0, glia::RS_A0);
  ME_HIP(h, hipMemcpyAsync(ra, c, sizeof(float) * ng, hipMemcpyDeviceToDevice, st));
  ME_SUB(h, e->apply_d_field(glia::RS_A0, glia::RS_A0), e->error());
  ME_HIP(h, hipMemcpyAsync(F(FS_DC), ra, sizeof(float) * ng, hipMemcpyDeviceToDevice, st));
  ME_HIP(h, (hipError_t)glia::mef::launch_flux(st, ng, c, v, G));
  if ((rc = feat_divergence(h, G, G[0]))) return rc;
  ME_HIP(h, (hipError_t)glia::mef::launch_dcdt(st, ng, F(FS_DC), c, e->field_ptr(0, glia::RS_RHO), G[0], F(FS_DCDT)));
  ME_HIP(h, (hipError_t)glia::mef::launch_cstar(st, ng, F(FS_GRADC), F(FS_SEG), c, h->fscratch, cstar));
  // |u|, |v|, grad u and the stress quantities with the elasticity solver's current mu and lam
  ME_HIP(h, (hipError_t)glia::mef::launch_mag3(st, ng, u[0], u[1], u[2], F(FS_UMAG)));
  ME_HIP(h, (hipError_t)glia::mef::launch_mag3(st, ng, v[0], v[1], v[2], F(FS_VMAG)));
  glia::mef::StressIn gi;
  for (int a = 0; a < 3; a++) {
    float* ga[3] = {F(FS_GU + 3 * a), F(FS_GU + 3 * a + 1), F(FS_GU + 3 * a + 2)};
    if ((rc = feat_gradient(h, u[a], ga))) return rc;
    for (int b = 0; b < 3; b++) gi.d[3 * a + b] = ga[b];
  }
  ME_SUB(h, glia_el_get_material(h->el, F(FS_MU), F(FS_LAM), nullptr, nullptr, GLIA_RD_DEVICE), glia_el_last_error(h->el));
  ME_HIP(h, (hipError_t)glia::mef::launch_stress(st, ng, gi, F(FS_MU), F(FS_LAM), F(FS_JAC), F(FS_TR), F(FS_TAU)));
  // writeStats of c, dcdt, |grad c|, |u|, det F, tr S, max shear, |v|
  const float* fields[kFeatures] = {c, F(FS_DCDT), F(FS_GRADC), F(FS_UMAG), F(FS_JAC), F(FS_TR), F(FS_TAU), F(FS_VMAG)};
  for (int k = 0; k < kFeatures; k++)
    ME_HIP(h, (hipError_t)glia::mef::launch_stats(st, ng, fields[k], F(FS_SEG), c, cstar, h->fscratch, F(FS_KEYS),
                                                   F(FS_KEYS2), h->acc, k * glia::mef::kPerField));
  ME_HIP(h, hipMemcpyAsync(h->hacc.data(), h->acc, sizeof(double) * h->hacc.size(), hipMemcpyDeviceToHost, st));
  ME_HIP(h, hipStreamSynchronize(st));
  const double h3 = (2.0 * M_PI / h->n[0]) * (2.0 * M_PI / h->n[1]) * (2.0 * M_PI / h->n[2]);
  const double sa_measure = h3 * h->n[0] / (2.0 * M_PI);
  const double* a0 = h->hacc.data();
  row[0] = i * dt;
  row[1] = a0[4] / (a0[5] + a0[4]);
  *ratio = (float)a0[4] / ((float)a0[5] + (float)a0[4]);
  for (int k = 0; k < kFeatures; k++) {
    const double* a = a0 + k * glia::mef::kPerField;
    double* r = row + 2 + 16 * k;
    r[0] = a[0] * h3;
    r[1] = a[1] * h3;
    r[2] = a[2] * h3;
    r[3] = a[3] * sa_measure;
    r[4] = std::sqrt(a[6] / (double)(ng - 1));
    r[5] = std::sqrt(a[7] / (a[4] - 1.0));
    r[6] = std::sqrt(a[8] / (a[5] - 1.0));
    for (int q = 0; q < 3; q++)
      for (int reg = 0; reg < 3; reg++) r[7 + 3 * q + reg] = a[9 + 3 * reg + q];
  }
  return 0;
}

}  // namespace

extern "C" {

void glia_me_params_init(glia_me_params* p) {
  if (!p) return;
  p->forcing_factor = 1e5;
  p->use_tanh = 1;
  p->max_cfl = 0.0;  // the model's own limit: 14 for model 4, 8 for model 5
  glia_el_params_init(&p->el);
}

int glia_me_destroy(glia_me_t h) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (h->st) (void)hipStreamSynchronize(h->st);
  if (h->work) (void)hipFree(h->work);
  if (h->feat) (void)hipFree(h->feat);
  if (h->acc) (void)hipFree(h->acc);
  if (h->fscratch) (void)hipFree(h->fscratch);
  if (h->sl) glia_sl_destroy(h->sl);
  if (h->el) glia_el_destroy(h->el);
  if (h->rd) glia_rd_destroy(h->rd);
  delete h;
  return 0;
}

int glia_me_create(const glia_rd_config* cfg, glia_me_t* out) {
  if (!cfg || !out) return GLIA_RD_ERR_ARG;
  *out = nullptr;
  if (cfg->nmembers != 1) return GLIA_RD_ERR_ARG;
  glia_me_s* h = new (std::nothrow) glia_me_s();
  if (!h) return GLIA_RD_ERR_ARG;
  // as glia_rd_create: a failed creation still hands out the handle, for glia_me_last_error, and
  // the caller destroys it
  *out = h;
  int rc = glia_rd_create(cfg, &h->rd);
  if (rc) {
    h->err = std::string("engine: ") + (h->rd ? glia_rd_last_error(h->rd) : "creation failed");
    return rc;
  }
  h->st = h->rd->e->stream();
  for (int d = 0; d < 3; d++) h->n[d] = cfg->n[d];
  h->ng = (int64_t)cfg->n[0] * cfg->n[1] * cfg->n[2];
  if ((rc = glia_el_create(cfg->n, (void*)h->st, &h->el))) {
    h->err = std::string("elasticity solver: ") + (h->el ? glia_el_last_error(h->el) : "creation failed");
    return rc;
  }
  if ((rc = glia_sl_create(cfg->n, (void*)h->st, &h->sl))) {
    h->err = std::string("advection solver: ") + (h->sl ? glia_sl_last_error(h->sl) : "creation failed");
    return rc;
  }
  const int np = glia::me::partial_count();
  h->hpart.resize(np);
  if (hipMalloc((void**)&h->work, sizeof(float) * (kWorkFields * h->ng + np)) != hipSuccess) {
    (void)hipGetLastError();
    h->work = nullptr;
    return fail(h, (int)hipErrorOutOfMemory, "out of device memory for the work fields");
  }
  return 0;
}

const char* glia_me_last_error(glia_me_t h) { return h ? h->err.c_str() : "null handle"; }
glia_rd_t glia_me_rd(glia_me_t h) { return h ? h->rd : nullptr; }
glia_el_t glia_me_el(glia_me_t h) { return h ? h->el : nullptr; }
glia_sl_t glia_me_sl(glia_me_t h) { return h ? h->sl : nullptr; }

int glia_me_solve_state(glia_me_t h, const glia_me_params* p, float* c, float* wm, float* gm, float* vt, float* csf,
                        float* mri, const float* bg, float* const u[3], float* const v[3], glia_me_info* info) {
  return glia_me_solve_state_features(h, p, c, wm, gm, vt, csf, mri, bg, u, v, info, nullptr, 0, nullptr);
}

const char* glia_me_feature_name(int k) {
  static const std::vector<std::string> names = []() {
    std::vector<std::string> out = {"t", "volc/volb"};
    const char* feats[kFeatures] = {"c", "dcdt", "gradc", "u", "jac", "trT", "tau", "vel"};
    const char* stats[6] = {"vol", "sa", "std", "quartmed", "95", "99"};
    for (const char* f : feats)
      for (const char* st : stats) {
        if (std::string(st) == "sa") {
          out.push_back(std::string(st) + "_" + f);
        } else {
          out.push_back(std::string(st) + "_" + f);
          out.push_back(std::string(st) + "_tc_" + f);
          out.push_back(std::string(st) + "_b_" + f);
        }
      }
    return out;
  }();
  return (k >= 0 && k < (int)names.size()) ? names[k].c_str() : nullptr;
}

int glia_me_solve_state_features(glia_me_t h, const glia_me_params* p, float* c, float* wm, float* gm, float* vt,
                                 float* csf, float* mri, const float* bg, float* const u[3], float* const v[3],
                                 glia_me_info* info, double* features, int max_rows, int* rows) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (rows) *rows = 0;
  if (features && (!rows || max_rows < h->rd->e->config().nt + 1))
    return fail(h, GLIA_RD_ERR_ARG, "features need rows and max_rows >= nt + 1");
  if (!p || !c || !wm || !gm || !vt || !csf || !bg || !u || !v)
    return fail(h, GLIA_RD_ERR_ARG, "null parameters or field");
  for (int d = 0; d < 3; d++)
    if (!u[d] || !v[d]) return fail(h, GLIA_RD_ERR_ARG, "null displacement or velocity component");
  if (!complete(h)) return fail(h, GLIA_RD_ERR_STATE, "the handle was not fully created: " + h->err);
  const double max_cfl = cfl_limit(p, 14.0);
  if (!(max_cfl > 0)) return fail(h, GLIA_RD_ERR_ARG, "max_cfl is NaN");
  glia::RDEnsemble* e = h->rd->e.get();
  const hipStream_t st = h->st;
  const int64_t ng = h->ng;
  const double dt = e->config().dt;
  const int nt = e->config().nt;
  const double h0 = 2.0 * M_PI / h->n[0], sigma = 2.0 * M_PI / h->n[0];
  float* uold[3] = {h->work, h->work + ng, h->work + 2 * ng};
  float* force[3] = {h->work + 3 * ng, h->work + 4 * ng, h->work + 5 * ng};
  float* adv[5];
  for (int k = 0; k < 5; k++) adv[k] = h->work + (6 + k) * ng;
  float* part = h->work + kWorkFields * ng;
  float* ct = e->field_ptr(0, glia::RS_CT);
  float* rho = e->field_ptr(0, glia::RS_RHO);
  float* kf = e->field_ptr(0, glia::RS_K);
  float* ws = e->field_ptr(0, glia::RS_CH);
  const int np = glia::me::partial_count();
  const size_t fb = sizeof(float) * ng;
  glia::MemberStats s0{};
  ME_SUB(h, e->stats(0, &s0), e->error());
  glia_me_info inf{};

  // filterTumor, computeForce, the elasticity solve into u_old from zero (:456-460)
  ME_HIP(h, (hipError_t)glia::me::launch_filter(st, ng, c, wm, gm));
  for (int d = 0; d < 3; d++) ME_HIP(h, hipMemsetAsync(uold[d], 0, fb, st));
  int its = 0;
  ME_SUB(h, glia_el_force(h->el, c, p->forcing_factor, p->use_tanh, force, GLIA_RD_DEVICE), glia_el_last_error(h->el));
  ME_SUB(h, glia_el_set_material(h->el, c, wm, gm, csf, vt, bg, &p->el, GLIA_RD_DEVICE), glia_el_last_error(h->el));
  ME_SUB(h, glia_el_solve(h->el, force, uold, GLIA_RD_DEVICE, &its, nullptr, nullptr), glia_el_last_error(h->el));
  inf.el_iterations += its;

  for (int i = 0; i < nt + 1; i++) {
    // CFL from the current velocity (:489-502)
    ME_HIP(h, (hipError_t)glia::me::launch_vmax(st, ng, v[0], v[1], v[2], part));
    ME_HIP(h, hipMemcpyAsync(h->hpart.data(), part, sizeof(float) * np, hipMemcpyDeviceToHost, st));
    ME_HIP(h, hipStreamSynchronize(st));
    float vmax = 0.0f;
    for (int b = 0; b < np; b++) vmax = std::max(vmax, h->hpart[b]);
    const double cfl = dt * (double)vmax / h0;
    inf.cfl_max = std::max(inf.cfl_max, cfl);
    if (cfl >= max_cfl) {
      inf.cfl_stop = 1;
      break;
    }
    // updateReacAndDiffCoefficients and precFactor (:578-582)
    ME_SUB(h, glia_ms_update_coefficients(ng, rho, kf, bg, gm, vt, csf, e->rho(0), e->kappa(0), e->config().r_gm_wm,
                                          e->config().k_gm_wm, (void*)st),
           "coefficient update");
    ME_SUB(h, e->prec_factor(), e->error());
    // feature_compute_ (:584-604): the features of this step, then the stop once the tumor core
    // holds about 5% of the brain (ratio within 5e-3 of 0.05, or above) after t = 1
    if (features) {
      float ratio = 0.0f;
      int rc = feat_step(h, i, dt, c, wm, gm, vt, csf, bg, u, v, features + (size_t)i * GLIA_ME_NFEATURES, &ratio);
      if (rc) return rc;
      (*rows)++;
      if ((std::fabs(ratio - 0.05f) < 5e-3f || ratio > 0.05f) && (float)i * (float)dt >= 1.0f) {
        inf.volume_stop = 1;
        break;
      }
    }
    // advection of gm, wm, vt, csf, c and the MRI on one set of trajectories (:607-621); c lands
    // in the engine's c(T) field, where the diffusion solve and the reaction continue
    ME_SUB(h, glia_sl_set_velocity(h->sl, v[0], v[1], v[2], GLIA_RD_DEVICE), glia_sl_last_error(h->sl));
    const float* in[6] = {gm, wm, vt, csf, c, mri};
    float* outp[6] = {adv[0], adv[1], adv[2], adv[3], ct, adv[4]};
    const int modes[6] = {GLIA_SL_MASS_CONSERVATION, GLIA_SL_MASS_CONSERVATION, GLIA_SL_PURE_ADVECTION,
                          GLIA_SL_PURE_ADVECTION,    GLIA_SL_MASS_CONSERVATION, GLIA_SL_PURE_ADVECTION};
    ME_SUB(h, glia_sl_advect(h->sl, mri ? 6 : 5, in, outp, modes, dt), glia_sl_last_error(h->sl));
    float* back[5] = {gm, wm, vt, csf, mri};
    for (int k = 0; k < (mri ? 5 : 4); k++)
      ME_HIP(h, hipMemcpyAsync(back[k], adv[k], fb, hipMemcpyDeviceToDevice, st));
    // diffusion over dt and the logistic reaction (:628-631)
    ME_SUB(h, e->diffusion_solve_field(glia::RS_CT, dt), e->error());
    ME_HIP(h, (hipError_t)glia::me::launch_reaction(st, ng, ct, rho, (float)dt));
    ME_HIP(h, hipMemcpyAsync(c, ct, fb, hipMemcpyDeviceToDevice, st));
    // force, displacement, velocity (:634-664)
    ME_SUB(h, glia_el_force(h->el, c, p->forcing_factor, p->use_tanh, force, GLIA_RD_DEVICE),
           glia_el_last_error(h->el));
    ME_SUB(h, glia_el_set_material(h->el, c, wm, gm, csf, vt, bg, &p->el, GLIA_RD_DEVICE), glia_el_last_error(h->el));
    ME_SUB(h, glia_el_solve(h->el, force, u, GLIA_RD_DEVICE, &its, nullptr, nullptr), glia_el_last_error(h->el));
    inf.el_iterations += its;
    ME_HIP(h, (hipError_t)glia::me::launch_velocity(st, ng, v, u, uold, (float)(1.0 / dt)));
    for (int d = 0; d < 3; d++) ME_SUB(h, e->smooth_ptr(v[d], ws, sigma), e->error());
    inf.steps++;
  }
  ME_HIP(h, hipStreamSynchronize(st));
  glia::MemberStats s1{};
  ME_SUB(h, e->stats(0, &s1), e->error());
  inf.diff_iterations = s1.its_total - s0.its_total;
  if (info) *info = inf;
  return 0;
}

int glia_me_solve_state_multispecies(glia_me_t h, const glia_me_params* p, const glia_ms_params* ms, const float* c0,
                                     float* prolif, float* infil, float* nec, float* ox, float* wm, float* gm, float* vt,
                                     float* csf, const float* bg, float* const u[3], float* const v[3],
                                     glia_me_info* info) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!p || !ms || !c0 || !prolif || !infil || !nec || !ox || !wm || !gm || !vt || !csf || !bg || !u || !v)
    return fail(h, GLIA_RD_ERR_ARG, "null parameters or field");
  for (int d = 0; d < 3; d++)
    if (!u[d] || !v[d]) return fail(h, GLIA_RD_ERR_ARG, "null displacement or velocity component");
  if (!complete(h)) return fail(h, GLIA_RD_ERR_STATE, "the handle was not fully created: " + h->err);
  const double max_cfl = cfl_limit(p, 8.0);
  if (!(max_cfl > 0)) return fail(h, GLIA_RD_ERR_ARG, "max_cfl is NaN");
  glia::RDEnsemble* e = h->rd->e.get();
  const hipStream_t st = h->st;
  const int64_t ng = h->ng;
  const double dt = e->config().dt;
  const int nt = e->config().nt;
  const bool moving = p->forcing_factor > 0;
  const double h0 = 2.0 * M_PI / h->n[0], sigma = 2.0 * M_PI / h->n[0];
  float* uold[3] = {h->work, h->work + ng, h->work + 2 * ng};
  float* force[3] = {h->work + 3 * ng, h->work + 4 * ng, h->work + 5 * ng};
  float* adv[6];
  for (int k = 0; k < 6; k++) adv[k] = h->work + (6 + k) * ng;
  float* core = h->work + 12 * ng;
  float* edema = h->work + kEdema * ng;
  float* seg = h->work + kSeg * ng;
  float* part = h->work + kWorkFields * ng;
  float* ib = e->field_ptr(0, glia::RS_CT);  // i through advection, clipping, smoothing and diffusion
  float* rho = e->field_ptr(0, glia::RS_RHO);
  float* kf = e->field_ptr(0, glia::RS_K);
  float* ws = e->field_ptr(0, glia::RS_CH);  // D i for the sources, then the velocity smoothing
  float* tissue[4] = {e->field_ptr(0, glia::RS_WM), e->field_ptr(0, glia::RS_GM), e->field_ptr(0, glia::RS_CSF),
                      e->field_ptr(0, glia::RS_VT)};
  const int np = glia::me::partial_count();
  const size_t fb = sizeof(float) * ng;
  glia::MemberStats s0{};
  ME_SUB(h, e->stats(0, &s0), e->error());
  glia_me_info inf{};
  int its = 0;

  // start (:251-283): species, clipTumor, filterTumor with p + i, the force of p and the
  // elasticity solve into u, u_old = u
  ME_HIP(h, (hipError_t)glia::me::launch_species_init(st, ng, c0, prolif, infil, ox));
  ME_SUB(h, glia_ms_clip(ng, nec, (void*)st), "clip");
  ME_HIP(h, (hipError_t)glia::me::launch_sum(st, ng, prolif, infil, nullptr, core));
  ME_HIP(h, (hipError_t)glia::me::launch_filter(st, ng, core, wm, gm));
  if (moving) {
    ME_SUB(h, glia_el_force(h->el, prolif, p->forcing_factor, p->use_tanh, force, GLIA_RD_DEVICE),
           glia_el_last_error(h->el));
    ME_SUB(h, glia_el_set_material(h->el, prolif, wm, gm, csf, vt, bg, &p->el, GLIA_RD_DEVICE),
           glia_el_last_error(h->el));
    ME_SUB(h, glia_el_solve(h->el, force, u, GLIA_RD_DEVICE, &its, nullptr, nullptr), glia_el_last_error(h->el));
    inf.el_iterations += its;
    for (int d = 0; d < 3; d++) ME_HIP(h, hipMemcpyAsync(uold[d], u[d], fb, hipMemcpyDeviceToDevice, st));
  }

  h->have_ms_outputs = false;
  for (int i = 0; i < nt + 1; i++) {
    // Tumor::computeEdema and computeSegmentation at the top of the step (:307-308): the edema
    // mask of i smoothed with sigma 2 pi / n0 (src/mat/Tumor.cpp:160-178) and the labels of the
    // current tissue and tumor c_t_ (p at the first step, then p + nec + i)
    ME_SUB(h, glia_ms_edema_mask(ng, infil, ms->invasive_threshold, edema, (void*)st), "edema mask");
    ME_SUB(h, e->smooth_ptr(edema, ws, sigma), e->error());
    ME_SUB(h, glia_ms_segmentation(ng, bg, gm, wm, vt, csf, i == 0 ? prolif : core, seg, (void*)st), "segmentation");
    h->have_ms_outputs = true;
    // CFL from the current velocity (:309-321)
    ME_HIP(h, (hipError_t)glia::me::launch_vmax(st, ng, v[0], v[1], v[2], part));
    ME_HIP(h, hipMemcpyAsync(h->hpart.data(), part, sizeof(float) * np, hipMemcpyDeviceToHost, st));
    ME_HIP(h, hipStreamSynchronize(st));
    float vmax = 0.0f;
    for (int b = 0; b < np; b++) vmax = std::max(vmax, h->hpart[b]);
    const double cfl = dt * (double)vmax / h0;
    inf.cfl_max = std::max(inf.cfl_max, cfl);
    if (cfl >= max_cfl) {
      inf.cfl_stop = 1;
      break;
    }
    // rho from the tissue, then k and its mean over the initial brain mask from the current wm,
    // gm and csf (updateIsotropicCoefficients), precFactor (:412-416)
    ME_SUB(h, glia_ms_update_coefficients(ng, rho, kf, bg, gm, vt, csf, e->rho(0), e->kappa(0), e->config().r_gm_wm,
                                          e->config().k_gm_wm, (void*)st),
           "coefficient update");
    const float* cur[4] = {wm, gm, csf, vt};
    for (int k = 0; k < 4; k++) ME_HIP(h, hipMemcpyAsync(tissue[k], cur[k], fb, hipMemcpyDeviceToDevice, st));
    ME_SUB(h, e->set_kappa(0, e->kappa(0)), e->error());
    ME_SUB(h, e->prec_factor(), e->error());
    // advection (:418-431): gm, wm, p, i, nec conserve mass, vt and csf are transported
    if (moving) {
      ME_SUB(h, glia_sl_set_velocity(h->sl, v[0], v[1], v[2], GLIA_RD_DEVICE), glia_sl_last_error(h->sl));
      const float* in[7] = {gm, wm, vt, csf, prolif, infil, nec};
      float* outp[7] = {adv[0], adv[1], adv[2], adv[3], adv[4], ib, adv[5]};
      const int modes[7] = {GLIA_SL_MASS_CONSERVATION, GLIA_SL_MASS_CONSERVATION, GLIA_SL_PURE_ADVECTION,
                            GLIA_SL_PURE_ADVECTION,    GLIA_SL_MASS_CONSERVATION, GLIA_SL_MASS_CONSERVATION,
                            GLIA_SL_MASS_CONSERVATION};
      ME_SUB(h, glia_sl_advect(h->sl, 7, in, outp, modes, dt), glia_sl_last_error(h->sl));
      float* back[6] = {gm, wm, vt, csf, prolif, nec};
      for (int k = 0; k < 6; k++) ME_HIP(h, hipMemcpyAsync(back[k], adv[k], fb, hipMemcpyDeviceToDevice, st));
    } else {
      ME_HIP(h, hipMemcpyAsync(ib, infil, fb, hipMemcpyDeviceToDevice, st));
    }
    // clipHealthyTissues, clipTumor (:435-437)
    float* clip[7] = {gm, wm, csf, vt, prolif, ib, nec};
    for (int k = 0; k < 7; k++) ME_SUB(h, glia_ms_clip(ng, clip[k], (void*)st), "clip");
    // i smoothed, D i, the diffusion solve of i (:440-447)
    ME_SUB(h, e->smooth_field(0, glia::RS_CT, sigma), e->error());
    ME_SUB(h, e->apply_d_field(glia::RS_CT, glia::RS_CH), e->error());
    ME_SUB(h, e->diffusion_solve_field(glia::RS_CT, dt), e->error());
    ME_HIP(h, hipMemcpyAsync(infil, ib, fb, hipMemcpyDeviceToDevice, st));
    // computeSources and the tumor core (:452-456)
    ME_SUB(h, glia_ms_sources(ng, prolif, infil, nec, ox, gm, wm, rho, ws, dt, ms, (void*)st), "sources");
    ME_HIP(h, (hipError_t)glia::me::launch_sum(st, ng, prolif, nec, infil, core));
    // force, displacement, velocity (:458-489)
    if (moving) {
      ME_SUB(h, glia_el_force(h->el, core, p->forcing_factor, p->use_tanh, force, GLIA_RD_DEVICE),
             glia_el_last_error(h->el));
      ME_SUB(h, glia_el_set_material(h->el, core, wm, gm, csf, vt, bg, &p->el, GLIA_RD_DEVICE),
             glia_el_last_error(h->el));
      ME_SUB(h, glia_el_solve(h->el, force, u, GLIA_RD_DEVICE, &its, nullptr, nullptr), glia_el_last_error(h->el));
      inf.el_iterations += its;
      ME_HIP(h, (hipError_t)glia::me::launch_velocity(st, ng, v, u, uold, (float)(1.0 / dt)));
      for (int d = 0; d < 3; d++) ME_SUB(h, e->smooth_ptr(v[d], ws, sigma), e->error());
    }
    inf.steps++;
  }
  ME_HIP(h, hipStreamSynchronize(st));
  glia::MemberStats s1{};
  ME_SUB(h, e->stats(0, &s1), e->error());
  inf.diff_iterations = s1.its_total - s0.its_total;
  if (info) *info = inf;
  return 0;
}

int glia_me_multispecies_outputs(glia_me_t h, float* edema, float* seg, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!edema && !seg) return fail(h, GLIA_RD_ERR_ARG, "no output requested");
  if (mem_kind != GLIA_RD_HOST && mem_kind != GLIA_RD_DEVICE) return fail(h, GLIA_RD_ERR_ARG, "bad mem_kind");
  if (!complete(h) || !h->have_ms_outputs)
    return fail(h, GLIA_RD_ERR_STATE, "no multispecies solve has computed the edema and segmentation");
  const hipMemcpyKind k = mem_kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  const size_t fb = sizeof(float) * h->ng;
  if (edema) ME_HIP(h, hipMemcpyAsync(edema, h->work + kEdema * h->ng, fb, k, h->st));
  if (seg) ME_HIP(h, hipMemcpyAsync(seg, h->work + kSeg * h->ng, fb, k, h->st));
  ME_HIP(h, hipStreamSynchronize(h->st));
  return 0;
}

}  // extern "C"
