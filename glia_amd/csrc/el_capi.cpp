// el_capi.cpp -- extern "C" boundary of the linear-elasticity solver (include/glia_el.h).
//
// The operator runs in spectral space on an RD engine of three members used as lanes, one per
// displacement component, so that every transform covers the three components (or three strain
// terms) in one launch.  Per application:
//   U_d = FFT(u_d) in lane d (one launch),
//   the strain terms d_j u_i + d_i u_j: the diagonal ones (lanes 0-2) and the off-diagonal ones
//   (xy, xz, yz in lanes 0-2) as cross-lane spectral combinations, each triple inverted, scaled by
//   mu and transformed back with the product in the Z pass between the two transforms; lam div u
//   in lane 0 alone the same way,
//   f_i = inverse of i (sum_j w_j S_ij + w_i L) in lane i, minus screen u_i in the inverse's Z pass,
// i.e. 20 transforms per application in 8 transform launches (the reference's sequence of gradient
// and divergence calls makes 48 single transforms).  Each lane's real slots RS_WM and RS_GM stage
// the force's transforms, six more hold the lane's component of the CG vectors x, b, r, z, p, w, so
// the transforms read and write them in place, and three hold the material (mu, lam, screen) where
// the Z passes read it.  Complex slots: U in CS_X, the combination in CS_W, the diagonal strains in
// CS_T0, the off-diagonal ones in CS_T1, L in CS_T2 of lane 0, the preconditioned spectra in CS_T3;
// CS_T7 is the transforms' work slot.
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
constexpr int CU = glia::CS_X, CS_C = glia::CS_W, CS_DIAG = glia::CS_T0, CS_OFF = glia::CS_T1, CS_L = glia::CS_T2,
              CS_PC = glia::CS_T3;
// the material in every lane's real slots, where the Z passes read it: mu (all lanes), lam (lane 0),
// the screening coefficient in RS_RHO (all lanes)
constexpr int RS_MU = glia::RS_K, RS_LAM = glia::RS_CH;
// real slot of each CG vector in every lane: x, b, r, z, p, w
enum { VX = 0, VB, VR, VZ, VP, VW, kVecs };
constexpr int kVecSlot[kVecs] = {glia::RS_CSF, glia::RS_VT, glia::RS_FILTER, glia::RS_OBS, glia::RS_D1, glia::RS_C0};
// where the strain spectrum S_ij lives: (lane, slot)
void strain_at(int i, int j, int* lane, int* slot) {
  if (i > j) std::swap(i, j);
  if (i == j) {
    *lane = i;
    *slot = CS_DIAG;
  } else {
    *lane = (i == 0) ? (j == 1 ? 0 : 1) : 2;  // xy -> 0, xz -> 1, yz -> 2
    *slot = CS_OFF;
  }
}
}  // namespace

struct glia_el_s {
  std::unique_ptr<glia::RDEnsemble> e;
  int64_t ng = 0;
  hipStream_t st = nullptr;
  float* v[kVecs][3] = {};   // lane d's real slot of each CG vector (the engine's memory)
  double* part = nullptr;    // block partials, then (at part + 3 np) the summed scalars
  std::vector<double> hpart;
  int* hflag_host = nullptr;  // mapped: set by the device when the CG stops
  int* hflag_dev = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // ends of the last enqueued iterations
  double avg[3] = {0, 0, 0};
  bool have_material = false;
  std::string err;
  ~glia_el_s() {
    if (part) (void)hipFree(part);
    if (hflag_host) (void)hipHostFree(hflag_host);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
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

float* lane_slot(glia_el_s* h, int lane, int rs) { return h->e->field_ptr(lane, rs); }

// lanes 0..k-1 take part in the engine's launches
struct Lanes {
  glia::RDEnsemble* e;
  Lanes(glia::RDEnsemble* en, int k) : e(en) { e->set_lanes(k); }
  ~Lanes() { e->set_lanes(0); }
};

Vec3 vec(glia_el_s* h, int k) { return Vec3{{h->v[k][0], h->v[k][1], h->v[k][2]}}; }
Vec3C cvec(glia_el_s* h, int k) { return Vec3C{{h->v[k][0], h->v[k][1], h->v[k][2]}}; }
Vec3 lane_field(glia_el_s* h, int rs) { return Vec3{{lane_slot(h, 0, rs), lane_slot(h, 1, rs), lane_slot(h, 2, rs)}}; }
Vec3C lane_cfield(glia_el_s* h, int rs) {
  return Vec3C{{lane_slot(h, 0, rs), lane_slot(h, 1, rs), lane_slot(h, 2, rs)}};
}

// one triple (or lane 0 alone) of strain-like terms: FFT(coef IFFT(combination)) into slot cs_out,
// the combination formed in the first inverse pass and the coefficient applied in the Z pass
int strain_channel(glia_el_s* h, int nl, const int* nterms, const int (*lane)[4], const int (*slot)[4],
                   const int (*dir)[4], int coef_rs, int cs_out) {
  Lanes use(h->e.get(), nl);
  EL_ENG(h, h->e->lin_coef_transform(nterms, lane, slot, dir, coef_rs, cs_out));
  return 0;
}

// f = A u on the lane vectors from U = FFT(u) in CU (vector ids; f must differ from u)
int apply_op_spectral(glia_el_s* h, int u, int f) {
  int rc;
  Lanes use(h->e.get(), 3);
  // d_j u_i + d_i u_j, diagonal terms (ii) in lane i, then xy, xz, yz in lanes 0, 1, 2
  {
    const int n[3] = {2, 2, 2};
    const int lane[3][4] = {{0, 0}, {1, 1}, {2, 2}}, slot[3][4] = {{CU, CU}, {CU, CU}, {CU, CU}};
    const int dir[3][4] = {{0, 0}, {1, 1}, {2, 2}};
    if ((rc = strain_channel(h, 3, n, lane, slot, dir, RS_MU, CS_DIAG))) return rc;
  }
  {
    const int n[3] = {2, 2, 2};
    const int lane[3][4] = {{0, 1}, {0, 2}, {1, 2}}, slot[3][4] = {{CU, CU}, {CU, CU}, {CU, CU}};
    const int dir[3][4] = {{1, 0}, {2, 0}, {2, 1}};
    if ((rc = strain_channel(h, 3, n, lane, slot, dir, RS_MU, CS_OFF))) return rc;
  }
  // lam div u in lane 0
  {
    const int n[3] = {3, 0, 0};
    const int lane[3][4] = {{0, 1, 2}}, slot[3][4] = {{CU, CU, CU}}, dir[3][4] = {{0, 1, 2}};
    if ((rc = strain_channel(h, 1, n, lane, slot, dir, RS_LAM, CS_L))) return rc;
  }
  // f_i = div_j(S_ij) + d_i L - screen u_i in lane i
  {
    const int n[3] = {4, 4, 4};
    int lane[3][4], slot[3][4], dir[3][4];
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) {
        strain_at(i, j, &lane[i][j], &slot[i][j]);
        dir[i][j] = j;
      }
      lane[i][3] = 0;
      slot[i][3] = CS_L;
      dir[i][3] = i;
    }
    EL_ENG(h, h->e->lin_screen_transform(n, lane, slot, dir, kVecSlot[u], kVecSlot[f]));
  }
  return 0;
}

// f = A u on the lane vectors (vector ids; f must differ from u)
int apply_op(glia_el_s* h, int u, int f) {
  {
    Lanes use(h->e.get(), 3);
    EL_ENG(h, h->e->fft_r2c(kVecSlot[u], CU));
  }
  return apply_op_spectral(h, u, f);
}

// the CG's p = z + (beta / betaold) p folded into the first pass of w = A p
int apply_op_new_p(glia_el_s* h) {
  {
    Lanes use(h->e.get(), 3);
    EL_ENG(h, h->e->fft_r2c_axpby(kVecSlot[VP], kVecSlot[VZ], CU));
  }
  return apply_op_spectral(h, VP, VW);
}

// u = B f on the lane vectors; the inverse's Z pass leaves the block partials of <u, u> and <u, f>
int apply_prec(glia_el_s* h, int f, int u) {
  Lanes use(h->e.get(), 3);
  EL_ENG(h, h->e->fft_r2c(kVecSlot[f], CU));
  EL_ENG(h, h->e->el_prec_lanes(CU, CS_PC, h->avg[0], h->avg[1], h->avg[2]));
  EL_ENG(h, h->e->c2r_dots(CS_PC, kVecSlot[u], kVecSlot[f]));
  return 0;
}

// the CG's x += t p, r -= t w (t = beta / sc[0] on the device) folded into the first pass of z = B r
int apply_prec_updated_r(glia_el_s* h, const double* sc) {
  Lanes use(h->e.get(), 3);
  EL_ENG(h, h->e->fft_r2c_cgupd(kVecSlot[VX], kVecSlot[VR], kVecSlot[VP], kVecSlot[VW], sc, CU));
  EL_ENG(h, h->e->el_prec_lanes(CU, CS_PC, h->avg[0], h->avg[1], h->avg[2]));
  EL_ENG(h, h->e->c2r_dots(CS_PC, kVecSlot[VZ], kVecSlot[VR]));
  return 0;
}

// the partials the last Z pass left (nq per block over the three lanes), summed on the device into
// dst[0..nq)
int lane_sums(glia_el_s* h, int nq, double* dst) {
  Lanes use(h->e.get(), 3);
  EL_ENG(h, h->e->lane_part_sum(nq, dst));
  return 0;
}

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
  sc.nmembers = 3;  // one lane per displacement component
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
  for (int k = 0; k < kVecs; k++)
    for (int d = 0; d < 3; d++) h->v[k][d] = lane_slot(h, d, kVecSlot[k]);
  const int np = glia::el::partial_count();
  if (hipMalloc((void**)&h->part, sizeof(double) * (3 * np + 8)) != hipSuccess) {
    delete h;
    return (int)hipErrorOutOfMemory;
  }
  h->hpart.assign(3 * np, 0.0);
  if (hipHostMalloc((void**)&h->hflag_host, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&h->hflag_dev, h->hflag_host, 0) != hipSuccess) {
    delete h;
    return (int)hipErrorOutOfMemory;
  }
  for (hipEvent_t& e : h->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      delete h;
      return (int)hipErrorOutOfMemory;
    }
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
  // device inputs are read where they are; host inputs are staged in the CG vectors (free outside
  // a solve).  The material goes straight to the lanes' slots where the Z passes read it: mu and
  // the screening in every lane, lam in lane 0.
  const float* src[6] = {c, wm, gm, csf, vt, bg};
  const float* dev[6];
  float* stage[6] = {h->v[VX][0], h->v[VX][1], h->v[VX][2], h->v[VB][0], h->v[VB][1], h->v[VB][2]};
  for (int k = 0; k < 6; k++) {
    if (mem_kind == GLIA_RD_DEVICE) {
      dev[k] = src[k];
    } else {
      EL_HIP(h, hipMemcpyAsync(stage[k], src[k], sizeof(float) * h->ng, hipMemcpyHostToDevice, h->st));
      dev[k] = stage[k];
    }
  }
  EL_HIP(h, glia::el::launch_material(h->st, h->ng, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], m,
                                      lane_field(h, RS_MU), lane_slot(h, 0, RS_LAM), lane_field(h, glia::RS_RHO),
                                      h->part));
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
  const float* src[3] = {lane_slot(h, 0, RS_MU), lane_slot(h, 0, RS_LAM), lane_slot(h, 0, glia::RS_RHO)};
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
  if ((rc = apply_op(h, VP, VW))) return rc;
  return stage_out(h, f, mem_kind, VW);
}

int glia_el_precondition(glia_el_t h, const float* const f[3], float* const u[3], int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!f || !u || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null vectors or bad mem_kind");
  if (!h->have_material) return fail(h, GLIA_RD_ERR_STATE, "material not set");
  int rc;
  if ((rc = stage_in(h, f, mem_kind, VR))) return rc;
  if ((rc = apply_prec(h, VR, VZ))) return rc;
  return stage_out(h, u, mem_kind, VZ);
}

// computeForce (src/mat/Tumor.cpp:114-158): f = ff grad(S c) (times tanh(S c)) with S the smoother
// of sigma 2 pi / n0; the smoothed c in lane 0's RS_WM, its gradient through the three lanes
int glia_el_force(glia_el_t h, const float* c, double forcing_factor, int use_tanh, float* const f[3], int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!c || !f || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null field or bad mem_kind");
  for (int d = 0; d < 3; d++)
    if (!f[d]) return fail(h, GLIA_RD_ERR_ARG, "null component");
  float* cs = lane_slot(h, 0, glia::RS_WM);
  EL_HIP(h, hipMemcpyAsync(cs, c, sizeof(float) * h->ng,
                           mem_kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
  const int n0 = h->e->dims().n0;
  EL_ENG(h, h->e->smooth_field(0, glia::RS_WM, 2.0 * M_PI / n0));  // sigma_smooth = 2 pi / n0
  {
    Lanes one(h->e.get(), 1);
    EL_ENG(h, h->e->fft_r2c(glia::RS_WM, CU));
  }
  {
    Lanes use(h->e.get(), 3);
    const int n[3] = {1, 1, 1};
    const int lane[3][4] = {{0}, {0}, {0}}, slot[3][4] = {{CU}, {CU}, {CU}}, dir[3][4] = {{0}, {1}, {2}};
    EL_ENG(h, h->e->lane_lin(n, lane, slot, dir, CS_C));
    EL_ENG(h, h->e->fft_c2r(CS_C, glia::RS_GM));
  }
  float* dst[3];
  for (int d = 0; d < 3; d++) dst[d] = mem_kind == GLIA_RD_DEVICE ? f[d] : h->v[VW][d];
  for (int d = 0; d < 3; d++)
    EL_HIP(h, glia::el::launch_force_scale(h->st, h->ng, lane_slot(h, d, glia::RS_GM), cs, (float)forcing_factor,
                                           use_tanh ? 1 : 0, dst[d]));
  if (mem_kind == GLIA_RD_DEVICE) {
    EL_HIP(h, hipStreamSynchronize(h->st));
    return 0;
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
  // KSPSolve_CG (PETSc, KSP_NORM_PRECONDITIONED, nonzero initial guess) with KSPConvergedDefault.
  // The set-up reads ||B b||, ||B r|| and <B r, r> on the host; the iterations then run without the
  // host: the scalars stay on the device, where k_el_cg_fin (lane 0's PcgState) makes the stopping
  // tests of KSPSolve_CG and the update takes its step; the host enqueues iterations two ahead of the
  // device and stops enqueueing when the device raises the mapped flag.  Iterations enqueued past the
  // stop return at once in every launch.  The vector updates run inside the first pass of the
  // transform that reads the vector next: p = z + beta p in the operator's, x and r in the
  // preconditioner's; <p, A p>, <z, z> and <z, r> come from the Z passes that write w and z.
  if ((rc = apply_op(h, VX, VR))) return rc;
  EL_HIP(h, glia::el::launch_axpby3(h->st, h->ng, vec(h, VR), cvec(h, VB), 0.0, 1));  // r = b - A x
  double rnorm0, dp;
  const int np = glia::el::partial_count();
  double* sc_dev = h->part + 3 * np;  // device scalars: <p, A p>, then <z, z>, <z, r>
  double sc[3] = {0, 0, 0};
  auto read3 = [&]() -> int {
    EL_HIP(h, hipMemcpyAsync(sc, sc_dev, sizeof(sc), hipMemcpyDeviceToHost, h->st));
    EL_HIP(h, hipStreamSynchronize(h->st));
    return 0;
  };
  // ||B b||, then ||B r|| and <B r, r>, each from the preconditioner's inverse pass
  if ((rc = apply_prec(h, VB, VZ)) || (rc = lane_sums(h, 2, sc_dev + 1)) || (rc = read3())) return rc;
  rnorm0 = std::sqrt(sc[1]);
  if ((rc = apply_prec(h, VR, VZ)) || (rc = lane_sums(h, 2, sc_dev + 1)) || (rc = read3())) return rc;
  dp = std::sqrt(sc[1]);
  if (rnorm0 == 0) rnorm0 = dp;
  const double ttol = std::fmax(rtol * rnorm0, atol);
  int its = 0, why = 0;
  if (std::isnan(dp) || std::isinf(dp))
    why = KSP_DIVERGED_NANORINF;
  else if (dp <= ttol)
    why = dp < atol ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL;
  else if (dp >= dtol * rnorm0)
    why = KSP_DIVERGED_DTOL;
  if (!why) {
    *h->hflag_host = 0;
    {
      Lanes use(h->e.get(), 3);
      EL_ENG(h, h->e->el_cg_begin(sc[2], dp, rnorm0, ttol, atol, dtol, maxit));
    }
    h->e->set_honor_lanes(true);
    constexpr int kAhead = 2;
    for (int k = 0; k < maxit && !rc; k++) {
      if ((rc = apply_op_new_p(h)) || (rc = lane_sums(h, 1, sc_dev)) || (rc = apply_prec_updated_r(h, sc_dev)) ||
          (rc = lane_sums(h, 2, sc_dev + 1)))
        break;
      {
        Lanes use(h->e.get(), 3);
        if ((rc = h->e->el_cg_fin(sc_dev, h->hflag_dev))) {
          fail(h, rc, h->e->error());
          break;
        }
      }
      hipError_t e = hipEventRecord(h->ev[k % 3], h->st);
      if (e == hipSuccess && k >= kAhead) e = hipEventSynchronize(h->ev[(k - kAhead) % 3]);
      if (e != hipSuccess) {
        rc = hip_fail(h, (int)e, "enqueueing the CG iterations");
        break;
      }
      if (*(volatile int*)h->hflag_host) break;
    }
    {
      Lanes use(h->e.get(), 3);
      const int rc2 = h->e->el_cg_state(&its, &why, &dp);
      const int rc3 = h->e->el_cg_end();
      if (!rc && (rc2 || rc3)) rc = fail(h, rc2 ? rc2 : rc3, h->e->error());
    }
    if (rc) return rc;
  }
  if ((rc = stage_out(h, u, mem_kind, VX))) return rc;
  if (iterations) *iterations = its;
  if (reason) *reason = why;
  if (rnorm) *rnorm = dp;
  return 0;
}

}  // extern "C"
