// rd_solver.cpp -- host side of the spectral reaction-diffusion solver (see rd_solver.h).
//
// The path mirrors the reference call sequence:
//   DerivativeOperatorsRD::evaluateObjectiveAndGradient (src/grad/DerivativeOperatorsRD.cpp:130-226)
//     updateReactionAndDiffusion -> Phi::apply -> PdeOperatorsRD::solveState -> misfit ->
//     PdeOperatorsRD::solveAdjoint -> Phi::applyTranspose, gradDiffusion, gradReaction
// with every CN half step (DiffusionSolver::solve, src/pde/DiffusionSolver.cpp:217-250) solved by
// preconditioned CG in spectral space.  The CG iterates are the Fourier transforms of the
// reference's real-space iterates (the preconditioner is diagonal there and inner products follow
// from Parseval), and the stopping logic is PETSc 3.11 KSPCG with KSPConvergedDefault, evaluated
// on the device.  The host only enqueues work and watches a mapped completion flag a few
// iterations behind the device.
#include "rd_solver.h"

#include <algorithm>
#include <atomic>
#include <fstream>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace glia {

#define GLIA_TRYC(x)            \
  do {                          \
    int rc_ = (x);              \
    if (rc_ != 0) return fail(rc_, std::string("launch failed: ") + #x); \
  } while (0)
#define GLIA_HIP(x)                                      \
  do {                                                   \
    hipError_t e_ = (x);                                 \
    if (e_ != hipSuccess) return check(e_, #x);          \
  } while (0)

static bool supported_z(int n2) {
  return n2 == 32 || n2 == 64 || n2 == 96 || n2 == 128 || n2 == 160 || n2 == 192 || n2 == 256;
}

RDEnsemble::RDEnsemble(const SolverConfig& cfg) : cfg_(cfg) {
  static std::atomic<unsigned long long> next_uid{1};
  uid_ = next_uid++;
  nm_ = cfg.nmembers;
  d_.n0 = cfg.n[0];
  d_.n1 = cfg.n[1];
  d_.n2 = cfg.n[2];
  d_.m2 = cfg.n[2] / 2;
  d_.nspec = (int64_t)d_.n0 * d_.n1 * d_.m2;
  d_.ng = (int64_t)d_.n0 * d_.n1 * d_.n2;
  d_.inv_ng = 1.0 / (double)d_.ng;
  for (int i = 0; i < 3; i++) h_[i] = 2.0 * M_PI / cfg.n[i];
  ctx_dt_ = cfg.dt;  // Ctx::dt_ = params dt at construction (src/pde/DiffusionSolver.cpp:11)
}

// blocking copy ordered on the solver stream (the stream is non-blocking w.r.t. the null stream)
int RDEnsemble::copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  return check(e, "hipMemcpyAsync");
}

int RDEnsemble::fail(int code, const std::string& msg) {
  err_ = msg;
  return code ? code : -1;
}

int RDEnsemble::check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  std::ostringstream s;
  s << what << ": " << hipGetErrorString(e);
  err_ = s.str();
  return (int)e;
}

LaunchCtx RDEnsemble::ctx() const {
  LaunchCtx c;
  c.d = d_;
  c.mem = md_;
  c.nm = (lanes_ > 0 && lanes_ < nm_) ? lanes_ : nm_;
  c.stream = stream_;
  c.tw0 = tw_[0];
  c.tw1 = tw_[1];
  c.twm = tw_[2];
  c.tw2 = tw_[3];
  c.yf_split = yf_split_;
  c.xf_split = xf_split_;
  return c;
}

// launch context of member m alone
static LaunchCtx member_ctx(LaunchCtx c, const MemberDev* md, int m) {
  c.mem = md + m;
  c.nm = 1;
  return c;
}

int RDEnsemble::init() {
  if (nm_ < 1) return fail(-2, "nmembers must be >= 1");
  // captured steps unroll each solve's CG loop to the last count of its kind plus this margin;
  // GLIA_STEP_MARGIN < 0 makes every captured step run again eagerly (a check of that path)
  if (const char* v = std::getenv("GLIA_STEP_MARGIN")) margin_ = std::atoi(v);
  // forward divergence variants: the X pass in two thread groups on power-of-two lengths (measured
  // faster at 128 and 256 points) and in one on the others (faster at 160 and 192), the Y pass in
  // one.  GLIA_SPLIT_DIV overrides with a comma-separated list of none, xf (X pass in two groups),
  // xf3 (X pass in three groups), yf (Y pass in three groups), yfb / yfb2 (Y pass split over blocks
  // at 3 / 2 waves per SIMD), both (xf and yf); glia_rd_autotune measures them.
  xf_split_ = (d_.n0 & (d_.n0 - 1)) == 0 ? 1 : 0;
  yf_split_ = 0;
  if (const char* v = std::getenv("GLIA_SPLIT_DIV")) {
    xf_split_ = 0;
    std::istringstream toks(v);
    std::string tok;
    while (std::getline(toks, tok, ',')) {
      if (tok == "xf" || tok == "both") xf_split_ = 1;
      if (tok == "xf3") xf_split_ = 2;
      if (tok == "yf" || tok == "both") yf_split_ = 1;
      if (tok == "yfb") yf_split_ = 2;
      if (tok == "yfb2") yf_split_ = 3;
    }
  }
  if (!supported_line(d_.n0) || !supported_line(d_.n1) || !supported_z(d_.n2))
    return fail(-4, "unsupported grid: n0, n1 in {32,48,64,80,96,128,160,192,256}, n2 in {32,64,96,128,160,192,256}");
  if (z_blocks(d_) <= 0 || ((int64_t)d_.n0 * d_.n1) % z_lines_per_block(d_.m2) != 0)
    return fail(-4, "unsupported grid: n0*n1 must be a multiple of the z-pass line group");
  if (cfg_.nt < 1 || !(cfg_.dt > 0)) return fail(-2, "nt >= 1 and dt > 0 required");
  if (cfg_.maxit < 1) return fail(-2, "pcg maxit >= 1 required");
  if (cfg_.checkpoint_slots < 0) return fail(-2, "checkpoint_slots >= 0 required");
  if (cfg_.stream) {
    stream_ = cfg_.stream;
  } else {
    GLIA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  // twiddle tables in FP64, stored as FP32: n0, n1, n2/2, n2
  {
    const int lens[4] = {d_.n0, d_.n1, d_.m2, d_.n2};
    int64_t tot = 0;
    for (int i = 0; i < 4; i++) tot += lens[i];
    std::vector<float2> h(tot);
    int64_t off = 0;
    int64_t offs[4];
    for (int i = 0; i < 4; i++) {
      offs[i] = off;
      for (int m = 0; m < lens[i]; m++) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / lens[i];
        h[off + m] = make_float2((float)cosl(a), (float)sinl(a));
      }
      off += lens[i];
    }
    GLIA_HIP(hipMalloc(&twbuf_, sizeof(float2) * tot));
    GLIA_TRYC(copy_sync(twbuf_, h.data(), sizeof(float2) * tot, hipMemcpyHostToDevice));
    for (int i = 0; i < 4; i++) tw_[i] = twbuf_ + offs[i];
  }
  GLIA_HIP(hipHostMalloc((void**)&hflag_host_, sizeof(int) * (kFlagSlots + kStepFlags) * nm_,
                         hipHostMallocMapped | hipHostMallocCoherent));
  GLIA_HIP(hipHostGetDevicePointer((void**)&hflag_dev_, hflag_host_, 0));
  // The host reads the convergence flags from coherent host memory that the finalize kernels
  // write with a system-scope fence, so these events only need a device-scope release (a
  // system-scope release would write back L2 on every CG iteration).
  for (int i = 0; i < kFlagSlots; i++)
    GLIA_HIP(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming | hipEventReleaseToDevice));
  GLIA_HIP(hipEventCreateWithFlags(&ev_step_, hipEventDisableTiming | hipEventReleaseToDevice));
  ev_ok_ = true;
  GLIA_HIP(hipMalloc(&coefs_dev_, sizeof(double) * 3 * nm_));

  hist_all_ = (cfg_.checkpoint_slots == 0);
  npool_ = hist_all_ ? cfg_.nt + 1 : std::max(1, cfg_.checkpoint_slots);
  mh_.resize(nm_);
  pool_.resize(nm_);
  chalf_.resize(nm_);
  apool_.resize(nm_);
  hist_dev_.assign(nm_, nullptr);
  chalf_dev_.assign(nm_, nullptr);
  st_.assign(nm_, nullptr);
  part_.assign(nm_, nullptr);
  np_.assign(nm_, 0);
  phimax_.assign(nm_, {});
  centers_.assign(nm_, {});
  sigma_.assign(nm_, 0.0);
  bricks_dev_.assign(nm_, nullptr);
  brick_data_.assign(nm_, nullptr);
  phi_scale_dev_.assign(nm_, nullptr);
  phit_dev_.assign(nm_, nullptr);
  kappa_.assign(nm_, 0.0);
  rho_.assign(nm_, 0.0);
  active_.assign(nm_, 0);
  for (int m = 0; m < nm_; m++) {
    int rc = alloc_member(m);
    if (rc) return rc;
  }
  GLIA_HIP(hipMalloc(&md_, sizeof(MemberDev) * nm_));
  return upload_members();
}

int RDEnsemble::alloc_member(int m) {
  MemberDev& M = mh_[m];
  std::memset(&M, 0, sizeof(M));
  const size_t cb = sizeof(float2) * d_.nspec, rb = sizeof(float) * d_.ng;
  for (int s = 0; s < CS_NUM; s++) {
    GLIA_HIP(hipMalloc(&M.c[s], cb));
    GLIA_HIP(hipMemsetAsync(M.c[s], 0, cb, stream_));
  }
  for (int s = 0; s < RS_NUM; s++) {
    GLIA_HIP(hipMalloc(&M.r[s], rb));
    GLIA_HIP(hipMemsetAsync(M.r[s], 0, rb, stream_));
  }
  GLIA_HIP(hipMalloc(&st_[m], sizeof(PcgState)));
  PcgState s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.rtol = cfg_.rtol;
  s0.atol = 1e-50;  // PETSc defaults
  s0.dtol = 1e5;
  s0.maxit = cfg_.maxit;
  s0.active = 0;
  s0.filter_sum = (double)d_.ng;
  s0.done = 1;
  GLIA_TRYC(copy_sync(st_[m], &s0, sizeof(s0), hipMemcpyHostToDevice));
  M.st = st_[m];
  // partial sums: the largest reduction grid times its width
  int64_t nb = std::max<int64_t>((int64_t)xf_div_blocks(d_) * 3, (int64_t)z_blocks(d_) * 2);
  nb = std::max<int64_t>(nb, (int64_t)pcg_update_blocks() * 2);
  nb = std::max<int64_t>(nb, (int64_t)ew_blocks() * 2);
  GLIA_HIP(hipMalloc(&part_[m], sizeof(double) * nb));
  M.part = part_[m];
  int rc = alloc_history(m);
  if (rc) return rc;
  M.kr_gm = (float)cfg_.k_gm_wm;
  M.rr_gm = (float)cfg_.r_gm_wm;
  // MatProp::setValues defaults (src/mat/MatProp.cpp:20-33): wm = 1, filter = 1; Obs filter = 1
  std::vector<float> ones(d_.ng, 1.0f);
  GLIA_TRYC(copy_sync(M.r[RS_WM], ones.data(), rb, hipMemcpyHostToDevice));
  GLIA_TRYC(copy_sync(M.r[RS_FILTER], ones.data(), rb, hipMemcpyHostToDevice));
  GLIA_TRYC(copy_sync(M.r[RS_OBS], ones.data(), rb, hipMemcpyHostToDevice));
  return 0;
}

// the time history of member m for the current nt: the spectral state pool (store-all: c_[0..nt];
// checkpointing: the checkpoint slots), the half-step states c_half_ (adjoint_store_, store-all)
// and, when kept, the adjoint history p_[0..nt] (PdeOperatorsRD ctor, src/pde/PdeOperators.cpp:26-53)
int RDEnsemble::alloc_history(int m) {
  MemberDev& M = mh_[m];
  const size_t cb = sizeof(float2) * d_.nspec, rb = sizeof(float) * d_.ng;
  pool_[m].assign(npool_, nullptr);
  for (int i = 0; i < npool_; i++) GLIA_HIP(hipMalloc(&pool_[m][i], cb));
  GLIA_HIP(hipMalloc(&hist_dev_[m], sizeof(float2*) * npool_));
  GLIA_TRYC(copy_sync(hist_dev_[m], pool_[m].data(), sizeof(float2*) * npool_, hipMemcpyHostToDevice));
  M.hist = hist_dev_[m];
  if (cfg_.adjoint_store && hist_all_) {
    chalf_[m].assign(cfg_.nt, nullptr);
    for (int i = 0; i < cfg_.nt; i++) GLIA_HIP(hipMalloc(&chalf_[m][i], rb));
    GLIA_HIP(hipMalloc(&chalf_dev_[m], sizeof(float*) * cfg_.nt));
    GLIA_TRYC(copy_sync(chalf_dev_[m], chalf_[m].data(), sizeof(float*) * cfg_.nt, hipMemcpyHostToDevice));
  }
  M.chalf = chalf_dev_[m];
  if (keep_adj_ && hist_all_) {
    apool_[m].assign(cfg_.nt + 1, nullptr);
    for (int i = 0; i <= cfg_.nt; i++) GLIA_HIP(hipMalloc(&apool_[m][i], cb));
  }
  return 0;
}

void RDEnsemble::free_history(int m) {
  for (auto p : pool_[m]) (void)hipFree(p);
  for (auto p : chalf_[m]) (void)hipFree(p);
  for (auto p : apool_[m]) (void)hipFree(p);
  (void)hipFree(hist_dev_[m]);
  (void)hipFree(chalf_dev_[m]);
  pool_[m].clear();
  chalf_[m].clear();
  apool_[m].clear();
  hist_dev_[m] = nullptr;
  chalf_dev_[m] = nullptr;
  mh_[m].hist = nullptr;
  mh_[m].chalf = nullptr;
}

// TumorParameters dt_ / nt_ and PdeOperatorsRD::resizeTimeHistory (src/pde/PdeOperators.cpp:56-103):
// the history is reallocated for the new nt; stored states are dropped.  The diffusion solver's
// context dt is left as it is, as in the reference.
int RDEnsemble::set_time(double dt, int nt) {
  if (nt < 1 || !(dt > 0)) return fail(-2, "nt >= 1 and dt > 0 required");
  GLIA_HIP(hipStreamSynchronize(stream_));
  cfg_.dt = dt;
  if (nt == cfg_.nt) return 0;
  for (int m = 0; m < nm_; m++) free_history(m);
  cfg_.nt = nt;
  npool_ = hist_all_ ? cfg_.nt + 1 : std::max(1, cfg_.checkpoint_slots);
  stored_.clear();
  free_pool_.clear();
  for (int m = 0; m < nm_; m++) {
    const int rc = alloc_history(m);
    if (rc) return rc;
  }
  return upload_members();
}

// keep the adjoint history p_[0..nt] during solveAdjoint (store-all only): nt + 1 spectral fields
// per member
int RDEnsemble::keep_adjoint_history(bool on) {
  if (on && !hist_all_) return fail(-4, "the adjoint history is kept in store-all mode only (checkpoint_slots = 0)");
  if (on == keep_adj_) return 0;
  GLIA_HIP(hipStreamSynchronize(stream_));
  keep_adj_ = on;
  const size_t cb = sizeof(float2) * d_.nspec;
  for (int m = 0; m < nm_; m++) {
    for (auto p : apool_[m]) (void)hipFree(p);
    apool_[m].clear();
    if (on) {
      apool_[m].assign(cfg_.nt + 1, nullptr);
      for (int i = 0; i <= cfg_.nt; i++) GLIA_HIP(hipMalloc(&apool_[m][i], cb));
    }
  }
  adj_valid_ = false;
  return 0;
}

int RDEnsemble::store_adjoint(int i) {
  if (!keep_adj_) return 0;
  const size_t cb = sizeof(float2) * d_.nspec;
  for (int m = 0; m < nm_; m++)
    GLIA_HIP(hipMemcpyAsync(apool_[m][i], mh_[m].c[CS_A], cb, hipMemcpyDeviceToDevice, stream_));
  return 0;
}

// the history entries the reference keeps as public members of PdeOperatorsRD: c_[index]
// (0..nt), c_half_[index] (0..nt-1, adjoint_store_) and p_[index] (0..nt, when kept), as a real
// field of member m.  Under checkpointing only the stored time indices are available.
int RDEnsemble::get_history(int m, int kind, int index, float* dst, bool device) {
  if (m < 0 || m >= nm_) return fail(-2, "bad member");
  const int nt = cfg_.nt;
  const size_t cb = sizeof(float2) * d_.nspec;
  LaunchCtx c = member_ctx(ctx(), md_, m);
  if (kind == HIST_STATE) {
    if (index < 0 || index > nt) return fail(-2, "state history index must be in [0, nt]");
    auto it = stored_.find(index);
    if (it == stored_.end()) return fail(-6, "this time index of the state history is not stored");
    GLIA_HIP(hipMemcpyAsync(mh_[m].c[CS_T6], pool_[m][it->second], cb, hipMemcpyDeviceToDevice, stream_));
  } else if (kind == HIST_HALF) {
    if (index < 0 || index >= nt) return fail(-2, "half-step history index must be in [0, nt)");
    if (!(cfg_.adjoint_store && hist_all_) || stored_.empty())
      return fail(-6, "the half-step history is kept with adjoint_store in store-all mode, after solveState");
    GLIA_HIP(hipMemcpyAsync(dst, chalf_[m][index], sizeof(float) * d_.ng,
                            device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream_));
    GLIA_HIP(hipStreamSynchronize(stream_));
    return 0;
  } else if (kind == HIST_ADJOINT) {
    if (index < 0 || index > nt) return fail(-2, "adjoint history index must be in [0, nt]");
    if (!keep_adj_ || !adj_valid_) return fail(-6, "the adjoint history is not kept (keep it, then run solveAdjoint)");
    GLIA_HIP(hipMemcpyAsync(mh_[m].c[CS_T6], apool_[m][index], cb, hipMemcpyDeviceToDevice, stream_));
  } else {
    return fail(-2, "unknown history kind");
  }
  GLIA_TRYC(launch_col_plain(c, 0, +1, CS_T6, CS_T5, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_T5, CS_T5, false));
  GLIA_TRYC(launch_z_point(c, ZPM_C2R, CS_T5, 0, RS_CH, -1, 0.f, false));
  return get_field(m, RS_CH, dst, device);
}

// DiffCoef::setSecondaryCoefficients: k~ = kt (m_wm + k_gm m_gm + k_glm m_csf) of member m, the
// coefficient field of the solveState(2) sources
int RDEnsemble::set_secondary(int m, double kt) {
  if (m < 0 || m >= nm_) return fail(-2, "bad member");
  LaunchCtx cm = member_ctx(ctx(), md_, m);
  GLIA_TRYC(launch_axpby_real(cm, RS_WM, kt, RS_GM, kt * cfg_.k_gm_wm, RS_KS));
  GLIA_TRYC(launch_axpby_real(cm, RS_KS, 1.0, RS_CSF, kt * cfg_.k_glm_wm, RS_KS));
  return 0;
}

// gradDiffusion / gradReaction (src/grad/DerivativeOperators.cpp:189-321) of member m, nk = nr = 1,
// from the trapezoid sums the last adjoint sweep accumulated
int RDEnsemble::grad_kappa_rho(int m, double* gk, double* gr) {
  if (m < 0 || m >= nm_) return fail(-2, "bad member");
  PcgState s;
  GLIA_TRYC(copy_sync(&s, st_[m], sizeof(s), hipMemcpyDeviceToHost));
  const double h3 = lebesgue();
  if (gk) *gk = h3 * cfg_.dt * s.gk;
  if (gr) *gr = h3 * cfg_.dt * s.gr;
  return 0;
}

int RDEnsemble::upload_members() {
  GLIA_TRYC(copy_sync(md_, mh_.data(), sizeof(MemberDev) * nm_, hipMemcpyHostToDevice));
  return 0;
}

RDEnsemble::~RDEnsemble() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (smooth_taps_) (void)hipFree(smooth_taps_);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  for (auto& r : prof_recs_) {
    (void)hipEventDestroy(r.a);
    if (r.b) (void)hipEventDestroy(r.b);
  }
  for (auto e : prof_pool_) (void)hipEventDestroy(e);
  for (int m = 0; m < (int)mh_.size(); m++) {
    for (int s = 0; s < CS_NUM; s++) (void)hipFree(mh_[m].c[s]);
    for (int s = 0; s < RS_NUM; s++) (void)hipFree(mh_[m].r[s]);
    free_history(m);
    (void)hipFree(st_[m]);
    (void)hipFree(part_[m]);
    (void)hipFree(bricks_dev_[m]);
    (void)hipFree(brick_data_[m]);
    (void)hipFree(phi_scale_dev_[m]);
    (void)hipFree(phit_dev_[m]);
  }
  (void)hipFree(md_);
  (void)hipFree(twbuf_);
  (void)hipFree(coefs_dev_);
  if (hflag_host_) (void)hipHostFree(hflag_host_);
  drop_step_graphs();
  if (ev_ok_) {
    for (int i = 0; i < kFlagSlots; i++) (void)hipEventDestroy(ev_[i]);
    (void)hipEventDestroy(ev_step_);
  }
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

int RDEnsemble::sync() {
  GLIA_HIP(hipStreamSynchronize(stream_));
  return 0;
}

// ---------------------------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------------------------

int RDEnsemble::set_field(int m, int rslot, const float* src, bool device) {
  if (m < 0 || m >= nm_ || rslot < 0 || rslot >= RS_NUM) return fail(-2, "bad member or field");
  const size_t rb = sizeof(float) * d_.ng;
  if (!src) {
    GLIA_HIP(hipMemsetAsync(mh_[m].r[rslot], 0, rb, stream_));
  } else {
    GLIA_HIP(hipMemcpyAsync(mh_[m].r[rslot], src, rb, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  }
  if (!device) GLIA_HIP(hipStreamSynchronize(stream_));
  return 0;
}

int RDEnsemble::get_field(int m, int rslot, float* dst, bool device) {
  if (m < 0 || m >= nm_ || rslot < 0 || rslot >= RS_NUM) return fail(-2, "bad member or field");
  GLIA_HIP(hipMemcpyAsync(dst, mh_[m].r[rslot], sizeof(float) * d_.ng, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          stream_));
  GLIA_HIP(hipStreamSynchronize(stream_));
  return 0;
}

// MatProp::setValuesCustom + the recomputation SolverInterface::updateTumorCoefficients
// (src/SolverInterface.cpp:986-1014) performs: k, rho, Phi and the preconditioner
int RDEnsemble::set_tissue(int m, const float* wm, const float* gm, const float* csf, const float* vt, bool device) {
  int rc;
  if ((rc = set_field(m, RS_WM, wm, device))) return rc;
  if ((rc = set_field(m, RS_GM, gm, device))) return rc;
  if ((rc = set_field(m, RS_CSF, csf, device))) return rc;
  if ((rc = set_field(m, RS_VT, vt, device))) return rc;
  return refresh_tissue(m);  // Phi::setValues runs again on the new mask
}

// MatProp::setValuesSinusoidal (src/mat/MatProp.cpp:34-68)
int RDEnsemble::set_tissue_sinusoidal(int m) {
  std::vector<float> wm(d_.ng), gm(d_.ng);
  const double f = 4.0;
  for (int x = 0; x < d_.n0; x++)
    for (int y = 0; y < d_.n1; y++)
      for (int z = 0; z < d_.n2; z++) {
        const int64_t i = ((int64_t)x * d_.n1 + y) * d_.n2 + z;
        const double w = 0.5 + 0.5 * std::sin(f * 2.0 * M_PI / d_.n0 * x) * std::sin(f * 2.0 * M_PI / d_.n1 * y) *
                                   std::sin(f * 2.0 * M_PI / d_.n2 * z);
        wm[i] = (float)w;
        gm[i] = (float)(1 - w);
      }
  int rc;
  if ((rc = set_field(m, RS_WM, wm.data(), false))) return rc;
  if ((rc = set_field(m, RS_GM, gm.data(), false))) return rc;
  if ((rc = set_field(m, RS_CSF, nullptr, false))) return rc;
  if ((rc = set_field(m, RS_VT, nullptr, false))) return rc;
  std::vector<float> ones(d_.ng, 1.0f);
  if ((rc = set_field(m, RS_FILTER, ones.data(), false))) return rc;
  const double fs = (double)d_.ng;
  GLIA_TRYC(copy_sync(&st_[m]->filter_sum, &fs, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

// DiffCoef::setValues (src/mat/DiffCoef.cpp:77-132): k = dk_gm gm + dk_wm wm + dk_glm csf,
// kbar = sum(k)/sum(filter)
int RDEnsemble::set_kappa(int m, double kappa) {
  kappa_[m] = kappa;
  double dk_gm = kappa * cfg_.k_gm_wm, dk_wm = kappa, dk_glm = kappa * cfg_.k_glm_wm;
  dk_gm = (dk_gm <= 0) ? 0.0 : dk_gm;
  dk_glm = (dk_glm <= 0) ? 0.0 : dk_glm;
  const double co[3] = {dk_gm, dk_wm, dk_glm};
  GLIA_TRYC(copy_sync(coefs_dev_ + 3 * m, co, sizeof(co), hipMemcpyHostToDevice));
  LaunchCtx c = member_ctx(ctx(), md_, m);
  GLIA_TRYC(launch_coef(c, RS_K, coefs_dev_ + 3 * m));
  GLIA_TRYC(launch_fin_state(c, ew_blocks(), FINK_KSUM, 0.0));
  active_[m] = (kappa != 0.0);
  const int a = active_[m];
  GLIA_HIP(hipMemcpyAsync(&st_[m]->active, &a, sizeof(int), hipMemcpyHostToDevice, stream_));
  return sync();
}

// ReacCoef::setValues (src/mat/ReacCoef.cpp:13-38)
int RDEnsemble::set_rho(int m, double rho) {
  rho_[m] = rho;
  double dr_gm = rho * cfg_.r_gm_wm, dr_wm = rho, dr_glm = rho * cfg_.r_glm_wm;
  dr_gm = (dr_gm <= 0) ? 0.0 : dr_gm;
  dr_glm = (dr_glm <= 0) ? 0.0 : dr_glm;
  const double co[3] = {dr_gm, dr_wm, dr_glm};
  GLIA_TRYC(copy_sync(coefs_dev_ + 3 * m, co, sizeof(co), hipMemcpyHostToDevice));
  LaunchCtx c = member_ctx(ctx(), md_, m);
  GLIA_TRYC(launch_coef(c, RS_RHO, coefs_dev_ + 3 * m));
  return sync();
}

int RDEnsemble::set_coefficients(int m, double kappa, double rho) {
  int rc = set_kappa(m, kappa);
  if (rc) return rc;
  return set_rho(m, rho);
}

// DiffCoef::setValuesSinusoidal (src/mat/DiffCoef.cpp:134-189): kbar = sum(k)/ng
int RDEnsemble::set_k_sinusoidal(int m, double scale) {
  std::vector<float> k(d_.ng);
  const double f = 4.0;
  double s = 0;
  for (int x = 0; x < d_.n0; x++)
    for (int y = 0; y < d_.n1; y++)
      for (int z = 0; z < d_.n2; z++) {
        const int64_t i = ((int64_t)x * d_.n1 + y) * d_.n2 + z;
        const double v = scale * (0.5 + 0.5 * std::sin(f * 2.0 * M_PI / d_.n0 * x) * std::sin(f * 2.0 * M_PI / d_.n1 * y) *
                                            std::sin(f * 2.0 * M_PI / d_.n2 * z));
        k[i] = (float)v;
        s += (double)k[i];
      }
  int rc = set_field(m, RS_K, k.data(), false);
  if (rc) return rc;
  const double kavg = s / (double)d_.ng;
  kappa_[m] = scale;
  active_[m] = (scale != 0.0);
  GLIA_TRYC(copy_sync(&st_[m]->kavg, &kavg, sizeof(double), hipMemcpyHostToDevice));
  GLIA_TRYC(copy_sync(&st_[m]->active, &active_[m], sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

// DiffusionSolver::precFactor (src/pde/DiffusionSolver.cpp:119-180): fixes (ctx dt, kbar)
int RDEnsemble::prec_factor() {
  GLIA_TRYC(launch_prec_factor(ctx(), ctx_dt_));
  return 0;
}

// Phi::setGaussians + Phi::setValues (src/mat/Phi.cpp:76-144): bricks of the masked, smoothed,
// truncated Gaussians; the smoothing is the separable 8-image periodic Gaussian of
// computeWeierstrassFilter (src/cuda/SpectralOperators.cu:12-51), normalized to unit sum.
int RDEnsemble::set_basis(int m, int np, const double* centers, double sigma) {
  if (m < 0 || m >= nm_) return fail(-2, "bad member");
  if (np < 0 || (np > 0 && !centers) || !(sigma > 0)) return fail(-2, "bad basis");
  GLIA_HIP(hipStreamSynchronize(stream_));
  (void)hipFree(bricks_dev_[m]);
  (void)hipFree(brick_data_[m]);
  (void)hipFree(phi_scale_dev_[m]);
  (void)hipFree(phit_dev_[m]);
  bricks_dev_[m] = nullptr;
  brick_data_[m] = nullptr;
  phi_scale_dev_[m] = nullptr;
  phit_dev_[m] = nullptr;
  np_[m] = np;
  phimax_[m].assign(np, 0.0);
  centers_[m].assign(centers, centers + 3 * np);
  sigma_[m] = sigma;
  const int n[3] = {d_.n0, d_.n1, d_.n2};
  const double sig_s = cfg_.smoothing_factor * 2.0 * M_PI / d_.n0;
  int kr[3] = {0, 0, 0}, tap_stride = 1;
  std::vector<double> taps1;
  smoother_taps(sig_s, taps1, kr, tap_stride);
  std::vector<Brick> bk(np);
  std::vector<BrickJob> jobs(np);
  int64_t off = 0;
  for (int i = 0; i < np; i++) {
    for (int a = 0; a < 3; a++) {
      const double c = centers[3 * i + a];
      int lo = (int)std::floor((c - 5.0 * sigma) / h_[a]);
      int hi = (int)std::ceil((c + 5.0 * sigma) / h_[a]);
      lo = std::max(lo, 0);
      hi = std::min(hi, n[a] - 1);
      if (hi < lo) hi = lo;
      bk[i].o[a] = lo;
      bk[i].b[a] = hi - lo + 1;
      jobs[i].o[a] = lo;
      jobs[i].b[a] = hi - lo + 1;
      jobs[i].ctr[a] = c;
    }
    bk[i].off = off;
    jobs[i].out_off = off;
    for (int a = 0; a < 3; a++) jobs[i].kr[a] = kr[a];
    jobs[i].member = 0;
    off += (int64_t)bk[i].b[0] * bk[i].b[1] * bk[i].b[2];
  }
  const int64_t total = std::max<int64_t>(off, 1);
  GLIA_HIP(hipMalloc(&brick_data_[m], sizeof(float) * total));
  GLIA_HIP(hipMalloc(&bricks_dev_[m], sizeof(Brick) * std::max(np, 1)));
  GLIA_HIP(hipMalloc(&phi_scale_dev_[m], sizeof(double) * std::max(np, 1)));
  GLIA_HIP(hipMalloc(&phit_dev_[m], sizeof(double) * std::max(np, 1)));
  GLIA_HIP(hipMemsetAsync(phi_scale_dev_[m], 0, sizeof(double) * std::max(np, 1), stream_));
  if (np > 0) GLIA_TRYC(copy_sync(bricks_dev_[m], bk.data(), sizeof(Brick) * np, hipMemcpyHostToDevice));
  // build in chunks bounded by ~512 MB of FP64 scratch
  LaunchCtx c = member_ctx(ctx(), md_, m);
  int j0 = 0;
  while (j0 < np) {
    int64_t ext = 0;
    int j1 = j0;
    while (j1 < np) {
      const int64_t ne = (int64_t)(jobs[j1].b[0] + 2 * kr[0]) * (jobs[j1].b[1] + 2 * kr[1]) * (jobs[j1].b[2] + 2 * kr[2]);
      if (j1 > j0 && (ext + ne) * 8 > (int64_t)512 << 20) break;
      jobs[j1].ext_off = ext;
      ext += ne;
      j1++;
    }
    std::vector<double> taps(3 * tap_stride * (j1 - j0));
    for (int j = j0; j < j1; j++) std::copy(taps1.begin(), taps1.end(), taps.begin() + (j - j0) * 3 * tap_stride);
    GLIA_TRYC(launch_bricks(c, jobs.data() + j0, j1 - j0, taps.data(), tap_stride, h_[0], h_[1], h_[2], sigma, brick_data_[m],
                            phimax_[m].data() + j0));
    j0 = j1;
  }
  mh_[m].bricks = bricks_dev_[m];
  mh_[m].brick_data = brick_data_[m];
  mh_[m].phi_scale = phi_scale_dev_[m];
  mh_[m].phit_out = phit_dev_[m];
  mh_[m].np = np;
  return upload_members();
}

// Obs::setDefaultFilter / setCustomFilter (src/mat/Obs.cpp:26-69)
int RDEnsemble::set_data(int m, const float* d1, const float* obs_mask, double thr, bool device) {
  int rc = set_field(m, RS_D1, d1, device);
  if (rc) return rc;
  if (obs_mask) return set_field(m, RS_OBS, obs_mask, device);
  GLIA_TRYC(launch_obs_default(member_ctx(ctx(), md_, m), (float)thr));
  return sync();
}

// ---------------------------------------------------------------------------------------------
// transforms and the CN / PCG solve
// ---------------------------------------------------------------------------------------------
int RDEnsemble::clear_done() {
  GLIA_TRYC(launch_set_done(ctx(), 0));
  return 0;
}

int RDEnsemble::forward_transform_real(int rslot, int cslot) {
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_z_point(c, ZPM_R2C, 0, cslot, rslot, -1, 0.f, false));
  GLIA_TRYC(launch_col_plain(c, 1, -1, cslot, cslot, false));
  GLIA_TRYC(launch_col_plain(c, 0, -1, cslot, cslot, false));
  return 0;
}

int RDEnsemble::inverse_to_real(int cslot, int work, int rslot) {
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_col_plain(c, 0, +1, cslot, work, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, work, work, false));
  GLIA_TRYC(launch_z_point(c, ZPM_C2R, work, 0, rslot, -1, 0.f, false));
  return 0;
}

int RDEnsemble::fft_r2c(int rslot, int cslot) { return forward_transform_real(rslot, cslot); }
int RDEnsemble::fft_c2r(int cslot, int rslot) { return inverse_to_real(cslot, CS_T7, rslot); }

int RDEnsemble::spec_lin(int nterms, const int* slots, const int* dirs, int cout) {
  GLIA_TRYC(launch_spec_lin(ctx(), nterms, slots, dirs, cout));
  return 0;
}

int RDEnsemble::lane_lin(const int* nterms, const int (*lane)[4], const int (*slot)[4], const int (*dir)[4],
                         int cout) {
  GLIA_TRYC(launch_lane_lin(ctx(), nterms, lane, slot, dir, cout));
  return 0;
}

int RDEnsemble::el_prec_lanes(int cf, int cu, double mu, double lam, double scr) {
  if (nm_ < 3) return fail(-2, "the lane preconditioner needs three members");
  GLIA_TRYC(launch_el_prec_lanes(ctx(), cf, cu, mu, lam, scr, honor_lanes_));
  return 0;
}

int RDEnsemble::lin_coef_transform(const int* nterms, const int (*lane)[4], const int (*slot)[4], const int (*dir)[4],
                                   int coef_rs, int cs_out) {
  LaunchCtx c = ctx();
  const bool hd = honor_lanes_;
  GLIA_TRYC(launch_col_lin(c, nterms, lane, slot, dir, CS_T7, hd));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_T7, CS_T7, hd));
  GLIA_TRYC(launch_z_point(c, ZPM_COEF, CS_T7, CS_T7, coef_rs, -1, 0.f, hd));
  GLIA_TRYC(launch_col_plain(c, 1, -1, CS_T7, CS_T7, hd));
  GLIA_TRYC(launch_col_plain(c, 0, -1, CS_T7, cs_out, hd));
  return 0;
}

int RDEnsemble::lin_screen_transform(const int* nterms, const int (*lane)[4], const int (*slot)[4],
                                     const int (*dir)[4], int u_rs, int out_rs) {
  LaunchCtx c = ctx();
  const bool hd = honor_lanes_;
  GLIA_TRYC(launch_col_lin(c, nterms, lane, slot, dir, CS_T7, hd));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_T7, CS_T7, hd));
  GLIA_TRYC(launch_z_point(c, ZPM_SCREEN, CS_T7, u_rs, out_rs, -1, 0.f, hd));
  return 0;
}

int RDEnsemble::c2r_dots(int cs_in, int out_rs, int r_rs) {
  LaunchCtx c = ctx();
  const bool hd = honor_lanes_;
  GLIA_TRYC(launch_col_plain(c, 0, +1, cs_in, CS_T7, hd));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_T7, CS_T7, hd));
  GLIA_TRYC(launch_z_point(c, ZPM_C2R_DOT, CS_T7, r_rs, out_rs, -1, 0.f, hd));
  return 0;
}

int RDEnsemble::lane_part_sum(int nq, double* out) {
  GLIA_TRYC(launch_lane_part_sum(ctx(), nq, out));
  return 0;
}

int RDEnsemble::fft_r2c_axpby(int p_rs, int z_rs, int cs_out) {
  LaunchCtx c = ctx();
  const bool hd = honor_lanes_;
  GLIA_TRYC(launch_z_upd_r2c(c, ZUM_AXPBY, p_rs, z_rs, -1, -1, nullptr, cs_out, hd));
  GLIA_TRYC(launch_col_plain(c, 1, -1, cs_out, cs_out, hd));
  GLIA_TRYC(launch_col_plain(c, 0, -1, cs_out, cs_out, hd));
  return 0;
}

int RDEnsemble::fft_r2c_cgupd(int x_rs, int r_rs, int p_rs, int w_rs, const double* sc, int cs_out) {
  if (!sc) return fail(-2, "fft_r2c_cgupd needs the device scalar <p, A p>");
  LaunchCtx c = ctx();
  const bool hd = honor_lanes_;
  GLIA_TRYC(launch_z_upd_r2c(c, ZUM_CGUPD, x_rs, r_rs, p_rs, w_rs, sc, cs_out, hd));
  GLIA_TRYC(launch_col_plain(c, 1, -1, cs_out, cs_out, hd));
  GLIA_TRYC(launch_col_plain(c, 0, -1, cs_out, cs_out, hd));
  return 0;
}

int RDEnsemble::el_cg_begin(double beta0, double dp0, double rnorm0, double ttol, double atol, double dtol, int maxit) {
  GLIA_TRYC(launch_el_cg_begin(ctx(), beta0, dp0, rnorm0, ttol, atol, dtol, maxit));
  return 0;
}

int RDEnsemble::el_cg_fin(const double* sc, int* hflag) {
  GLIA_TRYC(launch_el_cg_fin(ctx(), sc, hflag));
  return 0;
}

int RDEnsemble::el_cg_end() {
  honor_lanes_ = false;
  GLIA_TRYC(launch_set_done(ctx(), 0));
  return 0;
}

int RDEnsemble::el_cg_state(int* its, int* reason, double* dp) {
  PcgState s;
  GLIA_HIP(hipMemcpyAsync(&s, mh_[0].st, sizeof(PcgState), hipMemcpyDeviceToHost, stream_));
  GLIA_HIP(hipStreamSynchronize(stream_));
  if (its) *its = s.its;
  if (reason) *reason = s.reason;
  if (dp) *dp = s.dp;
  return 0;
}

int RDEnsemble::el_prec(const int f[3], const int u[3], double mu, double lam, double scr) {
  GLIA_TRYC(launch_el_prec(ctx(), make_int4(f[0], f[1], f[2], 0), make_int4(u[0], u[1], u[2], 0), mu, lam, scr));
  return 0;
}

int RDEnsemble::divergence(int rx, int ry, int rz, int rout) {
  int rc;
  if ((rc = forward_transform_real(rx, CS_T0)) || (rc = forward_transform_real(ry, CS_T1)) ||
      (rc = forward_transform_real(rz, CS_T2)))
    return rc;
  GLIA_TRYC(launch_spec_div(ctx(), CS_T0, CS_T1, CS_T2, CS_T3));
  return inverse_to_real(CS_T3, CS_T7, rout);
}

int RDEnsemble::get_spectrum(int m, int cslot, float2* host) {
  GLIA_HIP(hipMemcpyAsync(host, mh_[m].c[cslot], sizeof(float2) * d_.nspec, hipMemcpyDeviceToHost, stream_));
  GLIA_HIP(hipStreamSynchronize(stream_));
  return 0;
}

int RDEnsemble::set_spectrum(int m, int cslot, const float2* host) {
  if (m < 0 || m >= nm_ || cslot < 0 || cslot >= CS_NUM) return fail(-2, "bad member or slot");
  return copy_sync(mh_[m].c[cslot], host, sizeof(float2) * d_.nspec, hipMemcpyHostToDevice);
}

// b - A x0 for x0 = c: r = tau D c, with the partial norms of B b and B r (KSPConvergedDefault)
int RDEnsemble::matvec_setup(int sx) {
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_xi_grad(c, sx, 0, CS_T0, CS_T1, false));
  GLIA_TRYC(launch_yi_grad(c, CS_T0, CS_T1, CS_T2, CS_T3, CS_T4, -1));
  GLIA_TRYC(launch_z_diff(c, CS_T2, CS_T3, CS_T4, RS_K));
  GLIA_TRYC(launch_yf_div(c, CS_T2, CS_T3, CS_T4, CS_T0, CS_T1));
  GLIA_TRYC(launch_xf_div(c, XFM_SETUP, CS_T0, CS_T1, sx, CS_R));
  return 0;
}

// DiffusionSolver::solve(c, tau) on the spectral state in slot sx, in place.  kind >= 0 names the
// solve's place in a time step (SK_*): its iteration count becomes the budget of the next capture.
int RDEnsemble::cn_solve(int sx, double tau, int kind) {
  ctx_dt_ = tau;  // ctx->dt_ = dt (src/pde/DiffusionSolver.cpp:226)
  LaunchCtx c = ctx();
  const int nbx = xf_div_blocks(d_);
  GLIA_TRYC(launch_pcg_begin(c, tau));
  int rc = matvec_setup(sx);
  if (rc) return rc;
  int stage = 0;
  GLIA_TRYC(launch_pcg_fin_setup(c, nbx, hflag_dev_, stage % kFlagSlots));
  GLIA_HIP(hipEventRecord(ev_[stage % kFlagSlots], stream_));
  int checked = -1;
  bool all_done = false;
  int its_max = 0;
  auto check_stage = [&](int s) -> int {
    GLIA_HIP(hipEventSynchronize(ev_[s % kFlagSlots]));
    volatile int* f = hflag_host_ + (s % kFlagSlots) * nm_;
    bool all = true;
    for (int m = 0; m < nm_; m++) {
      all = all && (f[m] != 0);
      its_max = std::max(its_max, f[m] - 1);
    }
    if (all) all_done = true;
    return 0;
  };
  for (int it = 0; it < cfg_.maxit && !all_done; it++) {
    stage = it + 1;
    if (use_graphs_ && !prof_on_) {
      hipGraphExec_t g = nullptr;
      rc = iteration_graph(sx, stage % kFlagSlots, &g);
      if (rc) return rc;
      GLIA_HIP(hipGraphLaunch(g, stream_));
    } else {
      rc = launch_iteration(sx, stage % kFlagSlots);
      if (rc) return rc;
    }
    GLIA_HIP(hipEventRecord(ev_[stage % kFlagSlots], stream_));
    while (!all_done && checked < stage - kLookahead) {
      checked++;
      rc = check_stage(checked);
      if (rc) return rc;
    }
  }
  if (kind >= 0 && all_done) budget_[kind] = its_max;
  return 0;
}

// The launches of cn_solve with the iteration loop unrolled K times, for capture: CG iterations
// past convergence exit at once, and the last iteration's convergence flag (0 = still running,
// 1 + its otherwise, per member) lands in host flag slot fslot.
int RDEnsemble::capture_solve(int sx, double tau, int K, int fslot) {
  ctx_dt_ = tau;
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_pcg_begin(c, tau));
  int rc = matvec_setup(sx);
  if (rc) return rc;
  GLIA_TRYC(launch_pcg_fin_setup(c, xf_div_blocks(d_), hflag_dev_, K == 0 ? fslot : 0));
  for (int k = 1; k <= K; k++)
    if ((rc = launch_iteration(sx, k == K ? fslot : k % kFlagSlots))) return rc;
  return 0;
}

bool RDEnsemble::step_graphs_on() const { return use_graphs_ && !prof_on_ && hist_all_; }

// capture the launches of 'body' into a graph cached under key (the graph and its executable form)
int RDEnsemble::step_graph(const StepKey& key, const std::function<int()>& body, const StepGraph** out) {
  auto it = step_graphs_.find(key);
  if (it != step_graphs_.end()) {
    *out = &it->second;
    return 0;
  }
  if (step_graphs_.size() >= (size_t)(8 * cfg_.nt + 64)) {  // budgets moved: drop stale captures
    GLIA_HIP(hipStreamSynchronize(stream_));
    drop_step_graphs();
  }
  GLIA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream_, &g);
  if (rc) {
    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
    return rc;
  }
  GLIA_HIP(e);
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) (void)hipGraphDestroy(g);
  GLIA_HIP(ei);
  StepGraph& sg = step_graphs_[key];
  sg.g = g;  // kept: a group evaluation adds it to one graph of several handles' steps
  sg.ex = ex;
  *out = &sg;
  return 0;
}

void RDEnsemble::drop_step_graphs() {
  for (auto& kv : step_graphs_) {
    (void)hipGraphExecDestroy(kv.second.ex);
    (void)hipGraphDestroy(kv.second.g);
  }
  step_graphs_.clear();
  graph_epoch_++;
}

// after a step graph: wait for it and read the flags of its nsolve budgeted solves; ok when every
// member of every solve converged within the budget, its[j] the largest iteration count of solve j
int RDEnsemble::read_step_flags(int nsolve, int* its, bool* ok) {
  GLIA_HIP(hipEventRecord(ev_step_, stream_));
  GLIA_HIP(hipEventSynchronize(ev_step_));
  *ok = true;
  for (int j = 0; j < nsolve; j++) {
    volatile int* f = hflag_host_ + (kFlagSlots + j) * nm_;
    its[j] = 0;
    for (int m = 0; m < nm_; m++) {
      if (f[m] == 0) *ok = false;
      its[j] = std::max(its[j], f[m] - 1);
    }
  }
  return 0;
}

// one forward Strang step (forward_step) from c_i in CS_X, with c_{i+1} stored in the history:
// replayed from a graph captured with iteration budgets one above (margin_) the last solves' counts; when
// a solve needs more, the counters return to the step's start and the step runs again eagerly
// from c_i (stored in the history), so the result is that of forward_step.  fwd_prepare gives the
// step's graph (null when the step runs eagerly: no budgets yet, or graphs off); fwd_finish runs
// after its launch, or in place of it.
int RDEnsemble::fwd_prepare(int i, bool store_half, const StepGraph** out) {
  *out = nullptr;
  if (!(step_graphs_on() && budget_[SK_FWD1] >= 0 && budget_[SK_FWD2] >= 0)) return 0;
  const double dt = cfg_.dt;
  const int K1 = std::max(0, budget_[SK_FWD1] + margin_), K2 = std::max(0, budget_[SK_FWD2] + margin_);
  const StepKey key{0, i, K1, K2, 0, store_half ? 1 : 0};
  return step_graph(key, [&]() -> int {
    LaunchCtx c = ctx();
    GLIA_TRYC(launch_state_reset(c, 4));
    int r = capture_solve(CS_X, dt / 2.0, K1, kFlagSlots + 0);
    if (r) return r;
    GLIA_TRYC(launch_col_plain(c, 0, +1, CS_X, CS_X, false));
    GLIA_TRYC(launch_col_plain(c, 1, +1, CS_X, CS_X, false));
    GLIA_TRYC(launch_z_point(c, ZPM_REACT, CS_X, CS_X, -1, store_half ? i : -1, (float)dt, false));
    GLIA_TRYC(launch_col_plain(c, 1, -1, CS_X, CS_X, false));
    GLIA_TRYC(launch_col_plain(c, 0, -1, CS_X, CS_X, false));
    if ((r = capture_solve(CS_X, dt / 2.0, K2, kFlagSlots + 1))) return r;
    GLIA_TRYC(launch_copy(c, CS_X, state_slot(i + 1)));
    return 0;
  }, out);
}

int RDEnsemble::fwd_finish(int i, bool store_half, bool launched) {
  int rc;
  if (launched) {
    step_launches_++;
    int its[2];
    bool ok = false;
    if ((rc = read_step_flags(2, its, &ok))) return rc;
    if (ok) {
      budget_[SK_FWD1] = its[0];
      budget_[SK_FWD2] = its[1];
      return 0;
    }
    step_reruns_++;
    GLIA_TRYC(launch_state_reset(ctx(), 8));
    GLIA_TRYC(launch_copy(ctx(), state_slot(i), CS_X));
  }
  if ((rc = forward_step(CS_X, i, store_half))) return rc;
  GLIA_TRYC(launch_copy(ctx(), CS_X, state_slot(i + 1)));
  return 0;
}

int RDEnsemble::forward_step_stored(int i, bool store_half) {
  const StepGraph* sg = nullptr;
  int rc = fwd_prepare(i, store_half, &sg);
  if (rc) return rc;
  if (sg) GLIA_HIP(hipGraphLaunch(sg->ex, stream_));
  return fwd_finish(i, store_half, sg != nullptr);
}

// one backward step of the adjoint (adjoint_step at time i, store-all history) replayed from a
// captured graph in the same way; the adjoint state at the step's start is kept in CS_BK
int RDEnsemble::adj_prepare(int i, const StepGraph** out) {
  *out = nullptr;
  const bool half = cfg_.adjoint_store != 0;  // c_half_ stored, no recomputation solve
  if (!(step_graphs_on() && budget_[SK_ADJ1] >= 0 && budget_[SK_ADJ2] >= 0 && (half || budget_[SK_HALF] >= 0)))
    return 0;
  const double dt = cfg_.dt;
  const int K1 = std::max(0, budget_[SK_ADJ1] + margin_), K2 = std::max(0, budget_[SK_ADJ2] + margin_);
  const int Kh = half ? 0 : std::max(0, budget_[SK_HALF] + margin_);
  const int flags = (half ? 1 : 0) | (quad_on_ ? 2 : 0) | (keep_adj_ ? 4 : 0);
  const StepKey key{1, i, K1, K2, Kh, flags};
  return step_graph(key, [&]() -> int {
    LaunchCtx c = ctx();
    GLIA_TRYC(launch_state_reset(c, 4));
    GLIA_TRYC(launch_copy(c, CS_A, CS_BK));
    int r = capture_solve(CS_A, dt / 2.0, K1, kFlagSlots + 0);
    if (r) return r;
    if (!half) {
      GLIA_TRYC(launch_copy(c, state_slot(i), CS_H));
      if ((r = capture_solve(CS_H, dt / 2.0, Kh, kFlagSlots + 2))) return r;
      if ((r = inverse_to_real(CS_H, CS_H, RS_CH))) return r;
    }
    GLIA_TRYC(launch_col_plain(c, 0, +1, CS_A, CS_A, false));
    GLIA_TRYC(launch_col_plain(c, 1, +1, CS_A, CS_A, false));
    GLIA_TRYC(launch_z_point(c, ZPM_REACT_ADJ, CS_A, CS_A, half ? -1 : RS_CH, half ? i : -1, (float)dt, false));
    GLIA_TRYC(launch_col_plain(c, 1, -1, CS_A, CS_A, false));
    GLIA_TRYC(launch_col_plain(c, 0, -1, CS_A, CS_A, false));
    if ((r = capture_solve(CS_A, dt / 2.0, K2, kFlagSlots + 1))) return r;
    if ((r = store_adjoint(i))) return r;
    return quad_on_ ? quad(state_slot(i), CS_A, (i == 0) ? 0.5 : 1.0) : 0;
  }, out);
}

int RDEnsemble::adj_finish(int i, bool launched) {
  if (launched) {
    step_launches_++;
    const bool half = cfg_.adjoint_store != 0;
    int its[3] = {0, 0, 0};
    bool ok = false;
    int rc = read_step_flags(half ? 2 : 3, its, &ok);
    if (rc) return rc;
    if (ok) {
      budget_[SK_ADJ1] = its[0];
      budget_[SK_ADJ2] = its[1];
      if (!half) budget_[SK_HALF] = its[2];
      return 0;
    }
    step_reruns_++;
    GLIA_TRYC(launch_state_reset(ctx(), 8));
    GLIA_TRYC(launch_copy(ctx(), CS_BK, CS_A));
  }
  return adjoint_step(i, state_slot(i));
}

int RDEnsemble::adjoint_step_stored(int i) {
  const StepGraph* sg = nullptr;
  int rc = adj_prepare(i, &sg);
  if (rc) return rc;
  if (sg) GLIA_HIP(hipGraphLaunch(sg->ex, stream_));
  return adj_finish(i, sg != nullptr);
}

// One CG iteration: matvec w = A p (five transform kernels), alpha, x/r update, the PETSc
// convergence test with its flag written to host slot 'slot'.  Every argument is fixed for a
// handle except the solved slot sx and the flag slot, so the sequence can be captured once per
// (sx, slot) and replayed.
int RDEnsemble::launch_iteration(int sx, int slot) {
  LaunchCtx c = ctx();
  const int nbx = xf_div_blocks(d_);
  prof_mark(PK_XI_PUPD, true);
  GLIA_TRYC(launch_xi_grad(c, CS_P, CS_R, CS_T0, CS_T1, true));
  prof_mark(PK_XI_PUPD, false);
  prof_mark(PK_YI_GRAD, true);
  GLIA_TRYC(launch_yi_grad(c, CS_T0, CS_T1, CS_T2, CS_T3, CS_T4, -1));
  prof_mark(PK_YI_GRAD, false);
  prof_mark(PK_Z_DIFF, true);
  GLIA_TRYC(launch_z_diff(c, CS_T2, CS_T3, CS_T4, RS_K));
  prof_mark(PK_Z_DIFF, false);
  prof_mark(PK_YF_DIV, true);
  GLIA_TRYC(launch_yf_div(c, CS_T2, CS_T3, CS_T4, CS_T0, CS_T1));
  prof_mark(PK_YF_DIV, false);
  prof_mark(PK_XF_MATVEC, true);
  GLIA_TRYC(launch_xf_div(c, XFM_MATVEC, CS_T0, CS_T1, CS_P, CS_W));
  prof_mark(PK_XF_MATVEC, false);
  GLIA_TRYC(launch_pcg_fin_dpi(c, nbx));
  prof_mark(PK_PCG_UPDATE, true);
  GLIA_TRYC(launch_pcg_update(c, sx));
  prof_mark(PK_PCG_UPDATE, false);
  GLIA_TRYC(launch_pcg_fin_update(c, pcg_update_blocks(), hflag_dev_, slot));
  return 0;
}

int RDEnsemble::iteration_graph(int sx, int slot, hipGraphExec_t* out) {
  const auto key = std::make_pair(sx, slot);
  auto it = graphs_.find(key);
  if (it != graphs_.end()) {
    *out = it->second;
    return 0;
  }
  GLIA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  const int rc = launch_iteration(sx, slot);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream_, &g);
  if (rc) {
    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
    return rc;
  }
  GLIA_HIP(e);
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  GLIA_HIP(ei);
  graphs_[key] = ex;
  *out = ex;
  return 0;
}

int RDEnsemble::set_use_graphs(bool on) {
  use_graphs_ = on;
  return 0;
}

// one Strang step (src/pde/PdeOperators.cpp:264-300): D_{dt/2}, R_dt, D_{dt/2}
int RDEnsemble::forward_step(int sx, int step, bool store_half) {
  const double dt = cfg_.dt;
  int rc = cn_solve(sx, dt / 2.0, sx == CS_X ? SK_FWD1 : -1);
  if (rc) return rc;
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_col_plain(c, 0, +1, sx, sx, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, sx, sx, false));
  GLIA_TRYC(launch_z_point(c, ZPM_REACT, sx, sx, -1, store_half ? step : -1, (float)dt, false));
  GLIA_TRYC(launch_col_plain(c, 1, -1, sx, sx, false));
  GLIA_TRYC(launch_col_plain(c, 0, -1, sx, sx, false));
  return cn_solve(sx, dt / 2.0, sx == CS_X ? SK_FWD2 : -1);
}

// the start of solveState(0): c(0) from RS_C0 into CS_X, and into the history at index 0
int RDEnsemble::state_begin(bool keep_history) {
  int rc = forward_transform_real(RS_C0, CS_X);
  if (rc) return rc;
  stored_.clear();
  free_pool_.clear();
  if (keep_history) {
    GLIA_TRYC(launch_copy(ctx(), CS_X, state_slot(0)));
    stored_[0] = 0;
  }
  return 0;
}

// PdeOperatorsRD::solveState (src/pde/PdeOperators.cpp:235-316) from c0 in RS_C0
int RDEnsemble::solve_state(int linearized, bool keep_history) {
  if (linearized == 1 || linearized == 2) return solve_tangent(linearized == 2);
  if (linearized != 0) return fail(-2, "solveState: linearized must be 0, 1 or 2");
  int rc = state_begin(keep_history);
  if (rc) return rc;
  std::vector<int> spine;  // time indices stored during the first sweep (checkpoint mode)
  if (keep_history) {
    if (!hist_all_) {
      for (int i = npool_ - 1; i >= 1; i--) free_pool_.push_back(i);
      // the right spine of the recursive schedule: positions the reversal will checkpoint first
      int a = 0, s = npool_ - 1;
      while (cfg_.nt - a > 1 && s > 0) {
        // same rule as reverse(): split point from the cost recursion
        int l = cfg_.nt - a;
        // cost table is small: recompute the best first step k for (l, s)
        std::vector<std::vector<double>> T(l + 1, std::vector<double>(s + 1, 0.0));
        std::vector<std::vector<int>> K(l + 1, std::vector<int>(s + 1, 1));
        for (int ll = 1; ll <= l; ll++) T[ll][0] = 0.5 * ll * (ll - 1);
        for (int ss = 1; ss <= s; ss++)
          for (int ll = 2; ll <= l; ll++) {
            double best = 1e300;
            int bk = 1;
            for (int k = 1; k < ll; k++) {
              const double v = k + T[ll - k][ss - 1] + T[k][ss];
              if (v < best) { best = v; bk = k; }
            }
            T[ll][ss] = best;
            K[ll][ss] = bk;
          }
        const int mpos = a + K[l][s];
        spine.push_back(mpos);
        a = mpos;
        s--;
      }
    }
  }
  size_t sp = 0;
  for (int i = 0; i < cfg_.nt; i++) {
    const bool store_half = keep_history && hist_all_ && cfg_.adjoint_store;
    if (keep_history && hist_all_) {
      if ((rc = forward_step_stored(i, store_half))) return rc;
      stored_[i + 1] = i + 1;
      continue;
    }
    rc = forward_step(CS_X, i, store_half);
    if (rc) return rc;
    if (keep_history) {
      if (hist_all_) {
        GLIA_TRYC(launch_copy(ctx(), CS_X, state_slot(i + 1)));
        stored_[i + 1] = i + 1;
      } else if (sp < spine.size() && spine[sp] == i + 1) {
        const int pidx = free_pool_.back();
        free_pool_.pop_back();
        GLIA_TRYC(launch_copy(ctx(), CS_X, state_slot(pidx)));
        stored_[i + 1] = pidx;
        sp++;
      }
    }
  }
  return inverse_to_real(CS_X, CS_T5, RS_CT);
}

// PdeOperatorsRD::solveState(1) / solveState(2) (src/pde/PdeOperators.cpp:235-316): the tangent-linear
// forward from c~(0) in RS_C0 around the state history of the last solveState(0) (reaction(1, i)
// linearized at c_[i]); with 2 also the solveIncremental sources of the secondary coefficients
// (set_secondary).  c~(T) lands in RS_CT; the state history is left as it is.
int RDEnsemble::solve_tangent(bool sources) {
  if (stored_.find(0) == stored_.end()) return fail(-6, "solveState(1|2) linearizes around the history of solveState(0)");
  int rc = forward_transform_real(RS_C0, CS_A);
  if (rc) return rc;
  int ct = CS_X;
  if ((rc = tangent_forward(CS_A, sources, &ct))) return rc;
  return inverse_to_real(CS_A, CS_T5, RS_CT);
}

// alpha(T) = -O^T (O c(T) - d1) and m1 = ||O c(T) - d1||^2 (src/grad/DerivativeOperatorsRD.cpp:154-162)
int RDEnsemble::misfit_and_final_adjoint(int cslot_c, int out) {
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_col_plain(c, 0, +1, cslot_c, CS_T5, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_T5, CS_T5, false));
  GLIA_TRYC(launch_z_point(c, ZPM_MISFIT, CS_T5, CS_T5, -1, -1, 0.f, false));
  GLIA_TRYC(launch_col_plain(c, 1, -1, CS_T5, CS_T5, false));
  GLIA_TRYC(launch_col_plain(c, 0, -1, CS_T5, out, false));
  GLIA_TRYC(launch_fin_state(c, z_blocks(d_), FINK_MISFIT, 0.0));
  return 0;
}

// trapezoid term w * (<m_k, grad c . grad alpha>, <m_rho, (c^2 - c) alpha>) at one time point
int RDEnsemble::quad(int cslot_c, int cslot_a, double w) {
  LaunchCtx c = ctx();
  int rc = clear_done();
  if (rc) return rc;
  GLIA_TRYC(launch_xi_grad(c, cslot_c, 0, CS_T0, CS_T1, false));
  GLIA_TRYC(launch_yi_grad(c, CS_T0, CS_T1, CS_T2, CS_T3, CS_T4, CS_T5));
  GLIA_TRYC(launch_xi_grad(c, cslot_a, 0, CS_T0, CS_T1, false));
  GLIA_TRYC(launch_yi_grad(c, CS_T0, CS_T1, CS_T6, CS_T7, CS_R, CS_P));
  GLIA_TRYC(launch_z_gradq(c, make_int4(CS_T5, CS_T2, CS_T3, CS_T4), make_int4(CS_P, CS_T6, CS_T7, CS_R)));
  GLIA_TRYC(launch_fin_state(c, z_blocks(d_), FINK_QUAD, w));
  return 0;
}

// one backward step of PdeOperatorsRD::solveAdjoint (src/pde/PdeOperators.cpp:372-420) with
// reactionAdjoint (:318-370), then the trapezoid term at time i
int RDEnsemble::adjoint_step(int i, int cslot_c) {
  const double dt = cfg_.dt;
  int rc = cn_solve(CS_A, dt / 2.0, SK_ADJ1);
  if (rc) return rc;
  LaunchCtx c = ctx();
  int ch = -1;
  if (hist_all_ && cfg_.adjoint_store) {
    ch = i;
  } else {
    GLIA_TRYC(launch_copy(c, cslot_c, CS_H));
    rc = cn_solve(CS_H, dt / 2.0, SK_HALF);
    if (rc) return rc;
    rc = inverse_to_real(CS_H, CS_H, RS_CH);
    if (rc) return rc;
  }
  GLIA_TRYC(launch_col_plain(c, 0, +1, CS_A, CS_A, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_A, CS_A, false));
  GLIA_TRYC(launch_z_point(c, ZPM_REACT_ADJ, CS_A, CS_A, ch >= 0 ? -1 : RS_CH, ch, (float)dt, false));
  GLIA_TRYC(launch_col_plain(c, 1, -1, CS_A, CS_A, false));
  GLIA_TRYC(launch_col_plain(c, 0, -1, CS_A, CS_A, false));
  rc = cn_solve(CS_A, dt / 2.0, SK_ADJ2);
  if (rc) return rc;
  if ((rc = store_adjoint(i))) return rc;
  return quad_on_ ? quad(cslot_c, CS_A, (i == 0) ? 0.5 : 1.0) : 0;
}

// PdeOperatorsRD::solveState(1) step (src/pde/PdeOperators.cpp:264-300 with reaction(1, i),
// :140-190): D_{dt/2}; c~ *= e^{rho dt} / (c_i e^{rho dt} + 1 - c_i)^2 with c_i the state at the
// start of the step (not the half-step state the adjoint linearizes around); D_{dt/2}
int RDEnsemble::tangent_step(int sx, int cslot_c) {
  const double dt = cfg_.dt;
  int rc = cn_solve(sx, dt / 2.0);
  if (rc) return rc;
  rc = inverse_to_real(cslot_c, CS_H, RS_CH);
  if (rc) return rc;
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_col_plain(c, 0, +1, sx, sx, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, sx, sx, false));
  GLIA_TRYC(launch_z_point(c, ZPM_REACT_ADJ, sx, sx, RS_CH, -1, (float)dt, false));
  GLIA_TRYC(launch_col_plain(c, 1, -1, sx, sx, false));
  GLIA_TRYC(launch_col_plain(c, 0, -1, sx, sx, false));
  return cn_solve(sx, dt / 2.0);
}

// out = FFT(div(k_sec grad IFFT(in))) on packed spectra: DiffCoef::applyDWithSecondaryCoeffs
// (src/mat/DiffCoef.cpp:273-300) with the secondary coefficient field RS_KS
int RDEnsemble::apply_dsec(int in, int out) {
  int rc = clear_done();
  if (rc) return rc;
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_xi_grad(c, in, 0, CS_T0, CS_T1, false));
  GLIA_TRYC(launch_yi_grad(c, CS_T0, CS_T1, CS_T2, CS_T3, CS_T4, -1));
  GLIA_TRYC(launch_z_diff(c, CS_T2, CS_T3, CS_T4, RS_KS));
  GLIA_TRYC(launch_yf_div(c, CS_T2, CS_T3, CS_T4, CS_T0, CS_T1));
  GLIA_TRYC(launch_xf_div(c, XFM_D, CS_T0, CS_T1, in, out));
  return 0;
}

// PdeOperatorsRD::solveIncremental (src/pde/PdeOperators.cpp:192-233), called with dt/2:
// c~ += dt/4 D_sec(wa c_i + wb c_{i+1}), (wa, wb) = (1.5, 0.5) before the step, (0.5, 1.5) after
int RDEnsemble::incremental_source(int sx, int ci, int ci1, double wa, double wb) {
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_axpby_c(c, ci, wa, ci1, wb, CS_W));
  int rc = apply_dsec(CS_W, CS_P);
  if (rc) return rc;
  GLIA_TRYC(launch_axpby_c(c, sx, 1.0, CS_P, cfg_.dt / 4.0, sx));
  return 0;
}

// tangent-linear forward in slot sx (solveState(1), or solveState(2) with the secondary sources);
// under checkpointing the base states c_i, c_{i+1} come from a lockstep forward sweep in CS_X and
// CS_T6.  Leaves c(T) in slot *ct_slot.
int RDEnsemble::tangent_forward(int sx, bool sources, int* ct_slot) {
  int rc = 0;
  if (hist_all_) {
    for (int i = 0; i < cfg_.nt && rc == 0; i++) {
      if (sources && (rc = incremental_source(sx, state_slot(i), state_slot(i + 1), 1.5, 0.5))) return rc;
      if ((rc = tangent_step(sx, state_slot(i)))) return rc;
      if (sources) rc = incremental_source(sx, state_slot(i), state_slot(i + 1), 0.5, 1.5);
    }
    *ct_slot = state_slot(cfg_.nt);
    return rc;
  }
  GLIA_TRYC(launch_copy(ctx(), state_slot(stored_[0]), CS_X));
  for (int i = 0; i < cfg_.nt; i++) {
    GLIA_TRYC(launch_copy(ctx(), CS_X, CS_T6));
    if ((rc = forward_step(CS_X, i, false))) return rc;
    if (sources && (rc = incremental_source(sx, CS_T6, CS_X, 1.5, 0.5))) return rc;
    if ((rc = tangent_step(sx, CS_T6))) return rc;
    if (sources && (rc = incremental_source(sx, CS_T6, CS_X, 0.5, 1.5))) return rc;
  }
  *ct_slot = CS_X;
  return 0;
}

// the incremental adjoint after a tangent forward: a~(T) = -O^T O c~(T), optionally the kappa
// quadratures, then the reverse sweep (solveAdjoint(2)); leaves a~0 in RS_A0.  With quadratures,
// the i = nt term uses the adjoint history entry p_[nt] that solveAdjoint(2) leaves at the last
// gradient's a(T) = -O^T (O c(T) - d1) (src/pde/PdeOperators.cpp:384-386), as the reference does.
int RDEnsemble::incremental_adjoint(int ct_slot, bool quads) {
  LaunchCtx c = ctx();
  int rc;
  if (quads) {
    if ((rc = misfit_and_final_adjoint(ct_slot, CS_W))) return rc;
    if ((rc = quad(ct_slot, CS_W, 0.5))) return rc;
  }
  GLIA_TRYC(launch_col_plain(c, 0, +1, CS_A, CS_T5, false));
  GLIA_TRYC(launch_col_plain(c, 1, +1, CS_T5, CS_T5, false));
  GLIA_TRYC(launch_z_point(c, ZPM_OBS2, CS_T5, CS_T5, -1, -1, 0.f, false));
  GLIA_TRYC(launch_col_plain(c, 1, -1, CS_T5, CS_T5, false));
  GLIA_TRYC(launch_col_plain(c, 0, -1, CS_T5, CS_A, false));
  quad_on_ = quads;
  rc = 0;
  if (hist_all_) {
    for (int i = cfg_.nt - 1; i >= 0 && rc == 0; i--) rc = adjoint_step_stored(i);
  } else {
    rc = reverse(0, cfg_.nt, stored_[0], npool_ - 1);
  }
  quad_on_ = true;
  if (rc) return rc;
  return inverse_to_real(CS_A, CS_T5, RS_A0);
}

// DerivativeOperatorsRD::evaluateHessian, L2 regularization, nk = 1
// (src/grad/DerivativeOperatorsRD.cpp:229-438).  xt = [p~ | k~ | rho~] per member; the linearization
// point is the state history of the last evaluation.
//   without diffusivity: y_p = beta h^3 Phi^T Phi p~ - h^3 Phi^T a~0 (solveState(1), solveAdjoint(2))
//   with diffusivity:    y = [Hpp p~ + Hpk k~ ; Hkp p~ + Hkk k~ ; 0]; the kappa row is the
//   gradDiffusion quadrature after each incremental adjoint, the k~ part is solveState(2) from
//   c~0 = 0 with the secondary coefficient k~ (wm + k_gm gm + k_glm csf).
// Under checkpointing the base states are recomputed in lockstep with the tangent states.
int RDEnsemble::hessian(const std::vector<std::vector<double>>& xt, bool diffusivity,
                        std::vector<std::vector<double>>& y) {
  if (stored_.find(0) == stored_.end()) return fail(-6, "the Hessian needs the state history of an evaluation first");
  if ((int)xt.size() != nm_) return fail(-2, "one direction per member");
  for (int m = 0; m < nm_; m++)
    if ((int)xt[m].size() != np_[m] + 2) return fail(-2, "direction must have np + 2 entries");
  std::vector<std::vector<double>> pt(nm_);
  for (int m = 0; m < nm_; m++) pt[m].assign(xt[m].begin(), xt[m].begin() + np_[m]);
  LaunchCtx c = ctx();
  const double h3 = lebesgue();
  int rc, ct_slot = CS_X;
  // Hpp p~ (and Hkp p~)
  if (diffusivity) GLIA_TRYC(launch_state_reset(c, 2));
  if ((rc = phi_apply(pt))) return rc;
  if ((rc = forward_transform_real(RS_C0, CS_A))) return rc;
  if ((rc = tangent_forward(CS_A, false, &ct_slot))) return rc;
  if ((rc = incremental_adjoint(ct_slot, diffusivity))) return rc;
  std::vector<std::vector<double>> t1, t2;
  if ((rc = phi_apply_transpose(RS_C0, t1))) return rc;
  if ((rc = phi_apply_transpose(RS_A0, t2))) return rc;
  y.assign(nm_, {});
  for (int m = 0; m < nm_; m++) {
    y[m].assign(np_[m] + 2, 0.0);
    for (int i = 0; i < np_[m]; i++) y[m][i] = cfg_.beta * h3 * t1[m][i] - h3 * t2[m][i];
  }
  if (!diffusivity) return 0;
  auto add_kappa_row = [&]() -> int {
    for (int m = 0; m < nm_; m++) {
      PcgState s;
      GLIA_TRYC(copy_sync(&s, st_[m], sizeof(s), hipMemcpyDeviceToHost));
      y[m][np_[m]] += h3 * cfg_.dt * s.gk;
    }
    return 0;
  };
  if ((rc = add_kappa_row())) return rc;
  // Hpk k~ and Hkk k~
  GLIA_TRYC(launch_state_reset(c, 2));
  for (int m = 0; m < nm_; m++) {
    const double kt = xt[m][np_[m]];
    LaunchCtx cm = member_ctx(c, md_, m);
    GLIA_TRYC(launch_axpby_real(cm, RS_WM, kt, RS_GM, kt * cfg_.k_gm_wm, RS_KS));
    GLIA_TRYC(launch_axpby_real(cm, RS_KS, 1.0, RS_CSF, kt * cfg_.k_glm_wm, RS_KS));
  }
  GLIA_TRYC(launch_axpby_c(c, CS_A, 0.0, CS_A, 0.0, CS_A));  // c~0 = 0
  if ((rc = tangent_forward(CS_A, true, &ct_slot))) return rc;
  if ((rc = incremental_adjoint(ct_slot, true))) return rc;
  if ((rc = phi_apply_transpose(RS_A0, t2))) return rc;
  for (int m = 0; m < nm_; m++)
    for (int i = 0; i < np_[m]; i++) y[m][i] += -h3 * t2[m][i];
  return add_kappa_row();
}

// reverse the steps b-1 ... a with c_a in pool entry pa and free_slots further entries free
int RDEnsemble::reverse(int a, int b, int pa, int free_slots) {
  if (b - a <= 0) return 0;
  if (b - a == 1) return adjoint_step(a, state_slot(pa));
  if (free_slots <= 0) {
    for (int j = b - 1; j >= a; j--) {
      GLIA_TRYC(launch_copy(ctx(), state_slot(pa), CS_X));
      for (int t = a; t < j; t++) {
        int rc = forward_step(CS_X, t, false);
        if (rc) return rc;
      }
      int rc = adjoint_step(j, CS_X);
      if (rc) return rc;
    }
    return 0;
  }
  // split point from the recomputation-cost recursion T(l,s) = min_k k + T(l-k,s-1) + T(k,s)
  const int l = b - a, s = free_slots;
  std::vector<std::vector<double>> T(l + 1, std::vector<double>(s + 1, 0.0));
  std::vector<std::vector<int>> K(l + 1, std::vector<int>(s + 1, 1));
  for (int ll = 1; ll <= l; ll++) T[ll][0] = 0.5 * ll * (ll - 1);
  for (int ss = 1; ss <= s; ss++)
    for (int ll = 2; ll <= l; ll++) {
      double best = 1e300;
      int bk = 1;
      for (int k = 1; k < ll; k++) {
        const double v = k + T[ll - k][ss - 1] + T[k][ss];
        if (v < best) { best = v; bk = k; }
      }
      T[ll][ss] = best;
      K[ll][ss] = bk;
    }
  const int mpos = a + K[l][s];
  int pm;
  auto it = stored_.find(mpos);
  if (it != stored_.end()) {
    pm = it->second;
  } else {
    if (free_pool_.empty()) return fail(-5, "checkpoint pool exhausted");
    pm = free_pool_.back();
    free_pool_.pop_back();
    GLIA_TRYC(launch_copy(ctx(), state_slot(pa), CS_X));
    for (int t = a; t < mpos; t++) {
      int rc = forward_step(CS_X, t, false);
      if (rc) return rc;
    }
    GLIA_TRYC(launch_copy(ctx(), CS_X, state_slot(pm)));
    stored_[mpos] = pm;
  }
  int rc = reverse(mpos, b, pm, free_slots - 1);
  if (rc) return rc;
  stored_.erase(mpos);
  free_pool_.push_back(pm);
  return reverse(a, mpos, pa, free_slots);
}

// PdeOperatorsRD::solveAdjoint (src/pde/PdeOperators.cpp:372-420) with the gradient quadratures of
// gradDiffusion / gradReaction accumulated along the sweep.
//   linearized 0 or 1: from alpha(T) = -O^T (O c(T) - d1) of the current data and the state history
//   linearized 2: the incremental adjoint from a~(T) = -O^T O c~(T), c~(T) in RS_CT (after
//     solveState(1|2)); its i = nt quadrature term uses the adjoint final condition of the last
//     gradient, as the reference's p_[nt] (which solveAdjoint(2) does not overwrite) does
// alpha(0) lands in RS_A0; with the adjoint history kept, p_[i] is stored along the way.
int RDEnsemble::solve_adjoint(int linearized) {
  if (linearized < 0 || linearized > 2) return fail(-2, "solveAdjoint: linearized must be 0, 1 or 2");
  if (stored_.find(0) == stored_.end()) return fail(-6, "solveAdjoint needs solveState with history first");
  GLIA_TRYC(launch_state_reset(ctx(), 2));
  int rc;
  if (linearized == 2) {
    int base = CS_X;  // c(T) of the state history
    if (hist_all_) {
      base = state_slot(cfg_.nt);
    } else {
      GLIA_TRYC(launch_copy(ctx(), state_slot(stored_[0]), CS_X));
      for (int t = 0; t < cfg_.nt; t++)
        if ((rc = forward_step(CS_X, t, false))) return rc;
    }
    if ((rc = forward_transform_real(RS_CT, CS_A))) return rc;
    if ((rc = incremental_adjoint(base, true))) return rc;
    adj_valid_ = keep_adj_;
    return 0;
  }
  if ((rc = adjoint_begin())) return rc;
  if (hist_all_) {
    for (int i = cfg_.nt - 1; i >= 0; i--) {
      rc = adjoint_step_stored(i);
      if (rc) return rc;
    }
  } else {
    // the spine checkpoints of the first sweep are the allocations the recursion makes first
    rc = reverse(0, cfg_.nt, stored_[0], npool_ - 1);
    if (rc) return rc;
  }
  return adjoint_end();
}

// the start of solveAdjoint(0|1) from c(T) in CS_X: alpha(T), the misfit and the quadrature at T
int RDEnsemble::adjoint_begin() {
  int rc = misfit_and_final_adjoint(CS_X);
  if (rc) return rc;
  if ((rc = store_adjoint(cfg_.nt))) return rc;
  return quad(CS_X, CS_A, 0.5);
}

// the end of solveAdjoint(0|1): alpha(0) as a real field in RS_A0
int RDEnsemble::adjoint_end() {
  adj_valid_ = keep_adj_;
  return inverse_to_real(CS_A, CS_T5, RS_A0);
}

// ---------------------------------------------------------------------------------------------
// Autotuning of the kernel variants
// ---------------------------------------------------------------------------------------------
// The variants of the two forward divergence passes are timed on this handle's own grid and
// members with HIP events: two warm-up launches, then the median of five.  X pass: one thread group
// holding both fields (0) or one group per field (1).  Y pass: one group holding three fields (0),
// one group per field (1), or the fields split over blocks at 3 (2) or 2 (3) waves per SIMD.  The
// choice is kept in a text cache (one line per device, grid and member count: "<arch> n0 n1 n2 nm
// xf yf xf_ms0 xf_ms1 yf_ms0 yf_ms1 yf_ms2 yf_ms3") that later handles read instead of timing again.
int RDEnsemble::autotune(const char* cache_path, int* choice) {
  hipDeviceProp_t prop;
  int dev = 0;
  GLIA_HIP(hipGetDevice(&dev));
  GLIA_HIP(hipGetDeviceProperties(&prop, dev));
  std::string arch(prop.gcnArchName);
  for (char& ch : arch)
    if (ch == ' ' || ch == ':') ch = '_';
  std::ostringstream key;
  key << arch << ' ' << d_.n0 << ' ' << d_.n1 << ' ' << d_.n2 << ' ' << nm_;
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      if (line.compare(0, key.str().size() + 1, key.str() + " ") != 0) continue;
      std::istringstream ls(line.substr(key.str().size() + 1));
      int xf = -1, yf = -1;
      if (ls >> xf >> yf && xf >= 0 && xf < kXfVariants && yf >= 0 && yf < kYfVariants) {
        xf_split_ = xf;
        yf_split_ = yf;
        if (choice) { choice[0] = xf; choice[1] = yf; }
        return 0;
      }
    }
  }
  int rc = clear_done();  // every member's kernels run during the timing
  if (rc) return rc;
  hipEvent_t a, b;
  GLIA_HIP(hipEventCreate(&a));
  GLIA_HIP(hipEventCreate(&b));
  auto time_it = [&](const std::function<int()>& launch, double* ms) -> int {
    for (int w = 0; w < 2; w++)
      if (int r = launch()) return r;
    std::vector<float> t;
    for (int k = 0; k < 5; k++) {
      GLIA_HIP(hipEventRecord(a, stream_));
      if (int r = launch()) return r;
      GLIA_HIP(hipEventRecord(b, stream_));
      GLIA_HIP(hipEventSynchronize(b));
      float x = 0;
      GLIA_HIP(hipEventElapsedTime(&x, a, b));
      t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    *ms = t[2];
    return 0;
  };
  double xf_ms[kXfVariants], yf_ms[kYfVariants];
  for (int v = 0; v < kXfVariants && !rc; v++) {
    xf_split_ = v;
    rc = time_it([&]() { return launch_xf_div(ctx(), XFM_MATVEC, CS_T0, CS_T1, CS_P, CS_W); }, &xf_ms[v]);
  }
  for (int v = 0; v < kYfVariants && !rc; v++) {
    yf_split_ = v;
    rc = time_it([&]() { return launch_yf_div(ctx(), CS_T2, CS_T3, CS_T4, CS_T0, CS_T1); }, &yf_ms[v]);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (!rc) GLIA_TRYC(launch_set_done(ctx(), 1));  // no solve in progress
  if (!rc) rc = sync();
  if (rc) return rc;
  xf_split_ = (int)(std::min_element(xf_ms, xf_ms + kXfVariants) - xf_ms);
  yf_split_ = (int)(std::min_element(yf_ms, yf_ms + kYfVariants) - yf_ms);
  if (choice) { choice[0] = xf_split_; choice[1] = yf_split_; }
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    out << key.str() << ' ' << xf_split_ << ' ' << yf_split_;
    for (int v = 0; v < kXfVariants; v++) out << ' ' << xf_ms[v];
    for (int v = 0; v < kYfVariants; v++) out << ' ' << yf_ms[v];
    out << '\n';
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Phi apply / transpose, derivative operators
// ---------------------------------------------------------------------------------------------
int RDEnsemble::phi_apply(const std::vector<std::vector<double>>& p) {
  for (int m = 0; m < nm_; m++) {
    const int np = np_[m];
    std::vector<double> sc(std::max(np, 1), 0.0);
    double pmax = 0;
    bool nnz = false;
    double all = 0;
    for (int i = 0; i < np; i++) all = std::max(all, phimax_[m][i]);
    for (int i = 0; i < np; i++)
      if (p[m][i] != 0) {
        nnz = true;
        pmax = std::max(pmax, phimax_[m][i]);
      }
    // phi_store_: all bases normalized by the global max; on the fly: by the max over active bases
    double norm = cfg_.phi_store ? all : (nnz ? pmax : 1.0);
    if (norm == 0) norm = 1.0;
    for (int i = 0; i < np; i++) sc[i] = p[m][i] / norm;
    if (np > 0) GLIA_TRYC(copy_sync(phi_scale_dev_[m], sc.data(), sizeof(double) * np, hipMemcpyHostToDevice));
  }
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_phi_apply(c));
  GLIA_TRYC(launch_fin_state(c, ew_blocks(), FINK_C0SQ, 0.0));
  return 0;
}

int RDEnsemble::phi_apply_transpose(int rslot, std::vector<std::vector<double>>& out) {
  out.assign(nm_, {});
  for (int m = 0; m < nm_; m++) {
    const int np = np_[m];
    double all = 0;
    for (int i = 0; i < np; i++) all = std::max(all, phimax_[m][i]);
    if (all == 0) all = 1.0;
    GLIA_TRYC(launch_phi_applyT(member_ctx(ctx(), md_, m), np, rslot, 1.0 / all));
  }
  GLIA_HIP(hipStreamSynchronize(stream_));
  for (int m = 0; m < nm_; m++) {
    out[m].assign(np_[m], 0.0);
    if (np_[m] > 0) GLIA_TRYC(copy_sync(out[m].data(), phit_dev_[m], sizeof(double) * np_[m], hipMemcpyDeviceToHost));
  }
  return 0;
}

int RDEnsemble::diffusion_solve_field(int rslot, double tau) {
  int rc = forward_transform_real(rslot, CS_X);
  if (rc) return rc;
  rc = cn_solve(CS_X, tau);
  if (rc) return rc;
  rc = inverse_to_real(CS_X, CS_T5, rslot);
  if (rc) return rc;
  return sync();
}

int RDEnsemble::apply_d_field(int rin, int rout) {
  int rc = forward_transform_real(rin, CS_X);
  if (rc) return rc;
  rc = clear_done();
  if (rc) return rc;
  LaunchCtx c = ctx();
  GLIA_TRYC(launch_xi_grad(c, CS_X, 0, CS_T0, CS_T1, false));
  GLIA_TRYC(launch_yi_grad(c, CS_T0, CS_T1, CS_T2, CS_T3, CS_T4, -1));
  GLIA_TRYC(launch_z_diff(c, CS_T2, CS_T3, CS_T4, RS_K));
  GLIA_TRYC(launch_yf_div(c, CS_T2, CS_T3, CS_T4, CS_T0, CS_T1));
  GLIA_TRYC(launch_xf_div(c, XFM_D, CS_T0, CS_T1, CS_X, CS_T5));
  rc = inverse_to_real(CS_T5, CS_T6, rout);
  if (rc) return rc;
  return sync();
}

int RDEnsemble::reset_counters() {
  GLIA_TRYC(launch_state_reset(ctx(), 1));
  return 0;
}

// DerivativeOperatorsRD::evaluateObjective / evaluateObjectiveAndGradient
//   mode 0: p and kappa (flag_reaction_inv_ = false); mode 1: kappa and rho (c0 fixed);
//   mode 2: p, kappa and rho together.  x[m] = [p (np) | kappa | rho].
int RDEnsemble::objective_and_gradient(int mode, const std::vector<std::vector<double>>& x, bool want_grad,
                                       std::vector<double>& J, std::vector<std::vector<double>>& dJ) {
  int rc = eval_setup(mode, x);
  if (rc) return rc;
  if ((rc = solve_state(0, want_grad))) return rc;
  if (want_grad) {
    if ((rc = solve_adjoint(1))) return rc;
  } else {
    if ((rc = misfit_and_final_adjoint(CS_X))) return rc;
  }
  return eval_result(mode, want_grad, J, dJ);
}

// the start of an evaluation: updateReactionAndDiffusion (src/grad/DerivativeOperators.cpp:323-367,
// nk = nr = 1), the preconditioner, c(0) = Phi p (or ||c(0)||^2 for mode 1)
int RDEnsemble::eval_setup(int mode, const std::vector<std::vector<double>>& x) {
  if (mode < 0 || mode > 2) return fail(-2, "mode must be 0, 1 or 2");
  if ((int)x.size() != nm_) return fail(-2, "x must hold one vector per member");
  for (int m = 0; m < nm_; m++)
    if ((int)x[m].size() != np_[m] + 2) return fail(-2, "x[m] must have np + 2 entries");
  int rc;
  for (int m = 0; m < nm_; m++) {
    if ((rc = set_kappa(m, x[m][np_[m]] * cfg_.k_scale))) return rc;
    if (mode != 0 && (rc = set_rho(m, x[m][np_[m] + 1] * cfg_.rho_scale))) return rc;
  }
  if ((rc = prec_factor())) return rc;
  LaunchCtx c = ctx();
  if (mode != 1) {
    if ((rc = phi_apply(x))) return rc;
  } else {
    GLIA_TRYC(launch_dot_real(c, RS_C0, RS_C0));
    GLIA_TRYC(launch_fin_state(c, ew_blocks(), FINK_C0SQ, 0.0));
  }
  GLIA_TRYC(launch_state_reset(c, 2));  // gradient accumulators; CG counters reset by the caller
  return 0;
}

// the end of an evaluation: Phi^T (alpha(0) - beta c(0)) for the p block, then J and dJ from the
// members' reductions
int RDEnsemble::eval_result(int mode, bool want_grad, std::vector<double>& J, std::vector<std::vector<double>>& dJ) {
  int rc;
  std::vector<std::vector<double>> phit;
  if (want_grad && mode != 1) {
    GLIA_TRYC(launch_axpby_real(ctx(), RS_A0, 1.0, RS_C0, -cfg_.beta, RS_CH));
    if ((rc = phi_apply_transpose(RS_CH, phit))) return rc;
  }
  if ((rc = sync())) return rc;
  const double h3 = lebesgue();
  J.assign(nm_, 0.0);
  dJ.assign(nm_, {});
  for (int m = 0; m < nm_; m++) {
    PcgState s;
    GLIA_TRYC(copy_sync(&s, st_[m], sizeof(s), hipMemcpyDeviceToHost));
    const double reg = s.c0sq * 0.5 * cfg_.beta * h3;
    J[m] = h3 * 0.5 * s.m1 + reg;
    dJ[m].assign(np_[m] + 2, 0.0);
    if (want_grad) {
      if (mode != 1)
        for (int i = 0; i < np_[m]; i++) dJ[m][i] = -h3 * phit[m][i];
      dJ[m][np_[m]] = h3 * cfg_.dt * s.gk;
      dJ[m][np_[m] + 1] = (mode == 0) ? 0.0 : h3 * cfg_.dt * s.gr;
    }
  }
  return 0;
}

// objective_and_gradient(mode, x, want_grad = true) of a store-all handle in phases, so that the
// time steps of several handles can run as one graph: eval_begin; for i < nt: the forward step i
// (fwd_prepare, launch, fwd_finish); eval_middle; for i = nt - 1 .. 0: the adjoint step i
// (adj_prepare, launch, adj_finish); eval_end.  The launches are those of objective_and_gradient.
int RDEnsemble::eval_begin(int mode, const std::vector<std::vector<double>>& x) {
  if (!hist_all_) return fail(-6, "a phased evaluation needs the store-all history");
  int rc = eval_setup(mode, x);
  if (rc) return rc;
  eval_mode_ = mode;
  return state_begin(true);
}

int RDEnsemble::eval_middle() {
  for (int i = 1; i <= cfg_.nt; i++) stored_[i] = i;
  int rc = inverse_to_real(CS_X, CS_T5, RS_CT);  // the end of solveState(0)
  if (rc) return rc;
  GLIA_TRYC(launch_state_reset(ctx(), 2));        // the start of solveAdjoint(1)
  return adjoint_begin();
}

int RDEnsemble::eval_end(std::vector<double>& J, std::vector<std::vector<double>>& dJ) {
  int rc = adjoint_end();
  if (rc) return rc;
  return eval_result(eval_mode_, true, J, dJ);
}

// Per-axis taps of the reference's periodic smoothing kernel (computeWeierstrassFilter,
// src/cuda/SpectralOperators.cu:12-51): g(x) + g(x - 2 pi) per axis, normalized by its discrete
// sum; the 3D kernel is their product.  Axis a keeps the taps |s| <= kr[a] = min(n_a / 2 - 1,
// ceil(12 sigma / h_a)) at taps[a * stride + s + kr[a]], stride = 2 max(kr) + 1: taps beyond 12
// sigma (< e^-72) are dropped, and an index wraps at most once, so no residue of a short axis is
// counted twice.  sigma <= 0 gives the identity (kr = 0).
void RDEnsemble::smoother_taps(double sigma, std::vector<double>& taps, int kr[3], int& stride) const {
  const int n[3] = {d_.n0, d_.n1, d_.n2};
  int kmax = 0;
  for (int a = 0; a < 3; a++) {
    kr[a] = (sigma > 0) ? std::min(n[a] / 2 - 1, (int)std::ceil(12.0 * sigma / h_[a])) : 0;
    kmax = std::max(kmax, kr[a]);
  }
  stride = 2 * kmax + 1;
  taps.assign(3 * stride, 0.0);
  for (int a = 0; a < 3; a++) {
    if (sigma <= 0) {
      taps[a * stride] = 1.0;
      continue;
    }
    std::vector<double> f1(n[a]);
    double sum = 0;
    for (int i = 0; i < n[a]; i++) {
      const double X = i * h_[a], Xp = X - 2.0 * M_PI;
      double v = std::exp(-X * X / (2 * sigma * sigma)) + std::exp(-Xp * Xp / (2 * sigma * sigma));
      if (v != v) v = 0;
      f1[i] = v;
      sum += v;
    }
    for (int s = -kr[a]; s <= kr[a]; s++) taps[a * stride + s + kr[a]] = f1[((s % n[a]) + n[a]) % n[a]] / sum;
  }
}

// SpectralOperators::weierstrassSmoother (src/grad/SpectralOperators.cpp:264-381) of a real field,
// in place: three periodic axis convolutions through a scratch buffer.  The device taps are kept
// while the width repeats (the mass-effect loops smooth with one width many times).
int RDEnsemble::smooth_taps_for(double sigma) {
  if (sigma == smooth_sigma_) return 0;
  std::vector<double> taps;
  int kra[3], stride = 1;
  smoother_taps(sigma, taps, kra, stride);
  if (taps.size() > smooth_cap_) {
    GLIA_HIP(hipStreamSynchronize(stream_));
    if (smooth_taps_) (void)hipFree(smooth_taps_);
    smooth_taps_ = nullptr;
    smooth_cap_ = 0;
    GLIA_HIP(hipMalloc((void**)&smooth_taps_, sizeof(double) * taps.size()));
    smooth_cap_ = taps.size();
  }
  const int rc = copy_sync(smooth_taps_, taps.data(), sizeof(double) * taps.size(), hipMemcpyHostToDevice);
  if (rc) return rc;
  smooth_sigma_ = sigma;
  smooth_stride_ = stride;
  for (int a = 0; a < 3; a++) smooth_kr_[a] = kra[a];
  return 0;
}

int RDEnsemble::smooth_field(int m, int rslot, double sigma) {
  if (m < 0 || m >= nm_ || rslot < 0 || rslot >= RS_NUM) return fail(-2, "bad member or field");
  if (!(sigma > 0)) return 0;
  int rc = smooth_taps_for(sigma);
  if (rc) return rc;
  const double* td = smooth_taps_;
  const int* kr = smooth_kr_;
  const int stride = smooth_stride_;
  float* f = mh_[m].r[rslot];
  float* scratch = reinterpret_cast<float*>(mh_[m].c[CS_T7]);  // n0*n1*n2 floats
  LaunchCtx c = ctx();
  if (!rc) rc = launch_smooth_axis(c, f, scratch, 0, td, kr[0]);
  if (!rc) rc = launch_smooth_axis(c, scratch, f, 1, td + stride, kr[1]);
  if (!rc) rc = launch_smooth_axis(c, f, scratch, 2, td + 2 * stride, kr[2]);
  if (!rc) rc = check(hipMemcpyAsync(f, scratch, sizeof(float) * d_.ng, hipMemcpyDeviceToDevice, stream_), "smooth copy");
  return rc;
}

int RDEnsemble::smooth_ptr(float* f, float* tmp, double sigma) {
  if (!f || !tmp || f == tmp) return fail(-2, "smoothing needs two distinct device fields");
  if (!(sigma > 0)) return 0;
  int rc = smooth_taps_for(sigma);
  if (rc) return rc;
  const double* td = smooth_taps_;
  const int* kr = smooth_kr_;
  const int stride = smooth_stride_;
  float* scratch = reinterpret_cast<float*>(mh_[0].c[CS_T7]);  // n0*n1*n2 floats
  LaunchCtx c = ctx();
  if (!rc) rc = launch_smooth_axis(c, f, scratch, 0, td, kr[0]);
  if (!rc) rc = launch_smooth_axis(c, scratch, tmp, 1, td + stride, kr[1]);
  if (!rc) rc = launch_smooth_axis(c, tmp, f, 2, td + 2 * stride, kr[2]);
  return rc;
}

// Phi::setGaussians(Vec data) (src/mat/Phi.cpp:615-940) on one device.  The candidates are the
// lattice the reference's loops walk (:686-720): per axis from 0 to the largest index that is an
// exact multiple of space = 2 sigma / h0, in steps x = (int)(x + space).  Radius r = sigma / h0 on
// every axis; the ball holds the integer offsets with sqrt(|d|^2) <= r.
int RDEnsemble::data_support(int m, int rslot, double thr, double sigma, double gvf, std::vector<double>* centers) {
  if (m < 0 || m >= nm_ || rslot < 0 || rslot >= RS_NUM || !centers) return fail(-2, "bad member, field or output");
  const int n[3] = {d_.n0, d_.n1, d_.n2};
  const double h[3] = {2.0 * M_PI / n[0], 2.0 * M_PI / n[1], 2.0 * M_PI / n[2]};
  const double r = sigma / h[0], space = 2.0 * sigma / h[0];
  if (!(space >= 1.0) || !(r < 64.0)) return fail(-2, "data support: sigma gives a lattice step below one cell or a radius above 64 cells");
  const int R = (int)std::floor(r);
  int d2max = -1, interior = 0;
  for (int q = 0; q <= 3 * R * R; q++)
    if (std::sqrt((double)q) <= r) d2max = q;
  for (int i = -R; i <= R; i++)
    for (int j = -R; j <= R; j++)
      for (int k = -R; k <= R; k++)
        if (i * i + j * j + k * k <= d2max) interior++;
  // the step sequence x -> (int)(x + space) from 0 is the same on every axis; each axis ends at its
  // largest exact multiple of space
  int end[3], len[3];
  for (int a = 0; a < 3; a++) {
    end[a] = 0;
    for (int x = n[a] - 1; x >= 0; x--)
      if (std::fmod((double)x, space) == 0) { end[a] = x; break; }
  }
  const int step = (int)(0 + space);
  for (int a = 0; a < 3; a++) {
    len[a] = 0;
    for (int x = 0; x <= end[a]; x = (int)(x + space), len[a]++)
      if (x != len[a] * step) return fail(-2, "data support: uneven lattice");
  }
  const int64_t nl = (int64_t)len[0] * len[1] * len[2];
  int* cnt = reinterpret_cast<int*>(mh_[0].c[CS_T7]);  // n0*n1*n2 ints; nl <= that
  LaunchCtx c = ctx();
  int rc = launch_support_count(c, mh_[m].r[rslot], step, len[0], len[1], len[2], R, d2max, (float)thr, cnt);
  if (rc) return rc < 0 ? fail(-2, "data support: lattice does not fit") : rc;
  std::vector<int> host((size_t)nl);
  if ((rc = check(hipMemcpyAsync(host.data(), cnt, sizeof(int) * nl, hipMemcpyDeviceToHost, stream_), "support counts")))
    return rc;
  if ((rc = check(hipStreamSynchronize(stream_), "support sync"))) return rc;
  centers->clear();
  int64_t t = 0;
  for (int a = 0; a < len[0]; a++)
    for (int b = 0; b < len[1]; b++)
      for (int cc = 0; cc < len[2]; cc++, t++)
        if (host[(size_t)t] > gvf * interior) {
          centers->push_back(a * step * h[0]);
          centers->push_back(b * step * h[1]);
          centers->push_back(cc * step * h[2]);
        }
  return 0;
}

// the tail of set_tissue: mask, k, rho, preconditioner and Phi from the stored tissue fields
int RDEnsemble::refresh_tissue(int m) {
  if (m < 0 || m >= nm_) return fail(-2, "bad member");
  int rc;
  LaunchCtx c = member_ctx(ctx(), md_, m);
  GLIA_TRYC(launch_tissue(c));
  GLIA_TRYC(launch_fin_state(c, ew_blocks(), FINK_FILTER, 0.0));
  if ((rc = set_kappa(m, kappa_[m]))) return rc;
  if ((rc = set_rho(m, rho_[m]))) return rc;
  GLIA_TRYC(launch_prec_factor(c, ctx_dt_));
  if (np_[m] > 0) {
    std::vector<double> ctr = centers_[m];
    return set_basis(m, np_[m], ctr.data(), sigma_[m]);
  }
  return sync();
}

// ---------------------------------------------------------------------------------------------
// profiling: event pairs around the per-iteration kernels, summed when queried
// ---------------------------------------------------------------------------------------------
hipEvent_t RDEnsemble::prof_event() {
  if (!prof_pool_.empty()) {
    hipEvent_t e = prof_pool_.back();
    prof_pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  // device-scope release: timing events must not add an L2 write-back between kernels
  if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) return nullptr;
  return e;
}

void RDEnsemble::prof_mark(int kind, bool begin) {
  if (!prof_on_) return;
  hipEvent_t e = prof_event();
  if (!e) return;
  (void)hipEventRecord(e, stream_);
  if (begin) {
    prof_recs_.push_back({kind, e, nullptr});
  } else if (!prof_recs_.empty() && prof_recs_.back().b == nullptr && prof_recs_.back().kind == kind) {
    prof_recs_.back().b = e;
  } else {
    prof_pool_.push_back(e);
  }
  if (prof_recs_.size() > 4096) (void)prof_collect();
}

int RDEnsemble::prof_collect() {
  if (prof_recs_.empty()) return 0;
  GLIA_HIP(hipStreamSynchronize(stream_));
  for (auto& r : prof_recs_) {
    if (r.b) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        prof_ms_[r.kind] += ms;
        prof_n_[r.kind] += 1;
      }
      prof_pool_.push_back(r.b);
    }
    prof_pool_.push_back(r.a);
  }
  prof_recs_.clear();
  return 0;
}

int RDEnsemble::profile_enable(bool on) {
  int rc = prof_collect();
  if (rc) return rc;
  prof_on_ = on;
  for (int k = 0; k < PK_NUM; k++) {
    prof_ms_[k] = 0;
    prof_n_[k] = 0;
  }
  return 0;
}

int RDEnsemble::profile_get(int kind, double* ms, long long* count) {
  if (kind < 0 || kind >= PK_NUM) return fail(-2, "bad profile kind");
  int rc = prof_collect();
  if (rc) return rc;
  *ms = prof_ms_[kind];
  *count = prof_n_[kind];
  return 0;
}

int RDEnsemble::stats(int m, MemberStats* out) {
  if (m < 0 || m >= nm_) return fail(-2, "bad member");
  GLIA_HIP(hipStreamSynchronize(stream_));
  PcgState s;
  GLIA_TRYC(copy_sync(&s, st_[m], sizeof(s), hipMemcpyDeviceToHost));
  out->its_total = s.its_total;
  out->solves = s.solves;
  out->worst_reason = s.worst_reason;
  out->last_its = s.its;
  out->last_reason = s.reason;
  out->kavg = s.kavg;
  out->pc = s.pc;
  out->filter_sum = s.filter_sum;
  out->m1 = s.m1;
  out->c0sq = s.c0sq;
  return 0;
}

}  // namespace glia
