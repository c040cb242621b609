// rd_solver.h -- host-side solver of the reaction-diffusion forward / adjoint / gradient path for
// an ensemble of members that share one grid and one time discretization.  A plain solve is an
// ensemble of one.  All device work is enqueued on one HIP stream; kernels of a launch cover
// every member (grid.y = member), and converged members' workgroups exit at once.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "rd_device.h"
#include "rd_launch.h"

namespace glia {

struct SolverConfig {
  int n[3] = {64, 64, 64};
  int nmembers = 1;
  double dt = 0.5;          // TumorParameters::dt_ default
  int nt = 1;               // TumorParameters::nt_ default
  double rtol = 1e-6;       // DiffusionSolver KSP tolerances
  int maxit = 5000;
  int checkpoint_slots = 0; // 0: store every time step; s > 0: binomial checkpointing with s states
  int adjoint_store = 0;    // keep half-step states (store-every-step mode only)
  int phi_store = 0;        // Phi normalization of phi_store_ (global max) vs on-the-fly
  double k_gm_wm = 0.0, k_glm_wm = 0.0, r_gm_wm = 0.0, r_glm_wm = 0.0;
  double beta = 0.0;        // L2 regularization weight
  double k_scale = 1.0, rho_scale = 1.0;  // OptimizerSettings k_scale_, rho_scale_
  double smoothing_factor = 1.0;
  hipStream_t stream = nullptr;
};

// kernel kinds timed by the profiling mode (HIP events around each launch on the solver stream)
enum ProfKind { PK_XI_PUPD = 0, PK_YI_GRAD, PK_Z_DIFF, PK_YF_DIV, PK_XF_MATVEC, PK_PCG_UPDATE, PK_NUM };

struct MemberStats {
  int its_total, solves, worst_reason, last_its, last_reason;
  double kavg, pc, filter_sum, m1, c0sq;
};

class RDEnsemble {
 public:
  explicit RDEnsemble(const SolverConfig& cfg);
  ~RDEnsemble();
  int init();  // allocation; returns 0 or an error code (message in error())

  const std::string& error() const { return err_; }
  const Dims& dims() const { return d_; }
  int nmembers() const { return nm_; }
  const SolverConfig& config() const { return cfg_; }

  // --- setup (per member; host or device source pointers, FP32 fields of n0*n1*n2) ---
  int set_tissue(int m, const float* wm, const float* gm, const float* csf, const float* vt, bool device);
  int set_tissue_sinusoidal(int m);
  int set_field(int m, int rslot, const float* src, bool device);
  int get_field(int m, int rslot, float* dst, bool device);
  float* field_ptr(int m, int rslot) const { return mh_[m].r[rslot]; }  // device address of a real slot
  hipStream_t stream() const { return stream_; }
  int set_coefficients(int m, double kappa, double rho);     // DiffCoef/ReacCoef::setValues
  int set_kappa(int m, double kappa);
  int set_rho(int m, double rho);
  int set_k_sinusoidal(int m, double scale);                 // DiffCoef::setValuesSinusoidal
  int prec_factor();                                         // DiffusionSolver::precFactor
  void set_ctx_dt(double dt) { ctx_dt_ = dt; }
  void set_beta(double beta) { cfg_.beta = beta; }
  void set_ratios(double k_gm, double k_glm, double r_gm, double r_glm) {
    cfg_.k_gm_wm = k_gm;
    cfg_.k_glm_wm = k_glm;
    cfg_.r_gm_wm = r_gm;
    cfg_.r_glm_wm = r_glm;
  }
  double ctx_dt() const { return ctx_dt_; }
  double kappa(int m) const { return kappa_[m]; }  // TumorParameters::k_ of member m
  double rho(int m) const { return rho_[m]; }      // TumorParameters::rho_ of member m
  int set_basis(int m, int np, const double* centers, double sigma);
  int set_data(int m, const float* d1, const float* obs_mask, double thr, bool device);

  // --- operations on every member ---
  int phi_apply(const std::vector<std::vector<double>>& p);  // c0 = Phi p (RS_C0)
  int phi_apply_transpose(int rslot, std::vector<std::vector<double>>& out);  // Phi^T v
  int diffusion_solve_field(int rslot, double tau);          // real field in place (test/API)
  int apply_d_field(int rslot_in, int rslot_out);            // real D c (test/API)
  int fft_r2c(int rslot, int cslot);                         // packed spectrum of a real field
  int fft_c2r(int cslot, int rslot);                         // real field of a packed spectrum (1/N^3 included)
  // out = div (x, y, z) of three real fields of every member, spectrally (uses CS_T0..CS_T3, CS_T7)
  int divergence(int rx, int ry, int rz, int rout);
  // out = i sum_k w_{dirs[k]} X_{slots[k]} on packed spectra (1 to 4 terms)
  int spec_lin(int nterms, const int* slots, const int* dirs, int cout);
  // the constant-coefficient elasticity preconditioner on packed spectra, (fx, fy, fz) -> (ux, uy, uz)
  int el_prec(const int f[3], const int u[3], double mu, double lam, double scr);
  // members as lanes: launches cover the first k members (0: all); see launch_lane_lin and
  // launch_el_prec_lanes for the cross-lane kernels
  void set_lanes(int k) { lanes_ = k; }
  int lane_lin(const int* nterms, const int (*lane)[4], const int (*slot)[4], const int (*dir)[4], int cout);
  int el_prec_lanes(int cf, int cu, double mu, double lam, double scr);
  // out spectrum cs_out = FFT(m IFFT(X)) with m the real slot coef_rs and X the cross-lane
  // combination of the terms (launch_lane_lin's), formed in the first inverse pass; the product is in
  // the Z pass between the two transforms (CS_T7 as work)
  int lin_coef_transform(const int* nterms, const int (*lane)[4], const int (*slot)[4], const int (*dir)[4],
                         int coef_rs, int cs_out);
  // real slot out_rs = IFFT(X) - s u with X the combination, s in RS_RHO and u the real slot u_rs
  int lin_screen_transform(const int* nterms, const int (*lane)[4], const int (*slot)[4], const int (*dir)[4],
                           int u_rs, int out_rs);
  // real slot out_rs = IFFT(cs_in), with per-block partials of <out, out> and <out, r> (r the real
  // slot r_rs) left for lane_part_sum (CS_T7 as work)
  int c2r_dots(int cs_in, int out_rs, int r_rs);
  // out[q] (device), q < nq: the last Z pass's block partials of the lanes in use, summed in a fixed
  // order (lin_screen_transform leaves <u, out>, c2r_dots <out, out> and <out, r>)
  int lane_part_sum(int nq, double* out);
  // The elasticity CG on the lanes, its scalars in lane 0's PcgState (launch_el_cg_begin /
  // launch_el_cg_fin).  With honor_lanes set, the lane launches below return at once for lanes whose
  // CG has finished, so iterations enqueued past the stop cost only their launches.
  void set_honor_lanes(bool v) { honor_lanes_ = v; }
  // packed spectrum cs_out of the real slot p_rs after p = z + (beta / betaold) p in place (p = z in
  // the first iteration)
  int fft_r2c_axpby(int p_rs, int z_rs, int cs_out);
  // packed spectrum cs_out of the real slot r_rs after x += t p, r -= t w in place, t = beta / sc[0]
  // (skipped when sc[0] is 0, not finite, or of the sign opposite to the previous one)
  int fft_r2c_cgupd(int x_rs, int r_rs, int p_rs, int w_rs, const double* sc, int cs_out);
  int el_cg_begin(double beta0, double dp0, double rnorm0, double ttol, double atol, double dtol, int maxit);
  // closes an iteration from sc = (<p, A p>, <z, z>, <z, r>); sets *hflag (mapped) when the CG stops
  int el_cg_fin(const double* sc, int* hflag);
  int el_cg_state(int* its, int* reason, double* dp);  // synchronizes the stream
  // after a CG: lane launches no longer honor done, and the lanes' done flags are cleared
  int el_cg_end();
  int get_spectrum(int m, int cslot, float2* host);
  int set_spectrum(int m, int cslot, const float2* host);
  int solve_state(int linearized, bool keep_history);
  int solve_adjoint(int linearized);
  int objective_and_gradient(int mode, const std::vector<std::vector<double>>& x, bool want_grad,
                             std::vector<double>& J, std::vector<std::vector<double>>& dJ);
  // Gauss-Newton Hessian matvec (evaluateHessian; kappa blocks with diffusivity): needs the state
  // history of a previous evaluation at the linearization point
  int hessian(const std::vector<std::vector<double>>& xt, bool diffusivity, std::vector<std::vector<double>>& y);
  int stats(int m, MemberStats* s);
  // TumorParameters dt_/nt_ with PdeOperatorsRD::resizeTimeHistory (the history is reallocated)
  int set_time(double dt, int nt);
  // keep the adjoint history p_[0..nt] during solveAdjoint (store-all mode; nt + 1 spectral fields)
  int keep_adjoint_history(bool on);
  // c_[i] (HIST_STATE), c_half_[i] (HIST_HALF) or p_[i] (HIST_ADJOINT) of member m as a real field
  enum { HIST_STATE = 0, HIST_HALF = 1, HIST_ADJOINT = 2 };
  int get_history(int m, int kind, int index, float* dst, bool device);
  // DiffCoef::setSecondaryCoefficients for solveState(2): k~ = kt (wm + k_gm gm + k_glm csf)
  int set_secondary(int m, double kt);
  // gradDiffusion / gradReaction of member m from the last adjoint sweep's quadratures
  int grad_kappa_rho(int m, double* gk, double* gr);
  int basis_size(int m) const { return (m >= 0 && m < nm_) ? np_[m] : -1; }
  int reset_counters();
  int sync();

  double lebesgue() const { return h_[0] * h_[1] * h_[2]; }

  // profiling mode: per-kind launch count and summed event time (ms) since enabled
  int profile_enable(bool on);
  int set_use_graphs(bool on);
  // time the forward divergence variants on this handle (or read them from the cache file) and
  // keep the faster; choice (2 ints, may be NULL) receives (xf_split, yf_split)
  int autotune(const char* cache_path, int* choice);
  void variants(int* xf, int* yf) const { *xf = xf_split_; *yf = yf_split_; }
  int smooth_field(int m, int rslot, double sigma);  // weierstrassSmoother, in place
  // the same smoother on any device field f of the grid, in place, through tmp (a distinct field
  // of the grid) and the spectral work slot CS_T7 of member 0; no copies, no host synchronization
  int smooth_ptr(float* f, float* tmp, double sigma);
  // Phi::setGaussians(Vec data): the lattice centers whose ball holds more than gvf of its voxels
  // above thr in real field rslot of member m (counted on the device, selected in index order)
  int data_support(int m, int rslot, double thr, double sigma, double gvf, std::vector<double>* centers);
  int refresh_tissue(int m);                         // derived quantities from stored tissue fields
  int profile_get(int kind, double* ms, long long* count);

  // --- captured time steps and the phases of a gradient evaluation (store-all history) ---
  using StepKey = std::array<int, 6>;  // phase, step, budgets, variant flags
  struct StepGraph {
    hipGraph_t g = nullptr;       // the captured step, for graphs of several handles
    hipGraphExec_t ex = nullptr;  // its executable form, replayed alone
  };
  bool step_graphs_on() const;
  // one forward step / adjoint step: prepare gives the captured step to launch (null: the step
  // runs eagerly in finish), finish reads the budgets' flags after the launch and runs the step
  // again eagerly when a solve needed more iterations
  int fwd_prepare(int i, bool store_half, const StepGraph** out);
  int fwd_finish(int i, bool store_half, bool launched);
  int adj_prepare(int i, const StepGraph** out);
  int adj_finish(int i, bool launched);
  // objective_and_gradient (want_grad, store-all) in phases: begin runs up to c(0) in the history,
  // middle from c(T) through alpha(T) and the quadrature at T, end from alpha(0) to J and dJ
  int eval_begin(int mode, const std::vector<std::vector<double>>& x);
  int eval_middle();
  int eval_end(std::vector<double>& J, std::vector<std::vector<double>>& dJ);
  bool store_half() const { return hist_all_ && cfg_.adjoint_store != 0; }
  bool store_all() const { return hist_all_; }
  int nt() const { return cfg_.nt; }
  // process-unique handle number and the count of step-graph cache flushes: with the address of
  // a cached step they identify it in the key of a graph spanning handles
  unsigned long long uid() const { return uid_; }
  unsigned long long graph_epoch() const { return graph_epoch_; }

 private:
  unsigned long long uid_ = 0;
  unsigned long long graph_epoch_ = 0;
  int lanes_ = 0;
  bool honor_lanes_ = false;  // lane launches skip lanes whose elasticity CG has finished
  int eval_mode_ = 2;
  int eval_setup(int mode, const std::vector<std::vector<double>>& x);
  int eval_result(int mode, bool want_grad, std::vector<double>& J, std::vector<std::vector<double>>& dJ);
  int state_begin(bool keep_history);
  int adjoint_begin();
  int adjoint_end();
  void drop_step_graphs();
  int alloc_member(int m);
  int alloc_history(int m);
  void free_history(int m);
  int store_adjoint(int i);
  int solve_tangent(bool sources);
  int upload_members();
  int check(hipError_t e, const char* what);
  int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
  int fail(int code, const std::string& msg);
  LaunchCtx ctx() const;

  void smoother_taps(double sigma, std::vector<double>& taps, int kr[3], int& stride) const;
  int smooth_taps_for(double sigma);
  // the solves of a time step whose iteration counts budget its capture
  enum { SK_FWD1 = 0, SK_FWD2, SK_ADJ1, SK_HALF, SK_ADJ2, SK_NUM };
  int cn_solve(int sx, double tau, int kind = -1);
  int capture_solve(int sx, double tau, int K, int fslot);
  int step_graph(const StepKey& key, const std::function<int()>& body, const StepGraph** out);
  int read_step_flags(int nsolve, int* its, bool* ok);
  int forward_step_stored(int i, bool store_half);
  int adjoint_step_stored(int i);
  int launch_iteration(int sx, int slot);                       // the launches of one CG iteration
  int iteration_graph(int sx, int slot, hipGraphExec_t* out);   // the same, captured once and cached
  int matvec_setup(int sx);
  int forward_step(int sx, int step, bool store_half);
  int forward_transform_real(int rslot, int cslot);  // R2C of a real slot into a complex slot
  int inverse_to_real(int cslot, int work, int rslot);
  int adjoint_step(int i, int cslot_c);
  int tangent_step(int sx, int cslot_c);  // linearized forward step around the state in cslot_c
  int quad(int cslot_c, int cslot_a, double w);
  int misfit_and_final_adjoint(int cslot_c, int out = CS_A);
  int apply_dsec(int in, int out);
  int incremental_source(int sx, int ci, int ci1, double wa, double wb);
  int tangent_forward(int sx, bool sources, int* ct_slot);
  int incremental_adjoint(int ct_slot, bool quads);
  int clear_done();
  int reverse(int a, int b, int pa, int free_slots);
  int state_slot(int pool_index) const { return CS_NUM + pool_index; }

  SolverConfig cfg_;
  Dims d_;
  int nm_ = 0;
  double h_[3];
  std::string err_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  double ctx_dt_ = 0.5;

  // per-member device memory
  std::vector<MemberDev> mh_;           // host mirror of the member descriptors
  MemberDev* md_ = nullptr;             // device array
  std::vector<std::vector<float2*>> pool_;   // spectral state pool (history / checkpoints)
  std::vector<std::vector<float*>> chalf_;   // real half-step states (adjoint_store)
  std::vector<std::vector<float2*>> apool_;  // adjoint history p_[0..nt] when kept
  bool keep_adj_ = false;
  bool adj_valid_ = false;
  std::vector<float2**> hist_dev_;
  std::vector<float**> chalf_dev_;
  std::vector<PcgState*> st_;
  std::vector<double*> part_;
  std::vector<int> np_;
  std::vector<std::vector<double>> phimax_;  // per-basis maxima of the raw bricks
  std::vector<std::vector<double>> centers_;
  std::vector<double> sigma_;
  std::vector<Brick*> bricks_dev_;
  std::vector<float*> brick_data_;
  std::vector<double*> phi_scale_dev_;
  std::vector<double*> phit_dev_;
  std::vector<double> kappa_, rho_;
  std::vector<int> active_;
  double* coefs_dev_ = nullptr;

  const float2* tw_[4] = {nullptr, nullptr, nullptr, nullptr};
  float2* twbuf_ = nullptr;

  int* hflag_host_ = nullptr;
  int* hflag_dev_ = nullptr;
  static constexpr int kFlagSlots = 4;
  static constexpr int kStepFlags = 3;  // host flag slots of the budgeted solves of a captured step
  hipEvent_t ev_step_ = nullptr;
  int budget_[SK_NUM] = {-1, -1, -1, -1, -1};  // iterations of the last solve of each kind (-1: none yet)
  std::map<StepKey, StepGraph> step_graphs_;
  long long step_reruns_ = 0;  // captured steps that ran again eagerly
  long long step_launches_ = 0;  // captured steps launched
  int margin_ = 1;             // CG iterations captured beyond the last count of a solve kind

 public:
  void step_info(long long* launched, long long* reruns, int* cached) const {
    *launched = step_launches_;
    *reruns = step_reruns_;
    *cached = (int)step_graphs_.size();
  }

 private:
  static constexpr int kLookahead = 2;
  hipEvent_t ev_[kFlagSlots];
  bool ev_ok_ = false;

  struct ProfRec {
    int kind;
    hipEvent_t a, b;
  };
  int yf_split_ = 0, xf_split_ = 0;  // variants of the forward divergence passes (autotune)
  static constexpr int kYfVariants = 4;
  static constexpr int kXfVariants = 3;
  bool quad_on_ = true;     // adjoint steps accumulate the kappa/rho gradient quadratures
  bool use_graphs_ = true;  // replay CG iterations from captured graphs (not while profiling)
  double* smooth_taps_ = nullptr;  // device taps of the last smoothing width (three axes)
  size_t smooth_cap_ = 0;          // doubles allocated at smooth_taps_
  double smooth_sigma_ = -1.0;
  int smooth_kr_[3] = {0, 0, 0};   // per-axis radius of the last width
  int smooth_stride_ = 1;          // doubles per axis in smooth_taps_
  std::map<std::pair<int, int>, hipGraphExec_t> graphs_;  // (state slot, flag slot) -> graph
  bool prof_on_ = false;
  std::vector<ProfRec> prof_recs_;
  std::vector<hipEvent_t> prof_pool_;
  double prof_ms_[PK_NUM] = {0};
  long long prof_n_[PK_NUM] = {0};
  hipEvent_t prof_event();
  void prof_mark(int kind, bool begin);
  int prof_collect();

  int npool_ = 0;       // number of pool buffers per member
  bool hist_all_ = true;
  std::map<int, int> stored_;  // time index -> pool index (checkpoint mode)
  std::vector<int> free_pool_;
};

}  // namespace glia
