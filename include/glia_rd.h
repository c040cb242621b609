/*
 * glia_rd.h -- C ABI of the MI355X-native reaction-diffusion forward / adjoint / gradient path.
 *
 * This is the drop-in boundary for the reference's RD hot path.  Each entry point names the
 * reference interface it replaces (file:line in the GLIA sources).  PETSc Vec/Mat/KSP objects
 * become plain pointers and sizes; a handle owns every device buffer it computes with and copies
 * its inputs.  Return value: 0 on success, < 0 for argument / state errors, > 0 for a HIP runtime
 * error code; glia_rd_last_error() gives the message.  All device work of a handle is ordered on
 * the handle's stream; use one handle per host thread.
 *
 * Fields are FP32, n0*n1*n2 values, row-major with z fastest: index x*n1*n2 + y*n2 + z
 * (reference src/grad/SpectralOperators.cpp:3-17).  mem_kind says whether a pointer is host or
 * device memory.  A handle holds an ensemble of 'nmembers' members that share the grid and the
 * time discretization (nmembers = 1 is the plain solve); per-member calls take a member index.
 */
#ifndef GLIA_RD_H_
#define GLIA_RD_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glia_rd_s* glia_rd_t;

enum { GLIA_RD_HOST = 0, GLIA_RD_DEVICE = 1 };

/* error codes (negative); positive values are hipError_t */
enum {
  GLIA_RD_OK = 0,
  GLIA_RD_ERR_ARG = -2,
  GLIA_RD_ERR_STATE = -3,
  GLIA_RD_ERR_UNSUPPORTED = -4,
  GLIA_RD_ERR_POOL = -5,
  GLIA_RD_ERR_NOHISTORY = -6
};

/* fields addressable by set_field / get_field */
enum {
  GLIA_RD_FIELD_K = 0,      /* DiffCoef::kxx_ (isotropic k) */
  GLIA_RD_FIELD_RHO = 1,    /* ReacCoef::rho_vec_ */
  GLIA_RD_FIELD_WM = 2,     /* MatProp::wm_ */
  GLIA_RD_FIELD_GM = 3,
  GLIA_RD_FIELD_CSF = 4,
  GLIA_RD_FIELD_VT = 5,
  GLIA_RD_FIELD_FILTER = 6, /* MatProp::filter_ (brain mask) */
  GLIA_RD_FIELD_OBS = 7,    /* Obs::filter_1_ */
  GLIA_RD_FIELD_D1 = 8,     /* Data::dt1() */
  GLIA_RD_FIELD_C0 = 9,     /* Tumor::c_0_ */
  GLIA_RD_FIELD_CT = 10,    /* Tumor::c_t_ after solveState */
  GLIA_RD_FIELD_ALPHA0 = 11,/* Tumor::p_0_ after solveAdjoint */
  GLIA_RD_FIELD_WORK = 12   /* scratch field (half-step state) */
};

/* evaluation modes of the derivative operators: which parameters the inversion updates */
enum {
  GLIA_RD_MODE_P_KAPPA = 0,   /* flag_reaction_inv_ = false: c0 = Phi p; dJ = [g_p | g_kappa | 0] */
  GLIA_RD_MODE_KAPPA_RHO = 1, /* flag_reaction_inv_ = true: c0 as set; dJ = [0 | g_kappa | g_rho] */
  GLIA_RD_MODE_ALL = 2        /* p, kappa and rho together: dJ = [g_p | g_kappa | g_rho] */
};

/* Parameters, TumorParameters, Grid and the DiffusionSolver KSP settings
   (include/Parameters.h:155-451; src/pde/DiffusionSolver.cpp:25-39) */
typedef struct glia_rd_config {
  int n[3];              /* grid; n0, n1 in {32,48,64,80,96,128,160,192,256},
                            n2 in {32,64,96,128,160,192,256} */
  int nmembers;          /* ensemble members sharing this grid (1 = plain solve) */
  double dt;             /* TumorParameters::dt_ */
  int nt;                /* TumorParameters::nt_ */
  int order;             /* TumorParameters::order_; only 2 (Strang) */
  double pcg_rtol;       /* KSPSetTolerances rtol (1e-6) */
  int pcg_maxit;         /* KSPSetTolerances maxits (5000) */
  int checkpoint_slots;  /* 0: keep every time step (reference); s > 0: binomial checkpointing, s states */
  int adjoint_store;     /* TumorParameters::adjoint_store_ (keep half-step states; store-all only) */
  int phi_store;         /* TumorParameters::phi_store_ (Phi normalization semantics) */
  double k_gm_wm, k_glm_wm, r_gm_wm, r_glm_wm; /* tissue ratios of k and rho */
  double beta;           /* OptimizerSettings::beta_ (L2 regularization) */
  double k_scale, rho_scale; /* OptimizerSettings::k_scale_, rho_scale_ */
  double smoothing_factor;   /* TumorParameters::smoothing_factor_ (Phi smoothing) */
  void* stream;          /* hipStream_t to use, or NULL for a private stream */
} glia_rd_config;

typedef struct glia_rd_stats_s {
  int pcg_iterations;    /* CG iterations since the last reset (diff_ksp_itr_state_ + _adj_) */
  int cn_solves;         /* DiffusionSolver::solve calls since the last reset */
  int worst_reason;      /* most negative KSPConvergedReason since the last reset (0 if none) */
  int last_iterations;   /* DiffusionSolver::ksp_itr_ of the last solve */
  int last_reason;       /* KSPConvergedReason of the last solve */
  double kbar;           /* DiffCoef::kxx_avg_ */
  double prec_coef;      /* 0.25 * dt * kbar baked into the preconditioner */
  double filter_sum;     /* sum of MatProp::filter_ */
  double misfit;         /* ||O c(T) - d1||^2 of the last evaluation */
  double c0_norm2;       /* ||c(0)||^2 */
} glia_rd_stats_t;

/* defaults of TumorParameters / DiffusionSolver (include/Parameters.h:157-258) */
void glia_rd_config_init(glia_rd_config* cfg);

/* construction of Tumor, SpectralOperators, PdeOperatorsRD, DerivativeOperatorsRD
   (SolverInterface::initializeOperators, src/SolverInterface.cpp:929-965; PdeOperatorsRD ctor,
   src/pde/PdeOperators.cpp:17-54; DiffusionSolver ctor, src/pde/DiffusionSolver.cpp:5-45) */
int glia_rd_create(const glia_rd_config* cfg, glia_rd_t* out);
int glia_rd_destroy(glia_rd_t h);
const char* glia_rd_last_error(glia_rd_t h);
int glia_rd_sync(glia_rd_t h);
int glia_rd_num_members(glia_rd_t h);
/* the grid of the handle (Parameters::grid_->n_): n[3] */
int glia_rd_get_grid(glia_rd_t h, int* n);

/* MatProp::setValuesCustom + SolverInterface::updateTumorCoefficients
   (src/mat/MatProp.cpp:134-203; src/SolverInterface.cpp:986-1014); null fields are zero */
int glia_rd_set_tissue(glia_rd_t h, int member, const float* wm, const float* gm, const float* csf, const float* vt,
                       int mem_kind);
/* MatProp::setValuesSinusoidal (src/mat/MatProp.cpp:34-68) */
int glia_rd_set_tissue_sinusoidal(glia_rd_t h, int member);
/* DiffCoef::setValues + ReacCoef::setValues (src/mat/DiffCoef.cpp:77-132; src/mat/ReacCoef.cpp:13-38) */
int glia_rd_set_coefficients(glia_rd_t h, int member, double kappa, double rho);
/* DiffCoef::setValues alone (src/mat/DiffCoef.cpp:77-132) and ReacCoef::setValues alone
   (src/mat/ReacCoef.cpp:13-38) */
int glia_rd_set_kappa(glia_rd_t h, int member, double kappa);
int glia_rd_set_rho(glia_rd_t h, int member, double rho);
/* DiffCoef::setValuesSinusoidal (src/mat/DiffCoef.cpp:134-189) */
int glia_rd_set_k_sinusoidal(glia_rd_t h, int member, double scale);
/* DiffusionSolver::precFactor (src/pde/DiffusionSolver.cpp:119-180), all members */
int glia_rd_prec_factor(glia_rd_t h);
/* DiffusionSolver Ctx::dt_ (src/pde/DiffusionSolver.cpp:11,226) */
int glia_rd_set_ctx_dt(glia_rd_t h, double dt);
int glia_rd_get_ctx_dt(glia_rd_t h, double* dt);
/* OptimizerSettings::beta_ (L2 regularization of c0), which the reference's optimizers switch
   to 0 for gradient thresholding and reaction/diffusion solves (src/opt/SparseTILOptimizer.cpp:328,
   src/opt/RDOptimizer.cpp:170); beta >= 0 */
int glia_rd_set_beta(glia_rd_t h, double beta);
int glia_rd_get_beta(glia_rd_t h, double* beta);
/* the tissue ratios that DiffCoef::setValues / ReacCoef::setValues take with each call
   (src/mat/DiffCoef.cpp:77-92, src/mat/ReacCoef.cpp:13-38); they apply from the next
   glia_rd_set_coefficients or objective evaluation on */
int glia_rd_set_ratios(glia_rd_t h, double k_gm_wm, double k_glm_wm, double r_gm_wm, double r_glm_wm);
/* SpectralOperators::weierstrassSmoother (src/grad/SpectralOperators.cpp:264-381) of a stored real
   field of one member, in place, with standard deviation sigma (the reference smooths atlas
   tissue with sigma = smoothing_factor * 2 pi / n0, SolverInterface::readAtlas,
   src/SolverInterface.cpp:502-575) */
int glia_rd_smooth_field(glia_rd_t h, int member, int which, double sigma);
/* The smoother of glia_rd_smooth_field on caller-owned device fields, allocating nothing.  It has
   no counterpart in the reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops).  in and out are distinct device fields of
   n[0]*n[1]*n[2] floats (any lengths >= 4, at most 2^23 = 8,388,608 points); sigma > 0 with a
   smoother radius (ceil(12 sigma / h) per axis) of at most 127 points; stream may be NULL (the
   null stream).  The taps and the intermediate field live in one 32 MiB device array per GPU
   that every call on that GPU shares: calls that overlap on one GPU (other streams, other host
   threads) race.  Returns 0, GLIA_RD_ERR_ARG, GLIA_RD_ERR_UNSUPPORTED (grid or radius above the
   limits) or a hipError_t; err (may be NULL) receives the message. */
int glia_rd_smooth_noalloc(const int n[3], const float* in, float* out, double sigma, void* stream, char* err,
                           int err_len);
/* Phi::setGaussians(Vec data) (src/mat/Phi.cpp:615-940, one process): the data-driven basis
   support from real field `which` of a member (the reference passes Data::dt1() after smoothing, or
   a support map read with dataIn, src/SolverInterface.cpp:622-633).  Candidates lie on the lattice
   of the reference's loops: per axis from 0 to the largest index that is an exact multiple of
   space = 2 sigma / h0, in steps x = (int)(x + space).  A candidate counts the voxels with
   field > threshold (single precision) at integer offsets d with sqrt(|d|^2) <= sigma / h0, clipped
   to the grid without wrapping; it is kept when the count exceeds gvf times the full ball's count.
   The kept centers (x, y, z in [0, 2 pi)) are written in index order with *np their number;
   centers may be NULL with max_np = 0 to query the count.  The counting runs on the device.
   Returns GLIA_RD_ERR_ARG (with *np set) when more than max_np are kept, or when space < 1 (the
   reference's loop would not advance) or the radius sigma / h0 reaches 64 cells.  The reference's defaults: sigma = 2 pi / 256
   (phi_sigma_data_driven), threshold 0.05 (data_threshold), gvf 0 (gaussian_vol_frac). */
int glia_rd_data_support(glia_rd_t h, int member, int which, double threshold, double sigma, double gvf, int max_np,
                         double* centers, int* np);
/* MatProp::setValuesCustom on the tissue fields already stored (e.g. after glia_rd_set_field and
   glia_rd_smooth_field): mask, k, rho, preconditioner and Phi of the member */
int glia_rd_refresh_tissue(glia_rd_t h, int member);
/* Phi::setGaussians(centers) + Phi::setValues (src/mat/Phi.cpp:32-144): np centers (x,y,z in [0,2pi)) */
int glia_rd_set_basis(glia_rd_t h, int member, int np, const double* centers, double sigma);
int glia_rd_basis_size(glia_rd_t h, int member);
/* Data::set + Obs::setDefaultFilter / setCustomFilter (src/mat/Obs.cpp:26-69):
   obs_mask NULL -> O = 1[d1 > obs_thr] */
int glia_rd_set_data(glia_rd_t h, int member, const float* d1, const float* obs_mask, double obs_thr, int mem_kind);
int glia_rd_set_field(glia_rd_t h, int member, int which, const float* src, int mem_kind);
int glia_rd_get_field(glia_rd_t h, int member, int which, float* dst, int mem_kind);

/* Phi::apply (src/mat/Phi.cpp:313-377): c0 = Phi p for every member; p holds np(m) values per member */
int glia_rd_phi_apply(glia_rd_t h, const double* p);
/* Phi::applyTranspose (src/mat/Phi.cpp:379-439) of field 'which'; out holds np(m) values per member */
int glia_rd_phi_apply_transpose(glia_rd_t h, int which, double* out);
/* DiffusionSolver::solve (src/pde/DiffusionSolver.cpp:217-250) on field 'which', in place, all members */
int glia_rd_diffusion_solve(glia_rd_t h, int which, double tau);
/* DiffCoef::applyD (src/mat/DiffCoef.cpp:249-271): out = div(k grad in), all members */
int glia_rd_apply_d(glia_rd_t h, int which_in, int which_out);
/* SpectralOperators::executeFFTR2C / executeFFTC2R (src/grad/SpectralOperators.cpp:68-98) for one
   member, host arrays; spectrum in the reference layout n0*n1*(n2/2+1) interleaved complex.
   c2r returns the unnormalized inverse (N^3 times the field for a forward spectrum). */
int glia_rd_fft_r2c(glia_rd_t h, int member, const float* f, float* fhat);
int glia_rd_fft_c2r(glia_rd_t h, int member, const float* fhat, float* f);

/* PdeOperatorsRD::solveState (src/pde/PdeOperators.cpp:235-316).
   linearized 0: the state equation from c0, keeping the history the checkpoint policy asks for.
   linearized 1: the tangent-linear equation from c~0 (the c0 field) around the state history of
     the last solveState(0), reaction(1, i) linearized at c_[i]; c~(T) in the ct field.
   linearized 2: as 1, with the solveIncremental sources (:192-233) of the secondary coefficients
     set by glia_rd_set_secondary (the Gauss-Newton Hessian with diffusivity inversion).
   1 and 2 leave the state history as it is. */
int glia_rd_solve_state(glia_rd_t h, int linearized);
/* PdeOperatorsRD::solveAdjoint (src/pde/PdeOperators.cpp:372-420), accumulating the quadratures of
   gradDiffusion / gradReaction along the sweep; alpha(0) in the alpha0 field.
   linearized 0 or 1: from -O^T(O c(T) - d1) of DerivativeOperatorsRD
     (src/grad/DerivativeOperatorsRD.cpp:154-162) and the state history.
   linearized 2: the incremental adjoint from -O^T O c~(T), c~(T) = the ct field after
     solveState(1|2) (evaluateHessian, :404-420); the i = nt quadrature term uses the last
     gradient's final condition, as the reference's p_[nt] does. */
int glia_rd_solve_adjoint(glia_rd_t h, int linearized);
/* TumorParameters dt_ / nt_ with PdeOperatorsRD::resizeTimeHistory (src/pde/PdeOperators.cpp:56-103,
   called through PdeOperatorsRD::reset, :6-15): the history is reallocated for the new nt and the
   stored states dropped; the diffusion solver's context dt is unchanged. */
int glia_rd_set_time(glia_rd_t h, double dt, int nt);
/* Keep the adjoint history p_[0..nt] during solveAdjoint (enable = 1; store-all mode only: nt + 1
   more spectral fields per member) or release it (0).  The reference always keeps it
   (src/pde/PdeOperators.cpp:26-44); the gradient evaluation here does not need it. */
int glia_rd_keep_adjoint_history(glia_rd_t h, int enable);
/* The reference's public history members of PdeOperators (include/pde/PdeOperators.h:24-29) as a
   real field of one member: kind GLIA_RD_HIST_STATE c_[index] (0..nt; under checkpointing only
   the stored indices), GLIA_RD_HIST_HALF c_half_[index] (0..nt-1, adjoint_store in store-all
   mode), GLIA_RD_HIST_ADJOINT p_[index] (0..nt, when kept).  GLIA_RD_ERR_NOHISTORY when the entry
   is not held.  Overwrites the work field. */
enum { GLIA_RD_HIST_STATE = 0, GLIA_RD_HIST_HALF = 1, GLIA_RD_HIST_ADJOINT = 2 };
int glia_rd_get_history(glia_rd_t h, int member, int kind, int index, float* dst, int mem_kind);
/* DiffCoef::setSecondaryCoefficients (src/mat/DiffCoef.cpp:134-168) for solveState(2):
   k~ = kt (m_wm + k_gm_wm m_gm + k_glm_wm m_csf) of one member */
int glia_rd_set_secondary(glia_rd_t h, int member, double kt);
/* DerivativeOperators::gradDiffusion / gradReaction (src/grad/DerivativeOperators.cpp:189-321),
   nk = nr = 1: g_kappa = h^3 sum_i w_i dt <m_wm + k_gm_wm m_gm, grad c_i . grad alpha_i> and
   g_rho = h^3 sum_i w_i dt <m_wm + r_gm_wm m_gm, (c_i^2 - c_i) alpha_i> over the history of the
   last solveState(0) / solveAdjoint; one value per member */
int glia_rd_grad_diffusion(glia_rd_t h, double* gk);
int glia_rd_grad_reaction(glia_rd_t h, double* gr);
/* DerivativeOperators::checkGradient (src/grad/DerivativeOperators.cpp:369-450): the Taylor test
   at x along dir (laid out like x; NULL: entries uniform in [0, 1) from a fixed seed, as the
   reference's VecSetRandom draws them): with J0, dJ the objective and gradient at x and
   J_i = J(x + h_i dir), h_i = 10^-i for i = 0..5, err[6 m + i] = |J_i - J0 - h_i <dJ, dir>| and
   gdiff[6 m + i] = (J_i - J0) / h_i - <dJ, dir> for member m (either output may be NULL). */
int glia_rd_check_gradient(glia_rd_t h, int mode, const double* x, const double* dir, double* err, double* gdiff);

/* DerivativeOperatorsRD::evaluateObjective / evaluateGradient / evaluateObjectiveAndGradient
   (src/grad/DerivativeOperatorsRD.cpp:7-226).  x holds [p (np) | kappa | rho] per member,
   concatenated over members; J one value per member; dJ like x. */
int glia_rd_objective(glia_rd_t h, int mode, const double* x, double* J);
int glia_rd_gradient(glia_rd_t h, int mode, const double* x, double* dJ);
int glia_rd_objective_and_gradient(glia_rd_t h, int mode, const double* x, double* J, double* dJ);
/* DerivativeOperatorsRD::evaluateHessian with L2 regularization, nk = 1
   (src/grad/DerivativeOperatorsRD.cpp:229-438): the Gauss-Newton Hessian applied to a direction
   x = [p~ | kappa~ | rho~] per member (laid out like dJ).
   diffusivity = 0 (:420-436): y_p = beta h^3 Phi^T Phi p~ - h^3 Phi^T a~0 with the tangent-linear
     forward (solveState(1), src/pde/PdeOperators.cpp:235-316) and incremental adjoint
     (solveAdjoint(2), :372-420); kappa~ is ignored and y_kappa = 0.
   diffusivity = 1 (:244-418): y = [Hpp p~ + Hpk kappa~ ; Hkp p~ + Hkk kappa~]; the kappa row is the
     gradDiffusion quadrature after each incremental adjoint, including -- as in the reference --
     the i = nt term with the adjoint history entry that solveAdjoint(2) leaves from the last
     gradient; the kappa~ part is solveState(2) with the secondary sources of solveIncremental.
   y_rho = 0 and rho~ is ignored.  The linearization point is the state history of the last
   evaluation on this handle; the call overwrites c(0) with Phi p~. */
int glia_rd_hessian(glia_rd_t h, int diffusivity, const double* x, double* y);
/* The same evaluation for several handles at once, which may have different grids (an ensemble
   of uneven sizes, e.g. 128^3, 192^3 and 256^3 members).  Each handle is evaluated by its own
   host thread on its own stream, so the GPU overlaps them; handles must be distinct and should
   not share a stream.  x[i], J[i], dJ[i] are laid out as for glia_rd_objective_and_gradient on
   handles[i]; ms (optional, nh entries) receives each handle's wall time in milliseconds.  Returns
   the first nonzero code in handle order.  No single reference counterpart: it batches
   DerivativeOperatorsRD::evaluateObjectiveAndGradient (src/grad/DerivativeOperatorsRD.cpp:130-226)
   over independent problems. */
int glia_rd_objective_and_gradient_group(int nh, const glia_rd_t* handles, const int* modes, const double* const* x,
                                         double* const* J, double* const* dJ, double* ms);
/* The same evaluation for several store-all handles in lockstep, from one host thread: time step i
   of every handle (forward, then adjoint) is one graph launch holding each handle's captured step
   as an independent branch, where every handle has its step captured (from its second evaluation
   on; the first steps of a handle run eagerly).  A step whose CG solve needs more iterations than
   its capture holds runs again eagerly on that handle alone, so J and dJ equal those of
   glia_rd_objective_and_gradient on each handle bit for bit.  Handles with different nt are
   aligned at the start (forward) and at the end (adjoint).  Checkpointed handles are evaluated one
   after another.  spanned_steps (may be NULL) receives the number of steps launched as one graph
   across handles.  After an error the handles hold a partial evaluation.  No single reference
   counterpart (it batches evaluateObjectiveAndGradient, src/grad/DerivativeOperatorsRD.cpp:130-226,
   over independent problems). */
int glia_rd_objective_and_gradient_lockstep(int nh, const glia_rd_t* handles, const int* modes,
                                            const double* const* x, double* const* J, double* const* dJ,
                                            long long* spanned_steps);

int glia_rd_stats(glia_rd_t h, int member, glia_rd_stats_t* out);
int glia_rd_reset_stats(glia_rd_t h);

/* kernel timing (the reference's Event timers, 3rdparty/timings/EventTimings.hpp:31-140, for the
   CG kernels): with profiling on, HIP events on the handle's stream bracket every launch of the
   kinds below; glia_rd_profile_get returns launches and summed milliseconds since enabling. */
enum {
  GLIA_RD_K_XI_PUPD = 0,   /* X pass: p update + inverse transforms of p and i w_x p */
  GLIA_RD_K_YI_GRAD = 1,   /* Y pass: inverse, three gradient components */
  GLIA_RD_K_Z_DIFF = 2,    /* Z pass: C2R, k * grad, R2C for three components */
  GLIA_RD_K_YF_DIV = 3,    /* Y pass: forward, partial divergence */
  GLIA_RD_K_XF_MATVEC = 4, /* X pass: forward, w = p - tau/2 D p, <p, w> */
  GLIA_RD_K_PCG_UPDATE = 5 /* x, r update with ||B r||^2 and <B r, r> */
};
/* Replay each CG iteration from a captured HIP graph (default on; one graph per solved slot and
   flag slot, built on first use).  Profiling runs always launch kernel by kernel.  No reference
   counterpart: the reference launches every kernel from the host. */
int glia_rd_use_graphs(glia_rd_t h, int enable);
/* Autotuning of the kernel variants (no reference counterpart: the reference leaves plans to
   cuFFT): the variants of each forward divergence pass are timed on this handle's grid and
   members and the fastest is kept.  X pass: 0 = one thread group holding both fields, 1 = one
   group per field, 2 = one group per field and one for the output.  Y pass: 0 = one group
   holding three fields, 1 = one group per field, 2 and 3 = the fields split over blocks at 3 and
   2 waves per SIMD.  Untuned handles take X pass 1 on power-of-two n0, else 0, and Y pass 0.
   cache_path (may be NULL) names a text
   file of earlier choices, read first and appended to.  choice (2 ints, may be NULL) receives
   (X-pass variant, Y-pass variant).  Call before the first solve; the results do not change
   beyond FP32 rounding. */
int glia_rd_autotune(glia_rd_t h, const char* cache_path, int* choice);
/* Time steps replayed from captured graphs (glia_rd_use_graphs on, store-all history): each
   forward and adjoint step runs as one graph whose CG loops are unrolled to the last iteration
   counts plus one, and the host reads each solve's end-of-solve flag once per step.  A step
   whose solve needed more iterations runs again eagerly from its stored start.  launched and
   reruns count such steps since creation, cached the graphs held. */
int glia_rd_step_info(glia_rd_t h, long long* launched, long long* reruns, int* cached);
int glia_rd_profile(glia_rd_t h, int enable);
int glia_rd_profile_get(glia_rd_t h, int kind, double* ms, long long* launches);

#ifdef __cplusplus
}
#endif

#endif /* GLIA_RD_H_ */
