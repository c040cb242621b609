/*
 * glia_opt.h -- the inversion loop around the gradient: a bound-constrained limited-memory
 * quasi-Newton method with an Armijo line search (the reference drives TAO's BLMVM with the
 * Armijo line search, src/opt/Optimizer.cpp:116-260, and stops with its own test,
 * checkConvergenceGrad, src/opt/TaoInterface.cpp:362-520), and the CoSaMP sparse-TIL driver
 * around it (src/opt/SparseTILOptimizer.cpp).
 *
 * PETSc/TAO is not available, so this is a restatement of the published BLMVM algorithm and of the
 * More-Thuente line search, not a bitwise reproduction of TAO: same inputs, options and stopping
 * rule, own quasi-Newton matrix (two-loop L-BFGS with the scalar initial scaling s'y / y'y).
 */
#ifndef GLIA_OPT_H_
#define GLIA_OPT_H_

#include "glia_rd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* TaoConvergedReason values */
enum {
  GLIA_OPT_CONTINUE = 0,
  GLIA_OPT_CONVERGED_GATOL = 3,
  GLIA_OPT_CONVERGED_GTTOL = 5,
  GLIA_OPT_CONVERGED_STEPTOL = 6,
  GLIA_OPT_DIVERGED_MAXITS = -2,
  GLIA_OPT_DIVERGED_NAN = -4,
  GLIA_OPT_DIVERGED_LS_FAILURE = -6
};

/* line searches (OptimizerSettings::linesearch_) */
enum { GLIA_OPT_LS_ARMIJO = 0, GLIA_OPT_LS_MORE_THUENTE = 1 };
/* Newton solvers, numbered as the reference's GAUSSNEWTON and QUASINEWTON
   (include/utils/TypeDefs.h:12, OptimizerSettings::newton_solver_), and Krylov forcing sequences
   numbered as QDFS and SLFS (:10, fseqtype_) */
enum { GLIA_OPT_GAUSS_NEWTON = 0, GLIA_OPT_QUASI_NEWTON = 1 };
enum { GLIA_OPT_FORCING_QUADRATIC = 0, GLIA_OPT_FORCING_SUPERLINEAR = 1 };

typedef struct glia_opt_settings {
  int max_it;         /* OptimizerSettings::newton_maxit_ (30) */
  int min_it;         /* OptimizerSettings::newton_minit_ (0) */
  double gatol;       /* gatol_ (1e-4): ||g|| < gatol */
  double gttol;       /* opttolgrad_ (1e-3): ||g|| < gttol ||g_0|| */
  int lbfgs_vectors;  /* lbfgs_vectors_ (10) */
  int ls_max_funcs;   /* ls_max_func_evals (10) */
  double ls_sigma;    /* Armijo sufficient-decrease constant (1e-4) */
  double ls_minstep;  /* ls_minstep_ (machine epsilon): a smaller accepted step stops (STEPTOL) */
  int line_search;    /* linesearch_: GLIA_OPT_LS_MORE_THUENTE (the reference's default) or _ARMIJO */
  double ls_gtol;     /* curvature constant of the More-Thuente search (0.9) */
  double ls_xtol;     /* relative interval tolerance of the More-Thuente search (1e-10) */
  double gnorm_ref;   /* > 0: the reference gradient norm of the relative test (gradnorm0_) for
                         every problem; 0 (default): each problem's first projected gradient norm */
  int newton_solver;  /* newton_solver_: GLIA_OPT_QUASI_NEWTON (BLMVM, default) or _GAUSS_NEWTON */
  int krylov_maxit;   /* krylov_maxit_ (30): CG iterations per Gauss-Newton step */
  int forcing;        /* fseqtype_: CG tolerance max(1e-10, min(0.5, sqrt(||g_F|| / ||g_F,0||)))
                         (superlinear, default) or without the square root (quadratic) */
} glia_opt_settings;

typedef struct glia_opt_result {
  int iterations;     /* accepted quasi-Newton steps */
  int fevals;         /* objective/gradient evaluations */
  int reason;         /* GLIA_OPT_* */
  double f, gnorm, gnorm0;
} glia_opt_result;

void glia_opt_settings_init(glia_opt_settings* s);

/* Objective and gradient of nprob problems of n unknowns each, at x (nprob*n): f (nprob), g
   (nprob*n).  Returns 0, or nonzero to abort the minimization with that code. */
typedef int (*glia_fg_fn)(void* ctx, int nprob, const double* x, double* f, double* g);

/* Minimize nprob independent problems together: every evaluation round asks fg for all of them
   (problems that have stopped are evaluated at their final point).  x (nprob*n) holds the initial
   guesses and receives the results; lb/ub (nprob*n, NULL = unbounded) are the bounds; r (nprob,
   optional) receives the outcome of each problem.  Returns 0 or the first nonzero fg code. */
int glia_blmvm(int nprob, int n, double* x, const double* lb, const double* ub, glia_fg_fn fg, void* ctx,
               const glia_opt_settings* s, glia_opt_result* r);

/* Hessian-vector products of nprob problems at the point of the last fg call: y = H v (v, y:
   nprob*n).  Returns 0, or nonzero to abort with that code. */
typedef int (*glia_hv_fn)(void* ctx, int nprob, const double* v, double* y);

/* Bound-constrained Gauss-Newton-Krylov (the reference's GAUSSNEWTON solver, TAO's BNLS with a CG
   Krylov solve and the Eisenstat-Walker tolerance of preKrylovSolve, src/opt/TaoInterface.cpp:
   160-193): per iteration the variables on an active bound are held, H_FF s_F = -g_F is solved by
   CG (at most krylov_maxit iterations, the forcing tolerance of the settings), a non-descent
   direction is replaced by the scaled steepest descent, and the settings' line search runs along
   it; the stopping test is the one of glia_blmvm.  Batched like glia_blmvm: every CG iteration is
   one hv call and every line-search round one fg call for all problems. */
int glia_gauss_newton(int nprob, int n, double* x, const double* lb, const double* ub, glia_fg_fn fg, glia_hv_fn hv,
                      void* ctx, const glia_opt_settings* s, glia_opt_result* r);

/* The inversion of a handle: every member minimizes J over x = [p | kappa | rho] of its own
   (mode as for glia_rd_objective_and_gradient; the entries the mode does not invert keep their
   initial values since their gradient is 0).  All members need the same number of bases.
   GLIA_RD_MODE_KAPPA_RHO is RDOptimizer::solve (src/opt/RDOptimizer.cpp:145-250): beta is 0 during
   the solve and the relative test is against the first gradient.  The p modes are
   TILOptimizer::solve (src/opt/TILOptimizer.cpp:114-230): the relative test is against the
   gradient at x = 0 (src/opt/TaoInterface.cpp:285-300), taken over the entries with lb < ub (the
   others keep their values).  With newton_solver = GLIA_OPT_GAUSS_NEWTON, mode
   GLIA_RD_MODE_P_KAPPA runs glia_gauss_newton on glia_rd_hessian, with the kappa blocks unless every
   member's kappa is fixed (lb >= ub); the other modes return GLIA_RD_ERR_UNSUPPORTED (the Hessian
   has no rho row). */
int glia_rd_invert(glia_rd_t h, int mode, double* x, const double* lb, const double* ub, const glia_opt_settings* s,
                   glia_opt_result* r);

/* hardThreshold (src/utils/Utils.cpp:339-367): indices of the s largest entries of x[0..n) by
   signed value (ties: larger index first), stopping at the first entry with |x| = 0.  support
   receives them in that order; returns their count. */
int glia_hard_threshold(const double* x, int n, int s, int* support);
/* hardThreshold by connected component (src/utils/Utils.cpp:370-463): labels[i] in 1..ncomp,
   weights[ncomp]; components of weight >= thresh get 3 + floor(w (s - 3 ncc - ncc_t)) entries,
   those of weight in [thresh/10, thresh) get 1, the last component the rest; double_mode halves
   s for that split and doubles every share (the CoSaMP gradient step).  Returns the count. */
int glia_hard_threshold_components(const double* x, int n, int s, const int* labels, const double* weights,
                                   int ncomp, double thresh, int double_mode, int* support);

/* The operations sparse TIL needs from an RD model of nprob problems, x = [p | kappa | rho] per
   problem (np + 2 entries):
   eval: J (nprob) and, when g != NULL, dJ (nprob*(np+2)) at x with the mode of
         glia_rd_objective_and_gradient and regularization beta;
   set_c0: c0 = Phi p for every problem (p: nprob*np); cmax (nprob) receives max c0 over the grid. */
typedef int (*glia_eval_fn)(void* ctx, int mode, double beta, const double* x, double* f, double* g);
typedef int (*glia_c0_fn)(void* ctx, const double* p, double* cmax);

typedef struct glia_sparse_til_settings {
  int sparsity;                    /* sparsity_level_ (5) */
  int max_it;                      /* gist_maxit_ (5): CoSaMP iterations */
  double ftol;                     /* 1e-5: |J_old - J| < ftol |1 + J_ref| stops */
  int prune_components;            /* prune_components_ (1): component-wise pruning when labels are given */
  double thresh_component_weight;  /* thresh_component_weight_ (1e-3) */
  int diffusivity_inversion;       /* diffusivity_inversion_ (1): kappa is an unknown of the TIL solves */
  int reaction_inversion;          /* reaction_inversion_ (0): final kappa/rho solve */
  double beta;                     /* beta_p of the TIL solves */
  double k_lb, k_ub, rho_lb, rho_ub; /* OptimizerSettings bounds (1e-3, 1, 1, 15) */
  glia_opt_settings til, rd;       /* settings of the TIL and kappa/rho solves */
} glia_sparse_til_settings;

typedef struct glia_sparse_til_result {
  int iterations;           /* CoSaMP iterations */
  int stop;                 /* 1: max_it, 2: |J| < 1e-5, 3: relative objective change */
  int nnz;                  /* size of the final support */
  double J;                 /* mismatch (beta = 0) after the last CoSaMP iteration */
  glia_opt_result til, rd;  /* the final TIL solve and the kappa/rho solve */
} glia_sparse_til_result;

void glia_sparse_til_settings_init(glia_sparse_til_settings* s);

/* SparseTILOptimizer::solve (src/opt/SparseTILOptimizer.cpp:220-595) for nprob problems in
   lockstep.  Per problem: the support starts as the nonzero entries of the initial p (the
   multilevel warm start of src/Solver.cpp:285-305); each CoSaMP iteration thresholds |g_p| of the
   beta = 0 gradient to 2s entries, merges them into the support, runs the TIL solve restricted to
   the support (p >= 0 on it, 0 off it; kappa in [k_lb, k_ub] when inverted), prunes p to s entries
   and stops on max_it, |J| < 1e-5 or the relative change.  A final TIL solve follows; with
   reaction_inversion, p is rescaled so that max Phi p = 1, rho starts at the best of
   {0, 3, 6, 9, 10, 12, 15} and kappa/rho are inverted with c0 fixed.  The restricted solves keep
   the full basis with the off-support coefficients fixed at 0, which is the same problem.
   x (nprob*(np+2)) holds the initial guess and receives the result; support (nprob*np, optional)
   receives each final support, padded with -1; labels (nprob*np) and weights (nprob*ncomp) are
   optional. */
int glia_sparse_til(int nprob, int np, double* x, int* support, const int* labels, const double* weights, int ncomp,
                    glia_eval_fn eval, glia_c0_fn set_c0, void* ctx, const glia_sparse_til_settings* s,
                    glia_sparse_til_result* r);
/* glia_sparse_til on the members of a handle (all with the same number of bases).  The TIL
   stages use rho = x[np + 1] of each member; the handle's beta is restored at the end. */
int glia_rd_sparse_til(glia_rd_t h, double* x, int* support, const int* labels, const double* weights, int ncomp,
                       const glia_sparse_til_settings* s, glia_sparse_til_result* r);

/* The optimizer keys of a reference config file (setParameter, src/utils/IO.cpp:82-100, 171-172):
   newton_solver (GN / QN), line_search (armijo / mt), opttol_grad, newton_maxit, krylov_maxit,
   lbfgs_vectors, ls_max_func_evals into *settings and into sparse->til and sparse->rd; gist_maxit,
   kappa_lb/ub, rho_lb/ub, beta_p, invert_diff, invert_reac, sparsity_level and
   thresh_component_weight into *sparse.  Values the file does not set keep what the structs hold
   (initialize them with the *_init functions first); either pointer may be NULL.  A missing file
   or an unreadable value is GLIA_RD_ERR_ARG. */
int glia_opt_parse_config(const char* path, glia_opt_settings* settings, glia_sparse_til_settings* sparse);

#ifdef __cplusplus
}
#endif

#endif
