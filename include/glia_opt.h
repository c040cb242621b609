/*
 * glia_opt.h -- the inversion loop around the gradient: a bound-constrained limited-memory
 * quasi-Newton method with an Armijo line search (the reference drives TAO's BLMVM with the
 * Armijo line search, src/opt/Optimizer.cpp:116-260, and stops with its own test,
 * checkConvergenceGrad, src/opt/TaoInterface.cpp:362-520).
 *
 * PETSc/TAO is not available, so this is a restatement of the published BLMVM algorithm, not a
 * bitwise reproduction of TAO: same inputs, options and stopping rule, own quasi-Newton matrix
 * (two-loop L-BFGS with the scalar initial scaling s'y / y'y).
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

typedef struct glia_opt_settings {
  int max_it;         /* OptimizerSettings::newton_maxit_ (30) */
  int min_it;         /* OptimizerSettings::newton_minit_ (0) */
  double gatol;       /* gatol_ (1e-4): ||g|| < gatol */
  double gttol;       /* opttolgrad_ (1e-3): ||g|| < gttol ||g_0|| */
  int lbfgs_vectors;  /* lbfgs_vectors_ (10) */
  int ls_max_funcs;   /* ls_max_func_evals (10) */
  double ls_sigma;    /* Armijo sufficient-decrease constant (1e-4) */
  double ls_minstep;  /* ls_minstep_ (machine epsilon): a smaller accepted step stops (STEPTOL) */
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

/* The inversion of a handle: every member minimizes J over x = [p | kappa | rho] of its own
   (mode as for glia_rd_objective_and_gradient; the entries the mode does not invert keep their
   initial values since their gradient is 0).  All members need the same number of bases. */
int glia_rd_invert(glia_rd_t h, int mode, double* x, const double* lb, const double* ub, const glia_opt_settings* s,
                   glia_opt_result* r);

#ifdef __cplusplus
}
#endif

#endif
