/*
 * glia_el.h -- C ABI of the linear-elasticity solver of the mass-effect model: the reference's
 * VariableLinearElasticitySolver (include/pde/ElasticitySolver.h, src/pde/ElasticitySolver.cpp:6-533)
 * with its CUDA kernels (src/cuda/ElasticitySolver.cu).
 *
 * The displacement u solves A u = f with
 *   A u = div(mu (grad u + grad u^T)) + grad(lam div u) - screen u,
 * spectral derivatives with the Nyquist modes zeroed, by preconditioned CG (PETSc KSPCG semantics:
 * preconditioned residual norm, rtol 1e-3 relative to ||B f||, atol 1e-50, dtol 1e5, at most 100
 * iterations, the given u as initial guess) with the exact inverse of the operator for the mean
 * mu, lam and screen as preconditioner B.  A and B are both negative definite.
 *
 * Conventions are those of glia_rd.h: FP32 fields of n0*n1*n2 values with z fastest, mem_kind
 * GLIA_RD_HOST or GLIA_RD_DEVICE, return 0, a negative GLIA_RD_ERR_* or a positive hipError_t, and
 * glia_el_last_error() for the message.  Vector fields are three component arrays.
 */
#ifndef GLIA_EL_H_
#define GLIA_EL_H_

#include "glia_rd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glia_el_s* glia_el_t;

/* Young's moduli, Poisson ratios and screening coefficients (TumorParameters,
   include/Parameters.h:175-184) */
typedef struct glia_el_params {
  double E_healthy, nu_healthy; /* wm and gm: 2100, 0.4 */
  double E_tumor, nu_tumor;     /* 8000, 0.45 */
  double E_bg, nu_bg;           /* 15000, 0.45 */
  double E_csf, nu_csf;         /* vt and csf: 100, 0.1 */
  double screen_low, screen_high; /* 1e-2 where c >= 0.005, 1e2 elsewhere (1e6 where bg > 0.95) */
} glia_el_params;

void glia_el_params_init(glia_el_params* p);

/* ElasticitySolver ctor (src/pde/ElasticitySolver.cpp:6-67): a solver on an n0 x n1 x n2 grid (the
   grids glia_rd_create accepts) with its device fields and a one-member spectral engine.  stream:
   a hipStream_t, or NULL for a private non-blocking stream at default priority. */
int glia_el_create(const int n[3], void* stream, glia_el_t* out);
int glia_el_destroy(glia_el_t h);
const char* glia_el_last_error(glia_el_t h);

/* computeMaterialProperties (src/pde/ElasticitySolver.cpp:376-455): mu, lam and screen from the
   tumor c and the tissue maps, and their means (the preconditioner's constants) */
int glia_el_set_material(glia_el_t h, const float* c, const float* wm, const float* gm, const float* csf,
                         const float* vt, const float* bg, const glia_el_params* p, int mem_kind);
/* mu, lam, screen (any may be NULL) and the means (avg[3], may be NULL) */
int glia_el_get_material(glia_el_t h, float* mu, float* lam, float* screen, double* avg, int mem_kind);
/* operatorVariableCoefficients (src/pde/ElasticitySolver.cpp:268-374): f = A u */
int glia_el_apply(glia_el_t h, const float* const u[3], float* const f[3], int mem_kind);
/* operatorConstantCoefficients (src/pde/ElasticitySolver.cpp:122-266): u = B f */
int glia_el_precondition(glia_el_t h, const float* const f[3], float* const u[3], int mem_kind);
/* Tumor::computeForce (src/mat/Tumor.cpp:114-158): f = forcing_factor grad(S c), times tanh(S c)
   when use_tanh (use_tanh_scaling_), with S the Weierstrass smoother of sigma 2 pi / n0 and the
   spectral gradient.  c is one field; f three. */
int glia_el_force(glia_el_t h, const float* c, double forcing_factor, int use_tanh, float* const f[3], int mem_kind);
/* VariableLinearElasticitySolver::solve (src/pde/ElasticitySolver.cpp:471-533) after
   glia_el_set_material: u (initial guess in, solution out) for the force f.  iterations, reason
   (KSPConvergedReason values: 2 rtol, 3 atol, -3 max iterations, -4 dtol, -8 indefinite
   preconditioner, -9 NaN/Inf, -10 indefinite matrix) and rnorm (last preconditioned residual norm)
   may be NULL. */
int glia_el_solve(glia_el_t h, const float* const f[3], float* const u[3], int mem_kind, int* iterations, int* reason,
                  double* rnorm);

#ifdef __cplusplus
}
#endif

#endif /* GLIA_EL_H_ */
