/*
 * glia_me.h -- C ABI of the mass-effect forwards: model 4, PdeOperatorsMassEffect::solveState
 * (src/pde/PdeOperatorsMassEffect.cpp:421-743), the coupled tumor growth, elasticity and tissue
 * advection; and model 5, PdeOperatorsMultiSpecies::solveState
 * (src/pde/PdeOperatorsMultiSpecies.cpp:235-508), the same coupling with proliferative,
 * infiltrative and necrotic species and oxygen.
 *
 * A glia_me_t owns one diffusion engine (glia_rd_t, one member), one elasticity solver (glia_el_t)
 * and one advection solver (glia_sl_t) on one stream.  The engine carries what the reference keeps
 * in TumorParameters and DiffCoef/ReacCoef: dt and nt (glia_rd_config), k and rho with their gm/wm
 * ratios, and the preconditioner's mean k, fixed by glia_rd_set_tissue + glia_rd_set_coefficients
 * on the initial tissue as DiffCoef::setValues fixes it at initialization (the loop rewrites the k
 * and rho fields but not that mean).
 *
 * Conventions are those of glia_rd.h: FP32 fields of n0*n1*n2 values with z fastest, return 0, a
 * negative GLIA_RD_ERR_* or a positive hipError_t, and glia_me_last_error() for the message.  The
 * fields of glia_me_solve_state are device arrays.
 */
#ifndef GLIA_ME_H_
#define GLIA_ME_H_

#include "glia_adv.h"
#include "glia_el.h"
#include "glia_ms.h"
#include "glia_rd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glia_me_s* glia_me_t;

/* the mass-effect parameters of TumorParameters (include/Parameters.h) that the loop reads */
typedef struct glia_me_params {
  double forcing_factor; /* forcing_factor_ (gamma); 1e5 */
  int use_tanh;          /* use_tanh_scaling_; 1 */
  double max_cfl;        /* the loop's CFL limit; <= 0 (the default) selects the reference's limit of
                            the model: 14 for model 4 (src/pde/PdeOperatorsMassEffect.cpp:472), 8
                            for model 5 (src/pde/PdeOperatorsMultiSpecies.cpp:296) */
  glia_el_params el;     /* Young's moduli, Poisson ratios, screening */
} glia_me_params;

void glia_me_params_init(glia_me_params* p);

typedef struct glia_me_info {
  int steps;          /* time steps completed (nt + 1 unless the CFL test stopped the loop) */
  int cfl_stop;       /* 1 when CFL = dt max|v| / h0 reached max_cfl at the top of a step */
  double cfl_max;     /* largest CFL seen */
  int el_iterations;  /* elasticity CG iterations over all solves */
  int diff_iterations;/* diffusion CG iterations over all steps */
  int volume_stop;    /* 1 when the feature branch stopped the loop: the tumor core reached about 5%
                         of the brain (ratio within 5e-3 of 0.05, or above) at t >= 1 */
} glia_me_info;

/* the engine (glia_rd_create with cfg->nmembers = 1; cfg->stream, or the engine's own stream, is
   shared by the elasticity and advection solvers) and the two other solvers.  As with
   glia_rd_create, a failed creation still sets *out (unless cfg or out is NULL or nmembers != 1)
   so that glia_me_last_error can name the cause; destroy it either way.  The solve entry points
   answer such a partly created handle with GLIA_RD_ERR_STATE. */
int glia_me_create(const glia_rd_config* cfg, glia_me_t* out);
int glia_me_destroy(glia_me_t h);
const char* glia_me_last_error(glia_me_t h);
/* the owned handles, valid until glia_me_destroy: the engine takes the tissue and coefficients */
glia_rd_t glia_me_rd(glia_me_t h);
glia_el_t glia_me_el(glia_me_t h);
glia_sl_t glia_me_sl(glia_me_t h);

/* PdeOperatorsMassEffect::solveState(0) without its output and feature branches (the feature
   branch is glia_me_solve_state_features).
   c: Tumor::c_0_ in, c_t_ out.  wm, gm, vt, csf: MatProp's tissue in, filtered (filterTumor) and
   advected out; mri may be NULL (else advected too); bg: MatProp::bg_, fixed.
   u: Tumor::displacement_ (initial guess of the first step's elasticity solve in, the last
   displacement out); v: Tumor::velocity_ (in and out; zero for a fresh tumor).
   Before the loop: filterTumor, computeForce and the elasticity solve into u_old from zero.  Each
   of the nt + 1 steps: the CFL test (stop when dt max|v| / h0 >= max_cfl), k and rho from the
   tissue (updateReacAndDiffCoefficients, :98-138), precFactor, semi-Lagrangian steps of gm, wm
   (mass conservation), vt, csf (pure advection), c (mass conservation) and mri (pure advection) on
   one set of RK2 trajectories, the diffusion solve of c over dt, the logistic reaction, the force,
   the elasticity solve into u, v = S((u - u_old) / dt) with S the smoother of sigma 2 pi / n0, and
   u_old = u.  info may be NULL. */
int glia_me_solve_state(glia_me_t h, const glia_me_params* p, float* c, float* wm, float* gm, float* vt, float* csf,
                        float* mri, const float* bg, float* const u[3], float* const v[3], glia_me_info* info);

/* The columns of a feature row: GLIA_ME_NFEATURES values, named as the reference's CSV header
   (biophysical_features.csv) names them. */
#define GLIA_ME_NFEATURES 130
/* name of column k (0 <= k < GLIA_ME_NFEATURES), else NULL */
const char* glia_me_feature_name(int k);
/* glia_me_solve_state with feature_compute_ on (src/pde/PdeOperatorsMassEffect.cpp:584-604): at
   every step, after the coefficient update and before the advection,
   computeBiophysicalFeatures (:323-418) writes row i of features (row-major, GLIA_ME_NFEATURES
   doubles per row, max_rows >= nt + 1; *rows receives the rows written):
     t = i dt, then tc / (tc + healthy) with tc the voxels labelled 1 by computeSegmentation and
     healthy those labelled neither 0 nor 1; then for each of c, dc/dt = D c + rho (c - c^2) -
     div(c v), |grad c|, |u|, det(I + grad u), tr S, the maximum shear stress (S = lam tr(E) I +
     2 mu E, E the linearized strain, mu and lam the elasticity solver's current material) and |v|,
     16 statistics: the integrals over the grid, tc and healthy; the surface integral
     (n0 / 2 pi) h^3 sum I x with I = [|c - c*| < 0.1], c* the value of c where |grad c| is largest
     outside tc; the standard deviations (unbiased, two-pass) over the grid, tc and healthy; and
     the 0.875, 0.95 and 0.99 quantiles (entry floor(q m) of the m sorted values) over the grid,
     tc and healthy, in the reference's column order.  An empty region's quantiles are 0.
   After the features of step i, the loop stops (info->volume_stop) when the ratio is within 5e-3
   of 0.05 or above and i dt >= 1.  With features NULL this is glia_me_solve_state.  The feature
   computation uses the engine's alpha(0) field as scratch. */
int glia_me_solve_state_features(glia_me_t h, const glia_me_params* p, float* c, float* wm, float* gm, float* vt,
                                 float* csf, float* mri, const float* bg, float* const u[3], float* const v[3],
                                 glia_me_info* info, double* features, int max_rows, int* rows);

/* PdeOperatorsMultiSpecies::solveState(0) (src/pde/PdeOperatorsMultiSpecies.cpp:235-508), the
   multispecies forward of model 5 without its file output branch; the edema and segmentation it
   computes at every step are read with glia_me_multispecies_outputs.  The reference's CFL limit for this model is 8 (p->max_cfl <= 0 selects it).
   c0: Tumor::c_0_ (in); prolif, infil, ox: the proliferative, infiltrative and oxygen species
   (out; set to c0 clipped at zero, 0 and 1 at the start); nec: the necrotic species (in and out,
   not reset, as in the reference); wm, gm, vt, csf: the tissue (in, filtered and advected out);
   bg fixed; u, v: Tumor::displacement_ and velocity_ (in and out).
   Start (:251-287): clipTumor, filterTumor with p + i, and when forcing_factor > 0 the force of p
   and the elasticity solve into u, u_old = u.  Each of the nt + 1 steps: the CFL test; rho from the
   tissue (updateReacAndDiffCoefficients); k from the current wm, gm and csf with its mean over the
   initial brain mask (updateIsotropicCoefficients, DiffCoef::setValues); precFactor; when
   forcing_factor > 0 the advection of gm, wm, p, i, nec (mass conservation), vt and csf (pure
   advection); clipHealthyTissues and clipTumor; i smoothed with sigma 2 pi / n0; di = D i; the
   diffusion solve of i over dt; computeSources (glia_ms_sources); c = p + nec + i; when
   forcing_factor > 0 the force of c, the elasticity solve, v = S((u - u_old) / dt), u_old = u.
   info may be NULL. */
int glia_me_solve_state_multispecies(glia_me_t h, const glia_me_params* p, const glia_ms_params* ms, const float* c0,
                                     float* prolif, float* infil, float* nec, float* ox, float* wm, float* gm, float* vt,
                                     float* csf, const float* bg, float* const u[3], float* const v[3],
                                     glia_me_info* info);
/* The outputs the multispecies loop computes at the top of every step (:307-308), as they stand
   after the last solve: Tumor::species_["edema"] (computeEdema, src/mat/Tumor.cpp:160-178: the mask
   1[i > ms->invasive_threshold] smoothed with sigma 2 pi / n0) and Tumor::seg_
   (computeSegmentation, labels of glia_ms_segmentation with c_t_ = p at the first step, then
   p + nec + i).  Either output may be NULL; GLIA_RD_ERR_STATE before a multispecies solve. */
int glia_me_multispecies_outputs(glia_me_t h, float* edema, float* seg, int mem_kind);

#ifdef __cplusplus
}
#endif

#endif /* GLIA_ME_H_ */
