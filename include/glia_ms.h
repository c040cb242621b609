/*
 * glia_ms.h -- C ABI of the pointwise terms of the multispecies model (PdeOperatorsMultiSpecies,
 * src/pde/PdeOperatorsMultiSpecies.cpp:10-234, with the kernels of src/cuda/PdeOperators.cu:27-124):
 * the reaction/transition/death sources of proliferative (p), infiltrative (i) and necrotic (n) cells,
 * oxygen and the healthy tissues, and the tissue-dependent reaction and diffusion coefficients.
 *
 * Arrays are FP32 device arrays of n values (one field each), updated in place; calls are ordered
 * on the given hipStream_t (NULL: the null stream) and return 0, GLIA_RD_ERR_ARG or a hipError_t.
 */
#ifndef GLIA_MS_H_
#define GLIA_MS_H_

#include <stdint.h>

#include "glia_rd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* TumorParameters of the multispecies model (include/Parameters.h:188-194) */
typedef struct glia_ms_params {
  double death_rate;     /* 4 */
  double ox_source;      /* 1 */
  double ox_consumption; /* 5 */
  double alpha_0;        /* 0.1: proliferative -> infiltrative */
  double beta_0;         /* 0.01: infiltrative -> proliferative */
  double ox_inv;         /* 0.8 */
  double ox_hypoxia;     /* 0.6 */
  double invasive_threshold; /* 0.001: TumorParameters::invasive_threshold_, the edema mask level */
} glia_ms_params;

void glia_ms_params_init(glia_ms_params* p);

/* computeSources (src/pde/PdeOperatorsMultiSpecies.cpp:111-190) with its rate fields
   (computeReactionRate, computeTransition, computeThesholder: src/cuda/PdeOperators.cu:47-78), in
   one pass over the grid:
     m = rho / (1 + exp(-100 (ox - ox_hypoxia))), alpha = alpha_0 / (1 + exp(100 (ox - ox_inv))),
     beta = beta_0 ox, h = 1 / (1 + exp(100 (ox - ox_hypoxia))),
   then one explicit step of p, i, nec, ox (clipped at 0), gm and wm (healthy cells give way to the
   tumor growth and to di = D(i), the diffusion of the infiltrative cells before their solve).
   rho is the reaction-coefficient field. */
int glia_ms_sources(int64_t n, float* p, float* i, float* nec, float* ox, float* gm, float* wm, const float* rho,
                    const float* di, double dt, const glia_ms_params* prm, void* stream);

/* updateReacAndDiffCoefficients (src/pde/PdeOperatorsMultiSpecies.cpp:193-234,
   src/cuda/PdeOperators.cu:113-124): rho_field = rho max(0, 1 - (bg + (1 - r_gm_wm) gm + vt + csf)),
   k_field = k max(0, 1 - (bg + (1 - k_gm_wm) gm + vt + csf)). */
int glia_ms_update_coefficients(int64_t n, float* rho_field, float* k_field, const float* bg, const float* gm,
                                const float* vt, const float* csf, double rho, double k, double r_gm_wm,
                                double k_gm_wm, void* stream);

/* Tumor::computeSegmentation (src/mat/Tumor.cpp:180-276, computeTumorSegmentation
   src/cuda/Tumor.cu:14-33): the label of the largest of bg (0), c (1), wm (6), gm (5), vt (7), csf (8),
   the tissues attenuated by (1 - c), first maximum wins in that order; seg receives the label as a
   float. */
int glia_ms_segmentation(int64_t n, const float* bg, const float* gm, const float* wm, const float* vt,
                         const float* csf, const float* c, float* seg, void* stream);
/* Tumor::clipTumor / MatProp::clipHealthyTissues (clipVectorCuda): x = max(x, 0), in place */
int glia_ms_clip(int64_t n, float* x, void* stream);
/* the mask of Tumor::computeEdema (src/mat/Tumor.cpp:160-178): ed = 1 where i > threshold
   (invasive_threshold_, 0.001), else 0; the reference then smooths it with sigma 2 pi / n0 */
int glia_ms_edema_mask(int64_t n, const float* i, double threshold, float* ed, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GLIA_MS_H_ */
