/*
 * Parameters.h -- the parts of the reference's Parameters (include/Parameters.h:155-451) that the
 * RD path reads, and the device handle shared by the classes built from one Parameters object.
 *
 * The handle (one glia_rd_t, one member) is created by the first object that needs it, from the
 * values the Parameters hold at that time (grid, dt, nt, ratios, beta, scales, storage flags), as
 * the reference's objects read them in their constructors.
 */
#ifndef GLIA_PARAMETERS_H_
#define GLIA_PARAMETERS_H_

#include <cmath>
#include <array>
#include <memory>

#include "../glia_opt.h"
#include "Vec.h"

namespace glia {

struct Grid {
  int n_[3] = {64, 64, 64};
  ScalarType h_[3] = {2 * M_PI / 64, 2 * M_PI / 64, 2 * M_PI / 64};
  int isize_[3] = {64, 64, 64};
  int istart_[3] = {0, 0, 0};
  int64_t ng_ = 64 * 64 * 64, nl_ = 64 * 64 * 64;
};

/* TumorParameters defaults (include/Parameters.h:157-258) */
struct TumorParameters {
  int model_ = 1;
  bool adv_velocity_set_ = false;
  // elasticity of the mass-effect models (include/Parameters.h:175-184)
  ScalarType E_csf_ = 100, E_healthy_ = 2100, E_tumor_ = 8000, E_bg_ = 15000;
  ScalarType nu_csf_ = 0.1, nu_healthy_ = 0.4, nu_tumor_ = 0.45, nu_bg_ = 0.45;
  ScalarType screen_low_ = 1E-2, screen_high_ = 1E2;
  int np_ = 1, nk_ = 1, nr_ = 1;
  int nt_ = 1;
  ScalarType dt_ = 0.5;
  ScalarType k_ = 0.0, rho_ = 10.0;
  ScalarType k_gm_wm_ratio_ = 0.0, k_glm_wm_ratio_ = 0.0;
  ScalarType r_gm_wm_ratio_ = 0.0, r_glm_wm_ratio_ = 0.0;
  ScalarType phi_sigma_ = 2 * M_PI / 64;
  // basis placement (include/Parameters.h:203-207, 244-246): data-driven sigma, bounding-box
  // spacing, data threshold, Gaussian volume fraction and the bounding box's center
  ScalarType phi_sigma_data_driven_ = 2 * M_PI / 256;
  ScalarType phi_spacing_factor_ = 1.5;
  ScalarType data_threshold_ = 0.05;
  ScalarType gaussian_vol_frac_ = 0.0;
  std::array<ScalarType, 3> user_cm_ = {2 * M_PI / 128 * 64, 2 * M_PI / 128 * 38, 2 * M_PI / 128 * 56};
  ScalarType obs_threshold_1_ = 0.0;
  ScalarType smoothing_factor_ = 1.0;
  int order_ = 2;
  bool phi_store_ = false, adjoint_store_ = false;
  int sparsity_level_ = 5;
  ScalarType thresh_component_weight_ = 1e-3;
};

/* the reference's forcing sequences and Newton solvers (include/utils/TypeDefs.h:10-12) */
enum { QDFS = GLIA_OPT_FORCING_QUADRATIC, SLFS = GLIA_OPT_FORCING_SUPERLINEAR };
enum { GAUSSNEWTON = GLIA_OPT_GAUSS_NEWTON, QUASINEWTON = GLIA_OPT_QUASI_NEWTON };

/* OptimizerSettings defaults (include/Parameters.h:18-111) */
struct OptimizerSettings {
  ScalarType beta_ = 1e-4;
  ScalarType k_scale_ = 1.0, rho_scale_ = 1.0;
  bool flag_reaction_inv_ = false;
  bool diffusivity_inversion_ = true;
  bool reaction_inversion_ = false;
  ScalarType opttolgrad_ = 1e-3, gatol_ = 1e-4;
  int newton_maxit_ = 30, newton_minit_ = 1;
  int linesearch_ = 1;  // 0 Armijo, 1 More-Thuente
  int newton_solver_ = QUASINEWTON, krylov_maxit_ = 30, fseqtype_ = SLFS;
  int ls_max_func_evals = 10, lbfgs_vectors_ = 10;
  ScalarType k_lb_ = 1e-3, k_ub_ = 1.0, rho_lb_ = 1.0, rho_ub_ = 15.0;
  bool rescale_init_cond_ = false;
  int gist_maxit_ = 5;
  bool prune_components_ = true;
};

/* the device handle of one Parameters object, and the coefficients last pushed to it */
struct Context {
  glia_rd_t h = nullptr;
  ScalarType kappa = 0.0, rho = 0.0;
  Context() = default;
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ~Context() {
    if (h) glia_rd_destroy(h);
  }
};

struct Parameters {
  std::shared_ptr<Grid> grid_ = std::make_shared<Grid>();
  std::shared_ptr<TumorParameters> tu_ = std::make_shared<TumorParameters>();
  std::shared_ptr<OptimizerSettings> opt_ = std::make_shared<OptimizerSettings>();
  /* not in the reference: binomial checkpointing with this many stored states (0 = store all),
     PCG tolerances, and the handle */
  int checkpoint_slots_ = 0;
  double pcg_rtol_ = 1e-6;
  int pcg_maxit_ = 5000;
  std::shared_ptr<Context> ctx_;

  int get_nk() const { return (opt_->diffusivity_inversion_ || opt_->reaction_inversion_) ? tu_->nk_ : 0; }
  int get_nr() const { return opt_->reaction_inversion_ ? tu_->nr_ : 0; }
};

/* initializeGrid (src/utils/Utils.cpp): an n0 x n1 x n2 periodic grid on [0, 2 pi)^3 */
inline PetscErrorCode initializeGrid(int n0, int n1, int n2, std::shared_ptr<Parameters> params) {
  if (!params || n0 <= 0 || n1 <= 0 || n2 <= 0) return GLIA_RD_ERR_ARG;
  if (params->ctx_) return GLIA_RD_ERR_STATE;  // the grid of a live handle cannot change
  Grid& g = *params->grid_;
  const int n[3] = {n0, n1, n2};
  for (int d = 0; d < 3; d++) {
    g.n_[d] = g.isize_[d] = n[d];
    g.istart_[d] = 0;
    g.h_[d] = 2 * M_PI / n[d];
  }
  g.ng_ = g.nl_ = (int64_t)n0 * n1 * n2;
  return 0;
}
inline PetscErrorCode initializeGrid(int n, std::shared_ptr<Parameters> params) {
  return initializeGrid(n, n, n, params);
}

/* the handle of params, created on first use */
inline PetscErrorCode getHandle(const std::shared_ptr<Parameters>& params, glia_rd_t* h) {
  if (!params || !h) return GLIA_RD_ERR_ARG;
  if (!params->ctx_) {
    glia_rd_config c;
    glia_rd_config_init(&c);
    for (int d = 0; d < 3; d++) c.n[d] = params->grid_->n_[d];
    c.nmembers = 1;
    c.dt = params->tu_->dt_;
    c.nt = params->tu_->nt_;
    c.order = params->tu_->order_;
    c.pcg_rtol = params->pcg_rtol_;
    c.pcg_maxit = params->pcg_maxit_;
    c.checkpoint_slots = params->checkpoint_slots_;
    c.adjoint_store = params->tu_->adjoint_store_;
    c.phi_store = params->tu_->phi_store_;
    c.k_gm_wm = params->tu_->k_gm_wm_ratio_;
    c.k_glm_wm = params->tu_->k_glm_wm_ratio_;
    c.r_gm_wm = params->tu_->r_gm_wm_ratio_;
    c.r_glm_wm = params->tu_->r_glm_wm_ratio_;
    c.beta = params->opt_->beta_;
    c.k_scale = params->opt_->k_scale_;
    c.rho_scale = params->opt_->rho_scale_;
    c.smoothing_factor = params->tu_->smoothing_factor_;
    auto ctx = std::make_shared<Context>();
    const int rc = glia_rd_create(&c, &ctx->h);
    if (rc) return rc;
    params->ctx_ = ctx;
  }
  *h = params->ctx_->h;
  return 0;
}

}  // namespace glia

#endif
