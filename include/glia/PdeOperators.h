/*
 * PdeOperators.h -- the reference's DiffusionSolver (include/pde/DiffusionSolver.h) and
 * PdeOperators / PdeOperatorsRD (include/pde/PdeOperators.h) over a glia_rd handle.
 *
 * The state and adjoint histories stay on the device in the engine's spectral layout (or its
 * checkpoints).  The reference's public history members c_, p_, c_half_ are read-only Vec views:
 * reading c_[i] transforms the stored state back to a real field (under checkpointing only the
 * stored indices exist), p_ is kept in store-all mode as the reference keeps it, c_half_ with
 * adjoint_store_.  c(0), c(T) and alpha(0) are Tumor::c_0_, c_t_, p_0_.
 */
#ifndef GLIA_PDEOPERATORS_H_
#define GLIA_PDEOPERATORS_H_

#include <memory>
#include <vector>

#include "Tumor.h"

namespace glia {

/* DiffusionSolver (src/pde/DiffusionSolver.cpp): Crank-Nicolson step by preconditioned CG */
class DiffusionSolver {
 public:
  DiffusionSolver(std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops,
                  std::shared_ptr<DiffCoef> k)
      : params_(params), spec_ops_(spec_ops), k_(k) {
    rc_ = getHandle(params, &h_);
  }
  /* precFactor (:119-180): the preconditioner from the current kxx_avg_ and solver-context dt */
  PetscErrorCode precFactor() { return rc_ ? rc_ : glia_rd_prec_factor(h_); }
  /* solve (:217-250): c <- (I - dt/2 D)^-1 (I + dt/2 D) c, in place; ksp_itr_ = CG iterations */
  PetscErrorCode solve(Vec c, ScalarType dt) {
    if (rc_) return rc_;
    if (!c || c->n != params_->grid_->ng_) return GLIA_RD_ERR_ARG;
    Vec w = nullptr;
    VecCreateField(h_, 0, GLIA_RD_FIELD_WORK, c->n, &w);
    PetscErrorCode rc = VecCopy(c, w);
    if (!rc) rc = glia_rd_diffusion_solve(h_, GLIA_RD_FIELD_WORK, dt);
    glia_rd_stats_t st;
    if (!rc && !(rc = glia_rd_stats(h_, 0, &st))) ksp_itr_ = st.last_iterations;
    if (!rc) rc = VecCopy(w, c);
    VecDestroy(&w);
    return rc;
  }

  int ksp_itr_ = 0;
  std::shared_ptr<Parameters> params_;
  std::shared_ptr<SpectralOperators> spec_ops_;
  std::shared_ptr<DiffCoef> k_;
  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

/* PdeOperators (include/pde/PdeOperators.h:10-52) */
class PdeOperators {
 public:
  PdeOperators(std::shared_ptr<Tumor> tumor, std::shared_ptr<Parameters> params,
               std::shared_ptr<SpectralOperators> spec_ops)
      : tumor_(tumor), params_(params), spec_ops_(spec_ops) {
    diff_solver_ = std::make_shared<DiffusionSolver>(params, spec_ops, tumor->k_);
  }
  virtual ~PdeOperators() = default;
  virtual PetscErrorCode solveState(int linearized) = 0;
  virtual PetscErrorCode solveAdjoint(int linearized) = 0;
  virtual PetscErrorCode resizeTimeHistory(std::shared_ptr<Parameters> params) = 0;
  virtual PetscErrorCode reset(std::shared_ptr<Parameters> params, std::shared_ptr<Tumor> tumor = {}) = 0;

  /* time history of the state, the adjoint and the half-step states (read-only views) */
  std::vector<Vec> c_, p_, c_half_;

  std::shared_ptr<Tumor> tumor_;
  std::shared_ptr<Parameters> params_;
  std::shared_ptr<SpectralOperators> spec_ops_;
  std::shared_ptr<DiffusionSolver> diff_solver_;
  int diff_ksp_itr_state_ = 0, diff_ksp_itr_adj_ = 0;
};

/* PdeOperatorsRD (src/pde/PdeOperators.cpp) */
class PdeOperatorsRD : public PdeOperators {
 public:
  PdeOperatorsRD(std::shared_ptr<Tumor> tumor, std::shared_ptr<Parameters> params,
                 std::shared_ptr<SpectralOperators> spec_ops)
      : PdeOperators(tumor, params, spec_ops) {
    rc_ = getHandle(params, &h_);
    // the reference keeps p_ (:26-44); here it needs store-all mode
    if (!rc_ && params->checkpoint_slots_ == 0) rc_ = glia_rd_keep_adjoint_history(h_, 1);
    if (!rc_) make_views();
  }
  ~PdeOperatorsRD() override { drop_views(); }
  PdeOperatorsRD(const PdeOperatorsRD&) = delete;
  PdeOperatorsRD& operator=(const PdeOperatorsRD&) = delete;

  /* resizeTimeHistory (:56-103): the history for params->tu_->nt_ (dt_ taken too) */
  PetscErrorCode resizeTimeHistory(std::shared_ptr<Parameters> params) override {
    if (rc_) return rc_;
    PetscErrorCode rc = glia_rd_set_time(h_, params->tu_->dt_, params->tu_->nt_);
    if (rc) return rc;
    drop_views();
    make_views();
    return 0;
  }
  /* reset (:6-15): resizeTimeHistory, then the new parameters (and tumor) */
  PetscErrorCode reset(std::shared_ptr<Parameters> params, std::shared_ptr<Tumor> tumor = {}) override {
    params_ = params;
    PetscErrorCode rc = resizeTimeHistory(params);
    if (tumor) tumor_ = tumor;
    return rc;
  }
  /* solveState (:235-316): linearized 0 from Tumor::c_0_ with the history kept; 1 and 2 the
     tangent-linear equations from c~0 = Tumor::c_0_ around that history (2 with the secondary
     sources); c(T) or c~(T) in Tumor::c_t_ */
  PetscErrorCode solveState(int linearized) override {
    if (rc_) return rc_;
    PetscErrorCode rc = glia_rd_reset_stats(h_);
    if (rc || (rc = glia_rd_solve_state(h_, linearized))) return rc;
    glia_rd_stats_t st;
    if ((rc = glia_rd_stats(h_, 0, &st))) return rc;
    diff_ksp_itr_state_ += st.pcg_iterations;
    return 0;
  }
  /* solveAdjoint (:372-420); the final condition -O^T (O c(T) - d1) (linearized 0, 1) or
     -O^T O c~(T) (2) comes from the data the last derivative-operator call placed on the handle;
     alpha(0) in Tumor::p_0_ */
  PetscErrorCode solveAdjoint(int linearized) override {
    if (rc_) return rc_;
    PetscErrorCode rc = glia_rd_reset_stats(h_);
    if (rc || (rc = glia_rd_solve_adjoint(h_, linearized))) return rc;
    glia_rd_stats_t st;
    if ((rc = glia_rd_stats(h_, 0, &st))) return rc;
    diff_ksp_itr_adj_ += st.pcg_iterations;
    return 0;
  }

  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;

 private:
  void make_views() {
    const int nt = params_->tu_->nt_;
    const int64_t ng = params_->grid_->ng_;
    c_.assign(nt + 1, nullptr);
    p_.assign(nt + 1, nullptr);
    for (int i = 0; i <= nt; i++) {
      VecCreateHistory(h_, 0, GLIA_RD_HIST_STATE, i, ng, &c_[i]);
      VecCreateHistory(h_, 0, GLIA_RD_HIST_ADJOINT, i, ng, &p_[i]);
    }
    c_half_.assign(params_->tu_->adjoint_store_ ? nt : 0, nullptr);
    for (int i = 0; i < (int)c_half_.size(); i++) VecCreateHistory(h_, 0, GLIA_RD_HIST_HALF, i, ng, &c_half_[i]);
  }
  void drop_views() {
    for (auto* v : {&c_, &p_, &c_half_})
      for (auto& x : *v) VecDestroy(&x);
    c_.clear();
    p_.clear();
    c_half_.clear();
  }
};

}  // namespace glia

#endif
