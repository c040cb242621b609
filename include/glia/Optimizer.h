/*
 * Optimizer.h -- the reference's inversion classes (include/opt/Optimizer.h, RDOptimizer.h,
 * TILOptimizer.h, SparseTILOptimizer.h) over the inversion loop of include/glia_opt.h.
 *
 * The solvers run the loop on the reference-layout parameter vector through this directory's
 * DerivativeOperatorsRD, as TAO calls the reference's evaluation callbacks
 * (src/opt/TaoInterface.cpp:25-93):
 *   TILOptimizer        x = [p (np_) | kappa]  p >= 0, kappa in [k_lb_, k_ub_]; relative test
 *                       against the gradient at x = 0
 *   RDOptimizer         x = [kappa | rho]      bounds [k_lb_, k_ub_] x [rho_lb_, rho_ub_]; c0 = Phi p
 *                       fixed (rescaled to max 1 with rescale_init_cond_); beta = 0 while it runs
 *   SparseTILOptimizer  CoSaMP over the TIL and RD solves (glia_rd_sparse_til)
 * Results: getSolution() (the reference's xout_ layout [p | kappa | rho]) and result_.
 */
#ifndef GLIA_OPTIMIZER_H_
#define GLIA_OPTIMIZER_H_

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "../glia_opt.h"
#include "DerivativeOperators.h"

namespace glia {

class Optimizer {
 public:
  virtual ~Optimizer() {
    VecDestroy(&xin_);
    VecDestroy(&xout_);
  }
  virtual PetscErrorCode initialize(std::shared_ptr<DerivativeOperators> derivative_operators,
                                    std::shared_ptr<PdeOperators> pde_operators, std::shared_ptr<Parameters> params,
                                    std::shared_ptr<Tumor> tumor) {
    derivative_operators_ = derivative_operators;
    pde_operators_ = pde_operators;
    params_ = params;
    tumor_ = tumor;
    data_ = std::make_shared<Data>();
    initialized_ = true;
    return 0;
  }
  virtual PetscErrorCode solve() = 0;
  virtual PetscErrorCode setInitialGuess(Vec x_init) = 0;
  virtual PetscErrorCode setVariableBounds() = 0;

  void setData(std::shared_ptr<Data> d) { data_ = d; }
  void setData(Vec d1, Vec d0 = nullptr) { data_->set(d1, d0); }
  void updateReferenceGradient(bool b) { update_reference_gradient_ = b; }
  void updateReferenceObjective(bool) {}
  bool initialized() { return initialized_; }
  Vec getSolution() { return xout_; }
  PetscErrorCode evalGradient(Vec x, Vec dJ) { return derivative_operators_->evaluateGradient(dJ, x, data_); }
  PetscErrorCode evalObjective(Vec x, PetscReal* J) { return derivative_operators_->evaluateObjective(J, x, data_); }
  PetscErrorCode evalHessian(Vec y, Vec x) { return derivative_operators_->evaluateHessian(y, x); }
  PetscErrorCode evalObjectiveAndGradient(Vec x, PetscReal* J, Vec dJ) {
    return derivative_operators_->evaluateObjectiveAndGradient(J, dJ, x, data_);
  }

  glia_opt_result result_ = {};

 protected:
  /* the loop on x (n_inv_ entries, reference layout) within lb_/ub_ */
  PetscErrorCode run(std::vector<double>& x, double gnorm_ref) {
    glia_opt_settings s;
    glia_opt_settings_init(&s);
    const auto& o = *params_->opt_;
    s.max_it = o.newton_maxit_;
    s.min_it = o.newton_minit_;
    s.gatol = o.gatol_;
    s.gttol = o.opttolgrad_;
    s.lbfgs_vectors = o.lbfgs_vectors_;
    s.ls_max_funcs = o.ls_max_func_evals;
    s.line_search = o.linesearch_;
    s.newton_solver = o.newton_solver_;
    s.krylov_maxit = o.krylov_maxit_;
    s.forcing = o.fseqtype_;
    s.gnorm_ref = gnorm_ref;
    Vec xv = nullptr, gv = nullptr;
    VecCreateSeq(PETSC_COMM_SELF, (PetscInt)x.size(), &xv);
    VecCreateSeq(PETSC_COMM_SELF, (PetscInt)x.size(), &gv);
    struct Ctx {
      Optimizer* self;
      Vec x, g;
    } ctx{this, xv, gv};
    // evaluateObjectiveFunctionAndGradient (src/opt/TaoInterface.cpp:71-93)
    auto fg = [](void* c, int, const double* xa, double* f, double* g) -> int {
      Ctx* k = static_cast<Ctx*>(c);
      k->x->ensure();
      std::copy(xa, xa + k->x->n, k->x->v.begin());
      PetscErrorCode rc = k->self->evalObjectiveAndGradient(k->x, f, k->g);
      if (rc) return rc;
      std::copy(k->g->v.begin(), k->g->v.end(), g);
      return 0;
    };
    PetscErrorCode rc;
    if (o.newton_solver_ == GAUSSNEWTON) {
      // hessianMatVec of the TAO BNLS solve (src/opt/TaoInterface.cpp:101-121), no preconditioner
      auto hv = [](void* c, int, const double* va, double* ya) -> int {
        Ctx* k = static_cast<Ctx*>(c);
        k->x->ensure();
        std::copy(va, va + k->x->n, k->x->v.begin());
        PetscErrorCode e = k->self->evalHessian(k->g, k->x);
        if (e) return e;
        std::copy(k->g->v.begin(), k->g->v.end(), ya);
        return 0;
      };
      rc = glia_gauss_newton(1, (int)x.size(), x.data(), lb_.data(), ub_.data(), fg, hv, &ctx, &s, &result_);
    } else {
      rc = glia_blmvm(1, (int)x.size(), x.data(), lb_.data(), ub_.data(), fg, &ctx, &s, &result_);
    }
    VecDestroy(&xv);
    VecDestroy(&gv);
    return rc;
  }
  /* ||dJ(0)|| over the unknowns (src/opt/TaoInterface.cpp:285-300) */
  PetscErrorCode gradient_norm_at_zero(int n, double* out) {
    Vec z = nullptr, g = nullptr;
    VecCreateSeq(PETSC_COMM_SELF, n, &z);
    VecCreateSeq(PETSC_COMM_SELF, n, &g);
    VecSet(z, 0.0);
    PetscErrorCode rc = evalGradient(z, g);
    if (!rc) rc = VecNorm(g, NORM_2, out);
    VecDestroy(&z);
    VecDestroy(&g);
    return rc;
  }

  std::shared_ptr<DerivativeOperators> derivative_operators_;
  std::shared_ptr<PdeOperators> pde_operators_;
  std::shared_ptr<Parameters> params_;
  std::shared_ptr<Tumor> tumor_;
  std::shared_ptr<Data> data_;
  bool initialized_ = false, update_reference_gradient_ = true;
  int n_inv_ = 0;
  Vec xin_ = nullptr, xout_ = nullptr;
  std::vector<double> lb_, ub_;
};

/* TILOptimizer (src/opt/TILOptimizer.cpp): p and kappa */
class TILOptimizer : public Optimizer {
 public:
  /* setInitialGuess (:50-110): [p | kappa (| rho)] -> xin_ = [p | kappa] */
  PetscErrorCode setInitialGuess(Vec x_init) override {
    if (!x_init || !initialized_) return GLIA_RD_ERR_STATE;
    const int np = params_->tu_->np_, nk = params_->get_nk();
    n_inv_ = np + nk;
    PetscErrorCode rc = x_init->pull();
    if (rc || x_init->n < n_inv_) return rc ? rc : GLIA_RD_ERR_ARG;
    VecDestroy(&xin_);
    VecCreateSeq(PETSC_COMM_SELF, n_inv_, &xin_);
    xin_->ensure();
    std::copy(x_init->v.begin(), x_init->v.begin() + n_inv_, xin_->v.begin());
    return 0;
  }
  /* setVariableBounds (:245-270): p >= 0, kappa in [k_lb_, k_ub_] */
  PetscErrorCode setVariableBounds() override {
    const int np = params_->tu_->np_;
    lb_.assign(n_inv_, 0.0);
    ub_.assign(n_inv_, std::numeric_limits<double>::infinity());
    for (int i = np; i < n_inv_; i++) {
      lb_[i] = params_->opt_->k_lb_;
      ub_[i] = params_->opt_->k_ub_;
    }
    return 0;
  }
  /* solve (:114-230) */
  PetscErrorCode solve() override {
    if (!xin_) return GLIA_RD_ERR_STATE;
    const bool flag = params_->opt_->flag_reaction_inv_;
    params_->opt_->flag_reaction_inv_ = false;
    PetscErrorCode rc = setVariableBounds();
    double gref = 0;
    if (!rc && update_reference_gradient_) rc = gradient_norm_at_zero(n_inv_, &gref);
    std::vector<double> x(xin_->v.begin(), xin_->v.end());
    if (!rc) rc = run(x, gref);
    params_->opt_->flag_reaction_inv_ = flag;
    if (rc) return rc;
    VecDestroy(&xout_);
    VecCreateSeq(PETSC_COMM_SELF, n_inv_, &xout_);
    xout_->ensure();
    std::copy(x.begin(), x.end(), xout_->v.begin());
    if (params_->get_nk() > 0) params_->tu_->k_ = x[params_->tu_->np_];
    return 0;
  }
};

/* RDOptimizer (src/opt/RDOptimizer.cpp): kappa and rho with c0 fixed */
class RDOptimizer : public Optimizer {
 public:
  /* setInitialGuess (:36-135) with c0 given as the parametrization [p | kappa | rho]: c0 = Phi p
     (p rescaled so that max c0 = 1 with rescale_init_cond_) */
  PetscErrorCode setInitialGuess(Vec x_init) override {
    if (!x_init || !initialized_) return GLIA_RD_ERR_STATE;
    const auto& tu = *params_->tu_;
    const int np = tu.np_, nk = tu.nk_, nr = tu.nr_;
    PetscErrorCode rc = x_init->pull();
    if (rc || x_init->n != np + nk + nr) return rc ? rc : GLIA_RD_ERR_ARG;
    VecDestroy(&xin_);
    VecDuplicate(x_init, &xin_);
    xin_->v = x_init->v;
    Vec p = nullptr;
    VecCreateSeq(PETSC_COMM_SELF, np, &p);
    p->ensure();
    std::copy(xin_->v.begin(), xin_->v.begin() + np, p->v.begin());
    if ((rc = tumor_->phi_->apply(tumor_->c_0_, p))) {
      VecDestroy(&p);
      return rc;
    }
    if (params_->opt_->rescale_init_cond_) {
      double mx = 0;
      if (!(rc = VecMax(tumor_->c_0_, nullptr, &mx)) && mx > 0) {
        for (int i = 0; i < np; i++) p->v[i] = xin_->v[i] = xin_->v[i] / mx;
        rc = tumor_->phi_->apply(tumor_->c_0_, p);
      }
    }
    VecDestroy(&p);
    n_inv_ = nk + nr;
    return rc;
  }
  /* setVariableBounds (:306-343) */
  PetscErrorCode setVariableBounds() override {
    const auto& o = *params_->opt_;
    lb_ = {o.k_lb_, o.rho_lb_};
    ub_ = {o.k_ub_, o.rho_ub_};
    return 0;
  }
  /* solve (:145-300): beta = 0, flag_reaction_inv_, np_ = 0 while the loop runs */
  PetscErrorCode solve() override {
    if (!xin_ || n_inv_ != 2) return GLIA_RD_ERR_STATE;
    auto& tu = *params_->tu_;
    auto& o = *params_->opt_;
    const int np = tu.np_;
    const double beta = o.beta_;
    const bool flag = o.flag_reaction_inv_, rinv = o.reaction_inversion_;
    o.beta_ = 0;
    o.flag_reaction_inv_ = true;
    o.reaction_inversion_ = true;
    tu.np_ = 0;
    PetscErrorCode rc = setVariableBounds();
    std::vector<double> x = {xin_->v[np], xin_->v[np + tu.nk_]};
    if (!rc) rc = run(x, 0.0);
    o.beta_ = beta;
    o.flag_reaction_inv_ = flag;
    o.reaction_inversion_ = rinv;
    tu.np_ = np;
    if (rc) return rc;
    VecDestroy(&xout_);
    VecDuplicate(xin_, &xout_);
    xout_->v = xin_->v;
    xout_->v[np] = x[0];
    xout_->v[np + tu.nk_] = x[1];
    tu.k_ = x[0];
    tu.rho_ = x[1];
    return 0;
  }
};

/* SparseTILOptimizer (src/opt/SparseTILOptimizer.cpp): CoSaMP over TIL and RD solves on the
   handle (glia_rd_sparse_til); x = [p | kappa | rho] with np_ = the basis size */
class SparseTILOptimizer : public Optimizer {
 public:
  PetscErrorCode setInitialGuess(Vec x_init) override {
    if (!x_init || !initialized_) return GLIA_RD_ERR_STATE;
    const auto& tu = *params_->tu_;
    PetscErrorCode rc = x_init->pull();
    if (rc || x_init->n != tu.np_ + tu.nk_ + tu.nr_) return rc ? rc : GLIA_RD_ERR_ARG;
    VecDestroy(&xin_);
    VecDuplicate(x_init, &xin_);
    xin_->v = x_init->v;
    n_inv_ = (int)x_init->n;
    return 0;
  }
  PetscErrorCode setVariableBounds() override { return 0; }
  /* component labels of the bases and component weights (Phi::gaussian_labels_, component_weights_) */
  void setComponents(const std::vector<int>& labels, const std::vector<double>& weights) {
    labels_ = labels;
    weights_ = weights;
  }
  PetscErrorCode solve() override {
    if (!xin_) return GLIA_RD_ERR_STATE;
    auto& tu = *params_->tu_;
    const auto& o = *params_->opt_;
    glia_rd_t h = nullptr;
    PetscErrorCode rc = getHandle(params_, &h);
    if (rc) return rc;
    if (tu.nk_ != 1 || tu.nr_ != 1 || glia_rd_basis_size(h, 0) != tu.np_) return GLIA_RD_ERR_UNSUPPORTED;
    // the data reach the handle the way the derivative operators place them
    Vec d1 = data_ ? data_->dt1() : nullptr;
    if (!d1) return GLIA_RD_ERR_ARG;
    std::vector<float> d, mask;
    if ((rc = detail::to_f32(d1, d)) || (rc = detail::to_f32(tumor_->obs_->custom_, mask))) return rc;
    if ((rc = glia_rd_set_data(h, 0, d.data(), mask.empty() ? nullptr : mask.data(), tumor_->obs_->thr_,
                               GLIA_RD_HOST)))
      return rc;
    glia_sparse_til_settings s;
    glia_sparse_til_settings_init(&s);
    s.sparsity = tu.sparsity_level_;
    s.max_it = o.gist_maxit_;
    s.prune_components = o.prune_components_ ? 1 : 0;
    s.thresh_component_weight = tu.thresh_component_weight_;
    s.diffusivity_inversion = o.diffusivity_inversion_ ? 1 : 0;
    s.reaction_inversion = o.reaction_inversion_ ? 1 : 0;
    s.beta = o.beta_;
    s.k_lb = o.k_lb_;
    s.k_ub = o.k_ub_;
    s.rho_lb = o.rho_lb_;
    s.rho_ub = o.rho_ub_;
    for (glia_opt_settings* t : {&s.til, &s.rd}) {
      t->max_it = o.newton_maxit_;
      t->min_it = o.newton_minit_;
      t->gatol = o.gatol_;
      t->gttol = o.opttolgrad_;
      t->lbfgs_vectors = o.lbfgs_vectors_;
      t->ls_max_funcs = o.ls_max_func_evals;
      t->line_search = o.linesearch_;
    }
    std::vector<double> x(xin_->v.begin(), xin_->v.end());
    support_.assign(tu.np_, -1);
    const bool comps = !labels_.empty() && !weights_.empty();
    rc = glia_rd_sparse_til(h, x.data(), support_.data(), comps ? labels_.data() : nullptr,
                            comps ? weights_.data() : nullptr, comps ? (int)weights_.size() : 0, &s, &sparse_result_);
    if (rc) return rc;
    support_.erase(std::remove(support_.begin(), support_.end(), -1), support_.end());
    VecDestroy(&xout_);
    VecDuplicate(xin_, &xout_);
    xout_->v = x;
    tu.k_ = x[tu.np_];
    tu.rho_ = x[tu.np_ + 1];
    params_->ctx_->kappa = x[tu.np_];
    params_->ctx_->rho = x[tu.np_ + 1];
    return 0;
  }

  std::vector<int> support_;
  glia_sparse_til_result sparse_result_ = {};

 private:
  std::vector<int> labels_;
  std::vector<double> weights_;
};

}  // namespace glia

#endif
