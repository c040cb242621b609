/*
 * DerivativeOperators.h -- the reference's DerivativeOperators / DerivativeOperatorsRD
 * (include/grad/DerivativeOperators.h:9-81) over a glia_rd handle.
 *
 * x is the reference's parameter vector [p (np_) | kappa (nk_) | rho (nr_)] (nk_ = nr_ = 1 here).
 * As in updateReactionAndDiffusion (src/grad/DerivativeOperators.cpp:323-367), kappa is taken
 * from x when diffusivity_inversion_ or flag_reaction_inv_ is set and rho when flag_reaction_inv_
 * is set; with flag_reaction_inv_ the objective keeps c(0) as it is (no Phi p) and the gradient
 * is [0 | g_kappa | g_rho], otherwise c(0) = Phi p and the gradient is [g_p | g_kappa | 0].
 * Each call places the data (Data::dt1 with the Obs filter) and beta on the handle first.
 */
#ifndef GLIA_DERIVATIVEOPERATORS_H_
#define GLIA_DERIVATIVEOPERATORS_H_

#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <vector>

#include "PdeOperators.h"

namespace glia {

class DerivativeOperators {
 public:
  DerivativeOperators(std::shared_ptr<PdeOperators> pde_operators, std::shared_ptr<Parameters> params,
                      std::shared_ptr<Tumor> tumor)
      : pde_operators_(pde_operators), params_(params), tumor_(tumor) {}
  virtual ~DerivativeOperators() = default;
  virtual PetscErrorCode evaluateObjective(PetscReal* J, Vec x, std::shared_ptr<Data> data) = 0;
  virtual PetscErrorCode evaluateGradient(Vec dJ, Vec x, std::shared_ptr<Data> data) = 0;
  virtual PetscErrorCode evaluateObjectiveAndGradient(PetscReal* J, Vec dJ, Vec x, std::shared_ptr<Data> data) = 0;
  virtual PetscErrorCode evaluateHessian(Vec y, Vec x) = 0;
  /* evaluateConstantHessianApproximation (include/grad/DerivativeOperators.h:32): none by default */
  virtual PetscErrorCode evaluateConstantHessianApproximation(Vec y, Vec x) { return (y && x) ? 0 : GLIA_RD_ERR_ARG; }
  virtual PetscErrorCode updateReactionAndDiffusion(Vec x) = 0;
  virtual PetscErrorCode gradDiffusion(Vec dJ) = 0;
  virtual PetscErrorCode gradReaction(Vec dJ) = 0;
  virtual PetscErrorCode checkGradient(Vec p, std::shared_ptr<Data> data) = 0;
  /* reset (src/grad/DerivativeOperators.cpp:168-186): adopt the new operators, parameters, tumor */
  virtual PetscErrorCode reset(Vec p, std::shared_ptr<PdeOperators> pde_operators, std::shared_ptr<Parameters> params,
                               std::shared_ptr<Tumor> tumor) {
    if (!p) return GLIA_RD_ERR_ARG;
    pde_operators_ = pde_operators;
    params_ = params;
    tumor_ = tumor;
    return 0;
  }

  std::shared_ptr<PdeOperators> pde_operators_;
  std::shared_ptr<Parameters> params_;
  std::shared_ptr<Tumor> tumor_;
};

class DerivativeOperatorsRD : public DerivativeOperators {
 public:
  DerivativeOperatorsRD(std::shared_ptr<PdeOperators> pde_operators, std::shared_ptr<Parameters> params,
                        std::shared_ptr<Tumor> tumor)
      : DerivativeOperators(pde_operators, params, tumor) {
    rc_ = getHandle(params, &h_);
  }

  /* updateReactionAndDiffusion (src/grad/DerivativeOperators.cpp:323-367) */
  PetscErrorCode updateReactionAndDiffusion(Vec x) override {
    if (rc_) return rc_;
    const auto& tu = *params_->tu_;
    const auto& opt = *params_->opt_;
    PetscErrorCode rc = check_layout(x);
    if (rc || (rc = x->pull())) return rc;
    if (opt.diffusivity_inversion_ || opt.flag_reaction_inv_) {
      if ((rc = tumor_->k_->updateIsotropicCoefficients(x->v[tu.np_] * opt.k_scale_, 0, 0, tumor_->mat_prop_, params_)))
        return rc;
      if ((rc = pde_operators_->diff_solver_->precFactor())) return rc;
    }
    if (opt.flag_reaction_inv_)
      rc = tumor_->rho_->updateIsotropicCoefficients(x->v[tu.np_ + tu.nk_] * opt.rho_scale_, 0, 0, tumor_->mat_prop_,
                                                     params_);
    return rc;
  }
  /* evaluateObjective (src/grad/DerivativeOperatorsRD.cpp:7-61) */
  PetscErrorCode evaluateObjective(PetscReal* J, Vec x, std::shared_ptr<Data> data) override {
    return evaluate(J, nullptr, x, data);
  }
  /* evaluateGradient (:63-127) */
  PetscErrorCode evaluateGradient(Vec dJ, Vec x, std::shared_ptr<Data> data) override {
    PetscReal J = 0;
    return evaluate(&J, dJ, x, data);
  }
  /* evaluateObjectiveAndGradient (:130-226) */
  PetscErrorCode evaluateObjectiveAndGradient(PetscReal* J, Vec dJ, Vec x, std::shared_ptr<Data> data) override {
    return evaluate(J, dJ, x, data);
  }
  /* evaluateHessian (:229-438): Gauss-Newton matvec at the linearization point of the last
     evaluation; the kappa blocks when diffusivity_inversion_ is set */
  PetscErrorCode evaluateHessian(Vec y, Vec x) override {
    if (rc_) return rc_;
    PetscErrorCode rc = check_layout(x);
    if (rc || !y || y->n != x->n) return rc ? rc : GLIA_RD_ERR_ARG;
    std::vector<double> ax, ay;
    if ((rc = to_abi(x, ax))) return rc;
    ay.assign(ax.size(), 0.0);
    if ((rc = glia_rd_hessian(h_, params_->opt_->diffusivity_inversion_ ? 1 : 0, ax.data(), ay.data()))) return rc;
    return from_abi(ay, y);
  }

  /* evaluateConstantHessianApproximation (src/grad/DerivativeOperatorsRD.cpp:440-448): y = beta
     Phi^T Phi x over the p block (x and y hold np entries; Phi x passes through c_0_) */
  PetscErrorCode evaluateConstantHessianApproximation(Vec y, Vec x) override {
    if (rc_) return rc_;
    if (!y || !x) return GLIA_RD_ERR_ARG;
    PetscErrorCode rc = tumor_->phi_->apply(tumor_->c_0_, x);
    if (rc || (rc = tumor_->phi_->applyTranspose(y, tumor_->c_0_))) return rc;
    return VecScale(y, params_->opt_->beta_);
  }

  /* gradDiffusion (src/grad/DerivativeOperators.cpp:189-256): dJ[np] = g_kappa over the history of
     the last solveState / solveAdjoint, when kappa is inverted */
  PetscErrorCode gradDiffusion(Vec dJ) override {
    if (rc_) return rc_;
    const auto& opt = *params_->opt_;
    if (!dJ || dJ->n <= params_->tu_->np_) return GLIA_RD_ERR_ARG;
    if (!(opt.diffusivity_inversion_ || opt.flag_reaction_inv_)) return 0;
    double gk = 0;
    PetscErrorCode rc = glia_rd_grad_diffusion(h_, &gk);
    if (rc || (rc = dJ->pull())) return rc;
    dJ->v[params_->tu_->np_] = gk;
    return dJ->push();
  }
  /* gradReaction (:258-321): dJ[np + nk] = g_rho when rho is inverted */
  PetscErrorCode gradReaction(Vec dJ) override {
    if (rc_) return rc_;
    const auto& tu = *params_->tu_;
    if (!dJ || dJ->n <= tu.np_ + tu.nk_) return GLIA_RD_ERR_ARG;
    if (!params_->opt_->flag_reaction_inv_) return 0;
    double gr = 0;
    PetscErrorCode rc = glia_rd_grad_reaction(h_, &gr);
    if (rc || (rc = dJ->pull())) return rc;
    dJ->v[tu.np_ + tu.nk_] = gr;
    return dJ->push();
  }
  /* checkGradient (:369-450): the Taylor test at p along a random direction with entries uniform
     in [0, 1) (seeded), h = 10^0 .. 10^-5; prints the reference's lines and keeps
     |J - J_taylor| in taylor_err_ and (J - J_p) / h - <dJ, p~> in taylor_gdiff_ */
  PetscErrorCode checkGradient(Vec p, std::shared_ptr<Data> data) override {
    if (rc_) return rc_;
    PetscErrorCode rc = check_layout(p);
    if (rc) return rc;
    std::vector<double> a;
    if ((rc = to_abi(p, a)) || (rc = place_data(data))) return rc;
    // the direction over the inverted entries of the reference's x, mapped to the ABI's layout
    const auto& tu = *params_->tu_;
    const auto& opt = *params_->opt_;
    const int np = abi_np();
    std::mt19937_64 gen(20260101);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> ref(p->n), d(a.size(), 0.0);
    for (double& e : ref) e = u(gen);
    if (mode() == GLIA_RD_MODE_P_KAPPA)
      for (int i = 0; i < np && i < tu.np_; i++) d[i] = ref[i];
    if ((opt.diffusivity_inversion_ || opt.flag_reaction_inv_) && p->n > tu.np_) d[np] = ref[tu.np_];
    if (opt.flag_reaction_inv_ && p->n > tu.np_ + tu.nk_) d[np + 1] = ref[tu.np_ + tu.nk_];
    double norm = 0;
    for (double e : p->v) norm += e * e;
    std::printf(" ----- Gradient check with taylor expansion ----- \nGradient check performed at x with norm: %g\n",
                std::sqrt(norm));
    taylor_err_.assign(6, 0.0);
    taylor_gdiff_.assign(6, 0.0);
    if ((rc = glia_rd_check_gradient(h_, mode(), a.data(), d.data(), taylor_err_.data(), taylor_gdiff_.data())))
      return rc;
    for (int i = 0; i < 6; i++)
      std::printf("h[i]: %g |J - J_taylor|: %g  log10(diff) : %g g_fd - xg_dot: %g\n", std::pow(10.0, -i),
                  taylor_err_[i], std::log10(taylor_err_[i]), taylor_gdiff_[i]);
    return 0;
  }

  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
  std::vector<double> taylor_err_, taylor_gdiff_;

 private:
  int abi_np() const { return glia_rd_basis_size(h_, 0); }
  int mode() const { return params_->opt_->flag_reaction_inv_ ? GLIA_RD_MODE_KAPPA_RHO : GLIA_RD_MODE_P_KAPPA; }

  PetscErrorCode check_layout(Vec x) const {
    const auto& tu = *params_->tu_;
    if (!x) return GLIA_RD_ERR_ARG;
    if (tu.nk_ != 1 || tu.nr_ != 1) return GLIA_RD_ERR_UNSUPPORTED;  // one kappa, one rho
    if (x->n != tu.np_ + params_->get_nk() + params_->get_nr() && x->n != tu.np_ + tu.nk_ + tu.nr_)
      return GLIA_RD_ERR_ARG;
    if (mode() == GLIA_RD_MODE_P_KAPPA && abi_np() != tu.np_) return GLIA_RD_ERR_STATE;  // Phi p needs the basis
    return 0;
  }
  /* the reference's x -> the ABI's [p (basis size) | kappa | rho] */
  PetscErrorCode to_abi(Vec x, std::vector<double>& a) const {
    const auto& tu = *params_->tu_;
    const auto& opt = *params_->opt_;
    PetscErrorCode rc = x->pull();
    if (rc) return rc;
    const int np = abi_np();
    a.assign(np + 2, 0.0);
    if (mode() == GLIA_RD_MODE_P_KAPPA)
      for (int i = 0; i < np; i++) a[i] = x->v[i];
    const bool upd_k = opt.diffusivity_inversion_ || opt.flag_reaction_inv_;
    a[np] = upd_k ? x->v[tu.np_] : params_->ctx_->kappa / opt.k_scale_;
    a[np + 1] = (opt.flag_reaction_inv_ && x->n > tu.np_ + tu.nk_) ? x->v[tu.np_ + tu.nk_]
                                                                     : params_->ctx_->rho / opt.rho_scale_;
    return 0;
  }
  /* the ABI's gradient -> the reference's layout */
  PetscErrorCode from_abi(const std::vector<double>& a, Vec dJ) const {
    const auto& tu = *params_->tu_;
    const int np = abi_np();
    dJ->ensure();
    std::fill(dJ->v.begin(), dJ->v.end(), 0.0);
    if (mode() == GLIA_RD_MODE_P_KAPPA)
      for (int i = 0; i < np && i < tu.np_; i++) dJ->v[i] = a[i];
    if (dJ->n > tu.np_) dJ->v[tu.np_] = a[np];
    if (dJ->n > tu.np_ + tu.nk_) dJ->v[tu.np_ + tu.nk_] = a[np + 1];
    return dJ->push();
  }
  PetscErrorCode place_data(std::shared_ptr<Data> data) {
    if (!data || !data->dt1()) return GLIA_RD_ERR_ARG;
    std::vector<float> d1, mask;
    PetscErrorCode rc = detail::to_f32(data->dt1(), d1);
    if (rc || (rc = detail::to_f32(tumor_->obs_->custom_, mask))) return rc;
    if ((rc = glia_rd_set_data(h_, 0, d1.data(), mask.empty() ? nullptr : mask.data(), tumor_->obs_->thr_, GLIA_RD_HOST)))
      return rc;
    return glia_rd_set_beta(h_, params_->opt_->beta_);
  }
  PetscErrorCode evaluate(PetscReal* J, Vec dJ, Vec x, std::shared_ptr<Data> data) {
    if (rc_) return rc_;
    PetscErrorCode rc = check_layout(x);
    if (rc || !J || (dJ && dJ->n != x->n)) return rc ? rc : GLIA_RD_ERR_ARG;
    std::vector<double> a, g;
    if ((rc = to_abi(x, a)) || (rc = place_data(data))) return rc;
    const int np = abi_np();
    if (dJ) {
      g.assign(a.size(), 0.0);
      rc = glia_rd_objective_and_gradient(h_, mode(), a.data(), J, g.data());
    } else {
      rc = glia_rd_objective(h_, mode(), a.data(), J);
    }
    if (rc) return rc;
    params_->ctx_->kappa = a[np] * params_->opt_->k_scale_;
    if (mode() == GLIA_RD_MODE_KAPPA_RHO) params_->ctx_->rho = a[np + 1] * params_->opt_->rho_scale_;
    return dJ ? from_abi(g, dJ) : 0;
  }
};

}  // namespace glia

#endif
