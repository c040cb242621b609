/*
 * ElasticitySolver.h -- the reference's VariableLinearElasticitySolver (include/pde/ElasticitySolver.h,
 * src/pde/ElasticitySolver.cpp:6-533) over the C ABI of include/glia_el.h.
 *
 * solve(displacement, rhs) computes the material properties from tumor->c_t_ and the tissue maps of
 * tumor->mat_prop_ (bg = 1 - (gm + wm + vt + csf), as MatProp::setValuesCustom forms it) and runs
 * the CG solve from the given displacement.  The MatShell and PCShell operations are
 * applyOperator and applyPreconditioner.
 */
#ifndef GLIA_ELASTICITY_SOLVER_H_
#define GLIA_ELASTICITY_SOLVER_H_

#include <memory>
#include <vector>

#include "../glia_el.h"
#include "Tumor.h"

namespace glia {

class VariableLinearElasticitySolver {
 public:
  VariableLinearElasticitySolver(std::shared_ptr<Parameters> params, std::shared_ptr<Tumor> tumor,
                                 std::shared_ptr<SpectralOperators> /*spec_ops*/ = {})
      : params_(params), tumor_(tumor) {
    rc_ = glia_el_create(params->grid_->n_, nullptr, &h_);
  }
  ~VariableLinearElasticitySolver() {
    if (h_) glia_el_destroy(h_);
  }
  VariableLinearElasticitySolver(const VariableLinearElasticitySolver&) = delete;
  VariableLinearElasticitySolver& operator=(const VariableLinearElasticitySolver&) = delete;

  /* computeMaterialProperties (:376-455) */
  PetscErrorCode computeMaterialProperties() {
    if (rc_) return rc_;
    auto mp = tumor_->mat_prop_;
    if (!mp) return GLIA_RD_ERR_STATE;
    std::vector<float> c, wm, gm, csf, vt;
    PetscErrorCode rc;
    if ((rc = to_float(tumor_->c_t_, c)) || (rc = to_float(mp->wm_, wm)) || (rc = to_float(mp->gm_, gm)) ||
        (rc = to_float(mp->csf_, csf)) || (rc = to_float(mp->vt_, vt)))
      return rc;
    std::vector<float> bg(c.size());
    for (size_t i = 0; i < bg.size(); i++) bg[i] = 1.0f - (gm[i] + wm[i] + vt[i] + csf[i]);
    const auto& tu = *params_->tu_;
    glia_el_params p;
    glia_el_params_init(&p);
    p.E_healthy = tu.E_healthy_;
    p.nu_healthy = tu.nu_healthy_;
    p.E_tumor = tu.E_tumor_;
    p.nu_tumor = tu.nu_tumor_;
    p.E_bg = tu.E_bg_;
    p.nu_bg = tu.nu_bg_;
    p.E_csf = tu.E_csf_;
    p.nu_csf = tu.nu_csf_;
    p.screen_low = tu.screen_low_;
    p.screen_high = tu.screen_high_;
    return glia_el_set_material(h_, c.data(), wm.data(), gm.data(), csf.data(), vt.data(), bg.data(), &p,
                                GLIA_RD_HOST);
  }
  /* solve (:471-533) */
  PetscErrorCode solve(std::shared_ptr<VecField> displacement, std::shared_ptr<VecField> rhs) {
    if (rc_) return rc_;
    if (!displacement || !rhs) return GLIA_RD_ERR_ARG;
    PetscErrorCode rc = computeMaterialProperties();
    if (rc) return rc;
    Field3 f, u;
    if ((rc = f.load(*rhs)) || (rc = u.load(*displacement))) return rc;
    double rnorm = 0;
    if ((rc = glia_el_solve(h_, f.cptr(), u.ptr(), GLIA_RD_HOST, &iterations_, &reason_, &rnorm))) return rc;
    return u.store(*displacement);
  }
  /* operatorVariableCoefficients (:268-374): f = A u, with the last material properties */
  PetscErrorCode applyOperator(std::shared_ptr<VecField> u_in, std::shared_ptr<VecField> f_out) {
    Field3 u, f;
    PetscErrorCode rc;
    if ((rc = u.load(*u_in)) || (rc = f.load(*f_out))) return rc;
    if ((rc = glia_el_apply(h_, u.cptr(), f.ptr(), GLIA_RD_HOST))) return rc;
    return f.store(*f_out);
  }
  /* operatorConstantCoefficients (:122-266): u = B f */
  PetscErrorCode applyPreconditioner(std::shared_ptr<VecField> f_in, std::shared_ptr<VecField> u_out) {
    Field3 f, u;
    PetscErrorCode rc;
    if ((rc = f.load(*f_in)) || (rc = u.load(*u_out))) return rc;
    if ((rc = glia_el_precondition(h_, f.cptr(), u.ptr(), GLIA_RD_HOST))) return rc;
    return u.store(*u_out);
  }
  const char* lastError() const { return h_ ? glia_el_last_error(h_) : "no handle"; }

  int iterations_ = 0, reason_ = 0;  // of the last solve (KSPGetIterationNumber, KSPConvergedReason)

 private:
  static PetscErrorCode to_float(Vec x, std::vector<float>& out) {
    if (!x) return GLIA_RD_ERR_ARG;
    ScalarType* p = nullptr;
    PetscErrorCode rc = VecGetArray(x, &p);
    if (rc) return rc;
    out.assign(p, p + x->n);
    return VecRestoreArray(x, &p);
  }
  struct Field3 {
    std::vector<float> c[3];
    PetscErrorCode load(VecField& v) {
      Vec comp[3] = {v.x_, v.y_, v.z_};
      for (int a = 0; a < 3; a++) {
        PetscErrorCode rc = to_float(comp[a], c[a]);
        if (rc) return rc;
      }
      return 0;
    }
    PetscErrorCode store(VecField& v) {
      Vec comp[3] = {v.x_, v.y_, v.z_};
      for (int a = 0; a < 3; a++) {
        ScalarType* p = nullptr;
        PetscErrorCode rc = VecGetArray(comp[a], &p);
        if (rc) return rc;
        for (size_t i = 0; i < c[a].size(); i++) p[i] = c[a][i];
        if ((rc = VecRestoreArray(comp[a], &p))) return rc;
      }
      return 0;
    }
    const float* const* cptr() {
      for (int a = 0; a < 3; a++) cp_[a] = c[a].data();
      return cp_;
    }
    float* const* ptr() {
      for (int a = 0; a < 3; a++) p_[a] = c[a].data();
      return p_;
    }
    const float* cp_[3];
    float* p_[3];
  };
  std::shared_ptr<Parameters> params_;
  std::shared_ptr<Tumor> tumor_;
  glia_el_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

}  // namespace glia

#endif
