/*
 * AdvectionSolver.h -- the reference's SemiLagrangianSolver (include/pde/AdvectionSolver.h,
 * src/pde/AdvectionSolver.cpp:279-704) over the C ABI of include/glia_adv.h.
 *
 * Same members and call as the reference: advection_mode_ (1 mass conservation, the default; 2 pure
 * advection), trajectoryIsComputed_, and solve(scalar, velocity, dt), which computes the
 * trajectories from the velocity it is given when trajectoryIsComputed_ is false and then advances
 * the scalar in place.
 */
#ifndef GLIA_ADVECTION_SOLVER_H_
#define GLIA_ADVECTION_SOLVER_H_

#include <memory>
#include <vector>

#include "../glia_adv.h"
#include "Tumor.h"

namespace glia {

class SemiLagrangianSolver {
 public:
  SemiLagrangianSolver(std::shared_ptr<Parameters> params, std::shared_ptr<Tumor> /*tumor*/,
                       std::shared_ptr<SpectralOperators> /*spec_ops*/ = {})
      : params_(params) {
    rc_ = glia_sl_create(params->grid_->n_, nullptr, &h_);
  }
  ~SemiLagrangianSolver() {
    if (h_) glia_sl_destroy(h_);
  }
  SemiLagrangianSolver(const SemiLagrangianSolver&) = delete;
  SemiLagrangianSolver& operator=(const SemiLagrangianSolver&) = delete;

  /* solve (src/pde/AdvectionSolver.cpp:592-680) */
  PetscErrorCode solve(Vec scalar, std::shared_ptr<VecField> velocity, ScalarType dt) {
    if (rc_) return rc_;
    if (!scalar || !velocity) return GLIA_RD_ERR_ARG;
    const int64_t n = params_->grid_->ng_;
    if (scalar->n != n) return GLIA_RD_ERR_ARG;
    PetscErrorCode rc;
    if (!trajectoryIsComputed_) {
      std::vector<float> v[3];
      Vec comp[3] = {velocity->x_, velocity->y_, velocity->z_};
      for (int a = 0; a < 3; a++)
        if ((rc = to_float(comp[a], v[a]))) return rc;
      if ((rc = glia_sl_set_velocity(h_, v[0].data(), v[1].data(), v[2].data(), GLIA_RD_HOST))) return rc;
      if ((rc = glia_sl_compute_trajectories(h_, dt))) return rc;
      trajectoryIsComputed_ = true;
    }
    std::vector<float> c;
    if ((rc = to_float(scalar, c))) return rc;
    if ((rc = glia_sl_solve(h_, c.data(), dt, advection_mode_, GLIA_RD_HOST))) return rc;
    ScalarType* p = nullptr;
    if ((rc = VecGetArray(scalar, &p))) return rc;
    for (int64_t i = 0; i < n; i++) p[i] = c[i];
    return VecRestoreArray(scalar, &p);
  }
  const char* lastError() const { return h_ ? glia_sl_last_error(h_) : "no handle"; }

  int advection_mode_ = GLIA_SL_MASS_CONSERVATION;
  bool trajectoryIsComputed_ = false;

 private:
  static PetscErrorCode to_float(Vec x, std::vector<float>& out) {
    ScalarType* p = nullptr;
    PetscErrorCode rc = VecGetArray(x, &p);
    if (rc) return rc;
    out.assign(p, p + x->n);
    return VecRestoreArray(x, &p);
  }
  std::shared_ptr<Parameters> params_;
  glia_sl_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

}  // namespace glia

#endif
