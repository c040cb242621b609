/*
 * Vec.h -- a PETSc-free vector for the reference's class interface over libglia_rd.
 *
 * The reference's RD classes take PETSc Vec handles.  PETSc is not part of this build, so the
 * classes in this directory take this host vector instead, with the same call names and meaning
 * (VecCreateSeq, VecDuplicate, VecSet, VecCopy, VecNorm, VecAXPY, VecGetArray, ...) and the same
 * error convention (0 = success).  A Vec can be bound to a field that a glia_rd handle owns
 * (Tumor::c_t_, DiffCoef::kxx_, ...): reading it pulls the field from the device, writing it pushes
 * the field back, so host code always sees what the device holds.  Host scalars are FP64; device
 * fields are FP32.
 */
#ifndef GLIA_VEC_H_
#define GLIA_VEC_H_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../glia_rd.h"

namespace glia {

using PetscErrorCode = int;
using PetscInt = int;
using PetscReal = double;
using PetscScalar = double;
using ScalarType = double;

enum NormType { NORM_1 = 0, NORM_2 = 1, NORM_INFINITY = 3 };
constexpr int PETSC_COMM_SELF = 0;

struct _p_Vec {
  int64_t n = 0;
  std::vector<ScalarType> v;   // allocated on first use
  glia_rd_t h = nullptr;        // bound field: handle, member, field id
  int member = 0, which = -1;
  int hist_kind = -1, hist_index = 0;  // bound history entry (GLIA_RD_HIST_*), read-only
  std::vector<float> stage;

  void ensure() {
    if ((int64_t)v.size() != n) v.assign(n, 0.0);
  }
  PetscErrorCode pull() {
    ensure();
    if (!h) return 0;
    stage.resize(n);
    const int rc = (hist_kind >= 0)
                       ? glia_rd_get_history(h, member, hist_kind, hist_index, stage.data(), GLIA_RD_HOST)
                       : glia_rd_get_field(h, member, which, stage.data(), GLIA_RD_HOST);
    if (rc) return rc;
    std::copy(stage.begin(), stage.end(), v.begin());
    return 0;
  }
  PetscErrorCode push() {
    if (!h || hist_kind >= 0) return 0;  // history views are read-only: writes stay on the host
    ensure();
    stage.assign(v.begin(), v.end());
    return glia_rd_set_field(h, member, which, stage.data(), GLIA_RD_HOST);
  }
};
typedef _p_Vec* Vec;

inline PetscErrorCode VecCreateSeq(int, PetscInt n, Vec* x) {
  if (!x || n < 0) return GLIA_RD_ERR_ARG;
  *x = new _p_Vec();
  (*x)->n = n;
  return 0;
}
/* a Vec bound to field 'which' of member 'member' of handle h (n = grid size) */
inline PetscErrorCode VecCreateField(glia_rd_t h, int member, int which, int64_t n, Vec* x) {
  if (!x || !h || n < 0) return GLIA_RD_ERR_ARG;
  *x = new _p_Vec();
  (*x)->n = n;
  (*x)->h = h;
  (*x)->member = member;
  (*x)->which = which;
  return 0;
}
/* a read-only view of history entry 'index' of kind GLIA_RD_HIST_* (c_, c_half_, p_) of a member */
inline PetscErrorCode VecCreateHistory(glia_rd_t h, int member, int kind, int index, int64_t n, Vec* x) {
  if (!x || !h || n < 0) return GLIA_RD_ERR_ARG;
  *x = new _p_Vec();
  (*x)->n = n;
  (*x)->h = h;
  (*x)->member = member;
  (*x)->hist_kind = kind;
  (*x)->hist_index = index;
  return 0;
}
inline PetscErrorCode VecDuplicate(Vec x, Vec* y) {
  if (!x) return GLIA_RD_ERR_ARG;
  return VecCreateSeq(PETSC_COMM_SELF, (PetscInt)x->n, y);
}
inline PetscErrorCode VecDestroy(Vec* x) {
  if (x && *x) {
    delete *x;
    *x = nullptr;
  }
  return 0;
}
inline PetscErrorCode VecGetSize(Vec x, PetscInt* n) {
  if (!x || !n) return GLIA_RD_ERR_ARG;
  *n = (PetscInt)x->n;
  return 0;
}
inline PetscErrorCode VecSet(Vec x, PetscScalar a) {
  if (!x) return GLIA_RD_ERR_ARG;
  x->ensure();
  std::fill(x->v.begin(), x->v.end(), a);
  return x->push();
}
inline PetscErrorCode VecCopy(Vec x, Vec y) {
  if (!x || !y || x->n != y->n) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc) return rc;
  if (x != y) y->v = x->v;
  return y->push();
}
inline PetscErrorCode VecNorm(Vec x, NormType t, PetscReal* r) {
  if (!x || !r) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc) return rc;
  double s = 0;
  for (double a : x->v) {
    if (t == NORM_1) s += std::fabs(a);
    else if (t == NORM_2) s += a * a;
    else s = std::max(s, std::fabs(a));
  }
  *r = (t == NORM_2) ? std::sqrt(s) : s;
  return 0;
}
inline PetscErrorCode VecDot(Vec x, Vec y, PetscScalar* r) {
  if (!x || !y || !r || x->n != y->n) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc || (rc = y->pull())) return rc;
  double s = 0;
  for (int64_t i = 0; i < x->n; i++) s += x->v[i] * y->v[i];
  *r = s;
  return 0;
}
inline PetscErrorCode VecSum(Vec x, PetscScalar* r) {
  if (!x || !r) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc) return rc;
  double s = 0;
  for (double a : x->v) s += a;
  *r = s;
  return 0;
}
inline PetscErrorCode VecMax(Vec x, PetscInt* idx, PetscReal* r) {
  if (!x || !r || x->n == 0) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc) return rc;
  const auto it = std::max_element(x->v.begin(), x->v.end());
  *r = *it;
  if (idx) *idx = (PetscInt)(it - x->v.begin());
  return 0;
}
/* y = y + a x */
inline PetscErrorCode VecAXPY(Vec y, PetscScalar a, Vec x) {
  if (!x || !y || x->n != y->n) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc || (rc = y->pull())) return rc;
  for (int64_t i = 0; i < y->n; i++) y->v[i] += a * x->v[i];
  return y->push();
}
inline PetscErrorCode VecScale(Vec x, PetscScalar a) {
  if (!x) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  if (rc) return rc;
  for (double& e : x->v) e *= a;
  return x->push();
}
inline PetscErrorCode VecGetArray(Vec x, ScalarType** p) {
  if (!x || !p) return GLIA_RD_ERR_ARG;
  PetscErrorCode rc = x->pull();
  *p = x->v.data();
  return rc;
}
inline PetscErrorCode VecRestoreArray(Vec x, ScalarType** p) {
  if (!x) return GLIA_RD_ERR_ARG;
  if (p) *p = nullptr;
  return x->push();
}

/* VecField (include/utils/Utils.h:127-148): three component vectors */
class VecField {
 public:
  VecField(int64_t nl, int64_t /*ng*/) {
    VecCreateSeq(PETSC_COMM_SELF, (PetscInt)nl, &x_);
    VecCreateSeq(PETSC_COMM_SELF, (PetscInt)nl, &y_);
    VecCreateSeq(PETSC_COMM_SELF, (PetscInt)nl, &z_);
  }
  ~VecField() {
    VecDestroy(&x_);
    VecDestroy(&y_);
    VecDestroy(&z_);
  }
  VecField(const VecField&) = delete;
  VecField& operator=(const VecField&) = delete;
  PetscErrorCode set(ScalarType a) {
    PetscErrorCode rc = VecSet(x_, a);
    if (!rc) rc = VecSet(y_, a);
    return rc ? rc : VecSet(z_, a);
  }
  PetscErrorCode scale(ScalarType a) {
    PetscErrorCode rc = VecScale(x_, a);
    if (!rc) rc = VecScale(y_, a);
    return rc ? rc : VecScale(z_, a);
  }
  Vec x_ = nullptr, y_ = nullptr, z_ = nullptr;
};

}  // namespace glia

#endif
