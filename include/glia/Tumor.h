/*
 * Tumor.h -- the reference's Tumor and its coefficient classes for the RD path (include/mat/),
 * over one glia_rd handle: MatProp, DiffCoef, ReacCoef, Phi, Obs, Data, Tumor.
 *
 * Same class names, method names and argument meaning as the reference; the fields these classes
 * expose (wm_, kxx_, rho_vec_, c_0_, c_t_, p_0_, ...) are Vecs bound to the handle's device fields.
 * Differences, each because the engine owns the data: work vectors are not lent
 * (setWorkVecs accepts and ignores them); Phi takes its Gaussian centers explicitly; Obs filters are
 * applied when the data reach the handle (in the derivative operators).
 */
#ifndef GLIA_TUMOR_H_
#define GLIA_TUMOR_H_

#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../glia_io.h"
#include "Parameters.h"

namespace glia {

enum { ACCFFT = 0, CUFFT = 1 };

/* SpectralOperators: the transforms live inside the engine; the object only carries the choice */
class SpectralOperators {
 public:
  explicit SpectralOperators(int fft_mode = CUFFT) : fft_mode_(fft_mode) {}
  int fft_mode_;
};

namespace detail {
inline PetscErrorCode to_f32(Vec x, std::vector<float>& out) {
  if (!x) return 0;
  PetscErrorCode rc = x->pull();
  if (rc) return rc;
  out.assign(x->v.begin(), x->v.end());
  return 0;
}
}  // namespace detail

/* MatProp (src/mat/MatProp.cpp) */
class MatProp {
 public:
  MatProp(std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops = {}) : params_(params) {
    (void)spec_ops;
    rc_ = getHandle(params, &h_);
    if (!rc_) {
      const int64_t n = params->grid_->ng_;
      VecCreateField(h_, 0, GLIA_RD_FIELD_WM, n, &wm_);
      VecCreateField(h_, 0, GLIA_RD_FIELD_GM, n, &gm_);
      VecCreateField(h_, 0, GLIA_RD_FIELD_CSF, n, &csf_);
      VecCreateField(h_, 0, GLIA_RD_FIELD_VT, n, &vt_);
      VecCreateField(h_, 0, GLIA_RD_FIELD_FILTER, n, &filter_);
    }
  }
  ~MatProp() {
    VecDestroy(&wm_);
    VecDestroy(&gm_);
    VecDestroy(&csf_);
    VecDestroy(&vt_);
    VecDestroy(&filter_);
  }
  MatProp(const MatProp&) = delete;
  MatProp& operator=(const MatProp&) = delete;

  /* setValues (:70-132) with no atlas: the handle's tissue stays as set */
  PetscErrorCode setValues(std::shared_ptr<Parameters>) { return rc_; }
  /* setValuesCustom (:134-203): bg is derived (1 - gm - wm - vt - csf) and ignored */
  PetscErrorCode setValuesCustom(Vec gm, Vec wm, Vec csf, Vec vt, Vec bg, std::shared_ptr<Parameters>) {
    (void)bg;
    if (rc_) return rc_;
    std::vector<float> g, w, c, v;
    PetscErrorCode rc;
    if ((rc = detail::to_f32(gm, g)) || (rc = detail::to_f32(wm, w)) || (rc = detail::to_f32(csf, c)) ||
        (rc = detail::to_f32(vt, v)))
      return rc;
    auto ptr = [](std::vector<float>& a) { return a.empty() ? nullptr : a.data(); };
    return glia_rd_set_tissue(h_, 0, ptr(w), ptr(g), ptr(c), ptr(v), GLIA_RD_HOST);
  }
  /* setValuesSinusoidal (:34-68) */
  PetscErrorCode setValuesSinusoidal(std::shared_ptr<Parameters>) {
    return rc_ ? rc_ : glia_rd_set_tissue_sinusoidal(h_, 0);
  }

  Vec gm_ = nullptr, wm_ = nullptr, csf_ = nullptr, vt_ = nullptr, filter_ = nullptr;
  std::shared_ptr<Parameters> params_;
  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

/* DiffCoef (src/mat/DiffCoef.cpp), isotropic */
class DiffCoef {
 public:
  DiffCoef(std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops = {}) : params_(params) {
    (void)spec_ops;
    rc_ = getHandle(params, &h_);
    if (!rc_) VecCreateField(h_, 0, GLIA_RD_FIELD_K, params->grid_->ng_, &kxx_);
  }
  ~DiffCoef() { VecDestroy(&kxx_); }
  DiffCoef(const DiffCoef&) = delete;
  DiffCoef& operator=(const DiffCoef&) = delete;

  /* setValues (:77-132): kxx = k_scale (m_wm + k_gm_wm m_gm + k_glm_wm m_csf), kxx_avg_ */
  PetscErrorCode setValues(ScalarType k_scale, ScalarType k_gm_wm_ratio, ScalarType k_glm_wm_ratio,
                           std::shared_ptr<MatProp>, std::shared_ptr<Parameters> params) {
    if (rc_) return rc_;
    const auto& tu = *params->tu_;
    PetscErrorCode rc = glia_rd_set_ratios(h_, k_gm_wm_ratio, k_glm_wm_ratio, tu.r_gm_wm_ratio_, tu.r_glm_wm_ratio_);
    if (rc || (rc = glia_rd_set_kappa(h_, 0, k_scale))) return rc;
    params->ctx_->kappa = k_scale;
    k_scale_ = k_scale;
    return update_avg();
  }
  /* setValuesSinusoidal (:134-189) */
  PetscErrorCode setValuesSinusoidal(std::shared_ptr<Parameters>, ScalarType scale) {
    if (rc_) return rc_;
    PetscErrorCode rc = glia_rd_set_k_sinusoidal(h_, 0, scale);
    return rc ? rc : update_avg();
  }
  /* updateIsotropicCoefficients (:191-204) with nk = 1: k1 is the scale, the ratios stay */
  PetscErrorCode updateIsotropicCoefficients(ScalarType k1, ScalarType, ScalarType, std::shared_ptr<MatProp> mp,
                                             std::shared_ptr<Parameters> params) {
    return setValues(k1, params->tu_->k_gm_wm_ratio_, params->tu_->k_glm_wm_ratio_, mp, params);
  }
  /* applyD (:249-271): dc = div(k grad c) */
  PetscErrorCode applyD(Vec dc, Vec c) {
    if (rc_) return rc_;
    Vec in = nullptr, out = nullptr;
    VecCreateField(h_, 0, GLIA_RD_FIELD_WORK, c->n, &in);
    VecCreateField(h_, 0, GLIA_RD_FIELD_WORK, c->n, &out);
    PetscErrorCode rc = VecCopy(c, in);
    if (!rc) rc = glia_rd_apply_d(h_, GLIA_RD_FIELD_WORK, GLIA_RD_FIELD_WORK);
    if (!rc) rc = VecCopy(out, dc);
    VecDestroy(&in);
    VecDestroy(&out);
    return rc;
  }
  /* the engine owns its work memory */
  PetscErrorCode setWorkVecs(Vec*) { return 0; }

  Vec kxx_ = nullptr;
  ScalarType kxx_avg_ = 0, k_scale_ = 0;
  std::shared_ptr<Parameters> params_;
  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;

 private:
  PetscErrorCode update_avg() {
    glia_rd_stats_t st;
    PetscErrorCode rc = glia_rd_stats(h_, 0, &st);
    if (!rc) kxx_avg_ = st.kbar;
    return rc;
  }
};

/* ReacCoef (src/mat/ReacCoef.cpp) */
class ReacCoef {
 public:
  ReacCoef(std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops = {}) : params_(params) {
    (void)spec_ops;
    rc_ = getHandle(params, &h_);
    if (!rc_) VecCreateField(h_, 0, GLIA_RD_FIELD_RHO, params->grid_->ng_, &rho_vec_);
  }
  ~ReacCoef() { VecDestroy(&rho_vec_); }
  ReacCoef(const ReacCoef&) = delete;
  ReacCoef& operator=(const ReacCoef&) = delete;

  /* setValues (:13-38): rho = rho_scale (m_wm + r_gm_wm m_gm + r_glm_wm m_csf) */
  PetscErrorCode setValues(ScalarType rho_scale, ScalarType r_gm_wm_ratio, ScalarType r_glm_wm_ratio,
                           std::shared_ptr<MatProp>, std::shared_ptr<Parameters> params) {
    if (rc_) return rc_;
    const auto& tu = *params->tu_;
    PetscErrorCode rc = glia_rd_set_ratios(h_, tu.k_gm_wm_ratio_, tu.k_glm_wm_ratio_, r_gm_wm_ratio, r_glm_wm_ratio);
    if (rc || (rc = glia_rd_set_rho(h_, 0, rho_scale))) return rc;
    params->ctx_->rho = rho_scale;
    rho_scale_ = rho_scale;
    return 0;
  }
  /* updateIsotropicCoefficients (:40-50) with nr = 1 */
  PetscErrorCode updateIsotropicCoefficients(ScalarType r1, ScalarType, ScalarType, std::shared_ptr<MatProp> mp,
                                             std::shared_ptr<Parameters> params) {
    return setValues(r1, params->tu_->r_gm_wm_ratio_, params->tu_->r_glm_wm_ratio_, mp, params);
  }

  Vec rho_vec_ = nullptr;
  ScalarType rho_scale_ = 0;
  std::shared_ptr<Parameters> params_;
  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

/* Phi (src/mat/Phi.cpp): Gaussian basis of the initial condition */
class Phi {
 public:
  Phi(std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops = {}) : params_(params) {
    (void)spec_ops;
    rc_ = getHandle(params, &h_);
  }
  /* setGaussians with explicit centers (x, y, z in [0, 2 pi), 3 np values) and sigma: the basis
     the reference builds from centers (Phi::setValues, :78-144) */
  PetscErrorCode setGaussians(const std::vector<ScalarType>& centers, ScalarType sigma) {
    if (rc_) return rc_;
    if (centers.size() % 3) return GLIA_RD_ERR_ARG;
    np_ = (int)(centers.size() / 3);
    centers_ = centers;
    sigma_ = sigma;
    params_->tu_->np_ = np_;
    params_->tu_->phi_sigma_ = sigma;
    return glia_rd_set_basis(h_, 0, np_, centers.data(), sigma);
  }
  /* setGaussians(cm, sigma, spacing_factor, np) (src/mat/Phi.cpp:33-73): the bounding-box basis
     of phiMesh (:148-234) around cm */
  PetscErrorCode setGaussians(std::array<ScalarType, 3>& user_cm, ScalarType sigma, ScalarType spacing_factor, int np) {
    if (rc_) return rc_;
    if (np < 1) return GLIA_RD_ERR_ARG;
    std::vector<ScalarType> c(3 * (size_t)np);
    const int k = glia_phi_mesh(np, user_cm.data(), sigma, spacing_factor, c.data());
    c.resize(3 * (size_t)k);
    spacing_factor_ = spacing_factor;
    params_->tu_->phi_spacing_factor_ = spacing_factor;
    return setGaussians(c, sigma);
  }
  /* setGaussians(Vec data) (:615-940): the data-driven support on the lattice of spacing
     2 phi_sigma_data_driven, candidates kept when more than gaussian_vol_frac of their ball is
     above data_threshold (counted on the device); every Gaussian gets component label 1 */
  PetscErrorCode setGaussians(Vec data) {
    if (rc_) return rc_;
    if (!data) return GLIA_RD_ERR_ARG;
    Vec w = nullptr;
    PetscErrorCode rc = 0;
    int which = data->which, member = data->member;
    if (!(data->h == h_ && which >= 0)) {  // stage a caller vector in the work field
      which = GLIA_RD_FIELD_WORK, member = 0;
      VecCreateField(h_, 0, which, data->n, &w);
      rc = VecCopy(data, w);
      VecDestroy(&w);
      if (rc) return rc;
    }
    const auto& tu = *params_->tu_;
    int k = 0;
    rc = glia_rd_data_support(h_, member, which, tu.data_threshold_, tu.phi_sigma_data_driven_, tu.gaussian_vol_frac_,
                              0, nullptr, &k);
    if (rc && k == 0) return rc;
    std::vector<ScalarType> c(3 * (size_t)k);
    if ((rc = glia_rd_data_support(h_, member, which, tu.data_threshold_, tu.phi_sigma_data_driven_,
                                   tu.gaussian_vol_frac_, k, c.data(), &k)))
      return rc;
    spacing_factor_ = 2.0;
    params_->tu_->phi_spacing_factor_ = 2.0;
    gaussian_labels_.assign((size_t)k, 1);
    return k ? setGaussians(c, tu.phi_sigma_data_driven_) : GLIA_RD_ERR_ARG;
  }
  /* setGaussians(file, read_comp_data) (:525-612): centers (and component labels) from a
     phi-mesh file (readPhiMesh, src/utils/IO.cpp:717-790), sigma = phi_sigma_data_driven */
  PetscErrorCode setGaussians(std::string file, bool read_comp_data = false) {
    if (rc_) return rc_;
    int k = 0;
    if (glia_read_phi_mesh(file.c_str(), 0, nullptr, nullptr, &k, nullptr, nullptr) == -1 || k < 1) return 1;
    std::vector<ScalarType> c(3 * (size_t)k);
    std::vector<int> comps((size_t)k, 1);
    if (glia_read_phi_mesh(file.c_str(), k, c.data(), comps.data(), &k, nullptr, nullptr)) return 1;
    spacing_factor_ = 2.0;
    params_->tu_->phi_spacing_factor_ = 2.0;
    if (read_comp_data) gaussian_labels_ = comps;
    return setGaussians(c, params_->tu_->phi_sigma_data_driven_);
  }
  /* the basis is built when the centers are set and rebuilt when the tissue changes */
  PetscErrorCode setValues(std::shared_ptr<MatProp>) { return rc_; }
  /* apply (:313-377): out = Phi p (the handle's c_0_ receives it too) */
  PetscErrorCode apply(Vec out, Vec p) {
    if (rc_) return rc_;
    if (!out || !p || p->n != np_) return GLIA_RD_ERR_ARG;
    PetscErrorCode rc = p->pull();
    if (rc || (rc = glia_rd_phi_apply(h_, p->v.data()))) return rc;
    if (out->h == h_ && out->which == GLIA_RD_FIELD_C0) return 0;
    Vec c0 = nullptr;
    VecCreateField(h_, 0, GLIA_RD_FIELD_C0, out->n, &c0);
    rc = VecCopy(c0, out);
    VecDestroy(&c0);
    return rc;
  }
  /* applyTranspose (:379-439): pout_i = <phi_i, in> */
  PetscErrorCode applyTranspose(Vec pout, Vec in) {
    if (rc_) return rc_;
    if (!pout || !in || pout->n != np_) return GLIA_RD_ERR_ARG;
    Vec w = nullptr;
    VecCreateField(h_, 0, GLIA_RD_FIELD_WORK, in->n, &w);
    PetscErrorCode rc = VecCopy(in, w);
    VecDestroy(&w);
    pout->ensure();
    if (rc || (rc = glia_rd_phi_apply_transpose(h_, GLIA_RD_FIELD_WORK, pout->v.data()))) return rc;
    return pout->push();
  }

  int np_ = 0;
  std::vector<ScalarType> centers_;
  ScalarType spacing_factor_ = 0;
  std::vector<int> gaussian_labels_;
  ScalarType sigma_ = 0;
  std::shared_ptr<Parameters> params_;
  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

/* Obs (src/mat/Obs.cpp): O = 1[data > thr] (setDefaultFilter) or a custom mask; all ones until set */
class Obs {
 public:
  Obs(std::shared_ptr<Parameters> params) : params_(params) {}
  PetscErrorCode setDefaultFilter(Vec, int time_point = 1, ScalarType thr = -1) {
    if (time_point != 1) return GLIA_RD_ERR_UNSUPPORTED;
    thr_ = thr;
    custom_ = nullptr;
    return 0;
  }
  PetscErrorCode setCustomFilter(Vec custom_filter, int time_point = 1) {
    if (time_point != 1) return GLIA_RD_ERR_UNSUPPORTED;
    custom_ = custom_filter;
    return 0;
  }
  ScalarType thr_ = -std::numeric_limits<ScalarType>::infinity();
  Vec custom_ = nullptr;  // borrowed
  std::shared_ptr<Parameters> params_;
};

/* Data (include/utils/Utils.h:99-124): borrowed observation Vecs */
class Data {
 public:
  void set(Vec dt1, Vec dt0) {
    t1 = dt1;
    if (dt0 != nullptr) {
      t0 = dt0;
      two_snapshot = true;
    }
  }
  Vec dt1() { return t1; }
  Vec dt0() { return two_snapshot ? t0 : nullptr; }
  Vec t1 = nullptr, t0 = nullptr;
  bool two_snapshot = false;
};

/* Tumor (src/mat/Tumor.cpp) */
class Tumor {
 public:
  Tumor(std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops = {})
      : params_(params), spec_ops_(spec_ops) {
    rc_ = getHandle(params, &h_);
    k_ = std::make_shared<DiffCoef>(params, spec_ops);
    rho_ = std::make_shared<ReacCoef>(params, spec_ops);
    obs_ = std::make_shared<Obs>(params);
    if (!rc_) {
      const int64_t n = params->grid_->ng_;
      VecCreateField(h_, 0, GLIA_RD_FIELD_C0, n, &c_0_);
      VecCreateField(h_, 0, GLIA_RD_FIELD_CT, n, &c_t_);
      VecCreateField(h_, 0, GLIA_RD_FIELD_ALPHA0, n, &p_0_);
      VecCreateSeq(PETSC_COMM_SELF, (PetscInt)n, &p_t_);
      // velocity and work field of the advection models (src/mat/Tumor.cpp:35-38)
      const int model = params->tu_->model_;
      if (params->tu_->adv_velocity_set_ || model == 4 || model == 5) {
        velocity_ = std::make_shared<VecField>(n, n);
        work_field_ = std::make_shared<VecField>(n, n);
      }
    }
  }
  ~Tumor() {
    VecDestroy(&c_0_);
    VecDestroy(&c_t_);
    VecDestroy(&p_0_);
    VecDestroy(&p_t_);
    VecDestroy(&p_);
  }
  Tumor(const Tumor&) = delete;
  Tumor& operator=(const Tumor&) = delete;

  /* initialize (:59-80): material properties, coefficients from params, p, basis */
  PetscErrorCode initialize(Vec p, std::shared_ptr<Parameters> params, std::shared_ptr<SpectralOperators> spec_ops,
                            std::shared_ptr<Phi> phi = {}, std::shared_ptr<MatProp> mat_prop = {}) {
    if (rc_) return rc_;
    PetscErrorCode rc;
    mat_prop_ = mat_prop ? mat_prop : std::make_shared<MatProp>(params, spec_ops);
    if (!mat_prop && (rc = mat_prop_->setValues(params))) return rc;
    const auto& tu = *params->tu_;
    if ((rc = k_->setValues(tu.k_, tu.k_gm_wm_ratio_, tu.k_glm_wm_ratio_, mat_prop_, params))) return rc;
    if ((rc = rho_->setValues(tu.rho_, tu.r_gm_wm_ratio_, tu.r_glm_wm_ratio_, mat_prop_, params))) return rc;
    if (p) {
      VecDestroy(&p_);
      if ((rc = VecDuplicate(p, &p_)) || (rc = VecCopy(p, p_))) return rc;
    }
    if (phi) {
      phi_ = phi;
      return phi_->setValues(mat_prop_);
    }
    // no basis given: the bounding box of np_ Gaussians around user_cm_ (Tumor.cpp:72-75)
    phi_ = std::make_shared<Phi>(params, spec_ops);
    std::array<ScalarType, 3> cm = tu.user_cm_;
    if ((rc = phi_->setGaussians(cm, tu.phi_sigma_, tu.phi_spacing_factor_, tu.np_))) return rc;
    return phi_->setValues(mat_prop_);
  }
  /* setSinusoidalCoefficients (:102-112) */
  PetscErrorCode setSinusoidalCoefficients(std::shared_ptr<Parameters> params) {
    if (!mat_prop_) mat_prop_ = std::make_shared<MatProp>(params, spec_ops_);
    PetscErrorCode rc = mat_prop_->setValuesSinusoidal(params);
    const auto& tu = *params->tu_;
    if (rc || (rc = k_->setValues(tu.k_, tu.k_gm_wm_ratio_, tu.k_glm_wm_ratio_, mat_prop_, params))) return rc;
    if ((rc = rho_->setValues(tu.rho_, tu.r_gm_wm_ratio_, tu.r_glm_wm_ratio_, mat_prop_, params))) return rc;
    return phi_ ? phi_->setValues(mat_prop_) : 0;
  }

  std::shared_ptr<VecField> velocity_, work_field_;
  std::shared_ptr<DiffCoef> k_;
  std::shared_ptr<ReacCoef> rho_;
  std::shared_ptr<Phi> phi_;
  std::shared_ptr<Obs> obs_;
  std::shared_ptr<MatProp> mat_prop_;
  std::shared_ptr<Parameters> params_;
  std::shared_ptr<SpectralOperators> spec_ops_;
  Vec c_t_ = nullptr, c_0_ = nullptr, p_t_ = nullptr, p_0_ = nullptr, p_ = nullptr;
  glia_rd_t h_ = nullptr;
  PetscErrorCode rc_ = 0;
};

}  // namespace glia

#endif
