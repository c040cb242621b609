// The reference's own unit tests of the reaction-diffusion path, run through the C++ interface in
// include/glia/ (the reference's class names over libglia_rd.so), plus checks of that interface
// against the C ABI and against finite differences.
//
//   "Running diffusion solver"       src/test/pdesolver.cpp:7-57 (||c||_2 = 2.0487, 5 CG iterations)
//   "Running advection solver"       src/test/pdesolver.cpp:59-108 (round trip below 1%)
//   VariableLinearElasticitySolver: the solve meets its stopping test through the interface
//   "Evaluating objective function"  src/test/grad.cpp:7-80 (J = 0 at the truth, J > 0 otherwise)
//
//   TILOptimizer (quasi- and Gauss-Newton), RDOptimizer, SparseTILOptimizer on exact data (src/opt/*.cpp through include/glia/Optimizer.h)
//
// Exit status: the number of failed checks.  Needs a GPU.
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "glia/glia.h"

using namespace glia;

static int failures = 0, checks = 0;
#define REQUIRE(cond)                                                        \
  do {                                                                       \
    checks++;                                                                \
    if (!(cond)) {                                                           \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);         \
      failures++;                                                            \
    }                                                                        \
  } while (0)

// Catch2's Approx with its default epsilon (100 FLT_EPSILON, relative to the larger magnitude)
struct Approx {
  double v;
  explicit Approx(double x) : v(x) {}
};
static bool operator==(double a, Approx b) {
  return std::fabs(a - b.v) <= 100 * FLT_EPSILON * std::max(std::fabs(a), std::fabs(b.v));
}

// helper.cpp: createPVec, createTestFunction (a Gaussian of radius sqrt(2) 2 pi / 64 at the centre)
static PetscErrorCode createPVec(Vec& x, std::shared_ptr<Parameters> params) {
  const int n_inv = 1 + params->get_nk() + params->get_nr();
  PetscErrorCode ierr = VecCreateSeq(PETSC_COMM_SELF, n_inv, &x);
  if (ierr) return ierr;
  return VecSet(x, 0);
}
static PetscErrorCode createTestFunction(Vec x, std::shared_ptr<Parameters> params) {
  ScalarType* x_ptr;
  const ScalarType R = std::sqrt(2.) * (2 * M_PI) / 64;
  PetscErrorCode ierr = VecGetArray(x, &x_ptr);
  if (ierr) return ierr;
  const Grid& g = *params->grid_;
  for (int X = 0; X < g.isize_[0]; X++)
    for (int Y = 0; Y < g.isize_[1]; Y++)
      for (int Z = 0; Z < g.isize_[2]; Z++) {
        const int64_t ptr = ((int64_t)X * g.isize_[1] + Y) * g.isize_[2] + Z;
        const double dx = g.h_[0] * X - M_PI, dy = g.h_[1] * Y - M_PI, dz = g.h_[2] * Z - M_PI;
        const double ratio = std::sqrt(dx * dx + dy * dy + dz * dz) / R;
        x_ptr[ptr] = std::exp(-ratio * ratio);
      }
  return VecRestoreArray(x, &x_ptr);
}

// ---- src/test/pdesolver.cpp:7-57 -------------------------------------------------------------
static void test_diffusion_solver() {
  std::printf("TEST Running diffusion solver\n");
  auto params = std::make_shared<Parameters>();
  auto spec_ops = std::make_shared<SpectralOperators>(CUFFT);
  initializeGrid(64, params);
  auto k = std::make_shared<DiffCoef>(params, spec_ops);
  Vec* work = new Vec[8];
  for (int i = 0; i < 8; i++) {
    VecDuplicate(k->kxx_, &work[i]);
    VecSet(work[i], 0);
  }
  k->setWorkVecs(work);
  REQUIRE(k->setValuesSinusoidal(params, 1E-2) == 0);
  auto diff_solver = std::make_shared<DiffusionSolver>(params, spec_ops, k);
  Vec c;
  VecDuplicate(work[0], &c);
  VecSet(c, 0);
  createTestFunction(c, params);
  REQUIRE(diff_solver->precFactor() == 0);
  for (int i = 0; i < 10; i++) REQUIRE(diff_solver->solve(c, 0.02) == 0);
  ScalarType nrm;
  VecNorm(c, NORM_2, &nrm);
  std::printf("  ||c||_2 = %.6f, last solve %d CG iterations\n", nrm, diff_solver->ksp_itr_);
  REQUIRE(nrm == Approx(2.0487));
  REQUIRE(diff_solver->ksp_itr_ == 5);
  for (int i = 0; i < 8; i++) VecDestroy(&work[i]);
  delete[] work;
  VecDestroy(&c);
}

// helper.cpp createTestField (src/test/helper.cpp:47-80)
static PetscErrorCode createTestField(std::shared_ptr<VecField> v, std::shared_ptr<Parameters> params) {
  ScalarType *x_ptr, *y_ptr, *z_ptr;
  PetscErrorCode ierr;
  if ((ierr = VecGetArray(v->x_, &x_ptr)) || (ierr = VecGetArray(v->y_, &y_ptr)) || (ierr = VecGetArray(v->z_, &z_ptr)))
    return ierr;
  const Grid& g = *params->grid_;
  const ScalarType freq = 4;
  for (int X = 0; X < g.isize_[0]; X++)
    for (int Y = 0; Y < g.isize_[1]; Y++)
      for (int Z = 0; Z < g.isize_[2]; Z++) {
        const int64_t i = ((int64_t)X * g.isize_[1] + Y) * g.isize_[2] + Z;
        const double a = freq * 2.0 * M_PI / g.n_[0] * X, b = freq * 2.0 * M_PI / g.n_[1] * Y,
                     c = freq * 2.0 * M_PI / g.n_[2] * Z;
        x_ptr[i] = 1 + 0.25 * std::sin(a) * std::sin(b) * std::sin(c);
        y_ptr[i] = 1.5 + 0.1 * std::sin(a) * std::cos(b) * std::sin(c);
        z_ptr[i] = 1 + std::cos(a) * std::cos(b) * std::sin(c);
      }
  VecRestoreArray(v->x_, &x_ptr);
  VecRestoreArray(v->y_, &y_ptr);
  return VecRestoreArray(v->z_, &z_ptr);
}

// ---- src/test/pdesolver.cpp:59-108 -----------------------------------------------------------
static void test_advection_solver() {
  std::printf("TEST Running advection solver\n");
  auto params = std::make_shared<Parameters>();
  auto spec_ops = std::make_shared<SpectralOperators>(CUFFT);
  ScalarType norm_c, norm;
  Vec c, c_copy;
  initializeGrid(256, params);
  params->tu_->model_ = 4;  // creates velocities
  auto tumor = std::make_shared<Tumor>(params, spec_ops);
  REQUIRE(tumor->velocity_ != nullptr);
  createTestField(tumor->velocity_, params);
  VecDuplicate(tumor->c_t_, &c);
  VecDuplicate(tumor->c_t_, &c_copy);
  VecSet(c, 0);
  createTestFunction(c, params);
  VecCopy(c, c_copy);
  VecNorm(c, NORM_2, &norm_c);
  auto adv_solver = std::make_shared<SemiLagrangianSolver>(params, tumor, spec_ops);
  adv_solver->advection_mode_ = 2;  // pure advection
  for (int i = 0; i < 4; i++) REQUIRE(adv_solver->solve(c, tumor->velocity_, 0.02) == 0);
  tumor->velocity_->scale(-1);
  adv_solver->trajectoryIsComputed_ = false;
  for (int i = 0; i < 4; i++) REQUIRE(adv_solver->solve(c, tumor->velocity_, 0.02) == 0);
  VecAXPY(c, -1, c_copy);
  VecNorm(c, NORM_2, &norm);
  std::printf("  ||c - c0|| / ||c0|| = %.4e\n", norm / norm_c);
  REQUIRE(norm / norm_c < 1E-2);
  VecDestroy(&c);
  VecDestroy(&c_copy);
}

// ---- VariableLinearElasticitySolver (src/pde/ElasticitySolver.cpp:6-533) ---------------------------
static void test_elasticity_solver() {
  std::printf("TEST VariableLinearElasticitySolver through the interface\n");
  auto params = std::make_shared<Parameters>();
  auto spec_ops = std::make_shared<SpectralOperators>(CUFFT);
  initializeGrid(32, params);
  params->tu_->model_ = 4;
  auto tumor = std::make_shared<Tumor>(params, spec_ops);
  tumor->mat_prop_ = std::make_shared<MatProp>(params, spec_ops);
  REQUIRE(tumor->mat_prop_->setValuesSinusoidal(params) == 0);
  createTestFunction(tumor->c_t_, params);
  // force along the tumor gradient: f = -100 grad c for c = exp(-|x - pi|^2 / R^2)
  const int64_t ng = params->grid_->ng_;
  auto force = std::make_shared<VecField>(ng, ng), disp = std::make_shared<VecField>(ng, ng);
  disp->set(0);
  {
    const Grid& g = *params->grid_;
    const double R = std::sqrt(2.) * (2 * M_PI) / 64;
    ScalarType *fx, *fy, *fz;
    VecGetArray(force->x_, &fx);
    VecGetArray(force->y_, &fy);
    VecGetArray(force->z_, &fz);
    for (int X = 0; X < g.n_[0]; X++)
      for (int Y = 0; Y < g.n_[1]; Y++)
        for (int Z = 0; Z < g.n_[2]; Z++) {
          const int64_t i = ((int64_t)X * g.n_[1] + Y) * g.n_[2] + Z;
          const double d[3] = {g.h_[0] * X - M_PI, g.h_[1] * Y - M_PI, g.h_[2] * Z - M_PI};
          const double c = std::exp(-(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / (R * R));
          fx[i] = 100 * 2 * d[0] / (R * R) * c;
          fy[i] = 100 * 2 * d[1] / (R * R) * c;
          fz[i] = 100 * 2 * d[2] / (R * R) * c;
        }
    VecRestoreArray(force->x_, &fx);
    VecRestoreArray(force->y_, &fy);
    VecRestoreArray(force->z_, &fz);
  }
  VariableLinearElasticitySolver el(params, tumor, spec_ops);
  REQUIRE(el.solve(disp, force) == 0);
  // the solve's stopping test, checked with the interface's own operator and preconditioner
  auto r = std::make_shared<VecField>(ng, ng), z = std::make_shared<VecField>(ng, ng), bf = std::make_shared<VecField>(ng, ng);
  REQUIRE(el.applyOperator(disp, r) == 0);
  VecScale(r->x_, -1);
  VecScale(r->y_, -1);
  VecScale(r->z_, -1);
  VecAXPY(r->x_, 1, force->x_);
  VecAXPY(r->y_, 1, force->y_);
  VecAXPY(r->z_, 1, force->z_);
  REQUIRE(el.applyPreconditioner(r, z) == 0);
  REQUIRE(el.applyPreconditioner(force, bf) == 0);
  auto norm3 = [](std::shared_ptr<VecField> v) {
    ScalarType a, b, c;
    VecNorm(v->x_, NORM_2, &a);
    VecNorm(v->y_, NORM_2, &b);
    VecNorm(v->z_, NORM_2, &c);
    return std::sqrt(a * a + b * b + c * c);
  };
  const double rz = norm3(z), rb = norm3(bf), un = norm3(disp);
  std::printf("  %d CG iterations, reason %d, ||B(f - A u)|| / ||B f|| = %.3e, ||u|| = %.3e\n", el.iterations_,
              el.reason_, rz / rb, un);
  REQUIRE(el.reason_ == 2);
  REQUIRE(el.iterations_ >= 1 && el.iterations_ < 100);
  REQUIRE(rz <= 1.01e-3 * rb);
  REQUIRE(un > 0);
}

// ---- src/test/grad.cpp:7-80 ------------------------------------------------------------------
static void test_objective() {
  std::printf("TEST Evaluating objective function\n");
  auto params = std::make_shared<Parameters>();
  params->tu_->dt_ = 0.01;
  params->tu_->nt_ = 100;
  params->tu_->k_ = 0.01;
  params->tu_->rho_ = 10;
  params->tu_->model_ = 1;
  params->opt_->flag_reaction_inv_ = true;  // disables phi apply in obj
  params->opt_->reaction_inversion_ = true;
  params->opt_->diffusivity_inversion_ = true;
  params->opt_->beta_ = 0;
  auto spec_ops = std::make_shared<SpectralOperators>(CUFFT);
  initializeGrid(64, params);

  Vec p;
  createPVec(p, params);
  auto tumor = std::make_shared<Tumor>(params, spec_ops);
  REQUIRE(tumor->initialize(p, params, spec_ops, nullptr, nullptr) == 0);
  REQUIRE(tumor->setSinusoidalCoefficients(params) == 0);
  std::shared_ptr<PdeOperators> pde_operators = std::make_shared<PdeOperatorsRD>(tumor, params, spec_ops);
  std::shared_ptr<DerivativeOperators> derivative_operators =
      std::make_shared<DerivativeOperatorsRD>(pde_operators, params, tumor);

  Vec x;
  VecDuplicate(tumor->c_t_, &x);
  VecSet(x, 0);
  createTestFunction(x, params);

  // test update ~ will change all coefficients to zero
  ScalarType normk, normr;
  ScalarType rt, kt;
  rt = params->tu_->k_;
  kt = params->tu_->k_;
  REQUIRE(derivative_operators->updateReactionAndDiffusion(p) == 0);
  VecNorm(tumor->k_->kxx_, NORM_2, &normk);
  VecNorm(tumor->rho_->rho_vec_, NORM_2, &normr);
  REQUIRE(normk == 0);
  REQUIRE(normr == 0);

  // set correct values to p
  ScalarType* p_ptr;
  VecGetArray(p, &p_ptr);
  p_ptr[1] = kt;
  p_ptr[params->tu_->nk_ + 1] = rt;
  VecRestoreArray(p, &p_ptr);

  // generate test data
  VecCopy(x, tumor->c_0_);
  REQUIRE(derivative_operators->updateReactionAndDiffusion(p) == 0);
  REQUIRE(pde_operators->solveState(0) == 0);
  VecCopy(tumor->c_t_, x);
  auto data = std::make_shared<Data>();
  data->set(x, nullptr);

  // compute objective with ground truth
  ScalarType j;
  REQUIRE(derivative_operators->evaluateObjective(&j, p, data) == 0);
  std::printf("  J(truth) = %.3e\n", j);
  REQUIRE(j < 1E-7);
  VecSet(x, 0);
  REQUIRE(derivative_operators->evaluateObjective(&j, p, data) == 0);
  std::printf("  J(zero data) = %.3e\n", j);
  REQUIRE(j > 0);
  VecDestroy(&x);
  VecDestroy(&p);
}

// ---- the interface against the C ABI, and g_p against central differences --------------------
struct TilSetup {
  std::shared_ptr<Parameters> params;
  std::shared_ptr<Tumor> tumor;
  std::shared_ptr<PdeOperators> pde;
  std::shared_ptr<DerivativeOperators> der;
  std::shared_ptr<Data> data;
  Vec d = nullptr;
  std::vector<double> centers;
  double sigma = 0;
};

static const int NP = 4;

static void make_til(TilSetup& s) {
  s.params = std::make_shared<Parameters>();
  auto& P = *s.params;
  P.tu_->dt_ = 0.02;
  P.tu_->nt_ = 5;
  P.tu_->k_ = 0.05;
  P.tu_->rho_ = 8;
  P.tu_->k_gm_wm_ratio_ = 0.2;
  P.tu_->r_gm_wm_ratio_ = 0.2;
  P.opt_->beta_ = 1e-3;
  auto spec_ops = std::make_shared<SpectralOperators>(CUFFT);
  initializeGrid(64, s.params);
  s.tumor = std::make_shared<Tumor>(s.params, spec_ops);
  s.tumor->initialize(nullptr, s.params, spec_ops);
  s.tumor->setSinusoidalCoefficients(s.params);
  const double off[NP][3] = {{0, 0, 0}, {0.3, 0, 0}, {0, 0.3, 0.1}, {-0.2, -0.2, 0.3}};
  for (int i = 0; i < NP; i++)
    for (int d = 0; d < 3; d++) s.centers.push_back(M_PI + off[i][d]);
  s.sigma = 2 * M_PI / 32;
  s.tumor->phi_->setGaussians(s.centers, s.sigma);
  s.pde = std::make_shared<PdeOperatorsRD>(s.tumor, s.params, spec_ops);
  s.der = std::make_shared<DerivativeOperatorsRD>(s.pde, s.params, s.tumor);
  // data: c(T) of p* = (0.8, 0.5, 0, 0.3), thresholded at 0.3
  Vec pstar;
  VecCreateSeq(PETSC_COMM_SELF, NP, &pstar);
  double* pp;
  VecGetArray(pstar, &pp);
  pp[0] = 0.8, pp[1] = 0.5, pp[2] = 0.0, pp[3] = 0.3;
  VecRestoreArray(pstar, &pp);
  s.tumor->phi_->apply(s.tumor->c_0_, pstar);
  s.pde->solveState(0);
  VecDuplicate(s.tumor->c_t_, &s.d);
  VecCopy(s.tumor->c_t_, s.d);
  double* dp;
  VecGetArray(s.d, &dp);
  for (int64_t i = 0; i < s.d->n; i++) dp[i] = dp[i] > 0.3 ? 1.0 : 0.0;
  VecRestoreArray(s.d, &dp);
  s.data = std::make_shared<Data>();
  s.data->set(s.d, nullptr);
  VecDestroy(&pstar);
}

static void test_interface_matches_abi_and_fd() {
  std::printf("TEST TIL objective and gradient through the interface\n");
  TilSetup s;
  make_til(s);
  const int n = NP + 2;
  Vec x, g;
  VecCreateSeq(PETSC_COMM_SELF, n, &x);
  VecCreateSeq(PETSC_COMM_SELF, n, &g);
  const double x0[n] = {0.6, 0.6, 0.2, 0.4, 0.05, 8.0};
  double* xp;
  VecGetArray(x, &xp);
  for (int i = 0; i < n; i++) xp[i] = x0[i];
  VecRestoreArray(x, &xp);
  // a fresh solver context for every evaluation compared (the preconditioner uses the context dt)
  double J = 0;
  REQUIRE(s.der->evaluateObjectiveAndGradient(&J, g, x, s.data) == 0);

  // the same evaluation through the C ABI on its own handle
  glia_rd_config c;
  glia_rd_config_init(&c);
  c.n[0] = c.n[1] = c.n[2] = 64;
  c.dt = 0.02;
  c.nt = 5;
  c.k_gm_wm = 0.2;
  c.r_gm_wm = 0.2;
  c.beta = 1e-3;
  glia_rd_t h = nullptr;
  REQUIRE(glia_rd_create(&c, &h) == 0);
  glia_rd_set_tissue_sinusoidal(h, 0);
  glia_rd_set_coefficients(h, 0, 0.05, 8.0);
  glia_rd_set_basis(h, 0, NP, s.centers.data(), s.sigma);
  std::vector<float> d1(s.d->n);
  double* dp;
  VecGetArray(s.d, &dp);
  for (int64_t i = 0; i < s.d->n; i++) d1[i] = (float)dp[i];
  VecRestoreArray(s.d, &dp);
  glia_rd_set_data(h, 0, d1.data(), nullptr, -INFINITY, GLIA_RD_HOST);
  double Ja = 0, ga[n];
  REQUIRE(glia_rd_objective_and_gradient(h, GLIA_RD_MODE_P_KAPPA, x0, &Ja, ga) == 0);
  glia_rd_destroy(h);
  double* gp;
  VecGetArray(g, &gp);
  double err = 0, nrm = 0;
  for (int i = 0; i < n; i++) {
    err += (gp[i] - ga[i]) * (gp[i] - ga[i]);
    nrm += ga[i] * ga[i];
  }
  std::printf("  J %.9e (ABI %.9e), |dJ - dJ_ABI| / |dJ_ABI| = %.2e, g_rho = %g\n", J, Ja, std::sqrt(err / nrm), gp[n - 1]);
  REQUIRE(std::fabs(J - Ja) <= 1e-6 * std::fabs(Ja));
  REQUIRE(std::sqrt(err / nrm) < 1e-5);
  REQUIRE(gp[n - 1] == 0.0);  // TIL: [g_p | g_kappa | 0]

  // g_p against central differences of J (the reference's checkGradient procedure)
  std::mt19937 rng(7);
  std::normal_distribution<double> nd;
  double gpd[NP];
  for (int i = 0; i < NP; i++) gpd[i] = gp[i];
  VecRestoreArray(g, &gp);
  for (int t = 0; t < 3; t++) {
    double dir[NP], gd = 0;
    for (int i = 0; i < NP; i++) {
      dir[i] = nd(rng);
      gd += gpd[i] * dir[i];
    }
    const double hstep = 1e-2;
    double Jp = 0, Jm = 0;
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      VecGetArray(x, &xp);
      for (int i = 0; i < NP; i++) xp[i] = x0[i] + sgn * hstep * dir[i];
      VecRestoreArray(x, &xp);
      REQUIRE(s.der->evaluateObjective(sgn < 0 ? &Jm : &Jp, x, s.data) == 0);
    }
    const double fd = (Jp - Jm) / (2 * hstep);
    std::printf("  direction %d: FD %.6e, g.d %.6e, rel %.2e\n", t, fd, gd, std::fabs(fd - gd) / std::fabs(fd));
    REQUIRE(std::fabs(fd - gd) <= 1e-3 * std::fabs(fd));
  }
  VecDestroy(&x);
  VecDestroy(&g);
  VecDestroy(&s.d);
}

// ---- the rest of the PdeOperators / DerivativeOperators surface -------------------------------
// gradDiffusion after an evaluation; the public history c_, p_ (views of the device history);
// solveState(1) against a finite difference of solveState(0); resizeTimeHistory through reset;
// checkGradient's Taylor remainders (src/grad/DerivativeOperators.cpp:369-450)
static void test_boundary_methods() {
  std::printf("TEST PdeOperators / DerivativeOperators boundary methods\n");
  TilSetup s;
  make_til(s);
  const int n = NP + 2, nt = s.params->tu_->nt_;
  Vec x, g, g2;
  VecCreateSeq(PETSC_COMM_SELF, n, &x);
  VecCreateSeq(PETSC_COMM_SELF, n, &g);
  VecCreateSeq(PETSC_COMM_SELF, n, &g2);
  const double x0[n] = {0.6, 0.6, 0.2, 0.4, 0.05, 8.0};
  double* xp;
  VecGetArray(x, &xp);
  for (int i = 0; i < n; i++) xp[i] = x0[i];
  VecRestoreArray(x, &xp);
  double J = 0;
  REQUIRE(s.der->evaluateObjectiveAndGradient(&J, g, x, s.data) == 0);
  VecSet(g2, 0.0);
  REQUIRE(s.der->gradDiffusion(g2) == 0);
  double *gp, *g2p;
  VecGetArray(g, &gp);
  VecGetArray(g2, &g2p);
  std::printf("  gradDiffusion %.9e, evaluation g_kappa %.9e\n", g2p[NP], gp[NP]);
  REQUIRE(g2p[NP] == gp[NP]);
  VecRestoreArray(g, &gp);
  VecRestoreArray(g2, &g2p);
  REQUIRE(s.der->gradReaction(g2) == 0);  // no rho inversion here: dJ is left as it is
  // history views: c_[nt] is c(T), p_[0] is alpha(0)
  REQUIRE((int)s.pde->c_.size() == nt + 1 && (int)s.pde->p_.size() == nt + 1);
  double a = 0, b = 0;
  VecNorm(s.pde->c_[nt], NORM_2, &a);
  VecNorm(s.tumor->c_t_, NORM_2, &b);
  REQUIRE(a == b && a > 0);
  VecNorm(s.pde->p_[0], NORM_2, &a);
  VecNorm(s.tumor->p_0_, NORM_2, &b);
  REQUIRE(a == b && a > 0);
  VecNorm(s.pde->c_[0], NORM_2, &a);
  VecNorm(s.tumor->c_0_, NORM_2, &b);
  REQUIRE(std::fabs(a - b) <= 1e-6 * b);
  // solveState(1) from c~0 = c(0) against (c(T; (1 + e) c0) - c(T; c0)) / e
  const int64_t ng = s.params->grid_->ng_;
  Vec c0, ct_base, ct_lin;
  VecDuplicate(s.tumor->c_0_, &c0);
  VecDuplicate(s.tumor->c_0_, &ct_base);
  VecDuplicate(s.tumor->c_0_, &ct_lin);
  VecCopy(s.tumor->c_0_, c0);
  VecCopy(s.tumor->c_t_, ct_base);
  REQUIRE(s.pde->solveState(1) == 0);
  VecCopy(s.tumor->c_t_, ct_lin);
  const double e = 1e-2;
  VecScale(c0, 1.0 + e);
  VecCopy(c0, s.tumor->c_0_);
  REQUIRE(s.pde->solveState(0) == 0);
  double *pl, *pb, *pp;
  VecGetArray(ct_lin, &pl);
  VecGetArray(ct_base, &pb);
  VecGetArray(s.tumor->c_t_, &pp);
  double err = 0, nrm = 0;
  for (int64_t i = 0; i < ng; i++) {
    const double fd = (pp[i] - pb[i]) / e;
    err += (fd - pl[i]) * (fd - pl[i]);
    nrm += pl[i] * pl[i];
  }
  std::printf("  solveState(1) vs finite difference of solveState(0): rel %.2e\n", std::sqrt(err / nrm));
  REQUIRE(std::sqrt(err / nrm) < 2e-2);
  VecDestroy(&c0);
  VecDestroy(&ct_base);
  VecDestroy(&ct_lin);
  // checkGradient: Taylor remainders fall for h = 1, 0.1, 0.01
  REQUIRE(s.der->checkGradient(x, s.data) == 0);
  auto* dro = dynamic_cast<DerivativeOperatorsRD*>(s.der.get());
  REQUIRE(dro && dro->taylor_err_.size() == 6);
  if (dro && dro->taylor_err_.size() == 6)
    REQUIRE(dro->taylor_err_[1] < 0.2 * dro->taylor_err_[0] && dro->taylor_err_[2] < 0.2 * dro->taylor_err_[1]);
  // resizeTimeHistory through reset: the views follow the new nt and evaluations run
  s.params->tu_->nt_ = 8;
  REQUIRE(s.pde->reset(s.params, s.tumor) == 0);
  REQUIRE((int)s.pde->c_.size() == 9);
  REQUIRE(s.der->reset(x, s.pde, s.params, s.tumor) == 0);
  double J8 = 0;
  REQUIRE(s.der->evaluateObjectiveAndGradient(&J8, g, x, s.data) == 0);
  VecNorm(s.pde->c_[8], NORM_2, &a);
  VecNorm(s.tumor->c_t_, NORM_2, &b);
  std::printf("  nt 5 -> 8: J %.6e -> %.6e\n", J, J8);
  REQUIRE(a == b && J8 != J);
  VecDestroy(&x);
  VecDestroy(&g);
  VecDestroy(&g2);
  VecDestroy(&s.d);
}

// ---- Phi apply / transpose and the diffusion operator through the interface -------------------
static void test_phi_and_apply_d() {
  std::printf("TEST Phi and DiffCoef::applyD through the interface\n");
  TilSetup s;
  make_til(s);
  const int64_t ng = s.params->grid_->ng_;
  Vec p, c0, v, pt;
  VecCreateSeq(PETSC_COMM_SELF, NP, &p);
  VecCreateSeq(PETSC_COMM_SELF, NP, &pt);
  VecCreateSeq(PETSC_COMM_SELF, (PetscInt)ng, &c0);
  VecCreateSeq(PETSC_COMM_SELF, (PetscInt)ng, &v);
  double* pp;
  // one active basis: Phi p is that basis normalized by its own maximum (Phi::apply normalizes by
  // the maximum over the active bases), so max c0 = 1
  VecSet(p, 0.0);
  VecGetArray(p, &pp);
  pp[1] = 1.0;
  VecRestoreArray(p, &pp);
  REQUIRE(s.tumor->phi_->apply(c0, p) == 0);
  double mx = 0;
  VecMax(c0, nullptr, &mx);
  REQUIRE(std::fabs(mx - 1.0) < 1e-6);
  // every basis active: Phi and Phi^T normalize by the same global maximum, so
  // <Phi^T v, p> = <v, Phi p> for any field v
  VecGetArray(p, &pp);
  pp[0] = 1.0, pp[1] = 0.5, pp[2] = 0.25, pp[3] = 2.0;
  VecRestoreArray(p, &pp);
  REQUIRE(s.tumor->phi_->apply(c0, p) == 0);
  createTestFunction(v, s.params);
  REQUIRE(s.tumor->phi_->applyTranspose(pt, v) == 0);
  double lhs = 0, rhs = 0;
  VecDot(pt, p, &lhs);
  VecDot(v, c0, &rhs);
  std::printf("  <Phi^T v, p> = %.8e, <v, Phi p> = %.8e\n", lhs, rhs);
  REQUIRE(std::fabs(lhs - rhs) <= 1e-5 * std::fabs(rhs));
  // applyD is linear: D(2 c0) = 2 D(c0)
  Vec d1, d2;
  VecCreateSeq(PETSC_COMM_SELF, (PetscInt)ng, &d1);
  VecCreateSeq(PETSC_COMM_SELF, (PetscInt)ng, &d2);
  REQUIRE(s.tumor->k_->applyD(d1, c0) == 0);
  VecScale(c0, 2.0);
  REQUIRE(s.tumor->k_->applyD(d2, c0) == 0);
  VecAXPY(d2, -2.0, d1);
  double e = 0, n1 = 0;
  VecNorm(d2, NORM_2, &e);
  VecNorm(d1, NORM_2, &n1);
  std::printf("  ||D(2c) - 2D(c)|| / ||D(c)|| = %.2e\n", e / n1);
  REQUIRE(n1 > 0 && e <= 1e-5 * n1);
  // the constant Hessian approximation H = beta Phi^T Phi is symmetric and
  // <u, H u> = beta ||Phi u||^2 (u = p from above, v a second coefficient vector)
  Vec q, hp, hq;
  VecDuplicate(p, &q);
  VecDuplicate(p, &hp);
  VecDuplicate(p, &hq);
  VecGetArray(q, &pp);
  pp[0] = -0.3, pp[1] = 1.5, pp[2] = 0.7, pp[3] = 0.1;
  VecRestoreArray(q, &pp);
  REQUIRE(s.tumor->phi_->apply(c0, p) == 0);
  double phin = 0;
  VecNorm(c0, NORM_2, &phin);
  REQUIRE(s.der->evaluateConstantHessianApproximation(hp, p) == 0);
  REQUIRE(s.der->evaluateConstantHessianApproximation(hq, q) == 0);
  double qhp = 0, phq = 0, php = 0;
  VecDot(q, hp, &qhp);
  VecDot(p, hq, &phq);
  VecDot(p, hp, &php);
  const double beta = s.params->opt_->beta_;
  std::printf("  <q, H p> = %.8e, <H q, p> = %.8e, <p, H p> = %.8e, beta ||Phi p||^2 = %.8e\n", qhp, phq, php,
              beta * phin * phin);
  REQUIRE(std::fabs(qhp - phq) <= 1e-5 * std::fabs(qhp));
  REQUIRE(std::fabs(php - beta * phin * phin) <= 1e-5 * php);
  VecDestroy(&q);
  VecDestroy(&hp);
  VecDestroy(&hq);
  VecDestroy(&p);
  VecDestroy(&pt);
  VecDestroy(&c0);
  VecDestroy(&v);
  VecDestroy(&d1);
  VecDestroy(&d2);
  VecDestroy(&s.d);
}

// Phi::setGaussians in its three placements (src/mat/Phi.cpp:33-73, 525-612, 615-940)
static void test_basis_placement() {
  std::printf("TEST Phi::setGaussians: bounding box, data-driven support, phi-mesh file\n");
  TilSetup s;
  make_til(s);
  auto& tu = *s.params->tu_;
  // bounding box: 27 Gaussians around cm, spaced 1.5 sigma
  std::array<ScalarType, 3> cm = {M_PI, M_PI, M_PI};
  REQUIRE(s.tumor->phi_->setGaussians(cm, 2 * M_PI / 64, 1.5, 27) == 0);
  REQUIRE(s.tumor->phi_->np_ == 27 && tu.np_ == 27);
  REQUIRE(std::fabs(s.tumor->phi_->centers_[3 * 13] - M_PI) < 1e-15);
  // data-driven: radius 2 cells, candidates every 4 cells; brute force over the host copy of d
  tu.phi_sigma_data_driven_ = 2 * (2 * M_PI / 64);
  tu.data_threshold_ = 0.05;
  tu.gaussian_vol_frac_ = 0.0;
  REQUIRE(s.tumor->phi_->setGaussians(s.d) == 0);
  double* dp = nullptr;
  VecGetArray(s.d, &dp);
  std::vector<double> want;
  const double h = 2 * M_PI / 64;
  for (int x = 0; x <= 60; x += 4)
    for (int y = 0; y <= 60; y += 4)
      for (int z = 0; z <= 60; z += 4) {
        int cnt = 0;
        for (int i = -2; i <= 2; i++)
          for (int j = -2; j <= 2; j++)
            for (int k = -2; k <= 2; k++) {
              const int a = x + i, b = y + j, c = z + k;
              if (i * i + j * j + k * k > 4 || a < 0 || b < 0 || c < 0 || a >= 64 || b >= 64 || c >= 64) continue;
              cnt += (float)dp[((size_t)a * 64 + b) * 64 + c] > 0.05f;
            }
        if (cnt > 0) want.insert(want.end(), {x * h, y * h, z * h});
      }
  VecRestoreArray(s.d, &dp);
  std::printf("  data-driven support: %d centers (brute force %d)\n", s.tumor->phi_->np_, (int)(want.size() / 3));
  REQUIRE(!want.empty() && s.tumor->phi_->centers_ == want && tu.np_ == (int)(want.size() / 3));
  REQUIRE(s.tumor->phi_->gaussian_labels_.size() == want.size() / 3);
  // phi-mesh file with component labels
  const char* path = "/tmp/glia_test_phi_mesh.dat";
  FILE* f = std::fopen(path, "w");
  REQUIRE(f != nullptr);
  std::fprintf(f, " sigma = %.9g, spacing = %.9g\n centers = [\n", tu.phi_sigma_data_driven_, 2 * tu.phi_sigma_data_driven_);
  for (int i = 0; i < 3; i++) std::fprintf(f, " %.17g, %.17g, %.17g, %d\n", M_PI + 0.1 * i, M_PI, M_PI - 0.1 * i, i + 1);
  std::fprintf(f, "];\n");
  std::fclose(f);
  REQUIRE(s.tumor->phi_->setGaussians(std::string(path), true) == 0);
  REQUIRE(s.tumor->phi_->np_ == 3 && s.tumor->phi_->gaussian_labels_ == std::vector<int>({1, 2, 3}));
  REQUIRE(std::fabs(s.tumor->phi_->centers_[3] - (M_PI + 0.1)) < 1e-15);
  std::remove(path);
  VecDestroy(&s.d);
}

// ---- the inversion classes on exact data ------------------------------------------------------
// the TIL problem of make_til with exact data d = c(T) of p* and the truth kappa = 0.05, rho = 8
static void make_exact(TilSetup& s) {
  make_til(s);
  Vec pstar;
  VecCreateSeq(PETSC_COMM_SELF, NP, &pstar);
  double* pp;
  VecGetArray(pstar, &pp);
  pp[0] = 0.8, pp[1] = 0.5, pp[2] = 0.0, pp[3] = 0.3;
  VecRestoreArray(pstar, &pp);
  s.tumor->phi_->apply(s.tumor->c_0_, pstar);
  s.pde->solveState(0);
  VecCopy(s.tumor->c_t_, s.d);
  VecDestroy(&pstar);
}

static void test_til_optimizer(int solver) {
  std::printf("TEST TILOptimizer (p and kappa, %s) against glia_rd_invert\n",
              solver == GAUSSNEWTON ? "Gauss-Newton" : "quasi-Newton");
  TilSetup s;
  make_exact(s);
  s.params->opt_->opttolgrad_ = 1e-5;
  s.params->opt_->newton_solver_ = solver;
  TILOptimizer til;
  til.initialize(s.der, s.pde, s.params, s.tumor);
  til.setData(s.data);
  Vec x0;
  VecCreateSeq(PETSC_COMM_SELF, NP + 2, &x0);
  const double start[NP + 2] = {0.5, 0.5, 0.5, 0.5, 0.05, 8.0};
  double* xp;
  VecGetArray(x0, &xp);
  for (int i = 0; i < NP + 2; i++) xp[i] = start[i];
  VecRestoreArray(x0, &xp);
  REQUIRE(til.setInitialGuess(x0) == 0);
  REQUIRE(til.solve() == 0);
  Vec xo = til.getSolution();
  REQUIRE(xo != nullptr && xo->n == NP + 1);
  std::printf("  TIL: p = (%.4f %.4f %.4f %.4f), kappa %.5f, reason %d, %d iterations, J %.3e\n", xo->v[0], xo->v[1],
              xo->v[2], xo->v[3], xo->v[4], til.result_.reason, til.result_.iterations, til.result_.f);
  REQUIRE(til.result_.reason > 0);
  // the same inversion through the C ABI on a handle of its own
  glia_rd_config c;
  glia_rd_config_init(&c);
  c.n[0] = c.n[1] = c.n[2] = 64;
  c.dt = 0.02;
  c.nt = 5;
  c.k_gm_wm = 0.2;
  c.r_gm_wm = 0.2;
  c.beta = 1e-3;
  glia_rd_t h = nullptr;
  REQUIRE(glia_rd_create(&c, &h) == 0);
  glia_rd_set_tissue_sinusoidal(h, 0);
  glia_rd_set_coefficients(h, 0, 0.05, 8.0);
  glia_rd_set_basis(h, 0, NP, s.centers.data(), s.sigma);
  std::vector<float> d1(s.d->n);
  double* dp;
  VecGetArray(s.d, &dp);
  for (int64_t i = 0; i < s.d->n; i++) d1[i] = (float)dp[i];
  VecRestoreArray(s.d, &dp);
  glia_rd_set_data(h, 0, d1.data(), nullptr, -INFINITY, GLIA_RD_HOST);
  double xa[NP + 2], lb[NP + 2], ub[NP + 2];
  for (int i = 0; i < NP + 2; i++) {
    xa[i] = start[i];
    lb[i] = 0.0;
    ub[i] = INFINITY;
  }
  lb[NP] = 1e-3, ub[NP] = 1.0, lb[NP + 1] = ub[NP + 1] = 8.0;
  glia_opt_settings os;
  glia_opt_settings_init(&os);
  os.gttol = 1e-5;
  os.newton_solver = solver;
  glia_opt_result r;
  REQUIRE(glia_rd_invert(h, GLIA_RD_MODE_P_KAPPA, xa, lb, ub, &os, &r) == 0);
  glia_rd_destroy(h);
  double dmax = 0;
  for (int i = 0; i <= NP; i++) dmax = std::max(dmax, std::fabs(xa[i] - xo->v[i]));
  std::printf("  ABI: p = (%.4f %.4f %.4f %.4f), kappa %.5f, %d iterations; max |difference| %.2e\n", xa[0], xa[1],
              xa[2], xa[3], xa[4], r.iterations, dmax);
  REQUIRE(dmax < 1e-3);
  VecDestroy(&x0);
  VecDestroy(&s.d);
}

static void test_rd_optimizer() {
  std::printf("TEST RDOptimizer (kappa and rho) recovers the truth\n");
  TilSetup s;
  make_exact(s);
  s.params->opt_->opttolgrad_ = 1e-6;
  s.params->opt_->gatol_ = 1e-12;  // the default absolute tolerance (1e-4) stops 0.3% short here
  RDOptimizer rdo;
  rdo.initialize(s.der, s.pde, s.params, s.tumor);
  rdo.setData(s.data);
  Vec x0;
  VecCreateSeq(PETSC_COMM_SELF, NP + 2, &x0);
  const double start[NP + 2] = {0.8, 0.5, 0.0, 0.3, 0.1, 5.0};
  double* xp;
  VecGetArray(x0, &xp);
  for (int i = 0; i < NP + 2; i++) xp[i] = start[i];
  VecRestoreArray(x0, &xp);
  REQUIRE(rdo.setInitialGuess(x0) == 0);
  REQUIRE(rdo.solve() == 0);
  Vec xo = rdo.getSolution();
  std::printf("  kappa %.6f rho %.5f, reason %d, %d iterations\n", xo->v[NP], xo->v[NP + 1], rdo.result_.reason,
              rdo.result_.iterations);
  REQUIRE(rdo.result_.reason > 0);
  REQUIRE(std::fabs(xo->v[NP] - 0.05) < 1e-3 * 0.05);
  REQUIRE(std::fabs(xo->v[NP + 1] - 8.0) < 1e-3 * 8.0);
  REQUIRE(s.params->opt_->beta_ == 1e-3 && s.params->tu_->np_ == NP);  // restored after the solve
  VecDestroy(&x0);
  VecDestroy(&s.d);
}

static void test_sparse_til_optimizer() {
  std::printf("TEST SparseTILOptimizer finds the active bases\n");
  TilSetup s;
  make_exact(s);
  s.params->tu_->sparsity_level_ = 3;
  s.params->opt_->opttolgrad_ = 1e-5;
  SparseTILOptimizer sp;
  sp.initialize(s.der, s.pde, s.params, s.tumor);
  sp.setData(s.data);
  Vec x0;
  VecCreateSeq(PETSC_COMM_SELF, NP + 2, &x0);
  VecSet(x0, 0.0);
  double* xp;
  VecGetArray(x0, &xp);
  xp[NP] = 0.05, xp[NP + 1] = 8.0;
  VecRestoreArray(x0, &xp);
  REQUIRE(sp.setInitialGuess(x0) == 0);
  REQUIRE(sp.solve() == 0);
  Vec xo = sp.getSolution();
  std::printf("  support:");
  for (int i : sp.support_) std::printf(" %d", i);
  std::printf("; p = (%.4f %.4f %.4f %.4f), CoSaMP iterations %d\n", xo->v[0], xo->v[1], xo->v[2], xo->v[3],
              sp.sparse_result_.iterations);
  REQUIRE((sp.support_ == std::vector<int>{0, 1, 3}));
  REQUIRE(xo->v[2] == 0.0);
  VecDestroy(&x0);
  VecDestroy(&s.d);
}

int main() {
  test_diffusion_solver();
  test_advection_solver();
  test_elasticity_solver();
  test_objective();
  test_interface_matches_abi_and_fd();
  test_boundary_methods();
  test_phi_and_apply_d();
  test_basis_placement();
  test_til_optimizer(QUASINEWTON);
  test_til_optimizer(GAUSSNEWTON);
  test_rd_optimizer();
  test_sparse_til_optimizer();
  std::printf("%d checks, %d failed\n", checks, failures);
  return failures;
}
