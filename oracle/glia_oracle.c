/*
 * glia_oracle.c -- FP64 CPU restatement of GLIA's reaction-diffusion forward,
 * adjoint and gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by
 * nothing in the product.  The product (glia_amd/, libglia_rd.so) never links,
 * imports or calls it.
 *
 * It restates the reference operation by operation, in the reference's own
 * real-space formulation (12 FFTs per CG iteration, PETSc KSPCG stopping
 * rules), so that its results are what the reference CPU path computes, in
 * FP64.  Every function cites the reference file:line it follows.
 *
 * Layout (reference src/grad/SpectralOperators.cpp:3-17): a real field is
 * n0 x n1 x n2, row-major with index x*n1*n2 + y*n2 + z; a spectrum is
 * n0 x n1 x (n2/2+1) interleaved complex, unnormalized forward transform.
 *
 * How it is pinned (tests/test_oracle.py, test_adv.py, test_elasticity.py,
 * test_mass_effect.py, test_multispecies.py, test_me_features.py,
 * test_sinusoid_simulator.py): the reference's own known answers
 * (src/test/pdesolver.cpp:7-108, src/test/grad.cpp:7-80 and the sinusoid
 * simulator src/test/simulator.cpp:5-58, which pins the model-4 loop), closed
 * forms, exact discrete adjoints against finite differences, invariants and
 * library cross-checks.  Parity unpinned: the model-5 loop (orc_ms_solve_state)
 * as a whole, for which the reference holds no known answer -- its parts and
 * its conservation and symmetry properties are checked; exact PETSc iteration
 * counts on brain-like k beyond the known-answer test; AccFFT's Nyquist
 * handling on the reference CPU path (the CUDA path's kernels are followed).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* 1D mixed-radix Stockham FFT (radix 4, 2, 3, 5) in FP64.                   */
/* The reference uses cuFFT / AccFFT+FFTW for this (src/grad/                */
/* SpectralOperators.cpp:43-50,68-98); any exact DFT is a valid restatement. */
/* ------------------------------------------------------------------------ */
typedef struct {
  int L;
  int nr;
  int radix[40];
  cplx *tw; /* tw[m] = exp(-2 pi i m / L) */
} fplan;

static int plan_init(fplan *p, int L) {
  p->L = L;
  p->nr = 0;
  int m = L;
  while (m % 4 == 0) { p->radix[p->nr++] = 4; m /= 4; }
  while (m % 2 == 0) { p->radix[p->nr++] = 2; m /= 2; }
  while (m % 3 == 0) { p->radix[p->nr++] = 3; m /= 3; }
  while (m % 5 == 0) { p->radix[p->nr++] = 5; m /= 5; }
  if (m != 1) return -1;
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)L);
  for (int i = 0; i < L; i++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)L;
    p->tw[i] = (double)cosl(a) + I * (double)sinl(a);
  }
  return 0;
}

static void plan_free(fplan *p) { free(p->tw); p->tw = NULL; }

/* small DFT of length R in natural order, sign = -1 forward, +1 inverse */
static void dft_small(cplx *v, int R, int sign) {
  cplx out[5];
  static const double s3 = 0.86602540378443864676372317075294;
  if (R == 2) {
    cplx a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
    return;
  }
  if (R == 4) {
    cplx a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = (v[1] - v[3]) * (sign < 0 ? -I : I);
    v[0] = a + c; v[2] = a - c; v[1] = b + d; v[3] = b - d;
    return;
  }
  if (R == 3) {
    cplx w1 = -0.5 + (sign < 0 ? -s3 : s3) * I;
    cplx w2 = conj(w1);
    out[0] = v[0] + v[1] + v[2];
    out[1] = v[0] + w1 * v[1] + w2 * v[2];
    out[2] = v[0] + w2 * v[1] + w1 * v[2];
    v[0] = out[0]; v[1] = out[1]; v[2] = out[2];
    return;
  }
  /* R == 5: direct */
  for (int q = 0; q < R; q++) {
    cplx s = 0;
    for (int r = 0; r < R; r++) {
      double a = sign * 2.0 * M_PI * (double)((r * q) % R) / (double)R;
      s += v[r] * (cos(a) + I * sin(a));
    }
    out[q] = s;
  }
  for (int q = 0; q < R; q++) v[q] = out[q];
}

/* in-place transform of a[0..L) using tmp[0..L) as ping-pong storage */
static void fft_exec(const fplan *p, cplx *a, cplx *tmp, int sign) {
  const int L = p->L;
  cplx *src = a, *dst = tmp;
  int Ns = 1;
  for (int s = 0; s < p->nr; s++) {
    const int R = p->radix[s];
    const int LR = L / R;
    const int stride = L / (Ns * R);
    for (int j = 0; j < LR; j++) {
      cplx v[5];
      const int k = j % Ns;
      for (int r = 0; r < R; r++) {
        cplx x = src[j + r * LR];
        if (k) {
          cplx w = p->tw[(k * r * stride) % L];
          x *= (sign < 0) ? w : conj(w);
        }
        v[r] = x;
      }
      dft_small(v, R, sign);
      const int idxD = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) dst[idxD + r * Ns] = v[r];
    }
    cplx *t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)L);
}

/* ------------------------------------------------------------------------ */
/* Oracle context: grid, coefficients, solver state, history.               */
/* Mirrors Parameters/Grid (include/Parameters.h:374-451), Tumor            */
/* (src/mat/Tumor.cpp:3-57), DiffCoef/ReacCoef/Obs/Phi/MatProp state.       */
/* ------------------------------------------------------------------------ */
typedef struct orc {
  int n[3];
  int64_t ng;   /* n0*n1*n2 */
  int n2h;      /* n2/2+1 */
  int64_t nh;   /* n0*n1*n2h complex entries */
  double h[3];
  double lebesgue;
  fplan pl[3];
  int nthreads;

  /* MatProp (src/mat/MatProp.cpp:134-203) */
  double *wm, *gm, *csf, *vt, *bg, *filter;
  double filter_sum;     /* VecSum(filter) used by DiffCoef::setValues */
  int k_filter_is_ng;    /* setValuesSinusoidal divides by ng (DiffCoef.cpp:170) */

  /* DiffCoef (isotropic: kxx=kyy=kzz, off-diagonals zero; DiffCoef.cpp:77-132) */
  double *k;
  double *ksec;  /* DiffCoef temp_[7]: secondary coefficient sum_i m_i k~_i (Hessian, diffusivity) */
  double kavg;
  double k_scale, k_gm_wm, k_glm_wm;
  /* ReacCoef (src/mat/ReacCoef.cpp:13-38) */
  double *rho;
  double rho_scale, r_gm_wm, r_glm_wm;

  /* DiffusionSolver (src/pde/DiffusionSolver.cpp:5-45) */
  double ctx_dt;    /* Ctx::dt_, params dt at construction, last tau after solve */
  double prec_dt;   /* dt value baked into the current precfactor_ */
  double *prec;     /* precfactor_ (nh real factors) */
  double rtol, atol, dtol;
  int maxit;
  int ksp_itr;      /* ksp_itr_ of last solve */
  int ksp_reason;   /* >0 converged, <0 diverged, 0 none */
  int64_t ksp_total_calls;

  /* time stepping (TumorParameters::dt_, nt_) */
  double dt;
  int nt;
  int adjoint_store;

  /* fields of Tumor */
  double *c_0, *c_t, *p_0, *p_t;
  /* history (PdeOperatorsRD ctor, src/pde/PdeOperators.cpp:17-54) */
  double **c_hist, **p_hist, **c_half;
  int hist_nt;
  int diff_ksp_itr_state, diff_ksp_itr_adj;

  /* Phi (src/mat/Phi.cpp) */
  int np;
  double *centers;
  double phi_sigma;
  double smoothing_factor;
  int phi_store;
  double **phi;      /* stored bases (normalized if phi_store, raw otherwise) */
  double phi_max_all;

  /* Obs (src/mat/Obs.cpp) and data */
  double *obs;
  double *d1;

  /* optimizer settings (OptimizerSettings) */
  double beta;
  double opt_k_scale, opt_rho_scale;

  /* work: w[0] operatorA/solve temp, w[1..6] applyD gradient/flux fields, w[7..9] CG r, z and the
     KSPConvergedDefault B*b vector, w[10] smoother filter, w[11] divergence accumulator */
  double *w[12];
  double *cg_p, *cg_w, *rhs, *adj_temp;
  cplx *xhat, *what;
} orc_t;

static double *dalloc(int64_t n) {
  double *p = (double *)calloc((size_t)n, sizeof(double));
  return p;
}

/* progress lines on stderr, one per time step of solveState / solveAdjoint (long oracle runs) */
static int orc_verbose = 0;
void orc_set_verbose(int v) { orc_verbose = v; }

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

orc_t *orc_create(int n0, int n1, int n2) {
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->n[0] = n0; o->n[1] = n1; o->n[2] = n2;
  o->ng = (int64_t)n0 * n1 * n2;
  o->n2h = n2 / 2 + 1;
  o->nh = (int64_t)n0 * n1 * o->n2h;
  for (int d = 0; d < 3; d++) {
    o->h[d] = 2.0 * M_PI / o->n[d];
    if (plan_init(&o->pl[d], o->n[d]) != 0) { free(o); return NULL; }
  }
  o->lebesgue = o->h[0] * o->h[1] * o->h[2];
  o->wm = dalloc(o->ng); o->gm = dalloc(o->ng); o->csf = dalloc(o->ng);
  o->vt = dalloc(o->ng); o->bg = dalloc(o->ng); o->filter = dalloc(o->ng);
  o->k = dalloc(o->ng); o->rho = dalloc(o->ng); o->ksec = dalloc(o->ng);
  o->prec = dalloc(o->nh);
  o->c_0 = dalloc(o->ng); o->c_t = dalloc(o->ng); o->p_0 = dalloc(o->ng); o->p_t = dalloc(o->ng);
  o->obs = dalloc(o->ng); o->d1 = dalloc(o->ng);
  for (int i = 0; i < 12; i++) o->w[i] = dalloc(o->ng);
  o->cg_p = dalloc(o->ng); o->cg_w = dalloc(o->ng); o->rhs = dalloc(o->ng); o->adj_temp = dalloc(o->ng);
  o->xhat = (cplx *)calloc((size_t)o->nh, sizeof(cplx));
  o->what = (cplx *)calloc((size_t)o->nh, sizeof(cplx));
  /* MatProp::setValues defaults (src/mat/MatProp.cpp:20-33): wm = 1, filter = 1 */
  for (int64_t i = 0; i < o->ng; i++) { o->wm[i] = 1.0; o->filter[i] = 1.0; o->obs[i] = 1.0; }
  o->filter_sum = (double)o->ng;
  /* defaults: TumorParameters (include/Parameters.h:157-258), DiffusionSolver ctor tolerances
     (src/pde/DiffusionSolver.cpp:26), PETSc KSP defaults atol = 1e-50, dtol = 1e5 */
  o->dt = 0.5; o->nt = 1;
  o->ctx_dt = o->dt; o->prec_dt = o->dt;
  o->rtol = 1e-6; o->atol = 1e-50; o->dtol = 1e5; o->maxit = 5000;
  o->k_scale = 1e-2; o->k_gm_wm = 1.0 / 5.0; o->k_glm_wm = 3.0 / 5.0;
  o->phi_sigma = 2 * M_PI / 64; o->smoothing_factor = 1.0;
  o->opt_k_scale = 1.0; o->opt_rho_scale = 1.0;
  return o;
}

static void free_hist(orc_t *o) {
  if (o->c_hist) {
    for (int i = 0; i <= o->hist_nt; i++) { free(o->c_hist[i]); free(o->p_hist[i]); }
    free(o->c_hist); free(o->p_hist);
  }
  if (o->c_half) {
    for (int i = 0; i < o->hist_nt; i++) free(o->c_half[i]);
    free(o->c_half);
  }
  o->c_hist = o->p_hist = o->c_half = NULL;
}

static void free_phi(orc_t *o) {
  if (o->phi) {
    for (int i = 0; i < o->np; i++) free(o->phi[i]);
    free(o->phi);
  }
  o->phi = NULL;
  free(o->centers); o->centers = NULL;
  o->np = 0;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  free_hist(o);
  free_phi(o);
  for (int d = 0; d < 3; d++) plan_free(&o->pl[d]);
  free(o->wm); free(o->gm); free(o->csf); free(o->vt); free(o->bg); free(o->filter);
  free(o->k); free(o->rho); free(o->prec); free(o->ksec);
  free(o->c_0); free(o->c_t); free(o->p_0); free(o->p_t); free(o->obs); free(o->d1);
  for (int i = 0; i < 12; i++) free(o->w[i]);
  free(o->cg_p); free(o->cg_w); free(o->rhs); free(o->adj_temp);
  free(o->xhat); free(o->what);
  free(o);
}

/* ------------------------------------------------------------------------ */
/* 3D R2C / C2R (cufftExecR2C / cufftExecC2R, unnormalized;                 */
/* src/grad/SpectralOperators.cpp:68-98).                                   */
/* ------------------------------------------------------------------------ */
void orc_r2c(orc_t *o, const double *f, cplx *F) {
  const int n0 = o->n[0], n1 = o->n[1], n2 = o->n[2], n2h = o->n2h;
#pragma omp parallel
  {
    int Lm = n0 > n1 ? n0 : n1; if (n2 > Lm) Lm = n2;
    cplx *a = (cplx *)malloc(sizeof(cplx) * Lm), *t = (cplx *)malloc(sizeof(cplx) * Lm);
#pragma omp for schedule(static)
    for (int64_t l = 0; l < (int64_t)n0 * n1; l++) { /* z lines */
      const double *src = f + l * n2;
      for (int z = 0; z < n2; z++) a[z] = src[z];
      fft_exec(&o->pl[2], a, t, -1);
      for (int kz = 0; kz < n2h; kz++) F[l * n2h + kz] = a[kz];
    }
#pragma omp for schedule(static)
    for (int64_t l = 0; l < (int64_t)n0 * n2h; l++) { /* y lines */
      int x = (int)(l / n2h), kz = (int)(l % n2h);
      cplx *base = F + (int64_t)x * n1 * n2h + kz;
      for (int y = 0; y < n1; y++) a[y] = base[(int64_t)y * n2h];
      fft_exec(&o->pl[1], a, t, -1);
      for (int y = 0; y < n1; y++) base[(int64_t)y * n2h] = a[y];
    }
#pragma omp for schedule(static)
    for (int64_t l = 0; l < (int64_t)n1 * n2h; l++) { /* x lines */
      cplx *base = F + l;
      for (int x = 0; x < n0; x++) a[x] = base[(int64_t)x * n1 * n2h];
      fft_exec(&o->pl[0], a, t, -1);
      for (int x = 0; x < n0; x++) base[(int64_t)x * n1 * n2h] = a[x];
    }
    free(a); free(t);
  }
}

/* F is destroyed (used as scratch), as cuFFT C2R is allowed to do. */
void orc_c2r(orc_t *o, cplx *F, double *f) {
  const int n0 = o->n[0], n1 = o->n[1], n2 = o->n[2], n2h = o->n2h;
#pragma omp parallel
  {
    int Lm = n0 > n1 ? n0 : n1; if (n2 > Lm) Lm = n2;
    cplx *a = (cplx *)malloc(sizeof(cplx) * Lm), *t = (cplx *)malloc(sizeof(cplx) * Lm);
#pragma omp for schedule(static)
    for (int64_t l = 0; l < (int64_t)n1 * n2h; l++) {
      cplx *base = F + l;
      for (int x = 0; x < n0; x++) a[x] = base[(int64_t)x * n1 * n2h];
      fft_exec(&o->pl[0], a, t, +1);
      for (int x = 0; x < n0; x++) base[(int64_t)x * n1 * n2h] = a[x];
    }
#pragma omp for schedule(static)
    for (int64_t l = 0; l < (int64_t)n0 * n2h; l++) {
      int x = (int)(l / n2h), kz = (int)(l % n2h);
      cplx *base = F + (int64_t)x * n1 * n2h + kz;
      for (int y = 0; y < n1; y++) a[y] = base[(int64_t)y * n2h];
      fft_exec(&o->pl[1], a, t, +1);
      for (int y = 0; y < n1; y++) base[(int64_t)y * n2h] = a[y];
    }
#pragma omp for schedule(static)
    for (int64_t l = 0; l < (int64_t)n0 * n1; l++) {
      const cplx *src = F + l * n2h;
      for (int kz = 0; kz < n2h; kz++) a[kz] = src[kz];
      for (int kz = n2h; kz < n2; kz++) a[kz] = conj(src[n2 - kz]);
      fft_exec(&o->pl[2], a, t, +1);
      double *dst = f + l * n2;
      for (int z = 0; z < n2; z++) dst[z] = creal(a[z]);
    }
    free(a); free(t);
  }
}

/* wave number with the Nyquist mode zeroed (src/cuda/SpectralOperators.cu:63-72) */
static inline double wnum(int64_t K, int n) {
  double w = (double)K;
  if ((double)K > n / 2.0) w -= n;
  if ((double)K == n / 2.0) w = 0;
  return w;
}

/* w_f = (i * w_d / N^3) * f  (multiply{X,Y,Z}WaveNumber, src/cuda/SpectralOperators.cu:54-127) */
static void mult_wave(orc_t *o, int d, const cplx *f, cplx *wf) {
  const int n0 = o->n[0], n1 = o->n[1], n2h = o->n2h;
  const double factor = 1.0 / ((double)o->n[0] * o->n[1] * o->n[2]);
#pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < o->nh; idx++) {
    int64_t x = idx / ((int64_t)n1 * n2h), y = (idx / n2h) % n1, z = idx % n2h;
    int64_t K = (d == 0) ? x : (d == 1) ? y : z;
    double w = wnum(K, o->n[d]);
    double re = creal(f[idx]), im = cimag(f[idx]);
    wf[idx] = (-factor * w * im) + I * (factor * w * re);
  }
  (void)n0;
}

/* SpectralOperators::computeGradient (src/grad/SpectralOperators.cpp:100-177) */
void orc_grad(orc_t *o, const double *x, double *gx, double *gy, double *gz) {
  orc_r2c(o, x, o->xhat);
  double *out[3] = {gx, gy, gz};
  for (int d = 0; d < 3; d++) {
    mult_wave(o, d, o->xhat, o->what);
    orc_c2r(o, o->what, out[d]);
  }
}

/* SpectralOperators::computeDivergence (src/grad/SpectralOperators.cpp:179-261):
   div = dx-term, then += dz-term, then += dy-term (cublas axpy order). */
void orc_div(orc_t *o, const double *dx, const double *dy, const double *dz, double *div) {
  double *d1 = o->w[11];
  orc_r2c(o, dx, o->xhat);
  mult_wave(o, 0, o->xhat, o->what);
  orc_c2r(o, o->what, div);
  orc_r2c(o, dz, o->xhat);
  mult_wave(o, 2, o->xhat, o->what);
  orc_c2r(o, o->what, d1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) div[i] += d1[i];
  orc_r2c(o, dy, o->xhat);
  mult_wave(o, 1, o->xhat, o->what);
  orc_c2r(o, o->what, d1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) div[i] += d1[i];
}

/* SpectralOperators::weierstrassSmoother (src/grad/SpectralOperators.cpp:295-381) with the
   8-image periodic Gaussian filter of computeWeierstrassFilter (src/cuda/SpectralOperators.cu:12-51). */
void orc_smooth(orc_t *o, const double *c, double *wc, double sigma) {
  if (sigma == 0) { if (wc != c) memcpy(wc, c, sizeof(double) * o->ng); return; }
  const int n0 = o->n[0], n1 = o->n[1], n2 = o->n[2];
  const double twopi = 2.0 * M_PI, hx = o->h[0], hy = o->h[1], hz = o->h[2];
  const double factor = 1.0 / ((double)n0 * n1 * n2);
  double *f = o->w[10];
  double sum = 0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (int64_t idx = 0; idx < o->ng; idx++) {
    int64_t i = idx / ((int64_t)n1 * n2), j = (idx / n2) % n1, k = idx % n2;
    double X = i * hx, Xp = X - twopi, Y = j * hy, Yp = Y - twopi, Z = k * hz, Zp = Z - twopi;
    double s2 = sigma * sigma * 2.0;
    double v = exp((-X * X - Y * Y - Z * Z) / s2) + exp((-Xp * Xp - Yp * Yp - Zp * Zp) / s2);
    v += exp((-Xp * Xp - Y * Y - Z * Z) / s2) + exp((-X * X - Yp * Yp - Z * Z) / s2);
    v += exp((-X * X - Y * Y - Zp * Zp) / s2) + exp((-Xp * Xp - Yp * Yp - Z * Z) / s2);
    v += exp((-Xp * Xp - Y * Y - Zp * Zp) / s2) + exp((-X * X - Yp * Yp - Zp * Zp) / s2);
    if (v != v) v = 0.;
    f[idx] = v;
    sum += v;
  }
  double nf = 1. / (sum * hx * hy * hz);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) f[i] *= nf;
  orc_r2c(o, f, o->xhat);
  orc_r2c(o, c, o->what);
  const double alp = factor * hx * hy * hz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->nh; i++) o->xhat[i] *= o->what[i] * alp;
  orc_c2r(o, o->xhat, wc);
}

/* ------------------------------------------------------------------------ */
/* Material properties and coefficients                                     */
/* ------------------------------------------------------------------------ */

/* MatProp::setValuesCustom (src/mat/MatProp.cpp:134-203); null inputs are zero fields. */
void orc_set_tissue(orc_t *o, const double *wm, const double *gm, const double *csf, const double *vt, const double *bg) {
  double fs = 0;
#pragma omp parallel for schedule(static) reduction(+ : fs)
  for (int64_t i = 0; i < o->ng; i++) {
    double w = wm ? wm[i] : 0, g = gm ? gm[i] : 0, c = csf ? csf[i] : 0, v = vt ? vt[i] : 0;
    (void)bg;
    w = (w <= 0.) ? 0. : w; g = (g <= 0.) ? 0. : g; v = (v <= 0.) ? 0. : v; c = (c <= 0.) ? 0. : c;
    o->wm[i] = w; o->gm[i] = g; o->csf[i] = c; o->vt[i] = v;
    o->bg[i] = -((g + w + v + c) - 1.0);
    o->filter[i] = ((w > 0.1 || g > 0.1) && v < 0.8) ? 1.0 : 0.0;
    fs += o->filter[i];
  }
  o->filter_sum = fs;
  o->k_filter_is_ng = 0;
}

/* MatProp::setValuesSinusoidal (src/mat/MatProp.cpp:34-68) */
void orc_set_tissue_sinusoidal(orc_t *o) {
  const int n0 = o->n[0], n1 = o->n[1], n2 = o->n[2];
  const double freq = 4;
  for (int64_t idx = 0; idx < o->ng; idx++) {
    int64_t X = idx / ((int64_t)n1 * n2), Y = (idx / n2) % n1, Z = idx % n2;
    o->wm[idx] = 0.5 + 0.5 * sin(freq * 2.0 * M_PI / n0 * X) * sin(freq * 2.0 * M_PI / n1 * Y) * sin(freq * 2.0 * M_PI / n2 * Z);
    o->gm[idx] = 1 - o->wm[idx];
    o->vt[idx] = 0; o->csf[idx] = 0; o->bg[idx] = 0; o->filter[idx] = 1;
  }
  o->filter_sum = (double)o->ng;
  o->k_filter_is_ng = 0;
}

/* DiffCoef::setValues (src/mat/DiffCoef.cpp:77-132): isotropic k, mask-averaged kbar */
void orc_set_k(orc_t *o, double k_scale, double k_gm_wm, double k_glm_wm) {
  o->k_scale = k_scale; o->k_gm_wm = k_gm_wm; o->k_glm_wm = k_glm_wm;
  double dk_gm = k_scale * k_gm_wm, dk_wm = k_scale, dk_glm = k_scale * k_glm_wm;
  dk_gm = (dk_gm <= 0) ? 0.0 : dk_gm;
  dk_glm = (dk_glm <= 0) ? 0.0 : dk_glm;
  double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < o->ng; i++) {
    double v = 0.0;
    v += dk_gm * o->gm[i];
    v += dk_wm * o->wm[i];
    v += dk_glm * o->csf[i];
    o->k[i] = v;
    s += v;
  }
  o->kavg = s * (1.0 / o->filter_sum);
}

/* DiffCoef::setValuesSinusoidal (src/mat/DiffCoef.cpp:134-189): kbar = sum(k)/ng */
void orc_set_k_sinusoidal(orc_t *o, double scale) {
  const int n0 = o->n[0], n1 = o->n[1], n2 = o->n[2];
  const double freq = 4;
  o->k_scale = scale;
  double s = 0;
  for (int64_t idx = 0; idx < o->ng; idx++) {
    int64_t X = idx / ((int64_t)n1 * n2), Y = (idx / n2) % n1, Z = idx % n2;
    o->k[idx] = scale * (0.5 + 0.5 * sin(freq * 2.0 * M_PI / n0 * X) * sin(freq * 2.0 * M_PI / n1 * Y) * sin(freq * 2.0 * M_PI / n2 * Z));
    s += o->k[idx];
  }
  o->kavg = s * (1.0 / (double)o->ng);
}

/* ReacCoef::setValues (src/mat/ReacCoef.cpp:13-38) */
void orc_set_rho(orc_t *o, double rho_scale, double r_gm_wm, double r_glm_wm) {
  o->rho_scale = rho_scale; o->r_gm_wm = r_gm_wm; o->r_glm_wm = r_glm_wm;
  double dr_gm = rho_scale * r_gm_wm, dr_wm = rho_scale, dr_glm = rho_scale * r_glm_wm;
  dr_gm = (dr_gm <= 0) ? 0.0 : dr_gm;
  dr_glm = (dr_glm <= 0) ? 0.0 : dr_glm;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) {
    double v = 0.0;
    v += dr_gm * o->gm[i];
    v += dr_wm * o->wm[i];
    v += dr_glm * o->csf[i];
    o->rho[i] = v;
  }
}

void orc_get_fields(orc_t *o, double *k, double *rho, double *filter, double *kavg) {
  if (k) memcpy(k, o->k, sizeof(double) * o->ng);
  if (rho) memcpy(rho, o->rho, sizeof(double) * o->ng);
  if (filter) memcpy(filter, o->filter, sizeof(double) * o->ng);
  if (kavg) *kavg = o->kavg;
}

/* ------------------------------------------------------------------------ */
/* Diffusion operator, preconditioner, CN + PETSc-KSPCG                      */
/* ------------------------------------------------------------------------ */

/* DiffCoef::applyD (src/mat/DiffCoef.cpp:249-271) with applyK (:206-247) for isotropic k */
static void apply_d_with(orc_t *o, const double *kf, const double *c, double *dc) {
  double *g0 = o->w[4], *g1 = o->w[5], *g2 = o->w[6];
  double *t1 = o->w[1], *t2 = o->w[2], *t3 = o->w[3];
  orc_grad(o, c, g0, g1, g2);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) {
    t1[i] = kf[i] * g0[i];
    t2[i] = kf[i] * g1[i];
    t3[i] = kf[i] * g2[i];
  }
  orc_div(o, t1, t2, t3, dc);
}

void orc_apply_d(orc_t *o, const double *c, double *dc) { apply_d_with(o, o->k, c, dc); }

/* DiffCoef::setSecondaryCoefficients (src/mat/DiffCoef.cpp:44-58), nk = 1 */
void orc_set_secondary(orc_t *o, double k1) {
  const double k2 = o->k_gm_wm * k1, k3 = o->k_glm_wm * k1;
  for (int64_t i = 0; i < o->ng; i++) o->ksec[i] = o->wm[i] * k1 + k2 * o->gm[i] + k3 * o->csf[i];
}

/* DiffusionSolver::precFactor (src/pde/DiffusionSolver.cpp:119-180) */
void orc_prec_factor(orc_t *o) {
  const int n1 = o->n[1], n2h = o->n2h;
  const double factor = 1.0 / ((double)o->n[0] * o->n[1] * o->n[2]);
  const double dt = o->ctx_dt, kk = o->kavg;
  o->prec_dt = dt;
#pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < o->nh; idx++) {
    int64_t x = idx / ((int64_t)n1 * n2h), y = (idx / n2h) % n1, z = idx % n2h;
    double wx = wnum(x, o->n[0]), wy = wnum(y, o->n[1]), wz = wnum(z, o->n[2]);
    double v = 1 + 0.25 * dt * (kk * wx * wx + kk * wy * wy + kk * wz * wz);
    o->prec[idx] = (v == 0) ? factor : factor / v;
  }
}

/* applyPC (src/pde/DiffusionSolver.cpp:182-215): y = C2R(prec .* R2C(x)) */
static void apply_pc(orc_t *o, const double *x, double *y) {
  orc_r2c(o, x, o->xhat);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->nh; i++) o->xhat[i] *= o->prec[i];
  orc_c2r(o, o->xhat, y);
}

/* operatorA (src/pde/DiffusionSolver.cpp:97-117): y = x - (ctx_dt/2) * D x */
static void operator_a(orc_t *o, const double *x, double *y) {
  double *t = o->w[0];
  orc_apply_d(o, x, t);
  const double alph = -1.0 / 2.0 * o->ctx_dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) y[i] = x[i] + alph * t[i];
}

static double vdot(orc_t *o, const double *a, const double *b) {
  double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < o->ng; i++) s += a[i] * b[i];
  return s;
}

enum { KSP_CONVERGED_RTOL = 2, KSP_CONVERGED_ATOL = 3, KSP_DIVERGED_ITS = -3, KSP_DIVERGED_DTOL = -4,
       KSP_DIVERGED_INDEFINITE_PC = -8, KSP_DIVERGED_NANORINF = -9, KSP_DIVERGED_INDEFINITE_MAT = -10 };

/* KSPConvergedDefault (PETSc 3.11, src/ksp/ksp/interface/iterativ.c), default preconditioned norm,
   nonzero initial guess: rnorm0 = ||B b||, ttol = max(rtol*rnorm0, atol). */
static int ksp_converged(orc_t *o, int n, double rnorm, double *rnorm0, double *ttol, const double *b) {
  if (n == 0) {
    double *z = o->w[9];
    apply_pc(o, b, z);
    double snorm = sqrt(vdot(o, z, z));
    if (snorm == 0) snorm = rnorm;
    *rnorm0 = snorm;
    *ttol = fmax(o->rtol * snorm, o->atol);
  }
  if (isnan(rnorm) || isinf(rnorm)) return KSP_DIVERGED_NANORINF;
  if (rnorm <= *ttol) return (rnorm < o->atol) ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL;
  if (rnorm >= o->dtol * (*rnorm0)) return KSP_DIVERGED_DTOL;
  return 0;
}

/* KSPSolve_CG (PETSc 3.11, src/ksp/ksp/impls/cg/cg.c) for the default KSP_NORM_PRECONDITIONED,
   nonzero initial guess (KSPSetInitialGuessNonzero, src/pde/DiffusionSolver.cpp:28). */
static int ksp_cg(orc_t *o, const double *b, double *x, int *reason_out) {
  double *r = o->w[7], *z = o->w[8], *p = o->cg_p, *wv = o->cg_w;
  double dp, beta = 0, betaold = 0, dpi = 0, dpiold, a, bb, rnorm0 = 0, ttol = 0;
  int its = 0, reason = 0, i;
  operator_a(o, x, r);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < o->ng; k++) r[k] = b[k] - r[k];
  apply_pc(o, r, z);
  dp = sqrt(vdot(o, z, z));
  reason = ksp_converged(o, 0, dp, &rnorm0, &ttol, b);
  if (reason) { *reason_out = reason; return 0; }
  beta = vdot(o, z, r);
  i = 0;
  do {
    its = i + 1;
    if (beta == 0.0) { reason = KSP_CONVERGED_ATOL; break; }
    else if (i > 0 && beta * betaold < 0.0) { reason = KSP_DIVERGED_INDEFINITE_PC; break; }
    if (!i) {
      memcpy(p, z, sizeof(double) * o->ng);
    } else {
      bb = beta / betaold;
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < o->ng; k++) p[k] = z[k] + bb * p[k];
    }
    dpiold = dpi;
    operator_a(o, p, wv);
    dpi = vdot(o, p, wv);
    betaold = beta;
    if (isnan(dpi) || isinf(dpi)) { reason = KSP_DIVERGED_NANORINF; break; }
    if ((dpi == 0.0) || ((i > 0) && (dpi * dpiold <= 0.0))) { reason = KSP_DIVERGED_INDEFINITE_MAT; break; }
    a = beta / dpi;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < o->ng; k++) { x[k] += a * p[k]; r[k] -= a * wv[k]; }
    apply_pc(o, r, z);
    dp = sqrt(vdot(o, z, z));
    reason = ksp_converged(o, i + 1, dp, &rnorm0, &ttol, b);
    if (reason) break;
    beta = vdot(o, z, r);
    if (isnan(beta) || isinf(beta)) { reason = KSP_DIVERGED_NANORINF; break; }
    i++;
  } while (i < o->maxit);
  if (i >= o->maxit) reason = KSP_DIVERGED_ITS;
  *reason_out = reason;
  return its;
}

/* DiffusionSolver::solve (src/pde/DiffusionSolver.cpp:217-250): CN half step, in place */
int orc_diffusion_solve(orc_t *o, double *c, double tau) {
  o->ctx_dt = tau;
  o->ksp_total_calls++;
  if (o->k_scale == 0) { o->ksp_itr = 0; o->ksp_reason = 0; return 0; }
  double *rhs = o->rhs;
  double *t = o->w[0];
  const double alph = 1.0 / 2.0 * o->ctx_dt;
  orc_apply_d(o, c, t);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o->ng; i++) rhs[i] = c[i] + alph * t[i];
  int reason = 0;
  o->ksp_itr = ksp_cg(o, rhs, c, &reason);
  o->ksp_reason = reason;
  return o->ksp_itr;
}

int orc_last_reason(orc_t *o) { return o->ksp_reason; }

void orc_set_solver(orc_t *o, double rtol, int maxit) { o->rtol = rtol; o->maxit = maxit; }

void orc_set_ctx_dt(orc_t *o, double dt) { o->ctx_dt = dt; }
double orc_get_ctx_dt(orc_t *o) { return o->ctx_dt; }
double orc_get_prec_dt(orc_t *o) { return o->prec_dt; }

/* ------------------------------------------------------------------------ */
/* Reaction (src/pde/PdeOperators.cpp:140-190; src/cuda/PdeOperators.cu:4-25) */
/* ------------------------------------------------------------------------ */
void orc_reaction(orc_t *o, double *ct, const double *clin, double dt) {
  if (!clin) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < o->ng; i++) {
      double factor = exp(o->rho[i] * dt);
      double alph = ct[i] / (1.0 - ct[i]);
      if (isinf(alph)) ct[i] = 1.0;
      else ct[i] = alph * factor / (alph * factor + 1.0);
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < o->ng; i++) {
      double factor = exp(o->rho[i] * dt);
      double alph = (clin[i] * factor + 1.0 - clin[i]);
      ct[i] = ct[i] * factor / (alph * alph);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Time history and the forward / adjoint solves                            */
/* ------------------------------------------------------------------------ */

/* TumorParameters dt_/nt_ and PdeOperatorsRD::resizeTimeHistory (src/pde/PdeOperators.cpp:56-103).
   Setting dt also resets the diffusion solver's context dt, as constructing DiffusionSolver does
   (src/pde/DiffusionSolver.cpp:11). */
void orc_set_time(orc_t *o, double dt, int nt, int adjoint_store) {
  free_hist(o);
  o->dt = dt; o->nt = nt; o->adjoint_store = adjoint_store;
  o->ctx_dt = dt;
  o->hist_nt = nt;
  o->c_hist = (double **)calloc(nt + 1, sizeof(double *));
  o->p_hist = (double **)calloc(nt + 1, sizeof(double *));
  for (int i = 0; i <= nt; i++) { o->c_hist[i] = dalloc(o->ng); o->p_hist[i] = dalloc(o->ng); }
  if (adjoint_store) {
    o->c_half = (double **)calloc(nt, sizeof(double *));
    for (int i = 0; i < nt; i++) o->c_half[i] = dalloc(o->ng);
  }
}

void orc_set_c0(orc_t *o, const double *c0) { memcpy(o->c_0, c0, sizeof(double) * o->ng); }
void orc_set_pt(orc_t *o, const double *pt) { memcpy(o->p_t, pt, sizeof(double) * o->ng); }
void orc_get_c0(orc_t *o, double *c) { memcpy(c, o->c_0, sizeof(double) * o->ng); }
void orc_get_ct(orc_t *o, double *c) { memcpy(c, o->c_t, sizeof(double) * o->ng); }
void orc_get_p0(orc_t *o, double *p) { memcpy(p, o->p_0, sizeof(double) * o->ng); }
void orc_get_c_hist(orc_t *o, int i, double *c) { memcpy(c, o->c_hist[i], sizeof(double) * o->ng); }
void orc_get_p_hist(orc_t *o, int i, double *p) { memcpy(p, o->p_hist[i], sizeof(double) * o->ng); }
void orc_get_c_half(orc_t *o, int i, double *c) { memcpy(c, o->c_half[i], sizeof(double) * o->ng); }
void orc_get_ksp(orc_t *o, int *state, int *adj, int *last) {
  if (state) *state = o->diff_ksp_itr_state;
  if (adj) *adj = o->diff_ksp_itr_adj;
  if (last) *last = o->ksp_itr;
}

/* PdeOperatorsRD::solveIncremental (src/pde/PdeOperators.cpp:192-233) with its dt argument equal
   to dt/2: c~ += dt/4 * D_sec(1.5 c_i + 0.5 c_{i+1}) (mode 1) or D_sec(0.5 c_i + 1.5 c_{i+1})
   (mode 2), D_sec = div(k_sec grad) (applyDWithSecondaryCoeffs, src/mat/DiffCoef.cpp:273-300) */
static void solve_incremental(orc_t *o, int iter, int mode) {
  double *temp = o->w[9];
  const double *a = o->c_hist[iter], *b = o->c_hist[iter + 1];
  for (int64_t i = 0; i < o->ng; i++) temp[i] = 0.5 * (a[i] + b[i]) + (mode == 1 ? a[i] : b[i]);
  apply_d_with(o, o->ksec, temp, temp);
  const double s = (o->dt / 2.0) / 2.0;
  for (int64_t i = 0; i < o->ng; i++) o->c_t[i] += s * temp[i];
}

/* PdeOperatorsRD::solveState (src/pde/PdeOperators.cpp:235-316), order 2 (Strang),
   linearized 0 (state), 1 (tangent linear around c_[i]) or 2 (tangent linear with the secondary
   diffusion sources of the Hessian with diffusivity inversion). */
int orc_solve_state(orc_t *o, int linearized) {
  const double dt = o->dt;
  memcpy(o->c_t, o->c_0, sizeof(double) * o->ng);
  if (linearized == 0) memcpy(o->c_hist[0], o->c_t, sizeof(double) * o->ng);
  o->diff_ksp_itr_state = 0;
  for (int i = 0; i < o->nt; i++) {
    if (linearized == 2) solve_incremental(o, i, 1);
    orc_diffusion_solve(o, o->c_t, dt / 2.0);
    o->diff_ksp_itr_state += o->ksp_itr;
    if (linearized == 0 && o->adjoint_store) memcpy(o->c_half[i], o->c_t, sizeof(double) * o->ng);
    orc_reaction(o, o->c_t, linearized ? o->c_hist[i] : NULL, dt);
    orc_diffusion_solve(o, o->c_t, dt / 2.0);
    o->diff_ksp_itr_state += o->ksp_itr;
    if (linearized == 2) solve_incremental(o, i, 2);
    if (linearized == 0) memcpy(o->c_hist[i + 1], o->c_t, sizeof(double) * o->ng);
    if (orc_verbose) fprintf(stderr, "oracle: state step %d of %d\n", i + 1, o->nt);
  }
  return o->diff_ksp_itr_state;
}

/* PdeOperatorsRD::reactionAdjoint (src/pde/PdeOperators.cpp:318-370) */
static void reaction_adjoint(orc_t *o, int iter) {
  double *temp = o->adj_temp; /* the role of Tumor::work_[11] */
  memcpy(temp, o->c_hist[iter], sizeof(double) * o->ng);
  if (o->adjoint_store) {
    memcpy(temp, o->c_half[iter], sizeof(double) * o->ng);
  } else {
    orc_diffusion_solve(o, temp, o->dt / 2.0);
    o->diff_ksp_itr_adj += o->ksp_itr;
  }
  orc_reaction(o, o->p_0, temp, o->dt);
}

/* PdeOperatorsRD::solveAdjoint (src/pde/PdeOperators.cpp:372-420), order 2 */
int orc_solve_adjoint(orc_t *o, int linearized) {
  const double dt = o->dt;
  const int nt = o->nt;
  memcpy(o->p_0, o->p_t, sizeof(double) * o->ng);
  if (linearized == 1) memcpy(o->p_hist[nt], o->p_0, sizeof(double) * o->ng);
  o->diff_ksp_itr_adj = 0;
  for (int i = 0; i < nt; i++) {
    orc_diffusion_solve(o, o->p_0, dt / 2.0);
    o->diff_ksp_itr_adj += o->ksp_itr;
    reaction_adjoint(o, nt - i - 1);
    orc_diffusion_solve(o, o->p_0, dt / 2.0);
    o->diff_ksp_itr_adj += o->ksp_itr;
    memcpy(o->p_hist[nt - i - 1], o->p_0, sizeof(double) * o->ng);
    if (orc_verbose) fprintf(stderr, "oracle: adjoint step %d of %d\n", i + 1, nt);
  }
  return o->diff_ksp_itr_adj;
}

/* ------------------------------------------------------------------------ */
/* Phi: Gaussian basis (src/mat/Phi.cpp:76-439, src/cuda/Phi.cu:11-83)       */
/* ------------------------------------------------------------------------ */

/* Phi::phiMesh (src/mat/Phi.cpp:148-234); returns the number of centers written */
int orc_phi_mesh(int np, const double *cm, double sigma, double spacing_factor, double *center) {
  int h = (int)round(pow(np, 1.0 / 3.0));
  double space = spacing_factor * sigma;
  if (np == 1) { center[0] = cm[0]; center[1] = cm[1]; center[2] = cm[2]; return 1; }
  int ptr = 0;
  if (np % 2 == 1) {
    for (int k = -(h - 1) / 2; k <= (h - 1) / 2; k++)
      for (int j = -(h - 1) / 2; j <= (h - 1) / 2; j++)
        for (int i = -(h - 1) / 2; i <= (h - 1) / 2; i++) {
          center[ptr + 0] = i * space + cm[0];
          center[ptr + 1] = j * space + cm[1];
          center[ptr + 2] = k * space + cm[2];
          ptr += 3;
        }
  } else {
    for (int k = -(h) / 2; k <= (h) / 2; k++)
      for (int j = -(h) / 2; j <= (h) / 2; j++)
        for (int i = -(h) / 2; i <= (h) / 2; i++) {
          double m0 = ((i < 0) ? -1 : ((i > 0) ? 1 : 0)) * 0.5;
          double m1 = ((j < 0) ? -1 : ((j > 0) ? 1 : 0)) * 0.5;
          double m2 = ((k < 0) ? -1 : ((k > 0) ? 1 : 0)) * 0.5;
          if ((i != 0) && (j != 0) && (k != 0)) {
            center[ptr + 0] = (i - m0) * space + cm[0];
            center[ptr + 1] = (j - m1) * space + cm[1];
            center[ptr + 2] = (k - m2) * space + cm[2];
            ptr += 3;
          }
        }
  }
  return ptr / 3;
}

/* Phi::setGaussians(Vec data) on one process (src/mat/Phi.cpp:615-940).  Candidates sit on the
   lattice that the loops of :686-720 walk: per axis from 0 to the largest index that is an exact
   multiple of space = 2 sigma / h0 (fmod test), in steps x = (int)(x + space).  A candidate counts
   the voxels with data > threshold (compared in single precision, as the GPU build) at integer
   offsets d with sqrt(|d|^2) <= r = sigma / h0, clipped to the grid without wrapping
   (checkTumorExistence / checkTumorExistenceOutOfProc, :457-522); it is kept when its count
   exceeds gvf times the full ball's count (:870-888).  Kept centers (X hx, Y hy, Z hz) in index
   order, at most max_np written; returns the number kept, or -1 when space < 1 (the reference's
   loop would not advance). */
int orc_data_support(const float *data, const int *n, double threshold, double sigma, double gvf, int max_np,
                     double *centers, int *counts_out) {
  const double twopi = 2.0 * M_PI;
  const double hx = twopi / n[0], hy = twopi / n[1], hz = twopi / n[2];
  const double r = sigma / hx, space = 2.0 * sigma / hx;
  if (!(space >= 1.0)) return -1;
  const float thr = (float)threshold;
  int interior = 0;
  for (int i = (int)(-r); i <= r; i++)
    for (int j = (int)(-r); j <= r; j++)
      for (int k = (int)(-r); k <= r; k++)
        if (sqrt((double)(i * i + j * j + k * k)) <= r) interior++;
  int end[3];
  for (int a = 0; a < 3; a++) {
    end[a] = 0;
    for (int x = n[a] - 1; x >= 0; x--)
      if (fmod((double)x, space) == 0) { end[a] = x; break; }
  }
  int np = 0, idx = 0;
  for (int x = 0; x <= end[0]; x = (int)(x + space))
    for (int y = 0; y <= end[1]; y = (int)(y + space))
      for (int z = 0; z <= end[2]; z = (int)(z + space), idx++) {
        int num = 0;
        for (int i = (int)(x - r); i <= x + r; i++)
          for (int j = (int)(y - r); j <= y + r; j++)
            for (int k = (int)(z - r); k <= z + r; k++) {
              if (i < 0 || j < 0 || k < 0 || i >= n[0] || j >= n[1] || k >= n[2]) continue;
              const double dist = sqrt((double)((i - x) * (i - x) + (j - y) * (j - y) + (k - z) * (k - z)));
              if (dist <= r && data[((int64_t)i * n[1] + j) * n[2] + k] > thr) num++;
            }
        if (counts_out) counts_out[idx] = num;
        if (num > gvf * interior) {
          if (np < max_np) {
            centers[3 * np + 0] = x * hx;
            centers[3 * np + 1] = y * hy;
            centers[3 * np + 2] = z * hz;
          }
          np++;
        }
      }
  return np;
}

/* Phi::initialize (src/mat/Phi.cpp:264-311) + mask + smoother + Phi::truncate (:238-260) */
static void phi_build_one(orc_t *o, const double *ctr, double *out) {
  const int n1 = o->n[1], n2 = o->n[2];
  const double hx = o->h[0], hy = o->h[1], hz = o->h[2];
  const double R = sqrt(2.) * o->phi_sigma;
  const double xc = ctr[0], yc = ctr[1], zc = ctr[2];
#pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < o->ng; idx++) {
    int64_t X = idx / ((int64_t)n1 * n2), Y = (idx / n2) % n1, Z = idx % n2;
    double dx = hx * X - xc, dy = hy * Y - yc, dz = hz * Z - zc;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double ratio = r / R;
    out[idx] = exp(-ratio * ratio) * o->filter[idx];
  }
  double sigma_smooth = o->smoothing_factor * 2.0 * M_PI / o->n[0];
  orc_smooth(o, out, out, sigma_smooth);
#pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < o->ng; idx++) {
    int64_t X = idx / ((int64_t)n1 * n2), Y = (idx / n2) % n1, Z = idx % n2;
    double dx = hx * X - xc, dy = hy * Y - yc, dz = hz * Z - zc;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    out[idx] = (r / o->phi_sigma <= 5) ? out[idx] : 0.0;
  }
}

static double vmax(orc_t *o, const double *v) {
  double m = -INFINITY;
  for (int64_t i = 0; i < o->ng; i++) if (v[i] > m) m = v[i];
  return m;
}

/* Phi::setGaussians + Phi::setValues (src/mat/Phi.cpp:76-144).  With phi_store the bases are
   stored normalized by the global max; without it the raw bases are cached here only to save
   oracle time -- apply/applyTranspose below normalize exactly as the on-the-fly path does. */
void orc_set_phi(orc_t *o, int np, const double *centers, double sigma, double smoothing_factor, int phi_store) {
  free_phi(o);
  o->np = np;
  o->phi_sigma = sigma;
  o->smoothing_factor = smoothing_factor;
  o->phi_store = phi_store;
  o->centers = (double *)malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
  memcpy(o->centers, centers, sizeof(double) * 3 * np);
  o->phi = (double **)calloc(np > 0 ? np : 1, sizeof(double *));
  double phi_max = 0;
  for (int i = 0; i < np; i++) {
    o->phi[i] = dalloc(o->ng);
    phi_build_one(o, &centers[3 * i], o->phi[i]);
    double m = vmax(o, o->phi[i]);
    if (m > phi_max) phi_max = m;
  }
  o->phi_max_all = phi_max;
  if (phi_store) {
    for (int i = 0; i < np; i++)
      for (int64_t k = 0; k < o->ng; k++) o->phi[i][k] *= (1.0 / phi_max);
  }
}

void orc_get_phi(orc_t *o, int i, double *out) { memcpy(out, o->phi[i], sizeof(double) * o->ng); }

/* Phi::apply (src/mat/Phi.cpp:313-377) */
void orc_phi_apply(orc_t *o, const double *p, double *out) {
  memset(out, 0, sizeof(double) * o->ng);
  if (o->phi_store) {
    for (int i = 0; i < o->np; i++)
      for (int64_t k = 0; k < o->ng; k++) out[k] += p[i] * o->phi[i][k];
    return;
  }
  double phi_max = 0;
  int nnz = 0;
  for (int i = 0; i < o->np; i++) {
    if (p[i] != 0) {
      nnz = 1;
      double m = vmax(o, o->phi[i]);
      if (m > phi_max) phi_max = m;
      for (int64_t k = 0; k < o->ng; k++) out[k] += p[i] * o->phi[i][k];
    }
  }
  if (!nnz) phi_max = 1;
  for (int64_t k = 0; k < o->ng; k++) out[k] *= (1.0 / phi_max);
}

/* Phi::applyTranspose (src/mat/Phi.cpp:379-439) */
void orc_phi_applyT(orc_t *o, const double *in, double *pout) {
  for (int i = 0; i < o->np; i++) pout[i] = vdot(o, o->phi[i], in);
  if (!o->phi_store) {
    for (int i = 0; i < o->np; i++) pout[i] *= (1.0 / o->phi_max_all);
  }
}

/* ------------------------------------------------------------------------ */
/* Observation operator and data (src/mat/Obs.cpp:26-132)                    */
/* ------------------------------------------------------------------------ */
void orc_set_data(orc_t *o, const double *d1, const double *obs_mask, double thr) {
  memcpy(o->d1, d1, sizeof(double) * o->ng);
  for (int64_t i = 0; i < o->ng; i++) o->obs[i] = obs_mask ? obs_mask[i] : (double)(d1[i] > thr);
}

/* ------------------------------------------------------------------------ */
/* Derivative operators (src/grad/DerivativeOperatorsRD.cpp:7-226,           */
/* src/grad/DerivativeOperators.cpp:189-367)                                  */
/* mode 0: invert p and kappa (flag_reaction_inv_ = false): c0 = Phi p, g_p, g_kappa
   mode 1: invert kappa and rho (flag_reaction_inv_ = true): c0 fixed, g_kappa, g_rho
   mode 2: all of p, kappa, rho in one pass (the same formulas; the build's own API)
   x = [p (np) | kappa | rho]                                                    */
/* ------------------------------------------------------------------------ */
void orc_set_opt(orc_t *o, double beta, double k_scale, double rho_scale) {
  o->beta = beta; o->opt_k_scale = k_scale; o->opt_rho_scale = rho_scale;
}

/* updateReactionAndDiffusion (src/grad/DerivativeOperators.cpp:323-367) with nk = nr = 1 */
static void update_reaction_and_diffusion(orc_t *o, const double *x, int mode) {
  double k1 = x[o->np] * o->opt_k_scale;
  orc_set_k(o, k1, o->k_gm_wm, o->k_glm_wm);
  orc_prec_factor(o);
  if (mode != 0) {
    double r1 = x[o->np + 1] * o->opt_rho_scale;
    orc_set_rho(o, r1, o->r_gm_wm, o->r_glm_wm);
  }
}

static double misfit(orc_t *o) {
  double *t = o->w[0];
  for (int64_t i = 0; i < o->ng; i++) t[i] = o->obs[i] * o->c_t[i] - o->d1[i];
  return vdot(o, t, t);
}

double orc_objective(orc_t *o, const double *x, int mode) {
  update_reaction_and_diffusion(o, x, mode);
  if (mode != 1) orc_phi_apply(o, x, o->c_0);
  orc_solve_state(o, 0);
  double m1 = misfit(o);
  double reg = vdot(o, o->c_0, o->c_0) * 0.5 * o->beta * o->lebesgue;
  return o->lebesgue * 0.5 * m1 + reg;
}

/* gradDiffusion (:189-256) and gradReaction (:258-321) for nk = nr = 1 */
static void grad_kappa_rho(orc_t *o, double *gk, double *gr) {
  double *g1 = o->w[1], *g2 = o->w[2], *g3 = o->w[3], *g4 = o->w[4], *g5 = o->w[5], *g6 = o->w[6];
  double *Q = o->w[7], *P = o->w[8];
  memset(Q, 0, sizeof(double) * o->ng);
  memset(P, 0, sizeof(double) * o->ng);
  for (int i = 0; i < o->nt + 1; i++) {
    double iw = (i == 0 || i == o->nt) ? 0.5 : 1.0;
    orc_grad(o, o->c_hist[i], g1, g2, g3);
    orc_grad(o, o->p_hist[i], g4, g5, g6);
    const double *c = o->c_hist[i], *a = o->p_hist[i];
    const double s = o->dt * iw;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < o->ng; k++) {
      double v = g1[k] * g4[k];
      v += g2[k] * g5[k];
      v += g3[k] * g6[k];
      Q[k] += s * v;
      double r = c[k] * c[k] - c[k];
      P[k] += s * (a[k] * r);
    }
  }
  *gk = vdot(o, o->wm, Q) * o->lebesgue + vdot(o, o->gm, Q) * o->lebesgue * o->k_gm_wm;
  *gr = vdot(o, o->wm, P) * o->lebesgue + vdot(o, o->gm, P) * o->lebesgue * o->r_gm_wm;
}

static void grad_kappa_rho(orc_t *o, double *gk, double *gr);

/* Gauss-Newton Hessian matvec: DerivativeOperatorsRD::evaluateHessian with L2 regularization
   (src/grad/DerivativeOperatorsRD.cpp:229-438), nk = 1.  x~ and y are [p (np) | kappa | rho].
   Needs the state history of a previous evaluation at the linearization point.
   Without diffusivity inversion (:420-436): y_p = beta h^3 Phi^T Phi p~ - h^3 Phi^T a~0 with
     c~0 = Phi p~, solveState(1), a~(T) = -O^T O c~(T), solveAdjoint(2).
   With diffusivity inversion (:244-418), y = [Hpp p~ + Hpk k~ ; Hkp p~ + Hkk k~]: the kappa row
   is the gradDiffusion quadrature of grad c_i . grad a~_i after each incremental adjoint, whose
   i = nt term reads the adjoint history entry p_[nt] that solveAdjoint(2) leaves from the last
   gradient; the k~ part runs solveState(2) from c~0 = 0 with secondary coefficients k~ m.
   The rho entry of y is 0. */
void orc_hessian(orc_t *o, const double *xt, double *y, int diffusivity) {
  const int np = o->np;
  double *ptemp = (double *)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
  for (int i = 0; i < np + 2; i++) y[i] = 0;
  orc_phi_apply(o, xt, o->c_0);
  orc_solve_state(o, 1);
  for (int64_t i = 0; i < o->ng; i++) o->p_t[i] = -(o->obs[i] * (o->obs[i] * o->c_t[i]));
  orc_solve_adjoint(o, 2);
  orc_phi_applyT(o, o->p_0, ptemp);
  for (int i = 0; i < np; i++) ptemp[i] *= o->lebesgue;
  orc_phi_applyT(o, o->c_0, y);
  for (int i = 0; i < np; i++) y[i] = y[i] * (o->beta * o->lebesgue) - ptemp[i];
  if (diffusivity) {
    double gk, gr;
    grad_kappa_rho(o, &gk, &gr);
    y[np] = gk;
    orc_set_secondary(o, xt[np]);
    memset(o->c_0, 0, sizeof(double) * o->ng);
    orc_solve_state(o, 2);
    for (int64_t i = 0; i < o->ng; i++) o->p_t[i] = -(o->obs[i] * (o->obs[i] * o->c_t[i]));
    orc_solve_adjoint(o, 2);
    orc_phi_applyT(o, o->p_0, ptemp);
    for (int i = 0; i < np; i++) y[i] += -o->lebesgue * ptemp[i];
    grad_kappa_rho(o, &gk, &gr);
    y[np] += gk;
  }
  free(ptemp);
}

void orc_hessian_p(orc_t *o, const double *ptilde, double *y) {
  const int np = o->np;
  double *xt = (double *)calloc((size_t)np + 2, sizeof(double));
  double *yy = (double *)calloc((size_t)np + 2, sizeof(double));
  memcpy(xt, ptilde, sizeof(double) * np);
  orc_hessian(o, xt, yy, 0);
  memcpy(y, yy, sizeof(double) * np);
  free(xt);
  free(yy);
}

/* evaluateObjectiveAndGradient (src/grad/DerivativeOperatorsRD.cpp:130-226) */
void orc_objective_and_gradient(orc_t *o, const double *x, int mode, double *J, double *dJ) {
  const int np = o->np;
  for (int i = 0; i < np + 2; i++) dJ[i] = 0;
  update_reaction_and_diffusion(o, x, mode);
  if (mode != 1) orc_phi_apply(o, x, o->c_0);
  orc_solve_state(o, 0);
  double *t = o->w[0];
  for (int64_t i = 0; i < o->ng; i++) t[i] = o->obs[i] * o->c_t[i] - o->d1[i];
  double m1 = vdot(o, t, t);
  for (int64_t i = 0; i < o->ng; i++) o->p_t[i] = -(o->obs[i] * t[i]);
  orc_solve_adjoint(o, 1);
  if (mode != 1) {
    for (int64_t i = 0; i < o->ng; i++) o->p_0[i] += -o->beta * o->c_0[i];
    orc_phi_applyT(o, o->p_0, dJ);
    for (int i = 0; i < np; i++) dJ[i] *= -o->lebesgue;
  }
  double reg = vdot(o, o->c_0, o->c_0) * 0.5 * o->beta * o->lebesgue;
  *J = o->lebesgue * 0.5 * m1 + reg;
  double gk, gr;
  grad_kappa_rho(o, &gk, &gr);
  dJ[np] = gk;
  dJ[np + 1] = (mode == 0) ? 0.0 : gr;
}

/* ------------------------------------------------------------------------ */
/* Semi-Lagrangian advection (src/pde/AdvectionSolver.cpp:279-704) with the */
/* tensor-product cubic Lagrange interpolation of the GPU path              */
/* (cubicTex3DFastLagrange, src/cuda/InterpolationTexture.cu:270-366, on a  */
/* periodic texture: cudaAddressModeWrap).  Query points are kept as        */
/* offsets from their grid point, in grid units.                            */
/* ------------------------------------------------------------------------ */

/* lagrange_weights (3rdparty/interp_poly_kernel.cu:56-69): the cubic Lagrange basis on the nodes
   -1, 0, 1, 2 at fraction t in [0, 1) */
static void lag_weights(double t, double w[4]) {
  w[0] = -t * (t - 1.0) * (t - 2.0) / 6.0;
  w[1] = (t + 1.0) * (t - 1.0) * (t - 2.0) / 2.0;
  w[2] = -(t + 1.0) * t * (t - 2.0) / 2.0;
  w[3] = (t + 1.0) * t * (t - 1.0) / 6.0;
}

static int64_t wrap_index(int64_t v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

/* f at grid coordinate (i + dx, j + dy, k + dz), periodic */
static double interp_at(const orc_t *o, const double *f, int64_t i, int64_t j, int64_t k, double dx, double dy,
                        double dz) {
  const int n0 = o->n[0], n1 = o->n[1], n2 = o->n[2];
  const double d[3] = {dx, dy, dz};
  const int64_t b[3] = {i, j, k};
  int64_t ix[3][4];
  double w[3][4];
  const int nn[3] = {n0, n1, n2};
  for (int a = 0; a < 3; a++) {
    const double fl = floor(d[a]);
    lag_weights(d[a] - fl, w[a]);
    for (int t = 0; t < 4; t++) ix[a][t] = wrap_index(b[a] + (int64_t)fl - 1 + t, nn[a]);
  }
  double s = 0.0;
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) {
      const double *row = f + (ix[0][p] * n1 + ix[1][q]) * n2;
      double r = 0.0;
      for (int t = 0; t < 4; t++) r += w[2][t] * row[ix[2][t]];
      s += w[0][p] * w[1][q] * r;
    }
  return s;
}

/* SemiLagrangianSolver::interpolate (src/pde/AdvectionSolver.cpp:364-417): out(x) = f(x + q(x)) */
void orc_sl_interp(orc_t *o, const double *f, const double *qx, const double *qy, const double *qz, double *out) {
  const int n1 = o->n[1], n2 = o->n[2];
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < o->ng; p++) {
    const int64_t i = p / ((int64_t)n1 * n2), j = (p / n2) % n1, k = p % n2;
    out[p] = interp_at(o, f, i, j, k, qx[p], qy[p], qz[p]);
  }
}

/* computeTrajectories (src/pde/AdvectionSolver.cpp:470-590, single-GPU branch): Euler points
   X* = x - dt v / h, v* = v(X*), departure points X = x - dt (v + v*) / (2 h); h = 2 pi / n per axis.
   Returns X - x in grid units. */
void orc_sl_trajectories(orc_t *o, const double *vx, const double *vy, const double *vz, double dt, double *qx,
                         double *qy, double *qz) {
  const int n1 = o->n[1], n2 = o->n[2];
  const double s0 = dt * o->n[0] / (2.0 * M_PI), s1 = dt * o->n[1] / (2.0 * M_PI), s2 = dt * o->n[2] / (2.0 * M_PI);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < o->ng; p++) {
    const int64_t i = p / ((int64_t)n1 * n2), j = (p / n2) % n1, k = p % n2;
    const double ex = -s0 * vx[p], ey = -s1 * vy[p], ez = -s2 * vz[p];
    const double wx = interp_at(o, vx, i, j, k, ex, ey, ez);
    const double wy = interp_at(o, vy, i, j, k, ex, ey, ez);
    const double wz = interp_at(o, vz, i, j, k, ex, ey, ez);
    qx[p] = -0.5 * s0 * (vx[p] + wx);
    qy[p] = -0.5 * s1 * (vy[p] + wy);
    qz[p] = -0.5 * s2 * (vz[p] + wz);
  }
}

/* SemiLagrangianSolver::solve (src/pde/AdvectionSolver.cpp:592-660) on the departure offsets q.
   mode 2 (pure advection): c <- c(X).  mode 1 (mass conservation, d_t c + v.grad c = -c div v):
   ci = c(X), di = (div v)(X), c* = ci - dt ci di, c <- ci + dt/2 (-ci di - c* div v). */
void orc_sl_solve(orc_t *o, double *c, const double *vx, const double *vy, const double *vz, const double *qx,
                  const double *qy, const double *qz, double dt, int mode) {
  double *ci = (double *)malloc(sizeof(double) * o->ng);
  orc_sl_interp(o, c, qx, qy, qz, ci);
  if (mode == 2) {
    memcpy(c, ci, sizeof(double) * o->ng);
  } else {
    double *dv = (double *)malloc(sizeof(double) * o->ng);
    double *di = (double *)malloc(sizeof(double) * o->ng);
    orc_div(o, vx, vy, vz, dv);
    orc_sl_interp(o, dv, qx, qy, qz, di);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < o->ng; p++) {
      const double src = -ci[p] * di[p];
      const double cstar = ci[p] + dt * src;
      c[p] = ci[p] + 0.5 * dt * (src - cstar * dv[p]);
    }
    free(dv);
    free(di);
  }
  free(ci);
}

/* ------------------------------------------------------------------------ */
/* Linear elasticity of the mass-effect model (VariableLinearElasticity-    */
/* Solver, src/pde/ElasticitySolver.cpp:6-533): material properties, the   */
/* variable-coefficient Navier-Lame operator, the constant-coefficient      */
/* spectral preconditioner and PETSc's KSPCG (rtol 1e-3, maxit 100).        */
/* Vectors of 3 ng entries hold the x, y and z components one after the     */
/* other.                                                                   */
/* ------------------------------------------------------------------------ */

/* CtxElasticity::computeMu / computeLam (include/pde/ElasticitySolver.h:39-41) */
static double el_mu(double E, double nu) { return E / (2 * (1 + nu)); }
static double el_lam(double E, double nu) { return nu * E / ((1 + nu) * (1 - 2 * nu)); }

/* computeMaterialProperties (src/pde/ElasticitySolver.cpp:376-455).  prm = E_healthy, nu_healthy,
   E_tumor, nu_tumor, E_bg, nu_bg, E_csf, nu_csf, screen_low, screen_high; avg = mean mu, lam,
   screen over the grid. */
void orc_el_material(orc_t *o, const double *c, const double *wm, const double *gm, const double *csf,
                     const double *vt, const double *bg, const double *prm, double *mu, double *lam,
                     double *screen, double *avg) {
  const double mu_h = el_mu(prm[0], prm[1]), mu_t = el_mu(prm[2], prm[3]), mu_b = el_mu(prm[4], prm[5]),
               mu_c = el_mu(prm[6], prm[7]);
  const double la_h = el_lam(prm[0], prm[1]), la_t = el_lam(prm[2], prm[3]), la_b = el_lam(prm[4], prm[5]),
               la_c = el_lam(prm[6], prm[7]);
  const double c_threshold = 0.005;
  double sm = 0, sl = 0, ss = 0;
  for (int64_t i = 0; i < o->ng; i++) {
    double m = mu_b * bg[i] + mu_h * wm[i] + mu_h * gm[i] + mu_c * vt[i] + mu_c * csf[i];
    double l = la_b * bg[i] + la_h * wm[i] + la_h * gm[i] + la_c * vt[i] + la_c * csf[i];
    m = (m > 0) ? m : 0;
    l = (l > 0) ? l : 0;
    m += (c[i] > 0) ? mu_t * c[i] : 0;
    l += (c[i] > 0) ? la_t * c[i] : 0;
    double s = (c[i] >= c_threshold) ? prm[8] : prm[9];
    if (bg[i] > 0.95) s = 1E6;
    mu[i] = m;
    lam[i] = l;
    screen[i] = s;
    sm += m;
    sl += l;
    ss += s;
  }
  avg[0] = sm / o->ng;
  avg[1] = sl / o->ng;
  avg[2] = ss / o->ng;
}

/* operatorVariableCoefficients (src/pde/ElasticitySolver.cpp:268-374):
   f = div(mu (grad u + grad u^T)) + grad(lam div u) - screen u, in the reference's order of
   gradient and divergence calls */
void orc_el_matvec(orc_t *o, const double *mu, const double *lam, const double *screen, const double *u,
                   double *f) {
  const int64_t ng = o->ng;
  double *buf = (double *)malloc(sizeof(double) * ng * 16);
  double *g[3], *du[3][3], *a = buf + 12 * ng, *b = buf + 13 * ng, *cc = buf + 14 * ng, *ld = buf + 15 * ng;
  for (int k = 0; k < 3; k++) g[k] = buf + k * ng;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) du[i][j] = buf + (3 + 3 * i + j) * ng;  /* du[i][j] = d_j u_i */
  const double *ux = u, *uy = u + ng, *uz = u + 2 * ng;
  orc_div(o, ux, uy, uz, ld);
  for (int64_t k = 0; k < ng; k++) ld[k] *= lam[k];
  orc_grad(o, ld, g[0], g[1], g[2]);
  orc_grad(o, ux, du[0][0], du[0][1], du[0][2]);
  orc_grad(o, uy, du[1][0], du[1][1], du[1][2]);
  orc_grad(o, uz, du[2][0], du[2][1], du[2][2]);
  for (int i = 0; i < 3; i++) {
    double *cols[3] = {a, b, cc};
    for (int j = 0; j < 3; j++)
      for (int64_t k = 0; k < ng; k++) cols[j][k] = mu[k] * (du[i][j][k] + du[j][i][k]);
    double *fi = f + i * ng;
    orc_div(o, a, b, cc, fi);
    const double *ui = u + i * ng;
    for (int64_t k = 0; k < ng; k++) fi[k] = fi[k] + g[i][k] - screen[k] * ui[k];
  }
  free(buf);
}

/* operatorConstantCoefficients with precFactorElasticity (src/pde/ElasticitySolver.cpp:122-266,
   src/cuda/ElasticitySolver.cu:36-106): the exact inverse of the operator with the mean mu, lam and
   screen, per wave number w (Nyquist modes zeroed): s1 = -screen - mu |w|^2, s2 = lam + mu,
   s3 = 1 / (1 - |w|^2 s2 / s1), u = f / s1 + (s2 s3 / s1^2) w (w.f) */
void orc_el_prec(orc_t *o, const double *avg, const double *f, double *u) {
  const int64_t ng = o->ng, nh = o->nh;
  const int n1 = o->n[1], n2h = o->n2h;
  cplx *F = (cplx *)malloc(sizeof(cplx) * nh * 3);
  for (int d = 0; d < 3; d++) orc_r2c(o, f + d * ng, F + d * nh);
  const double mu = avg[0], lam = avg[1], scr = avg[2];
  const double factor = 1.0 / ((double)o->n[0] * o->n[1] * o->n[2]);
  const double s2 = lam + mu;
  for (int64_t idx = 0; idx < nh; idx++) {
    const int64_t x = idx / ((int64_t)n1 * n2h), y = (idx / n2h) % n1, z = idx % n2h;
    const double w[3] = {wnum(x, o->n[0]), wnum(y, o->n[1]), wnum(z, o->n[2])};
    const double wTw = -(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double s1 = -scr + mu * wTw;
    const double s3 = 1.0 / (1.0 + (wTw * s2) / s1);
    const cplx wf = w[0] * F[idx] + w[1] * F[nh + idx] + w[2] * F[2 * nh + idx];
    for (int d = 0; d < 3; d++) {
      const cplx scale = -w[d] * wf;
      F[d * nh + idx] = factor * (F[d * nh + idx] / s1 - (1.0 / (s1 * s1)) * s2 * s3 * scale);
    }
  }
  for (int d = 0; d < 3; d++) orc_c2r(o, F + d * nh, u + d * ng);
  free(F);
}

static double el_dot(int64_t n, const double *a, const double *b) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* VariableLinearElasticitySolver::solve (src/pde/ElasticitySolver.cpp:471-533): KSPCG with the
   shell preconditioner, rtol 1e-3, PETSc's default atol 1e-50 and dtol 1e5, maxit 100, nonzero
   initial guess u; the loop and its stopping rules are those of ksp_cg above.  Returns the
   iteration count; *reason_out the KSPConvergedReason; *rnorm_out the last preconditioned
   residual norm. */
int orc_el_solve(orc_t *o, const double *mu, const double *lam, const double *screen, const double *avg,
                 const double *rhs, double *u, int *reason_out, double *rnorm_out) {
  const int64_t N = 3 * o->ng;
  const double rtol = 1e-3, atol = 1e-50, dtol = 1e5;
  const int maxit = 100;
  double *r = (double *)malloc(sizeof(double) * N * 4), *z = r + N, *p = r + 2 * N, *wv = r + 3 * N;
  double dp, beta = 0, betaold = 0, dpi = 0, dpiold, a, bb, rnorm0, ttol;
  int its = 0, reason = 0, i;
  orc_el_matvec(o, mu, lam, screen, u, r);
  for (int64_t k = 0; k < N; k++) r[k] = rhs[k] - r[k];
  orc_el_prec(o, avg, r, z);
  dp = sqrt(el_dot(N, z, z));
  orc_el_prec(o, avg, rhs, wv);
  rnorm0 = sqrt(el_dot(N, wv, wv));
  if (rnorm0 == 0) rnorm0 = dp;
  ttol = fmax(rtol * rnorm0, atol);
#define EL_CONVERGED(rn)                                                          \
  ((isnan(rn) || isinf(rn)) ? KSP_DIVERGED_NANORINF                               \
   : (rn) <= ttol          ? (((rn) < atol) ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL) \
   : (rn) >= dtol * rnorm0 ? KSP_DIVERGED_DTOL                                    \
                           : 0)
  reason = EL_CONVERGED(dp);
  if (!reason) {
    beta = el_dot(N, z, r);
    i = 0;
    do {
      its = i + 1;
      if (beta == 0.0) { reason = KSP_CONVERGED_ATOL; break; }
      else if (i > 0 && beta * betaold < 0.0) { reason = KSP_DIVERGED_INDEFINITE_PC; break; }
      if (!i) {
        memcpy(p, z, sizeof(double) * N);
      } else {
        bb = beta / betaold;
        for (int64_t k = 0; k < N; k++) p[k] = z[k] + bb * p[k];
      }
      dpiold = dpi;
      orc_el_matvec(o, mu, lam, screen, p, wv);
      dpi = el_dot(N, p, wv);
      betaold = beta;
      if (isnan(dpi) || isinf(dpi)) { reason = KSP_DIVERGED_NANORINF; break; }
      if ((dpi == 0.0) || ((i > 0) && (dpi * dpiold <= 0.0))) { reason = KSP_DIVERGED_INDEFINITE_MAT; break; }
      a = beta / dpi;
      for (int64_t k = 0; k < N; k++) { u[k] += a * p[k]; r[k] -= a * wv[k]; }
      orc_el_prec(o, avg, r, z);
      dp = sqrt(el_dot(N, z, z));
      reason = EL_CONVERGED(dp);
      if (reason) break;
      beta = el_dot(N, z, r);
      if (isnan(beta) || isinf(beta)) { reason = KSP_DIVERGED_NANORINF; break; }
      i++;
    } while (i < maxit);
    if (i >= maxit) reason = KSP_DIVERGED_ITS;
  }
#undef EL_CONVERGED
  *reason_out = reason;
  *rnorm_out = dp;
  free(r);
  return its;
}

/* ------------------------------------------------------------------------ */
/* Multispecies pointwise terms (PdeOperatorsMultiSpecies,                   */
/* src/pde/PdeOperatorsMultiSpecies.cpp:10-234; src/cuda/PdeOperators.cu:27-124) */
/* ------------------------------------------------------------------------ */

/* computeReactionRate, computeTransition, computeThesholder and computeSources.  prm =
   death_rate, ox_source, ox_consumption, alpha_0, beta_0, ox_inv, ox_hypoxia. */
void orc_ms_sources(int64_t n, double *p, double *iv, double *nec, double *ox, double *gm, double *wm,
                    const double *rho, const double *di, double dt, const double *prm) {
  const double death = prm[0], oxs = prm[1], oxc = prm[2], a0 = prm[3], b0 = prm[4], oxinv = prm[5], oxh = prm[6];
  for (int64_t k = 0; k < n; k++) {
    const double m = rho[k] * (1 / (1 + exp(-100 * (ox[k] - oxh))));
    const double al = a0 * (1 / (1 + exp(100 * (ox[k] - oxinv))));
    const double be = b0 * ox[k];
    const double h = 1 / (1 + exp(100 * (ox[k] - oxh)));
    const double pt = p[k], it = iv[k];
    p[k] += dt * (m * p[k] * (1. - p[k]) - al * p[k] + be * iv[k] - death * h * p[k]);
    iv[k] += dt * (m * iv[k] * (1. - iv[k]) + al * pt - be * iv[k] - death * h * iv[k]);
    nec[k] += dt * (h * death * (pt + it + gm[k] + wm[k]));
    ox[k] += dt * (-oxc * pt + oxs * (1 - ox[k]) * (gm[k] + wm[k]));
    ox[k] = (ox[k] <= 0.) ? 0. : ox[k];
    double f1 = 0, f2 = 0;
    if (gm[k] > 0.01 || wm[k] > 0.01) {
      f1 = gm[k] / (gm[k] + wm[k]);
      f2 = wm[k] / (gm[k] + wm[k]);
    }
    f1 = isnan(f1) ? 0. : f1;
    f2 = isnan(f2) ? 0. : f2;
    const double loss = m * pt * (1. - pt) + m * it * (1. - it) + di[k];
    gm[k] += -dt * (f1 * loss + h * death * gm[k]);
    wm[k] += -dt * (f2 * loss + h * death * wm[k]);
  }
}

/* updateReacAndDiffCoefficients */
void orc_ms_update_coefficients(int64_t n, double *rho_f, double *k_f, const double *bg, const double *gm,
                                const double *vt, const double *csf, double rho, double k, double r_gm_wm,
                                double k_gm_wm) {
  for (int64_t i = 0; i < n; i++) {
    double t = 1 - (bg[i] + (1 - r_gm_wm) * gm[i] + vt[i] + csf[i]);
    rho_f[i] = (t < 0 ? 0 : t) * rho;
    t = 1 - (bg[i] + (1 - k_gm_wm) * gm[i] + vt[i] + csf[i]);
    k_f[i] = (t < 0 ? 0 : t) * k;
  }
}

/* Tumor::computeForce (src/mat/Tumor.cpp:114-158): f = ff grad(S c) (times tanh(S c) when use_tanh),
   S the Weierstrass smoother with sigma = 2 pi / n0 */
void orc_el_force(orc_t *o, const double *c, double ff, int use_tanh, double *f) {
  const int64_t ng = o->ng;
  double *cs = (double *)malloc(sizeof(double) * ng);
  orc_smooth(o, c, cs, 2.0 * M_PI / o->n[0]);
  orc_grad(o, cs, f, f + ng, f + 2 * ng);
  for (int d = 0; d < 3; d++)
    for (int64_t i = 0; i < ng; i++) f[d * ng + i] *= use_tanh ? ff * tanh(cs[i]) : ff;
  free(cs);
}

/* Tumor::computeSegmentation (src/mat/Tumor.cpp:180-276): labels bg 0, c 1, wm 6, gm 5, vt 7, csf 8 */
void orc_ms_segmentation(int64_t n, const double *bg, const double *gm, const double *wm, const double *vt,
                         const double *csf, const double *c, double *seg) {
  for (int64_t i = 0; i < n; i++) {
    double mx = bg[i];
    int ct = 0;
    const double w = wm[i] * (1 - c[i]), g = gm[i] * (1 - c[i]), v = vt[i] * (1 - c[i]), cs = csf[i] * (1 - c[i]);
    if (c[i] > mx) { mx = c[i]; ct = 1; }
    if (w > mx) { mx = w; ct = 6; }
    if (g > mx) { mx = g; ct = 5; }
    if (v > mx) { mx = v; ct = 7; }
    if (cs > mx) { mx = cs; ct = 8; }
    seg[i] = ct;
  }
}

/* ------------------------------------------------------------------------ */
/* Biophysical features of the model-4 loop                                 */
/* ------------------------------------------------------------------------ */

/* eigenvalues of a symmetric 3x3 matrix a (row-major) by cyclic Jacobi rotations (Golub & Van Loan,
   Matrix Computations, 8.5): sweeps until the off-diagonal part is negligible */
static void sym3_eig_jacobi(const double *a_in, double ev[3]) {
  double a[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) a[r][c] = a_in[3 * r + c];
  for (int sweep = 0; sweep < 50; sweep++) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-30 * dia || off == 0) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (a[p][q] == 0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        const double cs = 1 / sqrt(t * t + 1), sn = t * cs;
        for (int k = 0; k < 3; k++) {  /* A <- A J (columns p, q) */
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = cs * akp - sn * akq;
          a[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; k++) {  /* A <- J^T A (rows p, q) */
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = cs * apk - sn * aqk;
          a[q][k] = sn * apk + cs * aqk;
        }
      }
  }
  for (int k = 0; k < 3; k++) ev[k] = a[k][k];
}

/* computeStressQuants (src/pde/PdeOperatorsMassEffect.cpp:140-212): at one voxel, from the nine
   gradient entries g[3 a + b] = d u_a / d x_b, det(I + grad u), the trace of S = lam tr(E) I + 2 mu E
   with the linearized strain E = (grad u + grad u^T) / 2, and half the spread of S's eigenvalues */
void orc_stress_point(const double *g, double mu, double lam, double *jac, double *tr, double *tau) {
  double f[9], e[9], sm[9], ev[3];
  for (int k = 0; k < 9; k++) f[k] = g[k];
  f[0] += 1; f[4] += 1; f[8] += 1;
  *jac = f[0] * (f[4] * f[8] - f[7] * f[5]) - f[1] * (f[3] * f[8] - f[6] * f[5]) + f[2] * (f[3] * f[7] - f[6] * f[4]);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) e[3 * r + c] = 0.5 * (g[3 * r + c] + g[3 * c + r]);
  const double tre = e[0] + e[4] + e[8];
  for (int k = 0; k < 9; k++) sm[k] = 2 * mu * e[k];
  sm[0] += lam * tre; sm[4] += lam * tre; sm[8] += lam * tre;
  *tr = sm[0] + sm[4] + sm[8];
  sym3_eig_jacobi(sm, ev);
  const double lo = fmin(ev[0], fmin(ev[1], ev[2])), hi = fmax(ev[0], fmax(ev[1], ev[2]));
  *tau = 0.5 * (hi - lo);
}

static int cmp_dbl(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x < y) ? -1 : (x > y ? 1 : 0);
}

/* computeQuantile (src/utils/Utils.cpp:828-880): entry floor(q m) of the m sorted values, q m in
   single precision as the reference multiplies a float quantile by the count; 0 when m = 0 */
static double quantile_of(double *vals, int64_t m, float q) {
  if (m <= 0) return 0.0;
  qsort(vals, (size_t)m, sizeof(double), cmp_dbl);
  int64_t idx = (int64_t)floorf(q * (float)m);
  if (idx >= m) idx = m - 1;
  return vals[idx];
}

/* writeStats (src/pde/PdeOperatorsMassEffect.cpp:261-321) of x: 16 values into r */
static void feature_stats(orc_t *o, const double *x, const double *tc, const double *hb, const double *ind,
                          double ntc, double nh, double *r, double *buf) {
  const int64_t ng = o->ng;
  const double h3 = o->h[0] * o->h[1] * o->h[2];
  double s = 0, st = 0, sh = 0, si = 0;
  for (int64_t i = 0; i < ng; i++) {
    s += x[i];
    st += tc[i] * x[i];
    sh += hb[i] * x[i];
    si += ind[i] * x[i];
  }
  r[0] = s * h3; r[1] = st * h3; r[2] = sh * h3;
  r[3] = si * h3 * o->n[0] / (2.0 * M_PI);
  const double m = s / ng, mt = st / ntc, mh = sh / nh;
  double q = 0, qt = 0, qh = 0;
  for (int64_t i = 0; i < ng; i++) {
    const double a = x[i] - m, b = (x[i] - mt) * tc[i], c = (x[i] - mh) * hb[i];
    q += a * a; qt += b * b; qh += c * c;
  }
  r[4] = sqrt(q / (double)(ng - 1)); r[5] = sqrt(qt / (ntc - 1)); r[6] = sqrt(qh / (nh - 1));
  const float qs[3] = {0.875f, 0.95f, 0.99f};
  for (int reg = 0; reg < 3; reg++) {
    int64_t m = 0;
    for (int64_t i = 0; i < ng; i++)
      if (reg == 0 || (reg == 1 ? tc[i] != 0 : hb[i] != 0)) buf[m++] = x[i];
    for (int k = 0; k < 3; k++) {
      r[7 + 3 * k + reg] = quantile_of(buf, m, qs[k]);  /* buf stays sorted after the first call */
    }
  }
}

/* PdeOperatorsMassEffect::computeBiophysicalFeatures (src/pde/PdeOperatorsMassEffect.cpp:323-418) at
   step it: row[0] = it dt, row[1] = tc / (tc + healthy), then 16 statistics (feature_stats) of each
   of c, dc/dt, |grad c|, |u|, det F, tr S, max shear and |v|.  tc = [seg == 1], healthy =
   [seg not in {0, 1}] with seg from computeSegmentation; dc/dt = D c + rho (c - c^2) - div(c v)
   with o->k and o->rho as the loop last set them (computeTumorQuants, :214-259); the indicator of
   the surface statistic is |c - c*| < 0.1 with c* = c where |grad c| (1 - tc) is largest (first
   index).  u, v: 3 ng; mu, lam: the elasticity material. */
void orc_me_features(orc_t *o, int it, double dt, const double *c, const double *wm, const double *gm,
                     const double *vt, const double *csf, const double *bg, const double *u, const double *v,
                     const double *mu, const double *lam, double *row) {
  const int64_t ng = o->ng;
  double *seg = dalloc(ng), *tc = dalloc(ng), *hb = dalloc(ng), *ind = dalloc(ng), *buf = dalloc(ng);
  double *g = dalloc(3 * ng), *gu = dalloc(9 * ng), *fl = dalloc(3 * ng), *dv = dalloc(ng), *dc = dalloc(ng);
  double *fld = dalloc(8 * ng);
  double *fc = fld, *fdcdt = fld + ng, *fgrad = fld + 2 * ng, *fu = fld + 3 * ng, *fjac = fld + 4 * ng,
         *ftr = fld + 5 * ng, *ftau = fld + 6 * ng, *fv = fld + 7 * ng;
  orc_ms_segmentation(ng, bg, gm, wm, vt, csf, c, seg);
  double ntc = 0, nh = 0;
  for (int64_t i = 0; i < ng; i++) {
    tc[i] = (seg[i] == 1) ? 1.0 : 0.0;
    hb[i] = (seg[i] != 0 && seg[i] != 1) ? 1.0 : 0.0;
    ntc += tc[i];
    nh += hb[i];
  }
  orc_grad(o, c, g, g + ng, g + 2 * ng);
  for (int64_t i = 0; i < ng; i++)
    fgrad[i] = sqrt(g[i] * g[i] + g[ng + i] * g[ng + i] + g[2 * ng + i] * g[2 * ng + i]);
  orc_apply_d(o, c, dc);
  for (int64_t i = 0; i < ng; i++)
    for (int d = 0; d < 3; d++) fl[d * ng + i] = c[i] * v[d * ng + i];
  orc_div(o, fl, fl + ng, fl + 2 * ng, dv);
  for (int64_t i = 0; i < ng; i++) fdcdt[i] = dc[i] + o->rho[i] * (c[i] - c[i] * c[i]) - dv[i];
  int64_t imax = 0;
  double best = -INFINITY;
  for (int64_t i = 0; i < ng; i++) {
    const double w = fgrad[i] * (1 - tc[i]);
    if (w > best) { best = w; imax = i; }
  }
  const double cstar = c[imax];
  for (int64_t i = 0; i < ng; i++) ind[i] = (fabs(c[i] - cstar) < 0.1) ? 1.0 : 0.0;
  for (int a = 0; a < 3; a++) orc_grad(o, u + a * ng, gu + (3 * a) * ng, gu + (3 * a + 1) * ng, gu + (3 * a + 2) * ng);
  for (int64_t i = 0; i < ng; i++) {
    fc[i] = c[i];
    fu[i] = sqrt(u[i] * u[i] + u[ng + i] * u[ng + i] + u[2 * ng + i] * u[2 * ng + i]);
    fv[i] = sqrt(v[i] * v[i] + v[ng + i] * v[ng + i] + v[2 * ng + i] * v[2 * ng + i]);
    double gi[9];
    for (int k = 0; k < 9; k++) gi[k] = gu[k * ng + i];
    orc_stress_point(gi, mu[i], lam[i], &fjac[i], &ftr[i], &ftau[i]);
  }
  row[0] = it * dt;
  row[1] = ntc / (nh + ntc);
  for (int k = 0; k < 8; k++) feature_stats(o, fld + k * ng, tc, hb, ind, ntc, nh, row + 2 + 16 * k, buf);
  free(seg); free(tc); free(hb); free(ind); free(buf); free(g); free(gu); free(fl); free(dv); free(dc); free(fld);
}

/* ------------------------------------------------------------------------ */
/* Mass-effect forward (model 4)                                            */
/* ------------------------------------------------------------------------ */

/* PdeOperatorsMassEffect::solveState(0) (src/pde/PdeOperatorsMassEffect.cpp:421-743) without its
   output and feature branches.  The diffusion part is o's: orc_set_tissue + orc_set_k + orc_set_rho
   fix k_scale, rho_scale, the ratios and kavg as DiffCoef/ReacCoef::setValues do at initialization,
   orc_set_time fixes dt and nt.  Each of the nt + 1 steps: CFL = dt max|v| / h0 from the velocity
   (stop when >= max_cfl), updateReacAndDiffCoefficients (:98-138) into o->k and o->rho (kavg stays),
   precFactor, RK2 trajectories of v, semi-Lagrangian steps of gm, wm (mass conservation), vt, csf
   (pure advection), c (mass conservation) and mri (pure advection), the diffusion solve of c over
   dt, the logistic reaction, the force and the elasticity solve, v = (u - u_old) / dt smoothed with
   sigma 2 pi / n0, u_old = u.  Before the loop: filterTumor (wm, gm *= 1 - c, src/mat/MatProp.cpp:
   205-224), force and elasticity solve into u_old (from zero).
   c: c0 in, c(T) out; wm, gm, vt, csf, mri (mri may be NULL): in and advected out; bg fixed;
   u (3 ng): Tumor::displacement_, the first step's initial guess in and the last displacement out;
   v (3 ng): Tumor::velocity_, in and out; el_prm: the ten parameters of orc_el_material.
   info: steps done, 1 if the CFL test stopped the loop, largest CFL seen, elasticity CG iterations,
   diffusion CG iterations. */
/* With features (may be NULL; max_rows >= nt + 1 rows of 130 values), the loop runs its feature
   branch (:584-604): after precFactor, orc_me_features of the step into row it, then the stop when
   the tumor ratio (single precision) is within 5e-3 of 0.05 or above and it dt >= 1; *rows receives
   the rows written and info[5] is 1 for that stop. */
int orc_me_solve_state_f(orc_t *o, double *c, double *wm, double *gm, double *vt, double *csf, double *mri,
                         const double *bg, double *u, double *v, double ff, int use_tanh, const double *el_prm,
                         double max_cfl, double *info, double *features, int max_rows, int *rows);

int orc_me_solve_state(orc_t *o, double *c, double *wm, double *gm, double *vt, double *csf, double *mri,
                       const double *bg, double *u, double *v, double ff, int use_tanh, const double *el_prm,
                       double max_cfl, double *info) {
  return orc_me_solve_state_f(o, c, wm, gm, vt, csf, mri, bg, u, v, ff, use_tanh, el_prm, max_cfl, info, NULL, 0,
                              NULL);
}

int orc_me_solve_state_f(orc_t *o, double *c, double *wm, double *gm, double *vt, double *csf, double *mri,
                         const double *bg, double *u, double *v, double ff, int use_tanh, const double *el_prm,
                         double max_cfl, double *info, double *features, int max_rows, int *rows) {
  if (rows) *rows = 0;
  if (features && (!rows || max_rows < o->nt + 1)) return -2;
  int vstop = 0;
  const int64_t ng = o->ng;
  const double dt = o->dt, h0 = 2.0 * M_PI / o->n[0], sigma = 2.0 * M_PI / o->n[0];
  double *uold = dalloc(3 * ng), *f = dalloc(3 * ng), *q = dalloc(3 * ng), *tmp = dalloc(ng);
  double *mu = dalloc(ng), *lam = dalloc(ng), *scr = dalloc(ng), avg[3], rn;
  int reason, steps = 0, stopped = 0, el_its = 0, diff_its = 0;
  double cfl_max = 0;
  for (int64_t i = 0; i < ng; i++) {
    wm[i] *= 1.0 - c[i];
    gm[i] *= 1.0 - c[i];
  }
  orc_el_force(o, c, ff, use_tanh, f);
  orc_el_material(o, c, wm, gm, csf, vt, bg, el_prm, mu, lam, scr, avg);
  el_its += orc_el_solve(o, mu, lam, scr, avg, f, uold, &reason, &rn);
  double *vx = v, *vy = v + ng, *vz = v + 2 * ng;
  for (int it = 0; it < o->nt + 1; it++) {
    double vmax = 0;
    for (int64_t i = 0; i < ng; i++) vmax = fmax(vmax, sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
    const double cfl = dt * vmax / h0;
    cfl_max = fmax(cfl_max, cfl);
    if (cfl >= max_cfl) { stopped = 1; break; }
    orc_ms_update_coefficients(ng, o->rho, o->k, bg, gm, vt, csf, o->rho_scale, o->k_scale, o->r_gm_wm, o->k_gm_wm);
    orc_prec_factor(o);
    if (features) {
      double *row = features + (size_t)it * 130;
      orc_me_features(o, it, dt, c, wm, gm, vt, csf, bg, u, v, mu, lam, row);
      (*rows)++;
      const float ratio = (float)row[1];
      if ((fabsf(ratio - 0.05f) < 5e-3f || ratio > 0.05f) && (float)it * (float)dt >= 1.0f) {
        vstop = 1;
        break;
      }
    }
    orc_sl_trajectories(o, vx, vy, vz, dt, q, q + ng, q + 2 * ng);
    orc_sl_solve(o, gm, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
    orc_sl_solve(o, wm, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
    orc_sl_solve(o, vt, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 2);
    orc_sl_solve(o, csf, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 2);
    orc_sl_solve(o, c, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
    if (mri) orc_sl_solve(o, mri, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 2);
    diff_its += orc_diffusion_solve(o, c, dt);
    orc_reaction(o, c, NULL, dt);
    orc_el_force(o, c, ff, use_tanh, f);
    orc_el_material(o, c, wm, gm, csf, vt, bg, el_prm, mu, lam, scr, avg);
    el_its += orc_el_solve(o, mu, lam, scr, avg, f, u, &reason, &rn);
    for (int64_t i = 0; i < 3 * ng; i++) v[i] = (u[i] - uold[i]) * (1.0 / dt);
    for (int d = 0; d < 3; d++) {
      orc_smooth(o, v + d * ng, tmp, sigma);
      memcpy(v + d * ng, tmp, sizeof(double) * ng);
    }
    memcpy(uold, u, sizeof(double) * 3 * ng);
    steps++;
  }
  info[0] = steps;
  info[1] = stopped;
  info[2] = cfl_max;
  info[3] = el_its;
  info[4] = diff_its;
  info[5] = vstop;
  free(uold); free(f); free(q); free(tmp); free(mu); free(lam); free(scr);
  return 0;
}

static void clip0(int64_t n, double *x) {
  for (int64_t i = 0; i < n; i++) x[i] = (x[i] <= 0.) ? 0. : x[i];
}

/* PdeOperatorsMultiSpecies::solveState(0) (src/pde/PdeOperatorsMultiSpecies.cpp:235-508), the
   multispecies mass-effect forward of model 5, without its output, edema and segmentation branches
   (they feed only the output).  The diffusion settings are o's, as for orc_me_solve_state.
   Start (:251-287): p = c0, i = 0, ox = 1, clipTumor (p, i, nec >= 0), filterTumor with p + i, and
   when ff > 0 the force of p and the elasticity solve into u (warm start), u_old = u.  Each of the
   nt + 1 steps: the CFL test, the rho rule of updateReacAndDiffCoefficients, then
   updateIsotropicCoefficients (k from the current wm, gm and csf and its mean over the initial
   brain mask: DiffCoef::setValues, src/mat/DiffCoef.cpp:60-132), precFactor; when ff > 0 the
   advection of gm, wm (mass conservation), vt, csf (pure advection), p, i, nec (mass
   conservation); clipHealthyTissues and clipTumor; i smoothed with sigma 2 pi / n0; di = D i; the
   diffusion solve of i; computeSources; c = p + nec + i; when ff > 0 the force of c, the
   elasticity solve, v = S((u - u_old) / dt), u_old = u.
   nec: in and out (the reference does not reset it); ox out; ms_prm: death_rate, ox_source,
   ox_consumption, alpha_0, beta_0, ox_inv, ox_hypoxia.  info as for orc_me_solve_state. */
/* ed and seg (may be NULL) receive Tumor::species_["edema"] and Tumor::seg_ as the loop leaves them:
   computeEdema (src/mat/Tumor.cpp:160-178: 1[i > ms_prm[7]] smoothed with sigma 2 pi / n0) and
   computeSegmentation at the top of every step (src/pde/PdeOperatorsMultiSpecies.cpp:307-308), with
   c_t_ = p at the first step and p + nec + i afterwards. */
int orc_ms_solve_state(orc_t *o, const double *c0, double *p, double *iv, double *nec, double *ox, double *wm,
                       double *gm, double *vt, double *csf, const double *bg, double *u, double *v, double ff,
                       int use_tanh, const double *el_prm, const double *ms_prm, double max_cfl, double *info,
                       double *ed, double *seg) {
  const int64_t ng = o->ng;
  const double dt = o->dt, h0 = 2.0 * M_PI / o->n[0], sigma = 2.0 * M_PI / o->n[0];
  double *uold = dalloc(3 * ng), *f = dalloc(3 * ng), *q = dalloc(3 * ng), *tmp = dalloc(ng);
  double *ct = dalloc(ng), *di = dalloc(ng), *mu = dalloc(ng), *lam = dalloc(ng), *scr = dalloc(ng), avg[3], rn;
  int reason, steps = 0, stopped = 0, el_its = 0, diff_its = 0;
  double cfl_max = 0;
  double *vx = v, *vy = v + ng, *vz = v + 2 * ng;
  for (int64_t i = 0; i < ng; i++) {
    p[i] = c0[i];
    iv[i] = 0.0;
    ox[i] = 1.0;
  }
  clip0(ng, p); clip0(ng, iv); clip0(ng, nec);
  for (int64_t i = 0; i < ng; i++) {
    ct[i] = p[i] + iv[i];
    wm[i] *= 1.0 - ct[i];
    gm[i] *= 1.0 - ct[i];
  }
  if (ff > 0) {
    orc_el_force(o, p, ff, use_tanh, f);
    orc_el_material(o, p, wm, gm, csf, vt, bg, el_prm, mu, lam, scr, avg);
    el_its += orc_el_solve(o, mu, lam, scr, avg, f, u, &reason, &rn);
    memcpy(uold, u, sizeof(double) * 3 * ng);
  }
  for (int it = 0; it < o->nt + 1; it++) {
    if (ed) {
      for (int64_t i = 0; i < ng; i++) tmp[i] = (iv[i] > ms_prm[7]) ? 1.0 : 0.0;
      orc_smooth(o, tmp, ed, sigma);
    }
    if (seg) orc_ms_segmentation(ng, bg, gm, wm, vt, csf, it == 0 ? p : ct, seg);
    double vmax = 0;
    for (int64_t i = 0; i < ng; i++) vmax = fmax(vmax, sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
    const double cfl = dt * vmax / h0;
    cfl_max = fmax(cfl_max, cfl);
    if (cfl >= max_cfl) { stopped = 1; break; }
    orc_ms_update_coefficients(ng, o->rho, o->k, bg, gm, vt, csf, o->rho_scale, o->k_scale, o->r_gm_wm, o->k_gm_wm);
    memcpy(o->wm, wm, sizeof(double) * ng);
    memcpy(o->gm, gm, sizeof(double) * ng);
    memcpy(o->csf, csf, sizeof(double) * ng);
    memcpy(o->vt, vt, sizeof(double) * ng);
    orc_set_k(o, o->k_scale, o->k_gm_wm, o->k_glm_wm);
    orc_prec_factor(o);
    if (ff > 0) {
      orc_sl_trajectories(o, vx, vy, vz, dt, q, q + ng, q + 2 * ng);
      orc_sl_solve(o, gm, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
      orc_sl_solve(o, wm, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
      orc_sl_solve(o, vt, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 2);
      orc_sl_solve(o, csf, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 2);
      orc_sl_solve(o, p, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
      orc_sl_solve(o, iv, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
      orc_sl_solve(o, nec, vx, vy, vz, q, q + ng, q + 2 * ng, dt, 1);
    }
    clip0(ng, gm); clip0(ng, wm); clip0(ng, csf); clip0(ng, vt);
    clip0(ng, p); clip0(ng, iv); clip0(ng, nec);
    orc_smooth(o, iv, tmp, sigma);
    memcpy(iv, tmp, sizeof(double) * ng);
    orc_apply_d(o, iv, di);
    diff_its += orc_diffusion_solve(o, iv, dt);
    orc_ms_sources(ng, p, iv, nec, ox, gm, wm, o->rho, di, dt, ms_prm);
    for (int64_t i = 0; i < ng; i++) ct[i] = p[i] + nec[i] + iv[i];
    if (ff > 0) {
      orc_el_force(o, ct, ff, use_tanh, f);
      orc_el_material(o, ct, wm, gm, csf, vt, bg, el_prm, mu, lam, scr, avg);
      el_its += orc_el_solve(o, mu, lam, scr, avg, f, u, &reason, &rn);
      for (int64_t i = 0; i < 3 * ng; i++) v[i] = (u[i] - uold[i]) * (1.0 / dt);
      for (int d = 0; d < 3; d++) {
        orc_smooth(o, v + d * ng, tmp, sigma);
        memcpy(v + d * ng, tmp, sizeof(double) * ng);
      }
      memcpy(uold, u, sizeof(double) * 3 * ng);
    }
    steps++;
  }
  info[0] = steps;
  info[1] = stopped;
  info[2] = cfl_max;
  info[3] = el_its;
  info[4] = diff_its;
  free(uold); free(f); free(q); free(tmp); free(ct); free(di); free(mu); free(lam); free(scr);
  return 0;
}
