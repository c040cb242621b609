// glia_opt.cpp -- bound-constrained limited-memory quasi-Newton minimization with an Armijo line
// search (the BLMVM algorithm the reference runs through TAO) and its use on an RD handle
// (include/glia_opt.h).
//
// P(g) is the projected gradient: components of g that would push x out through an active bound
// (x_i at its lower bound with g_i > 0, at its upper bound with g_i < 0) are zero.
// Per problem, one iteration:
//   d = -H g (two-loop L-BFGS, initial scaling s'y / y'y of the newest pair; without pairs
//   H = delta I with delta = 2 |f| / ||g||^2, or 2 / ||g||^2 when f = 0, TAO's first-step scaling);
//   components that point out of an active bound are zeroed; if g'd >= 0 the memory is dropped
//   and d = -delta g.  Armijo: x(a) = P(x + a d) for a = 1, 1/2, 1/4, ... up to ls_max_funcs
//   trials, accepted when f(x(a)) <= f + sigma g'(x(a) - x).  A failed search is retried once along
//   -delta g with an empty memory; a direction with g'd >= 0 even then (the projected gradient
//   vanishes) ends the search as failed, as an Armijo search along an ascent direction does.  The
//   pair (s, y) with y = P(g_new) - P(g_old) (BLMVM's update with projected gradients, so an active
//   bound adds no curvature) enters the memory when s'y > 0.
// Stopping test of the reference (checkConvergenceGrad) on ||g|| = ||P(g)|| (the gradient vector a
// BLMVM solver holds is the projected one); from iteration
// min_it on, an accepted step length a below ls_minstep (STEPTOL), ||g|| < gttol ||g_0|| (GTTOL),
// ||g|| < gatol (GATOL), it > max_it (MAXITS); before min_it only GATOL.  NaN objective or gradient:
// NAN.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <vector>

#include "../../include/glia_opt.h"

namespace {

struct Problem {
  int n = 0;
  double* x = nullptr;
  const double* lb = nullptr;
  const double* ub = nullptr;
  std::vector<double> g, gp, d, xt;
  double f = 0, gnorm = 0, gnorm0 = 1, gtd = 0, alpha = 1;
  int ls_k = 0, it = 0, fevals = 0, reason = GLIA_OPT_CONTINUE;
  bool started = false, reset_tried = false;
  std::deque<std::vector<double>> S, Y;
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

double lo(const Problem& p, int i) { return p.lb ? p.lb[i] : -std::numeric_limits<double>::infinity(); }
double hi(const Problem& p, int i) { return p.ub ? p.ub[i] : std::numeric_limits<double>::infinity(); }

// P(g) at x
void project_gradient(const Problem& p, const double* x, const double* g, std::vector<double>& gp) {
  for (int i = 0; i < p.n; i++)
    gp[i] = ((x[i] <= lo(p, i) && g[i] > 0) || (x[i] >= hi(p, i) && g[i] < 0)) ? 0.0 : g[i];
}

void direction(Problem& p, bool steepest) {
  const int n = p.n;
  p.d = p.g;
  if (!steepest && !p.S.empty()) {
    const int m = (int)p.S.size();
    std::vector<double> a(m), rho(m);
    for (int k = m - 1; k >= 0; k--) {
      rho[k] = 1.0 / dot(p.Y[k], p.S[k]);
      a[k] = rho[k] * dot(p.S[k], p.d);
      for (int i = 0; i < n; i++) p.d[i] -= a[k] * p.Y[k][i];
    }
    const double gamma = dot(p.S.back(), p.Y.back()) / dot(p.Y.back(), p.Y.back());
    for (int i = 0; i < n; i++) p.d[i] *= gamma;
    for (int k = 0; k < m; k++) {
      const double b = rho[k] * dot(p.Y[k], p.d);
      for (int i = 0; i < n; i++) p.d[i] += (a[k] - b) * p.S[k][i];
    }
  } else if (p.gnorm > 0) {
    const double delta = 2.0 * (p.f != 0 ? std::fabs(p.f) : 1.0) / (p.gnorm * p.gnorm);
    for (int i = 0; i < n; i++) p.d[i] *= delta;
  }
  for (int i = 0; i < n; i++) {
    p.d[i] = -p.d[i];
    // a component that leaves the box from an active bound is dropped
    if ((p.x[i] <= lo(p, i) && p.d[i] < 0) || (p.x[i] >= hi(p, i) && p.d[i] > 0)) p.d[i] = 0;
  }
  p.gtd = dot(p.g, p.d);
}

void trial(Problem& p) {
  for (int i = 0; i < p.n; i++) p.xt[i] = std::min(std::max(p.x[i] + p.alpha * p.d[i], lo(p, i)), hi(p, i));
}

void new_search(Problem& p) {
  direction(p, false);
  if (!(p.gtd < 0)) {
    p.S.clear();
    p.Y.clear();
    direction(p, true);
  }
  if (!(p.gtd < 0)) {
    p.reason = GLIA_OPT_DIVERGED_LS_FAILURE;
    return;
  }
  p.alpha = 1.0;
  p.ls_k = 0;
  trial(p);
}

// the reference's stopping test after an accepted step (or at the start)
bool converged(Problem& p, const glia_opt_settings& s, double step) {
  if (!std::isfinite(p.f) || !std::isfinite(p.gnorm)) {
    p.reason = GLIA_OPT_DIVERGED_NAN;
    return true;
  }
  if (p.it >= s.min_it) {
    if (p.it > 0 && step < s.ls_minstep) p.reason = GLIA_OPT_CONVERGED_STEPTOL;
    else if (p.gnorm < s.gttol * p.gnorm0) p.reason = GLIA_OPT_CONVERGED_GTTOL;
    else if (p.gnorm < s.gatol) p.reason = GLIA_OPT_CONVERGED_GATOL;
    else if (p.it > s.max_it) p.reason = GLIA_OPT_DIVERGED_MAXITS;
  } else if (p.gnorm < s.gatol) {
    p.reason = GLIA_OPT_CONVERGED_GATOL;
  }
  return p.reason != GLIA_OPT_CONTINUE;
}

// consume one evaluation (f, g at the problem's requested point)
void advance(Problem& p, const glia_opt_settings& s, double f, const double* g) {
  p.fevals++;
  if (!p.started) {
    p.started = true;
    p.f = f;
    std::copy(g, g + p.n, p.g.begin());
    project_gradient(p, p.x, g, p.gp);
    p.gnorm = std::sqrt(dot(p.gp, p.gp));
    p.gnorm0 = (p.gnorm > 0) ? p.gnorm : 1.0;
    if (!converged(p, s, 1.0)) new_search(p);
    return;
  }
  double decrease = 0;
  for (int i = 0; i < p.n; i++) decrease += p.g[i] * (p.xt[i] - p.x[i]);
  if (std::isfinite(f) && f <= p.f + s.ls_sigma * decrease) {
    std::vector<double> sv(p.n), yv(p.n), gp(p.n);
    project_gradient(p, p.xt.data(), g, gp);
    for (int i = 0; i < p.n; i++) {
      sv[i] = p.xt[i] - p.x[i];
      yv[i] = gp[i] - p.gp[i];
    }
    if (dot(sv, yv) > 0) {
      p.S.push_back(sv);
      p.Y.push_back(yv);
      if ((int)p.S.size() > std::max(1, s.lbfgs_vectors)) {
        p.S.pop_front();
        p.Y.pop_front();
      }
    }
    std::copy(p.xt.begin(), p.xt.end(), p.x);
    p.f = f;
    std::copy(g, g + p.n, p.g.begin());
    p.gp.swap(gp);
    p.gnorm = std::sqrt(dot(p.gp, p.gp));
    p.it++;
    p.reset_tried = false;
    if (!converged(p, s, p.alpha)) new_search(p);
    return;
  }
  if (++p.ls_k < s.ls_max_funcs) {
    p.alpha *= 0.5;
    trial(p);
    return;
  }
  if (!p.reset_tried) {
    p.reset_tried = true;
    p.S.clear();
    p.Y.clear();
    direction(p, true);
    if (p.gtd < 0) {
      p.alpha = 1.0;
      p.ls_k = 0;
      trial(p);
      return;
    }
  }
  p.reason = GLIA_OPT_DIVERGED_LS_FAILURE;
}

struct RdCtx {
  glia_rd_t h;
  int mode;
};

int rd_fg(void* vctx, int nprob, const double* x, double* f, double* g) {
  RdCtx* c = static_cast<RdCtx*>(vctx);
  (void)nprob;
  return glia_rd_objective_and_gradient(c->h, c->mode, x, f, g);
}

}  // namespace

extern "C" {

void glia_opt_settings_init(glia_opt_settings* s) {
  if (!s) return;
  s->max_it = 30;
  s->min_it = 1;
  s->gatol = 1e-4;
  s->gttol = 1e-3;
  s->lbfgs_vectors = 10;
  s->ls_max_funcs = 10;
  s->ls_sigma = 1e-4;
  s->ls_minstep = std::numeric_limits<double>::epsilon();
}

int glia_blmvm(int nprob, int n, double* x, const double* lb, const double* ub, glia_fg_fn fg, void* ctx,
               const glia_opt_settings* sp, glia_opt_result* r) {
  if (nprob <= 0 || n <= 0 || !x || !fg) return GLIA_RD_ERR_ARG;
  glia_opt_settings s;
  if (sp) {
    s = *sp;
  } else {
    glia_opt_settings_init(&s);
  }
  std::vector<Problem> P(nprob);
  for (int m = 0; m < nprob; m++) {
    Problem& p = P[m];
    p.n = n;
    p.x = x + (size_t)m * n;
    p.lb = lb ? lb + (size_t)m * n : nullptr;
    p.ub = ub ? ub + (size_t)m * n : nullptr;
    p.g.assign(n, 0.0);
    p.gp.assign(n, 0.0);
    p.xt.assign(n, 0.0);
    for (int i = 0; i < n; i++) p.x[i] = std::min(std::max(p.x[i], lo(p, i)), hi(p, i));
    std::copy(p.x, p.x + n, p.xt.begin());
  }
  std::vector<double> xe((size_t)nprob * n), fe(nprob), ge((size_t)nprob * n);
  for (;;) {
    // each live problem asks for its trial point; problems that stopped are evaluated at their
    // result (a batched evaluation covers every member)
    bool active = false;
    for (int m = 0; m < nprob; m++) {
      const Problem& p = P[m];
      const bool live = p.reason == GLIA_OPT_CONTINUE;
      active = active || live;
      if (live) {
        std::copy(p.xt.begin(), p.xt.end(), xe.begin() + (size_t)m * n);
      } else {
        std::copy(p.x, p.x + n, xe.begin() + (size_t)m * n);
      }
    }
    if (!active) break;
    const int rc = fg(ctx, nprob, xe.data(), fe.data(), ge.data());
    if (rc) return rc;
    for (int m = 0; m < nprob; m++)
      if (P[m].reason == GLIA_OPT_CONTINUE) advance(P[m], s, fe[m], ge.data() + (size_t)m * n);
  }
  if (r)
    for (int m = 0; m < nprob; m++) {
      r[m].iterations = P[m].it;
      r[m].fevals = P[m].fevals;
      r[m].reason = P[m].reason;
      r[m].f = P[m].f;
      r[m].gnorm = P[m].gnorm;
      r[m].gnorm0 = P[m].gnorm0;
    }
  return 0;
}

int glia_rd_invert(glia_rd_t h, int mode, double* x, const double* lb, const double* ub, const glia_opt_settings* s,
                   glia_opt_result* r) {
  const int nm = glia_rd_num_members(h);
  if (nm <= 0 || !x) return GLIA_RD_ERR_ARG;
  const int np = glia_rd_basis_size(h, 0);
  for (int m = 1; m < nm; m++)
    if (glia_rd_basis_size(h, m) != np) return GLIA_RD_ERR_ARG;
  RdCtx c{h, mode};
  return glia_blmvm(nm, np + 2, x, lb, ub, rd_fg, &c, s, r);
}

}  // extern "C"
