// glia_opt.cpp -- bound-constrained limited-memory quasi-Newton minimization with a More-Thuente or
// an Armijo line search (the BLMVM algorithm the reference runs through TAO, src/opt/Optimizer.cpp:
// 116-260) and its use on an RD handle (include/glia_opt.h).
//
// P(g) is the projected gradient: components of g that would push x out through an active bound
// (x_i at its lower bound with g_i > 0, at its upper bound with g_i < 0) are zero.
// Per problem, one iteration:
//   d = -H g (two-loop L-BFGS, initial scaling s'y / y'y of the newest pair; without pairs
//   H = delta I with delta = 2 |f| / ||g||^2, or 2 / ||g||^2 when f = 0, TAO's first-step scaling);
//   components that point out of an active bound are zeroed; if g'd >= 0 the memory is dropped
//   and d = -delta g.  Trial points are x(a) = P(x + a d).  More-Thuente (the default): the
//   strong Wolfe conditions phi(a) <= phi(0) + sigma a phi'(0), |phi'(a)| <= gtol |phi'(0)| on
//   phi(a) = f(x(a)), with phi'(a) = g(x(a))'(x(a) - x)/a along the projected path, a capped where
//   every moving variable sits on its bound, at most ls_max_funcs evaluations; a search that halts
//   first takes its best sufficient-decrease point, if any.  Armijo: a = 1, 1/2, 1/4, ... up to
//   ls_max_funcs trials, accepted when f(x(a)) <= f + sigma g'(x(a) - x).  A failed search is retried once along
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
#include <functional>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/glia_opt.h"

namespace {

// More-Thuente search state (the interval [stx, sty] that brackets a step satisfying the strong
// Wolfe conditions, with function values and derivatives of phi(a) = f(P(x + a d)) at its ends)
struct MtState {
  double stx = 0, fx = 0, gx = 0, sty = 0, fy = 0, gy = 0;
  double stmin = 0, stmax = 0, stpmin = 0, stpmax = 0, width = 0, width1 = 0;
  double finit = 0, ginit = 0, gtest = 0;
  int stage = 1;
  bool brackt = false;
  // the point at stx (the best sufficient-decrease point so far), kept for a halted search
  std::vector<double> xb, gb;
  double fb = 0, ab = 0;
  bool have_best = false;
};

struct Problem {
  int n = 0;
  double* x = nullptr;
  const double* lb = nullptr;
  const double* ub = nullptr;
  std::vector<double> g, gp, d, xt;
  double f = 0, gnorm = 0, gnorm0 = 0, gtd = 0, alpha = 1;
  int ls_k = 0, it = 0, fevals = 0, reason = GLIA_OPT_CONTINUE;
  bool started = false, reset_tried = false;
  std::deque<std::vector<double>> S, Y;
  MtState mt;
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

// ---- More-Thuente line search (J. J. More and D. J. Thuente, "Line search algorithms with
// guaranteed sufficient decrease", ACM TOMS 20 (1994) 286-307; the scheme of MINPACK-2's
// dcsrch/dcstep): the step update of the safeguarded interval
void mt_step(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp, double fp, double dp,
             bool& brackt, double stpmin, double stpmax) {
  const double sgnd = dp * (dx / std::fabs(dx));
  double stpf;
  auto cubic_gamma = [](double theta, double a, double b, double s, bool clamp) {
    double r = (theta / s) * (theta / s) - (a / s) * (b / s);
    if (clamp) r = std::max(0.0, r);
    return s * std::sqrt(r);
  };
  if (fp > fx) {
    // higher function value: the minimizer is bracketed; cubic step, or halfway to the quadratic
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
    double gamma = cubic_gamma(theta, dx, dp, s, false);
    if (stp < stx) gamma = -gamma;
    const double pp = (gamma - dx) + theta, q = ((gamma - dx) + gamma) + dp;
    const double stpc = stx + (pp / q) * (stp - stx);
    const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
    stpf = (std::fabs(stpc - stx) < std::fabs(stpq - stx)) ? stpc : stpc + (stpq - stpc) / 2.0;
    brackt = true;
  } else if (sgnd < 0.0) {
    // lower value, derivatives of opposite sign: bracketed; cubic or secant, the farther
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
    double gamma = cubic_gamma(theta, dx, dp, s, false);
    if (stp > stx) gamma = -gamma;
    const double pp = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dx;
    const double stpc = stp + (pp / q) * (stx - stp);
    const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
    stpf = (std::fabs(stpc - stp) > std::fabs(stpq - stp)) ? stpc : stpq;
    brackt = true;
  } else if (std::fabs(dp) < std::fabs(dx)) {
    // lower value, same-sign derivatives, derivative magnitude decreasing
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
    double gamma = cubic_gamma(theta, dx, dp, s, true);
    if (stp > stx) gamma = -gamma;
    const double pp = (gamma - dp) + theta, q = (gamma + (dx - dp)) + gamma;
    const double r = pp / q;
    double stpc;
    if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
    else if (stp > stx) stpc = stpmax;
    else stpc = stpmin;
    const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
    if (brackt) {
      stpf = (std::fabs(stpc - stp) < std::fabs(stpq - stp)) ? stpc : stpq;
      stpf = (stp > stx) ? std::min(stp + 0.66 * (sty - stp), stpf) : std::max(stp + 0.66 * (sty - stp), stpf);
    } else {
      stpf = (std::fabs(stpc - stp) > std::fabs(stpq - stp)) ? stpc : stpq;
      stpf = std::max(stpmin, std::min(stpmax, stpf));
    }
  } else {
    // lower value, same-sign derivatives, derivative magnitude not decreasing
    if (brackt) {
      const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
      double gamma = cubic_gamma(theta, dy, dp, s, false);
      if (stp > sty) gamma = -gamma;
      const double pp = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dy;
      stpf = stp + (pp / q) * (sty - stp);
    } else {
      stpf = (stp > stx) ? stpmax : stpmin;
    }
  }
  // the new interval
  if (fp > fx) {
    sty = stp;
    fy = fp;
    dy = dp;
  } else {
    if (sgnd < 0.0) {
      sty = stx;
      fy = fx;
      dy = dx;
    }
    stx = stp;
    fx = fp;
    dx = dp;
  }
  stp = stpf;
}

// largest step before every moving variable sits on its bound (the projected path is constant
// beyond it)
double max_feasible_step(const Problem& p) {
  double rmax = 0;
  for (int i = 0; i < p.n; i++) {
    if (p.d[i] > 0) {
      const double b = hi(p, i);
      if (!std::isfinite(b)) return std::numeric_limits<double>::infinity();
      rmax = std::max(rmax, (b - p.x[i]) / p.d[i]);
    } else if (p.d[i] < 0) {
      const double b = lo(p, i);
      if (!std::isfinite(b)) return std::numeric_limits<double>::infinity();
      rmax = std::max(rmax, (b - p.x[i]) / p.d[i]);
    }
  }
  return rmax;
}

void mt_begin(Problem& p, const glia_opt_settings& s) {
  MtState& m = p.mt;
  m = MtState();
  m.finit = p.f;
  m.ginit = p.gtd;
  m.gtest = s.ls_sigma * m.ginit;
  m.stpmin = s.ls_minstep;
  m.stpmax = std::min(max_feasible_step(p), 1e15);
  m.width = m.stpmax - m.stpmin;
  m.width1 = 2.0 * m.width;
  m.fx = m.fy = m.finit;
  m.gx = m.gy = m.ginit;
  p.alpha = std::max(m.stpmin, std::min(1.0, m.stpmax));
  m.stmin = 0;
  m.stmax = p.alpha + 4.0 * p.alpha;
}

enum { LS_NEXT = 0, LS_ACCEPT = 1, LS_ACCEPT_BEST = 2, LS_FAIL = 3 };

// one evaluation of the search at step p.alpha (point p.xt, value f, gradient g)
int mt_advance(Problem& p, const glia_opt_settings& s, double f, const double* g) {
  MtState& m = p.mt;
  double stp = p.alpha;
  if (!std::isfinite(f)) return m.have_best ? LS_ACCEPT_BEST : LS_FAIL;
  // derivative of phi along the projected path
  double dg = 0;
  for (int i = 0; i < p.n; i++) dg += g[i] * (p.xt[i] - p.x[i]);
  dg /= stp;
  const double ftest = m.finit + stp * m.gtest;
  if (m.stage == 1 && f <= ftest && dg >= 0) m.stage = 2;
  if (f <= ftest && std::fabs(dg) <= s.ls_gtol * (-m.ginit)) return LS_ACCEPT;
  // a point with sufficient decrease that lowers phi becomes the fallback of a halted search
  const bool record = f <= ftest && (!m.have_best || f < m.fb);
  if (record) {
    m.xb = p.xt;
    m.gb.assign(g, g + p.n);
    m.fb = f;
    m.ab = stp;
    m.have_best = true;
  }
  const bool halted = (m.brackt && (stp <= m.stmin || stp >= m.stmax)) ||
                      (m.brackt && m.stmax - m.stmin <= s.ls_xtol * m.stmax) ||
                      (stp == m.stpmax && f <= ftest && dg <= m.gtest) ||
                      (stp == m.stpmin && (f > ftest || dg >= m.gtest)) || (++p.ls_k >= s.ls_max_funcs);
  if (halted) return m.have_best ? LS_ACCEPT_BEST : LS_FAIL;
  if (m.stage == 1 && f <= m.fx && f > ftest) {
    // the modified function psi(a) = phi(a) - a gtest until a sufficient-decrease point is found
    double fm = f - stp * m.gtest, fxm = m.fx - m.stx * m.gtest, fym = m.fy - m.sty * m.gtest;
    double gm = dg - m.gtest, gxm = m.gx - m.gtest, gym = m.gy - m.gtest;
    mt_step(m.stx, fxm, gxm, m.sty, fym, gym, stp, fm, gm, m.brackt, m.stmin, m.stmax);
    m.fx = fxm + m.stx * m.gtest;
    m.fy = fym + m.sty * m.gtest;
    m.gx = gxm + m.gtest;
    m.gy = gym + m.gtest;
  } else {
    mt_step(m.stx, m.fx, m.gx, m.sty, m.fy, m.gy, stp, f, dg, m.brackt, m.stmin, m.stmax);
  }
  if (m.brackt) {
    if (std::fabs(m.sty - m.stx) >= 0.66 * m.width1) stp = m.stx + 0.5 * (m.sty - m.stx);
    m.width1 = m.width;
    m.width = std::fabs(m.sty - m.stx);
    m.stmin = std::min(m.stx, m.sty);
    m.stmax = std::max(m.stx, m.sty);
  } else {
    m.stmin = stp + 1.1 * (stp - m.stx);
    m.stmax = stp + 4.0 * (stp - m.stx);
  }
  stp = std::max(m.stpmin, std::min(m.stpmax, stp));
  if ((m.brackt && (stp <= m.stmin || stp >= m.stmax)) || (m.brackt && m.stmax - m.stmin <= s.ls_xtol * m.stmax))
    stp = m.stx;
  if (!(stp > 0)) return m.have_best ? LS_ACCEPT_BEST : LS_FAIL;
  p.alpha = stp;
  trial(p);
  return LS_NEXT;
}

void start_search(Problem& p, const glia_opt_settings& s) {
  p.ls_k = 0;
  if (s.line_search == GLIA_OPT_LS_MORE_THUENTE) {
    mt_begin(p, s);
  } else {
    p.alpha = 1.0;
  }
  trial(p);
}

void new_search(Problem& p, const glia_opt_settings& s) {
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
  start_search(p, s);
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

// take the point xn (value f, gradient g) reached with step length a as the next iterate
void accept(Problem& p, const glia_opt_settings& s, const std::vector<double>& xn, double f, const double* g,
            double a) {
  std::vector<double> sv(p.n), yv(p.n), gp(p.n);
  project_gradient(p, xn.data(), g, gp);
  for (int i = 0; i < p.n; i++) {
    sv[i] = xn[i] - p.x[i];
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
  std::copy(xn.begin(), xn.end(), p.x);
  p.f = f;
  std::copy(g, g + p.n, p.g.begin());
  p.gp.swap(gp);
  p.gnorm = std::sqrt(dot(p.gp, p.gp));
  p.it++;
  p.reset_tried = false;
  if (!converged(p, s, a)) new_search(p, s);
}

// a failed search is retried once along the scaled steepest descent with an empty memory
void fail_search(Problem& p, const glia_opt_settings& s) {
  if (!p.reset_tried) {
    p.reset_tried = true;
    p.S.clear();
    p.Y.clear();
    direction(p, true);
    if (p.gtd < 0) {
      start_search(p, s);
      return;
    }
  }
  p.reason = GLIA_OPT_DIVERGED_LS_FAILURE;
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
    if (!(p.gnorm0 > 0)) p.gnorm0 = (p.gnorm > 0) ? p.gnorm : 1.0;
    if (!converged(p, s, 1.0)) new_search(p, s);
    return;
  }
  if (s.line_search == GLIA_OPT_LS_MORE_THUENTE) {
    const int st = mt_advance(p, s, f, g);
    if (st == LS_ACCEPT) {
      accept(p, s, p.xt, f, g, p.alpha);
    } else if (st == LS_ACCEPT_BEST) {
      const std::vector<double> xb = p.mt.xb, gb = p.mt.gb;
      accept(p, s, xb, p.mt.fb, gb.data(), p.mt.ab);
    } else if (st == LS_FAIL) {
      fail_search(p, s);
    }
    return;
  }
  // Armijo
  double decrease = 0;
  for (int i = 0; i < p.n; i++) decrease += p.g[i] * (p.xt[i] - p.x[i]);
  if (std::isfinite(f) && f <= p.f + s.ls_sigma * decrease) {
    accept(p, s, p.xt, f, g, p.alpha);
    return;
  }
  if (++p.ls_k < s.ls_max_funcs) {
    p.alpha *= 0.5;
    trial(p);
    return;
  }
  fail_search(p, s);
}

using FG = std::function<int(int nprob, const double* x, double* f, double* g)>;

// the loop; gref (nprob, optional): reference gradient norms for the relative test (the
// reference's gradnorm0_); without them the initial projected gradient norm is used
int blmvm(int nprob, int n, double* x, const double* lb, const double* ub, const FG& fg, const glia_opt_settings* sp,
          const double* gref, glia_opt_result* r) {
  if (nprob <= 0 || n <= 0 || !x) return GLIA_RD_ERR_ARG;
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
    if (gref && gref[m] > 0) p.gnorm0 = gref[m];
    else if (s.gnorm_ref > 0) p.gnorm0 = s.gnorm_ref;
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
    const int rc = fg(nprob, xe.data(), fe.data(), ge.data());
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

using HV = std::function<int(int nprob, const double* v, double* y)>;

// take the point reached by the line search as the next Gauss-Newton iterate (no quasi-Newton
// memory); returns whether the problem continues
void accept_gn(Problem& p, const glia_opt_settings& s, const std::vector<double>& xn, double f, const double* g,
               double a) {
  std::copy(xn.begin(), xn.end(), p.x);
  p.f = f;
  std::copy(g, g + p.n, p.g.begin());
  project_gradient(p, p.x, g, p.gp);
  p.gnorm = std::sqrt(dot(p.gp, p.gp));
  p.it++;
  converged(p, s, a);
}

// Gauss-Newton-Krylov with bounds (the reference's GAUSSNEWTON solver runs TAO's BNLS,
// src/opt/Optimizer.cpp:127-131, with a CG Krylov solve of at most krylov_maxit iterations and the
// Eisenstat-Walker tolerance of preKrylovSolve, src/opt/TaoInterface.cpp:160-193).  Per iteration
// and problem: the variables on an active bound (P(g)_i = 0 there) are held, H_FF s_F = -g_F is
// solved by CG to relative residual max(1e-10, min(0.5, sqrt(||g_F|| / ||g_F,0||))) (quadratic
// forcing: without the square root), a direction that is not a descent direction is replaced by
// -delta P(g), and the line search of the settings runs along it.  The problems of a batch share
// every Hessian round (hv) and evaluation round (fg).
int gauss_newton(int nprob, int n, double* x, const double* lb, const double* ub, const FG& fg, const HV& hv,
                 const glia_opt_settings* sp, const double* gref, glia_opt_result* r) {
  if (nprob <= 0 || n <= 0 || !x) return GLIA_RD_ERR_ARG;
  glia_opt_settings s;
  if (sp) {
    s = *sp;
  } else {
    glia_opt_settings_init(&s);
  }
  const size_t N = (size_t)nprob * n;
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
    p.d.assign(n, 0.0);
    if (gref && gref[m] > 0) p.gnorm0 = gref[m];
    else if (s.gnorm_ref > 0) p.gnorm0 = s.gnorm_ref;
    for (int i = 0; i < n; i++) p.x[i] = std::min(std::max(p.x[i], lo(p, i)), hi(p, i));
  }
  std::vector<double> xe(N), fe(nprob), ge(N), v(N), y(N);
  int rc = fg(nprob, x, fe.data(), ge.data());
  if (rc) return rc;
  for (int m = 0; m < nprob; m++) {
    Problem& p = P[m];
    p.fevals++;
    p.started = true;
    p.f = fe[m];
    std::copy(ge.begin() + (size_t)m * n, ge.begin() + (size_t)(m + 1) * n, p.g.begin());
    project_gradient(p, p.x, p.g.data(), p.gp);
    p.gnorm = std::sqrt(dot(p.gp, p.gp));
    if (!(p.gnorm0 > 0)) p.gnorm0 = (p.gnorm > 0) ? p.gnorm : 1.0;
    converged(p, s, 1.0);
  }
  std::vector<double> bnorm0(nprob, 0.0);
  std::vector<std::vector<double>> sv(nprob), rv(nprob), dv(nprob);
  std::vector<std::vector<char>> held(nprob);
  std::vector<double> rr(nprob), tol(nprob);
  std::vector<char> cg_live(nprob), searching(nprob), took_best(nprob);
  for (;;) {
    bool any = false;
    for (int m = 0; m < nprob; m++) any = any || P[m].reason == GLIA_OPT_CONTINUE;
    if (!any) break;
    // ---- CG on the free variables, all live problems in lockstep
    for (int m = 0; m < nprob; m++) {
      Problem& p = P[m];
      cg_live[m] = p.reason == GLIA_OPT_CONTINUE;
      sv[m].assign(n, 0.0);
      rv[m].assign(n, 0.0);
      dv[m].assign(n, 0.0);
      took_best[m] = 0;
      if (!cg_live[m]) continue;
      // held: on a bound with a gradient that does not point into the box
      held[m].assign(n, 0);
      for (int i = 0; i < n; i++) {
        held[m][i] = (p.x[i] <= lo(p, i) && p.g[i] >= 0) || (p.x[i] >= hi(p, i) && p.g[i] <= 0);
        rv[m][i] = held[m][i] ? 0.0 : -p.g[i];
      }
      const double bn = std::sqrt(dot(rv[m], rv[m]));
      if (!(bnorm0[m] > 0)) bnorm0[m] = bn > 0 ? bn : 1.0;
      const double ratio = bn / bnorm0[m];
      const double rel = std::max(1e-10, std::min(0.5, s.forcing == GLIA_OPT_FORCING_QUADRATIC ? ratio : std::sqrt(ratio)));
      tol[m] = rel * bn;
      dv[m] = rv[m];
      rr[m] = dot(rv[m], rv[m]);
      if (!(bn > 0)) cg_live[m] = 0;
    }
    for (int k = 0; k < std::max(1, s.krylov_maxit); k++) {
      bool live = false;
      for (int m = 0; m < nprob; m++) {
        std::copy(dv[m].begin(), dv[m].end(), v.begin() + (size_t)m * n);
        if (!cg_live[m]) std::fill(v.begin() + (size_t)m * n, v.begin() + (size_t)(m + 1) * n, 0.0);
        live = live || cg_live[m];
      }
      if (!live) break;
      if ((rc = hv(nprob, v.data(), y.data()))) return rc;
      for (int m = 0; m < nprob; m++) {
        if (!cg_live[m]) continue;
        std::vector<double> hd(y.begin() + (size_t)m * n, y.begin() + (size_t)(m + 1) * n);
        for (int i = 0; i < n; i++)
          if (held[m][i]) hd[i] = 0.0;
        const double curv = dot(dv[m], hd);
        if (!(curv > 0)) {  // no positive curvature along d: keep what CG has, or the steepest step
          if (k == 0) sv[m] = rv[m];
          cg_live[m] = 0;
          continue;
        }
        const double alpha = rr[m] / curv;
        for (int i = 0; i < n; i++) {
          sv[m][i] += alpha * dv[m][i];
          rv[m][i] -= alpha * hd[i];
        }
        const double rrn = dot(rv[m], rv[m]);
        if (std::sqrt(rrn) <= tol[m]) {
          cg_live[m] = 0;
          continue;
        }
        const double beta = rrn / rr[m];
        rr[m] = rrn;
        for (int i = 0; i < n; i++) dv[m][i] = rv[m][i] + beta * dv[m][i];
      }
    }
    // ---- directions and line searches
    for (int m = 0; m < nprob; m++) {
      Problem& p = P[m];
      searching[m] = 0;
      if (p.reason != GLIA_OPT_CONTINUE) continue;
      p.d = sv[m];
      for (int i = 0; i < n; i++)
        if ((p.x[i] <= lo(p, i) && p.d[i] < 0) || (p.x[i] >= hi(p, i) && p.d[i] > 0)) p.d[i] = 0;
      p.gtd = dot(p.g, p.d);
      if (!(p.gtd < 0)) {
        direction(p, true);  // -delta P(g)
        if (!(p.gtd < 0)) {
          p.reason = GLIA_OPT_DIVERGED_LS_FAILURE;
          continue;
        }
      }
      p.reset_tried = false;
      start_search(p, s);
      searching[m] = 1;
    }
    for (;;) {
      bool any_search = false;
      for (int m = 0; m < nprob; m++) {
        const Problem& p = P[m];
        any_search = any_search || searching[m];
        if (searching[m]) std::copy(p.xt.begin(), p.xt.end(), xe.begin() + (size_t)m * n);
        else std::copy(p.x, p.x + n, xe.begin() + (size_t)m * n);
      }
      if (!any_search) break;
      if ((rc = fg(nprob, xe.data(), fe.data(), ge.data()))) return rc;
      for (int m = 0; m < nprob; m++) {
        if (!searching[m]) continue;
        Problem& p = P[m];
        p.fevals++;
        const double* g = ge.data() + (size_t)m * n;
        int st;
        if (s.line_search == GLIA_OPT_LS_MORE_THUENTE) {
          st = mt_advance(p, s, fe[m], g);
        } else {
          double dec = 0;
          for (int i = 0; i < n; i++) dec += p.g[i] * (p.xt[i] - p.x[i]);
          if (std::isfinite(fe[m]) && fe[m] <= p.f + s.ls_sigma * dec) st = LS_ACCEPT;
          else if (++p.ls_k < s.ls_max_funcs) {
            p.alpha *= 0.5;
            trial(p);
            st = LS_NEXT;
          } else {
            st = LS_FAIL;
          }
        }
        if (st == LS_ACCEPT) {
          accept_gn(p, s, p.xt, fe[m], g, p.alpha);
          searching[m] = 0;
        } else if (st == LS_ACCEPT_BEST) {
          const std::vector<double> xb = p.mt.xb, gb = p.mt.gb;
          accept_gn(p, s, xb, p.mt.fb, gb.data(), p.mt.ab);
          searching[m] = 0;
          took_best[m] = 1;
        } else if (st == LS_FAIL) {
          // one retry along the scaled steepest descent, as the quasi-Newton loop does
          if (!p.reset_tried) {
            p.reset_tried = true;
            direction(p, true);
            if (p.gtd < 0) {
              start_search(p, s);
              continue;
            }
          }
          p.reason = GLIA_OPT_DIVERGED_LS_FAILURE;
          searching[m] = 0;
        }
      }
    }
    // every problem was evaluated at its current point in the last round, so the Hessian's
    // linearization point (the last evaluation) is right for the next CG solve, except after an
    // accepted best point of a halted search: evaluate again there
    bool refresh = false;
    for (int m = 0; m < nprob; m++) refresh = refresh || (P[m].reason == GLIA_OPT_CONTINUE && took_best[m]);
    if (refresh && (rc = fg(nprob, x, fe.data(), ge.data()))) return rc;
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

// the reference's gradnorm0_ for a p-inversion (src/opt/TaoInterface.cpp:285-300): the gradient
// at the zero vector of the unknowns the solve may move (lb < ub) -- entries fixed by their
// bounds keep their values, as rho does outside the TIL vector -- and its norm over those unknowns
int gradient_norm_at_zero(int nprob, int n, const double* x, const double* lb, const double* ub, const FG& fg,
                          std::vector<double>& gref) {
  const size_t N = (size_t)nprob * n;
  auto fixed = [&](size_t k) { return lb && ub && lb[k] >= ub[k]; };
  std::vector<double> x0(N), f(nprob), g(N);
  for (size_t k = 0; k < N; k++) x0[k] = fixed(k) ? x[k] : 0.0;
  const int rc = fg(nprob, x0.data(), f.data(), g.data());
  if (rc) return rc;
  gref.assign(nprob, 0.0);
  for (int m = 0; m < nprob; m++) {
    double s = 0;
    for (int i = 0; i < n; i++) {
      const size_t k = (size_t)m * n + i;
      if (!fixed(k)) s += g[k] * g[k];
    }
    gref[m] = std::sqrt(s);
  }
  return 0;
}

// plain and component-wise hardThreshold (src/utils/Utils.cpp:339-463): a max-priority queue of
// (value, index) pairs -- signed values, ties to the larger index -- popped while |value| > 0
int threshold_simple(const double* x, int n, int s, std::vector<int>& out) {
  std::priority_queue<std::pair<double, int>> q;
  for (int i = 0; i < n; i++) q.push({x[i], i});
  int nnz = 0;
  for (int k = 0; k < s && !q.empty(); k++) {
    if (!(std::fabs(q.top().first) > 0.0)) break;
    out.push_back(q.top().second);
    nnz++;
    q.pop();
  }
  return nnz;
}

int threshold_components(const double* x, int n, int s, const int* labels, const double* w, int ncomp, double thresh,
                         bool double_mode, std::vector<int>& out) {
  if (double_mode) s /= 2;
  int ncc = 0, ncc_t = 0;
  const double reduce = thresh / 10.0;
  for (int c = 0; c < ncomp; c++) {
    if (w[c] >= thresh) ncc++;
    if (w[c] < thresh && w[c] >= reduce) ncc_t++;
  }
  std::vector<int> budget;
  int nnz = 0;
  for (int c = 0; c < ncomp; c++) {
    int sp;
    if (c != ncomp - 1) {
      if (w[c] >= thresh) {
        sp = 3 + (int)std::floor(w[c] * (s - 3 * ncc - ncc_t));
      } else if (w[c] >= reduce) {
        sp = 1;
      } else {
        sp = 0;
      }
      if (double_mode) sp *= 2;
    } else {  // the last component takes the rest of the budget
      int used = 0;
      for (int b : budget) used += b;
      if (double_mode) s *= 2;
      sp = s - used;
    }
    budget.push_back(sp);
    std::priority_queue<std::pair<double, int>> q;
    for (int i = 0; i < n; i++)
      if (labels[i] == c + 1) q.push({x[i], i});
    for (int k = 0; k < sp && !q.empty(); k++) {
      if (!(std::fabs(q.top().first) > 0.0)) break;
      out.push_back(q.top().second);
      nnz++;
      q.pop();
    }
  }
  return nnz;
}

void sort_unique(std::vector<int>& v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

struct TilOps {
  std::function<int(int mode, double beta, const double* x, double* f, double* g)> eval;
  std::function<int(const double* p, double* cmax)> set_c0;
};

// SparseTILOptimizer::solve (src/opt/SparseTILOptimizer.cpp:220-595) for nprob problems in lockstep
int sparse_til(int nprob, int np, double* x, int* support_out, const int* labels, const double* weights, int ncomp,
               const TilOps& ops, const glia_sparse_til_settings& s, glia_sparse_til_result* r) {
  const int n = np + 2;
  const size_t N = (size_t)nprob * n;
  std::vector<std::vector<int>> support(nprob);
  std::vector<int> its(nprob, 0), stop(nprob, 0);
  std::vector<double> J(nprob), Jref(nprob), f(nprob), g(N);
  const bool comps = labels && weights && ncomp > 0;
  // the injected support: nonzero entries of the initial p (a warm start sets p on its support)
  for (int m = 0; m < nprob; m++)
    for (int i = 0; i < np; i++)
      if (x[(size_t)m * n + i] != 0.0) support[m].push_back(i);
  auto til_fg = [&](int nprob_, const double* xe, double* fe, double* ge) {
    (void)nprob_;
    return ops.eval(GLIA_RD_MODE_P_KAPPA, s.beta, xe, fe, ge);
  };
  // corrective L2 solve (TILOptimizer) on each problem's support; problems in 'frozen' keep x
  auto til_solve = [&](const std::vector<int>& frozen, glia_opt_result* res) {
    std::vector<double> lb(N), ub(N);
    for (int m = 0; m < nprob; m++) {
      double* xm = x + (size_t)m * n;
      double* l = lb.data() + (size_t)m * n;
      double* u = ub.data() + (size_t)m * n;
      for (int i = 0; i < n; i++) l[i] = u[i] = xm[i];
      if (frozen[m]) continue;
      for (int i : support[m]) {
        l[i] = 0.0;
        u[i] = std::numeric_limits<double>::infinity();
      }
      if (s.diffusivity_inversion) {
        l[np] = s.k_lb;
        u[np] = s.k_ub;
      }
    }
    std::vector<double> gref;
    int rc = gradient_norm_at_zero(nprob, n, x, lb.data(), ub.data(), til_fg, gref);
    if (rc) return rc;
    return blmvm(nprob, n, x, lb.data(), ub.data(), til_fg, &s.til, gref.data(), res);
  };
  // (1) reference objective and gradient, beta = 0 for thresholding
  int rc = ops.eval(GLIA_RD_MODE_P_KAPPA, 0.0, x, Jref.data(), g.data());
  if (rc) return rc;
  J = Jref;
  std::vector<int> active(nprob, 1), frozen(nprob, 0);
  std::vector<double> tmp(np), x_old(N);
  std::vector<glia_opt_result> res(nprob);
  for (;;) {
    bool any = false;
    for (int m = 0; m < nprob; m++) any = any || active[m];
    if (!any) break;
    for (int m = 0; m < nprob; m++) {
      frozen[m] = !active[m];
      if (!active[m]) continue;
      its[m]++;
      // hard-threshold |g_p| to 2s and merge with the support
      for (int i = 0; i < np; i++) tmp[i] = std::fabs(g[(size_t)m * n + i]);
      std::vector<int> idx;
      if (comps) {
        threshold_components(tmp.data(), np, 2 * s.sparsity, labels + (size_t)m * np, weights + (size_t)m * ncomp,
                             ncomp, s.thresh_component_weight, true, idx);
      } else {
        threshold_simple(tmp.data(), np, 2 * s.sparsity, idx);
      }
      support[m].insert(support[m].end(), idx.begin(), idx.end());
      sort_unique(support[m]);
    }
    std::copy(x, x + N, x_old.begin());
    if ((rc = til_solve(frozen, res.data()))) return rc;
    for (int m = 0; m < nprob; m++) {
      if (!active[m]) continue;
      double* xm = x + (size_t)m * n;
      // prune the solution to s entries; everything off the support is zero
      std::vector<int> idx;
      if (comps && s.prune_components) {
        threshold_components(xm, np, s.sparsity, labels + (size_t)m * np, weights + (size_t)m * ncomp, ncomp,
                             s.thresh_component_weight, false, idx);
      } else {
        threshold_simple(xm, np, s.sparsity, idx);
      }
      support[m] = idx;
      sort_unique(support[m]);
      std::vector<double> keep(np, 0.0);
      for (int i : support[m]) keep[i] = xm[i];
      std::copy(keep.begin(), keep.end(), xm);
    }
    // mismatch and gradient (beta = 0) of the pruned solution; stopping rules of the reference
    std::vector<double> Jold = J;
    if ((rc = ops.eval(GLIA_RD_MODE_P_KAPPA, 0.0, x, f.data(), g.data()))) return rc;
    for (int m = 0; m < nprob; m++) {
      if (!active[m]) continue;
      J[m] = f[m];
      if (its[m] >= s.max_it) {
        stop[m] = 1;
      } else if (std::fabs(J[m]) < 1e-5) {
        stop[m] = 2;
      } else if (std::fabs(Jold[m] - J[m]) < s.ftol * std::fabs(1.0 + Jref[m])) {
        stop[m] = 3;
      }
      if (stop[m]) active[m] = 0;
    }
  }
  // (3) final L2 solve on the support
  std::fill(frozen.begin(), frozen.end(), 0);
  if ((rc = til_solve(frozen, res.data()))) return rc;
  std::vector<glia_opt_result> rd_res(nprob);
  std::memset(rd_res.data(), 0, sizeof(glia_opt_result) * nprob);
  // (4) reaction/diffusion inversion with c0 = Phi p rescaled to max 1 (RDOptimizer::setInitialGuess,
  // src/opt/RDOptimizer.cpp:80-135) and rho started from the best of a scan
  if (s.reaction_inversion) {
    std::vector<double> p((size_t)nprob * np), cmax(nprob);
    for (int m = 0; m < nprob; m++) std::copy(x + (size_t)m * n, x + (size_t)m * n + np, p.begin() + (size_t)m * np);
    if ((rc = ops.set_c0(p.data(), cmax.data()))) return rc;
    for (int m = 0; m < nprob; m++) {
      if (!(cmax[m] > 0)) return GLIA_RD_ERR_ARG;  // c0 = 0: nothing to rescale
      const double scale = 1.0 / cmax[m];
      for (int i = 0; i < np; i++) {
        p[(size_t)m * np + i] *= scale;
        x[(size_t)m * n + i] *= scale;
      }
    }
    if ((rc = ops.set_c0(p.data(), cmax.data()))) return rc;
    static const double scan[7] = {0, 3, 6, 9, 10, 12, 15};
    std::vector<double> best(nprob, 1e300), xs(x, x + N);
    std::vector<double> rho0(nprob, 0.0);
    for (double rho : scan) {
      for (int m = 0; m < nprob; m++) xs[(size_t)m * n + np + 1] = rho;
      if ((rc = ops.eval(GLIA_RD_MODE_KAPPA_RHO, 0.0, xs.data(), f.data(), nullptr))) return rc;
      for (int m = 0; m < nprob; m++)
        if (f[m] < best[m]) {
          best[m] = f[m];
          rho0[m] = rho;
        }
    }
    std::vector<double> lb(N), ub(N);
    for (int m = 0; m < nprob; m++) {
      double* xm = x + (size_t)m * n;
      xm[np + 1] = rho0[m];
      double* l = lb.data() + (size_t)m * n;
      double* u = ub.data() + (size_t)m * n;
      for (int i = 0; i < n; i++) l[i] = u[i] = xm[i];
      l[np] = s.k_lb;
      u[np] = s.k_ub;
      l[np + 1] = s.rho_lb;
      u[np + 1] = s.rho_ub;
    }
    auto rd_fg = [&](int nprob_, const double* xe, double* fe, double* ge) {
      (void)nprob_;
      return ops.eval(GLIA_RD_MODE_KAPPA_RHO, 0.0, xe, fe, ge);
    };
    if ((rc = blmvm(nprob, n, x, lb.data(), ub.data(), rd_fg, &s.rd, nullptr, rd_res.data()))) return rc;
  }
  for (int m = 0; m < nprob; m++) {
    if (support_out) {
      int* so = support_out + (size_t)m * np;
      for (int i = 0; i < np; i++) so[i] = (i < (int)support[m].size()) ? support[m][i] : -1;
    }
    if (r) {
      r[m].iterations = its[m];
      r[m].stop = stop[m];
      r[m].nnz = (int)support[m].size();
      r[m].J = J[m];
      r[m].til = res[m];
      r[m].rd = rd_res[m];
    }
  }
  return 0;
}

struct RdCtx {
  glia_rd_t h;
  int mode;
};

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
  s->line_search = GLIA_OPT_LS_MORE_THUENTE;
  s->ls_gtol = 0.9;
  s->ls_xtol = 1e-10;
  s->gnorm_ref = 0.0;
  s->newton_solver = GLIA_OPT_QUASI_NEWTON;
  s->krylov_maxit = 30;
  s->forcing = GLIA_OPT_FORCING_SUPERLINEAR;
}

int glia_blmvm(int nprob, int n, double* x, const double* lb, const double* ub, glia_fg_fn fg, void* ctx,
               const glia_opt_settings* sp, glia_opt_result* r) {
  if (!fg) return GLIA_RD_ERR_ARG;
  FG f = [&](int np_, const double* xe, double* fe, double* ge) { return fg(ctx, np_, xe, fe, ge); };
  return blmvm(nprob, n, x, lb, ub, f, sp, nullptr, r);
}

int glia_gauss_newton(int nprob, int n, double* x, const double* lb, const double* ub, glia_fg_fn fg, glia_hv_fn hv,
                      void* ctx, const glia_opt_settings* sp, glia_opt_result* r) {
  if (!fg || !hv) return GLIA_RD_ERR_ARG;
  FG f = [&](int np_, const double* xe, double* fe, double* ge) { return fg(ctx, np_, xe, fe, ge); };
  HV h = [&](int np_, const double* v, double* y) { return hv(ctx, np_, v, y); };
  return gauss_newton(nprob, n, x, lb, ub, f, h, sp, nullptr, r);
}

int glia_rd_invert(glia_rd_t h, int mode, double* x, const double* lb, const double* ub, const glia_opt_settings* s,
                   glia_opt_result* r) {
  const int nm = glia_rd_num_members(h);
  if (nm <= 0 || !x || mode < 0 || mode > 2) return GLIA_RD_ERR_ARG;
  const int np = glia_rd_basis_size(h, 0);
  for (int m = 1; m < nm; m++)
    if (glia_rd_basis_size(h, m) != np) return GLIA_RD_ERR_ARG;
  FG fg = [&](int, const double* xe, double* fe, double* ge) {
    return glia_rd_objective_and_gradient(h, mode, xe, fe, ge);
  };
  const int n = np + 2;
  if (mode == GLIA_RD_MODE_KAPPA_RHO) {
    if (s && s->newton_solver == GLIA_OPT_GAUSS_NEWTON) return GLIA_RD_ERR_UNSUPPORTED;
    // RDOptimizer::solve (src/opt/RDOptimizer.cpp:170-250): beta = 0 during the solve, reference
    // gradient = the first one
    double beta = 0;
    int rc = glia_rd_get_beta(h, &beta);
    if (rc || (rc = glia_rd_set_beta(h, 0.0))) return rc;
    rc = blmvm(nm, n, x, lb, ub, fg, s, nullptr, r);
    const int rc2 = glia_rd_set_beta(h, beta);
    return rc ? rc : rc2;
  }
  // TILOptimizer: reference gradient at x = 0
  std::vector<double> gref;
  int rc = gradient_norm_at_zero(nm, n, x, lb, ub, fg, gref);
  if (rc) return rc;
  if (s && s->newton_solver == GLIA_OPT_GAUSS_NEWTON) {
    // the Gauss-Newton Hessian with its kappa blocks (DerivativeOperatorsRD::evaluateHessian);
    // rho is not an unknown of the p modes' Hessian, so mode 2 holds it
    if (mode == GLIA_RD_MODE_ALL) return GLIA_RD_ERR_UNSUPPORTED;
    int diffusivity = 0;  // the kappa blocks only where some member's kappa is free
    for (int m = 0; m < nm; m++) {
      const size_t k = (size_t)m * n + np;
      diffusivity |= !(lb && ub && lb[k] >= ub[k]);
    }
    HV hv = [&](int, const double* v, double* y) { return glia_rd_hessian(h, diffusivity, v, y); };
    return gauss_newton(nm, n, x, lb, ub, fg, hv, s, gref.data(), r);
  }
  return blmvm(nm, n, x, lb, ub, fg, s, gref.data(), r);
}

void glia_sparse_til_settings_init(glia_sparse_til_settings* s) {
  if (!s) return;
  s->sparsity = 5;
  s->max_it = 5;
  s->ftol = 1e-5;
  s->prune_components = 1;
  s->thresh_component_weight = 1e-3;
  s->diffusivity_inversion = 1;
  s->reaction_inversion = 0;
  s->beta = 0.0;
  s->k_lb = 1e-3;
  s->k_ub = 1.0;
  s->rho_lb = 1.0;
  s->rho_ub = 15.0;
  glia_opt_settings_init(&s->til);
  glia_opt_settings_init(&s->rd);
}

int glia_hard_threshold(const double* x, int n, int s, int* support) {
  if (!x || n < 0 || s < 0 || (!support && s > 0)) return GLIA_RD_ERR_ARG;
  std::vector<int> out;
  const int nnz = threshold_simple(x, n, s, out);
  std::copy(out.begin(), out.end(), support);
  return nnz;
}

int glia_hard_threshold_components(const double* x, int n, int s, const int* labels, const double* weights,
                                   int ncomp, double thresh, int double_mode, int* support) {
  if (!x || !labels || !weights || n < 0 || s < 0 || ncomp <= 0 || !support) return GLIA_RD_ERR_ARG;
  std::vector<int> out;
  const int nnz = threshold_components(x, n, s, labels, weights, ncomp, thresh, double_mode != 0, out);
  std::copy(out.begin(), out.end(), support);
  return nnz;
}

int glia_sparse_til(int nprob, int np, double* x, int* support, const int* labels, const double* weights, int ncomp,
                    glia_eval_fn eval, glia_c0_fn set_c0, void* ctx, const glia_sparse_til_settings* sp,
                    glia_sparse_til_result* r) {
  if (nprob <= 0 || np <= 0 || !x || !eval) return GLIA_RD_ERR_ARG;
  glia_sparse_til_settings s;
  if (sp) {
    s = *sp;
  } else {
    glia_sparse_til_settings_init(&s);
  }
  if (s.reaction_inversion && !set_c0) return GLIA_RD_ERR_ARG;
  TilOps ops;
  ops.eval = [&](int mode, double beta, const double* xe, double* f, double* g) {
    return eval(ctx, mode, beta, xe, f, g);
  };
  ops.set_c0 = [&](const double* p, double* cmax) { return set_c0(ctx, p, cmax); };
  return sparse_til(nprob, np, x, support, labels, weights, ncomp, ops, s, r);
}

int glia_rd_sparse_til(glia_rd_t h, double* x, int* support, const int* labels, const double* weights, int ncomp,
                       const glia_sparse_til_settings* sp, glia_sparse_til_result* r) {
  const int nm = glia_rd_num_members(h);
  if (nm <= 0 || !x) return GLIA_RD_ERR_ARG;
  const int np = glia_rd_basis_size(h, 0);
  for (int m = 1; m < nm; m++)
    if (glia_rd_basis_size(h, m) != np) return GLIA_RD_ERR_ARG;
  glia_sparse_til_settings s;
  if (sp) {
    s = *sp;
  } else {
    glia_sparse_til_settings_init(&s);
  }
  int grid[3];
  int rc = glia_rd_get_grid(h, grid);
  if (rc) return rc;
  const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
  double beta0 = 0;
  if ((rc = glia_rd_get_beta(h, &beta0))) return rc;
  // the TIL stages run with the given rho (tu_->rho_)
  for (int m = 0; m < nm; m++)
    if ((rc = glia_rd_set_coefficients(h, m, x[(size_t)m * (np + 2) + np], x[(size_t)m * (np + 2) + np + 1])))
      return rc;
  std::vector<float> c0;
  TilOps ops;
  ops.eval = [&](int mode, double beta, const double* xe, double* f, double* g) {
    int e = glia_rd_set_beta(h, beta);
    if (e) return e;
    return g ? glia_rd_objective_and_gradient(h, mode, xe, f, g) : glia_rd_objective(h, mode, xe, f);
  };
  ops.set_c0 = [&](const double* p, double* cmax) {
    int e = glia_rd_phi_apply(h, p);
    if (e) return e;
    c0.resize(ng);
    for (int m = 0; m < nm; m++) {
      if ((e = glia_rd_get_field(h, m, GLIA_RD_FIELD_C0, c0.data(), GLIA_RD_HOST))) return e;
      cmax[m] = *std::max_element(c0.begin(), c0.end());
    }
    return 0;
  };
  rc = sparse_til(nm, np, x, support, labels, weights, ncomp, ops, s, r);
  const int rc2 = glia_rd_set_beta(h, beta0);
  return rc ? rc : rc2;
}

}  // extern "C"
