"""The inversion loop (include/glia_opt.h) on the CPU: closed-form bounded problems, the
Rosenbrock function, independence of batched problems, the reference's stopping reasons, and a
kappa/rho inversion driven by the FP64 oracle's gradient that recovers the truth.

There is no PETSc/TAO here: the loop restates BLMVM (bound-constrained L-BFGS) with the reference's
stopping test; its iterates are not compared with TAO's ("parity unpinned" against TAO)."""
import numpy as np
import pytest

from glia_amd import opt
from opt_cases import SPARSE_RD, SPARSE_STARTS, oracle_ops, rd_sparse_problem

A = np.array([1.0, 10.0, 100.0])


def quad(c):
    def fg(X):
        return 0.5 * np.sum(A * (X - c) ** 2, axis=1), A * (X - c)
    return fg


def rosen(X):
    x, y = X[:, 0], X[:, 1]
    f = (1 - x) ** 2 + 100 * (y - x * x) ** 2
    g = np.stack([-2 * (1 - x) - 400 * x * (y - x * x), 200 * (y - x * x)], axis=1)
    return f, g


def test_bounded_quadratic_solution_is_the_clipped_minimizer():
    c = np.array([[2.0, -3.0, 0.5]])
    x, r = opt.minimize(quad(c), np.zeros(3), lb=-1.0, ub=1.0, gatol=1e-10, gttol=1e-14, max_it=100)
    assert np.allclose(x, np.clip(c, -1, 1), atol=1e-9)
    # the projected gradient vanishes at the bounded optimum (the full gradient does not)
    assert r[0]["gnorm"] < 1e-10 and r[0]["reason"] in ("gttol", "gatol")


def test_unbounded_quadratic_converges_in_few_iterations():
    c = np.array([[0.3, -0.2, 0.7]])
    x, r = opt.minimize(quad(c), np.zeros(3), gatol=1e-12, gttol=1e-14, max_it=100)
    assert np.allclose(x, c, atol=1e-10)
    assert r[0]["iterations"] <= 25


def test_rosenbrock_from_two_starts():
    x, r = opt.minimize(rosen, [[-1.2, 1.0], [2.0, 2.0]], gatol=1e-8, gttol=1e-12, max_it=500)
    assert np.allclose(x, 1.0, atol=1e-6), x
    assert all(m["reason"] in ("gatol", "gttol") for m in r)


def test_bounded_rosenbrock_active_upper_bound():
    # min over x <= 0.5: x = 0.5 on the bound, y = x^2
    x, r = opt.minimize(rosen, [[-1.2, 1.0]], lb=[-2, -2], ub=[0.5, 2], gatol=1e-8, gttol=1e-12, max_it=500)
    assert np.allclose(x, [[0.5, 0.25]], atol=1e-7), x
    assert r[0]["f"] == pytest.approx(0.25, abs=1e-12)


def test_batched_problems_are_independent():
    starts = np.array([[-1.2, 1.0], [2.0, 2.0], [0.0, 0.0]])
    xb, rb = opt.minimize(rosen, starts, gatol=1e-8, gttol=1e-12, max_it=500)
    for m in range(3):
        x1, r1 = opt.minimize(rosen, starts[m], gatol=1e-8, gttol=1e-12, max_it=500)
        assert np.array_equal(x1[0], xb[m])
        assert r1[0] == rb[m]


def test_iteration_limit_and_min_iterations():
    _, r = opt.minimize(rosen, [[-1.2, 1.0]], gatol=0.0, gttol=0.0, max_it=3)
    assert r[0]["reason"] == "maxits" and r[0]["iterations"] == 4  # stops once it > max_it
    # at the optimum, ||g|| < gatol stops before any step even with min_it = 1
    _, r = opt.minimize(rosen, [[1.0, 1.0]], max_it=3)
    assert r[0]["reason"] == "gatol" and r[0]["iterations"] == 0
    # a relative test needs min_it accepted steps first
    c = np.array([[0.3, -0.2, 0.7]])
    _, r = opt.minimize(quad(c), np.zeros(3), gatol=0.0, gttol=0.5, min_it=3, max_it=50)
    assert r[0]["iterations"] >= 3 and r[0]["reason"] == "gttol"


def test_nan_objective_stops_with_reason_nan():
    def fg(X):
        return np.full(len(X), np.nan), np.zeros_like(X)
    _, r = opt.minimize(fg, [[0.0, 0.0]])
    assert r[0]["reason"] == "nan"


def test_initial_guess_is_projected_and_errors_propagate():
    seen = []

    def fg(X):
        seen.append(X.copy())
        return quad(np.array([[0.0, 0.0, 0.0]]))(X)
    opt.minimize(fg, [[5.0, -5.0, 0.0]], lb=-1.0, ub=1.0, max_it=2)
    assert np.array_equal(seen[0], [[1.0, -1.0, 0.0]])

    def bad(X):
        raise ValueError("boom")
    with pytest.raises(ValueError):
        opt.minimize(bad, [[0.0]])


def test_oracle_kappa_rho_inversion_recovers_truth(oracle_mod):
    """RDOptimizer's kappa/rho inversion (src/opt/RDOptimizer.cpp) on exact data of a 32^3 problem:
    from (kappa, rho) = (0.1, 5) with the reference's bounds [1e-3, 1] x [1, 15] the loop
    reaches the truth (0.05, 8)."""
    from glia_amd import synth
    n, nt, dt = 32, 8, 0.05
    pr = synth.problem(n, 5, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 32)
    o = oracle_mod.Oracle(n)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_solver(1e-6, 5000)
    o.set_time(dt, nt)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_c0(o.phi_apply(pr["p"]))
    o.solve_state(0)
    o.set_data(o.ct(), None, -0.99)
    o.set_time(dt, nt)
    J = []

    def fg(X):
        f, g = o.objective_and_gradient(X[0], mode=1)
        J.append(f)
        return np.array([f]), g[None]
    x0 = np.concatenate([pr["p"], [0.1, 5.0]])
    lb = np.concatenate([np.full(6, -np.inf), [1e-3, 1.0]])
    ub = np.concatenate([np.full(6, np.inf), [1.0, 15.0]])
    x, r = opt.minimize(fg, x0, lb, ub, gatol=1e-9, gttol=1e-5, max_it=50)
    assert r[0]["reason"] == "gttol", r
    assert abs(x[0][-2] - 0.05) < 1e-4 and abs(x[0][-1] - 8.0) < 1e-3, x[0][-2:]
    assert np.array_equal(x[0][:-2], pr["p"])  # p is not an unknown of this mode
    assert J[-1] < 1e-8 * J[0]


# ------------------------------------------------------------------ hard thresholds ----
def test_hard_threshold_signed_order_ties_and_zero_stop():
    # src/utils/Utils.cpp:339-367: a max-heap of (value, index): signed values, the larger index
    # first on ties, and the scan ends at the first zero (so negative entries are never reached)
    x = [0.1, 0.5, 0.5, -2.0, 0.0, 0.3]
    assert opt.hard_threshold(x, 3) == [2, 1, 5]
    assert opt.hard_threshold(x, 10) == [2, 1, 5, 0]
    assert opt.hard_threshold([0.0, 0.0], 2) == []


def test_hard_threshold_matches_brute_force_on_positive_vectors():
    rng = np.random.default_rng(4)
    for _ in range(20):
        x = np.round(rng.uniform(0, 1, 30), 1)  # rounding makes ties
        x[rng.integers(0, 30, 5)] = 0.0
        s = int(rng.integers(1, 12))
        order = sorted(range(30), key=lambda i: (x[i], i), reverse=True)
        want = [i for i in order if x[i] > 0][:s]
        assert opt.hard_threshold(x, s) == want


def test_hard_threshold_components_budgets():
    # src/utils/Utils.cpp:370-463: weight >= thresh -> 3 + floor(w (s - 3 ncc - ncc_t)),
    # thresh/10 <= w < thresh -> 1, the last component takes the rest
    x = [0.9, 0.5, 0.4, 0.2, 0.8, 0.3, 0.7]
    lab = [1, 1, 1, 2, 2, 2, 2]
    # s = 5, ncc = 2: component 1 gets 3 + floor(0.6 * (5 - 6)) = 2, component 2 the other 3
    assert opt.hard_threshold(x, 5, lab, [0.6, 0.4]) == [0, 1, 4, 6, 5]
    # s = 8, weights (0.5, 5e-4, 0.4995): ncc = 2, ncc_t = 1 -> budgets 3 + floor(0.5 * 1) = 3, 1, 4
    x = np.arange(1, 13, dtype=float)
    lab = [1] * 4 + [2] * 4 + [3] * 4
    got = opt.hard_threshold(x, 8, lab, [0.5, 5e-4, 0.4995])
    assert got == [3, 2, 1, 7, 11, 10, 9, 8]
    # double mode (the CoSaMP gradient step): s = 10 is split as s = 5 and every share doubled:
    # weights (0.7, 0.2, 0.1): 3 + floor(0.7 * (5 - 9)) = 0, 3 + floor(0.2 * (-4)) = 2 -> 4, rest 6
    got = opt.hard_threshold(x, 10, lab, [0.7, 0.2, 0.1], double_mode=True)
    assert got == [7, 6, 5, 4, 11, 10, 9, 8]


# ------------------------------------------------------------------------ sparse TIL ----
def linear_model(A, b, kappa_star=0.3, rho_star=7.0):
    """A stand-in model with the shape of the RD operations: J = 1/2 |A p - b|^2 + 1/2 (kappa -
    kappa*)^2 (+ 1/2 (rho - rho*)^2 in the kappa/rho mode), c0 = A p."""
    np_ = A.shape[1]

    def ev(mode, beta, X, want):
        f, G = np.zeros(len(X)), np.zeros_like(X)
        for m, x in enumerate(X):
            p, k, r = x[:np_], x[np_], x[np_ + 1]
            res = A @ p - b[m]
            f[m] = 0.5 * res @ res + 0.5 * (k - kappa_star) ** 2 + 0.5 * beta * p @ p
            G[m, np_] = k - kappa_star
            if mode != 1:
                G[m, :np_] = A.T @ res + beta * p
            if mode != 0:
                f[m] += 0.5 * (r - rho_star) ** 2
                G[m, np_ + 1] = r - rho_star
        return f, (G if want else None)

    def c0(P):
        return np.array([(A @ p).max() for p in P])
    return ev, c0


def test_cosamp_recovers_sparse_vector_from_compressed_measurements():
    """CoSaMP's defining property: a 3-sparse nonnegative p* in R^40 is recovered from 20 random
    Gaussian measurements, with its exact support."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((20, 40)) / np.sqrt(20)
    supports, ps = [[3, 17, 31], [5, 6, 38]], []
    for s in supports:
        p = np.zeros(40)
        p[s] = rng.uniform(0.5, 1.5, 3)
        ps.append(p)
    b = np.array([A @ p for p in ps])
    ev, c0 = linear_model(A, b)
    x0 = np.tile(np.concatenate([np.zeros(40), [0.5, 7.0]]), (2, 1))
    kw = dict(sparsity=3, max_it=10, til=dict(max_it=200, gatol=1e-12, gttol=1e-12))
    x, sup, res = opt.sparse_til(ev, c0, x0, 40, **kw)
    for m in range(2):
        assert sup[m] == supports[m], (sup, res)
        assert np.allclose(x[m][:40], ps[m], atol=1e-6)
        assert abs(x[m][40] - 0.3) < 1e-6
    # batched problems in lockstep give what separate runs give
    for m in range(2):
        ev1, c01 = linear_model(A, b[m:m + 1])
        x1, sup1, res1 = opt.sparse_til(ev1, c01, x0[m:m + 1], 40, **kw)
        assert sup1[0] == sup[m] and np.array_equal(x1[0], x[m]) and res1[0] == res[m]


def test_cosamp_final_reaction_diffusion_stage_rescales_and_scans_rho():
    """With reaction_inversion: p is rescaled so that max c0 = 1, rho starts from the best of
    the scan {0, 3, 6, 9, 10, 12, 15}, then kappa/rho are solved within their bounds."""
    rng = np.random.default_rng(12)
    A = np.abs(rng.standard_normal((20, 40))) / np.sqrt(20)
    p = np.zeros(40)
    p[[2, 9]] = [1.0, 0.5]
    b = (A @ p)[None]
    ev, c0 = linear_model(A, b, rho_star=11.0)
    x0 = np.concatenate([np.zeros(40), [0.5, 4.0]])[None]
    x, sup, res = opt.sparse_til(ev, c0, x0, 40, sparsity=2, reaction_inversion=1,
                                 til=dict(max_it=200, gatol=1e-12, gttol=1e-12),
                                 rd=dict(max_it=50, gatol=1e-10, gttol=1e-12))
    assert sup[0] == [2, 9]
    assert (A @ x[0][:40]).max() == pytest.approx(1.0, rel=1e-12)
    assert np.allclose(x[0][:40], p / (A @ p).max(), rtol=1e-6)
    assert abs(x[0][41] - 11.0) < 1e-6 and abs(x[0][40] - 0.3) < 1e-6


def test_oracle_sparse_til_recovers_support_and_coefficients(oracle_mod):
    """SparseTILOptimizer on exact data: from p = 0, kappa = 0.1 the CoSaMP loop finds the two
    active bases, and the final kappa/rho solve (rho from the scan) recovers (0.05, 8)."""
    pr, o, p, _ = rd_sparse_problem(oracle_mod)
    ev, c0 = oracle_ops(o)
    x0 = np.concatenate([np.zeros(12), SPARSE_STARTS[0]])
    x, sup, res = opt.sparse_til(ev, c0, x0, 12, **SPARSE_RD)
    assert sup[0] == sorted(np.flatnonzero(p).tolist()), (sup, res)
    assert np.allclose(x[0][:12], p, atol=3e-3)
    assert abs(x[0][12] - 0.05) < 5e-4 and abs(x[0][13] - 8.0) < 0.05, x[0][12:]


# ------------------------------------------------------------------ More-Thuente search ----
def quartic(X):
    return np.sum(X ** 4 / 4 + X ** 2 / 2, axis=1), X ** 3 + X


def offset_quadratic(x0):
    """f = |x|^2 / 2 - 0.99 |x0|^2 / 2: at x0, |f| is small against |g|^2, so the first step of
    length delta = 2 |f0| / |g0|^2 is 100 times too short and breaks the curvature condition."""
    c = 0.99 * 0.5 * float(x0 @ x0)

    def fg(X):
        return 0.5 * np.sum(X * X, axis=1) - c, X.copy()
    return fg


def first_step(fg, x0, ls):
    f0, g0 = fg(x0[None])
    d = -2 * abs(f0[0]) / (g0[0] @ g0[0]) * g0[0]
    x1, r = opt.minimize(fg, x0[None], max_it=0, min_it=0, gatol=0.0, gttol=0.0, line_search=ls)
    assert r[0]["iterations"] == 1 and r[0]["reason"] == "maxits"
    step = x1[0] - x0
    a = step @ d / (d @ d)
    assert np.allclose(step, a * d, rtol=0, atol=1e-12)  # along d
    f1, g1 = fg(x1)
    return a, f0[0], g0[0] @ d, f1[0], g1[0] @ d


def test_more_thuente_first_step_satisfies_strong_wolfe():
    """With max_it = 0 the loop stops after one accepted step along the first direction
    d = -delta g0.  More-Thuente's step satisfies phi(a) <= phi(0) + 1e-4 a phi'(0) and
    |phi'(a)| <= 0.9 |phi'(0)| (More and Thuente, 1994); the Armijo step, accepted at a = 1, does
    not meet the curvature condition here."""
    for x0 in ([1.0, 1.0], [3.0, -0.5, 2.0]):
        x0 = np.array(x0)
        fg = offset_quadratic(x0)
        a, f0, dphi0, f1, dphi1 = first_step(fg, x0, 1)
        assert f1 <= f0 + 1e-4 * a * dphi0
        assert abs(dphi1) <= 0.9 * abs(dphi0)
        a, f0, dphi0, f1, dphi1 = first_step(fg, x0, 0)
        assert abs(a - 1.0) < 1e-9 and abs(dphi1) > 0.9 * abs(dphi0)
    for x0 in ([2.0, -1.5], [3.0, 0.5, -2.0]):  # a well-scaled first step is kept as it is
        a, f0, dphi0, f1, dphi1 = first_step(quartic, np.array(x0), 1)
        assert f1 <= f0 + 1e-4 * a * dphi0 and abs(dphi1) <= 0.9 * abs(dphi0)


def test_more_thuente_and_armijo_both_reach_the_minimum():
    for ls in (0, 1):
        x, r = opt.minimize(rosen, [[-1.2, 1.0], [2.0, 2.0]], gatol=1e-8, gttol=1e-12, max_it=500, line_search=ls)
        assert np.allclose(x, 1.0, atol=1e-6), (ls, x)
        x, r = opt.minimize(rosen, [[-1.2, 1.0]], lb=[-2, -2], ub=[0.5, 2], gatol=1e-8, gttol=1e-12, max_it=500,
                            line_search=ls)
        assert np.allclose(x, [[0.5, 0.25]], atol=1e-7), (ls, x)


def test_more_thuente_defaults_match_the_reference():
    s = opt.settings()
    assert s.line_search == 1 and s.ls_gtol == 0.9 and s.ls_max_funcs == 10 and s.ls_sigma == 1e-4


# ------------------------------------------------------------------ Gauss-Newton-Krylov ----
def lsq_rosen(X):
    """Rosenbrock as least squares, r = [10 (y - x^2), 1 - x], f = |r|^2: the Gauss-Newton
    Hessian 2 J^T J is exact at the zero-residual minimum (1, 1)."""
    x, y = X[:, 0], X[:, 1]
    r = np.stack([10 * (y - x * x), 1 - x], axis=1)
    J = np.zeros((len(x), 2, 2))
    J[:, 0, 0], J[:, 0, 1], J[:, 1, 0] = -20 * x, 10.0, -1.0
    return r, J


def lsq_fg_hv():
    state = {}

    def fg(X):
        r, J = lsq_rosen(X)
        state["J"] = J
        return np.sum(r * r, axis=1), 2 * np.einsum("mij,mi->mj", J, r)

    def hv(V):
        J = state["J"]
        return 2 * np.einsum("mij,mik,mk->mj", J, J, V)
    return fg, hv


def test_gauss_newton_one_dimensional_problems_stop_after_the_exact_step():
    """In one unknown the first conjugate-gradient step solves H s = -g exactly, and the full
    step of the line search is the minimizer: the first iterate is c (clipped to the bounds)."""
    a = np.array([[0.5], [3.0], [40.0]])
    c = np.array([[2.0], [-1.5], [0.25]])
    fg = lambda X: (0.5 * np.sum(a * (X - c) ** 2, axis=1), a * (X - c))
    hv = lambda V: a * V
    x, r = opt.gauss_newton(fg, hv, np.zeros((3, 1)), max_it=0, min_it=0, gatol=0.0, gttol=0.0)
    assert np.allclose(x, c, rtol=1e-14, atol=1e-15)
    assert all(q["iterations"] == 1 for q in r)
    x, r = opt.gauss_newton(fg, hv, np.zeros((3, 1)), lb=[-1.0], ub=[1.0], max_it=0, min_it=0, gatol=0.0, gttol=0.0)
    assert np.allclose(x, np.clip(c, -1.0, 1.0), rtol=1e-14, atol=1e-15)


def test_gauss_newton_single_krylov_step_is_the_cauchy_point():
    """krylov_maxit = 1 leaves one conjugate-gradient step, s = -(g.g / g.A g) g, the exact
    minimizer of a quadratic along -g, which the line search takes at unit length."""
    Q = np.array([[4.0, 1.0, 0.0], [1.0, 3.0, 0.5], [0.0, 0.5, 2.0]])
    b = np.array([1.0, -2.0, 0.5])
    fg = lambda X: (0.5 * np.einsum("mi,ij,mj->m", X, Q, X) - X @ b, X @ Q - b)
    hv = lambda V: V @ Q
    x0 = np.array([1.0, 1.0, -1.0])
    g = Q @ x0 - b
    want = x0 - (g @ g) / (g @ Q @ g) * g
    x, r = opt.gauss_newton(fg, hv, x0, max_it=0, min_it=0, gatol=0.0, gttol=0.0, krylov_maxit=1)
    assert r[0]["iterations"] == 1
    assert np.allclose(x[0], want, rtol=1e-13, atol=1e-14)


def test_gauss_newton_least_squares_rosenbrock():
    """Unbounded from two starts (batched) to (1, 1), and with x <= 0.5 to (0.5, 0.25); the
    inexact Newton steps (forcing term min(0.5, sqrt(|g| / |g0|))) take fewer iterations than
    the quasi-Newton loop from the same start."""
    fg, hv = lsq_fg_hv()
    starts = [[-1.2, 1.0], [2.0, 2.0]]
    x, r = opt.gauss_newton(fg, hv, starts, gatol=1e-10, gttol=1e-14, max_it=200)
    assert np.allclose(x, 1.0, atol=1e-7), x
    assert all(q["reason"] in ("gatol", "gttol") for q in r), r
    _, rq = opt.minimize(lambda X: fg(X), starts, gatol=1e-10, gttol=1e-14, max_it=500)
    assert r[1]["iterations"] < rq[1]["iterations"], (r, rq)
    x, r = opt.gauss_newton(fg, hv, starts[0], lb=[-2, -2], ub=[0.5, 2], gatol=1e-10, gttol=1e-14, max_it=200)
    assert np.allclose(x, [[0.5, 0.25]], atol=1e-8), x


def test_gauss_newton_settings_follow_the_reference_numbering():
    """GAUSSNEWTON = 0, QUASINEWTON = 1 and QDFS = 0, SLFS = 1 (include/utils/TypeDefs.h:10-12);
    OptimizerSettings defaults: quasi-Newton, 30 Krylov iterations, superlinear forcing."""
    s = opt.settings()
    assert s.newton_solver == 1 and s.krylov_maxit == 30 and s.forcing == 1
    fg, hv = lsq_fg_hv()
    for forcing in (0, 1):
        x, r = opt.gauss_newton(fg, hv, [[-1.2, 1.0]], gatol=1e-10, gttol=1e-14, max_it=200, forcing=forcing)
        assert np.allclose(x, 1.0, atol=1e-7), (forcing, x)


def test_gauss_newton_callback_errors_propagate():
    fg, hv = lsq_fg_hv()

    def bad(V):
        raise ValueError("hv")
    with pytest.raises(ValueError):
        opt.gauss_newton(fg, bad, [[-1.2, 1.0]])


def test_oracle_til_gauss_newton_recovers_the_initial_condition(oracle_mod):
    """TILOptimizer with newton_solver = GAUSSNEWTON (src/opt/TILOptimizer.cpp) on exact data of a
    32^3 problem: the Gauss-Newton Hessian matvec of the FP64 oracle (evaluateHessian, kappa
    fixed) drives the loop from p = 0 to the truth within a few Newton steps."""
    from glia_amd import synth
    n, nt, dt = 32, 4, 0.05
    pr = synth.problem(n, 5, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 32)
    o = oracle_mod.Oracle(n)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_solver(1e-6, 5000)
    o.set_time(dt, nt)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_c0(o.phi_apply(pr["p"]))
    o.solve_state(0)
    o.set_data(o.ct(), None, -0.99)
    o.set_time(dt, nt)
    o.set_opt(beta=1e-6)

    def fg(X):
        f, g = o.objective_and_gradient(X[0], mode=0)
        return np.array([f]), g[None]

    def hv(V):
        return o.hessian(V[0], False)[None]
    x0 = np.concatenate([np.zeros(6), [0.05, 8.0]])
    fixed = np.concatenate([np.full(6, -np.inf), [0.05, 8.0]])
    x, r = opt.gauss_newton(fg, hv, x0, fixed, np.concatenate([np.full(6, np.inf), [0.05, 8.0]]),
                            gatol=1e-9, gttol=1e-5, max_it=30)
    assert r[0]["reason"] == "gttol" and r[0]["iterations"] <= 8, r
    assert np.allclose(x[0][:6], pr["p"], atol=5e-5), x[0][:6] - pr["p"]
    assert x[0][6] == 0.05 and x[0][7] == 8.0


def test_optimizer_keys_of_a_reference_config_file(tmp_path):
    """setParameter (src/utils/IO.cpp:82-100, 171-172) with the optimizer block of the reference's
    config/test_sparsetil_config.txt; absent keys keep the OptimizerSettings defaults."""
    p = tmp_path / "c.txt"
    p.write_text("# optimizer\nnewton_solver=QN\nline_search=armijo\nopttol_grad=1e-05\nnewton_maxit=10\n"
                 "krylov_maxit=1\ngist_maxit=2\nkappa_lb=0\nkappa_ub=0.1\nrho_lb=0\nrho_ub=15\n"
                 "lbfgs_vectors=50\nls_max_func_evals=10\nsparsity_level=5\nbeta_p=1\ninvert_diff=1\n"
                 "invert_reac=1\n")
    st, sp = opt.parse_config(p)
    for o in (st, sp.til, sp.rd):
        assert (o.newton_solver, o.line_search, o.gttol, o.max_it, o.krylov_maxit) == (1, 0, 1e-5, 10, 1)
        assert (o.lbfgs_vectors, o.ls_max_funcs, o.gatol) == (50, 10, 1e-4)
    assert (sp.max_it, sp.k_lb, sp.k_ub, sp.rho_lb, sp.rho_ub, sp.sparsity, sp.beta) == (2, 0.0, 0.1, 0.0, 15.0, 5, 1.0)
    assert sp.diffusivity_inversion == 1 and sp.reaction_inversion == 1 and sp.thresh_component_weight == 1e-3
    p.write_text("newton_solver = GN  # Gauss-Newton\nline_search=mt\n")
    st, sp = opt.parse_config(p)
    assert st.newton_solver == 0 and st.line_search == 1 and sp.til.newton_solver == 0
    p.write_text("newton_maxit=ten\n")
    with pytest.raises(OSError):
        opt.parse_config(p)
    with pytest.raises(OSError):
        opt.parse_config(tmp_path / "missing.txt")
