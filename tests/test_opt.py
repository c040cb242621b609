"""The inversion loop (include/glia_opt.h) on the CPU: closed-form bounded problems, the
Rosenbrock function, independence of batched problems, the reference's stopping reasons, and a
kappa/rho inversion driven by the FP64 oracle's gradient that recovers the truth.

There is no PETSc/TAO here: the loop restates BLMVM (bound-constrained L-BFGS) with the reference's
stopping test; its iterates are not compared with TAO's ("parity unpinned" against TAO)."""
import numpy as np
import pytest

from glia_amd import opt

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
