"""The oracle's linear elasticity (VariableLinearElasticitySolver, src/pde/ElasticitySolver.cpp:6-533)
against what the mathematics fixes: the operator is symmetric; with constant coefficients a plane
wave gets the Navier-Lame symbol (-(2 mu + lam) k^2 - screen longitudinally, -mu k^2 - screen
transversally), the preconditioner is the operator's exact inverse and CG stops after one step;
the material rule gives the hand-computed Lame values."""
import numpy as np

from el_cases import tissue_and_tumor, tumor_force


def test_operator_is_symmetric_and_has_the_navier_lame_symbol(oracle_mod):
    n = 32
    o = oracle_mod.Oracle(n)
    X = np.arange(n) * 2 * np.pi / n
    prof = (1 + 0.3 * np.sin(X))[:, None, None] * (1 + 0.2 * np.cos(2 * X))[None, :, None] * np.ones((1, 1, n))
    mu, lam, scr = 500 * prof, 800 * prof, 0.5 * prof
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal((3, n, n, n)), rng.standard_normal((3, n, n, n))
    a, b = (o.el_matvec(mu, lam, scr, u) * v).sum(), (u * o.el_matvec(mu, lam, scr, v)).sum()
    assert abs(a - b) < 1e-12 * abs(a)
    one = np.ones((n, n, n))
    m0, l0, s0, k = 500.0, 800.0, 0.5, 3
    for comp, symbol in ((0, -(2 * m0 + l0) * k * k - s0), (1, -m0 * k * k - s0)):
        w = np.zeros((3, n, n, n))
        w[comp] = np.cos(k * X)[:, None, None]
        f = o.el_matvec(m0 * one, l0 * one, s0 * one, w)
        assert np.abs(f[comp] - symbol * w[comp]).max() < 1e-9 * abs(symbol)
        assert np.abs(np.delete(f, comp, axis=0)).max() < 1e-9 * abs(symbol)


def test_preconditioner_inverts_the_constant_operator(oracle_mod):
    n = 32
    o = oracle_mod.Oracle(n)
    one = np.ones((n, n, n))
    m0, l0, s0 = 300.0, 900.0, 2.0
    rng = np.random.default_rng(1)
    f = rng.standard_normal((3, n, n, n))
    z = o.el_prec([m0, l0, s0], f)
    assert np.abs(o.el_matvec(m0 * one, l0 * one, s0 * one, z) - f).max() < 1e-10 * np.abs(f).max()
    u, its, reason, rn = o.el_solve(m0 * one, l0 * one, s0 * one, [m0, l0, s0], f, np.zeros((3, n, n, n)))
    assert its == 1 and reason == 2


def test_material_rule_and_variable_solve(oracle_mod):
    """computeMaterialProperties by hand at single voxels, and a variable-coefficient solve whose
    residual meets the preconditioned tolerance."""
    n = 32
    o = oracle_mod.Oracle(n)
    c, wm, gm, csf, vt, bg = tissue_and_tumor(n)
    mu, lam, scr, avg = o.el_material(c, wm, gm, csf, vt, bg)
    E = dict(h=(2100, 0.4), t=(8000, 0.45), b=(15000, 0.45), c=(100, 0.1))
    m = {k: e / (2 * (1 + nu)) for k, (e, nu) in E.items()}
    la = {k: nu * e / ((1 + nu) * (1 - 2 * nu)) for k, (e, nu) in E.items()}
    for idx in [(0, 0, 0), (16, 16, 16), (5, 20, 9)]:
        want_mu = max(m["b"] * bg[idx] + m["h"] * (wm[idx] + gm[idx]) + m["c"] * (vt[idx] + csf[idx]), 0)
        want_mu += m["t"] * c[idx] if c[idx] > 0 else 0
        want_la = max(la["b"] * bg[idx] + la["h"] * (wm[idx] + gm[idx]) + la["c"] * (vt[idx] + csf[idx]), 0)
        want_la += la["t"] * c[idx] if c[idx] > 0 else 0
        want_s = 1e6 if bg[idx] > 0.95 else (1e-2 if c[idx] >= 0.005 else 1e2)
        assert abs(mu[idx] - want_mu) < 1e-9 * want_mu and abs(lam[idx] - want_la) < 1e-9 * want_la
        assert scr[idx] == want_s
    assert np.allclose(avg, [mu.mean(), lam.mean(), scr.mean()], rtol=1e-12)
    f = tumor_force(c)
    u, its, reason, rn = o.el_solve(mu, lam, scr, avg, f, np.zeros((3, n, n, n)))
    bf = o.el_prec(avg, f)
    assert reason == 2 and 1 <= its < 100
    assert rn <= 1e-3 * np.linalg.norm(bf)
