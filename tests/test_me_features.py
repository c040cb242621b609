"""The oracle's biophysical features of the model-4 loop (orc_me_features and orc_stress_point,
restating PdeOperatorsMassEffect::computeBiophysicalFeatures, src/pde/PdeOperatorsMassEffect.cpp:
323-418, and computeStressQuants, :140-212) against closed forms (CPU only).

The state is a tumor with a sine profile across x on a 16^3 grid of white matter: c = 0.5 + 0.5 sin x,
no displacement, no velocity.  Everything the features depend on follows by hand:
- computeSegmentation labels the planes with c >= 0.5 as tumor core (x index 0..8) and the rest
  white matter, so the tumor ratio is 9/16;
- |grad c| = 0.5 |cos x|, and with k = kappa, rho constant, dc/dt = -0.5 kappa sin x + rho c (1 - c);
- the largest |grad c| outside the core sits on planes 9 and 15 (a tie of equal c), so c* =
  0.5 - 0.5 sin(pi/8) and the surface indicator |c - c*| < 0.1 holds on exactly those two planes;
- quantiles are plane values at rank floor(q m) / 256 of the sorted plane values.
"""
import numpy as np

N = 16


def col(feature, stat):
    """Column of (feature, stat) in a row: 2 + 16 f + offset (vol 0-2, sa 3, std 4-6, q875 7-9,
    q95 10-12, q99 13-15; within a triple: grid, tc, healthy)."""
    return 2 + 16 * feature + stat


def test_stress_point_closed_forms_and_eigenvalues():
    from oracle import oracle
    mu, lam = 2.0, 3.0
    # simple shear u_x = a y: det 1, trace 0, principal stresses +-mu a, 0
    a = 0.3
    jac, tr, tau = oracle.stress_point([0, a, 0, 0, 0, 0, 0, 0, 0], mu, lam)
    assert abs(jac - 1.0) < 1e-15 and abs(tr) < 1e-15 and abs(tau - mu * a) < 1e-14
    # dilation along the axes: det is the product of stretches, trace (3 lam + 2 mu) tr(E)
    d = (0.1, -0.2, 0.05)
    jac, tr, tau = oracle.stress_point([d[0], 0, 0, 0, d[1], 0, 0, 0, d[2]], mu, lam)
    assert abs(jac - (1 + d[0]) * (1 + d[1]) * (1 + d[2])) < 1e-14
    assert abs(tr - (3 * lam + 2 * mu) * sum(d)) < 1e-14
    assert abs(tau - mu * (max(d) - min(d))) < 1e-14
    # a general gradient: the Jacobi eigenvalues against LAPACK's symmetric solver
    rng = np.random.default_rng(4)
    for _ in range(20):
        g = rng.standard_normal(9) * 0.1
        G = g.reshape(3, 3)
        E = 0.5 * (G + G.T)
        S = lam * np.trace(E) * np.eye(3) + 2 * mu * E
        ev = np.linalg.eigvalsh(S)
        jac, tr, tau = oracle.stress_point(g, mu, lam)
        assert abs(jac - np.linalg.det(np.eye(3) + G)) < 1e-13
        assert abs(tr - np.trace(S)) < 1e-13
        assert abs(tau - 0.5 * (ev[-1] - ev[0])) < 1e-12


def test_features_of_a_sine_tumor_closed_form(oracle_mod):
    h = 2 * np.pi / N
    X = np.arange(N) * h
    s = np.sin(X)[:, None, None] * np.ones((N, N, N))
    c = 0.5 + 0.5 * s
    wm = np.ones((N, N, N))
    zero = np.zeros((N, N, N))
    kappa, rho, mu, lam, dt, it = 0.03, 6.0, 2.0, 3.0, 0.25, 3
    o = oracle_mod.Oracle(N)
    o.set_tissue(wm, zero, zero, zero, zero)
    o.set_k(kappa)
    o.set_rho(rho)
    row = o.me_features(it, dt, c, wm, zero, zero, zero, zero, np.zeros(3 * N ** 3), np.zeros(3 * N ** 3),
                        np.full((N, N, N), mu), np.full((N, N, N), lam))
    V = (2 * np.pi) ** 3          # grid volume
    plane = N * N * h ** 3        # volume of one x-plane
    sa_plane = N * N * h ** 3 * N / (2 * np.pi)
    assert row[0] == it * dt
    assert abs(row[1] - 9 / 16) < 1e-15
    cstar = 0.5 - 0.5 * np.sin(np.pi / 8)
    # c: integrals, surface term, spread and order statistics
    sin_sum_core = 1 / np.tan(np.pi / 16)  # sum of sin(k pi / 8), k = 0..8
    assert abs(row[col(0, 0)] - 0.5 * V) < 1e-12 * V
    assert abs(row[col(0, 1)] - plane * (4.5 + 0.5 * sin_sum_core)) < 1e-12 * V
    assert abs(row[col(0, 2)] - (0.5 * V - plane * (4.5 + 0.5 * sin_sum_core))) < 1e-12 * V
    assert abs(row[col(0, 3)] - 2 * sa_plane * cstar) < 1e-12 * sa_plane
    ng = N ** 3
    assert abs(row[col(0, 4)] - np.sqrt(0.25 * (ng / 2) / (ng - 1))) < 1e-13
    q = [0.5 + 0.5 * np.cos(np.pi / 8), 1.0, 1.0]                     # grid: plane ranks 14, 15, 15
    qt = [0.5 + 0.5 * np.cos(np.pi / 8), 1.0, 1.0]                    # core: ranks 7, 8, 8 of 9
    qh = [cstar, cstar, cstar]                                        # healthy: rank 6 of 7
    for k in range(3):
        assert abs(row[col(0, 7 + 3 * k)] - q[k]) < 1e-14, k
        assert abs(row[col(0, 8 + 3 * k)] - qt[k]) < 1e-14, k
        assert abs(row[col(0, 9 + 3 * k)] - qh[k]) < 1e-14, k
    # |grad c| = 0.5 |cos x|: the grid's largest value is 0.5 (ranks 14, 15 of the plane values)
    cos_abs_sum = np.sum(np.abs(np.cos(X)))
    assert abs(row[col(2, 0)] - plane * 0.5 * cos_abs_sum) < 1e-11 * V
    for k in range(3):
        assert abs(row[col(2, 7 + 3 * k)] - 0.5) < 1e-12
    # dc/dt = -0.5 kappa sin x + rho c (1 - c) = -0.5 kappa sin x + 0.25 rho cos^2 x: integral 2 rho
    # over the planes (the sine part sums to zero, cos^2 to 8), surface term on planes 9 and 15
    assert abs(row[col(1, 0)] - plane * 2 * rho) < 1e-11 * V * rho
    d9 = -0.5 * kappa * np.sin(9 * np.pi / 8) + 0.25 * rho * np.cos(9 * np.pi / 8) ** 2
    assert abs(row[col(1, 3)] - 2 * sa_plane * d9) < 1e-11 * sa_plane * rho
    # no displacement: |u|, tr S and the shear vanish, det F = 1 everywhere; no velocity
    for f in (3, 5, 6, 7):
        assert np.all(np.abs(row[2 + 16 * f: 2 + 16 * (f + 1)]) < 1e-14), f
    jac = row[2 + 16 * 4: 2 + 16 * 5]
    assert abs(jac[0] - V) < 1e-12 * V and abs(jac[1] - 9 * plane) < 1e-12 * V and abs(jac[2] - 7 * plane) < 1e-12 * V
    assert abs(jac[3] - 2 * sa_plane) < 1e-12 * sa_plane
    assert np.all(np.abs(jac[4:7]) < 1e-14) and np.all(np.abs(jac[7:] - 1.0) < 1e-14)


def test_feature_branch_rows_and_volume_stop(oracle_mod):
    """The loop with its feature branch writes one row per step, each row's first entries the time
    and the tumor ratio, and stops once the ratio is near 5% or above after t = 1 (the rule of
    :592-603); a run that reaches that ratio before t = 1 keeps going."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from me_cases import mass_effect_case
    n, dt, nt = 16, 0.25, 6
    c0, wm, gm, vt, csf, bg = mass_effect_case(n)
    o = oracle_mod.Oracle(n)
    o.set_tissue(wm, gm, csf, vt, bg)
    o.set_k(0.01, 0.2)
    o.set_rho(12.0, 0.2)
    o.set_time(dt, nt)
    out = o.me_solve_state(c0, wm, gm, vt, csf, bg, ff=2e4, features=True)
    F = out["features"]
    ratios = F[:, 1]
    print("ratios", ratios, out["info"])
    assert np.allclose(F[:, 0], np.arange(len(F)) * dt)
    stop = [i for i in range(len(F)) if (abs(np.float32(ratios[i]) - np.float32(0.05)) < np.float32(5e-3)
                                          or np.float32(ratios[i]) > np.float32(0.05)) and np.float32(i) * np.float32(dt) >= 1]
    if stop:
        assert out["info"]["volume_stop"] and len(F) == stop[0] + 1
    else:
        assert not out["info"]["volume_stop"] and len(F) == nt + 1
    # the first row sees the initial state: c is the filtered initial tumor
    assert abs(F[0, col(0, 0)] - np.sum(c0) * (2 * np.pi / n) ** 3) < 1e-9 * np.sum(c0)
