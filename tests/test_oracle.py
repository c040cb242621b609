"""The FP64 oracle pinned against the reference's own known-answer test, numpy golden vectors,
brute-force DFTs, closed forms and invariants (CPU only)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def trig_field(n, kx=2, ky=3, kz=1):
    x = np.arange(n) * 2 * np.pi / n
    return np.sin(kx * x)[:, None, None] * np.cos(ky * x)[None, :, None] * np.sin(kz * x)[None, None, :], x


# ---------------------------------------------------------------------------------- FFT ----
def test_fft_vs_numpy_golden(oracle_mod):
    g = np.load(os.path.join(HERE, "golden", "fft_golden.npz"))
    keys = [k[3:] for k in g.files if k.startswith("in_")]
    assert len(keys) >= 5
    for key in keys:
        f, F = g["in_" + key], g["out_" + key]
        o = oracle_mod.Oracle(f.shape)
        assert rel(o.r2c(f), F) < 1e-14, key
        assert rel(o.c2r(F), f * f.size) < 1e-14, key


def test_fft_vs_bruteforce_dft(oracle_mod):
    shape = (4, 6, 10)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(shape)
    idx = [np.arange(s) for s in shape]
    F = np.zeros((4, 6, 6), complex)
    for a in range(4):
        for b in range(6):
            for c in range(6):
                ph = np.exp(-2j * np.pi * (a * idx[0][:, None, None] / 4 + b * idx[1][None, :, None] / 6
                                           + c * idx[2][None, None, :] / 10))
                F[a, b, c] = (f * ph).sum()
    o = oracle_mod.Oracle(shape)
    assert rel(o.r2c(f), F) < 1e-13


# ------------------------------------------------------------------- spectral operators ----
def test_gradient_closed_form(oracle_mod):
    n = 16
    f, x = trig_field(n)
    o = oracle_mod.Oracle(n)
    gx, gy, gz = o.grad(f)
    ex = 2 * np.cos(2 * x)[:, None, None] * np.cos(3 * x)[None, :, None] * np.sin(x)[None, None, :]
    ey = -3 * np.sin(2 * x)[:, None, None] * np.sin(3 * x)[None, :, None] * np.sin(x)[None, None, :]
    ez = np.sin(2 * x)[:, None, None] * np.cos(3 * x)[None, :, None] * np.cos(x)[None, None, :]
    assert rel(gx, ex) < 1e-13 and rel(gy, ey) < 1e-13 and rel(gz, ez) < 1e-13


def test_nyquist_mode_has_zero_gradient(oracle_mod):
    # multiplyXWaveNumber zeroes k = n/2 (src/cuda/SpectralOperators.cu:69-72)
    n = 16
    x = np.arange(n) * 2 * np.pi / n
    f = np.cos(n / 2 * x)[:, None, None] * np.ones((1, n, n))
    o = oracle_mod.Oracle(n)
    gx, gy, gz = o.grad(f)
    assert np.abs(gx).max() < 1e-12 and np.abs(gy).max() < 1e-12


def test_div_grad_is_laplacian(oracle_mod):
    n = 16
    f, _ = trig_field(n)
    o = oracle_mod.Oracle(n)
    lap = o.div(*o.grad(f))
    assert rel(lap, -(4 + 9 + 1) * f) < 1e-12


def test_apply_d_constant_k_is_k_laplacian(oracle_mod):
    n = 16
    f, _ = trig_field(n)
    o = oracle_mod.Oracle(n)
    o.set_tissue(np.ones((n, n, n)), None, None, None)  # wm = 1 everywhere
    o.set_k(0.3, 0.0, 0.0)
    assert rel(o.apply_d(f), -0.3 * 14 * f) < 1e-12


def test_smoother_preserves_constants(oracle_mod):
    n = 16
    o = oracle_mod.Oracle(n)
    c = np.full((n, n, n), 0.7)
    assert np.abs(o.smooth(c, 2 * np.pi / n) - 0.7).max() < 1e-12


# ---------------------------------------------------------------- diffusion solver ------
def test_pdesolver_kat(oracle_mod):
    """src/test/pdesolver.cpp:7-57: 64^3, sinusoidal k (scale 1e-2), preconditioner built with the
    default params dt 0.5, ten CN solves of 0.02: ||c||_2 = Approx(2.0487), last ksp_itr_ == 5."""
    n = 64
    o = oracle_mod.Oracle(n)
    o.set_k_sinusoidal(1e-2)
    o.set_time(0.5, 1)
    o.prec_factor()
    h = 2 * np.pi / n
    X = np.arange(n) * h - np.pi
    R = np.sqrt(2.0) * 2 * np.pi / 64
    c = np.exp(-(X[:, None, None] ** 2 + X[None, :, None] ** 2 + X[None, None, :] ** 2) / R ** 2)
    for _ in range(10):
        c, its = o.diffusion_solve(c, 0.02)
    nrm = np.linalg.norm(c)
    assert abs(nrm - 2.0487) <= 100 * np.finfo(np.float32).eps * 2.0487
    assert its == 5


def test_cn_closed_form_constant_k(oracle_mod):
    """Constant k: each Fourier mode is multiplied by (1 - tau k |w|^2 / 2) / (1 + tau k |w|^2 / 2);
    with the matched preconditioner (dt = 2 tau) CG converges in one iteration."""
    n, k, tau = 16, 0.05, 0.3
    f, _ = trig_field(n)
    o = oracle_mod.Oracle(n)
    o.set_tissue(np.ones((n, n, n)), None, None, None)
    o.set_k(k, 0.0, 0.0)
    o.set_solver(1e-12, 100)
    o.set_ctx_dt(2 * tau)
    o.prec_factor()
    c, its = o.diffusion_solve(f, tau)
    w2 = 14.0
    fac = (1 - tau * k * w2 / 2) / (1 + tau * k * w2 / 2)
    assert rel(c, fac * f) < 1e-12
    assert its == 1


def test_ksp_zero_k_is_noop(oracle_mod):
    n = 16
    o = oracle_mod.Oracle(n)
    o.set_k(0.0, 0.0, 0.0)
    f, _ = trig_field(n)
    c, its = o.diffusion_solve(f, 0.1)
    assert its == 0 and np.array_equal(c, f)


# ------------------------------------------------------------------------- reaction ------
def test_reaction_closed_form(oracle_mod):
    n, dt = 16, 0.05
    o = oracle_mod.Oracle(n)
    rng = np.random.default_rng(1)
    rho = rng.uniform(0, 12, (n, n, n))
    o.set_tissue(rho / 12.0, None, None, None)
    o.set_rho(12.0, 0.0, 0.0)
    rho = o.fields()["rho"]
    c = rng.uniform(0.0, 1.0, (n, n, n))
    c.flat[0] = 1.0  # the isinf branch keeps c = 1
    c.flat[1] = 0.0
    out = o.reaction(c, dt)
    e = np.exp(rho * dt)
    assert rel(out, c * e / (c * e + 1 - c)) < 1e-14
    assert out.flat[0] == 1.0 and out.flat[1] == 0.0
    # linearized reaction = derivative of the exact step (central difference in c)
    h = 1e-6
    dfd = (o.reaction(c + h, dt) - o.reaction(c - h, dt)) / (2 * h)
    lin = o.reaction(np.ones_like(c), dt, clin=c)
    assert rel(lin.ravel()[2:], dfd.ravel()[2:]) < 1e-7


# ------------------------------------------------------------------- forward / adjoint ----
def small_problem(oracle_mod, n=32, nt=4, dt=0.05, kappa=0.05, rho=8.0, rtol=1e-6, seed=5):
    from glia_amd import synth
    pr = synth.problem(n, seed, nbases=6, kappa=kappa, rho=rho, sigma=2 * np.pi / 32)
    o = oracle_mod.Oracle(n)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_solver(rtol, 5000)
    o.set_time(dt, nt)
    o.set_k(kappa, 0.2, 0.0)
    o.set_rho(rho, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    return pr, o


def test_forward_kappa_zero_is_pointwise_logistic(oracle_mod):
    pr, o = small_problem(oracle_mod, kappa=0.0, nt=5, dt=0.04)
    c0 = o.phi_apply(pr["p"])
    o.set_c0(c0)
    o.solve_state(0)
    rho = o.fields()["rho"]
    e = np.exp(rho * 0.2)
    assert rel(o.ct(), c0 * e / (c0 * e + 1 - c0)) < 1e-13


def test_forward_rho_zero_conserves_mass(oracle_mod):
    pr, o = small_problem(oracle_mod, rho=0.0, nt=4, dt=0.05, rtol=1e-10)
    c0 = o.phi_apply(pr["p"])
    o.set_c0(c0)
    o.solve_state(0)
    assert abs(o.ct().sum() - c0.sum()) / c0.sum() < 1e-10


def test_objective_invariants(oracle_mod):
    """src/test/grad.cpp:7-80 (J at the truth ~ 0, J > 0 for zero data)."""
    pr, o = small_problem(oracle_mod, nt=5, dt=0.02)
    x = np.concatenate([pr["p"], [0.05, 8.0]])
    o.set_c0(o.phi_apply(pr["p"]))
    o.set_time(0.02, 5)
    o.solve_state(0)
    d = o.ct()
    o.set_time(0.02, 5)
    o.set_data(d, None, -0.99)
    assert o.objective(x, mode=2) < 1e-7
    o.set_data(np.zeros_like(d), None, -0.99)
    assert o.objective(x, mode=2) > 0


def central_fd(o, x, d, h, mode=2):
    o.set_time(o.dt, o.nt)
    jp = o.objective(x + h * d, mode)
    o.set_time(o.dt, o.nt)
    jm = o.objective(x - h * d, mode)
    return (jp - jm) / (2 * h)


def test_gradient_p_is_exact_discrete_adjoint(oracle_mod):
    """g_p is the exact adjoint of the discrete Strang/CN scheme: central differences agree to
    the CG tolerance (rtol 1e-12 here)."""
    pr, o = small_problem(oracle_mod, nt=4, dt=0.05, rtol=1e-12)
    x = np.concatenate([pr["p"], [0.05, 8.0]])
    o.set_c0(o.phi_apply(pr["p"] * 1.2))
    o.solve_state(0)
    o.set_data((o.ct() > 0.3).astype(float), None, -0.99)
    o.set_opt(beta=1e-2)
    o.set_time(0.05, 4)
    J, g = o.objective_and_gradient(x, mode=2)
    rng = np.random.default_rng(9)
    for _ in range(3):
        d = np.zeros_like(x)
        d[:-2] = rng.standard_normal(len(x) - 2)
        fd = central_fd(o, x, d, 1e-4)
        assert abs(fd - g @ d) / abs(fd) < 1e-6


def test_gradient_kappa_rho_converge_in_dt(oracle_mod):
    """g_kappa and g_rho are trapezoid quadratures of the continuous-adjoint integrands: their
    gap to finite differences shrinks as dt is refined at fixed T."""
    gaps = []
    for nt in (2, 4, 8):
        pr, o = small_problem(oracle_mod, nt=nt, dt=0.2 / nt, rtol=1e-12)
        x = np.concatenate([pr["p"], [0.05, 8.0]])
        o.set_c0(o.phi_apply(pr["p"] * 1.2))
        o.solve_state(0)
        o.set_data((o.ct() > 0.3).astype(float), None, -0.99)
        o.set_time(0.2 / nt, nt)
        J, g = o.objective_and_gradient(x, mode=2)
        gap = []
        for k in (-2, -1):
            d = np.zeros_like(x)
            d[k] = 1.0
            fd = central_fd(o, x, d, 1e-5 * max(abs(x[k]), 1))
            gap.append(abs(g[k] - fd) / abs(fd))
        gaps.append(gap)
    gaps = np.array(gaps)
    assert np.all(gaps[1:] < gaps[:-1]), gaps
    assert np.all(gaps[-1] < 0.1), gaps


def test_mode_p_kappa_matches_all_for_shared_entries(oracle_mod):
    pr, o = small_problem(oracle_mod, nt=3, dt=0.05)
    x = np.concatenate([pr["p"], [0.05, 8.0]])
    o.set_c0(o.phi_apply(pr["p"] * 1.2))
    o.solve_state(0)
    o.set_data((o.ct() > 0.3).astype(float), None, -0.99)
    o.set_time(0.05, 3)
    _, g2 = o.objective_and_gradient(x, mode=2)
    o.set_time(0.05, 3)
    _, g0 = o.objective_and_gradient(x, mode=0)
    # OpenMP reductions combine per-thread partials in an unspecified order: equal to rounding
    assert np.allclose(g0[:-1], g2[:-1], rtol=1e-11, atol=0) and g0[-1] == 0.0


# ----------------------------------------------------------------------------- Phi -------
def test_phi_unsmoothed_is_masked_truncated_gaussian(oracle_mod):
    n = 32
    o = oracle_mod.Oracle(n)
    h = 2 * np.pi / n
    ctr = np.array([[16 * h, 15.5 * h, 17 * h]])
    sig = 2 * h
    o.set_phi(ctr, sig, 0.0, True)  # smoothing factor 0 -> no smoothing
    X = np.arange(n) * h
    r = np.sqrt((X[:, None, None] - ctr[0, 0]) ** 2 + (X[None, :, None] - ctr[0, 1]) ** 2 + (X[None, None, :] - ctr[0, 2]) ** 2)
    g = np.exp(-r ** 2 / (2 * sig ** 2))
    g[r / sig > 5] = 0
    assert rel(o.phi(0), g / g.max()) < 1e-14
    assert abs(o.phi(0).max() - 1.0) < 1e-15


def test_phi_on_the_fly_normalization_closed_form(oracle_mod):
    """Phi::apply without phi_store divides by the largest maximum among the ACTIVE bases and
    applyTranspose by the largest among all bases (src/mat/Phi.cpp:313-439).  Unsmoothed bases with
    hand-known maxima pin both: A on a grid point (max 1), B half a cell off (max e^(-1/32)), C
    inside a 3-cell hole of the mask, so its largest value sits 2 cells out (max e^(-1/2))."""
    n = 32
    h = 2 * np.pi / n
    sig = 2 * h
    o = oracle_mod.Oracle(n)
    wm = np.ones((n, n, n))
    wm[23:26, 23:26, 23:26] = 0.0  # mask hole around C
    o.set_tissue(wm, None, None, None)
    ctr = np.array([[16 * h, 16 * h, 16 * h], [8 * h, 8.5 * h, 8 * h], [24 * h, 24 * h, 24 * h]])
    o.set_phi(ctr, sig, 0.0, False)
    X = np.arange(n) * h
    mask = (wm > 0.1).astype(np.float64)
    g = []
    for c in ctr:
        r = np.sqrt((X[:, None, None] - c[0]) ** 2 + (X[None, :, None] - c[1]) ** 2 + (X[None, None, :] - c[2]) ** 2)
        gi = np.exp(-r ** 2 / (2 * sig ** 2)) * mask
        gi[r / sig > 5] = 0
        g.append(gi)
    maxima = (1.0, np.exp(-1.0 / 32.0), np.exp(-0.5))
    for gi, m in zip(g, maxima):
        assert abs(gi.max() - m) < 1e-15
    for p, norm in (([0, 0, 0.7], maxima[2]), ([0, 0.4, 0.7], maxima[1]), ([0.2, 0.4, 0.7], 1.0)):
        want = sum(pi * gi for pi, gi in zip(p, g)) / norm
        assert rel(o.phi_apply(np.array(p, dtype=np.float64)), want) < 1e-13
    assert np.all(o.phi_apply(np.zeros(3)) == 0.0)
    v = np.random.default_rng(2).standard_normal((n, n, n))
    want_t = np.array([np.sum(gi * v) for gi in g]) / 1.0
    assert rel(o.phi_applyT(v), want_t) < 1e-12
    # without basis A the largest maximum over all bases is B's
    o.set_phi(ctr[1:], sig, 0.0, False)
    assert rel(o.phi_applyT(v), want_t[1:] / maxima[1]) < 1e-12


def test_kbar_is_sum_of_k_over_mask_count(oracle_mod):
    """DiffCoef::setValues' kxx_avg_ is sum(k) over the whole grid divided by the number of mask
    voxels (src/mat/DiffCoef.cpp:103-116), so k outside the mask still counts in the numerator.  Half
    the grid white matter (mask 1, k = ks), a quarter ventricle with CSF 0.5 (mask 0 as vt >= 0.8,
    k = ks * k_glm * 0.5), a quarter empty: kbar = ks (1 + 0.25 k_glm)."""
    n = 16
    o = oracle_mod.Oracle(n)
    wm = np.zeros((n, n, n)); csf = np.zeros((n, n, n)); vt = np.zeros((n, n, n))
    wm[: n // 2] = 1.0
    csf[n // 2: 3 * n // 4] = 0.5
    vt[n // 2: 3 * n // 4] = 0.9
    o.set_tissue(wm, None, csf, vt)
    ks, kglm = 0.1, 0.2
    o.set_k(ks, 0.0, kglm)
    f = o.fields()
    assert f["filter"].sum() == n ** 3 // 2
    assert abs(f["kavg"] - ks * (1 + 0.25 * kglm)) < 1e-15


def test_phi_mesh_counts(oracle_mod):
    for np_, want in ((1, 1), (8, 8), (27, 27), (64, 64)):
        c = oracle_mod.phi_mesh(np_, np.array([3.0, 3.0, 3.0]), 0.1, 1.5)
        assert len(c) == want


def test_matprop_mask_rule(oracle_mod):
    n = 8
    o = oracle_mod.Oracle(n)
    wm = np.zeros((n, n, n)); gm = np.zeros((n, n, n)); vt = np.zeros((n, n, n))
    wm[0, 0, 0] = 0.2; gm[0, 0, 1] = 0.2; wm[0, 0, 2] = 0.05; gm[0, 0, 3] = 0.5; vt[0, 0, 3] = 0.9
    wm[0, 0, 4] = -0.3
    o.set_tissue(wm, gm, None, vt)
    f = o.fields()["filter"]
    assert list(f[0, 0, :5]) == [1.0, 1.0, 0.0, 0.0, 0.0]


def test_synthetic_tissue_properties():
    from glia_amd import synth
    a = synth.brain(32, 1)
    b = synth.brain(32, 1)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    wm, gm, csf, _ = a
    s = wm + gm + csf
    inside = s > 0
    assert np.allclose(s[inside], 1.0)
    assert wm.mean() > 0.01 and gm.mean() > 0.01 and csf.mean() > 0.005


# ---- Gauss-Newton Hessian matvec, p block (DerivativeOperatorsRD::evaluateHessian, no diffusivity
# inversion).  The reference's tangent-linear forward linearizes the reaction around the state at
# the start of each step while its adjoint uses the half-step state, so the operator is the exact
# Hessian only when one of the two processes is off; otherwise the gap is O(dt).

def grad_p(o, x, mode=0):
    o.set_time(o.dt, o.nt)
    return o.objective_and_gradient(x, mode)[1][:o.np]


def hess_p(o, x, v, mode=0):
    grad_p(o, x, mode)  # the linearization point: state history at x
    return o.hessian_p(v)


def test_hessian_linear_problem_is_exact_and_symmetric(oracle_mod):
    """rho = 0: J is quadratic in p, so H p~ equals central differences of the gradient and the
    operator is symmetric (to the CG tolerance, rtol 1e-12)."""
    pr, o = small_problem(oracle_mod, rho=0.0, nt=4, dt=0.05, rtol=1e-12)
    x = np.concatenate([pr["p"], [0.05, 0.0]])
    o.set_c0(o.phi_apply(pr["p"] * 1.2))
    o.solve_state(0)
    o.set_data((o.ct() > 0.3).astype(float), None, -0.99)
    o.set_opt(beta=1e-2)
    rng = np.random.default_rng(21)
    u, v = rng.standard_normal(o.np), rng.standard_normal(o.np)
    hv = hess_p(o, x, v)
    h = 1e-3
    dv = np.concatenate([v, [0.0, 0.0]])
    fd = (grad_p(o, x + h * dv) - grad_p(o, x - h * dv)) / (2 * h)
    assert rel(hv, fd) < 1e-7
    hu = hess_p(o, x, u)
    assert abs(u @ hv - v @ hu) / abs(u @ hv) < 1e-9
    assert v @ hv > 0  # Gauss-Newton with beta > 0 is positive definite on the active bases


def test_hessian_pure_reaction_matches_fd_at_zero_residual(oracle_mod):
    """kappa = 0 (the half-step state equals the step's start state) with data = O c(T): the
    Gauss-Newton operator is the exact Hessian; central differences of the gradient agree."""
    pr, o = small_problem(oracle_mod, kappa=0.0, nt=4, dt=0.05, rtol=1e-12)
    x = np.concatenate([pr["p"], [0.0, 8.0]])
    o.set_c0(o.phi_apply(pr["p"]))
    o.solve_state(0)
    o.set_data(o.ct(), np.ones(o.ng), 0.0)
    o.set_opt(beta=1e-3)
    v = np.random.default_rng(22).standard_normal(o.np)
    hv = hess_p(o, x, v)
    h = 1e-4
    dv = np.concatenate([v, [0.0, 0.0]])
    fd = (grad_p(o, x + h * dv) - grad_p(o, x - h * dv)) / (2 * h)
    assert rel(hv, fd) < 1e-6


def test_hessian_gap_to_exact_hessian_shrinks_with_dt(oracle_mod):
    """With diffusion and reaction, the reference operator differs from the exact Hessian by its
    linearization point (start of step vs half step); the gap shrinks as dt is refined."""
    gaps = []
    for nt in (2, 4, 8):
        pr, o = small_problem(oracle_mod, nt=nt, dt=0.2 / nt, rtol=1e-12)
        x = np.concatenate([pr["p"], [0.05, 8.0]])
        o.set_c0(o.phi_apply(pr["p"]))
        o.solve_state(0)
        o.set_data(o.ct(), np.ones(o.ng), 0.0)
        o.set_opt(beta=1e-3)
        v = np.random.default_rng(23).standard_normal(o.np)
        hv = hess_p(o, x, v)
        h = 1e-4
        dv = np.concatenate([v, [0.0, 0.0]])
        fd = (grad_p(o, x + h * dv) - grad_p(o, x - h * dv)) / (2 * h)
        gaps.append(rel(hv, fd))
    assert gaps[1] < gaps[0] and gaps[2] < gaps[1], gaps
    assert gaps[2] < 0.2, gaps


def test_hessian_with_diffusivity_is_affine_with_stale_offset(oracle_mod):
    """The kappa blocks enter through the secondary coefficient k~ m (setSecondaryCoefficients),
    linearly in the direction.  The kappa row also holds, from each of its two quadratures, the
    i = nt term with the adjoint history entry p_[nt], which solveAdjoint(2) leaves at the last
    gradient's a(T): so H(0) = [0 | 2 * (dt/2) h^3 <wm + k_gm gm, grad c(T) . grad a(T)> | 0] and
    H is affine (to the CG tolerance)."""
    pr, o = small_problem(oracle_mod, nt=3, dt=0.05, rtol=1e-12)
    x = np.concatenate([pr["p"], [0.05, 8.0]])
    o.set_c0(o.phi_apply(pr["p"] * 1.2))
    o.solve_state(0)
    o.set_data((o.ct() > 0.3).astype(float), None, -0.99)
    o.set_opt(beta=1e-3)
    rng = np.random.default_rng(41)
    u = np.concatenate([rng.standard_normal(o.np), [0.02, 0.0]])
    v = np.concatenate([rng.standard_normal(o.np), [-0.01, 0.0]])

    def H(d):
        o.set_time(o.dt, o.nt)
        o.objective_and_gradient(x, 0)
        return o.hessian(d, diffusivity=True)

    h0, hu, hv, huv = H(np.zeros_like(u)), H(u), H(v), H(2 * u - 3 * v)
    assert rel(huv - h0, 2 * (hu - h0) - 3 * (hv - h0)) < 1e-9
    assert np.all(h0[:-2] == 0.0) and hu[-1] == 0.0
    # the offset from the stale adjoint entry, recomputed from the oracle's own history
    o.set_time(o.dt, o.nt)
    o.objective_and_gradient(x, 0)
    gc = o.grad(o.c_hist(o.nt))
    ga = o.grad(o.p_hist(o.nt))
    q = sum(a * b for a, b in zip(gc, ga))
    h3 = (2 * np.pi) ** 3 / o.ng
    expect = 2 * 0.5 * o.dt * h3 * np.sum((np.clip(pr["wm"], 0, None) + 0.2 * np.clip(pr["gm"], 0, None)) * q)
    assert abs(h0[-2] - expect) <= 1e-10 * abs(expect)


def test_hessian_kappa_blocks_gap_shrinks_with_dt(oracle_mod):
    """At zero residual, the reference operator (tangent linearized at the step start, pseudo-time
    averaged secondary sources, the stale adjoint entry p_[nt]) differs from central differences
    of [g_p | g_kappa] by terms that vanish with dt."""
    gaps = []
    for nt in (2, 4, 8):
        pr, o = small_problem(oracle_mod, nt=nt, dt=0.2 / nt, rtol=1e-12)
        x = np.concatenate([pr["p"], [0.05, 8.0]])
        o.set_c0(o.phi_apply(pr["p"]))
        o.solve_state(0)
        o.set_data(o.ct(), np.ones(o.ng), 0.0)
        o.set_opt(beta=1e-3)
        rng = np.random.default_rng(42)
        d = np.concatenate([rng.standard_normal(o.np), [0.02, 0.0]])
        o.set_time(o.dt, o.nt)
        o.objective_and_gradient(x, 0)
        hd = o.hessian(d, diffusivity=True)[:-1]
        h = 1e-4

        def g(xx):
            o.set_time(o.dt, o.nt)
            return o.objective_and_gradient(xx, 0)[1][:-1]

        fd = (g(x + h * d) - g(x - h * d)) / (2 * h)
        gaps.append(rel(hd, fd))
    assert gaps[1] < gaps[0] and gaps[2] < gaps[1], gaps
    assert gaps[2] < 0.2, gaps
