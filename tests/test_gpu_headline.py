"""GPU parity on the headline grids: the 128- and 256-point line plans that the 128^3 and 256^3
workloads run, against the FP64 oracle (through the C ABI).

Every 256-point kernel is reached on grids with one long axis: 256 x 32 x 32 runs the X passes
(k_xi_grad<256>, k_xf_div<256>), 32 x 256 x 32 the Y passes (k_yi_grad<256>, k_yf_div<256>) and
32 x 32 x 256 the Z passes on 128-point packed lines (k_z_diff<128>, k_z_point<128>,
k_z_gradq<128>); the 128 analogues do the same for the 128-point plans.  The oracle finishes
those in seconds.  The full 256^3 and 128^3 grids are checked against fixtures the oracle wrote
(tests/golden/make_headline_golden.py): J and dJ, ||c(T)|| and c(T) on a 32^3 sub-lattice.

Bars (relative L2 unless stated): transforms 5e-6, applyD 1e-5, c(T) 1e-5, J 1e-5, dJ 1e-4.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LONG_AXIS = [(256, 32, 32), (32, 256, 32), (32, 32, 256), (128, 32, 32), (32, 128, 32), (32, 32, 128)]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def rd():
    from glia_amd import rd as rdmod
    return rdmod


@pytest.mark.parametrize("n", [(256, 256, 256), (128, 128, 128)] + LONG_AXIS)
def test_fft_round_trip_headline_plans(rd, n):
    """R2C against numpy's FP64 rfftn and C2R of numpy's spectrum (times N, the unnormalized
    inverse) on the 128- and 256-point plans."""
    print("case fft", n, flush=True)
    s = rd.RDSolver(n)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(n).astype(np.float32)
    F = s.fft_r2c(0, f)
    ref = np.fft.rfftn(f.astype(np.float64))
    er, ei = rel(F.real, ref.real), rel(F.imag, ref.imag)
    g = s.fft_c2r(0, ref)
    eb = rel(g, np.fft.irfftn(ref, s=n, axes=(0, 1, 2)) * np.prod(n))
    s.close()
    print("r2c %.2e %.2e c2r %.2e" % (er, ei, eb), flush=True)
    assert er < 5e-6 and ei < 5e-6 and eb < 5e-6


def problem(n, seed, nbases=6, nnz=None):
    from glia_amd import synth
    pr = synth.problem(n, seed, nbases=nbases, nnz=nnz, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(k_gm_wm=0.2, r_gm_wm=0.2)
    return pr


def oracle_for(oracle_mod, pr, dt, nt):
    o = oracle_mod.Oracle(pr["n"])
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(dt, nt)
    o.set_k(pr["kappa"], pr["k_gm_wm"], 0.0)
    o.set_rho(pr["rho"], pr["r_gm_wm"], 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    return o


def solver_for(rd, pr, dt, nt, **kw):
    s = rd.RDSolver(pr["n"], dt=dt, nt=nt, k_gm_wm=pr["k_gm_wm"], r_gm_wm=pr["r_gm_wm"], **kw)
    s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    s.set_coefficients(0, pr["kappa"], pr["rho"])
    s.set_basis(0, pr["centers"], pr["sigma"])
    return s


@pytest.mark.parametrize("n", LONG_AXIS)
def test_apply_d_headline_plans_vs_oracle(rd, oracle_mod, n):
    """D c = div(k grad c) with the brain-like k = kappa (m_wm + 0.2 m_gm) (DiffCoef::applyD,
    src/mat/DiffCoef.cpp:206-271) through the gradient and divergence kernels of the long axis."""
    print("case apply_d", n, flush=True)
    pr = problem(n, 21)
    o = oracle_for(oracle_mod, pr, 0.02, 1)
    rng = np.random.default_rng(5)
    c = o.phi_apply(rng.uniform(0.2, 1.0, 6)) + 0.1 * np.cos(np.arange(np.prod(n)) * 1e-3).reshape(n)
    ref = o.apply_d(c).reshape(n)
    s = solver_for(rd, pr, 0.02, 1)
    s.set_field(0, "work", c)
    s.apply_d("work", "ct")
    got = s.get_field(0, "ct")
    s.close()
    print("applyD rel %.2e" % rel(got, ref), flush=True)
    assert rel(got, ref) < 1e-5


@pytest.mark.parametrize("n", LONG_AXIS)
def test_gradient_headline_plans_vs_oracle(rd, oracle_mod, n):
    """c(T), J and dJ = [g_p | g_kappa | g_rho] with every CG, reaction, adjoint and quadrature
    kernel of the long axis's plan (nt = 4, dt = 0.02, data thresholded from 1.3 p)."""
    print("case gradient", n, flush=True)
    dt, nt = 0.02, 4
    pr = problem(n, 23)
    o = oracle_for(oracle_mod, pr, dt, nt)
    o.set_c0(o.phi_apply(pr["p"] * 1.3))
    o.solve_state(0)
    d1 = (o.ct() > 0.4).astype(np.float64)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    o.set_data(d1, None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(dt, nt)
    Jo, go = o.objective_and_gradient(x, mode=2)
    cto = o.ct()
    s = solver_for(rd, pr, dt, nt, beta=1e-3)
    s.set_data(0, d1, None, -0.99)
    J, g = s.objective_and_gradient([x])
    ct = s.get_field(0, "ct")
    s.close()
    print("c(T) %.2e J %.2e dJ %.2e" % (rel(ct, cto), abs(J[0] - Jo) / abs(Jo), rel(g[0], go)), flush=True)
    assert rel(ct, cto) < 1e-5
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4
    for k in (-2, -1):
        assert abs(g[0][k] - go[k]) / max(abs(go[k]), 1e-2 * np.linalg.norm(go)) < 1e-4


def load_golden(name):
    path = os.path.join(GOLDEN, name)
    z = np.load(path)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["headline_256_nt2.npz", "headline_128_nt50.npz"])
def test_headline_grid_gradient_vs_oracle_fixture(rd, name):
    """The full 256^3 grid (nt = 2, 128 bases) and the 128^3 gradient configuration (nt = 50, 64
    bases 8-sparse, data from a ground-truth p*, O = 1): c(T), J and dJ against the oracle's
    values.  c(T) is compared on the fixture's 32^3 sub-lattice and through ||c(T)||."""
    from golden.make_headline_golden import identity, problem as golden_problem
    gd = load_golden(name)
    n, nt, dt = int(gd["n"]), int(gd["nt"]), float(gd["dt"])
    print("case headline", name, flush=True)
    pr = golden_problem(n, int(gd["nbases"]), int(gd["nnz"]), int(gd["seed"]), int(gd["seed_star"]))
    # the synthetic inputs are those the oracle ran on
    assert np.allclose(identity(pr), gd["identity"], rtol=1e-12, atol=0)
    assert np.array_equal(pr["p"], gd["p"])
    d1 = np.unpackbits(gd["d1_bits"])[: n ** 3].reshape(n, n, n).astype(np.float32)
    s = rd.RDSolver(n, dt=dt, nt=nt, k_gm_wm=float(gd["k_gm_wm"]), r_gm_wm=float(gd["r_gm_wm"]),
                    beta=float(gd["beta"]))
    s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    s.set_coefficients(0, float(gd["kappa"]), float(gd["rho"]))
    s.set_basis(0, pr["centers"], pr["sigma"])
    s.set_data(0, d1, None, -0.99)
    x = np.concatenate([pr["p"], [float(gd["kappa"]), float(gd["rho"])]])
    J, g = s.objective_and_gradient([x])
    ct = s.get_field(0, "ct")
    st = s.stats(0)
    s.close()
    stride = int(gd["stride"])
    e_sub = rel(ct[::stride, ::stride, ::stride], gd["ct_sub"])
    e_norm = abs(float(np.linalg.norm(ct.astype(np.float64))) - float(gd["ct_norm"])) / float(gd["ct_norm"])
    eJ = abs(J[0] - float(gd["J"])) / abs(float(gd["J"]))
    eg = rel(g[0], gd["dJ"])
    print("c(T) sub %.2e norm %.2e J %.2e dJ %.2e its %s" % (e_sub, e_norm, eJ, eg, st["pcg_iterations"]), flush=True)
    assert e_sub < 1e-5 and e_norm < 1e-5
    assert eJ < 1e-5
    assert eg < 1e-4
    go = gd["dJ"]
    for k in (-2, -1):
        assert abs(g[0][k] - go[k]) / max(abs(go[k]), 1e-2 * np.linalg.norm(go)) < 1e-4


@pytest.mark.parametrize("split", ["none", "xf", "xf3", "yf", "yfb", "yfb2"])
@pytest.mark.parametrize("n", [(32, 256, 32), (256, 32, 32), (32, 128, 32)])
def test_divergence_variants_headline_plans_vs_oracle(rd, oracle_mod, n, split):
    """Every variant of the forward divergence passes the autotuner may choose (GLIA_SPLIT_DIV) on
    the long-axis grids that run it at 128 or 256 points: applyD within 1e-5 and J, dJ within the
    oracle bars (nt = 2)."""
    import os
    print("case variant", n, split, flush=True)
    dt, nt = 0.02, 2
    pr = problem(n, 25)
    o = oracle_for(oracle_mod, pr, dt, nt)
    rng = np.random.default_rng(6)
    c = o.phi_apply(rng.uniform(0.2, 1.0, 6))
    ref_d = o.apply_d(c).reshape(n)
    o.set_c0(o.phi_apply(pr["p"] * 1.3))
    o.solve_state(0)
    d1 = (o.ct() > 0.4).astype(np.float64)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    o.set_data(d1, None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(dt, nt)
    Jo, go = o.objective_and_gradient(x, mode=2)
    old = os.environ.get("GLIA_SPLIT_DIV")
    os.environ["GLIA_SPLIT_DIV"] = split
    try:
        s = solver_for(rd, pr, dt, nt, beta=1e-3)
    finally:
        if old is None:
            os.environ.pop("GLIA_SPLIT_DIV", None)
        else:
            os.environ["GLIA_SPLIT_DIV"] = old
    s.set_field(0, "work", c)
    s.apply_d("work", "ct")
    ed = rel(s.get_field(0, "ct"), ref_d)
    s.set_data(0, d1, None, -0.99)
    J, g = s.objective_and_gradient([x])
    s.close()
    print("applyD %.2e J %.2e dJ %.2e" % (ed, abs(J[0] - Jo) / abs(Jo), rel(g[0], go)), flush=True)
    assert ed < 1e-5
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4
