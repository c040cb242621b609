"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (relative L2 unless stated): spectral operators 1e-5; c(T) 1e-5; objective 1e-5;
gradient 1e-4 -- the bars of the FP32-bulk / FP64-reduction design against the FP64 restatement
of the reference path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def rd():
    from glia_amd import rd as rdmod
    return rdmod


def gauss_test_function(n):
    # helper.cpp createTestFunction: exp(-(r/R)^2), R = sqrt(2) 2 pi / 64, centered at (pi, pi, pi)
    h = 2 * np.pi / n
    X = np.arange(n) * h - np.pi
    R = np.sqrt(2.0) * 2 * np.pi / 64
    return np.exp(-((X[:, None, None] ** 2 + X[None, :, None] ** 2 + X[None, None, :] ** 2)) / R ** 2)


@pytest.mark.parametrize("n", [(32, 32, 32), (64, 64, 64), (48, 96, 64), (96, 64, 96), (192, 64, 192), (64, 192, 64),
                               (160, 80, 160), (80, 160, 32)])
def test_fft_r2c_c2r_vs_numpy(rd, n):
    print("case fft", n, flush=True)
    s = rd.RDSolver(n)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(n).astype(np.float32)
    F = s.fft_r2c(0, f)
    ref = np.fft.rfftn(f.astype(np.float64))
    assert rel(F.real, ref.real) < 5e-6 and rel(F.imag, ref.imag) < 5e-6
    g = s.fft_c2r(0, ref)
    assert rel(g, np.fft.irfftn(ref, s=n, axes=(0, 1, 2)) * np.prod(n)) < 5e-6
    s.close()


def test_apply_d_vs_oracle(rd, oracle_mod):
    n = 64
    print("case apply_d", n, flush=True)
    o = oracle_mod.Oracle(n)
    o.set_k_sinusoidal(1e-2)
    c = gauss_test_function(n)
    ref = o.apply_d(c)
    s = rd.RDSolver(n)
    s.set_k_sinusoidal(0, 1e-2)
    s.set_field(0, "work", c)
    s.apply_d("work", "ct")
    got = s.get_field(0, "ct")
    assert rel(got, ref) < 1e-5


def test_diffusion_kat(rd, oracle_mod):
    """Reference known-answer test src/test/pdesolver.cpp:7-57: ||c||_2 = 2.0487, last solve 5 its."""
    n = 64
    print("case kat", flush=True)
    s = rd.RDSolver(n)  # params dt default 0.5 -> preconditioner dt
    s.set_k_sinusoidal(0, 1e-2)
    s.precFactor()
    ds = rd.DiffusionSolver(s)
    c = gauss_test_function(n)
    o = oracle_mod.Oracle(n)
    o.set_k_sinusoidal(1e-2)
    o.set_time(0.5, 1)
    o.prec_factor()
    co = c.copy()
    its = []
    for _ in range(10):
        c = ds.solve(c, 0.02)
        its.append(ds.ksp_itr_)
        co, _ = o.diffusion_solve(co, 0.02)
    nrm = float(np.linalg.norm(np.asarray(c, dtype=np.float64)))
    # Catch2 Approx(2.0487): epsilon 100 * FLT_EPSILON relative to the expected value
    assert abs(nrm - 2.0487) <= 100 * np.finfo(np.float32).eps * 2.0487, nrm
    assert its[-1] == 5, its
    assert rel(c, co) < 1e-5


def c1_problem(n=64, nt=50, dt=0.02, nb=8, seed=1):
    from glia_amd import synth
    pr = synth.problem(n, seed, nbases=nb, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(nt=nt, dt=dt, k_gm_wm=0.2, r_gm_wm=0.2)
    return pr


def setup_gpu(rd, pr, **kw):
    s = rd.RDSolver(pr["n"], dt=pr["dt"], nt=pr["nt"], k_gm_wm=pr["k_gm_wm"], r_gm_wm=pr["r_gm_wm"], **kw)
    s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    s.set_coefficients(0, pr["kappa"], pr["rho"])
    s.set_basis(0, pr["centers"], pr["sigma"])
    return s


def setup_oracle(oracle_mod, pr, phi_store=False, adjoint_store=False):
    o = oracle_mod.Oracle(pr["n"])
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(pr["dt"], pr["nt"], adjoint_store)
    o.set_k(pr["kappa"], pr["k_gm_wm"], 0.0)
    o.set_rho(pr["rho"], pr["r_gm_wm"], 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, phi_store)
    return o


def test_phi_vs_oracle(rd, oracle_mod):
    pr = c1_problem(nt=1)
    print("case phi", flush=True)
    s = setup_gpu(rd, pr)
    o = setup_oracle(oracle_mod, pr)
    p = pr["p"].copy()
    p[2] = 0.0
    s.phi_apply([p])
    c0 = s.get_field(0, "c0")
    c0o = o.phi_apply(p)
    assert rel(c0, c0o) < 1e-6
    v = np.cos(np.linspace(0, 7, c0o.size)).reshape(c0o.shape)
    s.set_field(0, "work", v)
    pt = s.phi_apply_transpose("work")[0]
    pto = o.phi_applyT(v)
    assert rel(pt, pto) < 1e-6


def test_forward_c1_vs_oracle(rd, oracle_mod):
    """The 64^3 forward workload (nt = 50, dt = 0.02, FP64 reference): c(T) within 1e-5 of the oracle."""
    pr = c1_problem()
    print("case forward c1", flush=True)
    s = setup_gpu(rd, pr)
    o = setup_oracle(oracle_mod, pr)
    s.phi_apply([pr["p"]])
    s.precFactor()
    s.solve_state(0)
    ct = s.get_field(0, "ct")
    o.set_c0(o.phi_apply(pr["p"]))
    o.solve_state(0)
    cto = o.ct()
    st = s.stats(0)
    print("c(T) rel err", rel(ct, cto), "its gpu", st["pcg_iterations"], "oracle", o.ksp()["state"], flush=True)
    assert rel(ct, cto) < 1e-5


def grad_problem(oracle_mod, nt=10, seed=2):
    pr = c1_problem(nt=nt, nb=8, seed=seed)
    # observed data from a different ground-truth p, thresholded at 0.5; O = 1 everywhere
    o = setup_oracle(oracle_mod, pr)
    pstar = pr["p"] * 1.3
    o.set_c0(o.phi_apply(pstar))
    o.solve_state(0)
    d1 = (o.ct() > 0.5).astype(np.float64)
    pr["d1"] = d1
    return pr, o


def test_gradient_vs_oracle(rd, oracle_mod):
    pr, o = grad_problem(oracle_mod)
    print("case gradient", flush=True)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    o.set_data(pr["d1"], None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(pr["dt"], pr["nt"])  # fresh DiffusionSolver context (ctx dt = dt)
    Jo, gJo = o.objective_and_gradient(x, mode=2)
    s = setup_gpu(rd, pr, beta=1e-3)
    s.set_data(0, pr["d1"], None, -0.99)
    J, g = s.objective_and_gradient([x])
    g = g[0]
    print("J", J[0], Jo, "g", g[-3:], gJo[-3:], flush=True)
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g, gJo) < 1e-4
    assert rel(g[:-2], gJo[:-2]) < 1e-4
    assert abs(g[-2] - gJo[-2]) / abs(gJo[-2]) < 1e-4
    assert abs(g[-1] - gJo[-1]) / abs(gJo[-1]) < 1e-4


def test_checkpointing_bitwise(rd, oracle_mod):
    """Binomial checkpointing recomputes bit-identical states: the gradient equals store-all."""
    pr, _ = grad_problem(oracle_mod, nt=12)
    print("case checkpoint", flush=True)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    out = []
    for kw in (dict(), dict(checkpoint_slots=3), dict(checkpoint_slots=1), dict(adjoint_store=True)):
        s = setup_gpu(rd, pr, **kw)
        s.set_data(0, pr["d1"], None, -0.99)
        J, g = s.objective_and_gradient([x])
        out.append((J[0], g[0]))
        s.close()
    for J, g in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(g, out[0][1])


def test_ensemble_equals_single(rd, oracle_mod):
    """A batched ensemble computes each member exactly as a plain solve would."""
    from glia_amd import synth
    n, nt = 64, 6
    print("case ensemble", flush=True)
    members = []
    for seed, (kap, rho) in zip((3, 4, 5), ((0.02, 6.0), (0.08, 12.0), (0.05, 9.0))):
        pr = synth.problem(n, seed, nbases=8, nnz=4, kappa=kap, rho=rho)
        pr["d1"] = (pr["wm"] > 0.6).astype(np.float32)
        members.append(pr)
    ens = rd.RDSolver(n, nmembers=3, dt=0.02, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2)
    xs = []
    for m, pr in enumerate(members):
        ens.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        ens.set_coefficients(m, pr["kappa"], pr["rho"])
        ens.set_basis(m, pr["centers"], pr["sigma"])
        ens.set_data(m, pr["d1"], None, -0.99)
        xs.append(np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]]))
    J, g = ens.objective_and_gradient(xs)
    for m, pr in enumerate(members):
        s = rd.RDSolver(n, dt=0.02, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2)
        s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(0, pr["kappa"], pr["rho"])
        s.set_basis(0, pr["centers"], pr["sigma"])
        s.set_data(0, pr["d1"], None, -0.99)
        J1, g1 = s.objective_and_gradient([xs[m]])
        assert J1[0] == J[m]
        assert np.array_equal(g1[0], g[m])
        s.close()


def test_objective_invariants(rd):
    """Reference src/test/grad.cpp:7-80: zero update gives zero coefficients, J ~ 0 at the truth,
    J > 0 with zero data."""
    n = 64
    print("case invariants", flush=True)
    s = rd.RDSolver(n, dt=0.01, nt=100, k_gm_wm=0.0, r_gm_wm=0.0)
    s.set_tissue_sinusoidal(0)
    s.set_coefficients(0, 0.0, 0.0)
    assert np.linalg.norm(s.get_field(0, "k")) == 0 and np.linalg.norm(s.get_field(0, "rho")) == 0
    s.set_field(0, "c0", gauss_test_function(n))
    x = [np.array([0.01, 0.01])]
    s.set_data(0, np.zeros((n, n, n), np.float32), None, -0.99)
    J0 = s.objective(x, mode=1)[0]
    assert J0 > 0
    s.set_coefficients(0, 0.01, 0.01)
    s.precFactor()
    s.solve_state(0)
    d = s.get_field(0, "ct")
    s.set_data(0, d, None, -0.99)
    J = s.objective(x, mode=1)[0]
    assert J < 1e-7


def test_group_of_uneven_grids_equals_separate(rd, oracle_mod):
    """Concurrent evaluation of handles with different grids (glia_rd_objective_and_gradient_group)
    gives bit-identical J and dJ to evaluating each handle alone.  Fresh handles on both sides: an
    evaluation's preconditioner depends on the solver context's dt, which the first solve changes."""
    print("case group", flush=True)
    from glia_amd import synth
    specs = ((32, 5, 2), (64, 6, 1))
    problems = [[synth.problem(n, seed + m, nbases=4, nnz=None) for m in range(nm)] for n, seed, nm in specs]

    def handles():
        out = []
        for (n, _, nm), members in zip(specs, problems):
            s = rd.RDSolver(n, nmembers=nm, dt=0.05, nt=6, k_gm_wm=0.2, r_gm_wm=0.2)
            for m, pr in enumerate(members):
                s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
                s.set_coefficients(m, pr["kappa"], pr["rho"])
                s.set_basis(m, pr["centers"], pr["sigma"])
                s.set_data(m, (pr["wm"] > 0.5).astype(np.float32), None, -0.99)
            out.append(s)
        return out

    xs = [[np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]]) for pr in members] for members in problems]
    group = handles()
    Jg, gg, ms = rd.objective_and_gradient_group(group, xs)
    assert np.all(ms > 0)
    for s, x, J, g in zip(handles(), xs, Jg, gg):
        J1, g1 = s.objective_and_gradient(x)
        assert np.array_equal(J, J1)
        for a, b in zip(g, g1):
            assert np.array_equal(a, b)
        s.close()
    for s in group:
        s.close()


def test_graph_replay_equals_stream_launches(rd, oracle_mod):
    """CG iterations replayed from captured HIP graphs (many replays, with the solver's own eager
    work and synchronizations between them) give bit-identical J and dJ to launching every kernel
    from the host."""
    print("case graphs", flush=True)
    pr, _ = grad_problem(oracle_mod, nt=6)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    out = []
    for graphs in (False, True):
        s = setup_gpu(rd, pr)
        s.use_graphs(graphs)
        s.set_data(0, pr["d1"], None, -0.99)
        res = [s.objective_and_gradient([x]) for _ in range(2)]
        out.append(res)
        s.close()
    for (J0, g0), (J1, g1) in zip(out[0], out[1]):
        assert J0[0] == J1[0]
        assert np.array_equal(g0[0], g1[0])


def test_hessian_vs_oracle(rd, oracle_mod):
    """Gauss-Newton Hessian matvec in p (evaluateHessian without diffusivity inversion) against the
    FP64 oracle at the same linearization point (1e-4 relative L2, the gradient bar), and
    binomial checkpointing bit-identical to store-all."""
    pr, o = grad_problem(oracle_mod, nt=10)
    print("case hessian", flush=True)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    v = np.random.default_rng(31).standard_normal(len(pr["p"]))
    o.set_data(pr["d1"], None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(pr["dt"], pr["nt"])
    o.objective_and_gradient(x, mode=2)
    ho = o.hessian_p(v)
    out = []
    for kw in (dict(), dict(checkpoint_slots=3)):
        s = setup_gpu(rd, pr, beta=1e-3, **kw)
        s.set_data(0, pr["d1"], None, -0.99)
        s.objective_and_gradient([x])
        hv = s.hessian([np.concatenate([v, [0.0, 0.0]])])[0]
        out.append(hv)
        s.close()
    print("hessian rel", rel(out[0][:-2], ho), flush=True)
    assert rel(out[0][:-2], ho) < 1e-4
    assert out[0][-2] == 0.0 and out[0][-1] == 0.0
    assert np.array_equal(out[0], out[1])


def test_hessian_with_diffusivity_vs_oracle(rd, oracle_mod):
    """Gauss-Newton Hessian with the kappa blocks (evaluateHessian with diffusivity inversion,
    nk = 1) against the FP64 oracle (1e-4 relative L2), checkpointing bit-identical to store-all."""
    pr, o = grad_problem(oracle_mod, nt=8)
    print("case hessian diffusivity", flush=True)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    rng = np.random.default_rng(32)
    d = np.concatenate([rng.standard_normal(len(pr["p"])), [0.01, 0.0]])
    o.set_data(pr["d1"], None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(pr["dt"], pr["nt"])
    o.objective_and_gradient(x, mode=2)
    ho = o.hessian(d, diffusivity=True)
    out = []
    for kw in (dict(), dict(checkpoint_slots=3)):
        s = setup_gpu(rd, pr, beta=1e-3, **kw)
        s.set_data(0, pr["d1"], None, -0.99)
        s.objective_and_gradient([x])
        out.append(s.hessian([d], diffusivity=True)[0])
        s.close()
    print("hessian diffusivity rel", rel(out[0], ho), out[0][-2], ho[-2], flush=True)
    assert rel(out[0][:-2], ho[:-2]) < 1e-4
    assert abs(out[0][-2] - ho[-2]) <= 1e-4 * abs(ho[-2])
    assert out[0][-1] == 0.0
    assert np.array_equal(out[0], out[1])


def test_gradient_non_cubic_grid_vs_oracle(rd, oracle_mod):
    """J and dJ = [g_p | g_kappa | g_rho] on a 48 x 96 x 64 grid (every axis a different length
    and FFT plan) against the FP64 oracle."""
    from glia_amd import synth
    n = (48, 96, 64)
    print("case gradient 48x96x64", flush=True)
    pr = synth.problem(n, 3, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(n=n, nt=6, dt=0.02, k_gm_wm=0.2, r_gm_wm=0.2)
    o = setup_oracle(oracle_mod, pr)
    o.set_c0(o.phi_apply(pr["p"] * 1.3))
    o.solve_state(0)
    d1 = (o.ct() > 0.4).astype(np.float64)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    o.set_data(d1, None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(pr["dt"], pr["nt"])
    Jo, go = o.objective_and_gradient(x, mode=2)
    s = setup_gpu(rd, pr, beta=1e-3)
    s.set_data(0, d1, None, -0.99)
    J, g = s.objective_and_gradient([x])
    s.close()
    print("J", J[0], Jo, "rel g", rel(g[0], go), flush=True)
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4
    # g_rho nearly cancels here (|g_rho| ~ 3e-5 |g|): each coefficient entry is held to 1e-4 of
    # max(|entry|, 1e-2 |g|)
    for k in (-2, -1):
        assert abs(g[0][k] - go[k]) / max(abs(go[k]), 1e-2 * np.linalg.norm(go)) < 1e-4


def test_gradient_at_zero_vs_oracle(rd, oracle_mod):
    """The gradient at x = 0 (p = 0 so c = 0, kappa = 0 so the diffusion solves return at once),
    which the TIL solves take as their reference norm, in modes 0 and 2."""
    pr, o = grad_problem(oracle_mod, nt=6)
    print("case gradient at zero", flush=True)
    o.set_data(pr["d1"], None, -0.99)
    o.set_opt(beta=1e-3)
    s = setup_gpu(rd, pr, beta=1e-3)
    s.set_data(0, pr["d1"], None, -0.99)
    x = np.zeros(len(pr["p"]) + 2)
    for mode in (0, 2):
        o.set_time(pr["dt"], pr["nt"])
        o.set_rho(pr["rho"], pr["r_gm_wm"], 0.0)
        s.set_coefficients(0, pr["kappa"], pr["rho"])
        Jo, go = o.objective_and_gradient(x, mode=mode)
        J, g = s.objective_and_gradient([x], mode=mode)
        print("mode", mode, "J", J[0], Jo, "rel g", rel(g[0], go), flush=True)
        assert abs(J[0] - Jo) / abs(Jo) < 1e-5
        assert rel(g[0], go) < 1e-4
    s.close()


def test_gradient_on_192_point_lines_vs_oracle(rd, oracle_mod):
    """J and dJ on a 32 x 32 x 192 grid: every Z-pass kernel on 96-point packed lines (the 3 * 2^k
    line plan with 16 lines per block) against the FP64 oracle."""
    from glia_amd import synth
    n = (32, 32, 192)
    print("case gradient 32x32x192", flush=True)
    pr = synth.problem(n, 4, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(n=n, nt=4, dt=0.02, k_gm_wm=0.2, r_gm_wm=0.2)
    o = setup_oracle(oracle_mod, pr)
    o.set_c0(o.phi_apply(pr["p"] * 1.3))
    o.solve_state(0)
    d1 = (o.ct() > 0.4).astype(np.float64)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    o.set_data(d1, None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(pr["dt"], pr["nt"])
    Jo, go = o.objective_and_gradient(x, mode=2)
    s = setup_gpu(rd, pr, beta=1e-3)
    s.set_data(0, d1, None, -0.99)
    J, g = s.objective_and_gradient([x])
    s.close()
    print("J", J[0], Jo, "rel g", rel(g[0], go), flush=True)
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4
    for k in (-2, -1):
        assert abs(g[0][k] - go[k]) / max(abs(go[k]), 1e-2 * np.linalg.norm(go)) < 1e-4


def test_forward_on_192_point_columns_vs_oracle(rd, oracle_mod):
    """c(T) on a 192 x 32 x 32 grid: the CG matvec's X passes with the 192-point column plan."""
    from glia_amd import synth
    n = (192, 32, 32)
    print("case forward 192x32x32", flush=True)
    pr = synth.problem(n, 5, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(n=n, nt=6, dt=0.02, k_gm_wm=0.2, r_gm_wm=0.2)
    s = setup_gpu(rd, pr)
    o = setup_oracle(oracle_mod, pr)
    s.phi_apply([pr["p"]])
    s.precFactor()
    s.solve_state(0)
    ct = s.get_field(0, "ct")
    st = s.stats(0)
    s.close()
    o.set_c0(o.phi_apply(pr["p"]))
    o.solve_state(0)
    print("c(T) rel err", rel(ct, o.ct()), "its gpu", st["pcg_iterations"], "oracle", o.ksp()["state"], flush=True)
    assert rel(ct, o.ct()) < 1e-5



def test_gradient_on_160_point_lines_vs_oracle(rd, oracle_mod):
    """J and dJ on a 32 x 32 x 160 grid: the Z passes on 80-point packed lines (the radix-5 plan
    20 x 4) against the FP64 oracle's mixed-radix transform."""
    from glia_amd import synth
    n = (32, 32, 160)
    print("case gradient 32x32x160", flush=True)
    pr = synth.problem(n, 4, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(n=n, nt=4, dt=0.02, k_gm_wm=0.2, r_gm_wm=0.2)
    o = setup_oracle(oracle_mod, pr)
    o.set_c0(o.phi_apply(pr["p"] * 1.3))
    o.solve_state(0)
    d1 = (o.ct() > 0.4).astype(np.float64)
    x = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    o.set_data(d1, None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(pr["dt"], pr["nt"])
    Jo, go = o.objective_and_gradient(x, mode=2)
    s = setup_gpu(rd, pr, beta=1e-3)
    s.set_data(0, d1, None, -0.99)
    J, g = s.objective_and_gradient([x])
    s.close()
    print("J", J[0], Jo, "rel g", rel(g[0], go), flush=True)
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4
    for k in (-2, -1):
        assert abs(g[0][k] - go[k]) / max(abs(go[k]), 1e-2 * np.linalg.norm(go)) < 1e-4


@pytest.mark.parametrize("n", [(160, 32, 32), (32, 160, 32), (80, 80, 32)])
def test_forward_on_radix5_columns_vs_oracle(rd, oracle_mod, n):
    """c(T) with the X and Y passes on 160-point (20 x 8) and 80-point (20 x 4) column plans."""
    from glia_amd import synth
    print("case forward", n, flush=True)
    pr = synth.problem(n, 5, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(n=n, nt=6, dt=0.02, k_gm_wm=0.2, r_gm_wm=0.2)
    s = setup_gpu(rd, pr)
    o = setup_oracle(oracle_mod, pr)
    s.phi_apply([pr["p"]])
    s.precFactor()
    s.solve_state(0)
    ct = s.get_field(0, "ct")
    st = s.stats(0)
    s.close()
    o.set_c0(o.phi_apply(pr["p"]))
    o.solve_state(0)
    print("c(T) rel err", rel(ct, o.ct()), "its gpu", st["pcg_iterations"], "oracle", o.ksp()["state"], flush=True)
    assert rel(ct, o.ct()) < 1e-5
