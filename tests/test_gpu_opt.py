"""The inversion loop on the GPU gradient (glia_rd_invert, include/glia_opt.h): a two-member
kappa/rho inversion (RDOptimizer, src/opt/RDOptimizer.cpp) on exact data reaches the truth from
two starts, and the FP64 oracle driving the same loop from the first start agrees with it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_kappa_rho_inversion_on_gpu_recovers_truth(oracle_mod):
    from glia_amd import opt, rd, synth
    n, nt, dt = 32, 8, 0.05
    print("case invert kappa/rho 32^3 nt 8, 2 members", flush=True)
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
    d = o.ct()
    o.set_data(d, None, -0.99)
    o.set_time(dt, nt)

    s = rd.RDSolver(n, nmembers=2, dt=dt, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2)
    for m in range(2):
        s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(m, 0.05, 8.0)
        s.set_basis(m, pr["centers"], pr["sigma"])
        s.set_data(m, d, None, -0.99)
    s.phi_apply([pr["p"], pr["p"]])
    starts = [np.concatenate([pr["p"], [0.1, 5.0]]), np.concatenate([pr["p"], [0.02, 12.0]])]
    lb = np.concatenate([np.full(6, -np.inf), [1e-3, 1.0]])
    ub = np.concatenate([np.full(6, np.inf), [1.0, 15.0]])
    settings = dict(gatol=1e-9, gttol=1e-5, max_it=50)
    xs, res = opt.invert(s, starts, lb, ub, mode=1, **settings)
    s.close()
    print("gpu", [x[-2:] for x in xs], res, flush=True)
    for x, r in zip(xs, res):
        assert r["reason"] == "gttol", r
        assert abs(x[-2] - 0.05) < 1e-4 and abs(x[-1] - 8.0) < 2e-3, x[-2:]
        assert np.array_equal(x[:-2], pr["p"])

    def fg(X):
        f, g = o.objective_and_gradient(X[0], mode=1)
        return np.array([f]), g[None]
    xo, ro = opt.minimize(fg, starts[0], lb, ub, **settings)
    print("oracle", xo[0][-2:], ro, flush=True)
    assert abs(xs[0][-2] - xo[0][-2]) < 1e-4 and abs(xs[0][-1] - xo[0][-1]) < 2e-3
    assert abs(res[0]["iterations"] - ro[0]["iterations"]) <= 3


def test_sparse_til_on_gpu_matches_oracle_driven_run(oracle_mod):
    """glia_rd_sparse_til (SparseTILOptimizer with the final kappa/rho solve) on a two-member
    handle: both members find the truth's support and coefficients, and member 0 agrees with the
    same driver run on the FP64 oracle."""
    from glia_amd import opt, rd
    from opt_cases import SPARSE_RD, SPARSE_STARTS, oracle_ops, rd_sparse_problem
    print("case sparse TIL 32^3 nt 4, 2 members", flush=True)
    pr, o, p, d = rd_sparse_problem(oracle_mod)
    s = rd.RDSolver(32, nmembers=2, dt=0.05, nt=4, k_gm_wm=0.2, r_gm_wm=0.2)
    for m in range(2):
        s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(m, 0.05, 8.0)
        s.set_basis(m, pr["centers"], pr["sigma"])
        s.set_data(m, d, None, -0.99)
    starts = [np.concatenate([np.zeros(12), st]) for st in SPARSE_STARTS]
    xs, sups, res = opt.rd_sparse_til(s, starts, **SPARSE_RD)
    s.close()
    print("gpu", [np.round(x, 5) for x in xs], sups, res, flush=True)
    want = sorted(np.flatnonzero(p).tolist())
    for x, sup in zip(xs, sups):
        assert sup == want
        assert np.allclose(x[:12], p, atol=3e-3)
        assert abs(x[12] - 0.05) < 5e-4 and abs(x[13] - 8.0) < 0.05, x[12:]
    ev, c0 = oracle_ops(o)
    xo, supo, reso = opt.sparse_til(ev, c0, starts[0], 12, **SPARSE_RD)
    print("oracle", np.round(xo[0], 5), supo, reso, flush=True)
    assert supo[0] == sups[0] and reso[0]["iterations"] == res[0]["iterations"]
    assert np.allclose(xs[0][:12], xo[0][:12], atol=1e-4)
    assert abs(xs[0][12] - xo[0][12]) < 1e-4 and abs(xs[0][13] - xo[0][13]) < 5e-3


def test_til_gauss_newton_on_gpu_matches_oracle_driven_run(oracle_mod):
    """glia_rd_invert with newton_solver = GAUSSNEWTON (TILOptimizer over TAO BNLS,
    src/opt/TILOptimizer.cpp) on p and kappa, rho fixed, from p = 0 and kappa = 0.1: three Newton
    steps on the GPU's Hessian matvec (evaluateHessian with the kappa blocks) agree with the same
    loop over the FP64 oracle, and a full run on the GPU reaches the truth."""
    from glia_amd import opt, rd, synth
    n, nt, dt = 32, 4, 0.05
    print("case Gauss-Newton TIL p/kappa 32^3 nt 4, 2 members", flush=True)
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
    d = o.ct()
    o.set_data(d, None, -0.99)
    o.set_time(dt, nt)
    o.set_opt(beta=1e-6)

    s = rd.RDSolver(n, nmembers=2, dt=dt, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2, beta=1e-6)
    for m in range(2):
        s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(m, 0.05, 8.0)
        s.set_basis(m, pr["centers"], pr["sigma"])
        s.set_data(m, d, None, -0.99)
    x0 = np.concatenate([np.zeros(6), [0.1, 8.0]])
    lb = np.concatenate([np.full(6, -np.inf), [1e-3, 8.0]])
    ub = np.concatenate([np.full(6, np.inf), [1.0, 8.0]])
    settings = dict(gatol=1e-9, gttol=1e-5, newton_solver=0)  # GAUSSNEWTON
    x3, r3 = opt.invert(s, [x0, x0], lb, ub, mode=0, max_it=3, **settings)
    print("gpu 3 steps", [np.round(x, 5) for x in x3], r3, flush=True)
    xs, rs = opt.invert(s, [x0, x0], lb, ub, mode=0, max_it=30, **settings)
    s.close()
    print("gpu", [np.round(x, 5) for x in xs], rs, flush=True)
    for x, r in zip(xs, rs):
        assert r["reason"] == "gttol", r
        assert np.allclose(x[:6], pr["p"], atol=2e-3), x[:6] - pr["p"]
        assert abs(x[6] - 0.05) < 2e-3 and x[7] == 8.0, x[6:]
    assert np.array_equal(x3[0], x3[1]) and r3[0] == r3[1]

    def fg(X):
        f, g = o.objective_and_gradient(X[0], mode=0)
        return np.array([f]), g[None]

    def hv(V):
        return o.hessian(V[0], True)[None]
    xz = x0.copy()
    xz[6] = 0.0
    gref = np.linalg.norm(fg(xz[None])[1][0][:7])  # TILOptimizer's reference gradient at p = 0
    xo, ro = opt.gauss_newton(fg, hv, x0, lb, ub, max_it=3, gnorm_ref=gref, **settings)
    print("oracle 3 steps", np.round(xo[0], 5), ro, flush=True)
    assert ro[0]["iterations"] == r3[0]["iterations"] and ro[0]["fevals"] == r3[0]["fevals"]
    assert abs(ro[0]["gnorm0"] - r3[0]["gnorm0"]) < 1e-4 * ro[0]["gnorm0"]
    assert abs(ro[0]["f"] - r3[0]["f"]) < 1e-4 * ro[0]["f"]
    assert np.allclose(x3[0][:7], xo[0][:7], atol=1e-4), x3[0][:7] - xo[0][:7]
