"""The torch.autograd binding of the gradient (glia_amd.autograd) on the GPU.

The differentiation configuration at a test-sized grid: 64^3, nt = 100 steps of dt = 0.01, 128
Gaussian bases, gradient w.r.t. p, kappa and rho.  Checks, in order of strictness:
  * x.grad from J.backward() equals the gradient of RDSolver.objective_and_gradient exactly;
  * binomial checkpointing with 10 stored states gives bit-identical J and dJ to store-all;
  * dJ agrees with the FP64 oracle's adjoint gradient (the reference algorithm) within 1e-4
    relative L2 -- the gradient bar of the FP32-bulk design;
  * central finite differences of J (h = 1e-3) along 10 random unit directions agree with
    <dJ, v>: |fd - <dJ, v>| <= 2e-2 |<dJ, v>| + 1e-3 ||dJ||.  The floor covers the FP32 bulk and
    the PCG tolerance (rtol 1e-6), whose noise in J is amplified by 1 / (2h).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NT, DT, NB = 64, 100, 0.01, 128


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def problem(oracle_mod):
    from glia_amd import synth
    pr = synth.problem(N, 11, nbases=NB, nnz=8, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    pr.update(nt=NT, dt=DT, k_gm_wm=0.2, r_gm_wm=0.2)
    # observed data: thresholded c(T) of a ground truth with other coefficients (oracle, FP64)
    o = oracle_mod.Oracle(N)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(DT, NT)
    o.set_k(0.06, 0.2, 0.0)
    o.set_rho(9.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_c0(o.phi_apply(synth.sparse_p(NB, 8, 12)))
    o.solve_state(0)
    pr["d1"] = (o.ct() > 0.5).astype(np.float64)
    pr["x"] = np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])
    return pr


def make_solver(pr, **kw):
    from glia_amd import rd
    s = rd.RDSolver(N, dt=DT, nt=NT, k_gm_wm=0.2, r_gm_wm=0.2, beta=1e-4, **kw)
    s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    s.set_coefficients(0, pr["kappa"], pr["rho"])
    s.set_basis(0, pr["centers"], pr["sigma"])
    s.set_data(0, pr["d1"], None, -0.99)
    return s


def test_autograd_backward_equals_gradient(problem):
    import torch
    from glia_amd.autograd import rd_objective
    print("case autograd backward", flush=True)
    # two fresh handles: the preconditioner of an evaluation depends on the solver context's dt,
    # which (as in the reference) is the last CN step size after the first solve
    s = make_solver(problem)
    x = torch.tensor(problem["x"][None, :], dtype=torch.float64, requires_grad=True)
    J = rd_objective(x, s)
    J.sum().backward()
    s2 = make_solver(problem)
    J2, g2 = s2.objective_and_gradient([problem["x"]])
    assert float(J.detach()[0]) == J2[0]
    assert np.array_equal(x.grad.numpy()[0], g2[0])
    s.close()
    s2.close()


def test_checkpointing_10_states_bitwise(problem):
    print("case checkpoint 10 of 100", flush=True)
    out = []
    for kw in (dict(), dict(checkpoint_slots=10)):
        s = make_solver(problem, **kw)
        J, g = s.objective_and_gradient([problem["x"]])
        out.append((J[0], g[0]))
        s.close()
    assert out[1][0] == out[0][0]
    assert np.array_equal(out[1][1], out[0][1])


def test_gradient_vs_oracle_nt100(problem, oracle_mod):
    print("case gradient vs oracle nt=100", flush=True)
    pr = problem
    o = oracle_mod.Oracle(N)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(DT, NT)
    o.set_k(pr["kappa"], 0.2, 0.0)
    o.set_rho(pr["rho"], 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_data(pr["d1"], None, -0.99)
    o.set_opt(beta=1e-4)
    o.set_time(DT, NT)
    Jo, go = o.objective_and_gradient(pr["x"], mode=2)
    s = make_solver(pr, checkpoint_slots=10)
    J, g = s.objective_and_gradient([pr["x"]])
    print("J", J[0], Jo, "rel g", rel(g[0], go), flush=True)
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4
    s.close()


def test_finite_differences_10_directions(problem):
    import torch
    from glia_amd.autograd import rd_objective
    print("case finite differences", flush=True)
    s = make_solver(problem, checkpoint_slots=10)
    x0 = problem["x"]
    s.objective([x0])  # every later evaluation then starts from the same solver-context dt
    x = torch.tensor(x0[None, :], dtype=torch.float64, requires_grad=True)
    rd_objective(x, s).sum().backward()
    g = x.grad.numpy()[0]
    rng = np.random.default_rng(2024)
    h = 1e-3
    for k in range(10):
        v = rng.standard_normal(x0.size)
        v /= np.linalg.norm(v)
        jp = s.objective([x0 + h * v])[0]
        jm = s.objective([x0 - h * v])[0]
        fd = (jp - jm) / (2 * h)
        ad = float(g @ v)
        print("direction %d: fd %.6e adjoint %.6e" % (k, fd, ad), flush=True)
        assert abs(fd - ad) <= 2e-2 * abs(ad) + 1e-3 * np.linalg.norm(g)
    s.close()
