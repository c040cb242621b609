"""The rest of the reference's PdeOperators / DerivativeOperators surface through the C ABI, against
the FP64 oracle: solveState(1) and solveState(2) (the tangent-linear equations,
src/pde/PdeOperators.cpp:235-316), solveAdjoint(2) (the incremental adjoint, :372-420), the public
time history c_, p_, c_half_ (include/pde/PdeOperators.h:24-29), gradDiffusion / gradReaction
(src/grad/DerivativeOperators.cpp:189-321), resizeTimeHistory (:56-103 of PdeOperators.cpp) and
the checkGradient Taylor test (src/grad/DerivativeOperators.cpp:369-450).

Bars: fields 1e-5 relative L2 (the c(T) bar), quadratures 1e-4 (the gradient bar)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NT, DT = 64, 6, 0.02


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def case(oracle_mod):
    """A 64^3 problem at its linearization point: x, data d1 from 1.3 p, and the oracle after one
    evaluation (state and adjoint histories in place)."""
    from glia_amd import synth
    pr = synth.problem(N, 31, nbases=6, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64)
    o = oracle_mod.Oracle(N)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(DT, NT, True)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_c0(o.phi_apply(pr["p"] * 1.3))
    o.solve_state(0)
    d1 = (o.ct() > 0.4).astype(np.float64)
    x = np.concatenate([pr["p"], [0.05, 8.0]])
    o.set_data(d1, None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(DT, NT, True)
    Jo, go = o.objective_and_gradient(x, mode=2)
    return dict(pr=pr, o=o, d1=d1, x=x, Jo=Jo, go=go)


def gpu_solver(rd, case, **kw):
    pr = case["pr"]
    s = rd.RDSolver(N, dt=DT, nt=NT, k_gm_wm=0.2, r_gm_wm=0.2, beta=1e-3, **kw)
    s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    s.set_coefficients(0, 0.05, 8.0)
    s.set_basis(0, pr["centers"], pr["sigma"])
    s.set_data(0, case["d1"], None, -0.99)
    return s


@pytest.fixture(scope="module")
def rd():
    from glia_amd import rd as rdmod
    return rdmod


def test_history_and_grad_quadratures(rd, case):
    """After an evaluation: c_[i], c_half_[i] and p_[i] against the oracle's histories; gradDiffusion
    and gradReaction equal the evaluation's g_kappa and g_rho; with checkpointing the stored
    indices of c_ are available and the others answer GLIA_RD_ERR_NOHISTORY."""
    o = case["o"]
    print("case history", flush=True)
    s = gpu_solver(rd, case, adjoint_store=True)
    s.keep_adjoint_history(True)
    J, g = s.objective_and_gradient([case["x"]])
    print("J rel %.2e dJ rel %.2e" % (abs(J[0] - case["Jo"]) / case["Jo"], rel(g[0], case["go"])), flush=True)
    assert rel(g[0], case["go"]) < 1e-4
    for i in (0, 1, NT // 2, NT):
        ec, ep = rel(s.history(0, "state", i), o.c_hist(i)), rel(s.history(0, "adjoint", i), o.p_hist(i))
        print("i %d: c_ %.2e p_ %.2e" % (i, ec, ep), flush=True)
        assert ec < 1e-5 and ep < 1e-5
    for i in (0, NT - 1):
        assert rel(s.history(0, "half", i), o.c_half(i)) < 1e-5
    # p_[0] is alpha(0), c_[nt] is c(T), bit for bit
    assert np.array_equal(s.history(0, "adjoint", 0), s.get_field(0, "alpha0"))
    assert np.array_equal(s.history(0, "state", NT), s.get_field(0, "ct"))
    assert s.grad_diffusion()[0] == g[0][-2] and s.grad_reaction()[0] == g[0][-1]
    s.close()
    # checkpointing: c_[0] is stored, an unstored index is reported as such
    s = gpu_solver(rd, case, checkpoint_slots=2)
    s.objective_and_gradient([case["x"]])
    assert rel(s.history(0, "state", 0), o.c_hist(0)) < 1e-5
    with pytest.raises(rd.GliaError) as ei:
        s.history(0, "state", NT - 1)
    assert ei.value.code == -6
    with pytest.raises(rd.GliaError):
        s.keep_adjoint_history(True)  # store-all mode only
    s.close()


def test_explicit_state_adjoint_and_quadratures_vs_oracle(rd, oracle_mod, case):
    """solveState(0) then solveAdjoint(1) called one by one (as PdeOperatorsRD users do), then
    gradDiffusion / gradReaction, against the oracle's quadratures (the evaluation's g_kappa,
    g_rho) and alpha(0) (the oracle's p_[0]: its evaluation adds -beta c(0) to alpha(0) afterwards
    for Phi^T)."""
    print("case explicit solves", flush=True)
    o = case["o"]
    s = gpu_solver(rd, case)
    s.objective([case["x"]])  # the same coefficients and solver-context dt as the evaluation
    s.set_field(0, "c0", o.c0())
    s.solve_state(0)
    s.solve_adjoint(1)
    gk, gr = s.grad_diffusion()[0], s.grad_reaction()[0]
    print("g_kappa %.6e (oracle %.6e) g_rho %.6e (oracle %.6e)" % (gk, case["go"][-2], gr, case["go"][-1]), flush=True)
    assert abs(gk - case["go"][-2]) <= 1e-4 * abs(case["go"][-2])
    assert abs(gr - case["go"][-1]) <= 1e-4 * max(abs(case["go"][-1]), 1e-2 * np.linalg.norm(case["go"]))
    assert rel(s.get_field(0, "alpha0"), o.p_hist(0)) < 1e-5
    s.close()


@pytest.mark.parametrize("linearized", [1, 2])
def test_tangent_linear_state_and_incremental_adjoint_vs_oracle(rd, oracle_mod, case, linearized):
    """solveState(1|2) from c~0 around the state history, then solveAdjoint(2) from -O^T O c~(T),
    against the oracle (solveState(2) with the secondary coefficient kt = 0.01)."""
    print("case tangent", linearized, flush=True)
    pr = case["pr"]
    o = oracle_mod.Oracle(N)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(DT, NT)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_data(case["d1"], None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(DT, NT)
    o.objective_and_gradient(case["x"], mode=2)
    rng = np.random.default_rng(5)
    ct0 = o.phi_apply(rng.uniform(-1, 1, 6)) if linearized == 1 else np.zeros((N, N, N))
    s = gpu_solver(rd, case)
    s.objective_and_gradient([case["x"]])
    if linearized == 2:
        o.set_secondary(0.01)
        s.set_secondary(0, 0.01)
    o.set_c0(ct0)
    o.solve_state(linearized)
    s.set_field(0, "c0", ct0)
    s.solve_state(linearized)
    ect = rel(s.get_field(0, "ct"), o.ct())
    obs = np.ones((N, N, N))
    o.set_pt(-(obs * obs * o.ct()))
    o.solve_adjoint(2)
    s.solve_adjoint(2)
    ea = rel(s.get_field(0, "alpha0"), o.p0())
    print("c~(T) rel %.2e  a~(0) rel %.2e" % (ect, ea), flush=True)
    assert ect < 1e-5
    assert ea < 1e-5
    s.close()


def test_resize_time_history_equals_fresh_handle(rd, case):
    """resizeTimeHistory (set_time) to nt = 9 gives bit for bit the evaluation of a handle created
    with nt = 9 (one warm-up evaluation on each first: the preconditioner uses the solver-context
    dt that the first solve sets)."""
    print("case resize", flush=True)
    a = gpu_solver(rd, case)
    a.objective_and_gradient([case["x"]])
    a.set_time(DT, 9)
    a.objective_and_gradient([case["x"]])
    Ja, ga = a.objective_and_gradient([case["x"]])
    pr = case["pr"]
    b = rd.RDSolver(N, dt=DT, nt=9, k_gm_wm=0.2, r_gm_wm=0.2, beta=1e-3)
    b.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    b.set_coefficients(0, 0.05, 8.0)
    b.set_basis(0, pr["centers"], pr["sigma"])
    b.set_data(0, case["d1"], None, -0.99)
    b.objective_and_gradient([case["x"]])
    Jb, gb = b.objective_and_gradient([case["x"]])
    assert Ja[0] == Jb[0] and np.array_equal(ga[0], gb[0])
    a.close()
    b.close()


def test_check_gradient_taylor(rd, case):
    """The Taylor test along a direction in the p block (mode p, kappa: there the gradient is the
    discrete adjoint up to the CG tolerance): the remainder |J(x + h d) - J - h <dJ, d>| falls
    quadratically for h = 1, 0.1, 0.01 and the finite-difference slope at h = 1e-3 matches
    <dJ, d> to 2e-2 (the FP64 oracle gives 4.5e-3 here: O(h) truncation)."""
    print("case taylor", flush=True)
    s = gpu_solver(rd, case)
    s.objective([case["x"]])
    d = np.concatenate([np.random.default_rng(9).uniform(0, 1, 6), [0.0, 0.0]])
    _, g = s.objective_and_gradient([case["x"]], mode=0)
    gd = float(g[0] @ d)
    err, gdiff = s.check_gradient([case["x"]], mode=0, direction=[d])
    print("err", err[0], "gdiff", gdiff[0], "<dJ,d>", gd, flush=True)
    assert err[0][1] < 0.2 * err[0][0] and err[0][2] < 0.2 * err[0][1]
    assert abs(gdiff[0][3]) < 2e-2 * abs(gd)
    # the seeded default direction runs as well
    err2, _ = s.check_gradient([case["x"]], mode=0)
    assert np.all(np.isfinite(err2))
    s.close()
