"""Shared inversion problems of the optimizer tests (CPU and GPU)."""
import numpy as np


def rd_sparse_problem(oracle_mod, nt=4, dt=0.05):
    """32^3, 12 bases, truth with 2 nonzeros normalized to max Phi p = 1, exact data."""
    from glia_amd import synth
    pr = synth.problem(32, 7, nbases=12, nnz=2, kappa=0.05, rho=8.0, sigma=2 * np.pi / 32)
    o = oracle_mod.Oracle(32)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_solver(1e-6, 5000)
    o.set_time(dt, nt)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    p = pr["p"] / o.phi_apply(pr["p"]).max()
    o.set_c0(o.phi_apply(p))
    o.solve_state(0)
    d = o.ct()
    o.set_data(d, None, -0.99)
    o.set_time(dt, nt)
    return pr, o, p, d


def oracle_ops(o, r_gm_wm=0.2):
    def ev(mode, beta, X, want):
        o.set_opt(beta=beta)
        if mode == 0:  # the p stages run with rho = x[np + 1] (the handle takes it from x0 too)
            o.set_rho(X[0][-1], r_gm_wm, 0.0)
        if want:
            f, g = o.objective_and_gradient(X[0], mode=mode)
            return np.array([f]), g[None]
        return np.array([o.objective(X[0], mode=mode)]), None

    def c0(P):
        c = o.phi_apply(P[0])
        o.set_c0(c)
        return np.array([c.max()])
    return ev, c0


SPARSE_RD = dict(sparsity=2, beta=1e-4, reaction_inversion=1, til=dict(max_it=50), rd=dict(max_it=50, gttol=1e-5))

# (kappa, rho) starts of the sparse-TIL runs: the TIL stages use rho as given
SPARSE_STARTS = [(0.1, 8.0), (0.02, 8.0)]
