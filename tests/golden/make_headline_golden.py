"""Writes the FP64 oracle's results on the headline grids as small fixtures (data only):

  headline_128_nt50.npz  the 128^3 gradient configuration: nt = 50 at dt = 0.02, 64 Gaussian bases,
                         p 8-sparse U(0, 1), observed data d1 = 1[c*(T) > 0.5] from a seeded
                         ground-truth p*, O = 1 everywhere (obs threshold -0.99)
  headline_256_nt2.npz   one 256^3 member of the differentiation configuration shortened to nt = 2
                         (dt = 0.01, 128 bases, p 8-sparse), data as above

Each file holds the inputs' identity (seeds, sizes, checksums of the generated tissue and basis
centers), d1 as packed bits, the oracle's J and dJ = [g_p | g_kappa | g_rho] (mode 2: p, kappa
and rho in one pass), ||c(T)||_2 and c(T) on a regular 32^3 sub-lattice (every 4th or 8th point of
each axis), rounded to FP32.  The oracle takes minutes at these sizes, so the GPU tests read the
fixtures instead of running it.

    python tests/golden/make_headline_golden.py [128] [256]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# name: (n, nt, dt, nbases, nnz, seed of the member, seed of p*, sub-lattice stride)
CASES = {
    "128": ("headline_128_nt50.npz", 128, 50, 0.02, 64, 8, 202, 1202, 4),
    "256": ("headline_256_nt2.npz", 256, 2, 0.01, 128, 8, 700, 1700, 8),
}
KAPPA, RHO, K_GM_WM, R_GM_WM, BETA = 0.05, 8.0, 0.2, 0.2, 1e-3


def problem(n, nb, nnz, seed, seed_star):
    from glia_amd import synth
    pr = synth.problem(n, seed, nbases=nb, nnz=nnz, kappa=KAPPA, rho=RHO, sigma=2 * np.pi / 64)
    pr["pstar"] = synth.sparse_p(nb, nnz, seed_star)
    return pr


def identity(pr):
    """Checksums that tie a fixture to the synthetic inputs it was computed from."""
    return np.array([pr["wm"].sum(), pr["gm"].sum(), pr["csf"].sum(), (pr["wm"] ** 2).sum(),
                     pr["centers"].sum(), (pr["centers"] ** 2).sum()])


def make(key):
    from oracle import oracle
    fname, n, nt, dt, nb, nnz, seed, seed_star, stride = CASES[key]
    t0 = time.time()
    pr = problem(n, nb, nnz, seed, seed_star)
    o = oracle.Oracle(n)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(dt, nt)
    o.set_k(KAPPA, K_GM_WM, 0.0)
    o.set_rho(RHO, R_GM_WM, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    print("%s: setup %.1f s" % (fname, time.time() - t0), flush=True)
    o.set_c0(o.phi_apply(pr["pstar"]))
    o.solve_state(0)
    d1 = (o.ct() > 0.5)
    print("%s: data %.1f s, %d observed voxels" % (fname, time.time() - t0, int(d1.sum())), flush=True)
    o.set_data(d1.astype(np.float64), None, -0.99)
    o.set_opt(beta=BETA)
    o.set_time(dt, nt)  # a fresh solver context, as a new DiffusionSolver would have
    x = np.concatenate([pr["p"], [KAPPA, RHO]])
    J, dJ = o.objective_and_gradient(x, mode=2)
    ct = o.ct()
    ksp = o.ksp()
    print("%s: gradient %.1f s, J %.10e, its %s" % (fname, time.time() - t0, J, ksp), flush=True)
    np.savez_compressed(
        os.path.join(HERE, fname),
        n=n, nt=nt, dt=dt, nbases=nb, nnz=nnz, seed=seed, seed_star=seed_star, stride=stride,
        kappa=KAPPA, rho=RHO, k_gm_wm=K_GM_WM, r_gm_wm=R_GM_WM, beta=BETA,
        identity=identity(pr), p=pr["p"], pstar=pr["pstar"],
        d1_bits=np.packbits(d1.reshape(-1)),
        J=J, dJ=dJ, ct_norm=float(np.linalg.norm(ct)),
        ct_sub=ct[::stride, ::stride, ::stride].astype(np.float32),
        its_state=ksp["state"], its_adjoint=ksp["adjoint"])


def main():
    keys = sys.argv[1:] or list(CASES)
    for k in keys:
        make(k)


if __name__ == "__main__":
    main()
