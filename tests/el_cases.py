"""Shared inputs of the elasticity tests: tissue maps with a background and a tumor."""
import numpy as np


def tissue_and_tumor(n, seed=4):
    from glia_amd import synth
    pr = synth.problem(n, seed, nbases=4, kappa=0.05, rho=8.0, sigma=2 * np.pi / 32)
    wm, gm, csf, vt = (np.asarray(pr[k], np.float64) for k in ("wm", "gm", "csf", "vt"))
    bg = np.clip(1.0 - (wm + gm + csf + vt), 0.0, 1.0)
    x = np.arange(n) * 2 * np.pi / n - np.pi
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    c = np.exp(-r2 / 0.3)
    return c, wm, gm, csf, vt, bg


def tumor_force(c):
    """A force along the tumor gradient, the shape of the mass-effect forcing."""
    n = c.shape[0]
    g = np.gradient(c, 2 * np.pi / n)
    return -1e2 * np.stack(g)
