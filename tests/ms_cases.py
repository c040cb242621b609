"""Random multispecies states in their ranges, as flat float32 arrays (rounded once so that the GPU
and the oracle see the same inputs)."""
import numpy as np


def state(n=20000, seed=6, di_zero=False):
    rng = np.random.default_rng(seed)
    f = lambda lo, hi: rng.uniform(lo, hi, n).astype(np.float32).astype(np.float64)
    p, i, nec = f(0, 0.6), f(0, 0.3), f(0, 0.2)
    ox = f(0, 1)
    gm, wm = f(0, 0.5), f(0, 0.5)
    rho = f(0, 12)
    di = np.zeros(n) if di_zero else f(-1, 1)
    return p, i, nec, ox, gm, wm, rho, di
