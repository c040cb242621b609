"""Shared inputs of the mass-effect forward tests: tissue with all four classes, a tumor seed and
a background that closes the partition of unity."""
import numpy as np

from el_cases import tissue_and_tumor


def mass_effect_case(n, symmetric=False, amplitude=0.5):
    c, wm, gm, csf, vt, bg = tissue_and_tumor(n)
    vt = 0.4 * csf
    csf = 0.6 * csf
    if symmetric:  # invariant under the exchange of the x and y axes
        sym = lambda a: 0.5 * (a + a.transpose(1, 0, 2))  # noqa: E731
        wm, gm, csf, vt = sym(wm), sym(gm), sym(csf), sym(vt)
    bg = np.clip(1.0 - (wm + gm + csf + vt), 0.0, 1.0)
    c = amplitude * c
    return [a.astype(np.float32).astype(np.float64) for a in (c, wm, gm, vt, csf, bg)]
