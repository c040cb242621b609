"""Seeded synthetic inputs of the benchmark configurations (no real data is used or shipped).

Tissue: white matter, gray matter and CSF probability maps from nested random ellipsoids, each
smoothed by a periodic Gaussian of sigma = 2 voxels and normalized so the three sum to 1 inside
the brain mask (0 outside).  Gaussian basis centers sit at seeded white-matter voxels.  Everything
is numpy with explicit PCG64 seeds, so every process (tests, bench, smoke) builds identical inputs.
"""
import numpy as np


def _ellipsoid(n, center, radii):
    idx = [np.arange(v, dtype=np.float64) for v in n]
    X = ((idx[0] - center[0]) / radii[0])[:, None, None] ** 2
    Y = ((idx[1] - center[1]) / radii[1])[None, :, None] ** 2
    Z = ((idx[2] - center[2]) / radii[2])[None, None, :] ** 2
    return (X + Y + Z) <= 1.0


try:
    import scipy.fft as _fft

    def _rfftn(a):
        return _fft.rfftn(a, workers=-1)

    def _irfftn(a, n):
        return _fft.irfftn(a, s=n, axes=(0, 1, 2), workers=-1)
except ImportError:  # numpy only
    def _rfftn(a):
        return np.fft.rfftn(a)

    def _irfftn(a, n):
        return np.fft.irfftn(a, s=n, axes=(0, 1, 2))


def gaussian_smooth(f, sigma_vox):
    """Periodic Gaussian smoothing (sigma in voxels) through the FFT."""
    n = f.shape
    F = _rfftn(f)
    k = [np.fft.fftfreq(n[0]) * n[0], np.fft.fftfreq(n[1]) * n[1], np.fft.rfftfreq(n[2]) * n[2]]
    w2 = ((2 * np.pi * k[0] / n[0]) ** 2)[:, None, None] + ((2 * np.pi * k[1] / n[1]) ** 2)[None, :, None] \
        + ((2 * np.pi * k[2] / n[2]) ** 2)[None, None, :]
    F *= np.exp(-0.5 * sigma_vox ** 2 * w2)
    return _irfftn(F, n)


def brain(n, seed):
    """(wm, gm, csf, vt) float64 maps on an n^3 (or n0 x n1 x n2) grid."""
    if np.isscalar(n):
        n = (int(n),) * 3
    rng = np.random.default_rng([int(seed), 1001])
    N = np.array(n, dtype=np.float64)
    center = N / 2 + rng.uniform(-0.02, 0.02, 3) * N
    rb = np.array([0.36, 0.30, 0.34]) * N * rng.uniform(0.95, 1.05, 3)
    B = _ellipsoid(n, center, rb)
    I = _ellipsoid(n, center, rb * 0.90)
    W = _ellipsoid(n, center + rng.uniform(-0.02, 0.02, 3) * N, rb * rng.uniform(0.55, 0.65, 3))
    off = np.array([0.06, 0.0, 0.0]) * N
    rv = rb * np.array([0.16, 0.10, 0.12]) * rng.uniform(0.9, 1.1, 3)
    V = _ellipsoid(n, center + off, rv) | _ellipsoid(n, center - off, rv)
    sig = 2.0
    wm = gaussian_smooth((W & ~V).astype(np.float64), sig)
    gm = gaussian_smooth((I & ~W & ~V).astype(np.float64), sig)
    csf = gaussian_smooth(((B & ~I) | V).astype(np.float64), sig)
    mask = gaussian_smooth(B.astype(np.float64), sig) > 0.5
    wm, gm, csf = (np.clip(a, 0.0, None) for a in (wm, gm, csf))
    s = wm + gm + csf
    ok = mask & (s > 1e-6)
    for a in (wm, gm, csf):
        a[ok] /= s[ok]
        a[~ok] = 0.0
    return wm, gm, csf, np.zeros(n)


def basis_centers(wm, nb, seed, thr=0.5):
    """nb centers (x, y, z in [0, 2pi)) at distinct seeded voxels with wm > thr."""
    n = wm.shape
    rng = np.random.default_rng([int(seed), 2002])
    cand = np.flatnonzero(wm.reshape(-1) > thr)
    pick = rng.choice(cand, size=nb, replace=False)
    ijk = np.stack(np.unravel_index(pick, n), axis=1).astype(np.float64)
    h = 2 * np.pi / np.array(n, dtype=np.float64)
    return ijk * h


def sparse_p(nb, nnz, seed):
    rng = np.random.default_rng([int(seed), 3003])
    p = np.zeros(nb)
    idx = rng.choice(nb, size=nnz, replace=False)
    p[idx] = rng.uniform(0.0, 1.0, nnz)
    return p


def problem(n, seed, nbases=8, nnz=None, kappa=0.05, rho=8.0, sigma=2 * np.pi / 64):
    """One member: tissue, basis centers, coefficients p (nnz nonzero, all if None), kappa, rho."""
    wm, gm, csf, vt = brain(n, seed)
    centers = basis_centers(wm, nbases, seed)
    if nnz is None:
        rng = np.random.default_rng([int(seed), 4004])
        p = rng.uniform(0.0, 1.0, nbases)
    else:
        p = sparse_p(nbases, nnz, seed)
    return dict(n=wm.shape, wm=wm, gm=gm, csf=csf, vt=vt, centers=centers, p=p, kappa=kappa, rho=rho, sigma=sigma)


def ensemble_params(count, seed):
    """kappa ~ U(0.01, 0.1) and rho ~ U(4, 15) per member."""
    rng = np.random.default_rng([int(seed), 5005])
    return rng.uniform(0.01, 0.1, count), rng.uniform(4.0, 15.0, count)


def tissue_and_tumor(n, seed=4):
    """Tissue maps with a background and a centred Gaussian tumor (the inputs of the elasticity and
    mass-effect configurations): c, wm, gm, csf, vt, bg on an n (or n0 x n1 x n2) grid."""
    shape = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
    pr = problem(shape, seed, nbases=4, kappa=0.05, rho=8.0, sigma=2 * np.pi / 32)
    wm, gm, csf, vt = (np.asarray(pr[k], np.float64) for k in ("wm", "gm", "csf", "vt"))
    bg = np.clip(1.0 - (wm + gm + csf + vt), 0.0, 1.0)
    x = [np.arange(m) * 2 * np.pi / m - np.pi for m in shape]
    r2 = x[0][:, None, None] ** 2 + x[1][None, :, None] ** 2 + x[2][None, None, :] ** 2
    c = np.exp(-r2 / 0.3)
    return c, wm, gm, csf, vt, bg


def tumor_force(c):
    """A force along the tumor gradient, the shape of the mass-effect forcing (3 components)."""
    h = [2 * np.pi / m for m in c.shape]
    g = np.gradient(c, *h)
    return -1e2 * np.stack(g)


def mass_effect_case(n, symmetric=False, amplitude=0.5):
    """Inputs of the mass-effect forwards: tissue with all four classes (part of the CSF map as
    ventricles), a tumor seed of the given amplitude and a background that closes the partition of
    unity; c, wm, gm, vt, csf, bg rounded to FP32."""
    c, wm, gm, csf, vt, bg = tissue_and_tumor(n)
    vt = 0.4 * csf
    csf = 0.6 * csf
    if symmetric:  # invariant under the exchange of the x and y axes
        sym = lambda a: 0.5 * (a + a.transpose(1, 0, 2))  # noqa: E731
        wm, gm, csf, vt = sym(wm), sym(gm), sym(csf), sym(vt)
    bg = np.clip(1.0 - (wm + gm + csf + vt), 0.0, 1.0)
    c = amplitude * c
    return [a.astype(np.float32).astype(np.float64) for a in (c, wm, gm, vt, csf, bg)]


def reference_velocity(n):
    """createTestField (src/test/helper.cpp:47-80): frequency-4 sinusoids around (1, 1.5, 1)."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    h = 2 * np.pi / np.array(n)
    X, Y, Z = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    a, b, c = 4 * h[0] * X, 4 * h[1] * Y, 4 * h[2] * Z
    vx = 1 + 0.25 * np.sin(a) * np.sin(b) * np.sin(c)
    vy = 1.5 + 0.1 * np.sin(a) * np.cos(b) * np.sin(c)
    vz = 1 + np.cos(a) * np.cos(b) * np.sin(c)
    return [vx, vy, vz]
