"""Shared inputs of the advection tests: the reference's test velocity (glia_amd.synth, where the
benchmark configuration takes it too) and test function."""
import numpy as np

from glia_amd.synth import reference_velocity  # noqa: F401


def gauss_test_function(n):
    """createTestFunction (src/test/helper.cpp): exp(-(r/R)^2), R = sqrt(2) 2 pi / 64, centred."""
    h = 2 * np.pi / n
    X = np.arange(n) * h - np.pi
    R = np.sqrt(2.0) * 2 * np.pi / 64
    return np.exp(-((X[:, None, None] ** 2 + X[None, :, None] ** 2 + X[None, None, :] ** 2)) / R ** 2)


def compressive(n):
    """A velocity with nonzero divergence and a field whose mass it moves."""
    X = np.arange(n) * 2 * np.pi / n
    vx = np.broadcast_to(0.5 * np.sin(X)[:, None, None], (n, n, n)).copy()
    vy = np.broadcast_to(0.3 * np.cos(X)[None, :, None], (n, n, n)).copy()
    vz = np.zeros((n, n, n))
    c = np.broadcast_to(1.0 + 0.5 * np.cos(X)[:, None, None] + 0.2 * np.sin(2 * X)[None, None, :], (n, n, n)).copy()
    return [vx, vy, vz], c
