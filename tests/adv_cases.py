"""Shared inputs of the advection tests: the reference's test velocity and test function."""
import numpy as np


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
