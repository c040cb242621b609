"""The oracle's semi-Lagrangian advection (SemiLagrangianSolver, src/pde/AdvectionSolver.cpp:279-704)
against what the mathematics fixes: the cubic Lagrange interpolant keeps constants, shifts by whole
cells exactly and reproduces cubics; a constant velocity has the closed-form trajectory -dt v / h;
the mass-conservation mode keeps the total mass where pure advection moves it; and the reference's
own round-trip test (src/test/pdesolver.cpp:59-108) converges at fourth order."""
import numpy as np
import pytest

from adv_cases import compressive, gauss_test_function, reference_velocity


def test_interpolant_keeps_constants_shifts_exactly_and_reproduces_cubics(oracle_mod):
    n = 32
    o = oracle_mod.Oracle(n)
    rng = np.random.default_rng(1)
    q = [rng.uniform(-3, 3, (n, n, n)) for _ in range(3)]
    assert np.abs(o.sl_interp(np.ones((n, n, n)), q) - 1).max() < 1e-14
    f = rng.standard_normal((n, n, n))
    whole = [np.full((n, n, n), 2.0), np.full((n, n, n), -1.0), np.full((n, n, n), 5.0)]
    assert np.array_equal(o.sl_interp(f, whole), np.roll(f, (-2, 1, -5), axis=(0, 1, 2)))
    z = np.arange(n, dtype=np.float64)
    for axis in range(3):
        shape = [1, 1, 1]
        shape[axis] = n
        g = np.broadcast_to((0.01 * z ** 3 - 0.2 * z ** 2 + z).reshape(shape), (n, n, n)).copy()
        qa = [np.zeros((n, n, n)) for _ in range(3)]
        qa[axis][:] = 0.37
        got = np.moveaxis(o.sl_interp(g, qa), axis, 0)[2:n - 3, 0, 0]
        zz = z[2:n - 3] + 0.37
        assert np.abs(got - (0.01 * zz ** 3 - 0.2 * zz ** 2 + zz)).max() < 1e-12


def test_constant_velocity_trajectory_is_closed_form(oracle_mod):
    n, dt = 32, 0.03
    o = oracle_mod.Oracle(n)
    v = [np.full((n, n, n), 1.2), np.full((n, n, n), -0.7), np.full((n, n, n), 2.5)]
    q = o.sl_trajectories(v, dt)
    for a in range(3):
        assert np.allclose(q[a], -dt * v[a] * n / (2 * np.pi), rtol=0, atol=1e-13)


def test_mass_conservation_mode_keeps_the_mass(oracle_mod):
    """d_t c + div(c v) = 0 keeps the integral of c; pure advection (d_t c + v.grad c = 0) moves it
    by the integral of c div v."""
    n, dt = 64, 0.05
    o = oracle_mod.Oracle(n)
    v, c = compressive(n)
    q = o.sl_trajectories(v, dt)
    m0 = c.sum()
    c1, c2 = c.copy(), c.copy()
    for _ in range(10):
        c1 = o.sl_solve(c1, v, q, dt, 1)
        c2 = o.sl_solve(c2, v, q, dt, 2)
    assert abs(c1.sum() / m0 - 1) < 1e-5
    assert abs(c2.sum() / m0 - 1) > 1e-2
    # a divergence-free velocity makes the two modes agree
    vc = [np.full((n, n, n), 0.8), np.zeros((n, n, n)), np.full((n, n, n), -0.4)]
    qc = o.sl_trajectories(vc, dt)
    assert np.allclose(o.sl_solve(c, vc, qc, dt, 1), o.sl_solve(c, vc, qc, dt, 2), rtol=0, atol=1e-12)


def test_reference_round_trip_converges_at_fourth_order(oracle_mod):
    """Four steps along the test velocity and four back (src/test/pdesolver.cpp:59-108, there at
    256^3 with a 1% bound): the relative error falls by more than 4x per doubling of n."""
    errs = []
    for n in (64, 128):
        o = oracle_mod.Oracle(n)
        v = reference_velocity(n)
        c0 = gauss_test_function(n)
        q = o.sl_trajectories(v, 0.02)
        c = c0
        for _ in range(4):
            c = o.sl_solve(c, v, q, 0.02, 2)
        vm = [-a for a in v]
        q = o.sl_trajectories(vm, 0.02)
        for _ in range(4):
            c = o.sl_solve(c, vm, q, 0.02, 2)
        errs.append(np.linalg.norm(c - c0) / np.linalg.norm(c0))
    assert errs[1] < 0.05 and errs[0] / errs[1] > 4, errs
