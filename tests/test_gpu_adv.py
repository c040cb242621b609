"""The semi-Lagrangian solver on the GPU (include/glia_adv.h) against the FP64 oracle and the
reference's round-trip test (src/test/pdesolver.cpp:59-108)."""
import numpy as np
import pytest

from adv_cases import compressive, gauss_test_function, reference_velocity

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_trajectories_and_steps_match_the_oracle(oracle_mod):
    from glia_amd import adv
    n, dt = 64, 0.02
    print("case advection vs oracle 64^3", flush=True)
    o = oracle_mod.Oracle(n)
    v = reference_velocity(n)
    c0 = gauss_test_function(n)
    s = adv.SemiLagrangianSolver(n)
    s.set_velocity(*v)
    s.computeTrajectories(dt)
    q = s.offsets()
    qo = o.sl_trajectories(v, dt)
    dq = max(float(np.abs(q[a] - qo[a]).max()) for a in range(3))
    c = c0.astype(np.float32)
    co = c0.copy()
    for _ in range(4):
        c = s.solve(c, dt, mode=2)
        co = o.sl_solve(co, v, qo, dt, 2)
    print("max |q - q_oracle|", dq, "rel c", rel(c, co), flush=True)
    assert dq < 2e-6
    assert rel(c, co) < 1e-5
    s.close()


def test_mass_conservation_mode_matches_the_oracle(oracle_mod):
    from glia_amd import adv
    n, dt = 64, 0.05
    print("case advection mode 1 vs oracle 64^3", flush=True)
    o = oracle_mod.Oracle(n)
    v, c0 = compressive(n)
    s = adv.SemiLagrangianSolver(n)
    s.set_velocity(*v)
    qo = o.sl_trajectories(v, dt)
    c, co = c0.astype(np.float32), c0.copy()
    for _ in range(10):
        c = s.solve(c, dt, mode=1)
        co = o.sl_solve(co, v, qo, dt, 1)
    s.close()
    print("rel c", rel(c, co), "mass", c.sum(dtype=np.float64) / c0.sum() - 1, flush=True)
    assert rel(c, co) < 1e-5
    assert abs(c.sum(dtype=np.float64) / c0.sum() - 1) < 1e-5


def test_whole_cell_shift_and_rest_are_exact():
    """v = (1, 0, -1) with dt = 4 pi / n moves every point by exactly (-2, 0, 2) cells: the step is
    a periodic roll, bit for bit; v = 0 leaves the field unchanged."""
    from glia_amd import adv
    n = (32, 48, 64)
    print("case advection whole-cell shift", n, flush=True)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(n).astype(np.float32)
    s = adv.SemiLagrangianSolver(n)
    one = np.ones(n, np.float32)
    s.set_velocity(one * 0, one * 0, one * 0)
    assert np.array_equal(s.solve(f, 0.1, mode=2), f)
    for k in range(3):
        dt = 2 * 2 * np.pi / n[k]  # 2 cells along axis k at unit speed
        v = [one * 0, one * 0, one * 0]
        v[k] = one * (1.0 if k != 2 else -1.0)
        s.set_velocity(*v)
        got = s.solve(f, dt, mode=2)
        shift = [0, 0, 0]
        shift[k] = 2 if k != 2 else -2
        assert np.array_equal(got, np.roll(f, shift, axis=(0, 1, 2))), k
    s.close()


def test_reference_round_trip_at_256():
    """src/test/pdesolver.cpp:59-108: four steps of dt = 0.02 along the test velocity, the velocity
    reversed, trajectories recomputed, four steps back; relative error below 1e-2."""
    from glia_amd import adv
    n = 256
    print("case advection round trip 256^3", flush=True)
    v = [a.astype(np.float32) for a in reference_velocity(n)]
    c0 = gauss_test_function(n).astype(np.float32)
    s = adv.SemiLagrangianSolver(n)
    s.set_velocity(*v)
    c = c0.copy()
    for _ in range(4):
        c = s.solve(c, 0.02, mode=2)
    s.set_velocity(*[-a for a in v])
    for _ in range(4):
        c = s.solve(c, 0.02, mode=2)
    s.close()
    e = rel(c, c0)
    print("round trip rel error", e, flush=True)
    assert e < 1e-2


def test_batched_advection_equals_single_solves():
    """glia_sl_advect over six fields (gm, wm, vt, csf, c, mri of the mass-effect step) with
    mixed modes gives the single solves' results bit for bit."""
    import torch
    from glia_amd import adv
    n = 64
    print("case batched advection 6 fields 64^3", flush=True)
    v, _ = compressive(n)
    rng = np.random.default_rng(5)
    fields = [rng.uniform(0, 1, (n, n, n)).astype(np.float32) for _ in range(6)]
    modes = [1, 1, 2, 2, 1, 2]
    s = adv.SemiLagrangianSolver(n)
    s.set_velocity(*v)
    dev = [torch.from_numpy(f.reshape(-1)).cuda() for f in fields]
    outs = s.advect(dev, modes, 0.05)
    torch.cuda.synchronize()
    for f, m, o in zip(fields, modes, outs):
        ref = s.solve(f, 0.05, mode=m)
        assert np.array_equal(o.cpu().numpy().reshape(n, n, n), ref)
    with pytest.raises(Exception):
        s.advect([dev[0]], [1], 0.05, out=[dev[0]])
    s.close()


def test_large_displacements_wrap_and_match_the_oracle(oracle_mod):
    """A rough velocity of tens of cells per step sends departure points far from their grid
    points and across the periodic boundary several times; the values still match the oracle's
    interpolant at the GPU's own departure offsets."""
    from glia_amd import adv
    n = 32
    print("case advection large displacements 32^3", flush=True)
    rng = np.random.default_rng(9)
    v = [rng.uniform(-40, 40, (n, n, n)) for _ in range(3)]
    f = rng.standard_normal((n, n, n))
    s = adv.SemiLagrangianSolver(n)
    s.set_velocity(*v)
    s.computeTrajectories(2 * np.pi / n)  # dt / h = 1: offsets of tens of cells
    q = s.offsets()
    got = s.interpolate(f.astype(np.float32))
    s.close()
    o = oracle_mod.Oracle(n)
    want = o.sl_interp(f.astype(np.float32).astype(np.float64), [a.astype(np.float64) for a in q])
    span = max(float(a.max() - a.min()) for a in q)
    print("offset span", span, "rel", rel(got, want), flush=True)
    assert span > 40
    assert rel(got, want) < 1e-6
