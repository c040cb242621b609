"""The oracle's multispecies pointwise terms (src/pde/PdeOperatorsMultiSpecies.cpp:10-234,
src/cuda/PdeOperators.cu:27-124) against what the equations fix: with no diffusion loss (di = 0)
and healthy tissue present, every transfer between p, i, n, gm and wm cancels, so their sum is
conserved at each voxel; the rate switches take their limits far from the oxygen thresholds; the
coefficient update matches the formula by hand."""
import numpy as np

from ms_cases import state


def test_sources_conserve_cells_per_voxel_without_diffusion_loss(oracle_mod):
    p, i, nec, ox, gm, wm, rho, di = state(di_zero=True)
    before = p + i + nec + gm + wm
    out = oracle_mod.ms_sources(p, i, nec, ox, gm, wm, rho, di, 0.01)
    after = out[0] + out[1] + out[2] + out[4] + out[5]
    tissue = (gm > 0.01) | (wm > 0.01)
    assert np.abs(after - before)[tissue].max() < 1e-14
    # di removes healthy cells one for one
    di2 = np.full_like(p, 0.3)
    out2 = oracle_mod.ms_sources(p, i, nec, ox, gm, wm, rho, di2, 0.01)
    after2 = out2[0] + out2[1] + out2[2] + out2[4] + out2[5]
    assert np.allclose((after2 - before)[tissue], -0.01 * 0.3, atol=1e-14)


def test_rate_switches_take_their_limits(oracle_mod):
    """With ox = 1 the cells grow at rho (m = rho, alpha ~ 0, h ~ 0); with ox = 0 growth stops and
    the death switch is on (m ~ 0, h ~ 1)."""
    dt, rho = 1e-3, 5.0
    one = np.ones(1)
    for oxv, m, al, h in ((1.0, rho, 0.0, 0.0), (0.0, 0.0, 0.1, 1.0)):
        p, i = np.array([0.3]), np.array([0.2])
        out = oracle_mod.ms_sources(p, i, 0 * one, oxv * one, 0.2 * one, 0.3 * one, rho * one, 0 * one, dt)
        be = 0.01 * oxv
        want_p = 0.3 + dt * (m * 0.3 * 0.7 - al * 0.3 + be * 0.2 - 4.0 * h * 0.3)
        assert abs(out[0][0] - want_p) < 1e-9, (oxv, out[0][0], want_p)


def test_coefficients_by_hand(oracle_mod):
    bg, gm, vt, csf = (np.array(v) for v in ([0.1, 0.9], [0.5, 0.3], [0.1, 0.0], [0.0, 0.2]))
    rf, kf = oracle_mod.ms_update_coefficients(bg, gm, vt, csf, 10.0, 0.05, 0.2, 0.2)
    assert np.allclose(rf, [10 * (1 - (0.1 + 0.8 * 0.5 + 0.1)), 0.0])
    assert np.allclose(kf, [0.05 * (1 - (0.1 + 0.8 * 0.5 + 0.1)), 0.0])


def test_segmentation_labels_by_hand(oracle_mod):
    """computeSegmentation: the largest of bg, c and the tissues times (1 - c); ties keep the earlier
    label in the order bg, c, wm, gm, vt, csf."""
    bg = np.array([0.9, 0.1, 0.1, 0.0, 0.2, 0.3])
    c = np.array([0.05, 0.8, 0.0, 0.0, 0.0, 0.3])
    wm = np.array([0.0, 0.1, 0.7, 0.1, 0.0, 0.3 / 0.7])
    gm = np.array([0.0, 0.0, 0.2, 0.6, 0.0, 0.0])
    vt = np.array([0.0, 0.0, 0.0, 0.3, 0.7, 0.0])
    csf = np.array([0.05, 0.0, 0.0, 0.0, 0.1, 0.0])
    seg = oracle_mod.ms_segmentation(bg, gm, wm, vt, csf, c)
    # last voxel: bg 0.3, c 0.3, wm (1 - 0.3) * 0.3 / 0.7 = 0.3: bg keeps the first maximum
    assert seg.tolist() == [0, 1, 6, 5, 7, 0]
