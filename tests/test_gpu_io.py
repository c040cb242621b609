"""The atlas flow on the GPU (SolverInterface::readAtlas, src/SolverInterface.cpp:502-575): the
reference's 64^3 atlas segmentation read from NetCDF, split per tissue, smoothed on the device and
used as tissue of a forward solve; compared with the FP64 oracle fed the same label image."""
import gzip
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
LABELS = dict(wm=6, gm=5, vt=7, csf=8)  # scripts/forward.py atlas_labels


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_atlas_tissue_and_forward_vs_oracle(oracle_mod, tmp_path):
    from glia_amd import io, rd, synth
    print("case atlas", flush=True)
    path = tmp_path / "atlas.nc"
    path.write_bytes(gzip.decompress(open(os.path.join(HERE, "golden", "testdata", "atlas.nc.gz"), "rb").read()))
    n, nt, dt = 64, 10, 0.02
    s = rd.RDSolver(n, dt=dt, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2)
    maps = io.load_atlas(s, 0, str(path), LABELS, smoothing_factor=1.0)
    o = oracle_mod.Oracle(n)
    sigma = 2 * np.pi / n
    sm = {k: o.smooth(maps[k].astype(np.float64), sigma) for k in ("wm", "gm", "csf", "vt")}
    for k in ("wm", "gm", "csf", "vt"):
        assert rel(s.get_field(0, k), sm[k]) < 1e-6, k
    centers = synth.basis_centers(sm["wm"], 8, 3)
    p = np.random.default_rng(5).uniform(0, 1, 8)
    s.set_coefficients(0, 0.05, 8.0)
    s.set_basis(0, centers, sigma)
    s.precFactor()
    s.phi_apply([p])
    s.solve_state(0)
    o.set_tissue(sm["wm"], sm["gm"], sm["csf"], sm["vt"])
    o.set_time(dt, nt)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(centers, sigma, 1.0, False)
    o.set_c0(o.phi_apply(p))
    o.solve_state(0)
    err = rel(s.get_field(0, "ct"), o.ct())
    print("atlas forward rel err", err, flush=True)
    assert err < 1e-5
    s.close()


def test_patient_data_from_segmentation_vs_oracle(oracle_mod):
    """SolverInterface::readData with the target taken from the segmentation (src/SolverInterface.cpp:
    577-640): d1 = tumor core (necrotic and enhancing labels), the edema-based observation mask
    (obs_lambda >= 0), d1 smoothed with sigma = smoothing_factor_data 2 pi / n, the threshold
    relative to max d1; the member's objective and gradient then match the oracle fed the same d1,
    mask and threshold."""
    from glia_amd import io, rd, synth
    print("case patient data", flush=True)
    n, nt, dt = 32, 2, 0.02
    lab = dict(wm=6, gm=5, vt=7, csf=8, ed=2, nec=1, en=4)
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt((X - 16.0) ** 2 + (Y - 14.0) ** 2 + (Z - 15.0) ** 2)
    seg = np.zeros((n, n, n), np.float32)
    seg[r < 13] = 8
    seg[r < 12] = 5
    seg[r < 9] = 6
    seg[(np.abs(X - 16) < 2) & (np.abs(Y - 14) < 2) & (r < 11)] = 7
    r2 = np.sqrt((X - 12.0) ** 2 + (Y - 18.0) ** 2 + (Z - 15.0) ** 2)
    seg[r2 < 4.5] = 2
    seg[r2 < 3.0] = 4
    seg[r2 < 1.5] = 1
    cfg = dict(obs_lambda=0.3, smoothing_factor_data=1.0, obs_threshold_1=0.2, obs_threshold_rel=1)
    s = rd.RDSolver(n, dt=dt, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2)
    maps = io.load_atlas(s, 0, seg, lab, smoothing_factor=1.0)
    assert "tc" in maps and "ed" in maps
    d1, mask, thr = io.load_data(s, 0, cfg, maps)
    o = oracle_mod.Oracle(n)
    sigma = 2 * np.pi / n
    d1o = o.smooth(maps["tc"].astype(np.float64), sigma)
    assert rel(d1, d1o) < 1e-6
    assert abs(thr - 0.2 * d1o.max()) < 1e-6 * thr
    tc, ed = maps["tc"], maps["ed"]
    assert np.array_equal(mask, np.where(tc > 0.99, 1.0, np.where(ed > 0.99, 0.0, 0.3)).astype(np.float32))
    assert np.array_equal(s.get_field(0, "obs"), mask)
    sm = {k: o.smooth(maps[k].astype(np.float64), sigma) for k in ("wm", "gm", "csf", "vt")}
    centers = synth.basis_centers(sm["wm"], 6, 3)
    p = np.random.default_rng(2).uniform(0.2, 1.0, 6)
    s.set_basis(0, centers, sigma)
    xv = np.concatenate([p, [0.05, 8.0]])
    J, g = s.objective_and_gradient([xv])
    s.close()
    o.set_tissue(sm["wm"], sm["gm"], sm["csf"], sm["vt"])
    o.set_time(dt, nt)
    o.set_k(0.05, 0.2, 0.0)
    o.set_rho(8.0, 0.2, 0.0)
    o.prec_factor()
    o.set_phi(centers, sigma, 1.0, False)
    o.set_opt(beta=0.0)
    o.set_data(d1.astype(np.float64), mask.astype(np.float64), thr)
    Jo, go = o.objective_and_gradient(xv, mode=2)
    print("J", J[0], Jo, "dJ rel", rel(g[0], go), flush=True)
    assert abs(J[0] - Jo) / abs(Jo) < 1e-5
    assert rel(g[0], go) < 1e-4


@pytest.mark.parametrize("factor", [0.5, 1.0, 2.0])
def test_smoother_radii_vs_oracle(oracle_mod, factor):
    """glia_rd_smooth_field at sigma = factor * 2 pi / n: tap radius 6 and 12 (compile-time radius)
    and 24 (the general loop) against the oracle's FFT convolution of the reference's kernel."""
    from glia_amd import rd
    print("case smoother factor", factor, flush=True)
    n = 64
    rng = np.random.default_rng(11)
    f = rng.uniform(0.0, 1.0, (n, n, n))
    s = rd.RDSolver(n)
    try:
        s.set_field(0, "work", f)
        s.smooth_field(0, "work", factor * 2 * np.pi / n)
        got = s.get_field(0, "work")
    finally:
        s.close()
    ref = oracle_mod.Oracle(n).smooth(f.astype(np.float32).astype(np.float64), factor * 2 * np.pi / n)
    err = rel(got, ref)
    print("rel", err, flush=True)
    assert err < 1e-6


@pytest.mark.gpu
def test_init_gaussians_from_config(oracle_mod, tmp_path):
    """SolverInterface::initializeGaussians (src/SolverInterface.cpp:860-905) after readData: the
    data-driven support of the smoothed d1 (Phi::setGaussians(Vec data)) equals the oracle's on the
    same field, and is the basis the member then applies; a phi-mesh file (gaussian_cm_path)
    replaces it."""
    from glia_amd import io, rd
    print("case init gaussians", flush=True)
    n = 64
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    seg = np.full((n, n, n), 6, np.float32)
    seg[np.sqrt((X - 30.0) ** 2 + (Y - 35.0) ** 2 + (Z - 28.0) ** 2) < 6] = 4
    lab = dict(wm=6, gm=5, vt=7, csf=8, nec=1, en=4)
    cfg = dict(smoothing_factor_data=1.0, sigma_data_driven=2 * np.pi / 32, threshold_data_driven=0.05,
               gaussian_volume_fraction=0.0)
    s = rd.RDSolver(n)
    try:
        maps = io.load_atlas(s, 0, seg, lab, smoothing_factor=1.0)
        d1, _, _ = io.load_data(s, 0, cfg, maps)
        centers = io.init_gaussians(s, 0, cfg)
        want = oracle_mod.data_support(s.get_field(0, "d1"), 0.05, 2 * np.pi / 32)
        print("support centers", len(centers), flush=True)
        assert len(want) > 8 and np.array_equal(centers, want)
        p = np.zeros(len(centers))
        p[len(centers) // 2] = 1.0  # one active basis: Phi p is that basis scaled to maximum 1
        s.phi_apply([p])
        assert abs(float(s.get_field(0, "c0").max()) - 1.0) < 1e-5
        path = tmp_path / "phi-mesh.dat"
        path.write_text(" sigma = 0.1, spacing = 0.2\n centers = [\n" +
                        "".join(" %.17g, %.17g, %.17g, 1\n" % tuple(c) for c in centers[:3]) + "];\n")
        got = io.init_gaussians(s, 0, dict(cfg, gaussian_cm_path=str(path)))
        assert np.array_equal(got, centers[:3])
    finally:
        s.close()


@pytest.mark.gpu
def test_c0_input_vs_oracle(oracle_mod, tmp_path):
    """The c(0) half of readData (src/SolverInterface.cpp:654-690): c(0) from the reference's
    testdata/patient_tumor_IC.nc, and from a map with negative values (clipped at zero), smoothed
    with sigma = smoothing_factor_data_t0 * 2 pi / n and scaled to maximum 1, equals the oracle's
    smoothing of the same map (1e-6); the threshold is taken relative to the maximum before the
    scaling.  A kappa/rho evaluation from that c(0) as set (use_c0_) then matches the oracle's: J
    within 1e-5, (g_kappa, g_rho) within 1e-4."""
    from glia_amd import io, rd
    print("case c0 input", flush=True)
    n, nt, dt = 64, 4, 0.02
    path = tmp_path / "patient_tumor_IC.nc"
    path.write_bytes(gzip.decompress(open(os.path.join(HERE, "golden", "testdata", "patient_tumor_IC.nc.gz"), "rb").read()))
    raw = io.read_netcdf(str(path)).astype(np.float32)
    o = oracle_mod.Oracle(n)
    sigma = 2 * np.pi / n
    cfg = dict(d0_path=str(path), smoothing_factor_data_t0=1.0, obs_threshold_0=0.1, obs_threshold_rel=1)
    s = rd.RDSolver(n, dt=dt, nt=nt)
    try:
        c0, thr = io.load_c0(s, 0, cfg)
        sm = o.smooth(raw.astype(np.float64), sigma)
        assert rel(c0, sm / sm.max()) < 1e-6 and abs(c0.max() - 1.0) < 1e-6
        assert abs(thr - 0.1 * sm.max()) < 1e-6 * thr
        neg = raw - np.float32(0.05)
        c0n, _ = io.load_c0(s, 0, dict(cfg, d0_path=""), c0=neg)
        smn = o.smooth(np.maximum(neg, 0.0).astype(np.float64), sigma)
        assert rel(c0n, smn / smn.max()) < 1e-6
        c0, _ = io.load_c0(s, 0, cfg)
        s.set_tissue_sinusoidal(0)
        s.set_basis(0, np.array([[np.pi, np.pi, np.pi]]), sigma)
        s.set_field(0, "c0", c0)  # the initial condition the kappa/rho mode uses as set
        d1 = np.roll(c0, 2, axis=0)
        s.set_data(0, d1, None, -0.99)
        x = np.array([0.0, 0.02, 6.0])
        J, g = s.objective_and_gradient([x], mode=1)
    finally:
        s.close()
    o.set_tissue_sinusoidal()
    o.set_time(dt, nt)
    o.set_k(0.02, 0.0, 0.0)
    o.set_rho(6.0, 0.0, 0.0)
    o.prec_factor()
    o.set_phi(np.array([[np.pi, np.pi, np.pi]]), sigma, 1.0, False)
    o.set_c0(c0.astype(np.float64))
    o.set_opt(beta=0.0)
    o.set_data(d1.astype(np.float64), None, -0.99)
    Jo, go = o.objective_and_gradient(x, mode=1)
    print("J", J[0], Jo, "g", g[0][-2:], go[-2:], flush=True)
    assert abs(J[0] - Jo) < 1e-5 * abs(Jo)
    assert np.abs(g[0][-2:] - go[-2:]).max() < 1e-4 * np.abs(go[-2:]).max()
