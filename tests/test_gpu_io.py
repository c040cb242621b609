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
