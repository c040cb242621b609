"""Host-side input of the RD path (include/glia_io.h through glia_amd.io): NetCDF classic files
against scipy's independent reader/writer, the segmentation split of the reference, config
parsing, and the separable form of the reference's 8-image smoothing kernel (what the GPU smoother
computes) against the oracle's FFT convolution."""
import gzip
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TESTDATA = os.path.join(HERE, "golden", "testdata")


def unpacked(name, tmp_path):
    out = tmp_path / name
    out.write_bytes(gzip.decompress(open(os.path.join(TESTDATA, name + ".gz"), "rb").read()))
    return str(out)


def scipy_read(path):
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        return np.array(f.variables["data"][:]), f.version_byte


@pytest.mark.parametrize("name", ["atlas.nc", "sinusoid.nc", "patient_tumor_IC.nc"])
def test_read_reference_testdata_like_scipy(name, tmp_path):
    from glia_amd import io
    path = unpacked(name, tmp_path)
    ref, _ = scipy_read(path)
    got = io.read_netcdf(path, dtype=np.float64)
    assert got.shape == ref.shape == (64, 64, 64)
    assert np.array_equal(got, ref.astype(np.float64))


def test_atlas_is_cdf1_double_and_sinusoid_cdf2_float(tmp_path):
    _, v1 = scipy_read(unpacked("atlas.nc", tmp_path))
    _, v2 = scipy_read(unpacked("sinusoid.nc", tmp_path))
    assert (v1, v2) == (1, 2)


@pytest.mark.parametrize("as_double", [False, True])
def test_write_reads_back_with_scipy(as_double, tmp_path):
    from glia_amd import io
    a = np.random.default_rng(3).standard_normal((8, 12, 16)).astype(np.float32)
    p = str(tmp_path / "w.nc")
    io.write_netcdf(p, a, as_double=as_double)
    ref, ver = scipy_read(p)
    assert ver == 2 and ref.shape == a.shape and np.array_equal(ref.astype(np.float32), a)
    assert np.array_equal(io.read_netcdf(p), a)


def test_read_scipy_written_file(tmp_path):
    from scipy.io import netcdf_file
    from glia_amd import io
    a = np.arange(4 * 6 * 10, dtype=np.float64).reshape(4, 6, 10) * 0.25
    p = str(tmp_path / "s.nc")
    with netcdf_file(p, "w", version=1) as f:
        f.createDimension("x", 4)
        f.createDimension("y", 6)
        f.createDimension("z", 10)
        f.createVariable("other", "i", ("x",))[:] = np.arange(4)
        f.createVariable("data", "d", ("x", "y", "z"))[:] = a
    assert np.array_equal(io.read_netcdf(p, dtype=np.float64), a)


def test_read_rejects_missing_file_and_variable(tmp_path):
    from glia_amd import io
    with pytest.raises(OSError):
        io.read_netcdf(str(tmp_path / "none.nc"))
    with pytest.raises(OSError):
        io.read_netcdf(unpacked("sinusoid.nc", tmp_path), var="nope")


def test_split_segmentation_of_the_atlas(tmp_path):
    """scripts/forward.py labels [wm=6,gm=5,vt=7,csf=8]: one-hot maps, background elsewhere."""
    from glia_amd import io
    seg = io.read_netcdf(unpacked("atlas.nc", tmp_path))
    maps = io.split_segmentation(seg, dict(wm=6, gm=5, vt=7, csf=8))
    for k, lab in (("wm", 6), ("gm", 5), ("vt", 7), ("csf", 8)):
        assert np.array_equal(maps[k], (seg == lab).astype(np.float32))
    total = maps["wm"] + maps["gm"] + maps["vt"] + maps["csf"]
    assert np.array_equal(total, (seg != 0).astype(np.float32))


def test_split_segmentation_tumor_rules():
    """src/utils/Utils.cpp:625-646: tumor core (tc, or nec with en) and edema count as WM when a
    tumor map is requested; the edema map is filled only then."""
    from glia_amd import io
    seg = np.array([6, 5, 7, 8, 1, 4, 2, 0], dtype=np.float32)
    lab = dict(wm=6, gm=5, vt=7, csf=8, nec=1, en=4, ed=2)
    m = io.split_segmentation(seg, lab, with_tumor=True)
    assert m["wm"].tolist() == [1, 0, 0, 0, 1, 1, 1, 0]
    assert m["tc"].tolist() == [0, 0, 0, 0, 1, 1, 0, 0]
    assert m["ed"].tolist() == [0, 0, 0, 0, 0, 0, 1, 0]
    m2 = io.split_segmentation(seg, lab)
    assert m2["wm"].tolist() == [1, 0, 0, 0, 0, 0, 0, 0]
    with pytest.raises(ValueError):
        io.split_segmentation(seg, dict(wm=6, gm=5))


def test_parse_config(tmp_path):
    """parseConfig/setParameter (src/utils/IO.cpp:6-257): whitespace removed, '#' comments, the
    reference defaults for keys that are absent."""
    from glia_amd import io
    p = tmp_path / "c.txt"
    p.write_text("# comment\nn = 64\nmodel=1\nnt_inv=25   # steps\ndt_inv=0.04\ninit_rho=8\ninit_k=0.01\n"
                 "k_gm_wm=0.2\nobs_threshold_1=-0.99\nbeta_p=1\natlas_labels=[wm=6,gm=5,vt=7,csf=8]\n"
                 "a_seg_path=./testdata/atlas.nc\nunknown_key=7\nnumber_gaussians=8\n")
    c = io.parse_config(str(p))
    assert (c["n"], c["model"], c["nt"], c["np"]) == (64, 1, 25, 8)
    assert (c["dt"], c["rho"], c["k"], c["k_gm_wm"], c["obs_threshold_1"], c["beta"]) == (0.04, 8.0, 0.01, 0.2, -0.99, 1.0)
    assert c["atlas_labels"] == [6, 5, 7, 8, -1, -1, -1, -1]
    assert c["a_seg_path"] == "./testdata/atlas.nc"
    assert (c["r_gm_wm"], c["smoothing_factor"], c["sigma_spacing"]) == (0.0, 1.0, 1.5)  # defaults
    assert (c["obs_lambda"], c["smoothing_factor_data"], c["obs_threshold_rel"]) == (-1.0, 1.0, 0)  # defaults
    assert c["patient_labels"] == [-1] * 8 and c["p_seg_path"] == ""
    p.write_text("obs_lambda=0.3\nsmoothing_factor_data=0\nobs_threshold_rel=1\np_seg_path=p.nc\n"
                 "patient_labels=[wm=6,gm=5,vt=7,csf=8,ed=2,nec=1,en=4]\n")
    c = io.parse_config(str(p))
    assert (c["obs_lambda"], c["smoothing_factor_data"], c["obs_threshold_rel"], c["p_seg_path"]) == (0.3, 0.0, 1, "p.nc")
    assert c["patient_labels"] == [6, 5, 7, 8, -1, 1, 4, 2] and c["atlas_labels"] == [-1] * 8
    assert (c["threshold_data_driven"], c["gaussian_volume_fraction"]) == (0.05, 0.0)  # defaults
    assert c["sigma_data_driven"] == 2 * np.pi / 256 and c["support_data_path"] == c["gaussian_cm_path"] == ""
    # sigma_factor sets the data-driven sigma in units of 2 pi / n (src/utils/IO.cpp:175-177)
    p.write_text("n=128\nsigma_factor=2\nthreshold_data_driven=0.1\ngaussian_volume_fraction=0.25\n"
                 "support_data_path=s.nc\ngaussian_cm_path=phi-mesh.dat\ndata_comp_data_path=concomp.dat\n"
                 "d0_path=c0.nc\nsmoothing_factor_data_t0=0.5\nobs_threshold_0=0.2\n")
    c = io.parse_config(str(p))
    assert c["sigma_data_driven"] == 2 * 2.0 * np.pi / 128
    assert (c["d0_path"], c["smoothing_factor_data_t0"], c["obs_threshold_0"]) == ("c0.nc", 0.5, 0.2)
    assert (c["threshold_data_driven"], c["gaussian_volume_fraction"]) == (0.1, 0.25)
    assert (c["support_data_path"], c["gaussian_cm_path"], c["data_comp_data_path"]) == ("s.nc", "phi-mesh.dat",
                                                                                 "concomp.dat")
    with pytest.raises(OSError):
        io.parse_config(str(tmp_path / "missing.txt"))


def test_edema_observation_mask():
    """createEdemaBasedObservationMask (src/utils/Utils.cpp:567-590) on the maps of a BraTS-style
    label image: 1 on the tumor core (necrotic and enhancing labels), 0 on edema, lambda elsewhere
    (healthy tissue and background); the 0.99 cut holds for values just below it."""
    from glia_amd import io
    lab = dict(wm=6, gm=5, vt=7, csf=8, ed=2, nec=1, en=4)
    seg = np.array([0, 6, 5, 7, 8, 1, 4, 2], dtype=np.float32)
    m = io.split_segmentation(seg, lab, with_tumor=True)
    mask = io.edema_obs_mask(m["tc"], m["ed"], 0.25)
    assert mask.tolist() == [0.25, 0.25, 0.25, 0.25, 0.25, 1.0, 1.0, 0.0]
    tc = np.array([0.995, 0.98, 0.0, 0.0], np.float32)
    ed = np.array([1.0, 1.0, 0.995, 0.98], np.float32)
    assert io.edema_obs_mask(tc, ed, 0.5).tolist() == [1.0, 0.0, 0.0, 0.5]
    with pytest.raises(ValueError):
        io.edema_obs_mask(tc, ed[:3], 0.5)


def test_smoother_is_separable_truncated_convolution(oracle_mod):
    """The reference's 8-image kernel (src/cuda/SpectralOperators.cu:12-51) factors into per-axis
    kernels g(x) + g(x - 2 pi) normalized by their sums; three periodic 1D convolutions with taps
    cut at 12 sigma (what the GPU smoother does) equal the oracle's FFT convolution."""
    n = 32
    o = oracle_mod.Oracle(n)
    f = np.random.default_rng(4).random((n, n, n))
    sigma = 1.5 * 2 * np.pi / n
    ref = o.smooth(f, sigma)
    h = 2 * np.pi / n
    x = np.arange(n) * h
    g = np.exp(-x ** 2 / (2 * sigma ** 2)) + np.exp(-(x - 2 * np.pi) ** 2 / (2 * sigma ** 2))
    g /= g.sum()
    kr = min(n // 2 - 1, int(np.ceil(12 * sigma / h)))
    out = f.copy()
    for ax in range(3):
        acc = np.zeros_like(out)
        for s in range(-kr, kr + 1):
            acc += g[s % n] * np.roll(out, s, axis=ax)
        out = acc
    assert np.max(np.abs(out - ref)) < 1e-13
