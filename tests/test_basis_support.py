"""Placement of the Gaussian basis (Phi::setGaussians, src/mat/Phi.cpp:33-73, 148-234, 524-940;
readPhiMesh, src/utils/IO.cpp:717-790).

The oracle's data-driven support (orc_data_support) is pinned by closed forms: a single voxel
above the threshold is seen by exactly the lattice points within the radius of it; a full field
keeps every candidate at gvf = 0 and, at gvf just below 1, exactly the candidates whose ball lies
inside the grid (the count is clipped, not wrapped); the lattice ends at the largest exact multiple
of the spacing (the reference's fmod test), which leaves the last cells of a 160-point axis
without candidates at the default sigma; the comparison is strict.  The GPU counts are compared
with the oracle bit for bit (integer counts, centers index * h in FP64)."""
import numpy as np
import pytest

TWO_PI = 2 * np.pi


def ball_count(r):
    """Integer offsets with |d| <= r, by enumeration."""
    R = int(np.floor(r))
    g = np.arange(-R, R + 1)
    d2 = g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2
    return int(np.sum(np.sqrt(d2) <= r))


def test_single_voxel_is_seen_by_the_lattice_points_within_the_radius(oracle_mod):
    n = 64
    sigma = 2 * TWO_PI / n  # radius 2 cells, spacing 4
    d = np.zeros((n, n, n), np.float32)
    v = np.array([22, 28, 8])  # exactly 2 cells from the candidates (20, 28, 8) and (24, 28, 8)
    d[tuple(v)] = 1.0
    got = oracle_mod.data_support(d, 0.05, sigma)
    lat = np.arange(0, 61, 4)
    want = [(x, y, z) for x in lat for y in lat for z in lat
            if (x - v[0]) ** 2 + (y - v[1]) ** 2 + (z - v[2]) ** 2 <= 4]
    assert want == [(20, 28, 8), (24, 28, 8)]
    assert np.array_equal(got, np.array(want, np.float64) * (TWO_PI / n))


def test_full_field_and_volume_fraction(oracle_mod):
    n = 64
    sigma = 2 * TWO_PI / n
    assert ball_count(2.0) == 33
    d = np.ones((n, n, n), np.float32)
    c, cnt = oracle_mod.data_support(d, 0.05, sigma, 0.0, counts=True)
    assert len(c) == 16 ** 3
    # a candidate's ball is clipped at the grid's faces: counts 33 inside, fewer at the faces
    assert cnt[: 16 ** 3].max() == 33 and cnt[0] == 11  # corner: offsets >= 0 with |d|^2 in {0, 1, 2, 3, 4}: 1 + 3 + 3 + 1 + 3
    inner = oracle_mod.data_support(d, 0.05, sigma, 0.99)
    lat_in = np.arange(4, 61, 4)  # x - 2 >= 0 and x + 2 <= 63
    assert len(inner) == len(lat_in) ** 3
    assert np.array_equal(np.unique(np.round(inner[:, 0] * n / TWO_PI)), lat_in)


def test_lattice_end_and_strict_threshold(oracle_mod):
    n = 80
    sigma = TWO_PI / 128  # radius 0.625: the ball is the candidate itself; spacing 1.25, steps of 1
    d = np.full((n, n, n), 0.5, np.float32)
    c = oracle_mod.data_support(d, 0.25, sigma)
    # the lattice ends at 75, the largest exact multiple of 1.25 below 80
    assert len(c) == 76 ** 3
    assert np.round(c[-1] * n / TWO_PI).tolist() == [75, 75, 75]
    assert len(oracle_mod.data_support(d, 0.5, sigma)) == 0  # data > threshold, strictly
    with pytest.raises(ValueError):
        oracle_mod.data_support(np.zeros((64, 64, 64), np.float32), 0.05, TWO_PI / 256)  # spacing 0.5


def test_phi_mesh_layouts():
    from glia_amd import io
    cm = np.array([3.0, 3.1, 2.9])
    sp = 0.1 * 1.5
    assert np.array_equal(io.phi_mesh(1, cm, 0.1), cm[None, :])
    c27 = io.phi_mesh(27, cm, 0.1, 1.5)
    assert c27.shape == (27, 3)
    assert np.allclose(c27.mean(axis=0), cm, atol=1e-15)
    assert np.allclose(c27[1] - c27[0], [sp, 0, 0]) and np.allclose(c27[3] - c27[0], [0, sp, 0])
    c8 = io.phi_mesh(8, cm, 0.1, 1.5)
    assert np.allclose(np.abs(c8 - cm), sp / 2)  # the octants, half a spacing from cm
    c64 = io.phi_mesh(64, cm, 0.1, 1.5)
    assert np.allclose(np.unique(np.round((c64[:, 0] - cm[0]) / sp, 12)), [-1.5, -0.5, 0.5, 1.5])


def test_phi_mesh_matches_oracle(oracle_mod):
    from glia_amd import io
    for n_ in (1, 8, 27, 64, 125):
        a = io.phi_mesh(n_, [3.0, 3.1, 2.9], 0.1, 1.5)
        assert np.array_equal(a, oracle_mod.phi_mesh(n_, np.array([3.0, 3.1, 2.9]), 0.1, 1.5))


def test_read_phi_mesh(tmp_path):
    from glia_amd import io
    c = np.array([[1.5, 2.25, 3.0], [0.125, 6.0, 4.5], [2.0, 2.0, 2.0]])
    p = tmp_path / "phi-mesh-0.dat"
    # the layout the reference writes (src/mat/Phi.cpp:914-935), with a component column
    p.write_text(" sigma = 0.0245437, spacing = 0.0490874\n centers = [\n" +
                 "".join(" %.17g, %.17g, %.17g, %d\n" % (a, b, z, i + 1) for i, (a, b, z) in enumerate(c)) + "];\n")
    got, comps, sg, sp = io.read_phi_mesh(p)
    assert np.array_equal(got, c) and comps.tolist() == [1, 2, 3]
    assert sg == 0.0245437 and sp == 0.0490874
    with pytest.raises(ValueError):
        io.read_phi_mesh(p, max_np=2)
    with pytest.raises(FileNotFoundError):
        io.read_phi_mesh(tmp_path / "missing.dat")


def blob(n, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(n) * TWO_PI / n
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    f = np.exp(-((x - 2.9) ** 2 + (y - 3.3) ** 2 + (z - 3.1) ** 2) / 0.3)
    f += 0.02 * rng.standard_normal(f.shape)
    return f.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("n,sigma,gvf", [(256, TWO_PI / 256, 0.0), (256, TWO_PI / 128, 0.3), (160, TWO_PI / 256, 0.0),
                                         (128, 3 * TWO_PI / 128, 0.5)])
def test_gpu_data_support_vs_oracle(oracle_mod, n, sigma, gvf):
    from glia_amd.rd import RDSolver
    print("case data support n %d sigma %.5f gvf %.2f" % (n, sigma, gvf), flush=True)
    d = blob(n, n)
    r = RDSolver(n)
    try:
        r.set_field(0, "d1", d)
        got = r.data_support(0, "d1", threshold=0.05, sigma=sigma, gvf=gvf)
        f = r.get_field(0, "d1")
    finally:
        r.close()
    want = oracle_mod.data_support(f, 0.05, sigma, gvf)
    print("centers", len(got), len(want), flush=True)
    assert len(want) > 8
    assert np.array_equal(got, want)


def test_read_concomp_and_sparsity(tmp_path):
    """readConCompDat (src/utils/IO.cpp:929-990) on a file in its layout, and the sparsity level of
    src/Solver.cpp:231-248: s per component of weight >= thresh, one per component of weight in
    [thresh / 10, thresh), none below."""
    from glia_amd import io
    p = tmp_path / "concomp.dat"
    p.write_text("#components\n3\ncenter of mass:\n1.5,2.25,3\n0.5,0.75,1\n4,5,6\nrelative mass:\n"
                 "0.9\n0.0005\n0.00005\n")
    w, c = io.read_concomp(p)
    assert w.tolist() == [0.9, 0.0005, 0.00005]
    assert c.tolist() == [[1.5, 2.25, 3.0], [0.5, 0.75, 1.0], [4.0, 5.0, 6.0]]
    assert io.component_sparsity(5, w, 1e-3) == 5 * 1 + 1
    assert io.component_sparsity(5, [0.5, 0.3, 0.2], 1e-3) == 15
    with pytest.raises(ValueError):
        io.read_concomp(p, max_comp=2)
    with pytest.raises(FileNotFoundError):
        io.read_concomp(tmp_path / "missing.dat")
