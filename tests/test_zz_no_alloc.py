"""glia_rd_smooth_noalloc (the smoother on caller-owned device fields, scratch in a static device
array) against glia_rd_smooth_field (the handle's smoother, which allocates its taps): two calls in
a row, the second on a smaller grid, each within 1e-6 relative L2 (the FP32 bar; the taps are
computed in FP64 on the device on one side and on the host on the other).  Also the size limits,
which are rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def handle_smooth(n, f, sigma):
    from glia_amd import _lib, rd
    s = rd.RDSolver(n)
    s.set_field(0, "work", f)
    rc = s.L.glia_rd_smooth_field(s.h, 0, _lib.FIELD["work"], float(sigma))
    assert rc == 0, s.L.glia_rd_last_error(s.h)
    out = s.get_field(0, "work")
    s.close()
    return out


def noalloc_smooth(n, f, sigma):
    import torch
    from glia_amd import _lib
    L = _lib.load()
    fin = torch.from_numpy(np.ascontiguousarray(f, np.float32)).cuda()
    fout = torch.full_like(fin, float("nan"))
    torch.cuda.synchronize()
    nn = (ctypes.c_int * 3)(*n)
    err = ctypes.create_string_buffer(256)
    rc = L.glia_rd_smooth_noalloc(nn, ctypes.c_void_p(fin.data_ptr()), ctypes.c_void_p(fout.data_ptr()),
                                  float(sigma), None, err, 256)
    assert rc == 0, err.value
    torch.cuda.synchronize()
    return fout.cpu().numpy()


def test_noalloc_smoother_matches_handle_smoother_twice():
    rng = np.random.default_rng(17)
    for n, sigma in (((64, 64, 64), 2 * np.pi / 64), ((48, 32, 32), 3 * 2 * np.pi / 48)):
        print("case noalloc smooth", n, sigma, flush=True)
        f = rng.standard_normal(n).astype(np.float32)
        ref = handle_smooth(n, f, sigma)
        got = noalloc_smooth(n, f, sigma)
        print("rel %.2e" % rel(got, ref), flush=True)
        assert np.all(np.isfinite(got))
        assert rel(got, ref) < 1e-6


def test_noalloc_smoother_rejects_sizes_beyond_its_scratch():
    from glia_amd import _lib
    L = _lib.load()
    err = ctypes.create_string_buffer(256)
    big = (ctypes.c_int * 3)(256, 256, 256)
    p = ctypes.c_void_p(256)  # never dereferenced: the size checks come first
    q = ctypes.c_void_p(512)
    assert L.glia_rd_smooth_noalloc(big, p, q, 0.1, None, err, 256) == _lib.ERR_UNSUPPORTED
    assert b"scratch limit" in err.value
    long_axis = (ctypes.c_int * 3)(512, 16, 16)  # a radius of 255 points on the long axis
    assert L.glia_rd_smooth_noalloc(long_axis, p, q, 1.0, None, err, 256) == _lib.ERR_UNSUPPORTED
    assert b"radius" in err.value
    small = (ctypes.c_int * 3)(64, 64, 64)
    assert L.glia_rd_smooth_noalloc(small, p, p, 0.1, None, err, 256) == _lib.ERR_ARG
