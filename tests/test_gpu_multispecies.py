"""The fused multispecies kernels (include/glia_ms.h) against the FP64 oracle."""
import ctypes

import numpy as np
import pytest

from ms_cases import state

pytestmark = pytest.mark.gpu


def test_sources_and_coefficients_match_the_oracle(oracle_mod):
    import torch
    from glia_amd import _lib
    L = _lib.load()
    print("case multispecies sources 20000 points", flush=True)
    for di_zero in (False, True):
        fields = state(di_zero=di_zero)
        dev = [torch.from_numpy(a.astype(np.float32)).cuda() for a in fields]
        prm = _lib.MsParams()
        L.glia_ms_params_init(ctypes.byref(prm))
        rc = L.glia_ms_sources(dev[0].numel(), *[ctypes.c_void_p(t.data_ptr()) for t in dev], 0.01, ctypes.byref(prm),
                               None)
        assert rc == 0
        torch.cuda.synchronize()
        got = [t.cpu().numpy().astype(np.float64) for t in dev[:6]]
        want = oracle_mod.ms_sources(*fields, 0.01)
        for g, w, name in zip(got, want, ("p", "i", "n", "ox", "gm", "wm")):
            assert np.abs(g - w).max() < 1e-5 * max(1.0, np.abs(w).max()), name
        if di_zero:
            tissue = (fields[4] > 0.01) | (fields[5] > 0.01)
            before = fields[0] + fields[1] + fields[2] + fields[4] + fields[5]
            after = got[0] + got[1] + got[2] + got[4] + got[5]
            assert np.abs(after - before)[tissue].max() < 1e-6
    rng = np.random.default_rng(8)
    tis = [rng.uniform(0, 0.5, 5000).astype(np.float32) for _ in range(4)]
    dev = [torch.from_numpy(a).cuda() for a in tis]
    rf, kf = torch.zeros(5000, device="cuda"), torch.zeros(5000, device="cuda")
    assert L.glia_ms_update_coefficients(5000, ctypes.c_void_p(rf.data_ptr()), ctypes.c_void_p(kf.data_ptr()),
                                         *[ctypes.c_void_p(t.data_ptr()) for t in dev], 10.0, 0.05, 0.2, 0.1,
                                         None) == 0
    torch.cuda.synchronize()
    wr, wk = oracle_mod.ms_update_coefficients(*[a.astype(np.float64) for a in tis], 10.0, 0.05, 0.2, 0.1)
    assert np.abs(rf.cpu().numpy() - wr).max() < 1e-5 and np.abs(kf.cpu().numpy() - wk).max() < 1e-7


def test_segmentation_clip_and_edema_match_the_oracle(oracle_mod):
    import torch
    from glia_amd import _lib
    L = _lib.load()
    print("case multispecies segmentation/clip/edema", flush=True)
    rng = np.random.default_rng(12)
    f = [rng.uniform(0, 1, 8000).astype(np.float32) for _ in range(6)]  # bg, gm, wm, vt, csf, c
    dev = [torch.from_numpy(a).cuda() for a in f]
    seg = torch.zeros(8000, device="cuda")
    assert L.glia_ms_segmentation(8000, *[ctypes.c_void_p(t.data_ptr()) for t in dev], ctypes.c_void_p(seg.data_ptr()),
                                  None) == 0
    x = torch.from_numpy(rng.uniform(-1, 1, 8000).astype(np.float32)).cuda()
    x0 = x.cpu().numpy()
    assert L.glia_ms_clip(8000, ctypes.c_void_p(x.data_ptr()), None) == 0
    ed = torch.zeros(8000, device="cuda")
    iv = torch.from_numpy(rng.uniform(0, 0.002, 8000).astype(np.float32)).cuda()
    assert L.glia_ms_edema_mask(8000, ctypes.c_void_p(iv.data_ptr()), 0.001, ctypes.c_void_p(ed.data_ptr()), None) == 0
    torch.cuda.synchronize()
    got = seg.cpu().numpy()
    # the same rule in float32 arithmetic: exact; the FP64 oracle: equal except at near-ties
    bg, gm, wm, vt, csf, c = f
    a = np.float32(1) - c
    cand = [bg, c, wm * a, gm * a, vt * a, csf * a]
    labels = np.array([0, 1, 6, 5, 7, 8], np.float32)
    best, lab = cand[0].copy(), np.zeros(8000, np.float32)
    for v, l in zip(cand[1:], labels[1:]):
        take = v > best
        best = np.where(take, v, best)
        lab = np.where(take, l, lab)
    assert np.array_equal(got, lab)
    want = oracle_mod.ms_segmentation(*[x.astype(np.float64) for x in f])
    assert np.mean(got == want) > 0.999
    assert np.array_equal(x.cpu().numpy(), np.where(x0 <= 0, 0, x0).astype(np.float32))
    assert np.array_equal(ed.cpu().numpy(), (iv.cpu().numpy() > np.float32(0.001)).astype(np.float32))
