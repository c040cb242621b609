"""Host-side input of the RD path over libglia_rd.so (include/glia_io.h): NetCDF classic fields,
segmentation labels, config files, and the atlas flow of SolverInterface::readAtlas
(src/SolverInterface.cpp:502-575): read the label image, split it per tissue, smooth the maps on
the GPU and derive mask, k, rho and Phi from them."""
import ctypes

import numpy as np

from . import _lib


def _b(s):
    return s.encode() if isinstance(s, str) else s


def read_netcdf(path, var="data", dtype=np.float32):
    """dataIn (src/utils/IO.cpp:511-539): the whole variable as a numpy array."""
    L = _lib.load()
    nd, tp = ctypes.c_int(), ctypes.c_int()
    dims = (ctypes.c_int64 * 8)()
    rc = L.glia_nc_inquire(_b(path), _b(var), ctypes.byref(nd), dims, ctypes.byref(tp))
    if rc:
        raise OSError("cannot read NetCDF variable %r from %s (code %d)" % (var, path, rc))
    shape = tuple(dims[i] for i in range(nd.value))
    out = np.empty(shape, dtype=np.float64 if dtype == np.float64 else np.float32)
    f = L.glia_nc_read_f64 if out.dtype == np.float64 else L.glia_nc_read_f32
    rc = f(_b(path), _b(var), out.ctypes.data_as(ctypes.c_void_p), out.size)
    if rc:
        raise OSError("cannot read NetCDF data from %s (code %d)" % (path, rc))
    return out


def write_netcdf(path, a, var="data", as_double=False):
    """dataOut (src/utils/IO.cpp:541-601): a 3D field as a CDF-2 file with dimensions x, y, z."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("a 3D field is required")
    n = (ctypes.c_int * 3)(*a.shape)
    rc = _lib.load().glia_nc_write3_f32(_b(path), _b(var), n, a.ctypes.data_as(ctypes.c_void_p), int(bool(as_double)))
    if rc:
        raise OSError("cannot write %s (code %d)" % (path, rc))


def split_segmentation(seg, labels, with_tumor=False):
    """splitSegmentation (src/utils/Utils.cpp:592-660). labels: dict with keys wm, gm, vt (required),
    csf, tc, nec, en, ed.  Returns dict of float32 maps (wm, gm, vt, csf[, tc, ed])."""
    seg = np.ascontiguousarray(seg, dtype=np.float32)
    lab = (ctypes.c_int * 8)(*[int(labels.get(k, -1)) for k in ("wm", "gm", "vt", "csf", "tc", "nec", "en", "ed")])
    names = ["wm", "gm", "vt", "csf"] + (["tc", "ed"] if with_tumor else [])
    out = {k: np.zeros(seg.shape, np.float32) for k in names}
    ptr = lambda k: out[k].ctypes.data_as(ctypes.c_void_p) if k in out else None
    rc = _lib.load().glia_split_segmentation(seg.ctypes.data_as(ctypes.c_void_p), seg.size, lab, ptr("wm"), ptr("gm"),
                                             ptr("vt"), ptr("csf"), ptr("tc"), ptr("ed"))
    if rc:
        raise ValueError("segmentation needs wm, gm and vt labels (code %d)" % rc)
    return out


def parse_config(path):
    """The RD-relevant keys of a reference config file (src/utils/IO.cpp:6-257) as a dict."""
    c = _lib.FileConfig()
    rc = _lib.load().glia_rd_parse_config(_b(path), ctypes.byref(c))
    if rc:
        raise OSError("cannot parse config %s (code %d)" % (path, rc))
    out = {}
    for name, _ in c._fields_:
        v = getattr(c, name)
        if isinstance(v, bytes):
            v = v.decode()
        elif name == "atlas_labels":
            v = list(v)
        out[name] = v
    return out


def load_atlas(solver, member, seg, labels, smoothing_factor=1.0):
    """SolverInterface::readAtlas with a segmentation: split the label image (an array or a NetCDF
    path), upload the maps, smooth them on the GPU with sigma = smoothing_factor * 2 pi / n0, then
    derive mask, k, rho, the preconditioner and Phi (MatProp::setValuesCustom)."""
    if isinstance(seg, (str, bytes)):
        seg = read_netcdf(seg)
    maps = split_segmentation(seg, labels)
    for k in ("wm", "gm", "csf", "vt"):
        solver.set_field(member, k, maps[k])
    sigma = smoothing_factor * 2 * np.pi / solver.n[0]
    if smoothing_factor > 0:
        for k in ("gm", "wm", "vt", "csf"):
            solver.smooth_field(member, k, sigma)
    solver.refresh_tissue(member)
    return maps
