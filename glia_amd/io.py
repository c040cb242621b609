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
        elif name in ("atlas_labels", "patient_labels"):
            v = list(v)
        out[name] = v
    return out


def phi_mesh(np_, cm, sigma, spacing_factor=1.5):
    """Phi::phiMesh (src/mat/Phi.cpp:148-234): the bounding-box basis of np_ Gaussians around cm
    (Phi::setGaussians(cm, sigma, spacing_factor, np), :33-73) as a (k, 3) array."""
    c = np.ascontiguousarray(cm, dtype=np.float64)
    if c.size != 3 or int(np_) < 1:
        raise ValueError("phi_mesh needs a 3-vector center and np >= 1")
    out = np.zeros(3 * int(np_), dtype=np.float64)
    k = _lib.load().glia_phi_mesh(int(np_), c.ctypes.data_as(ctypes.c_void_p), float(sigma), float(spacing_factor),
                                  out.ctypes.data_as(ctypes.c_void_p))
    return out[: 3 * k].reshape(k, 3)


def read_phi_mesh(path, max_np=100000):
    """readPhiMesh (src/utils/IO.cpp:717-790): (centers (k, 3), component labels (k,), sigma,
    spacing) from a phi-mesh text file."""
    L = _lib.load()
    c = np.zeros(3 * max_np, dtype=np.float64)
    comps = np.zeros(max_np, dtype=np.int32)
    k, sg, sp = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_double(0)
    rc = L.glia_read_phi_mesh(str(path).encode(), int(max_np), c.ctypes.data_as(ctypes.c_void_p),
                              comps.ctypes.data_as(ctypes.c_void_p), ctypes.byref(k), ctypes.byref(sg),
                              ctypes.byref(sp))
    if rc == -1:
        raise FileNotFoundError(path)
    if rc:
        raise ValueError("%s holds %d centers, more than %d" % (path, k.value, max_np))
    return c[: 3 * k.value].reshape(k.value, 3), comps[: k.value], sg.value, sp.value


def read_concomp(path, max_comp=4096):
    """readConCompDat (src/utils/IO.cpp:929-990): (relative masses (k,), centers of mass (k, 3)) of
    the connected components in a concomp file."""
    w = np.zeros(max_comp, dtype=np.float64)
    c = np.zeros(3 * max_comp, dtype=np.float64)
    k = ctypes.c_int(0)
    rc = _lib.load().glia_read_concomp(str(path).encode(), int(max_comp), w.ctypes.data_as(ctypes.c_void_p),
                                       c.ctypes.data_as(ctypes.c_void_p), ctypes.byref(k))
    if rc == -1:
        raise FileNotFoundError(path)
    if rc:
        raise ValueError("%s holds %d components, more than %d" % (path, k.value, max_comp))
    return w[: k.value], c[: 3 * k.value].reshape(k.value, 3)


def component_sparsity(sparsity_level, weights, thresh_component_weight=1e-3):
    """The sparsity level the reference sets from the component weights (src/Solver.cpp:231-248):
    s * (components of weight >= thresh) + (components of weight in [thresh / 10, thresh))."""
    w = np.asarray(weights, dtype=np.float64)
    t = float(thresh_component_weight)
    nnc = int(np.sum(w >= t))
    nnc_t = int(np.sum((w < t) & (w >= t / 10.0)))
    return int(sparsity_level) * nnc + nnc_t


def edema_obs_mask(tc, ed, lam):
    """createEdemaBasedObservationMask (src/utils/Utils.cpp:567-590): OBS = 1[TC] + lambda 1[B/WT],
    i.e. 1 where tc > 0.99, else 0 where ed > 0.99, else lambda."""
    t = np.ascontiguousarray(tc, dtype=np.float32)
    e = np.ascontiguousarray(ed, dtype=np.float32)
    if t.shape != e.shape:
        raise ValueError("tumor core and edema maps differ in shape")
    out = np.empty_like(t)
    rc = _lib.load().glia_edema_obs_mask(t.ctypes.data_as(ctypes.c_void_p), e.ctypes.data_as(ctypes.c_void_p), t.size,
                                         float(lam), out.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise ValueError("bad tumor core / edema maps (code %d)" % rc)
    return out


def load_atlas(solver, member, seg, labels, smoothing_factor=1.0):
    """SolverInterface::readAtlas with a segmentation: split the label image (an array or a NetCDF
    path), upload the maps, smooth them on the GPU with sigma = smoothing_factor * 2 pi / n0, then
    derive mask, k, rho, the preconditioner and Phi (MatProp::setValuesCustom).  When the labels name
    a tumor core (tc, or nec and en) the split also returns it ("tc", counted as white matter, as the
    reference does) and, with an edema label, the edema ("ed"): the target data of load_data."""
    if isinstance(seg, (str, bytes)):
        seg = read_netcdf(seg)
    with_tumor = labels.get("tc", -1) > 0 or (labels.get("nec", -1) > 0 and labels.get("en", -1) > 0)
    maps = split_segmentation(seg, labels, with_tumor=with_tumor)
    if with_tumor and labels.get("ed", -1) <= 0:
        del maps["ed"]
    for k in ("wm", "gm", "csf", "vt"):
        solver.set_field(member, k, maps[k])
    sigma = smoothing_factor * 2 * np.pi / solver.n[0]
    if smoothing_factor > 0:
        for k in ("gm", "wm", "vt", "csf"):
            solver.smooth_field(member, k, sigma)
    solver.refresh_tissue(member)
    return maps


def load_data(solver, member, cfg, maps=None):
    """SolverInterface::readData (src/SolverInterface.cpp:577-640) for one member: the data d1 from
    cfg["d1_path"], else the tumor core of the segmentation load_atlas split (maps["tc"]) and, when
    that segmentation has edema and cfg["obs_lambda"] >= 0, the observation mask
    OBS = 1[TC] + lambda 1[B/WT] (createEdemaBasedObservationMask); d1 smoothed on the GPU with sigma =
    smoothing_factor_data * 2 pi / n0; obs_threshold_1 taken relative to max d1 when
    obs_threshold_rel; a mask file (obs_mask_path) replaces the mask.  cfg is parse_config's dict (or
    any dict with those keys).  Sets the member's data; returns (d1, mask or None, threshold)."""
    mask = None
    if cfg.get("d1_path"):
        d1 = read_netcdf(cfg["d1_path"]).astype(np.float32)
    elif maps is not None and "tc" in maps:
        d1 = np.asarray(maps["tc"], dtype=np.float32)
        if "ed" in maps and cfg.get("obs_lambda", -1.0) >= 0:
            mask = edema_obs_mask(maps["tc"], maps["ed"], cfg["obs_lambda"])
    else:
        raise ValueError("no target data: neither d1_path nor a tumor label in the segmentation")
    sfd = float(cfg.get("smoothing_factor_data", 1.0))
    if sfd > 0:
        solver.set_field(member, "d1", d1)
        solver.smooth_field(member, "d1", sfd * 2 * np.pi / solver.n[0])
        d1 = solver.get_field(member, "d1")
    thr = float(cfg.get("obs_threshold_1", 0.0))
    if cfg.get("obs_threshold_rel"):
        thr *= float(np.max(d1))
    if cfg.get("obs_mask_path"):
        mask = read_netcdf(cfg["obs_mask_path"]).astype(np.float32)
    solver.set_data(member, d1, mask, thr)
    return d1, mask, thr


def load_c0(solver, member, cfg, c0=None):
    """The c(0) half of SolverInterface::readData (src/SolverInterface.cpp:654-690) for one member:
    c(0) from cfg["d0_path"] (or the array c0), clipped at zero when it has negative values,
    smoothed on the GPU with sigma = smoothing_factor_data_t0 * 2 pi / n0, then scaled to maximum 1;
    stored as the member's initial condition, which the kappa/rho evaluation mode uses as set
    (use_c0_).  Returns (c(0), threshold), the threshold obs_threshold_0 taken relative to the
    maximum before scaling under obs_threshold_rel."""
    if c0 is None:
        if not cfg.get("d0_path"):
            raise ValueError("no c(0): neither d0_path nor an array")
        c0 = read_netcdf(cfg["d0_path"])
    c = np.asarray(c0, dtype=np.float32)
    if c.min() < 0:
        c = np.maximum(c, 0.0).astype(np.float32)
    solver.set_field(member, "c0", c)
    sf = float(cfg.get("smoothing_factor_data_t0", 1.0))
    if sf > 0:
        solver.smooth_field(member, "c0", sf * 2 * np.pi / solver.n[0])
    c = solver.get_field(member, "c0")
    mx = float(c.max())
    if not mx > 0:
        raise ValueError("c(0) has no positive value")
    c = (c * np.float32(1.0 / mx)).astype(np.float32)
    solver.set_field(member, "c0", c)
    thr = float(cfg.get("obs_threshold_0", 0.0))
    if cfg.get("obs_threshold_rel"):
        thr *= mx
    return c, thr


def init_gaussians(solver, member, cfg):
    """SolverInterface::initializeGaussians (src/SolverInterface.cpp:860-905) for one member, after
    load_data: the centers from the phi-mesh file cfg["gaussian_cm_path"] when given, else the
    data-driven support (Phi::setGaussians(Vec data), counted on the GPU) of the map
    cfg["support_data_path"] or, without one, of the member's data d1; sigma is
    cfg["sigma_data_driven"].  Sets the member's basis and returns the centers."""
    sigma = float(cfg.get("sigma_data_driven", 2 * np.pi / 256))
    if cfg.get("gaussian_cm_path"):
        centers = read_phi_mesh(cfg["gaussian_cm_path"])[0]
    else:
        which = "d1"
        if cfg.get("support_data_path"):
            solver.set_field(member, "work", read_netcdf(cfg["support_data_path"]).astype(np.float32))
            which = "work"
        centers = solver.data_support(member, which, float(cfg.get("threshold_data_driven", 0.05)), sigma,
                                      float(cfg.get("gaussian_volume_fraction", 0.0)))
    if len(centers) == 0:
        raise ValueError("no Gaussian centers: the support holds no voxel above the threshold")
    solver.set_basis(member, centers, sigma)
    return centers
