"""Python mirror of the reference's mass-effect forwards over libglia_rd.so (include/glia_me.h): model 4
(PdeOperatorsMassEffect::solveState, src/pde/PdeOperatorsMassEffect.cpp:421-743) and the multispecies
model 5 (PdeOperatorsMultiSpecies::solveState, src/pde/PdeOperatorsMultiSpecies.cpp:235-508).  Fields are
numpy arrays on the grid; they are staged through device memory for the call.  There is no CPU
path: every step runs in the HIP library."""
import ctypes

import numpy as np

from . import _lib
from .rd import GliaError

NFEATURES = 130  # GLIA_ME_NFEATURES: t, the tumor ratio, 16 statistics of 8 fields


def feature_names(L=None):
    """The column names of a feature row (glia_me_feature_name), as the reference's CSV header."""
    L = L or _lib.load()
    return [L.glia_me_feature_name(k).decode() for k in range(NFEATURES)]


class MassEffectSolver:
    """The engine (dt, nt, k, rho and their gm/wm ratios), the elasticity solver and the advection
    solver of one tumor on an n0 x n1 x n2 grid."""

    def __init__(self, n, dt, nt, k=0.0, rho=0.0, k_gm_wm=0.0, r_gm_wm=0.0, rtol=1e-6, maxit=5000):
        self.L = _lib.load()
        self.shape = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.ng = int(np.prod(self.shape))
        self.k, self.rho = float(k), float(rho)
        self.nt = int(nt)
        c = _lib.Config()
        self.L.glia_rd_config_init(ctypes.byref(c))
        c.n[0], c.n[1], c.n[2] = self.shape
        c.nmembers, c.dt, c.nt, c.order = 1, float(dt), int(nt), 2
        c.pcg_rtol, c.pcg_maxit = float(rtol), int(maxit)
        c.k_gm_wm, c.r_gm_wm = float(k_gm_wm), float(r_gm_wm)
        self.h = ctypes.c_void_p()
        rc = self.L.glia_me_create(ctypes.byref(c), ctypes.byref(self.h))
        if rc != 0:
            msg = self.L.glia_me_last_error(self.h).decode() if self.h.value else "glia_me_create failed"
            if self.h.value:
                self.L.glia_me_destroy(self.h)
            self.h = None
            raise GliaError(rc, msg)
        self.rd = self.L.glia_me_rd(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.L.glia_me_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _f(self, a):
        h = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if h.size != self.ng:
            raise ValueError("field has %d values, grid has %d" % (h.size, self.ng))
        return h

    def set_tissue(self, wm, gm, csf, vt):
        """The initial tissue and DiffCoef/ReacCoef::setValues on it (the preconditioner's mean k)."""
        arrs = [self._f(a) for a in (wm, gm, csf, vt)]
        rc = self.L.glia_rd_set_tissue(self.rd, 0, *[a.ctypes.data for a in arrs], 0)
        if rc == 0:
            rc = self.L.glia_rd_set_coefficients(self.rd, 0, self.k, self.rho)
        if rc != 0:
            raise GliaError(rc, self.L.glia_rd_last_error(self.rd).decode())

    def solve_state(self, c0, wm, gm, vt, csf, bg, u=None, v=None, mri=None, forcing_factor=1e5, use_tanh=True,
                    max_cfl=None, features=False, **el):
        """solveState(0) from c0 and the tissue (max_cfl None: the reference's limit 14); el overrides
        glia_el_params fields.  Returns a dict
        of c, wm, gm, vt, csf, mri (None when not given), u, v (shape (3,) + grid) and info.  With
        features the loop runs its feature branch (glia_me_solve_state_features): the dict also holds
        "features", one row of GLIA_ME_NFEATURES values per step, and "feature_names"."""
        import torch

        p = _lib.MeParams()
        self.L.glia_me_params_init(ctypes.byref(p))
        p.forcing_factor, p.use_tanh = float(forcing_factor), int(bool(use_tanh))
        p.max_cfl = 0.0 if max_cfl is None else float(max_cfl)
        for key, val in el.items():
            if not hasattr(p.el, key):
                raise KeyError("unknown elasticity parameter %s" % key)
            setattr(p.el, key, float(val))
        dev = torch.device("cuda")

        def put(a):
            return torch.from_numpy(self._f(a).copy()).to(dev)

        f = {k: put(a) for k, a in (("c", c0), ("wm", wm), ("gm", gm), ("vt", vt), ("csf", csf), ("bg", bg))}
        m = None if mri is None else put(mri)
        zero = np.zeros((3, self.ng), np.float32)
        uu = [put(a) for a in (zero if u is None else np.asarray(u, np.float32).reshape(3, -1))]
        vv = [put(a) for a in (zero if v is None else np.asarray(v, np.float32).reshape(3, -1))]
        torch.cuda.synchronize()
        up = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in uu])
        vp = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in vv])
        info = _lib.MeInfo()
        args = (self.h, ctypes.byref(p), f["c"].data_ptr(), f["wm"].data_ptr(), f["gm"].data_ptr(),
                f["vt"].data_ptr(), f["csf"].data_ptr(), None if m is None else m.data_ptr(), f["bg"].data_ptr(),
                up, vp, ctypes.byref(info))
        feats, rows = None, ctypes.c_int(0)
        if features:
            feats = np.zeros((self.nt + 1, NFEATURES))
            rc = self.L.glia_me_solve_state_features(*args, feats.ctypes.data_as(ctypes.c_void_p), self.nt + 1,
                                                     ctypes.byref(rows))
        else:
            rc = self.L.glia_me_solve_state(*args)
        if rc != 0:
            raise GliaError(rc, self.L.glia_me_last_error(self.h).decode())
        out = {k: f[k].cpu().numpy().reshape(self.shape) for k in ("c", "wm", "gm", "vt", "csf")}
        out["mri"] = None if m is None else m.cpu().numpy().reshape(self.shape)
        out["u"] = np.stack([t.cpu().numpy() for t in uu]).reshape((3,) + self.shape)
        out["v"] = np.stack([t.cpu().numpy() for t in vv]).reshape((3,) + self.shape)
        out["info"] = {"steps": info.steps, "cfl_stop": bool(info.cfl_stop), "cfl_max": info.cfl_max,
                       "el_iterations": info.el_iterations, "diff_iterations": info.diff_iterations,
                       "volume_stop": bool(info.volume_stop)}
        if features:
            out["features"] = feats[: rows.value]
            out["feature_names"] = feature_names(self.L)
        return out

    def solve_state_multispecies(self, c0, wm, gm, vt, csf, bg, nec=None, u=None, v=None, forcing_factor=1e5,
                                 use_tanh=True, max_cfl=None, ms=None, **el):
        """PdeOperatorsMultiSpecies::solveState(0) from c0 and the tissue (max_cfl None: the
        reference's limit 8); ms overrides glia_ms_params
        fields (death_rate, ox_source, ...), el the glia_el_params fields.  Returns a dict of p, i,
        nec, ox, wm, gm, vt, csf, u, v, edema and seg (Tumor::species_["edema"] and Tumor::seg_ as
        the loop leaves them) and info."""
        import torch

        p = _lib.MeParams()
        self.L.glia_me_params_init(ctypes.byref(p))
        p.forcing_factor, p.use_tanh = float(forcing_factor), int(bool(use_tanh))
        p.max_cfl = 0.0 if max_cfl is None else float(max_cfl)
        for key, val in el.items():
            if not hasattr(p.el, key):
                raise KeyError("unknown elasticity parameter %s" % key)
            setattr(p.el, key, float(val))
        mp = _lib.MsParams()
        self.L.glia_ms_params_init(ctypes.byref(mp))
        for key, val in (ms or {}).items():
            if not hasattr(mp, key):
                raise KeyError("unknown multispecies parameter %s" % key)
            setattr(mp, key, float(val))
        dev = torch.device("cuda")

        def put(a):
            return torch.from_numpy(self._f(a).copy()).to(dev)

        f = {k: put(a) for k, a in (("c0", c0), ("wm", wm), ("gm", gm), ("vt", vt), ("csf", csf), ("bg", bg))}
        for k in ("p", "i", "ox"):
            f[k] = torch.zeros(self.ng, dtype=torch.float32, device=dev)
        f["nec"] = torch.zeros(self.ng, dtype=torch.float32, device=dev) if nec is None else put(nec)
        zero = np.zeros((3, self.ng), np.float32)
        uu = [put(a) for a in (zero if u is None else np.asarray(u, np.float32).reshape(3, -1))]
        vv = [put(a) for a in (zero if v is None else np.asarray(v, np.float32).reshape(3, -1))]
        torch.cuda.synchronize()
        up = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in uu])
        vp = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in vv])
        info = _lib.MeInfo()
        rc = self.L.glia_me_solve_state_multispecies(
            self.h, ctypes.byref(p), ctypes.byref(mp), *[f[k].data_ptr() for k in
                                                         ("c0", "p", "i", "nec", "ox", "wm", "gm", "vt", "csf", "bg")],
            up, vp, ctypes.byref(info))
        if rc != 0:
            raise GliaError(rc, self.L.glia_me_last_error(self.h).decode())
        out = {k: f[k].cpu().numpy().reshape(self.shape) for k in ("p", "i", "nec", "ox", "wm", "gm", "vt", "csf")}
        ed = np.empty(self.ng, np.float32)
        seg = np.empty(self.ng, np.float32)
        rc = self.L.glia_me_multispecies_outputs(self.h, ed.ctypes.data_as(ctypes.c_void_p),
                                                 seg.ctypes.data_as(ctypes.c_void_p), _lib.HOST)
        if rc != 0:
            raise GliaError(rc, self.L.glia_me_last_error(self.h).decode())
        out["edema"] = ed.reshape(self.shape)
        out["seg"] = seg.reshape(self.shape)
        out["u"] = np.stack([t.cpu().numpy() for t in uu]).reshape((3,) + self.shape)
        out["v"] = np.stack([t.cpu().numpy() for t in vv]).reshape((3,) + self.shape)
        out["info"] = {"steps": info.steps, "cfl_stop": bool(info.cfl_stop), "cfl_max": info.cfl_max,
                       "el_iterations": info.el_iterations, "diff_iterations": info.diff_iterations}
        return out
