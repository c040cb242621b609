"""ctypes wrapper around the FP64 oracle (oracle/glia_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.

Each method names the reference function it restates; see glia_oracle.c for
the file:line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


def build():
    """Compile liborc.so from glia_oracle.c with the Makefile beside it."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(_HERE, "glia_oracle.c")):
        build()
    L = ctypes.CDLL(_LIB)
    vp = ctypes.c_void_p
    i, d = ctypes.c_int, ctypes.c_double
    sig = {
        "orc_create": (vp, [i, i, i]),
        "orc_destroy": (None, [vp]),
        "orc_max_threads": (i, []),
        "orc_set_threads": (None, [i]),
        "orc_set_verbose": (None, [i]),
        "orc_r2c": (None, [vp, _dp, _cp]),
        "orc_c2r": (None, [vp, _cp, _dp]),
        "orc_grad": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_div": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_sl_interp": (None, [vp, _dp, _dp, _dp, _dp, _dp]),
        "orc_ms_sources": (None, [ctypes.c_int64] + [_dp] * 8 + [d, _dp]),
        "orc_ms_update_coefficients": (None, [ctypes.c_int64] + [_dp] * 6 + [d, d, d, d]),
        "orc_ms_segmentation": (None, [ctypes.c_int64] + [_dp] * 7),
        "orc_el_material": (None, [vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_el_matvec": (None, [vp, _dp, _dp, _dp, _dp, _dp]),
        "orc_el_force": (None, [vp, _dp, d, i, _dp]),
        "orc_el_prec": (None, [vp, _dp, _dp, _dp]),
        "orc_el_solve": (i, [vp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_double)]),
        "orc_sl_trajectories": (None, [vp, _dp, _dp, _dp, d, _dp, _dp, _dp]),
        "orc_sl_solve": (None, [vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, d, i]),
        "orc_smooth": (None, [vp, _dp, _dp, d]),
        "orc_me_solve_state": (i, [vp, _dp, _dp, _dp, _dp, _dp, vp, _dp, _dp, _dp, d, i, _dp, d, _dp]),
        "orc_me_solve_state_f": (i, [vp, _dp, _dp, _dp, _dp, _dp, vp, _dp, _dp, _dp, d, i, _dp, d, _dp, vp, i,
                                     ctypes.POINTER(ctypes.c_int)]),
        "orc_me_features": (None, [vp, i, d] + [_dp] * 11),
        "orc_stress_point": (None, [_dp, d, d, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
        "orc_ms_solve_state": (i, [vp] + [_dp] * 12 + [d, i, _dp, _dp, d, _dp, _dp, _dp]),
        "orc_set_tissue": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_set_tissue_sinusoidal": (None, [vp]),
        "orc_set_k": (None, [vp, d, d, d]),
        "orc_set_k_sinusoidal": (None, [vp, d]),
        "orc_set_rho": (None, [vp, d, d, d]),
        "orc_get_fields": (None, [vp, vp, vp, vp, vp]),
        "orc_apply_d": (None, [vp, _dp, _dp]),
        "orc_prec_factor": (None, [vp]),
        "orc_diffusion_solve": (i, [vp, _dp, d]),
        "orc_last_reason": (i, [vp]),
        "orc_set_solver": (None, [vp, d, i]),
        "orc_set_ctx_dt": (None, [vp, d]),
        "orc_get_ctx_dt": (d, [vp]),
        "orc_get_prec_dt": (d, [vp]),
        "orc_reaction": (None, [vp, _dp, vp, d]),
        "orc_set_time": (None, [vp, d, i, i]),
        "orc_set_c0": (None, [vp, _dp]),
        "orc_set_pt": (None, [vp, _dp]),
        "orc_get_c0": (None, [vp, _dp]),
        "orc_get_ct": (None, [vp, _dp]),
        "orc_get_p0": (None, [vp, _dp]),
        "orc_get_c_hist": (None, [vp, i, _dp]),
        "orc_get_p_hist": (None, [vp, i, _dp]),
        "orc_get_c_half": (None, [vp, i, _dp]),
        "orc_get_ksp": (None, [vp, vp, vp, vp]),
        "orc_solve_state": (i, [vp, i]),
        "orc_solve_adjoint": (i, [vp, i]),
        "orc_phi_mesh": (i, [i, _dp, d, d, _dp]),
        "orc_data_support": (i, [vp, vp, d, d, d, i, vp, vp]),
        "orc_set_phi": (None, [vp, i, _dp, d, d, i]),
        "orc_get_phi": (None, [vp, i, _dp]),
        "orc_phi_apply": (None, [vp, _dp, _dp]),
        "orc_phi_applyT": (None, [vp, _dp, _dp]),
        "orc_set_data": (None, [vp, _dp, vp, d]),
        "orc_set_opt": (None, [vp, d, d, d]),
        "orc_objective": (d, [vp, _dp, i]),
        "orc_objective_and_gradient": (None, [vp, _dp, i, _dp, _dp]),
        "orc_hessian_p": (None, [vp, _dp, _dp]),
        "orc_hessian": (None, [vp, _dp, _dp, i]),
        "orc_set_secondary": (None, [vp, d]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def max_threads():
    return _load().orc_max_threads()


def set_threads(n):
    _load().orc_set_threads(int(n))


def set_verbose(on):
    """One stderr line per time step of the state and adjoint solves (progress of long runs)."""
    _load().orc_set_verbose(int(bool(on)))


def phi_mesh(np_, cm, sigma, spacing_factor):
    """Phi::phiMesh (src/mat/Phi.cpp:148-234)."""
    out = np.zeros(3 * max(np_, 1) * 2, dtype=np.float64)
    k = _load().orc_phi_mesh(int(np_), _c(cm), float(sigma), float(spacing_factor), out)
    return out[: 3 * k].reshape(k, 3)


def data_support(data, threshold, sigma, gvf=0.0, counts=False):
    """Phi::setGaussians(Vec data) (src/mat/Phi.cpp:615-940) on one process (orc_data_support):
    the kept lattice centers as a (k, 3) array, and with counts=True also the per-candidate voxel
    counts in lattice order."""
    f = np.ascontiguousarray(data, dtype=np.float32)
    n = np.array(f.shape, dtype=np.int32)
    lib = _load()
    cnt = np.zeros(f.size, dtype=np.int32)
    k = lib.orc_data_support(f.ctypes.data_as(ctypes.c_void_p), n.ctypes.data_as(ctypes.c_void_p),
                             float(threshold), float(sigma), float(gvf), 0, None,
                             cnt.ctypes.data_as(ctypes.c_void_p))
    if k < 0:
        raise ValueError("data support lattice spacing below one cell")
    out = np.zeros(3 * max(k, 1), dtype=np.float64)
    lib.orc_data_support(f.ctypes.data_as(ctypes.c_void_p), n.ctypes.data_as(ctypes.c_void_p), float(threshold),
                         float(sigma), float(gvf), k, out.ctypes.data_as(ctypes.c_void_p), None)
    c = out[: 3 * k].reshape(k, 3)
    return (c, cnt) if counts else c


def stress_point(g, mu, lam):
    """computeStressQuants at one voxel (orc_stress_point): g[3 a + b] = d u_a / d x_b; returns
    (det(I + grad u), tr S, max shear) with S = lam tr(E) I + 2 mu E, E the linearized strain."""
    jac, tr, tau = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _load().orc_stress_point(_c(g), float(mu), float(lam), ctypes.byref(jac), ctypes.byref(tr), ctypes.byref(tau))
    return jac.value, tr.value, tau.value


class Oracle:
    """FP64 restatement of the reference RD path on an n0 x n1 x n2 periodic grid."""

    # evaluation modes of evaluateObjective/Gradient: which parameters the reference inverts
    MODE_P_KAPPA = 0    # flag_reaction_inv_ = false: c0 = Phi p; g_p, g_kappa
    MODE_KAPPA_RHO = 1  # flag_reaction_inv_ = true: c0 fixed; g_kappa, g_rho
    MODE_ALL = 2        # p, kappa, rho in one pass

    def __init__(self, n):
        if np.isscalar(n):
            n = (int(n),) * 3
        self.n = tuple(int(v) for v in n)
        self.L = _load()
        self.h = self.L.orc_create(*self.n)
        if not self.h:
            raise ValueError("unsupported grid %r (lengths must be 2^a 3^b 5^c)" % (self.n,))
        self.ng = self.n[0] * self.n[1] * self.n[2]
        self.shape = self.n
        self.sshape = (self.n[0], self.n[1], self.n[2] // 2 + 1)
        self.np = 0

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _f(self, a):
        a = _c(a).reshape(-1)
        if a.size != self.ng:
            raise ValueError("field has %d entries, grid has %d" % (a.size, self.ng))
        return a

    def _new(self):
        return np.zeros(self.ng, dtype=np.float64)

    # ---- spectral operators -------------------------------------------------
    def r2c(self, f):
        out = np.zeros(self.sshape, dtype=np.complex128).reshape(-1)
        self.L.orc_r2c(self.h, self._f(f), out)
        return out.reshape(self.sshape)

    def c2r(self, F):
        F = np.ascontiguousarray(F, dtype=np.complex128).reshape(-1).copy()
        out = self._new()
        self.L.orc_c2r(self.h, F, out)
        return out.reshape(self.shape)

    def grad(self, x):
        g = [self._new() for _ in range(3)]
        self.L.orc_grad(self.h, self._f(x), *g)
        return [v.reshape(self.shape) for v in g]

    def div(self, dx, dy, dz):
        out = self._new()
        self.L.orc_div(self.h, self._f(dx), self._f(dy), self._f(dz), out)
        return out.reshape(self.shape)

    # elasticity: vectors are three fields stacked as (3,) + shape
    EL_DEFAULTS = (2100.0, 0.4, 8000.0, 0.45, 15000.0, 0.45, 100.0, 0.1, 1e-2, 1e2)

    def el_material(self, c, wm, gm, csf, vt, bg, prm=None):
        """computeMaterialProperties: (mu, lam, screen, avg[3])."""
        prm = np.array(self.EL_DEFAULTS if prm is None else prm, dtype=np.float64)
        mu, lam, scr = self._new(), self._new(), self._new()
        avg = np.zeros(3)
        self.L.orc_el_material(self.h, *[self._f(a) for a in (c, wm, gm, csf, vt, bg)], prm, mu, lam, scr, avg)
        return mu.reshape(self.shape), lam.reshape(self.shape), scr.reshape(self.shape), avg

    def _vec(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != 3 * self.ng:
            raise ValueError("vector field has %d entries, grid needs %d" % (u.size, 3 * self.ng))
        return u

    def el_matvec(self, mu, lam, scr, u):
        f = np.zeros(3 * self.ng)
        self.L.orc_el_matvec(self.h, self._f(mu), self._f(lam), self._f(scr), self._vec(u), f)
        return f.reshape((3,) + self.shape)

    def el_force(self, c, ff=1e5, use_tanh=True):
        """computeForce: ff grad(S c) (times tanh(S c))."""
        f = np.zeros(3 * self.ng)
        self.L.orc_el_force(self.h, self._f(c), float(ff), int(bool(use_tanh)), f)
        return f.reshape((3,) + self.shape)

    def el_prec(self, avg, f):
        u = np.zeros(3 * self.ng)
        self.L.orc_el_prec(self.h, np.ascontiguousarray(avg, dtype=np.float64), self._vec(f), u)
        return u.reshape((3,) + self.shape)

    def el_solve(self, mu, lam, scr, avg, rhs, u0):
        """KSPCG on the elasticity operator: (u, iterations, reason, last preconditioned residual norm)."""
        u = self._vec(u0).copy()
        reason, rn = ctypes.c_int(), ctypes.c_double()
        its = self.L.orc_el_solve(self.h, self._f(mu), self._f(lam), self._f(scr),
                                  np.ascontiguousarray(avg, dtype=np.float64), self._vec(rhs), u,
                                  ctypes.byref(reason), ctypes.byref(rn))
        return u.reshape((3,) + self.shape), its, reason.value, rn.value

    def me_solve_state(self, c0, wm, gm, vt, csf, bg, u=None, v=None, mri=None, ff=1e5, use_tanh=True,
                       el_prm=None, max_cfl=14.0, features=False):
        """PdeOperatorsMassEffect::solveState(0): the mass-effect forward of model 4 over nt + 1
        steps, with the diffusion settings of set_tissue/set_k/set_rho/set_time.  Returns a dict of
        c, wm, gm, vt, csf, mri (None when not given), u and v (3-component arrays) and info
        (steps, cfl_stop, cfl_max, el_iterations, diff_iterations, volume_stop).  With features the
        loop runs its feature branch and the dict holds "features" (one row of 130 per step)."""
        f = {k: np.array(a, dtype=np.float64, order="C").reshape(-1).copy()
             for k, a in (("c", c0), ("wm", wm), ("gm", gm), ("vt", vt), ("csf", csf))}
        m = None if mri is None else np.array(mri, dtype=np.float64, order="C").reshape(-1).copy()
        uu = np.zeros(3 * self.ng) if u is None else self._vec(u).copy()
        vv = np.zeros(3 * self.ng) if v is None else self._vec(v).copy()
        prm = np.array(self.EL_DEFAULTS if el_prm is None else el_prm, dtype=np.float64)
        info = np.zeros(6)
        feats, rows = None, ctypes.c_int(0)
        if features:
            feats = np.zeros((self.nt + 1, 130))
        self.L.orc_me_solve_state_f(self.h, f["c"], f["wm"], f["gm"], f["vt"], f["csf"],
                                    None if m is None else m.ctypes.data_as(ctypes.c_void_p), self._f(bg), uu, vv,
                                    float(ff), int(bool(use_tanh)), prm, float(max_cfl), info,
                                    None if feats is None else feats.ctypes.data_as(ctypes.c_void_p),
                                    0 if feats is None else self.nt + 1, ctypes.byref(rows))
        out = {k: a.reshape(self.shape) for k, a in f.items()}
        out["mri"] = None if m is None else m.reshape(self.shape)
        out["u"] = uu.reshape((3,) + self.shape)
        out["v"] = vv.reshape((3,) + self.shape)
        out["info"] = {"steps": int(info[0]), "cfl_stop": bool(info[1]), "cfl_max": float(info[2]),
                       "el_iterations": int(info[3]), "diff_iterations": int(info[4]),
                       "volume_stop": bool(info[5])}
        if features:
            out["features"] = feats[: rows.value]
        return out

    def me_features(self, it, dt, c, wm, gm, vt, csf, bg, u, v, mu, lam):
        """computeBiophysicalFeatures of one state (orc_me_features): a row of 130 values.  D and
        rho are the oracle's current k and rho fields."""
        row = np.zeros(130)
        self.L.orc_me_features(self.h, int(it), float(dt), self._f(c), self._f(wm), self._f(gm), self._f(vt),
                               self._f(csf), self._f(bg), self._vec(u), self._vec(v), self._f(mu), self._f(lam), row)
        return row

    MS_DEFAULTS = (4.0, 1.0, 5.0, 0.1, 0.01, 0.8, 0.6)

    def ms_solve_state(self, c0, wm, gm, vt, csf, bg, nec=None, u=None, v=None, ff=1e5, use_tanh=True,
                       el_prm=None, ms_prm=None, max_cfl=8.0):
        """PdeOperatorsMultiSpecies::solveState(0): the multispecies mass-effect forward of model 5
        over nt + 1 steps.  Returns a dict of p, i, nec, ox, wm, gm, vt, csf, u, v, edema, seg and info."""
        f = {k: np.array(a, dtype=np.float64, order="C").reshape(-1).copy()
             for k, a in (("wm", wm), ("gm", gm), ("vt", vt), ("csf", csf))}
        for k in ("p", "i", "ox"):
            f[k] = np.zeros(self.ng)
        f["nec"] = np.zeros(self.ng) if nec is None else np.array(nec, dtype=np.float64).reshape(-1).copy()
        uu = np.zeros(3 * self.ng) if u is None else self._vec(u).copy()
        vv = np.zeros(3 * self.ng) if v is None else self._vec(v).copy()
        prm = np.array(self.EL_DEFAULTS if el_prm is None else el_prm, dtype=np.float64)
        mprm = list(self.MS_DEFAULTS if ms_prm is None else ms_prm)
        if len(mprm) == 7:
            mprm.append(0.001)  # TumorParameters::invasive_threshold_
        mprm = np.array(mprm, dtype=np.float64)
        info = np.zeros(5)
        ed, seg = np.zeros(self.ng), np.zeros(self.ng)
        self.L.orc_ms_solve_state(self.h, self._f(c0), f["p"], f["i"], f["nec"], f["ox"], f["wm"], f["gm"], f["vt"],
                                  f["csf"], self._f(bg), uu, vv, float(ff), int(bool(use_tanh)), prm, mprm,
                                  float(max_cfl), info, ed, seg)
        out = {k: a.reshape(self.shape) for k, a in f.items()}
        out["edema"] = ed.reshape(self.shape)
        out["seg"] = seg.reshape(self.shape)
        out["u"] = uu.reshape((3,) + self.shape)
        out["v"] = vv.reshape((3,) + self.shape)
        out["info"] = {"steps": int(info[0]), "cfl_stop": bool(info[1]), "cfl_max": float(info[2]),
                       "el_iterations": int(info[3]), "diff_iterations": int(info[4])}
        return out

    def sl_interp(self, f, q):
        """f at the points x + q (q: three offset fields in grid units), cubic Lagrange, periodic."""
        out = self._new()
        self.L.orc_sl_interp(self.h, self._f(f), *[self._f(a) for a in q], out)
        return out.reshape(self.shape)

    def sl_trajectories(self, v, dt):
        """Offsets (grid units) of the RK2 departure points of the velocity v over dt."""
        q = [self._new() for _ in range(3)]
        self.L.orc_sl_trajectories(self.h, *[self._f(a) for a in v], float(dt), *q)
        return [a.reshape(self.shape) for a in q]

    def sl_solve(self, c, v, q, dt, mode=2):
        """One semi-Lagrangian step of c along v with departure offsets q (mode 2: pure advection,
        1: mass conservation)."""
        out = np.array(c, dtype=np.float64, order="C").reshape(-1).copy()
        self.L.orc_sl_solve(self.h, out, *[self._f(a) for a in v], *[self._f(a) for a in q], float(dt), int(mode))
        return out.reshape(self.shape)

    def smooth(self, c, sigma):
        out = self._new()
        self.L.orc_smooth(self.h, self._f(c), out, float(sigma))
        return out.reshape(self.shape)

    # ---- material / coefficients -------------------------------------------
    def set_tissue(self, wm=None, gm=None, csf=None, vt=None, bg=None):
        arrs = [None if a is None else self._f(a) for a in (wm, gm, csf, vt, bg)]
        self._keep = arrs
        self.L.orc_set_tissue(self.h, *[_ptr(a) for a in arrs])

    def set_tissue_sinusoidal(self):
        self.L.orc_set_tissue_sinusoidal(self.h)

    def set_k(self, k_scale, k_gm_wm=0.0, k_glm_wm=0.0):
        self.L.orc_set_k(self.h, float(k_scale), float(k_gm_wm), float(k_glm_wm))

    def set_k_sinusoidal(self, scale):
        self.L.orc_set_k_sinusoidal(self.h, float(scale))

    def set_rho(self, rho, r_gm_wm=0.0, r_glm_wm=0.0):
        self.L.orc_set_rho(self.h, float(rho), float(r_gm_wm), float(r_glm_wm))

    def fields(self):
        k, rho, filt = self._new(), self._new(), self._new()
        kavg = np.zeros(1)
        self.L.orc_get_fields(self.h, _ptr(k), _ptr(rho), _ptr(filt), _ptr(kavg))
        return dict(k=k.reshape(self.shape), rho=rho.reshape(self.shape), filter=filt.reshape(self.shape), kavg=float(kavg[0]))

    # ---- diffusion solver ---------------------------------------------------
    def set_solver(self, rtol=1e-6, maxit=5000):
        self.L.orc_set_solver(self.h, float(rtol), int(maxit))

    def set_time(self, dt, nt, adjoint_store=False):
        self.L.orc_set_time(self.h, float(dt), int(nt), int(bool(adjoint_store)))
        self.dt, self.nt = float(dt), int(nt)

    def set_ctx_dt(self, dt):
        self.L.orc_set_ctx_dt(self.h, float(dt))

    @property
    def ctx_dt(self):
        return self.L.orc_get_ctx_dt(self.h)

    @property
    def prec_dt(self):
        return self.L.orc_get_prec_dt(self.h)

    def prec_factor(self):
        self.L.orc_prec_factor(self.h)

    def apply_d(self, c):
        out = self._new()
        self.L.orc_apply_d(self.h, self._f(c), out)
        return out.reshape(self.shape)

    def diffusion_solve(self, c, tau):
        c = self._f(c).copy()
        its = self.L.orc_diffusion_solve(self.h, c, float(tau))
        return c.reshape(self.shape), its

    @property
    def last_reason(self):
        return self.L.orc_last_reason(self.h)

    def reaction(self, c, dt, clin=None):
        c = self._f(c).copy()
        cl = None if clin is None else self._f(clin)
        self.L.orc_reaction(self.h, c, _ptr(cl), float(dt))
        return c.reshape(self.shape)

    # ---- forward / adjoint --------------------------------------------------
    def set_c0(self, c0):
        self.L.orc_set_c0(self.h, self._f(c0))

    def set_pt(self, pt):
        self.L.orc_set_pt(self.h, self._f(pt))

    def solve_state(self, linearized=0):
        return self.L.orc_solve_state(self.h, int(linearized))

    def solve_adjoint(self, linearized=1):
        return self.L.orc_solve_adjoint(self.h, int(linearized))

    def _get(self, fn, *a):
        out = self._new()
        fn(self.h, *a, out)
        return out.reshape(self.shape)

    def c0(self):
        return self._get(self.L.orc_get_c0)

    def ct(self):
        return self._get(self.L.orc_get_ct)

    def p0(self):
        return self._get(self.L.orc_get_p0)

    def c_hist(self, i):
        return self._get(self.L.orc_get_c_hist, int(i))

    def p_hist(self, i):
        return self._get(self.L.orc_get_p_hist, int(i))

    def c_half(self, i):
        """c_half_[i] (kept with adjoint_store)."""
        return self._get(self.L.orc_get_c_half, int(i))

    def set_secondary(self, kt):
        """DiffCoef::setSecondaryCoefficients: k~ = kt (wm + k_gm gm + k_glm csf)."""
        self.L.orc_set_secondary(self.h, float(kt))

    def ksp(self):
        a, b, c = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int())
        self.L.orc_get_ksp(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return dict(state=a.value, adjoint=b.value, last=c.value)

    # ---- Phi / Obs / data ---------------------------------------------------
    def set_phi(self, centers, sigma, smoothing_factor=1.0, phi_store=False):
        centers = _c(centers).reshape(-1)
        self.np = centers.size // 3
        self.L.orc_set_phi(self.h, self.np, centers, float(sigma), float(smoothing_factor), int(bool(phi_store)))

    def phi(self, i):
        return self._get(self.L.orc_get_phi, int(i))

    def phi_apply(self, p):
        out = self._new()
        self.L.orc_phi_apply(self.h, _c(p), out)
        return out.reshape(self.shape)

    def phi_applyT(self, v):
        out = np.zeros(max(self.np, 1))
        self.L.orc_phi_applyT(self.h, self._f(v), out)
        return out[: self.np]

    def set_data(self, d1, obs_mask=None, thr=0.0):
        m = None if obs_mask is None else self._f(obs_mask)
        self.L.orc_set_data(self.h, self._f(d1), _ptr(m), float(thr))

    def set_opt(self, beta=0.0, k_scale=1.0, rho_scale=1.0):
        self.L.orc_set_opt(self.h, float(beta), float(k_scale), float(rho_scale))

    def objective(self, x, mode=MODE_ALL):
        return self.L.orc_objective(self.h, _c(x), int(mode))

    def objective_and_gradient(self, x, mode=MODE_ALL):
        J = np.zeros(1)
        dJ = np.zeros(self.np + 2)
        self.L.orc_objective_and_gradient(self.h, _c(x), int(mode), J, dJ)
        return float(J[0]), dJ

    def hessian(self, xt, diffusivity=False):
        """Gauss-Newton Hessian matvec on [p | kappa | rho] (kappa blocks with diffusivity)."""
        y = np.zeros(self.np + 2)
        self.L.orc_hessian(self.h, _c(xt), y, int(bool(diffusivity)))
        return y

    def hessian_p(self, ptilde):
        """Gauss-Newton Hessian matvec, p block (needs the state history of the last evaluation)."""
        y = np.zeros(self.np)
        self.L.orc_hessian_p(self.h, _c(ptilde), y)
        return y


MS_DEFAULTS = (4.0, 1.0, 5.0, 0.1, 0.01, 0.8, 0.6)  # death_rate, ox_source, ox_consumption, alpha_0, beta_0, ox_inv, ox_hypoxia


def ms_sources(p, i, nec, ox, gm, wm, rho, di, dt, prm=MS_DEFAULTS):
    """computeSources of the multispecies model on flat arrays: the updated (p, i, nec, ox, gm, wm)."""
    L = _load()
    out = [np.array(a, dtype=np.float64).reshape(-1).copy() for a in (p, i, nec, ox, gm, wm)]
    L.orc_ms_sources(out[0].size, *out, _c(rho).reshape(-1), _c(di).reshape(-1), float(dt),
                     np.ascontiguousarray(prm, dtype=np.float64))
    return out


def ms_update_coefficients(bg, gm, vt, csf, rho, k, r_gm_wm, k_gm_wm):
    """updateReacAndDiffCoefficients on flat arrays: (rho field, k field)."""
    L = _load()
    n = np.asarray(bg).size
    rf, kf = np.zeros(n), np.zeros(n)
    L.orc_ms_update_coefficients(n, rf, kf, *[_c(a).reshape(-1) for a in (bg, gm, vt, csf)], float(rho), float(k),
                                 float(r_gm_wm), float(k_gm_wm))
    return rf, kf


def ms_segmentation(bg, gm, wm, vt, csf, c):
    """computeSegmentation labels (bg 0, c 1, wm 6, gm 5, vt 7, csf 8) on flat arrays."""
    L = _load()
    n = np.asarray(bg).size
    seg = np.zeros(n)
    L.orc_ms_segmentation(n, *[_c(a).reshape(-1) for a in (bg, gm, wm, vt, csf, c)], seg)
    return seg

