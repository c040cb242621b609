"""Python mirror of the reference's reaction-diffusion operator interface over libglia_rd.so.

Class and method names follow the reference (GLIA): ``DiffusionSolver.solve / precFactor``
(src/pde/DiffusionSolver.cpp), ``PdeOperatorsRD.solveState / solveAdjoint``
(src/pde/PdeOperators.cpp), ``DerivativeOperatorsRD.evaluateObjective / evaluateGradient /
evaluateObjectiveAndGradient`` (src/grad/DerivativeOperatorsRD.cpp).  They share one
``RDSolver`` handle, the counterpart of the reference's Tumor + SpectralOperators + operators
object graph.  Every call goes to the HIP library; fields may be numpy arrays (host) or torch
tensors on the GPU (device pointers are passed through without copies).
"""
import ctypes

import numpy as np

from . import _lib

_F = _lib.FIELD


class GliaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("glia_rd error %d: %s" % (code, msg))
        self.code = code


def _is_device(a):
    return hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


class RDSolver:
    """Handle over an ensemble of members sharing one grid (nmembers = 1: the plain solve)."""

    def __init__(self, n, nmembers=1, dt=0.5, nt=1, rtol=1e-6, maxit=5000, checkpoint_slots=0,
                 adjoint_store=False, phi_store=False, k_gm_wm=0.0, k_glm_wm=0.0, r_gm_wm=0.0,
                 r_glm_wm=0.0, beta=0.0, k_scale=1.0, rho_scale=1.0, smoothing_factor=1.0, stream=None):
        self.L = _lib.load()
        if np.isscalar(n):
            n = (int(n),) * 3
        self.n = tuple(int(v) for v in n)
        self.shape = self.n
        self.ng = self.n[0] * self.n[1] * self.n[2]
        self.nmembers = int(nmembers)
        self.dt, self.nt = float(dt), int(nt)
        self.h3 = (2 * np.pi) ** 3 / self.ng
        c = _lib.Config()
        self.L.glia_rd_config_init(ctypes.byref(c))
        c.n[0], c.n[1], c.n[2] = self.n
        c.nmembers = self.nmembers
        c.dt, c.nt, c.order = self.dt, self.nt, 2
        c.pcg_rtol, c.pcg_maxit = float(rtol), int(maxit)
        c.checkpoint_slots = int(checkpoint_slots)
        c.adjoint_store, c.phi_store = int(bool(adjoint_store)), int(bool(phi_store))
        c.k_gm_wm, c.k_glm_wm, c.r_gm_wm, c.r_glm_wm = map(float, (k_gm_wm, k_glm_wm, r_gm_wm, r_glm_wm))
        c.beta, c.k_scale, c.rho_scale = float(beta), float(k_scale), float(rho_scale)
        c.smoothing_factor = float(smoothing_factor)
        c.stream = stream
        self.cfg = c
        h = ctypes.c_void_p()
        rc = self.L.glia_rd_create(ctypes.byref(c), ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.glia_rd_last_error(h).decode() if h.value else "create failed"
            if h.value:
                self.L.glia_rd_destroy(h)
            self.h = None
            raise GliaError(rc, msg)
        self.np = [0] * self.nmembers

    def close(self):
        if getattr(self, "h", None):
            self.L.glia_rd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise GliaError(rc, self.L.glia_rd_last_error(self.h).decode())

    # ---- field marshalling ---------------------------------------------------
    def _src(self, a):
        """(pointer, mem_kind, keepalive) of a field given as numpy or a torch GPU tensor."""
        if a is None:
            return None, _lib.HOST, None
        if _is_device(a):
            import torch
            t = a.detach().to(torch.float32).contiguous()
            if t.numel() != self.ng:
                raise ValueError("field has %d entries, grid has %d" % (t.numel(), self.ng))
            return ctypes.c_void_p(t.data_ptr()), _lib.DEVICE, t
        arr = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if arr.size != self.ng:
            raise ValueError("field has %d entries, grid has %d" % (arr.size, self.ng))
        return arr.ctypes.data_as(ctypes.c_void_p), _lib.HOST, arr

    def set_field(self, member, which, a):
        p, kind, keep = self._src(a)
        self._chk(self.L.glia_rd_set_field(self.h, member, _F[which], p, kind))

    def get_field(self, member, which, out=None):
        if out is not None and _is_device(out):
            self._chk(self.L.glia_rd_get_field(self.h, member, _F[which], ctypes.c_void_p(out.data_ptr()), _lib.DEVICE))
            return out
        o = np.empty(self.ng, dtype=np.float32)
        self._chk(self.L.glia_rd_get_field(self.h, member, _F[which], o.ctypes.data_as(ctypes.c_void_p), _lib.HOST))
        return o.reshape(self.shape)

    # ---- setup (MatProp, DiffCoef, ReacCoef, Phi, Obs) -------------------------
    def set_tissue(self, member, wm=None, gm=None, csf=None, vt=None):
        """MatProp::setValuesCustom + updateTumorCoefficients (src/mat/MatProp.cpp:134-203)."""
        ps = [self._src(a) for a in (wm, gm, csf, vt)]
        kinds = {k for (_, k, keep) in ps if keep is not None}
        if len(kinds) > 1:
            raise ValueError("tissue fields must all be host or all device arrays")
        kind = kinds.pop() if kinds else _lib.HOST
        self._chk(self.L.glia_rd_set_tissue(self.h, member, *[p for (p, _, _) in ps], kind))

    def set_tissue_sinusoidal(self, member=0):
        self._chk(self.L.glia_rd_set_tissue_sinusoidal(self.h, member))

    def set_coefficients(self, member, kappa, rho):
        self._chk(self.L.glia_rd_set_coefficients(self.h, member, float(kappa), float(rho)))

    def set_k_sinusoidal(self, member, scale):
        self._chk(self.L.glia_rd_set_k_sinusoidal(self.h, member, float(scale)))

    def precFactor(self):
        self._chk(self.L.glia_rd_prec_factor(self.h))

    @property
    def ctx_dt(self):
        v = ctypes.c_double()
        self._chk(self.L.glia_rd_get_ctx_dt(self.h, ctypes.byref(v)))
        return v.value

    @ctx_dt.setter
    def ctx_dt(self, dt):
        self._chk(self.L.glia_rd_set_ctx_dt(self.h, float(dt)))

    @property
    def beta(self):
        """OptimizerSettings::beta_ (L2 regularization of c0)."""
        v = ctypes.c_double()
        self._chk(self.L.glia_rd_get_beta(self.h, ctypes.byref(v)))
        return v.value

    @beta.setter
    def beta(self, beta):
        self._chk(self.L.glia_rd_set_beta(self.h, float(beta)))

    def set_basis(self, member, centers, sigma):
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1)
        self._chk(self.L.glia_rd_set_basis(self.h, member, c.size // 3, c.ctypes.data_as(ctypes.c_void_p), float(sigma)))
        self.np[member] = c.size // 3

    def set_data(self, member, d1, obs_mask=None, obs_thr=0.0):
        p, kind, k1 = self._src(d1)
        q, kind2, k2 = self._src(obs_mask)
        if obs_mask is not None and kind2 != kind:
            raise ValueError("d1 and obs_mask must both be host or both device arrays")
        self._chk(self.L.glia_rd_set_data(self.h, member, p, q, float(obs_thr), kind))

    # ---- operators ---------------------------------------------------------------
    def phi_apply(self, p):
        """Phi::apply for every member; p: list of per-member coefficient vectors."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).reshape(-1) for v in p]))
        self._chk(self.L.glia_rd_phi_apply(self.h, flat.ctypes.data_as(ctypes.c_void_p)))

    def phi_apply_transpose(self, which="work"):
        out = np.zeros(max(sum(self.np), 1))
        self._chk(self.L.glia_rd_phi_apply_transpose(self.h, _F[which], out.ctypes.data_as(ctypes.c_void_p)))
        res, off = [], 0
        for k in self.np:
            res.append(out[off:off + k].copy())
            off += k
        return res

    def diffusion_solve(self, which, tau):
        self._chk(self.L.glia_rd_diffusion_solve(self.h, _F[which], float(tau)))

    def apply_d(self, which_in, which_out):
        self._chk(self.L.glia_rd_apply_d(self.h, _F[which_in], _F[which_out]))

    def fft_r2c(self, member, f):
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(-1)
        out = np.zeros(self.n[0] * self.n[1] * (self.n[2] // 2 + 1) * 2, dtype=np.float32)
        self._chk(self.L.glia_rd_fft_r2c(self.h, member, f.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p)))
        out = out.reshape(self.n[0], self.n[1], self.n[2] // 2 + 1, 2)
        return out[..., 0] + 1j * out[..., 1]

    def fft_c2r(self, member, F):
        F = np.ascontiguousarray(F, dtype=np.complex64)
        buf = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1), dtype=np.float32).reshape(-1)
        out = np.zeros(self.ng, dtype=np.float32)
        self._chk(self.L.glia_rd_fft_c2r(self.h, member, buf.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p)))
        return out.reshape(self.shape)

    def solve_state(self, linearized=0):
        self._chk(self.L.glia_rd_solve_state(self.h, int(linearized)))

    def solve_adjoint(self, linearized=1):
        self._chk(self.L.glia_rd_solve_adjoint(self.h, int(linearized)))

    def _flat_x(self, x):
        xs = [np.asarray(v, dtype=np.float64).reshape(-1) for v in x]
        for m, v in enumerate(xs):
            if v.size != self.np[m] + 2:
                raise ValueError("member %d: x needs np + 2 = %d entries" % (m, self.np[m] + 2))
        return np.ascontiguousarray(np.concatenate(xs))

    def _split(self, flat):
        res, off = [], 0
        for k in self.np:
            res.append(flat[off:off + k + 2].copy())
            off += k + 2
        return res

    def objective(self, x, mode=_lib.MODE_ALL):
        flat = self._flat_x(x)
        J = np.zeros(self.nmembers)
        self._chk(self.L.glia_rd_objective(self.h, mode, flat.ctypes.data_as(ctypes.c_void_p), J.ctypes.data_as(ctypes.c_void_p)))
        return J

    def objective_and_gradient(self, x, mode=_lib.MODE_ALL):
        flat = self._flat_x(x)
        J = np.zeros(self.nmembers)
        dJ = np.zeros_like(flat)
        self._chk(self.L.glia_rd_objective_and_gradient(self.h, mode, flat.ctypes.data_as(ctypes.c_void_p),
                                                        J.ctypes.data_as(ctypes.c_void_p), dJ.ctypes.data_as(ctypes.c_void_p)))
        return J, self._split(dJ)

    def hessian(self, x, diffusivity=False):
        """Gauss-Newton Hessian matvec (DerivativeOperatorsRD::evaluateHessian; kappa blocks with
        diffusivity) at the linearization point of the last evaluation; x and the result are
        per-member [p | kappa | rho] vectors."""
        flat = self._flat_x(x)
        y = np.zeros_like(flat)
        self._chk(self.L.glia_rd_hessian(self.h, int(bool(diffusivity)), flat.ctypes.data_as(ctypes.c_void_p),
                                         y.ctypes.data_as(ctypes.c_void_p)))
        return self._split(y)

    # ---- time history (resizeTimeHistory, public c_ / p_ / c_half_) ------------------------
    def set_time(self, dt, nt):
        """TumorParameters dt_ / nt_ with PdeOperatorsRD::resizeTimeHistory: the history is
        reallocated for nt; stored states are dropped."""
        self._chk(self.L.glia_rd_set_time(self.h, float(dt), int(nt)))
        self.dt, self.nt = float(dt), int(nt)

    def keep_adjoint_history(self, enable=True):
        """Keep p_[0..nt] during solveAdjoint (store-all mode only)."""
        self._chk(self.L.glia_rd_keep_adjoint_history(self.h, int(bool(enable))))

    def history(self, member, kind, index):
        """c_[index] (kind "state"), c_half_[index] ("half") or p_[index] ("adjoint") of a member."""
        k = {"state": 0, "half": 1, "adjoint": 2}[kind]
        o = np.empty(self.ng, dtype=np.float32)
        self._chk(self.L.glia_rd_get_history(self.h, int(member), k, int(index), o.ctypes.data_as(ctypes.c_void_p),
                                             _lib.HOST))
        return o.reshape(self.shape)

    def set_secondary(self, member, kt):
        """DiffCoef::setSecondaryCoefficients: k~ = kt (wm + k_gm gm + k_glm csf), for solveState(2)."""
        self._chk(self.L.glia_rd_set_secondary(self.h, int(member), float(kt)))

    def grad_diffusion(self):
        """gradDiffusion of every member from the last adjoint sweep (g_kappa)."""
        out = np.zeros(self.nmembers)
        self._chk(self.L.glia_rd_grad_diffusion(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def grad_reaction(self):
        """gradReaction of every member from the last adjoint sweep (g_rho)."""
        out = np.zeros(self.nmembers)
        self._chk(self.L.glia_rd_grad_reaction(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def check_gradient(self, x, mode=_lib.MODE_ALL, direction=None):
        """DerivativeOperators::checkGradient: the Taylor test at x along direction (per-member
        vectors like x; None: seeded uniform [0, 1) entries) for h = 1 .. 1e-5.  Returns per member
        (|J(x + h d) - J(x) - h <dJ, d>|, (J(x + h d) - J(x)) / h - <dJ, d>), six values each."""
        flat = self._flat_x(x)
        d = None if direction is None else self._flat_x(direction)
        err = np.zeros(6 * self.nmembers)
        gd = np.zeros(6 * self.nmembers)
        self._chk(self.L.glia_rd_check_gradient(self.h, int(mode), flat.ctypes.data_as(ctypes.c_void_p),
                                                None if d is None else d.ctypes.data_as(ctypes.c_void_p),
                                                err.ctypes.data_as(ctypes.c_void_p), gd.ctypes.data_as(ctypes.c_void_p)))
        return err.reshape(self.nmembers, 6), gd.reshape(self.nmembers, 6)

    def stats(self, member=0):
        s = _lib.Stats()
        self._chk(self.L.glia_rd_stats(self.h, member, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.Stats._fields_}

    def smooth_field(self, member, which, sigma):
        """weierstrassSmoother of a stored real field, in place."""
        self._chk(self.L.glia_rd_smooth_field(self.h, int(member), _lib.FIELD[which], float(sigma)))

    def data_support(self, member, which="d1", threshold=0.05, sigma=2 * np.pi / 256, gvf=0.0):
        """Phi::setGaussians(Vec data) (src/mat/Phi.cpp:615-940): the lattice centers whose ball
        holds more than gvf of its voxels above threshold in a stored field, as a (k, 3) array in
        index order (counted on the device).  Defaults are the reference's phi_sigma_data_driven,
        data_threshold and gaussian_vol_frac."""
        k = ctypes.c_int(0)
        args = (self.h, int(member), _lib.FIELD[which], float(threshold), float(sigma), float(gvf))
        rc = self.L.glia_rd_data_support(*args, 0, None, ctypes.byref(k))
        if rc != 0 and k.value == 0:
            self._chk(rc)
        out = np.zeros((max(k.value, 1), 3), dtype=np.float64)
        self._chk(self.L.glia_rd_data_support(*args, k.value, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(k)))
        return out[: k.value]

    def refresh_tissue(self, member):
        """Mask, k, rho, preconditioner and Phi from the stored tissue fields."""
        self._chk(self.L.glia_rd_refresh_tissue(self.h, int(member)))

    def autotune(self, cache_path=None):
        """Time the variants of the forward divergence passes on this handle and keep the faster
        (cache_path: a text file of earlier choices, read first and appended to).  Returns
        (x_pass_split, y_pass_split)."""
        ch = (ctypes.c_int * 2)()
        path = None if cache_path is None else str(cache_path).encode()
        self._chk(self.L.glia_rd_autotune(self.h, path, ctypes.cast(ch, ctypes.c_void_p)))
        return int(ch[0]), int(ch[1])

    def step_info(self):
        """(steps launched from captured graphs, steps run again eagerly, graphs cached)."""
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        self._chk(self.L.glia_rd_step_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def use_graphs(self, enable=True):
        """Replay CG iterations from captured HIP graphs (default) or launch kernel by kernel."""
        self._chk(self.L.glia_rd_use_graphs(self.h, int(bool(enable))))

    def profile(self, enable=True):
        self._chk(self.L.glia_rd_profile(self.h, int(bool(enable))))

    def profile_get(self):
        """{kernel: (launches, total ms)} of the timed CG kernels since profile(True)."""
        out = {}
        for k, name in enumerate(_lib.KERNELS):
            ms = ctypes.c_double()
            cnt = ctypes.c_longlong()
            self._chk(self.L.glia_rd_profile_get(self.h, k, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name] = (cnt.value, ms.value)
        return out

    def reset_stats(self):
        self._chk(self.L.glia_rd_reset_stats(self.h))

    def sync(self):
        self._chk(self.L.glia_rd_sync(self.h))


def objective_and_gradient_group(solvers, xs, modes=None):
    """Evaluate J and dJ of several RDSolver handles (possibly of different grids) concurrently.

    solvers: list of distinct RDSolver; xs[i]: list of per-member parameter vectors of solvers[i];
    modes[i]: evaluation mode (default MODE_ALL).  Returns (J list, dJ list, ms) with ms[i] the
    wall time of handle i in milliseconds."""
    nh = len(solvers)
    if nh == 0 or len(xs) != nh:
        raise ValueError("one x list per solver")
    L = solvers[0].L
    modes = [_lib.MODE_ALL] * nh if modes is None else list(modes)
    flats = [s._flat_x(x) for s, x in zip(solvers, xs)]
    Js = [np.zeros(s.nmembers) for s in solvers]
    dJs = [np.zeros_like(f) for f in flats]
    ms = np.zeros(nh)
    vp = ctypes.c_void_p
    H = (vp * nh)(*[s.h.value for s in solvers])
    M = (ctypes.c_int * nh)(*modes)
    X = (vp * nh)(*[f.ctypes.data for f in flats])
    JP = (vp * nh)(*[j.ctypes.data for j in Js])
    DP = (vp * nh)(*[d.ctypes.data for d in dJs])
    rc = L.glia_rd_objective_and_gradient_group(nh, ctypes.cast(H, vp), ctypes.cast(M, vp), ctypes.cast(X, vp),
                                                ctypes.cast(JP, vp), ctypes.cast(DP, vp), ms.ctypes.data_as(vp))
    if rc != 0:
        msgs = [s.L.glia_rd_last_error(s.h).decode() for s in solvers]
        raise GliaError(rc, "; ".join(m for m in msgs if m))
    return Js, [s._split(d) for s, d in zip(solvers, dJs)], ms


def objective_and_gradient_lockstep(solvers, xs, modes=None):
    """The same evaluation of several store-all handles in lockstep from this thread: time step i of
    every handle runs as one graph launch once each handle has its steps captured.  J and dJ equal
    each handle's own objective_and_gradient bit for bit.  Returns (J list, dJ list, spanned) with
    spanned the number of time steps launched as one graph across handles."""
    nh = len(solvers)
    if nh == 0 or len(xs) != nh:
        raise ValueError("one x list per solver")
    L = solvers[0].L
    modes = [_lib.MODE_ALL] * nh if modes is None else list(modes)
    flats = [s._flat_x(x) for s, x in zip(solvers, xs)]
    Js = [np.zeros(s.nmembers) for s in solvers]
    dJs = [np.zeros_like(f) for f in flats]
    spanned = ctypes.c_longlong(0)
    vp = ctypes.c_void_p
    H = (vp * nh)(*[s.h.value for s in solvers])
    M = (ctypes.c_int * nh)(*modes)
    X = (vp * nh)(*[f.ctypes.data for f in flats])
    JP = (vp * nh)(*[j.ctypes.data for j in Js])
    DP = (vp * nh)(*[d.ctypes.data for d in dJs])
    rc = L.glia_rd_objective_and_gradient_lockstep(nh, ctypes.cast(H, vp), ctypes.cast(M, vp), ctypes.cast(X, vp),
                                                   ctypes.cast(JP, vp), ctypes.cast(DP, vp),
                                                   ctypes.cast(ctypes.byref(spanned), vp))
    if rc != 0:
        msgs = [s.L.glia_rd_last_error(s.h).decode() for s in solvers]
        raise GliaError(rc, "; ".join(m for m in msgs if m))
    return Js, [s._split(d) for s, d in zip(solvers, dJs)], spanned.value


class DiffusionSolver:
    """DiffusionSolver (src/pde/DiffusionSolver.cpp): CN half step by PCG on one field."""

    def __init__(self, solver, field="work"):
        self.s = solver
        self.field = field
        self.ksp_itr_ = 0

    def precFactor(self):
        self.s.precFactor()

    def solve(self, c, dt):
        """Solve in place on a host array (numpy) or GPU tensor c; returns c."""
        for m in range(self.s.nmembers):
            self.s.set_field(m, self.field, c if self.s.nmembers == 1 else c[m])
        self.s.diffusion_solve(self.field, dt)
        self.ksp_itr_ = self.s.stats(0)["last_iterations"]
        if _is_device(c):
            self.s.get_field(0, self.field, out=c)
            return c
        return self.s.get_field(0, self.field)


class _History:
    """A read-only view of one history member (c_, c_half_ or p_) of member 0: h[i] is a field."""

    def __init__(self, solver, kind, length):
        self.s, self.kind, self.length = solver, kind, length

    def __len__(self):
        return self.length()

    def __getitem__(self, i):
        n = self.length()
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError(i)
        return self.s.history(0, self.kind, i)


class PdeOperatorsRD:
    """PdeOperatorsRD (src/pde/PdeOperators.cpp): solveState / solveAdjoint on the handle's c0,
    the public time history c_, p_, c_half_ (views; p_ kept in store-all mode), resizeTimeHistory
    and reset."""

    def __init__(self, solver, keep_adjoint=True):
        self.s = solver
        self.diff_ksp_itr_state_ = 0
        self.diff_ksp_itr_adj_ = 0
        if keep_adjoint and solver.cfg.checkpoint_slots == 0:
            solver.keep_adjoint_history(True)  # the reference always keeps p_
        self.c_ = _History(solver, "state", lambda: self.s.nt + 1)
        self.p_ = _History(solver, "adjoint", lambda: self.s.nt + 1)
        self.c_half_ = _History(solver, "half", lambda: self.s.nt if self.s.cfg.adjoint_store else 0)

    def resizeTimeHistory(self, dt, nt):
        self.s.set_time(dt, nt)

    def reset(self, dt, nt):
        """PdeOperatorsRD::reset (src/pde/PdeOperators.cpp:6-15): resizeTimeHistory."""
        self.resizeTimeHistory(dt, nt)

    def solveState(self, linearized=0):
        self.s.reset_stats()
        self.s.solve_state(linearized)
        self.diff_ksp_itr_state_ = self.s.stats(0)["pcg_iterations"]

    def solveAdjoint(self, linearized=1):
        self.s.reset_stats()
        self.s.solve_adjoint(linearized)
        self.diff_ksp_itr_adj_ = self.s.stats(0)["pcg_iterations"]


class DerivativeOperatorsRD:
    """DerivativeOperatorsRD (src/grad/DerivativeOperatorsRD.cpp:7-226); x = [p | kappa | rho]."""

    def __init__(self, solver, mode=_lib.MODE_ALL):
        self.s = solver
        self.mode = mode

    def _x(self, x):
        return [x] if self.s.nmembers == 1 and np.ndim(x[0]) == 0 else x

    def evaluateObjective(self, x):
        J = self.s.objective(self._x(x), self.mode)
        return J[0] if self.s.nmembers == 1 else J

    def evaluateGradient(self, x):
        _, dJ = self.s.objective_and_gradient(self._x(x), self.mode)
        return dJ[0] if self.s.nmembers == 1 else dJ

    def evaluateObjectiveAndGradient(self, x):
        J, dJ = self.s.objective_and_gradient(self._x(x), self.mode)
        return (J[0], dJ[0]) if self.s.nmembers == 1 else (J, dJ)

    def gradDiffusion(self):
        """g_kappa of the last adjoint sweep (src/grad/DerivativeOperators.cpp:189-256)."""
        g = self.s.grad_diffusion()
        return g[0] if self.s.nmembers == 1 else g

    def gradReaction(self):
        """g_rho of the last adjoint sweep (src/grad/DerivativeOperators.cpp:258-321)."""
        g = self.s.grad_reaction()
        return g[0] if self.s.nmembers == 1 else g

    def checkGradient(self, x, direction=None):
        """The Taylor test of src/grad/DerivativeOperators.cpp:369-450 (see RDSolver.check_gradient)."""
        d = None if direction is None else self._x(direction)
        err, gd = self.s.check_gradient(self._x(x), self.mode, d)
        return (err[0], gd[0]) if self.s.nmembers == 1 else (err, gd)

    def reset(self, solver=None, mode=None):
        """DerivativeOperators::reset (src/grad/DerivativeOperators.cpp:168-186): adopt new operators."""
        if solver is not None:
            self.s = solver
        if mode is not None:
            self.mode = mode
