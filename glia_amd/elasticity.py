"""Python mirror of the reference's linear-elasticity solver of the mass-effect model
(VariableLinearElasticitySolver, src/pde/ElasticitySolver.cpp:6-533) over libglia_rd.so
(include/glia_el.h).  Vector fields are arrays of shape (3,) + grid; scalar fields numpy arrays.
There is no CPU path: every call goes to the HIP library."""
import ctypes

import numpy as np

from . import _lib
from .rd import GliaError

REASONS = {2: "rtol", 3: "atol", -3: "maxits", -4: "dtol", -8: "indefinite_pc", -9: "nan", -10: "indefinite_mat"}


class ElasticitySolver:
    def __init__(self, n, stream=None):
        self.L = _lib.load()
        self.shape = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.ng = int(np.prod(self.shape))
        self.h = ctypes.c_void_p()
        nn = (ctypes.c_int * 3)(*self.shape)
        rc = self.L.glia_el_create(nn, None if stream is None else ctypes.c_void_p(stream), ctypes.byref(self.h))
        if rc != 0:
            self.h = None
            raise GliaError(rc, "glia_el_create failed")

    def close(self):
        if getattr(self, "h", None):
            self.L.glia_el_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise GliaError(rc, self.L.glia_el_last_error(self.h).decode())

    def _f(self, a):
        h = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if h.size != self.ng:
            raise ValueError("field has %d values, grid has %d" % (h.size, self.ng))
        return h

    def _vec(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(3, -1)
        if u.shape[1] != self.ng:
            raise ValueError("vector field needs 3 x %d values" % self.ng)
        return [np.ascontiguousarray(u[d]) for d in range(3)]

    @staticmethod
    def _ptrs(comps):
        return (ctypes.c_void_p * 3)(*[c.ctypes.data for c in comps])

    def set_material(self, c, wm, gm, csf, vt, bg, **params):
        """computeMaterialProperties; params override glia_el_params (E_healthy, nu_tumor, ...)."""
        p = _lib.ElParams()
        self.L.glia_el_params_init(ctypes.byref(p))
        for k, v in params.items():
            setattr(p, k, float(v))
        f = [self._f(a) for a in (c, wm, gm, csf, vt, bg)]
        self._chk(self.L.glia_el_set_material(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in f],
                                              ctypes.byref(p), _lib.HOST))

    def material(self):
        """(mu, lam, screen, avg[3])"""
        out = [np.empty(self.ng, np.float32) for _ in range(3)]
        avg = np.zeros(3)
        self._chk(self.L.glia_el_get_material(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in out],
                                              avg.ctypes.data_as(ctypes.c_void_p), _lib.HOST))
        return [a.reshape(self.shape) for a in out] + [avg]

    def apply(self, u):
        """operatorVariableCoefficients: A u."""
        uin, out = self._vec(u), [np.empty(self.ng, np.float32) for _ in range(3)]
        self._chk(self.L.glia_el_apply(self.h, self._ptrs(uin), self._ptrs(out), _lib.HOST))
        return np.stack(out).reshape((3,) + self.shape)

    def precondition(self, f):
        """operatorConstantCoefficients: B f."""
        fin, out = self._vec(f), [np.empty(self.ng, np.float32) for _ in range(3)]
        self._chk(self.L.glia_el_precondition(self.h, self._ptrs(fin), self._ptrs(out), _lib.HOST))
        return np.stack(out).reshape((3,) + self.shape)

    def solve(self, f, u0=None):
        """solve(displacement, rhs): (u, iterations, reason, last preconditioned residual norm)."""
        fin = self._vec(f)
        u = self._vec(np.zeros((3,) + self.shape) if u0 is None else u0)
        u = [a.copy() for a in u]
        its, why, rn = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._chk(self.L.glia_el_solve(self.h, self._ptrs(fin), self._ptrs(u), _lib.HOST, ctypes.byref(its),
                                       ctypes.byref(why), ctypes.byref(rn)))
        return np.stack(u).reshape((3,) + self.shape), its.value, REASONS.get(why.value, why.value), rn.value

    def solve_device(self, f, u):
        """The same solve on device-resident vectors: f and u are float32 torch tensors on the GPU,
        each of three contiguous components (shape (3,) + grid); u holds the initial guess and
        receives the solution.  Returns (iterations, reason, last preconditioned residual norm)."""
        for t in (f, u):
            if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 4 and t.is_contiguous()
                    and tuple(t.shape) == (3,) + self.shape):
                raise ValueError("f and u must be contiguous float32 GPU tensors of shape (3,) + grid")
        fp = (ctypes.c_void_p * 3)(*[f[d].data_ptr() for d in range(3)])
        up = (ctypes.c_void_p * 3)(*[u[d].data_ptr() for d in range(3)])
        its, why, rn = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._chk(self.L.glia_el_solve(self.h, fp, up, _lib.DEVICE, ctypes.byref(its), ctypes.byref(why),
                                       ctypes.byref(rn)))
        return its.value, REASONS.get(why.value, why.value), rn.value

    def force(self, c, forcing_factor=1e5, use_tanh=True):
        """Tumor::computeForce: forcing_factor grad(S c) (times tanh(S c)), shape (3,) + grid."""
        cin, out = self._f(c), [np.empty(self.ng, np.float32) for _ in range(3)]
        self._chk(self.L.glia_el_force(self.h, cin.ctypes.data_as(ctypes.c_void_p), float(forcing_factor),
                                       int(bool(use_tanh)), self._ptrs(out), _lib.HOST))
        return np.stack(out).reshape((3,) + self.shape)

