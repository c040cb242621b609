"""Python mirror of the reference's semi-Lagrangian advection solver (SemiLagrangianSolver,
src/pde/AdvectionSolver.cpp:279-704) over libglia_rd.so (include/glia_adv.h).

Fields may be numpy arrays (host; copied) or float32 torch tensors on the GPU (device pointers are
passed through).  There is no CPU path: every call goes to the HIP library.
"""
import ctypes

import numpy as np

from . import _lib
from .rd import GliaError, _is_device

MASS_CONSERVATION = 1
PURE_ADVECTION = 2


class SemiLagrangianSolver:
    """Semi-Lagrangian transport on an n0 x n1 x n2 periodic grid (``advection_mode_`` 1 or 2)."""

    def __init__(self, n, stream=None):
        self.L = _lib.load()
        self.shape = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.ng = int(np.prod(self.shape))
        self.h = ctypes.c_void_p()
        nn = (ctypes.c_int * 3)(*self.shape)
        st = None if stream is None else ctypes.c_void_p(stream)
        rc = self.L.glia_sl_create(nn, st, ctypes.byref(self.h))
        if rc != 0:
            self.h = None
            raise GliaError(rc, "glia_sl_create failed")
        self.advection_mode_ = PURE_ADVECTION

    def close(self):
        if getattr(self, "h", None):
            self.L.glia_sl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise GliaError(rc, self.L.glia_sl_last_error(self.h).decode())

    def _arg(self, a, copy=False):
        """(pointer, mem_kind, keepalive) of a field: numpy (host; a private copy when copy is set)
        or a float32 torch GPU tensor."""
        if _is_device(a):
            import torch
            if a.dtype != torch.float32 or not a.is_contiguous() or a.numel() != self.ng:
                raise ValueError("device fields must be contiguous float32 tensors of %d values" % self.ng)
            return ctypes.c_void_p(a.data_ptr()), _lib.DEVICE, a
        h = np.array(a, dtype=np.float32, order="C", copy=True) if copy else np.ascontiguousarray(a, dtype=np.float32)
        h = h.reshape(-1)
        if h.size != self.ng:
            raise ValueError("field has %d values, grid has %d" % (h.size, self.ng))
        return h.ctypes.data_as(ctypes.c_void_p), _lib.HOST, h

    def set_velocity(self, vx, vy, vz):
        """Tumor::velocity_ for the next solves (marks the trajectories stale)."""
        a = [self._arg(v) for v in (vx, vy, vz)]
        kinds = {k for _, k, _ in a}
        if len(kinds) != 1:
            raise ValueError("velocity components must all be host or all device arrays")
        self._chk(self.L.glia_sl_set_velocity(self.h, a[0][0], a[1][0], a[2][0], kinds.pop()))

    def invalidate(self):
        """trajectoryIsComputed_ = false."""
        self._chk(self.L.glia_sl_invalidate(self.h))

    def computeTrajectories(self, dt):
        self._chk(self.L.glia_sl_compute_trajectories(self.h, float(dt)))

    def offsets(self):
        """Departure point minus grid point, in grid units, as three host arrays."""
        q = [np.empty(self.ng, dtype=np.float32) for _ in range(3)]
        self._chk(self.L.glia_sl_get_offsets(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in q], _lib.HOST))
        return [a.reshape(self.shape) for a in q]

    def interpolate(self, f, out=None):
        """interpolate(output, input): f at the departure points."""
        p, kind, keep = self._arg(f)
        if kind == _lib.DEVICE:
            import torch
            o = torch.empty_like(f) if out is None else out
            self._chk(self.L.glia_sl_interpolate(self.h, p, ctypes.c_void_p(o.data_ptr()), kind))
            return o
        o = np.empty(self.ng, dtype=np.float32)
        self._chk(self.L.glia_sl_interpolate(self.h, p, o.ctypes.data_as(ctypes.c_void_p), kind))
        return o.reshape(self.shape)

    def solve(self, scalar, dt, mode=None):
        """solve(scalar, velocity, dt): one step; in place for a device tensor, a new array for
        host input."""
        mode = self.advection_mode_ if mode is None else int(mode)
        p, kind, keep = self._arg(scalar, copy=True)
        self._chk(self.L.glia_sl_solve(self.h, p, float(dt), mode, kind))
        if kind == _lib.DEVICE:
            return scalar
        return keep.reshape(self.shape)

    def advect(self, fields, modes, dt, out=None):
        """One step of several device fields on one set of trajectories (out of place)."""
        import torch
        n = len(fields)
        outs = [torch.empty_like(f) for f in fields] if out is None else list(out)
        for f in list(fields) + outs:
            self._arg(f)
            if not _is_device(f):
                raise ValueError("advect takes device tensors")
        ins = (ctypes.c_void_p * n)(*[f.data_ptr() for f in fields])
        ots = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        md = (ctypes.c_int * n)(*[int(m) for m in modes])
        self._chk(self.L.glia_sl_advect(self.h, n, ctypes.cast(ins, ctypes.c_void_p), ctypes.cast(ots, ctypes.c_void_p),
                                        ctypes.cast(md, ctypes.c_void_p), float(dt)))
        return outs

    def sync(self):
        self._chk(self.L.glia_sl_sync(self.h))
