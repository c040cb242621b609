"""The inversion loop around the gradient (include/glia_opt.h): bound-constrained limited-memory
quasi-Newton with an Armijo line search and the reference's stopping test (the reference runs TAO's
BLMVM, src/opt/Optimizer.cpp:116-260, with checkConvergenceGrad, src/opt/TaoInterface.cpp:362)."""
import ctypes

import numpy as np

from . import _lib

REASONS = {3: "gatol", 5: "gttol", 6: "steptol", -2: "maxits", -4: "nan", -6: "ls_failure"}


def settings(**kw):
    """OptimizerSettings defaults (newton_maxit 30, gatol 1e-4, opttolgrad 1e-3, 10 L-BFGS vectors,
    10 line-search evaluations), overridden by keyword."""
    s = _lib.OptSettings()
    _lib.load().glia_opt_settings_init(ctypes.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _results(res, nprob):
    return [dict(iterations=res[m].iterations, fevals=res[m].fevals, reason=REASONS.get(res[m].reason, res[m].reason),
                 f=res[m].f, gnorm=res[m].gnorm, gnorm0=res[m].gnorm0) for m in range(nprob)]


def _bounds(b, shape):
    return None if b is None else np.ascontiguousarray(np.broadcast_to(np.asarray(b, dtype=np.float64), shape))


def minimize(fg, x0, lb=None, ub=None, **kw):
    """Minimize nprob problems together. fg(X) with X of shape (nprob, n) returns (f (nprob,),
    G (nprob, n)). x0: (nprob, n) or (n,). Returns (x, results)."""
    L = _lib.load()
    x = np.array(x0, dtype=np.float64, ndmin=2, copy=True)
    nprob, n = x.shape
    lo, hi = _bounds(lb, x.shape), _bounds(ub, x.shape)
    err = []

    def cb(_ctx, npb, xp, fp, gp):
        try:
            X = np.ctypeslib.as_array(xp, shape=(npb * n,)).reshape(npb, n).copy()
            f, g = fg(X)
            np.ctypeslib.as_array(fp, shape=(npb,))[:] = np.asarray(f, dtype=np.float64).reshape(npb)
            np.ctypeslib.as_array(gp, shape=(npb * n,))[:] = np.asarray(g, dtype=np.float64).reshape(-1)
            return 0
        except Exception as e:  # reported after the call
            err.append(e)
            return 1

    fn = _lib.FG_FN(cb)
    res = (_lib.OptResult * nprob)()
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    rc = L.glia_blmvm(nprob, n, x.ctypes.data_as(ctypes.c_void_p), ptr(lo), ptr(hi), fn, None,
                      ctypes.byref(settings(**kw)), ctypes.cast(res, ctypes.c_void_p))
    if err:
        raise err[0]
    if rc:
        raise RuntimeError("glia_blmvm failed with code %d" % rc)
    return x, _results(res, nprob)


def invert(solver, x0, lb=None, ub=None, mode=_lib.MODE_KAPPA_RHO, **kw):
    """Invert every member of an RDSolver handle on the GPU (glia_rd_invert).  x0: per-member
    [p | kappa | rho] vectors (all members with the same number of bases)."""
    flat = solver._flat_x(x0)
    n = len(flat) // solver.nmembers
    lo, hi = _bounds(lb, (solver.nmembers, n)), _bounds(ub, (solver.nmembers, n))
    res = (_lib.OptResult * solver.nmembers)()
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    rc = solver.L.glia_rd_invert(solver.h, int(mode), flat.ctypes.data_as(ctypes.c_void_p), ptr(lo), ptr(hi),
                                 ctypes.byref(settings(**kw)), ctypes.cast(res, ctypes.c_void_p))
    solver._chk(rc)
    return solver._split(flat), _results(res, solver.nmembers)
