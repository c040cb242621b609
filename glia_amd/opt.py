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


def hard_threshold(x, s, labels=None, weights=None, thresh=1e-3, double_mode=False):
    """hardThreshold: indices of the s largest entries (plain, or per connected component when
    labels (1..ncomp per entry) and weights (ncomp) are given)."""
    L = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.full(max(int(s), 1) * 2 + x.size, -1, dtype=np.int32)
    if labels is None:
        k = L.glia_hard_threshold(x.ctypes.data_as(ctypes.c_void_p), x.size, int(s), out.ctypes.data_as(ctypes.c_void_p))
    else:
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        k = L.glia_hard_threshold_components(x.ctypes.data_as(ctypes.c_void_p), x.size, int(s),
                                             lab.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p),
                                             w.size, float(thresh), int(bool(double_mode)),
                                             out.ctypes.data_as(ctypes.c_void_p))
    if k < 0:
        raise RuntimeError("hard threshold failed with code %d" % k)
    return out[:k].tolist()


STOPS = {0: None, 1: "max_it", 2: "jtol", 3: "ftol"}


def sparse_til_settings(til=None, rd=None, **kw):
    """SparseTILOptimizer defaults, overridden by keyword; til/rd: dicts for the inner solves."""
    s = _lib.SparseTilSettings()
    _lib.load().glia_sparse_til_settings_init(ctypes.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    for name, d in (("til", til), ("rd", rd)):
        for k, v in (d or {}).items():
            setattr(getattr(s, name), k, v)
    return s


def parse_config(path, settings_in=None, sparse_in=None):
    """The optimizer keys of a reference config file (setParameter, src/utils/IO.cpp:82-100,
    171-172) over the defaults (or the given structs): (glia_opt_settings, glia_sparse_til_settings)."""
    st = settings() if settings_in is None else settings_in
    sp = sparse_til_settings() if sparse_in is None else sparse_in
    rc = _lib.load().glia_opt_parse_config(str(path).encode(), ctypes.byref(st), ctypes.byref(sp))
    if rc:
        raise OSError("cannot parse config %s (code %d)" % (path, rc))
    return st, sp


def _til_results(res, nprob):
    one = lambda o: dict(iterations=o.iterations, fevals=o.fevals, reason=REASONS.get(o.reason, o.reason), f=o.f,
                         gnorm=o.gnorm, gnorm0=o.gnorm0)
    return [dict(iterations=res[m].iterations, stop=STOPS.get(res[m].stop), nnz=res[m].nnz, J=res[m].J,
                 til=one(res[m].til), rd=one(res[m].rd)) for m in range(nprob)]


def _support(sup, np_):
    return [[int(i) for i in row if i >= 0] for row in sup.reshape(-1, np_)]


def sparse_til(eval_fn, c0_fn, x0, np_, labels=None, weights=None, **kw):
    """Sparse TIL over Python operations: eval_fn(mode, beta, X, want_grad) -> (f (nprob,),
    G (nprob, np+2) or None); c0_fn(P (nprob, np)) -> max of Phi p per problem.  Returns
    (x, supports, results)."""
    L = _lib.load()
    x = np.array(x0, dtype=np.float64, ndmin=2, copy=True)
    nprob, n = x.shape
    assert n == np_ + 2
    err = []

    def ev(_ctx, mode, beta, xp, fp, gp):
        try:
            X = np.ctypeslib.as_array(xp, shape=(nprob * n,)).reshape(nprob, n).copy()
            f, g = eval_fn(mode, beta, X, bool(gp))
            np.ctypeslib.as_array(fp, shape=(nprob,))[:] = np.asarray(f, dtype=np.float64).reshape(nprob)
            if gp:
                np.ctypeslib.as_array(gp, shape=(nprob * n,))[:] = np.asarray(g, dtype=np.float64).reshape(-1)
            return 0
        except Exception as e:
            err.append(e)
            return 1

    def c0(_ctx, pp, cp):
        try:
            P = np.ctypeslib.as_array(pp, shape=(nprob * np_,)).reshape(nprob, np_).copy()
            np.ctypeslib.as_array(cp, shape=(nprob,))[:] = np.asarray(c0_fn(P), dtype=np.float64).reshape(nprob)
            return 0
        except Exception as e:
            err.append(e)
            return 1

    efn, cfn = _lib.EVAL_FN(ev), _lib.C0_FN(c0)
    sup = np.full(nprob * np_, -1, dtype=np.int32)
    lab = None if labels is None else np.ascontiguousarray(np.broadcast_to(labels, (nprob, np_)), dtype=np.int32)
    w = None if weights is None else np.ascontiguousarray(np.atleast_2d(weights), dtype=np.float64)
    ncomp = 0 if w is None else w.shape[1]
    res = (_lib.SparseTilResult * nprob)()
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    rc = L.glia_sparse_til(nprob, np_, x.ctypes.data_as(ctypes.c_void_p), sup.ctypes.data_as(ctypes.c_void_p),
                           ptr(lab), ptr(w), ncomp, efn, cfn, None, ctypes.byref(sparse_til_settings(**kw)),
                           ctypes.cast(res, ctypes.c_void_p))
    if err:
        raise err[0]
    if rc:
        raise RuntimeError("glia_sparse_til failed with code %d" % rc)
    return x, _support(sup, np_), _til_results(res, nprob)


def rd_sparse_til(solver, x0, labels=None, weights=None, **kw):
    """Sparse TIL on every member of an RDSolver handle (glia_rd_sparse_til)."""
    flat = solver._flat_x(x0)
    nm = solver.nmembers
    np_ = len(flat) // nm - 2
    sup = np.full(nm * np_, -1, dtype=np.int32)
    lab = None if labels is None else np.ascontiguousarray(np.broadcast_to(labels, (nm, np_)), dtype=np.int32)
    w = None if weights is None else np.ascontiguousarray(np.atleast_2d(weights), dtype=np.float64)
    ncomp = 0 if w is None else w.shape[1]
    res = (_lib.SparseTilResult * nm)()
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    rc = solver.L.glia_rd_sparse_til(solver.h, flat.ctypes.data_as(ctypes.c_void_p), sup.ctypes.data_as(ctypes.c_void_p),
                                     ptr(lab), ptr(w), ncomp, ctypes.byref(sparse_til_settings(**kw)),
                                     ctypes.cast(res, ctypes.c_void_p))
    solver._chk(rc)
    return solver._split(flat), _support(sup, np_), _til_results(res, nm)


def gauss_newton(fg, hv, x0, lb=None, ub=None, **kw):
    """Bound-constrained Gauss-Newton-Krylov over Python callbacks: fg(X) -> (f, G) as for
    minimize; hv(V) -> H V at the point of the last fg call (V, result: (nprob, n)).  Returns
    (x, results)."""
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
        except Exception as e:
            err.append(e)
            return 1

    def hcb(_ctx, npb, vp, yp):
        try:
            V = np.ctypeslib.as_array(vp, shape=(npb * n,)).reshape(npb, n).copy()
            np.ctypeslib.as_array(yp, shape=(npb * n,))[:] = np.asarray(hv(V), dtype=np.float64).reshape(-1)
            return 0
        except Exception as e:
            err.append(e)
            return 1

    fn, hn = _lib.FG_FN(cb), _lib.HV_FN(hcb)
    res = (_lib.OptResult * nprob)()
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    rc = L.glia_gauss_newton(nprob, n, x.ctypes.data_as(ctypes.c_void_p), ptr(lo), ptr(hi), fn, hn, None,
                             ctypes.byref(settings(**kw)), ctypes.cast(res, ctypes.c_void_p))
    if err:
        raise err[0]
    if rc:
        raise RuntimeError("glia_gauss_newton failed with code %d" % rc)
    return x, _results(res, nprob)
