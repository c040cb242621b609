"""Error behaviour of the C ABI on a live handle: every misuse returns a negative GLIA_RD_ERR_* code
with a message (the reference's methods return a PetscErrorCode that CHKERRQ propagates), and the
handle stays usable afterwards."""
import ctypes

import numpy as np
import pytest

from glia_amd import _lib

pytestmark = pytest.mark.gpu


def test_misuse_returns_typed_errors_and_the_handle_survives():
    from glia_amd import rd, synth
    L = _lib.load()
    print("case error paths 32^3", flush=True)
    pr = synth.problem(32, 3, nbases=4, kappa=0.05, rho=8.0, sigma=2 * np.pi / 32)
    s = rd.RDSolver(32, nmembers=2, dt=0.02, nt=3, k_gm_wm=0.2, r_gm_wm=0.2)
    for m in range(2):
        s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(m, 0.05, 8.0)
        s.set_basis(m, pr["centers"], pr["sigma"])
    h = s.h
    vp = ctypes.c_void_p
    x = np.zeros(2 * 6)
    J = np.zeros(2)
    g = np.zeros_like(x)
    buf = np.zeros(32 ** 3, dtype=np.float32)

    def err(rc, code, text=None):
        assert rc == code, (rc, code)
        msg = L.glia_rd_last_error(h).decode()
        assert msg and (text is None or text in msg), msg

    # the adjoint and the Hessian need a state history first
    err(L.glia_rd_solve_adjoint(h, 1), _lib.ERR_NOHISTORY)
    err(L.glia_rd_hessian(h, 0, x.ctypes.data_as(vp), g.ctypes.data_as(vp)), _lib.ERR_NOHISTORY)
    # bad member, field, mode, linearization, basis, beta
    err(L.glia_rd_set_coefficients(h, 2, 0.05, 8.0), _lib.ERR_ARG, "member")
    err(L.glia_rd_set_kappa(h, -1, 0.05), _lib.ERR_ARG, "member")
    err(L.glia_rd_get_field(h, 0, 99, buf.ctypes.data_as(vp), _lib.HOST), _lib.ERR_ARG, "field")
    err(L.glia_rd_get_field(h, 5, _lib.FIELD["ct"], buf.ctypes.data_as(vp), _lib.HOST), _lib.ERR_ARG)
    err(L.glia_rd_objective(h, 3, x.ctypes.data_as(vp), J.ctypes.data_as(vp)), _lib.ERR_ARG)
    err(L.glia_rd_solve_state(h, 3), _lib.ERR_ARG, "linearized")
    err(L.glia_rd_solve_adjoint(h, 3), _lib.ERR_ARG, "linearized")
    # the tangent-linear equations linearize around a state history
    err(L.glia_rd_solve_state(h, 1), _lib.ERR_NOHISTORY)
    err(L.glia_rd_solve_adjoint(h, 2), _lib.ERR_NOHISTORY)
    centers = np.ascontiguousarray(pr["centers"]).reshape(-1)
    err(L.glia_rd_set_basis(h, 0, 4, centers.ctypes.data_as(vp), -1.0), _lib.ERR_ARG)
    err(L.glia_rd_set_beta(h, -1.0), _lib.ERR_ARG, "beta")
    err(L.glia_rd_set_data(h, 0, None, None, 0.0, _lib.HOST), _lib.ERR_ARG, "data")
    # the handle still evaluates, and the errors left no state behind
    d1 = (pr["wm"] > 0.6).astype(np.float32)
    for m in range(2):
        s.set_data(m, d1, None, -0.99)
    xs = [np.concatenate([pr["p"], [0.05, 8.0]])] * 2
    Ja, ga = s.objective_and_gradient(xs)
    assert np.all(np.isfinite(Ja)) and Ja[0] == Ja[1] and np.array_equal(ga[0], ga[1])
    # after an evaluation the adjoint and the Hessian run
    assert L.glia_rd_solve_adjoint(h, 1) == 0
    y = s.hessian(xs)
    assert np.all(np.isfinite(np.concatenate(y)))
    s.close()


def test_python_layer_rejects_wrong_shapes_before_the_library():
    from glia_amd import rd
    s = rd.RDSolver(32, nt=1)
    with pytest.raises(ValueError):
        s.objective_and_gradient([np.zeros(5)])  # basis size 0: x needs 2 entries
    with pytest.raises(ValueError):
        s.set_field(0, "ct", np.zeros(10, dtype=np.float32))  # a field must cover the grid
    with pytest.raises(rd.GliaError):
        s.solve_adjoint(1)  # no state history yet
    s.close()
