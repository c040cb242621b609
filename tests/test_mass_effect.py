"""The oracle's mass-effect forward of model 4 (PdeOperatorsMassEffect::solveState,
src/pde/PdeOperatorsMassEffect.cpp:421-743) against what the mathematics fixes: the closed-form
logistic growth when nothing moves and nothing diffuses, the CFL stop, and the symmetry of a problem
invariant under the exchange of two axes.  Its pieces (advection, elasticity, force, diffusion,
reaction) are pinned by their own tests."""
import numpy as np
import pytest

from me_cases import mass_effect_case

N = 32


def _oracle(oracle_mod, fields, k, rho, dt, nt, r_gm_wm=0.0):
    c, wm, gm, vt, csf, bg = fields
    o = oracle_mod.Oracle(N)
    o.set_tissue(wm, gm, csf, vt, bg)
    o.set_k(k)
    o.set_rho(rho, r_gm_wm)
    o.set_time(dt, nt)
    return o


def test_logistic_growth_without_force_or_diffusion(oracle_mod):
    """ff = 0 and k = 0: u = v = 0, the tissue moves not at all, and each of the nt + 1 steps is one
    logistic step with rho(x) = rho max(0, 1 - (bg + (1 - r) gm + vt + csf)) of the filtered
    tissue, so c(T) is the logistic solution over (nt + 1) dt."""
    fields = mass_effect_case(N)
    c0, wm, gm, vt, csf, bg = fields
    dt, nt, rho, r = 0.05, 3, 6.0, 0.3
    o = _oracle(oracle_mod, fields, 0.0, rho, dt, nt, r)
    out = o.me_solve_state(c0, wm, gm, vt, csf, bg, ff=0.0)
    assert out["info"]["steps"] == nt + 1 and not out["info"]["cfl_stop"]
    assert np.abs(out["u"]).max() == 0 and np.abs(out["v"]).max() == 0
    wmf, gmf = wm * (1 - c0), gm * (1 - c0)
    assert np.allclose(out["wm"], wmf, rtol=0, atol=1e-15) and np.allclose(out["gm"], gmf, rtol=0, atol=1e-15)
    assert np.array_equal(out["vt"], vt) and np.array_equal(out["csf"], csf)
    rf = rho * np.maximum(0.0, 1.0 - (bg + (1 - r) * gmf + vt + csf))
    T = (nt + 1) * dt
    expected = c0 * np.exp(rf * T) / (1.0 - c0 + c0 * np.exp(rf * T))
    assert np.max(np.abs(out["c"] - expected)) < 1e-12


def test_cfl_stop_after_the_first_velocity(oracle_mod):
    """Zero initial velocity passes the first CFL test; the velocity of the first step then stops
    the loop at any limit below the CFL that step produced."""
    fields = mass_effect_case(N)
    c0, wm, gm, vt, csf, bg = fields
    o = _oracle(oracle_mod, fields, 0.01, 8.0, 0.04, 3)
    full = o.me_solve_state(c0, wm, gm, vt, csf, bg, ff=1.2e5)
    assert full["info"]["steps"] == 4 and full["info"]["cfl_max"] > 0
    o = _oracle(oracle_mod, fields, 0.01, 8.0, 0.04, 3)
    lim = 0.5 * full["info"]["cfl_max"]
    cut = o.me_solve_state(c0, wm, gm, vt, csf, bg, ff=1.2e5, max_cfl=lim)
    assert cut["info"]["cfl_stop"] and 1 <= cut["info"]["steps"] < 4
    assert cut["info"]["cfl_max"] >= lim


def test_axis_exchange_symmetry(oracle_mod):
    """Tissue and tumor invariant under x <-> y: c(T) and the tissue stay invariant and the
    displacement's x and y components exchange (u_x(x, y, z) = u_y(y, x, z))."""
    fields = mass_effect_case(N, symmetric=True)
    c0, wm, gm, vt, csf, bg = fields
    assert np.array_equal(c0, c0.transpose(1, 0, 2))
    o = _oracle(oracle_mod, fields, 0.01, 8.0, 0.04, 2)
    out = o.me_solve_state(c0, wm, gm, vt, csf, bg, ff=1.2e5)
    sw = lambda a: a.transpose(1, 0, 2)  # noqa: E731
    scale = np.abs(out["u"]).max()
    assert scale > 0
    for k in ("c", "wm", "gm", "vt", "csf"):
        assert np.max(np.abs(out[k] - sw(out[k]))) < 1e-9 * max(1.0, np.abs(out[k]).max()), k
    assert np.max(np.abs(out["u"][0] - sw(out["u"][1]))) < 1e-8 * scale
    assert np.max(np.abs(out["u"][2] - sw(out["u"][2]))) < 1e-8 * scale
    assert np.max(np.abs(out["v"][0] - sw(out["v"][1]))) < 1e-8 * np.abs(out["v"]).max()


def _full_tissue(n):
    """wm + gm = 1 everywhere (no background), so every growth and death transfer of the
    multispecies sources stays inside p + i + nec + gm + wm."""
    x = np.arange(n) * 2 * np.pi / n
    wm = 0.5 + 0.3 * np.sin(x)[:, None, None] * np.sin(x)[None, :, None] * np.ones(n)[None, None, :]
    gm = 1.0 - wm
    z = np.zeros_like(wm)
    c0 = mass_effect_case(n)[0]
    return c0, wm, gm, z, z.copy(), z.copy()


def test_multispecies_conserves_cells_without_motion_or_diffusion(oracle_mod):
    """Model 5 with no force and k = 0: nothing moves, the smoothing of i keeps its total, D i = 0,
    and with wm + gm = 1 the sources move cells between p, i, nec, gm and wm only, so their grid
    total stays at its value after filterTumor; i starts at 0 and oxygen at 1."""
    c0, wm, gm, vt, csf, bg = _full_tissue(N)
    o = _oracle(oracle_mod, (c0, wm, gm, vt, csf, bg), 0.0, 8.0, 0.02, 3)
    out = o.ms_solve_state(c0, wm, gm, vt, csf, bg, ff=0.0)
    assert out["info"]["steps"] == 4 and out["info"]["el_iterations"] == 0
    assert np.abs(out["u"]).max() == 0 and np.abs(out["v"]).max() == 0
    start = np.sum(c0 + (wm + gm) * (1 - c0))
    end = sum(np.sum(out[k]) for k in ("p", "i", "nec", "gm", "wm"))
    assert abs(end - start) < 1e-10 * start
    assert out["i"].max() > 0 and out["nec"].max() > 0 and out["ox"].max() <= 1.0
    # the last source step comes after the last clipping: round-off below zero only
    assert min(out["p"].min(), out["gm"].min(), out["wm"].min()) > -1e-20


def test_multispecies_axis_exchange_symmetry(oracle_mod):
    """Model 5 on a problem invariant under x <-> y: every species and tissue stays invariant and
    u_x, u_y exchange."""
    fields = mass_effect_case(N, symmetric=True)
    c0, wm, gm, vt, csf, bg = fields
    o = _oracle(oracle_mod, fields, 0.01, 8.0, 0.04, 2)
    out = o.ms_solve_state(c0, wm, gm, vt, csf, bg, ff=1.2e5)
    sw = lambda a: a.transpose(1, 0, 2)  # noqa: E731
    assert out["info"]["steps"] == 3 and out["info"]["el_iterations"] > 0
    for k in ("p", "i", "nec", "ox", "wm", "gm", "vt", "csf"):
        assert np.max(np.abs(out[k] - sw(out[k]))) < 1e-9 * max(1.0, np.abs(out[k]).max()), k
    scale = np.abs(out["u"]).max()
    assert np.max(np.abs(out["u"][0] - sw(out["u"][1]))) < 1e-8 * scale


def test_c_abi_rejects_missing_handles_without_a_gpu():
    """The mass-effect entry points answer a null handle or config with GLIA_RD_ERR_ARG before any
    device call, and name a null handle in the error text."""
    import ctypes

    from glia_amd import _lib
    L = _lib.load()
    out = ctypes.c_void_p()
    assert L.glia_me_create(None, ctypes.byref(out)) == -2 and not out.value
    assert L.glia_me_destroy(None) == -2
    assert L.glia_me_last_error(None) == b"null handle"
    assert L.glia_me_rd(None) is None and L.glia_me_el(None) is None and L.glia_me_sl(None) is None
    info = _lib.MeInfo()
    assert L.glia_me_solve_state(None, None, *([None] * 9), ctypes.byref(info)) == -2
    assert L.glia_me_solve_state_multispecies(None, None, None, *([None] * 12), ctypes.byref(info)) == -2
    p = _lib.MeParams()
    L.glia_me_params_init(ctypes.byref(p))
    # max_cfl 0 selects each model's reference limit (14 for model 4, 8 for model 5)
    assert p.forcing_factor == 1e5 and p.use_tanh == 1 and p.max_cfl == 0.0 and p.el.E_tumor == 8000.0


def test_c_abi_partly_created_handle_without_a_gpu():
    """A mass-effect handle whose engine could not be created (order 3 is rejected before any
    device call) is still handed out with its message; both solve entry points then answer
    GLIA_RD_ERR_STATE instead of touching the missing engine."""
    import ctypes

    from glia_amd import _lib
    L = _lib.load()
    cfg = _lib.Config()
    L.glia_rd_config_init(ctypes.byref(cfg))
    cfg.order = 3
    h = ctypes.c_void_p()
    assert L.glia_me_create(ctypes.byref(cfg), ctypes.byref(h)) == -4 and h.value
    assert b"engine" in L.glia_me_last_error(h)
    p = _lib.MeParams()
    L.glia_me_params_init(ctypes.byref(p))
    mp = _lib.MsParams()
    L.glia_ms_params_init(ctypes.byref(mp))
    dummy = ctypes.c_void_p(16)  # never dereferenced: the handle check comes first
    comp = (ctypes.c_void_p * 3)(dummy, dummy, dummy)
    info = _lib.MeInfo()
    rc = L.glia_me_solve_state(h, ctypes.byref(p), *([dummy] * 7), comp, comp, ctypes.byref(info))
    assert rc == -3 and b"not fully created" in L.glia_me_last_error(h)
    rc = L.glia_me_solve_state_multispecies(h, ctypes.byref(p), ctypes.byref(mp), *([dummy] * 10), comp, comp,
                                            ctypes.byref(info))
    assert rc == -3
    assert L.glia_me_destroy(h) == 0
