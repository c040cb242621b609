"""ctypes declarations of libglia_rd.so (the C ABI in include/glia_rd.h).

The library is built in-tree (glia_amd/libglia_rd.so) by glia_amd/build.py.  There is no fallback:
importing this module without the library raises ImportError.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GLIA_RD_LIBRARY may point at another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GLIA_RD_LIBRARY") or os.path.join(_HERE, "libglia_rd.so")

HOST, DEVICE = 0, 1
OK, ERR_ARG, ERR_STATE, ERR_UNSUPPORTED, ERR_POOL, ERR_NOHISTORY = 0, -2, -3, -4, -5, -6
FIELD = dict(k=0, rho=1, wm=2, gm=3, csf=4, vt=5, filter=6, obs=7, d1=8, c0=9, ct=10, alpha0=11, work=12)
MODE_P_KAPPA, MODE_KAPPA_RHO, MODE_ALL = 0, 1, 2


class Config(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("nmembers", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("nt", ctypes.c_int),
        ("order", ctypes.c_int),
        ("pcg_rtol", ctypes.c_double),
        ("pcg_maxit", ctypes.c_int),
        ("checkpoint_slots", ctypes.c_int),
        ("adjoint_store", ctypes.c_int),
        ("phi_store", ctypes.c_int),
        ("k_gm_wm", ctypes.c_double),
        ("k_glm_wm", ctypes.c_double),
        ("r_gm_wm", ctypes.c_double),
        ("r_glm_wm", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("k_scale", ctypes.c_double),
        ("rho_scale", ctypes.c_double),
        ("smoothing_factor", ctypes.c_double),
        ("stream", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("pcg_iterations", ctypes.c_int),
        ("cn_solves", ctypes.c_int),
        ("worst_reason", ctypes.c_int),
        ("last_iterations", ctypes.c_int),
        ("last_reason", ctypes.c_int),
        ("kbar", ctypes.c_double),
        ("prec_coef", ctypes.c_double),
        ("filter_sum", ctypes.c_double),
        ("misfit", ctypes.c_double),
        ("c0_norm2", ctypes.c_double),
    ]


# every entry point declared in include/glia_rd.h, with its ctypes signature
_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "glia_rd_config_init": (None, [ctypes.POINTER(Config)]),
    "glia_rd_create": (_i, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "glia_rd_destroy": (_i, [_vp]),
    "glia_rd_last_error": (ctypes.c_char_p, [_vp]),
    "glia_rd_sync": (_i, [_vp]),
    "glia_rd_num_members": (_i, [_vp]),
    "glia_rd_set_tissue": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i]),
    "glia_rd_set_tissue_sinusoidal": (_i, [_vp, _i]),
    "glia_rd_set_coefficients": (_i, [_vp, _i, _d, _d]),
    "glia_rd_set_kappa": (_i, [_vp, _i, _d]),
    "glia_rd_set_rho": (_i, [_vp, _i, _d]),
    "glia_rd_set_k_sinusoidal": (_i, [_vp, _i, _d]),
    "glia_rd_prec_factor": (_i, [_vp]),
    "glia_rd_set_ctx_dt": (_i, [_vp, _d]),
    "glia_rd_get_ctx_dt": (_i, [_vp, _pd]),
    "glia_rd_set_beta": (_i, [_vp, _d]),
    "glia_rd_get_grid": (_i, [_vp, _vp]),
    "glia_rd_get_beta": (_i, [_vp, _pd]),
    "glia_rd_set_ratios": (_i, [_vp, _d, _d, _d, _d]),
    "glia_rd_set_basis": (_i, [_vp, _i, _i, _vp, _d]),
    "glia_rd_data_support": (_i, [_vp, _i, _i, _d, _d, _d, _i, _vp, ctypes.POINTER(_i)]),
    "glia_rd_basis_size": (_i, [_vp, _i]),
    "glia_rd_set_data": (_i, [_vp, _i, _vp, _vp, _d, _i]),
    "glia_rd_set_field": (_i, [_vp, _i, _i, _vp, _i]),
    "glia_rd_get_field": (_i, [_vp, _i, _i, _vp, _i]),
    "glia_rd_phi_apply": (_i, [_vp, _vp]),
    "glia_rd_phi_apply_transpose": (_i, [_vp, _i, _vp]),
    "glia_rd_diffusion_solve": (_i, [_vp, _i, _d]),
    "glia_rd_apply_d": (_i, [_vp, _i, _i]),
    "glia_rd_fft_r2c": (_i, [_vp, _i, _vp, _vp]),
    "glia_rd_fft_c2r": (_i, [_vp, _i, _vp, _vp]),
    "glia_rd_solve_state": (_i, [_vp, _i]),
    "glia_rd_solve_adjoint": (_i, [_vp, _i]),
    "glia_rd_objective": (_i, [_vp, _i, _vp, _vp]),
    "glia_rd_gradient": (_i, [_vp, _i, _vp, _vp]),
    "glia_rd_objective_and_gradient": (_i, [_vp, _i, _vp, _vp, _vp]),
    "glia_rd_objective_and_gradient_group": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "glia_rd_objective_and_gradient_lockstep": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "glia_rd_hessian": (_i, [_vp, _i, _vp, _vp]),
    "glia_rd_stats": (_i, [_vp, _i, ctypes.POINTER(Stats)]),
    "glia_rd_reset_stats": (_i, [_vp]),
    "glia_rd_smooth_field": (_i, [_vp, _i, _i, _d]),
    "glia_rd_set_time": (_i, [_vp, _d, _i]),
    "glia_rd_autotune": (_i, [_vp, ctypes.c_char_p, _vp]),
    "glia_rd_step_info": (_i, [_vp, _vp, _vp, _vp]),
    "glia_rd_keep_adjoint_history": (_i, [_vp, _i]),
    "glia_rd_get_history": (_i, [_vp, _i, _i, _i, _vp, _i]),
    "glia_rd_set_secondary": (_i, [_vp, _i, _d]),
    "glia_rd_grad_diffusion": (_i, [_vp, _vp]),
    "glia_rd_grad_reaction": (_i, [_vp, _vp]),
    "glia_rd_check_gradient": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "glia_rd_smooth_noalloc": (_i, [ctypes.POINTER(ctypes.c_int), _vp, _vp, _d, _vp, ctypes.c_char_p, _i]),
    "glia_rd_refresh_tissue": (_i, [_vp, _i]),
    "glia_rd_use_graphs": (_i, [_vp, _i]),
    "glia_rd_profile": (_i, [_vp, _i]),
    "glia_rd_profile_get": (_i, [_vp, _i, _pd, ctypes.POINTER(ctypes.c_longlong)]),
}

_i64 = ctypes.c_int64
_cp = ctypes.c_char_p
# include/glia_io.h (host-side input); glia_rd_file_config is read through FileConfig
IO_SIGNATURES = {
    "glia_nc_inquire": (_i, [_cp, _cp, ctypes.POINTER(_i), ctypes.POINTER(_i64), ctypes.POINTER(_i)]),
    "glia_nc_read_f32": (_i, [_cp, _cp, _vp, _i64]),
    "glia_nc_read_f64": (_i, [_cp, _cp, _vp, _i64]),
    "glia_nc_write3_f32": (_i, [_cp, _cp, _vp, _vp, _i]),
    "glia_split_segmentation": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "glia_rd_parse_config": (_i, [_cp, _vp]),
    "glia_edema_obs_mask": (_i, [_vp, _vp, _i64, _d, _vp]),
    "glia_phi_mesh": (_i, [_i, _vp, _d, _d, _vp]),
    "glia_read_phi_mesh": (_i, [_cp, _i, _vp, _vp, ctypes.POINTER(_i), _pd, _pd]),
    "glia_read_concomp": (_i, [_cp, _i, _vp, _vp, ctypes.POINTER(_i)]),
}


class OptSettings(ctypes.Structure):
    _fields_ = [("max_it", ctypes.c_int), ("min_it", ctypes.c_int), ("gatol", ctypes.c_double),
                ("gttol", ctypes.c_double), ("lbfgs_vectors", ctypes.c_int), ("ls_max_funcs", ctypes.c_int),
                ("ls_sigma", ctypes.c_double), ("ls_minstep", ctypes.c_double), ("line_search", ctypes.c_int),
                ("ls_gtol", ctypes.c_double), ("ls_xtol", ctypes.c_double), ("gnorm_ref", ctypes.c_double),
                ("newton_solver", ctypes.c_int), ("krylov_maxit", ctypes.c_int), ("forcing", ctypes.c_int)]


class OptResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("fevals", ctypes.c_int), ("reason", ctypes.c_int),
                ("f", ctypes.c_double), ("gnorm", ctypes.c_double), ("gnorm0", ctypes.c_double)]


# int (*)(void* ctx, int nprob, const double* x, double* f, double* g)
FG_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


class SparseTilSettings(ctypes.Structure):
    _fields_ = [("sparsity", ctypes.c_int), ("max_it", ctypes.c_int), ("ftol", ctypes.c_double),
                ("prune_components", ctypes.c_int), ("thresh_component_weight", ctypes.c_double),
                ("diffusivity_inversion", ctypes.c_int), ("reaction_inversion", ctypes.c_int),
                ("beta", ctypes.c_double), ("k_lb", ctypes.c_double), ("k_ub", ctypes.c_double),
                ("rho_lb", ctypes.c_double), ("rho_ub", ctypes.c_double), ("til", OptSettings), ("rd", OptSettings)]


class SparseTilResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("stop", ctypes.c_int), ("nnz", ctypes.c_int), ("J", ctypes.c_double),
                ("til", OptResult), ("rd", OptResult)]


# int (*)(void* ctx, int mode, double beta, const double* x, double* f, double* g)
EVAL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_double))
# int (*)(void* ctx, const double* p, double* cmax)
C0_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double))
# int (*)(void* ctx, int nprob, const double* v, double* y)
HV_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double))
# include/glia_opt.h
OPT_SIGNATURES = {
    "glia_opt_settings_init": (None, [ctypes.POINTER(OptSettings)]),
    "glia_blmvm": (_i, [_i, _i, _vp, _vp, _vp, FG_FN, _vp, ctypes.POINTER(OptSettings), _vp]),
    "glia_rd_invert": (_i, [_vp, _i, _vp, _vp, _vp, ctypes.POINTER(OptSettings), _vp]),
    "glia_gauss_newton": (_i, [_i, _i, _vp, _vp, _vp, FG_FN, HV_FN, _vp, ctypes.POINTER(OptSettings), _vp]),
    "glia_opt_parse_config": (_i, [ctypes.c_char_p, ctypes.POINTER(OptSettings), ctypes.POINTER(SparseTilSettings)]),
    "glia_hard_threshold": (_i, [_vp, _i, _i, _vp]),
    "glia_hard_threshold_components": (_i, [_vp, _i, _i, _vp, _vp, _i, _d, _i, _vp]),
    "glia_sparse_til_settings_init": (None, [ctypes.POINTER(SparseTilSettings)]),
    "glia_sparse_til": (_i, [_i, _i, _vp, _vp, _vp, _vp, _i, EVAL_FN, C0_FN, _vp,
                             ctypes.POINTER(SparseTilSettings), _vp]),
    "glia_rd_sparse_til": (_i, [_vp, _vp, _vp, _vp, _vp, _i, ctypes.POINTER(SparseTilSettings), _vp]),
}


class FileConfig(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("model", ctypes.c_int), ("nt", ctypes.c_int), ("dt", ctypes.c_double),
        ("k", ctypes.c_double), ("rho", ctypes.c_double), ("k_gm_wm", ctypes.c_double), ("r_gm_wm", ctypes.c_double),
        ("beta", ctypes.c_double), ("smoothing_factor", ctypes.c_double), ("obs_threshold_1", ctypes.c_double),
        ("np", ctypes.c_int), ("sigma_factor", ctypes.c_double), ("sigma_spacing", ctypes.c_double),
        ("phi_store", ctypes.c_int), ("adjoint_store", ctypes.c_int), ("atlas_labels", ctypes.c_int * 8),
        ("a_seg_path", ctypes.c_char * 512), ("a_wm_path", ctypes.c_char * 512), ("a_gm_path", ctypes.c_char * 512),
        ("a_csf_path", ctypes.c_char * 512), ("a_vt_path", ctypes.c_char * 512), ("d1_path", ctypes.c_char * 512),
        ("obs_mask_path", ctypes.c_char * 512),
        ("obs_lambda", ctypes.c_double), ("smoothing_factor_data", ctypes.c_double), ("obs_threshold_rel", ctypes.c_int),
        ("patient_labels", ctypes.c_int * 8), ("p_seg_path", ctypes.c_char * 512),
        ("threshold_data_driven", ctypes.c_double), ("gaussian_volume_fraction", ctypes.c_double),
        ("sigma_data_driven", ctypes.c_double), ("support_data_path", ctypes.c_char * 512),
        ("gaussian_cm_path", ctypes.c_char * 512), ("data_comp_data_path", ctypes.c_char * 512),
        ("d0_path", ctypes.c_char * 512), ("smoothing_factor_data_t0", ctypes.c_double),
        ("obs_threshold_0", ctypes.c_double),
    ]


# include/glia_adv.h (semi-Lagrangian advection)
ADV_SIGNATURES = {
    "glia_sl_create": (_i, [_vp, _vp, ctypes.POINTER(_vp)]),
    "glia_sl_destroy": (_i, [_vp]),
    "glia_sl_last_error": (ctypes.c_char_p, [_vp]),
    "glia_sl_sync": (_i, [_vp]),
    "glia_sl_set_velocity": (_i, [_vp, _vp, _vp, _vp, _i]),
    "glia_sl_invalidate": (_i, [_vp]),
    "glia_sl_compute_trajectories": (_i, [_vp, _d]),
    "glia_sl_get_offsets": (_i, [_vp, _vp, _vp, _vp, _i]),
    "glia_sl_interpolate": (_i, [_vp, _vp, _vp, _i]),
    "glia_sl_solve": (_i, [_vp, _vp, _d, _i, _i]),
    "glia_sl_advect": (_i, [_vp, _i, _vp, _vp, _vp, _d]),
}


# include/glia_el.h (linear elasticity)
class ElParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("E_healthy", "nu_healthy", "E_tumor", "nu_tumor", "E_bg", "nu_bg",
                                                 "E_csf", "nu_csf", "screen_low", "screen_high")]


EL_SIGNATURES = {
    "glia_el_params_init": (None, [ctypes.POINTER(ElParams)]),
    "glia_el_create": (_i, [_vp, _vp, ctypes.POINTER(_vp)]),
    "glia_el_destroy": (_i, [_vp]),
    "glia_el_last_error": (ctypes.c_char_p, [_vp]),
    "glia_el_set_material": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(ElParams), _i]),
    "glia_el_get_material": (_i, [_vp, _vp, _vp, _vp, _vp, _i]),
    "glia_el_apply": (_i, [_vp, _vp, _vp, _i]),
    "glia_el_precondition": (_i, [_vp, _vp, _vp, _i]),
    "glia_el_solve": (_i, [_vp, _vp, _vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_d)]),
    "glia_el_force": (_i, [_vp, _vp, _d, _i, _vp, _i]),
}


# include/glia_ms.h (multispecies pointwise terms)
class MsParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("death_rate", "ox_source", "ox_consumption", "alpha_0", "beta_0",
                                                 "ox_inv", "ox_hypoxia", "invasive_threshold")]


MS_SIGNATURES = {
    "glia_ms_params_init": (None, [ctypes.POINTER(MsParams)]),
    "glia_ms_sources": (_i, [ctypes.c_int64] + [_vp] * 8 + [_d, ctypes.POINTER(MsParams), _vp]),
    "glia_ms_update_coefficients": (_i, [ctypes.c_int64] + [_vp] * 6 + [_d, _d, _d, _d, _vp]),
    "glia_ms_segmentation": (_i, [ctypes.c_int64] + [_vp] * 7 + [_vp]),
    "glia_ms_clip": (_i, [ctypes.c_int64, _vp, _vp]),
    "glia_ms_edema_mask": (_i, [ctypes.c_int64, _vp, _d, _vp, _vp]),
}


# include/glia_me.h (mass-effect forward of model 4)
class MeParams(ctypes.Structure):
    _fields_ = [("forcing_factor", ctypes.c_double), ("use_tanh", ctypes.c_int), ("max_cfl", ctypes.c_double),
                ("el", ElParams)]


class MeInfo(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("cfl_stop", ctypes.c_int), ("cfl_max", ctypes.c_double),
                ("el_iterations", ctypes.c_int), ("diff_iterations", ctypes.c_int), ("volume_stop", ctypes.c_int)]


ME_SIGNATURES = {
    "glia_me_params_init": (None, [ctypes.POINTER(MeParams)]),
    "glia_me_create": (_i, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "glia_me_destroy": (_i, [_vp]),
    "glia_me_last_error": (ctypes.c_char_p, [_vp]),
    "glia_me_rd": (_vp, [_vp]),
    "glia_me_el": (_vp, [_vp]),
    "glia_me_sl": (_vp, [_vp]),
    "glia_me_solve_state": (_i, [_vp, ctypes.POINTER(MeParams)] + [_vp] * 9 + [ctypes.POINTER(MeInfo)]),
    "glia_me_solve_state_features": (_i, [_vp, ctypes.POINTER(MeParams)] + [_vp] * 9 + [ctypes.POINTER(MeInfo), _vp, _i,
                                                                                        ctypes.POINTER(_i)]),
    "glia_me_feature_name": (ctypes.c_char_p, [_i]),
    "glia_me_solve_state_multispecies": (_i, [_vp, ctypes.POINTER(MeParams), ctypes.POINTER(MsParams)] + [_vp] * 12
                                         + [ctypes.POINTER(MeInfo)]),
    "glia_me_multispecies_outputs": (_i, [_vp, _vp, _vp, _i]),
}


KERNELS = ["xi_pupd", "yi_grad", "z_diff", "yf_div", "xf_matvec", "pcg_update"]
# algorithmic bytes of one launch per member, in units of one field (4 n0 n1 n2 bytes): every
# input read once and every output written once
KERNEL_FIELDS = {"xi_pupd": 5, "yi_grad": 5, "z_diff": 7, "yf_div": 5, "xf_matvec": 4, "pcg_update": 6}

_lib = None


def load():
    """Load libglia_rd.so (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch ships its own HIP runtime under the soname of ROCm's (libamdhip64.so.7), and the first
    # one loaded serves the whole process; loading torch first gives torch and this library one
    # runtime, so device tensors can be passed to the library in any call order
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libglia_rd.so not found at %s: build it with `python -m glia_amd.build` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback" % LIB_PATH
        )
    lib = ctypes.CDLL(LIB_PATH)
    tables = [SIGNATURES, IO_SIGNATURES, OPT_SIGNATURES, ADV_SIGNATURES, EL_SIGNATURES, MS_SIGNATURES,
              ME_SIGNATURES]
    for name, (res, args) in [kv for t in tables for kv in t.items()]:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
