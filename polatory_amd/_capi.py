"""ctypes binding of the C ABI in include/polatory_amd.h.

The shared library is built in-tree (``polatory_amd/libpolatory_amd.so``) by
``polatory_amd/build.py``. Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpolatory_amd.so")

PAM_RBF_BH3, PAM_RBF_TH3, PAM_RBF_EXP = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
MODE_AUTO, MODE_DIRECT, MODE_FMM = 0, 1, 2


class PamRbf(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("params", ctypes.c_double * 2),
                ("aniso", ctypes.c_double * 9)]


class PamFmmConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("order", ctypes.c_int),
                ("leaf_level", ctypes.c_int), ("mode", ctypes.c_int),
                ("svd_tolerance", ctypes.c_double)]


class PamFmmStats(ctypes.Structure):
    _fields_ = [("ms_tree", ctypes.c_double), ("ms_upward", ctypes.c_double),
                ("ms_m2l", ctypes.c_double), ("ms_l2l", ctypes.c_double),
                ("ms_l2p", ctypes.c_double), ("ms_p2p", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("precision", ctypes.c_int),
                ("order", ctypes.c_int), ("leaf_level", ctypes.c_int), ("direct", ctypes.c_int),
                ("n_src_leaves", ctypes.c_int64), ("n_trg_leaves", ctypes.c_int64),
                ("m2l_rank", ctypes.c_int)]


class PamFitOptions(ctypes.Structure):
    _fields_ = [("op", PamFmmConfig), ("residual", PamFmmConfig), ("level", PamFmmConfig),
                ("use_level_config", ctypes.c_int), ("use_preconditioner", ctypes.c_int),
                ("refine_rtol", ctypes.c_double), ("verbose", ctypes.c_int),
                ("use_graphs", ctypes.c_int), ("keep_unconverged", ctypes.c_int),
                ("tune_solves", ctypes.c_int), ("ras_fp32_factors", ctypes.c_int),
                ("tune_factor", ctypes.c_int),
                ("factor_panel", ctypes.c_int),
                ("cpd2_auto", ctypes.c_int)]


class PamFitStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("restarts", ctypes.c_int), ("n_levels", ctypes.c_int),
                ("residual", ctypes.c_double), ("seconds", ctypes.c_double),
                ("setup_seconds", ctypes.c_double), ("precond_setup_seconds", ctypes.c_double),
                ("factor_bytes", ctypes.c_int64), ("graph_replays", ctypes.c_int),
                ("escalations", ctypes.c_int), ("converged", ctypes.c_int), ("rounds", ctypes.c_int),
                ("grad_residual", ctypes.c_double)]


class PamServerOptions(ctypes.Structure):
    _fields_ = [("n_slots", ctypes.c_int), ("hbm_budget", ctypes.c_int64),
                ("rel_accuracy", ctypes.c_double)]


class PamServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("failed", ctypes.c_int64), ("points", ctypes.c_int64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("mean_queue_ms", ctypes.c_double), ("points_per_s", ctypes.c_double),
                ("wall_s", ctypes.c_double), ("max_in_flight", ctypes.c_int),
                ("resident_bytes", ctypes.c_int64), ("budget", ctypes.c_int64),
                ("peak_admitted_bytes", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("batched_requests", ctypes.c_int64), ("max_batch", ctypes.c_int),
                ("fmm_batches", ctypes.c_int64), ("fmm_batched_requests", ctypes.c_int64),
                ("max_fmm_batch", ctypes.c_int)]


class PamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class InvalidArgument(PamError, ValueError):
    """std::invalid_argument in the reference API."""


class NotImplementedByLibrary(PamError, NotImplementedError):
    """A polatory feature outside this library (an rbf or gradient data in a file)."""


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int

# name -> argtypes (all return int status)
_SIGS = {
    "pam_last_error": [ctypes.c_char_p, ctypes.c_size_t],
    "pam_device_count": [ctypes.POINTER(ctypes.c_int)],
    "pam_device_cache_release": [],
    "pam_device_cache_bytes": [ctypes.POINTER(ctypes.c_size_t)],
    "pam_fmm_evaluator_create": [ctypes.POINTER(PamRbf), _P, _P, ctypes.POINTER(_P)],
    "pam_fmm_symmetric_evaluator_create": [ctypes.POINTER(PamRbf), _P, _P, ctypes.POINTER(_P)],
    "pam_fmm_evaluator_destroy": [_P],
    "pam_fmm_evaluator_set_accuracy": [_P, _D],
    "pam_fmm_evaluator_set_search_precision": [_P, _INT],
    "pam_fmm_evaluator_set_source_points": [_P, _P, _I64],
    "pam_fmm_evaluator_set_target_points": [_P, _P, _I64],
    "pam_fmm_evaluator_set_points": [_P, _P, _I64],
    "pam_fmm_evaluator_set_weights": [_P, _P, _I64],
    "pam_fmm_evaluator_evaluate": [_P, _P, _I64],
    "pam_fmm_evaluator_evaluate_gradient": [_P, _P, _I64],
    "pam_fmm_hessian_evaluator_create": [ctypes.POINTER(PamRbf), _P, _P, ctypes.POINTER(_P)],
    "pam_fmm_hessian_symmetric_evaluator_create": [ctypes.POINTER(PamRbf), _P, _P, ctypes.POINTER(_P)],
    "pam_fmm_gradient_evaluator_create": [ctypes.POINTER(PamRbf), _P, _P, ctypes.POINTER(_P)],
    "pam_fmm_gradient_transpose_evaluator_create": [ctypes.POINTER(PamRbf), _P, _P, ctypes.POINTER(_P)],
    "pam_fmm_evaluator_evaluate_gradient_device": [_P, _P, _I64],
    "pam_fmm_evaluator_set_source_points_device": [_P, _P, _I64],
    "pam_fmm_evaluator_set_target_points_device": [_P, _P, _I64],
    "pam_fmm_evaluator_set_points_device": [_P, _P, _I64],
    "pam_fmm_evaluator_set_weights_device": [_P, _P, _I64],
    "pam_fmm_evaluator_evaluate_device": [_P, _P, _I64],
    "pam_fmm_evaluator_synchronize": [_P],
    "pam_fmm_evaluator_set_config": [_P, ctypes.POINTER(PamFmmConfig)],
    "pam_fmm_evaluator_get_config": [_P, ctypes.POINTER(PamFmmConfig)],
    "pam_fmm_evaluator_set_overlap": [_P, _INT],
    "pam_fmm_evaluator_stats": [_P, ctypes.POINTER(PamFmmStats)],
    "pam_fmm_evaluator_count_work": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)],
    "pam_fmm_evaluator_tree_info": [_P, _INT, ctypes.POINTER(_INT), _P, ctypes.POINTER(_D),
                                    ctypes.POINTER(_D), ctypes.POINTER(_I64), ctypes.POINTER(_I64)],
    "pam_fmm_evaluator_export_tree": [_P, _INT, _P, _P, _P, _P],
    "pam_fmm_evaluator_export_level": [_P, _INT, _INT, ctypes.POINTER(_I64), _P],
    "pam_fmm_evaluator_export_p2p": [_P, _P],
    "pam_fmm_evaluator_export_m2l": [_P, _INT, _P],
    "pam_direct_evaluate": [ctypes.POINTER(PamRbf), _P, _I64, _P, _P, _I64, _P],
    "pam_direct_evaluate_gradient": [ctypes.POINTER(PamRbf), _P, _I64, _P, _P, _I64, _P],
    "pam_direct_evaluate_device_noalloc": [ctypes.POINTER(PamRbf), _P, _I64, _P, _P, _I64, _P, _P],
    "pam_rbf_evaluate": [ctypes.POINTER(PamRbf), _P, _I64, _P],
    "pam_rbf_evaluate_gradient": [ctypes.POINTER(PamRbf), _P, _I64, _P],
    "pam_rbf_evaluate_hessian": [ctypes.POINTER(PamRbf), _P, _I64, _P],
    "pam_model_create": [_P, _INT, _INT, _D, ctypes.POINTER(_P)],
    "pam_model_destroy": [_P],
    "pam_model_poly_degree": [_P, ctypes.POINTER(_INT)],
    "pam_model_poly_basis_size": [_P, ctypes.POINTER(_INT)],
    "pam_model_num_rbfs": [_P, ctypes.POINTER(_INT)],
    "pam_model_rbf": [_P, _INT, _P],
    "pam_model_nugget": [_P, ctypes.POINTER(ctypes.c_double)],
    "pam_model_save": [_P, ctypes.c_char_p],
    "pam_model_load": [ctypes.c_char_p, ctypes.POINTER(_P)],
    "pam_fit_options_default": [ctypes.POINTER(PamFitOptions)],
    "pam_interpolant_create": [_P, ctypes.POINTER(_P)],
    "pam_interpolant_destroy": [_P],
    "pam_interpolant_fit": [_P, _P, _I64, _P, _D, _INT, _D, ctypes.POINTER(PamFitOptions), _P],
    "pam_interpolant_evaluate": [_P, _P, _I64, _D, _P],
    "pam_interpolant_evaluate_device": [_P, _P, _I64, _D, _P],
    "pam_interpolant_set_eval_config": [_P, ctypes.POINTER(PamFmmConfig)],
    "pam_interpolant_set_evaluation_bbox": [_P, _P, _P, _D, _D],
    "pam_interpolant_evaluate_impl": [_P, _P, _I64, _P],
    "pam_interpolant_evaluate_impl_device": [_P, _P, _I64, _P],
    "pam_interpolant_num_centers": [_P, ctypes.POINTER(_I64)],
    "pam_interpolant_centers": [_P, _P],
    "pam_interpolant_weights": [_P, _P],
    "pam_interpolant_set_fitted": [_P, _P, _I64, _P],
    "pam_interpolant_evaluate_gradient": [_P, _P, _I64, ctypes.c_double, _P],
    "pam_interpolant_fit_incrementally": [_P, _P, _I64, _P, ctypes.c_double, _INT, ctypes.c_double, _P],
    "pam_interpolant_fit_inequality": [_P, _P, _I64, _P, _P, _P, ctypes.c_double, _INT, ctypes.c_double, _P,
                                       _P],
    "pam_cross_validate": [_P, _P, _I64, _P, _P, ctypes.c_double, _INT, ctypes.c_double, _P, _P],
    "pam_distance_filter": [_P, _I64, ctypes.c_double, _P, _I64, _P, ctypes.POINTER(_I64)],
    "pam_interpolant_save": [_P, ctypes.c_char_p],
    "pam_interpolant_load": [ctypes.c_char_p, ctypes.POINTER(_P)],
    "pam_interpolant_model": [_P, ctypes.POINTER(_P)],
    "pam_interpolant_fitted": [_P, ctypes.POINTER(_INT)],
    "pam_interpolant_fit_stats": [_P, ctypes.POINTER(PamFitStats)],
    "pam_interpolant_residual_history": [_P, _P, ctypes.POINTER(_I64)],
    "pam_fgmres_dense": [_P, _P, _I64, _P, _P, _INT, _P, _P, ctypes.POINTER(_INT)],
    "pam_ras_create": [_P, _P, _I64, ctypes.POINTER(_P)],
    "pam_ras_create_with_options": [_P, _P, _I64, ctypes.POINTER(PamFitOptions), ctypes.POINTER(_P)],
    "pam_ras_destroy": [_P],
    "pam_ras_num_levels": [_P, ctypes.POINTER(_INT)],
    "pam_ras_level_points": [_P, _INT, _P, ctypes.POINTER(_I64)],
    "pam_ras_num_domains": [_P, _INT, ctypes.POINTER(_INT)],
    "pam_ras_domain": [_P, _INT, _INT, _P, _P, ctypes.POINTER(_I64)],
    "pam_ras_apply": [_P, _P, _P],
    "pam_ras_apply_device": [_P, _P, _P],
    "pam_ras_size": [_P, ctypes.POINTER(_I64)],
    "pam_ras_synchronize": [_P],
    "pam_operator_create": [_P, _P, _I64, ctypes.POINTER(PamFmmConfig), ctypes.POINTER(_P)],
    "pam_interpolant_fit_hermite": [_P, _P, _I64, _P, _I64, _P, _D, _D, _INT, _D, _D, _P, _P],
    "pam_interpolant_fit_incrementally_hermite": [_P, _P, _I64, _P, _I64, _P, _D, _D, _INT, _D, _D, _P],
    "pam_interpolant_evaluate_hermite": [_P, _P, _I64, _P, _I64, _P],
    "pam_interpolant_num_grad_centers": [_P, ctypes.POINTER(_I64)],
    "pam_interpolant_grad_centers": [_P, _P],
    "pam_domain_divider_create": [_P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _I64, ctypes.POINTER(_P)],
    "pam_domain_divider_destroy": [_P],
    "pam_domain_divider_num_domains": [_P, ctypes.POINTER(_I64)],
    "pam_domain_divider_domain": [_P, _I64, _P, _P, ctypes.POINTER(_I64), _P, _P, ctypes.POINTER(_I64)],
    "pam_domain_divider_choose_coarse_points": [_P, _I64, _P, ctypes.POINTER(_I64), _P, ctypes.POINTER(_I64)],
    "pam_ras_create_hermite": [_P, _P, _I64, _P, _I64, _P, ctypes.POINTER(_P)],
    "pam_operator_create_hermite": [_P, _P, _I64, _P, _I64, _P, ctypes.POINTER(_P)],
    "pam_operator_destroy": [_P],
    "pam_operator_size": [_P, ctypes.POINTER(_I64)],
    "pam_operator_apply_device": [_P, _P, _P],
    "pam_operator_apply": [_P, _P, _P],
    "pam_operator_synchronize": [_P],
    "pam_server_options_default": [ctypes.POINTER(PamServerOptions)],
    "pam_server_create": [ctypes.POINTER(PamServerOptions), ctypes.POINTER(_P)],
    "pam_server_destroy": [_P],
    "pam_server_add_model": [_P, _P, _P, _P, ctypes.POINTER(PamFmmConfig), ctypes.POINTER(_INT)],
    "pam_server_model_config": [_P, _INT, ctypes.POINTER(PamFmmConfig), ctypes.POINTER(_I64)],
    "pam_server_submit": [_P, _INT, _P, _I64, _P, ctypes.POINTER(_I64)],
    "pam_server_wait": [_P, _I64],
    "pam_server_test": [_P, _I64, ctypes.POINTER(_INT)],
    "pam_server_drain": [_P],
    "pam_server_request_times": [_P, _I64, ctypes.POINTER(_D), ctypes.POINTER(_D)],
    "pam_server_get_stats": [_P, ctypes.POINTER(PamServerStats)],
    "pam_server_reset_stats": [_P],
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python polatory_amd/build.py` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGS)


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().pam_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code):
    if code != 0:
        msg = last_error()
        if code == 1:
            raise InvalidArgument(code, msg)
        if code == 5:
            raise NotImplementedByLibrary(code, msg)
        raise PamError(code, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))
