"""polatory_amd: MI355X-native RBF fit/evaluate hot path behind polatory's API.

Compute runs in hand-written gfx950 HIP kernels inside libpolatory_amd.so,
reached through the C ABI declared in include/polatory_amd.h; this package is
the host-side mirror used by tests and benchmarks.
"""
from . import _capi
from .fmm import (FmmConfig, FmmEvaluator, FmmGradientEvaluator, FmmGradientTransposeEvaluator,
                  FmmHessianEvaluator, FmmHessianSymmetricEvaluator, FmmSymmetricEvaluator, direct_evaluate,
                  direct_evaluate_gradient,
                  make_fmm_evaluator, make_fmm_gradient_evaluator, make_fmm_gradient_transpose_evaluator,
                  make_fmm_hessian_evaluator, make_fmm_hessian_symmetric_evaluator, make_fmm_symmetric_evaluator)
from .interpolant import (DomainDivider, FitOptions, Interpolant, Model, Operator, RasPreconditioner, RbfFieldFunction, cross_validate,
                          distance_filter, fgmres_dense, POLY_DEGREE_MIN)
from .rbf import Biharmonic3D, CovExponential, Rbf, Triharmonic3D, make_rbf
from .serve import Server


def device_count():
    import ctypes
    n = ctypes.c_int()
    _capi.call("pam_device_count", ctypes.byref(n))
    return n.value


def device_cache_release():
    """Frees the device buffers the library keeps for reuse on the current GPU."""
    _capi.call("pam_device_cache_release")


def device_cache_bytes():
    """Bytes of released device buffers held for reuse on the current GPU."""
    import ctypes
    n = ctypes.c_size_t()
    _capi.call("pam_device_cache_bytes", ctypes.byref(n))
    return n.value


__all__ = ["Biharmonic3D", "Triharmonic3D", "CovExponential", "Rbf", "make_rbf",
           "FmmConfig", "FmmEvaluator", "FmmSymmetricEvaluator", "FmmGradientTransposeEvaluator", "FmmGradientEvaluator",
           "make_fmm_gradient_evaluator", "FmmHessianEvaluator", "make_fmm_hessian_evaluator",
           "make_fmm_evaluator", "make_fmm_symmetric_evaluator", "make_fmm_hessian_symmetric_evaluator", "make_fmm_gradient_transpose_evaluator", "direct_evaluate", "direct_evaluate_gradient",
           "device_count", "device_cache_release", "device_cache_bytes", "Model",
           "Interpolant", "FitOptions", "DomainDivider", "Operator", "RasPreconditioner", "fgmres_dense", "POLY_DEGREE_MIN", "distance_filter",
           "cross_validate", "RbfFieldFunction",
           "Server"]
