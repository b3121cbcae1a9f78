"""Host-side mirror of polatory's FMM seam over the C ABI.

``make_fmm_evaluator`` / ``make_fmm_symmetric_evaluator`` follow
include/polatory/fmm/fmm_evaluator.hpp:92-93 and fmm_symmetric_evaluator.hpp:80-82:
set points, set weights, ``evaluate()`` returns the FP64 values in target order.
"""
import ctypes

import numpy as np

from . import _capi
from .rbf import Rbf


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _points(p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("points must be an (n, 3) array")
    return p


def _bbox(bbox):
    if bbox is None:
        return None, None
    lo = np.ascontiguousarray(bbox[0], dtype=np.float64).reshape(3)
    hi = np.ascontiguousarray(bbox[1], dtype=np.float64).reshape(3)
    return lo, hi


class FmmConfig:
    def __init__(self, precision="mixed", order=6, leaf_level=-1, mode="auto", svd_tol=-1.0):
        self.precision = precision
        self.order = order
        self.leaf_level = leaf_level
        self.mode = mode
        self.svd_tol = svd_tol  # <0: default, 0: dense M2L, >0: compression cutoff

    def to_c(self):
        c = _capi.PamFmmConfig()
        c.precision = {"fp64": 0, "mixed": 1, "mixed_comp": 2}[self.precision]
        c.order = self.order
        c.leaf_level = self.leaf_level
        c.mode = {"auto": 0, "direct": 1, "fmm": 2}[self.mode]
        c.svd_tolerance = self.svd_tol
        return c

    @staticmethod
    def from_c(c):
        return FmmConfig({0: "fp64", 1: "mixed", 2: "mixed_comp"}[c.precision], c.order, c.leaf_level,
                         {0: "auto", 1: "direct", 2: "fmm"}[c.mode], c.svd_tolerance)

    def __repr__(self):
        return (f"FmmConfig({self.precision!r}, order={self.order}, leaf_level={self.leaf_level}, "
                f"mode={self.mode!r})")


class FmmEvaluator:
    """Source-target evaluator (FmmGenericEvaluatorBase<3>)."""

    _symmetric = False
    _create = "pam_fmm_evaluator_create"

    def __init__(self, rbf: Rbf, bbox=None):
        self._rbf_c = rbf.to_c()
        lo, hi = _bbox(bbox)
        self._bbox = (lo, hi)
        h = ctypes.c_void_p()
        create = self._create
        _capi.call(create, ctypes.byref(self._rbf_c), None if lo is None else _ptr(lo),
                   None if hi is None else _ptr(hi), ctypes.byref(h))
        self._h = h
        self.n_src = 0
        self.n_trg = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.load().pam_fmm_evaluator_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_accuracy(self, accuracy):
        _capi.call("pam_fmm_evaluator_set_accuracy", self._h, float(accuracy))

    def set_search_precision(self, precision):
        """Restrict the accuracy search to one precision ("fp64", "mixed",
        "mixed_comp"; None for the default list)."""
        c = -1 if precision is None else {"fp64": 0, "mixed": 1, "mixed_comp": 2}[precision]
        _capi.call("pam_fmm_evaluator_set_search_precision", self._h, c)

    def set_config(self, config: FmmConfig):
        c = config.to_c()
        _capi.call("pam_fmm_evaluator_set_config", self._h, ctypes.byref(c))

    def set_overlap(self, on):
        """Near field on a second stream beside the far field (default), or
        every stage in sequence (per-stage times of each stage alone)."""
        _capi.call("pam_fmm_evaluator_set_overlap", self._h, 1 if on else 0)

    def set_source_points(self, points):
        p = _points(points)
        self.n_src = p.shape[0]
        _capi.call("pam_fmm_evaluator_set_source_points", self._h, _ptr(p), p.shape[0])

    def set_target_points(self, points):
        p = _points(points)
        self.n_trg = p.shape[0]
        _capi.call("pam_fmm_evaluator_set_target_points", self._h, _ptr(p), p.shape[0])

    def set_weights(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        _capi.call("pam_fmm_evaluator_set_weights", self._h, _ptr(w), w.shape[0])

    def evaluate(self):
        n = self.n_src if self._symmetric else self.n_trg
        out = np.empty(n, dtype=np.float64)
        _capi.call("pam_fmm_evaluator_evaluate", self._h, _ptr(out), n)
        return out

    def evaluate_gradient(self):
        """Gradients at the targets (gradient-transpose evaluator): (n_trg, 3)."""
        if self._symmetric:
            raise RuntimeError("gradients are evaluated at target points: use a source-target evaluator")
        out = np.empty((self.n_trg, 3), dtype=np.float64)
        _capi.call("pam_fmm_evaluator_evaluate_gradient", self._h, _ptr(out), self.n_trg)
        return out

    def stats(self):
        s = _capi.PamFmmStats()
        _capi.call("pam_fmm_evaluator_stats", self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def config(self):
        c = _capi.PamFmmConfig()
        _capi.call("pam_fmm_evaluator_get_config", self._h, ctypes.byref(c))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def count_work(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _capi.call("pam_fmm_evaluator_count_work", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    # --- tree export (bit-exact checks) ---
    def tree_info(self, which=0):
        lvl = ctypes.c_int()
        lo = np.zeros(3)
        w, ic = ctypes.c_double(), ctypes.c_double()
        n, nl = ctypes.c_int64(), ctypes.c_int64()
        _capi.call("pam_fmm_evaluator_tree_info", self._h, which, ctypes.byref(lvl), _ptr(lo),
                   ctypes.byref(w), ctypes.byref(ic), ctypes.byref(n), ctypes.byref(nl))
        return dict(level=lvl.value, lo=lo, width=w.value, inv_cell=ic.value, n_points=n.value,
                    n_leaves=nl.value)

    def export_tree(self, which=0):
        info = self.tree_info(which)
        n, nl = info["n_points"], info["n_leaves"]
        codes = np.empty(n, dtype=np.uint64)
        perm = np.empty(n, dtype=np.int32)
        lc = np.empty(nl, dtype=np.uint64)
        ls = np.empty(nl + 1, dtype=np.int32)
        _capi.call("pam_fmm_evaluator_export_tree", self._h, which, _ptr(codes), _ptr(perm),
                   _ptr(lc), _ptr(ls))
        return codes, perm, lc, ls

    def export_level(self, which, level):
        n = ctypes.c_int64()
        _capi.call("pam_fmm_evaluator_export_level", self._h, which, level, ctypes.byref(n), None)
        codes = np.empty(n.value, dtype=np.uint64)
        _capi.call("pam_fmm_evaluator_export_level", self._h, which, level, ctypes.byref(n),
                   _ptr(codes))
        return codes

    def export_p2p(self):
        nl = self.tree_info(1)["n_leaves"]
        out = np.empty((nl, 27), dtype=np.int32)
        _capi.call("pam_fmm_evaluator_export_p2p", self._h, _ptr(out))
        return out

    def export_m2l(self, level):
        n = self.export_level(1, level).shape[0]
        out = np.empty((n, 343), dtype=np.int32)
        _capi.call("pam_fmm_evaluator_export_m2l", self._h, level, _ptr(out))
        return out


class FmmSymmetricEvaluator(FmmEvaluator):
    """Evaluator at the source points themselves (FmmGenericSymmetricEvaluatorBase<3>)."""

    _symmetric = True
    _create = "pam_fmm_symmetric_evaluator_create"

    def set_points(self, points):
        p = _points(points)
        self.n_src = p.shape[0]
        self.n_trg = p.shape[0]
        _capi.call("pam_fmm_evaluator_set_points", self._h, _ptr(p), p.shape[0])

    def set_source_points(self, points):
        raise TypeError("use set_points on a symmetric evaluator")

    def set_target_points(self, points):
        raise TypeError("use set_points on a symmetric evaluator")


def make_fmm_evaluator(rbf, bbox=None):
    return FmmEvaluator(rbf, bbox)


def make_fmm_symmetric_evaluator(rbf, bbox=None):
    return FmmSymmetricEvaluator(rbf, bbox)


def make_fmm_gradient_transpose_evaluator(rbf, bbox=None):
    return FmmGradientTransposeEvaluator(rbf, bbox)


def direct_evaluate_gradient(rbf: Rbf, src, weights, trg):
    """FP64 direct gradient sum on the GPU: (n_trg, 3) with row i the gradient
    of sum_j w_j phi(x - s_j) at x = t_i (the gradient kernel's full_direct)."""
    s = _points(src)
    t = _points(trg)
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    out = np.empty((t.shape[0], 3))
    rc = rbf.to_c()
    _capi.call("pam_direct_evaluate_gradient", ctypes.byref(rc), _ptr(s), s.shape[0], _ptr(w), _ptr(t),
               t.shape[0], _ptr(out))
    return out


def direct_evaluate(rbf: Rbf, src, weights, trg):
    """FP64 direct sum on the GPU (no polynomial terms)."""
    s = _points(src)
    t = _points(trg)
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    out = np.empty(t.shape[0])
    rc = rbf.to_c()
    _capi.call("pam_direct_evaluate", ctypes.byref(rc), _ptr(s), s.shape[0], _ptr(w), _ptr(t),
               t.shape[0], _ptr(out))
    return out


class FmmGradientTransposeEvaluator(FmmEvaluator):
    """make_fmm_gradient_transpose_evaluator (make_fmm_evaluator.cpp:119-148):
    evaluate() returns the gradients at the targets, (n_trg, 3), and the
    accuracy search measures them."""

    _create = "pam_fmm_gradient_transpose_evaluator_create"

    def evaluate(self):
        out = np.empty((self.n_trg, 3), dtype=np.float64)
        _capi.call("pam_fmm_evaluator_evaluate", self._h, _ptr(out), self.n_trg)
        return out


class FmmGradientEvaluator(FmmEvaluator):
    """make_fmm_gradient_evaluator (make_fmm_evaluator.cpp:78-107): sources carry
    gradient weights, (n_src, 3); evaluate() returns values at the targets,
    sum_j w_j . (-A^T grad phi(A (t_i - s_j)))."""

    _create = "pam_fmm_gradient_evaluator_create"

    def evaluate_gradient(self):
        raise RuntimeError("a gradient-kernel evaluator evaluates values")


def make_fmm_gradient_evaluator(rbf, bbox=None):
    return FmmGradientEvaluator(rbf, bbox)


class FmmHessianEvaluator(FmmEvaluator):
    """make_fmm_hessian_evaluator (make_fmm_evaluator.cpp:160-188): sources carry
    gradient weights, (n_src, 3); evaluate() returns (n_trg, 3),
    -sum_j A^T H(A (t_i - s_j)) A w_j."""

    _create = "pam_fmm_hessian_evaluator_create"

    def evaluate(self):
        out = np.empty((self.n_trg, 3), dtype=np.float64)
        _capi.call("pam_fmm_evaluator_evaluate", self._h, _ptr(out), self.n_trg)
        return out

    def evaluate_gradient(self):
        raise RuntimeError("a Hessian-kernel evaluator evaluates through evaluate()")


def make_fmm_hessian_evaluator(rbf, bbox=None):
    return FmmHessianEvaluator(rbf, bbox)


class FmmHessianSymmetricEvaluator(FmmHessianEvaluator):
    """make_fmm_hessian_symmetric_evaluator (make_fmm_evaluator.cpp:241-270):
    the Hessian kernel with the points as both sources and targets; weights
    (n, 3) in, (n, 3) out."""

    _create = "pam_fmm_hessian_symmetric_evaluator_create"

    def set_points(self, points):
        p = _points(points)
        self.n_src = p.shape[0]
        self.n_trg = p.shape[0]
        _capi.call("pam_fmm_evaluator_set_points", self._h, _ptr(p), p.shape[0])

    def set_source_points(self, points):
        raise TypeError("use set_points on a symmetric evaluator")

    def set_target_points(self, points):
        raise TypeError("use set_points on a symmetric evaluator")


def make_fmm_hessian_symmetric_evaluator(rbf, bbox=None):
    return FmmHessianSymmetricEvaluator(rbf, bbox)
