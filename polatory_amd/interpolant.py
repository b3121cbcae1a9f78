"""Host-side mirror of polatory's Model / Interpolant API over the C ABI.

Model          include/polatory/model.hpp:32-183
Interpolant    include/polatory/interpolant.hpp:34-367 (fit, evaluate, weights)
RasPreconditioner (test seam) include/polatory/preconditioner/ras_preconditioner.hpp:56-250
fgmres_dense   (test seam) src/krylov/fgmres.cpp, gmres.cpp
"""
import ctypes
import math
import os

import numpy as np

from . import _capi
from .fmm import FmmConfig, _points, _ptr
from .rbf import Rbf, from_c

POLY_DEGREE_MIN = -2  # Model::kMinRequiredPolyDegree


class Model:
    def __init__(self, rbfs, poly_degree=POLY_DEGREE_MIN, nugget=0.0):
        if isinstance(rbfs, Rbf):
            rbfs = [rbfs]
        if len(rbfs) == 0:
            raise ValueError("rbfs must not be empty")
        self.rbfs = list(rbfs)
        self._arr = (_capi.PamRbf * len(rbfs))(*[r.to_c() for r in rbfs])
        h = ctypes.c_void_p()
        _capi.call("pam_model_create", ctypes.cast(self._arr, ctypes.c_void_p), len(rbfs),
                   int(poly_degree), float(nugget), ctypes.byref(h))
        self._h = h
        self._nugget = float(nugget)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.load().pam_model_destroy(h)
            except Exception:
                pass
            self._h = None

    def nugget(self):
        return self._nugget

    def poly_degree(self):
        d = ctypes.c_int()
        _capi.call("pam_model_poly_degree", self._h, ctypes.byref(d))
        return d.value

    def poly_basis_size(self):
        l = ctypes.c_int()
        _capi.call("pam_model_poly_basis_size", self._h, ctypes.byref(l))
        return l.value

    def cpd_order(self):
        return max(r.cpd_order for r in self.rbfs)

    # Model::save / Model::load (model.hpp:169; polatory's binary model file)
    def save(self, filename):
        _capi.call("pam_model_save", self._h, os.fsencode(filename))

    @staticmethod
    def load(filename):
        h = ctypes.c_void_p()
        _capi.call("pam_model_load", os.fsencode(filename), ctypes.byref(h))
        try:
            return Model._from_handle(h)
        finally:
            _capi.load().pam_model_destroy(h)

    @staticmethod
    def _from_handle(h):
        n = ctypes.c_int()
        _capi.call("pam_model_num_rbfs", h, ctypes.byref(n))
        rbfs = []
        for i in range(n.value):
            r = _capi.PamRbf()
            _capi.call("pam_model_rbf", h, i, ctypes.byref(r))
            rbfs.append(from_c(r))
        d = ctypes.c_int()
        _capi.call("pam_model_poly_degree", h, ctypes.byref(d))
        nug = ctypes.c_double()
        _capi.call("pam_model_nugget", h, ctypes.byref(nug))
        return Model(rbfs, poly_degree=d.value, nugget=nug.value)


class FitOptions:
    """Configuration of the device fit (no reference counterpart; defaults follow
    pam_fit_options_default)."""

    def __init__(self, op=None, residual=None, level=None, use_preconditioner=True,
                 refine_rtol=None, verbose=False, use_graphs=True, keep_unconverged=False,
                 tune_solves=False, ras_fp32_factors=None, tune_factor=False, factor_panel=None,
                 cpd2_auto=True):
        d = _capi.PamFitOptions()
        _capi.call("pam_fit_options_default", ctypes.byref(d))
        self._c = d
        if op is not None:
            d.op = op.to_c()
        if residual is not None:
            d.residual = residual.to_c()
        if level is not None:
            d.level = level.to_c()
            d.use_level_config = 1
        d.use_preconditioner = 1 if use_preconditioner else 0
        if refine_rtol is not None:
            d.refine_rtol = refine_rtol
        d.verbose = 1 if verbose else 0
        d.use_graphs = 1 if use_graphs else 0
        d.keep_unconverged = 1 if keep_unconverged else 0
        d.tune_solves = 1 if tune_solves else 0
        d.tune_factor = 1 if tune_factor else 0
        if factor_panel is not None:  # None: the rule or tune_factor (-1)
            d.factor_panel = int(factor_panel)
        d.cpd2_auto = 1 if cpd2_auto else 0
        if ras_fp32_factors is not None:  # None: the library's rule (-1)
            d.ras_fp32_factors = 1 if ras_fp32_factors else 0


class Interpolant:
    def __init__(self, model: Model):
        self.model = model
        h = ctypes.c_void_p()
        _capi.call("pam_interpolant_create", model._h, ctypes.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.load().pam_interpolant_destroy(h)
            except Exception:
                pass
            self._h = None

    # Interpolant::save / Interpolant::load (interpolant.hpp:218; polatory's binary file)
    def save(self, filename):
        _capi.call("pam_interpolant_save", self._h, os.fsencode(filename))

    @staticmethod
    def load(filename):
        h = ctypes.c_void_p()
        _capi.call("pam_interpolant_load", os.fsencode(filename), ctypes.byref(h))
        ip = Interpolant.__new__(Interpolant)
        ip._h = h
        mh = ctypes.c_void_p()
        _capi.call("pam_interpolant_model", h, ctypes.byref(mh))
        try:
            ip.model = Model._from_handle(mh)
        finally:
            _capi.load().pam_model_destroy(mh)
        return ip

    def fitted(self):
        f = ctypes.c_int()
        _capi.call("pam_interpolant_fitted", self._h, ctypes.byref(f))
        return bool(f.value)

    def fit(self, points, values, tolerance, max_iter=100, accuracy=math.inf, initial=None,
            options: FitOptions = None):
        p = _points(points)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.shape[0] != p.shape[0]:
            raise ValueError(f"values.rows() must be equal to {p.shape[0]}")
        opt = None if options is None else ctypes.byref(options._c)
        _capi.call("pam_interpolant_fit", self._h, _ptr(p), p.shape[0], _ptr(v), float(tolerance),
                   int(max_iter), float(accuracy), opt, None if initial is None else initial._h)

    def fit_hermite(self, points, grad_points, values, tolerance, grad_tolerance, max_iter=100,
                    accuracy=math.inf, grad_accuracy=math.inf, initial=None, options: FitOptions = None):
        """fit(points, grad_points, values, tolerance, grad_tolerance, ...)
        (interpolant.hpp:83-113): values holds the values at points, then the
        gradients at grad_points (x y z per point, or an (n, 3) array appended
        after them). The weights are then [w ; g ; c]."""
        p = _points(points)
        g = _points(grad_points)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.shape[0] != p.shape[0] + 3 * g.shape[0]:
            raise ValueError("values.rows() must be equal to points.rows() + 3 * grad_points.rows()")
        opt = None if options is None else ctypes.byref(options._c)
        _capi.call("pam_interpolant_fit_hermite", self._h, _ptr(p), p.shape[0], _ptr(g), g.shape[0], _ptr(v),
                   float(tolerance), float(grad_tolerance), int(max_iter), float(accuracy), float(grad_accuracy),
                   opt, None if initial is None else initial._h)

    def fit_incrementally_hermite(self, points, grad_points, values, tolerance, grad_tolerance, max_iter=100,
                                  accuracy=math.inf, grad_accuracy=math.inf, options: FitOptions = None):
        """fit_incrementally with gradient data (interpolant.hpp:121-147)."""
        p = _points(points)
        g = _points(grad_points)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.shape[0] != p.shape[0] + 3 * g.shape[0]:
            raise ValueError("values.rows() must be equal to points.rows() + 3 * grad_points.rows()")
        opt = None if options is None else ctypes.byref(options._c)
        _capi.call("pam_interpolant_fit_incrementally_hermite", self._h, _ptr(p), p.shape[0], _ptr(g), g.shape[0],
                   _ptr(v), float(tolerance), float(grad_tolerance), int(max_iter), float(accuracy),
                   float(grad_accuracy), opt)

    def evaluate_hermite(self, points, grad_points):
        """Values at points, then gradients (x y z) at grad_points, of any
        interpolant (interpolant.hpp:58-67)."""
        p = _points(points)
        g = _points(grad_points)
        out = np.empty(p.shape[0] + 3 * g.shape[0])
        _capi.call("pam_interpolant_evaluate_hermite", self._h, _ptr(p), p.shape[0], _ptr(g), g.shape[0],
                   _ptr(out))
        return out

    def num_grad_centers(self):
        n = ctypes.c_int64()
        _capi.call("pam_interpolant_num_grad_centers", self._h, ctypes.byref(n))
        return n.value

    def grad_centers(self):
        out = np.empty((self.num_grad_centers(), 3))
        _capi.call("pam_interpolant_grad_centers", self._h, _ptr(out))
        return out

    def fit_incrementally(self, points, values, tolerance, max_iter=100, accuracy=math.inf,
                          options: FitOptions = None):
        """interpolant.hpp:115-147: fits on a growing subset of the points;
        centers() then holds the chosen points in the order they were added."""
        p = _points(points)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.shape[0] != p.shape[0]:
            raise ValueError(f"values.rows() must be equal to {p.shape[0]}")
        opt = None if options is None else ctypes.byref(options._c)
        _capi.call("pam_interpolant_fit_incrementally", self._h, _ptr(p), p.shape[0], _ptr(v),
                   float(tolerance), int(max_iter), float(accuracy), opt)

    def fit_inequality(self, points, values, values_lb, values_ub, tolerance, max_iter=100,
                       accuracy=math.inf, initial=None, options: FitOptions = None):
        """interpolant.hpp:149-190: NaN means no equality / lower / upper
        constraint at that point; centers() then holds the equality points and
        the active bounds."""
        p = _points(points)
        n = p.shape[0]
        arrs = []
        for name, a in (("values", values), ("values_lb", values_lb), ("values_ub", values_ub)):
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.shape[0] != n:
                raise ValueError(f"{name}.rows() must be equal to points.rows()")
            arrs.append(a)
        opt = None if options is None else ctypes.byref(options._c)
        _capi.call("pam_interpolant_fit_inequality", self._h, _ptr(p), n, _ptr(arrs[0]), _ptr(arrs[1]),
                   _ptr(arrs[2]), float(tolerance), int(max_iter), float(accuracy), opt,
                   None if initial is None else initial._h)

    def evaluate(self, points, grad_points=None, accuracy=math.inf, grad_accuracy=math.inf):
        """Values at points; with grad_points, the values followed by the
        gradients at grad_points (x y z per point), as interpolant.hpp:58-67
        returns them. Each half has its own accuracy search: `accuracy` for
        the values, `grad_accuracy` for the gradients."""
        p = _points(points)
        out = np.empty(p.shape[0])
        _capi.call("pam_interpolant_evaluate", self._h, _ptr(p), p.shape[0], float(accuracy), _ptr(out))
        if grad_points is None:
            return out
        return np.concatenate([out, self.evaluate_gradient(grad_points, grad_accuracy).reshape(-1)])

    def evaluate_gradient(self, points, accuracy=math.inf):
        """Gradients of the fitted function at points: (n, 3)."""
        p = _points(points)
        out = np.empty((p.shape[0], 3))
        _capi.call("pam_interpolant_evaluate_gradient", self._h, _ptr(p), p.shape[0], float(accuracy),
                   _ptr(out))
        return out

    def set_evaluation_bbox(self, bbox_min, bbox_max, accuracy=math.inf, grad_accuracy=math.inf):
        """set_evaluation_bbox_impl (interpolant.hpp:201-210): builds the
        evaluators for the union of the box and the centres' box, at these
        accuracies, for evaluate_impl to reuse."""
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64).reshape(3)
        hi = np.ascontiguousarray(bbox_max, dtype=np.float64).reshape(3)
        _capi.call("pam_interpolant_set_evaluation_bbox", self._h, _ptr(lo), _ptr(hi), float(accuracy),
                   float(grad_accuracy))

    def evaluate_impl(self, points):
        """evaluate_impl (interpolant.hpp:69-75): values at points with the
        evaluator of the last set_evaluation_bbox or evaluate call."""
        p = _points(points)
        out = np.empty(p.shape[0])
        _capi.call("pam_interpolant_evaluate_impl", self._h, _ptr(p), p.shape[0], _ptr(out))
        return out

    def set_eval_config(self, config: FmmConfig = None):
        if config is None:
            _capi.call("pam_interpolant_set_eval_config", self._h, None)
        else:
            c = config.to_c()
            _capi.call("pam_interpolant_set_eval_config", self._h, ctypes.byref(c))

    def num_centers(self):
        n = ctypes.c_int64()
        _capi.call("pam_interpolant_num_centers", self._h, ctypes.byref(n))
        return n.value

    def centers(self):
        out = np.empty((self.num_centers(), 3))
        _capi.call("pam_interpolant_centers", self._h, _ptr(out))
        return out

    def weights(self):
        out = np.empty(self.num_centers() + 3 * self.num_grad_centers() + self.model.poly_basis_size())
        _capi.call("pam_interpolant_weights", self._h, _ptr(out))
        return out

    def set_fitted(self, centers, weights):
        c = _points(centers)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        _capi.call("pam_interpolant_set_fitted", self._h, _ptr(c), c.shape[0], _ptr(w))

    def fit_stats(self):
        s = _capi.PamFitStats()
        _capi.call("pam_interpolant_fit_stats", self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def residual_history(self):
        n = ctypes.c_int64()
        _capi.call("pam_interpolant_residual_history", self._h, None, ctypes.byref(n))
        out = np.empty(n.value)
        _capi.call("pam_interpolant_residual_history", self._h, _ptr(out), ctypes.byref(n))
        return out


class RbfFieldFunction:
    """isosurface::RbfFieldFunction (isosurface/rbf_field_function.hpp:12-36):
    the field an isosurface extractor samples, a fitted interpolant evaluated
    in batches inside a box fixed beforehand."""

    def __init__(self, interpolant: Interpolant, accuracy=math.inf, grad_accuracy=math.inf):
        self.interpolant = interpolant
        self.accuracy = accuracy
        self.grad_accuracy = grad_accuracy

    def set_evaluation_bbox(self, bbox_min, bbox_max):
        self.interpolant.set_evaluation_bbox(bbox_min, bbox_max, self.accuracy, self.grad_accuracy)

    def __call__(self, points):
        return self.interpolant.evaluate_impl(points)


class RasPreconditioner:
    """Device multilevel RAS preconditioner (exposed for tests)."""

    def __init__(self, model: Model, points, options=None, grad_points=None):
        """options: a FitOptions, for the preconditioner that a fit with those
        options builds (precision and form of the factors, coarse grid).
        grad_points: the preconditioner of a fit with gradient data (vectors
        [w ; g ; c])."""
        self.model = model
        p = _points(points)
        self.n = p.shape[0]
        h = ctypes.c_void_p()
        if grad_points is not None:
            g = _points(grad_points)
            opt = None if options is None else ctypes.byref(options._c)
            _capi.call("pam_ras_create_hermite", model._h, _ptr(p), p.shape[0], _ptr(g), g.shape[0], opt,
                       ctypes.byref(h))
        elif options is None:
            _capi.call("pam_ras_create", model._h, _ptr(p), p.shape[0], ctypes.byref(h))
        else:
            _capi.call("pam_ras_create_with_options", model._h, _ptr(p), p.shape[0], ctypes.byref(options._c),
                       ctypes.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.load().pam_ras_destroy(h)
            except Exception:
                pass
            self._h = None

    def n_levels(self):
        n = ctypes.c_int()
        _capi.call("pam_ras_num_levels", self._h, ctypes.byref(n))
        return n.value

    def level_points(self, level):
        n = ctypes.c_int64()
        _capi.call("pam_ras_level_points", self._h, level, None, ctypes.byref(n))
        out = np.empty(n.value, dtype=np.int32)
        _capi.call("pam_ras_level_points", self._h, level, _ptr(out), ctypes.byref(n))
        return out

    def domains(self, level):
        nd = ctypes.c_int()
        _capi.call("pam_ras_num_domains", self._h, level, ctypes.byref(nd))
        res = []
        for d in range(nd.value):
            n = ctypes.c_int64()
            _capi.call("pam_ras_domain", self._h, level, d, None, None, ctypes.byref(n))
            idx = np.empty(n.value, dtype=np.int32)
            inner = np.empty(n.value, dtype=np.uint8)
            _capi.call("pam_ras_domain", self._h, level, d, _ptr(idx), _ptr(inner), ctypes.byref(n))
            res.append((idx, inner.astype(bool)))
        return res

    def apply(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        out = np.empty_like(v)
        _capi.call("pam_ras_apply", self._h, _ptr(v), _ptr(out))
        return out

    def size(self):
        n = ctypes.c_int64()
        _capi.call("pam_ras_size", self._h, ctypes.byref(n))
        return n.value

    def apply_device(self, d_v, d_out):
        """operator()(v) on device vectors (addresses as ints), enqueued on the
        handle's stream; complete after synchronize()."""
        _capi.call("pam_ras_apply_device", self._h, ctypes.c_void_p(d_v), ctypes.c_void_p(d_out))

    def synchronize(self):
        _capi.call("pam_ras_synchronize", self._h)


class Operator:
    """interpolation::Operator (operator.hpp:38-111) as a LinearOperator:
    y = [(A + nugget I) w + P c ; P^T w] for x = [w ; c] at the points."""

    def __init__(self, model: Model, points, config: FmmConfig = None, grad_points=None):
        """With grad_points (operator.hpp:30-36), x = [w ; g ; c] with g the 3-vector
        weights of the gradient points (x y z per point) and the output rows
        [values ; gradients ; P^T [w ; g]]; the default configuration is then FP64
        FMM of order 10."""
        self.model = model
        p = _points(points)
        h = ctypes.c_void_p()
        c = None if config is None else ctypes.byref(config.to_c())
        if grad_points is None:
            _capi.call("pam_operator_create", model._h, _ptr(p), p.shape[0], c, ctypes.byref(h))
        else:
            g = _points(grad_points)
            _capi.call("pam_operator_create_hermite", model._h, _ptr(p), p.shape[0], _ptr(g), g.shape[0], c,
                       ctypes.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.load().pam_operator_destroy(h)
            except Exception:
                pass
            self._h = None

    def size(self):
        n = ctypes.c_int64()
        _capi.call("pam_operator_size", self._h, ctypes.byref(n))
        return n.value

    def __call__(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.shape[0] != self.size():
            raise ValueError("x must have size() entries")
        y = np.empty_like(x)
        _capi.call("pam_operator_apply", self._h, _ptr(x), _ptr(y))
        return y

    def apply_device(self, d_x, d_y):
        """operator()(x) on device vectors (addresses as ints), enqueued on the
        handle's stream; complete after synchronize()."""
        _capi.call("pam_operator_apply_device", self._h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y))

    def synchronize(self):
        _capi.call("pam_operator_synchronize", self._h)


def fgmres_dense(a, rhs, n_iter, m=None, x0=None):
    """Device FGMRES on a dense operator; returns (x, |g_j| history, iterations)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    mm = None if m is None else np.ascontiguousarray(m, dtype=np.float64)
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    x = np.empty(n)
    hist = np.empty(n_iter + 1)
    it = ctypes.c_int()
    _capi.call("pam_fgmres_dense", _ptr(a), None if mm is None else _ptr(mm), n, _ptr(rhs),
               None if x0a is None else _ptr(x0a), int(n_iter), _ptr(x), _ptr(hist), ctypes.byref(it))
    return x, hist, it.value


def distance_filter(points, distance, indices=None):
    """DistanceFilter(points).filter(distance, indices).filtered_indices()
    (point_cloud/distance_filter.hpp:21-56)."""
    p = _points(points)
    idx = np.arange(p.shape[0], dtype=np.int64) if indices is None else \
        np.ascontiguousarray(indices, dtype=np.int64)
    out = np.empty(idx.shape[0], dtype=np.int64)
    n_out = ctypes.c_int64()
    _capi.call("pam_distance_filter", _ptr(p), p.shape[0], float(distance), _ptr(idx), idx.shape[0],
               _ptr(out), ctypes.byref(n_out))
    return out[:n_out.value]


def cross_validate(model: Model, points, values, set_ids, tolerance, max_iter=100, accuracy=math.inf,
                   options: FitOptions = None):
    """kriging::cross_validate (kriging/cross_validate.hpp:14-47): for each set
    id, the predictions at its points from the fit on the other points."""
    p = _points(points)
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    ids = np.ascontiguousarray(set_ids, dtype=np.int32).reshape(-1)
    if v.shape[0] != p.shape[0] or ids.shape[0] != p.shape[0]:
        raise ValueError("values and set_ids must have one entry per point")
    out = np.empty(p.shape[0])
    opt = None if options is None else ctypes.byref(options._c)
    _capi.call("pam_cross_validate", model._h, _ptr(p), p.shape[0], _ptr(v), _ptr(ids), float(tolerance),
               int(max_iter), float(accuracy), opt, _ptr(out))
    return out


class DomainDivider:
    """preconditioner::DomainDivider (domain_divider.hpp:30-125), host only:
    the overlapping domains of one RAS level and the choice of the coarser
    level's points; gradient points count 3 rows each."""

    def __init__(self, points, grad_points, point_indices, grad_point_indices, poly_point_indices):
        p = _points(points)
        g = _points(grad_points)
        self._idx = [np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
                     for a in (point_indices, grad_point_indices, poly_point_indices)]
        h = ctypes.c_void_p()
        i, gi, pi = self._idx
        _capi.call("pam_domain_divider_create", _ptr(p), p.shape[0], _ptr(g), g.shape[0], _ptr(i), i.shape[0],
                   _ptr(gi), gi.shape[0], _ptr(pi), pi.shape[0], ctypes.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.load().pam_domain_divider_destroy(h)
            except Exception:
                pass
            self._h = None

    def domains(self):
        """[(point_indices, inner_point, grad_point_indices, inner_grad_point)] per domain."""
        nd = ctypes.c_int64()
        _capi.call("pam_domain_divider_num_domains", self._h, ctypes.byref(nd))
        out = []
        for d in range(nd.value):
            n, ng = ctypes.c_int64(), ctypes.c_int64()
            _capi.call("pam_domain_divider_domain", self._h, d, None, None, ctypes.byref(n), None, None,
                       ctypes.byref(ng))
            idx = np.empty(n.value, dtype=np.int32)
            inner = np.empty(n.value, dtype=np.uint8)
            gidx = np.empty(ng.value, dtype=np.int32)
            ginner = np.empty(ng.value, dtype=np.uint8)
            _capi.call("pam_domain_divider_domain", self._h, d, _ptr(idx), _ptr(inner), ctypes.byref(n),
                       _ptr(gidx), _ptr(ginner), ctypes.byref(ng))
            out.append((idx, inner.astype(bool), gidx, ginner.astype(bool)))
        return out

    def choose_coarse_points(self, n_coarse):
        n, ng = ctypes.c_int64(), ctypes.c_int64()
        _capi.call("pam_domain_divider_choose_coarse_points", self._h, int(n_coarse), None, ctypes.byref(n), None,
                   ctypes.byref(ng))
        idx = np.empty(n.value, dtype=np.int32)
        gidx = np.empty(ng.value, dtype=np.int32)
        _capi.call("pam_domain_divider_choose_coarse_points", self._h, int(n_coarse), _ptr(idx), ctypes.byref(n),
                   _ptr(gidx), ctypes.byref(ng))
        return idx, gidx
