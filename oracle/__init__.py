"""CPU oracle for the RBF fit/evaluate hot path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``oracle/liboracle.so`` (a plain-C FP64 restatement of the
reference semantics, see ``oracle.h`` for the file:line each function follows).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker; the product package
``polatory_amd`` never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

BH3, TH3, EXP = 0, 1, 2


class OrcRbf(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("params", ctypes.c_double * 2),
                ("aniso", ctypes.c_double * 9)]


class OrcFmmStats(ctypes.Structure):
    _fields_ = [("t_tree", ctypes.c_double), ("t_upward", ctypes.c_double), ("t_m2l", ctypes.c_double),
                ("t_downward", ctypes.c_double), ("t_p2p", ctypes.c_double), ("t_total", ctypes.c_double),
                ("height", ctypes.c_int), ("order", ctypes.c_int), ("m2l_pairs", ctypes.c_int64),
                ("p2p_pairs", ctypes.c_int64), ("n_src_leaves", ctypes.c_int64),
                ("n_trg_leaves", ctypes.c_int64)]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run oracle.build())")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.orc_rbf_evaluate.restype = ctypes.c_double
        L.orc_rbf_evaluate.argtypes = [ctypes.POINTER(OrcRbf), ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double]
        L.orc_rbf_gradient.argtypes = [ctypes.POINTER(OrcRbf), ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, P]
        L.orc_rbf_hessian.argtypes = [ctypes.POINTER(OrcRbf), ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, P]
        L.orc_direct_gradient_sum.argtypes = [P, P, i64, P, P, i64, P]
        L.orc_transform_points.argtypes = [P, P, i64, P]
        L.orc_poly_basis_size.restype = i64
        L.orc_poly_basis_size.argtypes = [ctypes.c_int]
        L.orc_monomials.argtypes = [ctypes.c_int, P, i64, P]
        L.orc_direct_evaluate.argtypes = [P, ctypes.c_int, ctypes.c_int, P, i64, P, P, i64, P]
        L.orc_direct_rbf_sum.argtypes = [P, P, i64, P, P, i64, P]
        L.orc_direct_operator.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, i64,
                                          P, P]
        L.orc_monomial_gradients.argtypes = [ctypes.c_int, P, i64, P]
        L.orc_hermite_operator.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, i64, P, i64,
                                           P, P]
        L.orc_make_box.argtypes = [P, P, P, P, P]
        L.orc_morton_codes.argtypes = [P, i64, P, ctypes.c_double, ctypes.c_int, P]
        L.orc_stable_sort_codes.argtypes = [P, i64, P]
        L.orc_leaves.restype = i64
        L.orc_leaves.argtypes = [P, i64, P, P]
        L.orc_p2p_lists.argtypes = [P, i64, P, i64, ctypes.c_int, P]
        L.orc_m2l_lists.argtypes = [P, i64, P, i64, ctypes.c_int, P]
        L.orc_fgmres_dense.restype = ctypes.c_int
        L.orc_fgmres_dense.argtypes = [P, P, i64, P, P, ctypes.c_int, P, P]
        L.orc_cpu_fmm_evaluate.restype = ctypes.c_int
        L.orc_cpu_fmm_evaluate.argtypes = [P, P, P, P, i64, P, P, i64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, P, ctypes.POINTER(OrcFmmStats)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def make_rbf(kind, params=None, aniso=None):
    r = OrcRbf()
    r.kind = kind
    if params is None:
        params = [1.0, 0.0] if kind in (BH3, TH3) else [1.0, 1.0]
    params = list(params) + [0.0] * (2 - len(params))
    r.params[0], r.params[1] = params[0], params[1]
    a = np.eye(3) if aniso is None else np.asarray(aniso, dtype=np.float64).reshape(3, 3)
    for i in range(9):
        r.aniso[i] = a.flat[i]
    return r


def _rbf_array(rbfs):
    if isinstance(rbfs, OrcRbf):
        rbfs = [rbfs]
    arr = (OrcRbf * len(rbfs))(*rbfs)
    return arr, len(rbfs)


def rbf_evaluate(rbf, d):
    return lib().orc_rbf_evaluate(ctypes.byref(rbf), float(d[0]), float(d[1]), float(d[2]))


def rbf_gradient(rbf, d):
    g = np.empty(3)
    lib().orc_rbf_gradient(ctypes.byref(rbf), float(d[0]), float(d[1]), float(d[2]), _p(g))
    return g


def rbf_hessian(rbf, d):
    h = np.empty(9)
    lib().orc_rbf_hessian(ctypes.byref(rbf), float(d[0]), float(d[1]), float(d[2]), _p(h))
    return h.reshape(3, 3)


def direct_gradient_sum(rbf, src, w, trg):
    s = _f64(src, (-1, 3))
    t = _f64(trg, (-1, 3))
    w = _f64(w)
    out = np.empty((t.shape[0], 3))
    lib().orc_direct_gradient_sum(ctypes.byref(rbf), _p(s), s.shape[0], _p(w), _p(t), t.shape[0], _p(out))
    return out


def transform_points(aniso, pts):
    pts = _f64(pts, (-1, 3))
    a = _f64(aniso, (9,))
    out = np.empty_like(pts)
    lib().orc_transform_points(_p(a), _p(pts), pts.shape[0], _p(out))
    return out


def poly_basis_size(degree):
    return int(lib().orc_poly_basis_size(degree))


def monomials(degree, pts):
    pts = _f64(pts, (-1, 3))
    out = np.empty((pts.shape[0], poly_basis_size(degree)))
    lib().orc_monomials(degree, _p(pts), pts.shape[0], _p(out))
    return out


def direct_evaluate(rbfs, poly_degree, src, weights, trg):
    arr, n = _rbf_array(rbfs)
    src = _f64(src, (-1, 3))
    trg = _f64(trg, (-1, 3))
    w = _f64(weights)
    assert w.shape[0] == src.shape[0] + poly_basis_size(poly_degree)
    out = np.empty(trg.shape[0])
    lib().orc_direct_evaluate(ctypes.cast(arr, ctypes.c_void_p), n, poly_degree, _p(src),
                              src.shape[0], _p(w), _p(trg), trg.shape[0], _p(out))
    return out


def direct_rbf_sum(rbf, src, w, trg):
    src = _f64(src, (-1, 3))
    trg = _f64(trg, (-1, 3))
    w = _f64(w)
    out = np.empty(trg.shape[0])
    lib().orc_direct_rbf_sum(ctypes.byref(rbf), _p(src), src.shape[0], _p(w), _p(trg),
                             trg.shape[0], _p(out))
    return out


def direct_operator(rbfs, poly_degree, nugget, pts, x):
    arr, n = _rbf_array(rbfs)
    pts = _f64(pts, (-1, 3))
    x = _f64(x)
    y = np.empty_like(x)
    lib().orc_direct_operator(ctypes.cast(arr, ctypes.c_void_p), n, poly_degree, float(nugget),
                              _p(pts), pts.shape[0], _p(x), _p(y))
    return y


def monomial_gradients(degree, pts):
    """(3 n) x l: row 3 i + c holds the derivatives of the monomials along axis c at point i."""
    pts = _f64(pts, (-1, 3))
    l = lib().orc_poly_basis_size(degree)
    out = np.empty((3 * pts.shape[0], max(l, 1)))
    lib().orc_monomial_gradients(degree, _p(pts), pts.shape[0], _p(out))
    return out[:, :l]


def hermite_operator(rbfs, poly_degree, nugget, pts, gpts, x):
    """Operator of a fit with gradient data: x = [w (mu); g (3 sigma); c (l)]."""
    arr, n = _rbf_array(rbfs)
    pts = _f64(pts, (-1, 3))
    gpts = _f64(gpts, (-1, 3))
    x = _f64(x)
    y = np.empty_like(x)
    lib().orc_hermite_operator(ctypes.cast(arr, ctypes.c_void_p), n, poly_degree, float(nugget),
                               _p(pts), pts.shape[0], _p(gpts), gpts.shape[0], _p(x), _p(y))
    return y


def make_box(bmin, bmax, aniso=None):
    a = _f64(np.eye(3) if aniso is None else aniso, (9,))
    bmin = _f64(bmin, (3,))
    bmax = _f64(bmax, (3,))
    c = np.empty(3)
    w = np.empty(1)
    lib().orc_make_box(_p(bmin), _p(bmax), _p(a), _p(c), _p(w))
    return c, float(w[0])


def morton_codes(pts, lo, inv_cell, level):
    pts = _f64(pts, (-1, 3))
    lo = _f64(lo, (3,))
    out = np.empty(pts.shape[0], dtype=np.uint64)
    lib().orc_morton_codes(_p(pts), pts.shape[0], _p(lo), float(inv_cell), level, _p(out))
    return out


def stable_sort_codes(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint64)
    perm = np.empty(codes.shape[0], dtype=np.int32)
    lib().orc_stable_sort_codes(_p(codes), codes.shape[0], _p(perm))
    return perm


def leaves(sorted_codes):
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
    lc = np.empty(max(sc.shape[0], 1), dtype=np.uint64)
    ls = np.empty(sc.shape[0] + 1, dtype=np.int32)
    nl = int(lib().orc_leaves(_p(sc), sc.shape[0], _p(lc), _p(ls)))
    return lc[:nl].copy(), ls[:nl + 1].copy()


def p2p_lists(trg_codes, src_codes, level):
    t = np.ascontiguousarray(trg_codes, dtype=np.uint64)
    s = np.ascontiguousarray(src_codes, dtype=np.uint64)
    out = np.empty((t.shape[0], 27), dtype=np.int32)
    lib().orc_p2p_lists(_p(t), t.shape[0], _p(s), s.shape[0], level, _p(out))
    return out


def m2l_lists(trg_codes, src_codes, level):
    t = np.ascontiguousarray(trg_codes, dtype=np.uint64)
    s = np.ascontiguousarray(src_codes, dtype=np.uint64)
    out = np.empty((t.shape[0], 343), dtype=np.int32)
    lib().orc_m2l_lists(_p(t), t.shape[0], _p(s), s.shape[0], level, _p(out))
    return out


def fgmres_dense(a, rhs, n_iter, m=None, x0=None):
    a = _f64(a)
    n = a.shape[0]
    rhs = _f64(rhs)
    x0 = np.zeros(n) if x0 is None else _f64(x0)
    mm = None if m is None else _f64(m)
    x = np.empty(n)
    hist = np.zeros(n_iter + 1)
    it = lib().orc_fgmres_dense(_p(a), _p(mm), n, _p(rhs), _p(x0), n_iter, _p(x), _p(hist))
    return x, hist[:it + 1], it


def cpu_fmm_evaluate(rbf, src, w, trg, bbox, order, fh_degree=-1, height=0):
    """The reference-algorithm CPU FMM (cpu_fmm.c): polatory's evaluate path
    over a uniform-interpolation FMM, FP64, OpenMP. Returns (values, stats)."""
    src = _f64(src, (-1, 3))
    trg = _f64(trg, (-1, 3))
    w = _f64(w)
    lo = _f64(bbox[0], (3,))
    hi = _f64(bbox[1], (3,))
    out = np.empty(trg.shape[0])
    st = OrcFmmStats()
    rc = lib().orc_cpu_fmm_evaluate(ctypes.byref(rbf), _p(lo), _p(hi), _p(src), src.shape[0], _p(w), _p(trg),
                                    trg.shape[0], order, fh_degree, height, _p(out), ctypes.byref(st))
    if rc != 0:
        raise ValueError(f"cpu_fmm_evaluate: order or height out of range ({rc})")
    return out, {k: getattr(st, k) for k, _ in st._fields_}
