"""RBF descriptors mirroring polatory's rbf classes (include/polatory/rbf/).

Biharmonic3D  phi = -slope * sqrt(r^2 + c^2)   params [slope=1, c=0]  cpd order 1
Triharmonic3D phi = +slope * (r^2 + c^2)^1.5   params [slope=1, c=0]  cpd order 2
CovExponential phi = psill * exp(-3 r / range) params [psill, range]  cpd order 0
(rbf/polyharmonic_odd.hpp:13-141, rbf/cov_exponential.hpp:12-68, rbf_base.hpp:15-101)
"""
import numpy as np

from . import _capi


class Rbf:
    kind = None
    short_name = ""
    cpd_order = 0
    n_params = 2

    def __init__(self, params=None):
        self._params = []
        self._aniso = np.eye(3)
        self.set_parameters([] if params is None else list(params))

    # rbf_base.hpp:73-79
    def set_anisotropy(self, aniso):
        a = np.asarray(aniso, dtype=np.float64).reshape(3, 3)
        if not np.linalg.det(a) > 0.0:
            raise ValueError("aniso must have a positive determinant")
        self._aniso = a.copy()

    def anisotropy(self):
        return self._aniso.copy()

    def parameters(self):
        return list(self._params)

    def set_parameters(self, params):
        if len(params) != self.n_params:
            raise ValueError(f"params.size() must be {self.n_params}")
        self._params = [float(p) for p in params]

    def to_c(self):
        r = _capi.PamRbf()
        r.kind = self.kind
        r.params[0], r.params[1] = self._params[0], self._params[1]
        for i in range(9):
            r.aniso[i] = self._aniso.flat[i]
        return r

    def __eq__(self, other):
        return (type(self) is type(other) and self._params == other._params
                and np.array_equal(self._aniso, other._aniso))

    # rbf_base.hpp:40-53, evaluated on the GPU for a batch of differences
    def _point(self, diff, what, per):
        import ctypes
        d = np.ascontiguousarray(diff, dtype=np.float64).reshape(-1, 3)
        out = np.empty((d.shape[0], per) if per > 1 else d.shape[0])
        rc = self.to_c()
        _capi.call(what, ctypes.byref(rc), d.ctypes.data_as(ctypes.c_void_p), d.shape[0],
                   out.ctypes.data_as(ctypes.c_void_p))
        return out

    def evaluate(self, diff):
        """phi(A d) for each row d of diff (n, 3) -> (n,)."""
        return self._point(diff, "pam_rbf_evaluate", 1)

    def evaluate_gradient(self, diff):
        """A^T grad phi(A d) for each row -> (n, 3)."""
        return self._point(diff, "pam_rbf_evaluate_gradient", 3)

    def evaluate_hessian(self, diff):
        """A^T H(A d) A for each row -> (n, 3, 3)."""
        return self._point(diff, "pam_rbf_evaluate_hessian", 9).reshape(-1, 3, 3)


class _PolyharmonicOdd(Rbf):
    def set_parameters(self, params):
        # polyharmonic_odd.hpp:87-99: [] -> [1, 0], [slope] -> [slope, 0]
        if len(params) == 0:
            params = [1.0, 0.0]
        elif len(params) == 1:
            params = [params[0], 0.0]
        super().set_parameters(params)


class Biharmonic3D(_PolyharmonicOdd):
    kind = _capi.PAM_RBF_BH3
    short_name = "bh3"
    cpd_order = 1


class Triharmonic3D(_PolyharmonicOdd):
    kind = _capi.PAM_RBF_TH3
    short_name = "th3"
    cpd_order = 2


class CovExponential(Rbf):
    kind = _capi.PAM_RBF_EXP
    short_name = "exp"
    cpd_order = 0


def from_c(r):
    """The descriptor of a pam_rbf (kind, parameters, anisotropy)."""
    table = {_capi.PAM_RBF_BH3: Biharmonic3D, _capi.PAM_RBF_TH3: Triharmonic3D,
             _capi.PAM_RBF_EXP: CovExponential}
    rbf = table[r.kind]([r.params[0], r.params[1]])
    rbf.set_anisotropy(np.array([r.aniso[i] for i in range(9)]).reshape(3, 3))
    return rbf


def make_rbf(name, params):
    """make_rbf (rbf/make_rbf.hpp:23-50) for the kernels on this path."""
    table = {"bh3": Biharmonic3D, "th3": Triharmonic3D, "exp": CovExponential}
    if name not in table:
        raise ValueError(f"unknown RBF name: {name}")
    return table[name](params)
