"""CPU checks of the C ABI: the library loads (no GPU needed) and exports every
entry point declared in include/polatory_amd.h; argument errors are reported
as status codes with messages, never as crashes."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "polatory_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(pam_\w+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    from polatory_amd import _capi
    lib = _capi.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding covers the whole header
    assert set(syms) == set(_capi.exported_symbols())


def test_invalid_rbf_is_rejected_without_gpu():
    from polatory_amd import _capi
    r = _capi.PamRbf()
    r.kind = 7
    h = ctypes.c_void_p()
    rc = _capi.load().pam_fmm_evaluator_create(ctypes.byref(r), None, None, ctypes.byref(h))
    assert rc == 1
    assert "unknown rbf kind" in _capi.last_error()


def test_anisotropy_must_have_positive_determinant():
    import polatory_amd as pa
    r = pa.Biharmonic3D()
    with pytest.raises(ValueError):
        r.set_anisotropy(np.diag([1.0, -1.0, 1.0]))
    from polatory_amd import _capi
    c = r.to_c()
    c.aniso[0] = -1.0
    h = ctypes.c_void_p()
    rc = _capi.load().pam_fmm_evaluator_create(ctypes.byref(c), None, None, ctypes.byref(h))
    assert rc == 1 and "determinant" in _capi.last_error()


def test_rbf_parameter_defaults_follow_reference():
    import polatory_amd as pa
    assert pa.Biharmonic3D().parameters() == [1.0, 0.0]
    assert pa.Triharmonic3D([2.0]).parameters() == [2.0, 0.0]
    with pytest.raises(ValueError):
        pa.CovExponential([1.0])
    assert pa.make_rbf("exp", [1.0, 2.0]).short_name == "exp"


def test_cpp_api_mirror_links():
    """The C++ mirror compiles and links against the C ABI (runs only on a GPU)."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_api")
    assert os.path.exists(exe), "tests/cpp/test_api is built by polatory_amd/build.py"


def test_model_validation_follows_reference():
    import polatory_amd as pa
    # model.hpp:39-50: poly degree within cpd_order - 1 .. 2
    assert pa.Model(pa.Biharmonic3D()).poly_degree() == 0
    assert pa.Model(pa.Triharmonic3D()).poly_degree() == 1
    assert pa.Model(pa.CovExponential([1.0, 1.0])).poly_degree() == -1
    assert pa.Model(pa.Biharmonic3D(), 1).poly_basis_size() == 4
    with pytest.raises(ValueError, match="poly_degree"):
        pa.Model(pa.Triharmonic3D(), 0)
    with pytest.raises(ValueError, match="poly_degree"):
        pa.Model(pa.Biharmonic3D(), 3)
    with pytest.raises(ValueError, match="nugget"):
        pa.Model(pa.Biharmonic3D(), 1, nugget=-1.0)
