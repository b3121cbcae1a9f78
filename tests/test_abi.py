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
