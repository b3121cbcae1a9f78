"""The allocation-free FP64 direct sum (pam_direct_evaluate_device_noalloc)
against pam_direct_evaluate, through the C ABI: the same bits on two calls in
a row (the second smaller, so stale contents of the shared work area would
show), the empty cases, and sizes beyond the fixed work area refused."""
import ctypes
import sys

import numpy as np
import pytest

import polatory_amd as pa
from polatory_amd import _capi

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = (64 << 20) // 8


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(torch, rbf, src, w, trg, stream):
    dev = torch.device("cuda:0")
    d_src = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
    d_w = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    d_trg = torch.from_numpy(np.ascontiguousarray(trg)).to(dev)
    d_out = torch.full((max(len(trg), 1),), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    c = rbf.to_c()
    P = ctypes.c_void_p
    _capi.call("pam_direct_evaluate_device_noalloc", ctypes.byref(c), P(d_src.data_ptr()), len(src),
               P(d_w.data_ptr()), P(d_trg.data_ptr()), len(trg), P(d_out.data_ptr()), P(stream.cuda_stream))
    stream.synchronize()
    return d_out.cpu().numpy()[:len(trg)]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def test_noalloc_matches_direct_evaluate_twice(torch, orc):
    rng = np.random.default_rng(71)
    stream = torch.cuda.Stream()
    cases = [(pa.Biharmonic3D(), 20000, 30000), (pa.CovExponential([1.0, 0.5]), 5000, 7000)]
    cases[0][0].set_anisotropy(np.array([[1.0, 0.2, 0.0], [0.0, 0.9, 0.1], [0.0, 0.0, 1.2]]))
    for rbf, ns, nt in cases:
        log("noalloc", rbf.short_name, ns, "->", nt)
        src = rng.uniform(0, 1, (ns, 3))
        trg = rng.uniform(-0.1, 1.1, (nt, 3))
        w = rng.normal(size=ns)
        got = noalloc(torch, rbf, src, w, trg, stream)
        ref = pa.direct_evaluate(rbf, src, w, trg)
        np.testing.assert_array_equal(got, ref)  # same kernels, same order: same bits
        kind = {"bh3": orc.BH3, "exp": orc.EXP}[rbf.short_name]
        o = orc.direct_rbf_sum(orc.make_rbf(kind, rbf.parameters(), rbf.anisotropy()), src, w, trg)
        assert np.abs(got - o).max() / np.abs(o).max() < 1e-13


def test_noalloc_empty_inputs(torch):
    stream = torch.cuda.current_stream()
    rbf = pa.Triharmonic3D()
    trg = np.random.default_rng(72).uniform(0, 1, (1000, 3))
    got = noalloc(torch, rbf, np.zeros((0, 3)), np.zeros(0), trg, stream)
    np.testing.assert_array_equal(got, np.zeros(1000))
    assert noalloc(torch, rbf, trg, np.ones(1000), np.zeros((0, 3)), stream).shape == (0,)


def test_noalloc_refuses_sizes_beyond_work_area(torch):
    dev = torch.device("cuda:0")
    buf = torch.zeros(16, dtype=torch.float64, device=dev)
    c = pa.Biharmonic3D().to_c()
    P = ctypes.c_void_p
    p = P(buf.data_ptr())
    n = SCRATCH_DOUBLES // 6 + 1  # 3 n + 3 n alone exceed the area; nothing is launched
    rc = _capi.load().pam_direct_evaluate_device_noalloc(ctypes.byref(c), p, n, p, p, n, p, None)
    assert rc == 1
    assert "work area" in _capi.last_error()
