"""Gradient and Hessian of the kernels (rbf/polyharmonic_odd.hpp:45-72,
rbf/cov_exponential.hpp:42-62, rbf/rbf_base.hpp:45-53) and the direct gradient
sum. The oracle's restatement is pinned by central finite differences of its
own values (which the closed-form golden vectors pin), by the isotropic closed
forms and by the reference's special cases; the GPU kernels are then compared
with the oracle."""
import numpy as np
import pytest

A = np.array([[1.2, 0.1, -0.2], [0.05, 0.9, 0.1], [0.0, -0.15, 1.4]])  # det > 0, not symmetric

CASES = [("bh3", 0, [1.0, 0.0]), ("bh3c", 0, [2.5, 0.3]), ("th3", 1, [1.0, 0.0]),
         ("th3c", 1, [0.7, 0.2]), ("exp", 2, [1.0, 2.0]), ("exp2", 2, [0.4, 0.6])]


def _orc_rbf(orc, kind, params, aniso):
    return orc.make_rbf([orc.BH3, orc.TH3, orc.EXP][kind], params, aniso)


def _diffs(n, seed=0):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return d * np.random.default_rng(seed + 1).uniform(0.2, 3.0, size=(n, 1))


@pytest.mark.parametrize("name,kind,params", CASES)
@pytest.mark.parametrize("aniso", [None, "A"])
def test_oracle_gradient_matches_finite_differences(orc, name, kind, params, aniso):
    r = _orc_rbf(orc, kind, params, None if aniso is None else A)
    for d in _diffs(20):
        g = orc.rbf_gradient(r, d)
        fd = np.empty(3)
        h = 1e-6 * max(1.0, np.abs(d).max())
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd[k] = (orc.rbf_evaluate(r, d + e) - orc.rbf_evaluate(r, d - e)) / (2 * h)
        np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8 * np.abs(g).max())


@pytest.mark.parametrize("name,kind,params", CASES)
@pytest.mark.parametrize("aniso", [None, "A"])
def test_oracle_hessian_matches_finite_differences_and_is_symmetric(orc, name, kind, params, aniso):
    r = _orc_rbf(orc, kind, params, None if aniso is None else A)
    for d in _diffs(20, seed=5):
        H = orc.rbf_hessian(r, d)
        np.testing.assert_allclose(H, H.T, rtol=1e-12, atol=1e-14 * np.abs(H).max())
        fd = np.empty((3, 3))
        h = 1e-6 * max(1.0, np.abs(d).max())
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd[:, k] = (orc.rbf_gradient(r, d + e) - orc.rbf_gradient(r, d - e)) / (2 * h)
        np.testing.assert_allclose(H, fd, rtol=1e-5, atol=1e-7 * np.abs(H).max())


def test_oracle_isotropic_closed_forms(orc):
    d = np.array([3.0, 4.0, 12.0])  # |d| = 13
    # bh3: -d / |d|; th3: 3 |d| d; exp(psill, range): -3 psill e^(-3 r / range) / (range r) d
    np.testing.assert_allclose(orc.rbf_gradient(_orc_rbf(orc, 0, [1.0, 0.0], None), d), -d / 13.0, rtol=1e-15)
    np.testing.assert_allclose(orc.rbf_gradient(_orc_rbf(orc, 1, [1.0, 0.0], None), d), 39.0 * d, rtol=1e-15)
    f = -3.0 * 0.5 * np.exp(-3.0 * 13.0 / 2.0) / (2.0 * 13.0)
    np.testing.assert_allclose(orc.rbf_gradient(_orc_rbf(orc, 2, [0.5, 2.0], None), d), f * d, rtol=1e-14)
    # bh3 Hessian: -(I - u u^T / r^2) / r (the gradient of -d/|d|)
    H = orc.rbf_hessian(_orc_rbf(orc, 0, [1.0, 0.0], None), d)
    np.testing.assert_allclose(H, -(np.eye(3) - np.outer(d, d) / 169.0) / 13.0, rtol=1e-14, atol=1e-17)


def test_oracle_special_cases_follow_the_reference(orc):
    z = np.zeros(3)
    # polyharmonic: zero gradient and Hessian at rho = 0 (polyharmonic_odd.hpp:50-52, 63-65)
    for kind in (0, 1):
        r = _orc_rbf(orc, kind, [1.0, 0.0], A)
        assert np.all(orc.rbf_gradient(r, z) == 0.0)
        assert np.all(orc.rbf_hessian(r, z) == 0.0)
    # with c > 0, rho > 0 at d = 0: the gradient is zero, the Hessian is not
    r = _orc_rbf(orc, 1, [1.0, 0.5], None)
    assert np.all(orc.rbf_gradient(r, z) == 0.0)
    np.testing.assert_allclose(orc.rbf_hessian(r, z), 1.5 * np.eye(3), rtol=1e-15)
    # exponential: no guard at r = 0 (cov_exponential.hpp:42-50) -> NaN
    assert np.all(np.isnan(orc.rbf_gradient(_orc_rbf(orc, 2, [1.0, 2.0], None), z)))


def test_oracle_direct_gradient_sum_matches_finite_differences(orc):
    rng = np.random.default_rng(3)
    src = rng.uniform(0, 1, (200, 3))
    w = rng.normal(size=200)
    trg = rng.uniform(0, 1, (10, 3)) + 1.5
    for kind, params in ((0, [1.0, 0.0]), (1, [1.0, 0.1]), (2, [1.0, 0.7])):
        r = _orc_rbf(orc, kind, params, A)
        g = orc.direct_gradient_sum(r, src, w, trg)
        h = 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (orc.direct_rbf_sum(r, src, w, trg + e) - orc.direct_rbf_sum(r, src, w, trg - e)) / (2 * h)
            np.testing.assert_allclose(g[:, k], fd, rtol=1e-6, atol=1e-7 * np.abs(g).max())


# ---- GPU kernels against the oracle (through the C ABI) ----

def _pa_rbf(pa, kind, params, aniso):
    r = [pa.Biharmonic3D, pa.Triharmonic3D, pa.CovExponential][kind](params)
    if aniso is not None:
        r.set_anisotropy(aniso)
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("name,kind,params", CASES)
@pytest.mark.parametrize("aniso", [None, "A"])
def test_gpu_pointwise_value_gradient_hessian_match_oracle(orc, name, kind, params, aniso):
    import polatory_amd as pa
    a = None if aniso is None else A
    r = _pa_rbf(pa, kind, params, a)
    o = _orc_rbf(orc, kind, params, a)
    d = _diffs(300, seed=11)
    v = r.evaluate(d)
    g = r.evaluate_gradient(d)
    H = r.evaluate_hessian(d)
    v_ref = np.array([orc.rbf_evaluate(o, x) for x in d])
    g_ref = np.array([orc.rbf_gradient(o, x) for x in d])
    H_ref = np.array([orc.rbf_hessian(o, x) for x in d])
    # FP64 on both sides; the device may contract multiply-adds into FMAs
    np.testing.assert_allclose(v, v_ref, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(g, g_ref, rtol=1e-12, atol=1e-13 * np.abs(g_ref).max())
    np.testing.assert_allclose(H, H_ref, rtol=1e-11, atol=1e-12 * np.abs(H_ref).max())


@pytest.mark.gpu
def test_gpu_pointwise_special_cases_and_empty_input():
    import polatory_amd as pa
    z = np.zeros((1, 3))
    assert np.all(pa.Biharmonic3D().evaluate_gradient(z) == 0.0)
    assert np.all(np.isnan(pa.CovExponential([1.0, 2.0]).evaluate_gradient(z)))
    assert pa.Triharmonic3D().evaluate(np.zeros((0, 3))).shape == (0,)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,params", [(0, [1.0, 0.0]), (1, [0.5, 0.1]), (2, [1.0, 0.7])])
def test_gpu_direct_gradient_matches_oracle(orc, kind, params):
    import polatory_amd as pa
    rng = np.random.default_rng(21)
    src = rng.uniform(0, 1, (3000, 3))
    w = rng.normal(size=3000)
    trg = rng.uniform(-0.2, 1.2, (700, 3))
    r = _pa_rbf(pa, kind, params, A)
    g = pa.direct_evaluate_gradient(r, src, w, trg)
    g_ref = orc.direct_gradient_sum(_orc_rbf(orc, kind, params, A), src, w, trg)
    np.testing.assert_allclose(g, g_ref, rtol=1e-10, atol=1e-11 * np.abs(g_ref).max())
