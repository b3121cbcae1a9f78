"""Gradients through the FMM (gradient-transpose kernel,
fmm/gradient_transpose_kernel.hpp:47-62; make_fmm_gradient_transpose_evaluator,
make_fmm_evaluator.cpp:119-148) and the grad_points half of
Interpolant::evaluate (interpolant.hpp:58-67, interpolation/evaluator.hpp:68-74).

The FMM gradients are compared with the FP64 direct gradient sum on the GPU,
which is itself compared here with the oracle's direct gradient sum (pinned by
finite differences in test_rbf_gradient.py). The interpolant's gradients are
compared with the oracle's direct gradient sum plus central differences of the
oracle's polynomial values.
"""
import sys

import numpy as np
import pytest

import polatory_amd as pa

pytestmark = pytest.mark.gpu

A = np.array([[1.2, 0.1, -0.2], [0.05, 0.9, 0.1], [0.0, -0.15, 1.4]])  # det > 0, not symmetric
KINDS = {"bh3": lambda: pa.Biharmonic3D(), "th3": lambda: pa.Triharmonic3D(),
         "exp": lambda: pa.CovExponential([1.0, 0.5])}
# Relative L-inf error of the gradients at 2,000 targets, about 2x the worst
# measured over the kernels and anisotropies below (FP64 order 10: 1.0e-6;
# mixed order 7: 8.2e-5; both the exponential with A).
TOL = {("fp64", 10): 2e-6, ("mixed", 7): 2e-4}


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def orc_rbf(orc, rbf):
    kind = {"bh3": orc.BH3, "th3": orc.TH3, "exp": orc.EXP}[rbf.short_name]
    return orc.make_rbf(kind, rbf.parameters(), rbf.anisotropy())


@pytest.fixture(scope="module")
def cloud():
    rng = np.random.default_rng(11)
    src = rng.uniform(0, 1, (200_000, 3))
    trg = rng.uniform(0, 1, (50_000, 3))
    w = rng.normal(size=200_000)
    return src, trg, w


@pytest.mark.parametrize("name", ["bh3", "th3", "exp"])
@pytest.mark.parametrize("aniso", [False, True])
def test_fmm_gradient_matches_direct(orc, cloud, name, aniso):
    src, trg, w = cloud
    rbf = KINDS[name]()
    if aniso:
        rbf.set_anisotropy(A)
    ref = pa.direct_evaluate_gradient(rbf, src, w, trg[:2000])
    # the GPU direct gradient sum against the oracle's
    o = orc.direct_gradient_sum(orc_rbf(orc, rbf), src, w, trg[:100])
    assert np.abs(ref[:100] - o).max() <= 1e-12 * np.abs(o).max()
    for (prec, order), tol in TOL.items():
        ev = pa.FmmEvaluator(rbf)
        ev.set_config(pa.FmmConfig(prec, order, mode="fmm"))
        ev.set_source_points(src)
        ev.set_weights(w)
        ev.set_target_points(trg)
        g = ev.evaluate_gradient()
        assert ev.stats()["direct"] == 0, "the FMM path did not run"
        assert g.shape == (50_000, 3) and np.isfinite(g).all()
        err = np.abs(g[:2000] - ref).max() / np.abs(ref).max()
        log(f"{name} aniso={aniso} {prec}{order}: gradient rel err {err:.2e}")
        assert err <= tol
        # values from the same handle are unchanged by the gradient evaluation
        v = ev.evaluate()
        assert np.array_equal(v, ev.evaluate())


def test_gradient_transpose_evaluator_accuracy_search(cloud):
    src, trg, w = cloud
    rbf = pa.Biharmonic3D()
    rbf.set_anisotropy(A)
    ref = pa.direct_evaluate_gradient(rbf, src, w, trg[:2000])
    acc = 1e-6 * np.abs(ref).max()
    ev = pa.make_fmm_gradient_transpose_evaluator(rbf)
    ev.set_source_points(src)
    ev.set_weights(w)
    ev.set_accuracy(acc)
    ev.set_target_points(trg)
    g = ev.evaluate()
    cfg = ev.config()
    err = np.abs(g[:2000] - ref).max()
    log(f"gradient accuracy search: {cfg}, error {err:.2e} (accuracy {acc:.2e})")
    # the search measures on source points; other targets get a factor-2 margin
    assert err <= 2 * acc
    # a value evaluator at the same accuracy picks a configuration for values
    vev = pa.make_fmm_evaluator(rbf)
    vev.set_source_points(src)
    vev.set_weights(w)
    vev.set_accuracy(acc)
    vev.set_target_points(trg)
    vev.evaluate()
    log(f"value accuracy search at the same accuracy: {vev.config()}")


@pytest.mark.parametrize("rbf_name,deg", [("bh3", 1), ("bh3", 2), ("th3", 1), ("exp", 0)])
def test_interpolant_gradients_match_oracle(orc, rbf_name, deg):
    rng = np.random.default_rng(21)
    pts = rng.uniform(0, 1, (20_000, 3))
    vals = np.sin(3 * pts[:, 0]) * np.cos(2 * pts[:, 1]) + 0.5 * pts[:, 2] ** 2
    rbf = KINDS[rbf_name]()
    rbf.set_anisotropy(A)
    ip = pa.Interpolant(pa.Model(rbf, poly_degree=deg))
    ip.fit(pts, vals, 1e-4)
    q = rng.uniform(0.05, 0.95, (30_000, 3))
    gq = rng.uniform(0.05, 0.95, (30_000, 3))
    w = ip.weights()
    n = len(pts)
    orbf = [orc_rbf(orc, rbf)]
    sub = gq[:100]
    # rbf part: the oracle's direct gradient sum; polynomial part: central
    # differences of the oracle's polynomial values (exact for degree <= 2 up
    # to rounding), so that ill-conditioned weights do not enter a difference
    ref = orc.direct_gradient_sum(orbf[0], pts, w[:n], sub)
    wp = w.copy()
    wp[:n] = 0.0
    h = 1e-3
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        ref[:, k] += (orc.direct_evaluate(orbf, deg, pts, wp, sub + e)
                      - orc.direct_evaluate(orbf, deg, pts, wp, sub - e)) / (2 * h)
    gacc = 1e-5 * np.abs(ref).max()
    out = ip.evaluate(q, gq, accuracy=1e-6, grad_accuracy=gacc)
    assert out.shape == (len(q) + 3 * len(gq),)
    assert np.array_equal(out[:len(q)], ip.evaluate(q, accuracy=1e-6))
    g = out[len(q):].reshape(-1, 3)
    assert np.array_equal(g, ip.evaluate_gradient(gq, accuracy=gacc))
    err = np.abs(g[:len(sub)] - ref).max()
    log(f"interpolant {rbf_name} degree {deg}: gradient error {err:.2e} (grad accuracy {gacc:.2e})")
    # the accuracy search measures between the centres; other points get a factor-2 margin
    assert err <= 2 * gacc
    # the same through the FMM path (FP64, order 10), polynomial gradient included
    ip.set_eval_config(pa.FmmConfig("fp64", 10, mode="fmm"))
    g2 = ip.evaluate_gradient(gq)
    err2 = np.abs(g2[:len(sub)] - ref).max() / np.abs(ref).max()
    log(f"  through the FMM (fp64 order 10): gradient rel err {err2:.2e}")
    assert err2 <= 1e-5


@pytest.mark.parametrize("name", ["bh3", "exp"])
def test_gradient_kernel_matches_source_differences(orc, name):
    """Gradient-weighted sources (fmm/gradient_kernel.hpp:46-62): the value at a
    target is sum_j w_j . d/ds phi(A (t - s)) at s = s_j, compared with central
    differences of the oracle's direct value sums over shifted sources (no
    gradient formula enters the reference)."""
    rng = np.random.default_rng(51)
    src = rng.uniform(0, 1, (20_000, 3))
    trg = rng.uniform(0, 1, (5_000, 3))
    w3 = rng.normal(size=(20_000, 3))
    rbf = KINDS[name]()
    rbf.set_anisotropy(A)
    orbf = orc_rbf(orc, rbf)
    sub = trg[:300]
    h = 1e-5
    ref = np.zeros(len(sub))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        ref += (orc.direct_rbf_sum(orbf, src + e, w3[:, k], sub)
                - orc.direct_rbf_sum(orbf, src - e, w3[:, k], sub)) / (2 * h)
    ev = pa.make_fmm_gradient_evaluator(rbf)
    ev.set_config(pa.FmmConfig("fp64", 10, mode="fmm"))
    ev.set_source_points(src)
    ev.set_weights(w3)
    ev.set_target_points(trg)
    v = ev.evaluate()
    assert v.shape == (5_000,) and ev.stats()["direct"] == 0
    err = np.abs(v[:300] - ref).max() / np.abs(ref).max()
    log(f"gradient kernel {name}: rel err {err:.2e} vs source differences")
    assert err <= 5e-6
    with pytest.raises(RuntimeError):
        ev.evaluate_gradient()
    with pytest.raises(ValueError):
        ev.set_weights(w3[:, 0])


@pytest.mark.parametrize("name", ["bh3", "exp"])
def test_hessian_kernel_matches_source_differences(orc, name):
    """Hessian kernel (fmm/hessian_kernel.hpp:46-64): out_i = -sum_j A^T H A w_j,
    compared with central differences of the oracle's direct gradient sums
    over shifted sources: d/ds_k of A^T grad phi(A (t - s)) is -A^T H A e_k."""
    rng = np.random.default_rng(52)
    src = rng.uniform(0, 1, (20_000, 3))
    trg = rng.uniform(0, 1, (5_000, 3))
    w3 = rng.normal(size=(20_000, 3))
    rbf = KINDS[name]()
    rbf.set_anisotropy(A)
    orbf = orc_rbf(orc, rbf)
    sub = trg[:300]

    def diff(h):
        d = np.zeros((len(sub), 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            d += (orc.direct_gradient_sum(orbf, src + e, w3[:, k], sub)
                  - orc.direct_gradient_sum(orbf, src - e, w3[:, k], sub)) / (2 * h)
        return d

    # Richardson extrapolation of the central differences (the h^2 term of
    # the closest source-target pairs is otherwise ~1e-5 relative)
    ref = (4.0 * diff(1e-5) - diff(2e-5)) / 3.0
    ev = pa.make_fmm_hessian_evaluator(rbf)
    ev.set_config(pa.FmmConfig("fp64", 10, mode="fmm"))
    ev.set_source_points(src)
    ev.set_weights(w3)
    ev.set_target_points(trg)
    v = ev.evaluate()
    assert v.shape == (5_000, 3) and ev.stats()["direct"] == 0
    err = np.abs(v[:300] - ref).max() / np.abs(ref).max()
    log(f"hessian kernel {name}: rel err {err:.2e} vs source differences")
    assert err <= 5e-6  # measured 1.9e-6 (bh3), 1.3e-6 (exp)
    # the direct path gives the same within the FMM error
    ev.set_config(pa.FmmConfig("fp64", 10, mode="direct"))
    vd = ev.evaluate()
    errd = np.abs(vd[:300] - ref).max() / np.abs(ref).max()
    log(f"hessian kernel {name}, direct: rel err {errd:.2e}")
    assert errd <= 1e-7  # measured 1.1e-9, 1.4e-9


@pytest.mark.parametrize("name", ["bh3", "th3"])
def test_hessian_symmetric_evaluator(orc, name):
    """make_fmm_hessian_symmetric_evaluator (make_fmm_evaluator.cpp:241-270):
    the Hessian kernel with the points as sources and targets. It returns the
    bits of the source-target Hessian evaluator given the same points twice,
    and on a subsample matches central differences of the oracle's direct
    gradient sums over the other points (the zero-distance term of bh3 and th3
    is zero)."""
    rng = np.random.default_rng(53)
    pts = rng.uniform(0, 1, (20_000, 3))
    w3 = rng.normal(size=(20_000, 3))
    rbf = KINDS[name]()
    rbf.set_anisotropy(A)
    cfg = pa.FmmConfig("fp64", 10, mode="fmm")
    sym = pa.make_fmm_hessian_symmetric_evaluator(rbf)
    sym.set_config(cfg)
    sym.set_points(pts)
    sym.set_weights(w3)
    v = sym.evaluate()
    assert v.shape == (20_000, 3) and sym.stats()["direct"] == 0
    st = pa.make_fmm_hessian_evaluator(rbf)
    st.set_config(cfg)
    st.set_source_points(pts)
    st.set_weights(w3)
    st.set_target_points(pts)
    assert np.array_equal(st.evaluate(), v)
    with pytest.raises(TypeError):
        sym.set_target_points(pts)

    orbf = orc_rbf(orc, rbf)
    idx = rng.choice(20_000, 40, replace=False)

    def diff(h):
        d = np.zeros((len(idx), 3))
        for r, i in enumerate(idx):
            others = np.delete(np.arange(20_000), i)
            for k in range(3):
                e = np.zeros(3)
                e[k] = h
                d[r] += (orc.direct_gradient_sum(orbf, pts[others] + e, w3[others, k], pts[i:i + 1])
                         - orc.direct_gradient_sum(orbf, pts[others] - e, w3[others, k], pts[i:i + 1]))[0] / (2 * h)
        return d

    ref = (4.0 * diff(1e-5) - diff(2e-5)) / 3.0
    err = np.abs(v[idx] - ref).max() / np.abs(ref).max()
    log(f"hessian symmetric {name}: rel err {err:.2e} vs source differences")
    assert err <= 5e-6
    sym.set_config(pa.FmmConfig("fp64", 10, mode="direct"))
    vd = sym.evaluate()
    errd = np.abs(vd[idx] - ref).max() / np.abs(ref).max()
    log(f"hessian symmetric {name}, direct: rel err {errd:.2e}")
    assert errd <= 1e-7
