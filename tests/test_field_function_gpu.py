"""The field an isosurface extractor samples: isosurface::RbfFieldFunction
(isosurface/rbf_field_function.hpp:12-36) over Interpolant's
set_evaluation_bbox_impl and evaluate_impl (interpolant.hpp:69-75,201-210).
The lattice reaches past the centres' box, as an isosurface's bbox does, and
every value is checked against the oracle's FP64 direct evaluation of the
fitted weights."""
import numpy as np
import pytest

import polatory_amd as pa

pytestmark = pytest.mark.gpu


def lattice(lo, hi, g):
    ax = [np.linspace(lo[k], hi[k], g) for k in range(3)]
    x, y, z = np.meshgrid(*ax, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def test_rbf_field_function_on_a_lattice(orc):
    rng = np.random.default_rng(51)
    pts = rng.uniform(0, 1, (3000, 3))
    vals = np.sin(3 * pts[:, 0]) * np.cos(2 * pts[:, 1]) + 0.5 * pts[:, 2] ** 2
    ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
    ip.fit(pts, vals, 1e-6, accuracy=1e-8)
    lo, hi = np.full(3, -0.3), np.full(3, 1.3)
    field = pa.RbfFieldFunction(ip, accuracy=1e-8)
    field.set_evaluation_bbox(lo, hi)
    q = lattice(lo, hi, 20)
    f = field(q)
    assert np.array_equal(field(q), f)  # the evaluator is reused
    orbf = [orc.make_rbf(orc.BH3, [1.0], np.eye(3))]
    ref = orc.direct_evaluate(orbf, 1, ip.centers(), ip.weights(), q)
    err = np.abs(f - ref).max()
    print(f"field on a 20^3 lattice: max error {err:.3e} (scale {np.abs(ref).max():.3e})", flush=True)
    assert err <= 1e-8 * np.abs(ref).max() + 1e-12
    # a sub-box evaluated in pieces gives the values of the whole lattice
    part = field(q[:4000])
    assert np.abs(part - f[:4000]).max() <= 1e-8 * np.abs(ref).max() + 1e-12


def test_evaluate_impl_reuses_the_last_evaluator():
    rng = np.random.default_rng(52)
    pts = rng.uniform(-1, 1, (2000, 3))
    vals = pts[:, 0] ** 2 - pts[:, 1] + np.cos(pts[:, 2])
    ip = pa.Interpolant(pa.Model(pa.Triharmonic3D(), poly_degree=1))
    ip.fit(pts, vals, 1e-6, accuracy=1e-8)
    q = rng.uniform(-1, 1, (5000, 3))
    a = ip.evaluate(q, accuracy=1e-6)
    assert np.array_equal(ip.evaluate_impl(q), a)
    with pytest.raises(ValueError):
        ip.set_evaluation_bbox([0, 0, 0], [-1, 1, 1])
    with pytest.raises(ValueError):
        ip.set_evaluation_bbox([0, 0, 0], [1, 1, 1], accuracy=0.0)
    # a refit clears the box; the field of the new fit is evaluated afresh
    ip.set_evaluation_bbox([-2, -2, -2], [2, 2, 2], 1e-8)
    ip.fit(pts, -vals, 1e-6, accuracy=1e-8)
    b = ip.evaluate(q, accuracy=1e-6)
    assert np.abs(b + a).max() < 1e-4


def test_rbf_field_function_through_the_fmm(orc):
    """The same field with the evaluator forced onto the FMM (FP64, order 10)
    over the box enlarged past the centres: within 1e-8 of the oracle, and not
    the direct sum's bits (the FMM path ran)."""
    rng = np.random.default_rng(52)
    pts = rng.uniform(0, 1, (3000, 3))
    vals = np.sin(3 * pts[:, 0]) * np.cos(2 * pts[:, 1]) + 0.5 * pts[:, 2] ** 2
    ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
    ip.fit(pts, vals, 1e-6)
    lo, hi = np.full(3, -0.3), np.full(3, 1.3)
    q = lattice(lo, hi, 30)
    orbf = [orc.make_rbf(orc.BH3, [1.0], np.eye(3))]
    ref = orc.direct_evaluate(orbf, 1, ip.centers(), ip.weights(), q)
    ip.set_eval_config(pa.FmmConfig("fp64", 10, -1, "direct"))
    field = pa.RbfFieldFunction(ip)
    field.set_evaluation_bbox(lo, hi)
    direct = field(q)
    ip.set_eval_config(pa.FmmConfig("fp64", 10, -1, "fmm"))
    field = pa.RbfFieldFunction(ip)
    field.set_evaluation_bbox(lo, hi)
    f = field(q)
    err = np.abs(f - ref).max() / np.abs(ref).max()
    print(f"field through the FMM on a 30^3 lattice: rel error {err:.3e}", flush=True)
    assert err <= 1e-8
    assert np.abs(direct - ref).max() <= 1e-12 * np.abs(ref).max()
    assert not np.array_equal(f, direct)
