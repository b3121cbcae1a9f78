"""k-fold cross-validation (kriging/cross_validate.hpp:14-47): each set's
predictions come from the fit on the other sets. Checked against the same
loop written over Interpolant.fit / evaluate (the reference's loop), and each
fold's predictions against the oracle's FP64 direct evaluation of that fold's
fitted weights."""
import numpy as np
import pytest

import polatory_amd as pa

pytestmark = pytest.mark.gpu


def test_cross_validate_matches_per_fold_fits(orc):
    rng = np.random.default_rng(41)
    pts = rng.uniform(0, 1, (3000, 3))
    vals = np.sin(3 * pts[:, 0]) + pts[:, 1] * pts[:, 2]
    ids = rng.integers(0, 5, len(pts)).astype(np.int32)
    model = pa.Model(pa.CovExponential([1.0, 0.5]), poly_degree=0, nugget=0.0)
    tol = 1e-6
    pred = pa.cross_validate(model, pts, vals, ids, tol, accuracy=1e-8)
    orbf = [orc.make_rbf(orc.EXP, [1.0, 0.5], np.eye(3))]
    for k in range(5):
        test = ids == k
        ip = pa.Interpolant(model)
        ip.fit(pts[~test], vals[~test], tol, accuracy=1e-8)
        manual = ip.evaluate(pts[test], accuracy=1e-8)
        assert np.array_equal(pred[test], manual), f"fold {k}"
        ref = orc.direct_evaluate(orbf, 0, ip.centers(), ip.weights(), pts[test])
        assert np.abs(pred[test] - ref).max() <= 1e-8 * np.abs(ref).max() + 1e-12
    # held-out predictions of a smooth function are close to it (typically;
    # points near the box faces are extrapolated)
    assert np.median(np.abs(pred - vals)) < 0.01


def test_cross_validate_argument_errors():
    model = pa.Model(pa.Biharmonic3D(), poly_degree=1)
    with pytest.raises(ValueError):
        pa.cross_validate(model, np.zeros((4, 3)), np.zeros(3), np.zeros(4), 1e-4)
