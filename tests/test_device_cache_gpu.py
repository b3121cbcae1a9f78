"""The device buffer cache: buffers the library releases are kept and handed
out again, results do not depend on whether a buffer is fresh or reused, and
the cache empties on request. Calls go through the C ABI."""
import sys

import numpy as np
import pytest

import polatory_amd as pa

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def evaluate(src, trg, w):
    ev = pa.make_fmm_evaluator(pa.Biharmonic3D(), ((0, 0, 0), (1, 1, 1)))
    ev.set_config(pa.FmmConfig("mixed", 6, 3, "fmm"))
    ev.set_source_points(src)
    ev.set_weights(w)
    ev.set_target_points(trg)
    return ev.evaluate()


def fit_residual(pts, vals):
    ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
    ip.fit(pts, vals, 1e-4, max_iter=50)
    w = ip.weights()
    n = len(pts)
    y = pa.direct_evaluate(pa.Biharmonic3D(), pts, w[:n], pts) + w[n] + pts @ w[n + 1:n + 4]
    return float(np.abs(vals - y).max()), ip.fit_stats()["iterations"]


def test_reused_buffers_give_the_same_results():
    rng = np.random.default_rng(31)
    src = rng.uniform(0, 1, (30000, 3))
    trg = rng.uniform(0, 1, (20000, 3))
    w = rng.normal(size=30000)
    pts = rng.uniform(0, 1, (6000, 3))
    vals = np.sin(4 * pts[:, 0]) + pts[:, 1] * pts[:, 2]
    pa.device_cache_release()
    assert pa.device_cache_bytes() == 0
    log("fresh buffers")
    y0 = evaluate(src, trg, w)
    r0, it0 = fit_residual(pts, vals)
    held = pa.device_cache_bytes()
    log(f"  cached after release: {held} bytes; fit residual {r0:.2e} in {it0} iterations")
    assert held > 0
    log("reused buffers")
    y1 = evaluate(src, trg, w)
    r1, it1 = fit_residual(pts, vals)
    log(f"  fit residual {r1:.2e} in {it1} iterations")
    assert np.array_equal(y0, y1)
    assert r0 <= 1e-4 and r1 <= 1e-4
    pa.device_cache_release()
    assert pa.device_cache_bytes() == 0
    log("after release")
    assert np.array_equal(y0, evaluate(src, trg, w))
