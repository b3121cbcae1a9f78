"""Incremental fitting (interpolation/incremental_fitter.hpp:25-183,
Interpolant::fit_incrementally, interpolant.hpp:115-147) and its distance
filter (point_cloud/distance_filter.hpp:16-99).

The filter is pinned against a brute-force greedy walk on small inputs. The
fitter is checked as the reference's own test checks it (every point within
the tolerance afterwards, test/interpolation/test_incremental_fitter.cpp), and
its first rounds, which pick centres before any solve (the fit is zero until
the polynomial is determined), are replayed with the brute-force filter.
"""
import numpy as np
import pytest

import polatory_amd as pa

A = np.array([[1.2, 0.1, -0.2], [0.05, 0.9, 0.1], [0.0, -0.15, 1.4]])


def brute_filter(pts, distance, indices):
    removed = set()
    for i in indices:
        if i in removed:
            continue
        d2 = ((pts[indices] - pts[i]) ** 2).sum(axis=1)
        for j, dd in zip(indices, d2):
            if j != i and dd <= distance * distance:
                removed.add(j)
    return [i for i in indices if i not in removed]


@pytest.mark.parametrize("distance", [1e-3, 0.05, 0.2, 0.7, 10.0])
def test_distance_filter_matches_brute_force(distance):
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 1, (600, 3))
    pts[100] = pts[7]  # a duplicate
    order = rng.permutation(600)
    got = pa.distance_filter(pts, distance, order)
    assert list(got) == brute_filter(pts, distance, list(order))


def test_distance_filter_subset_and_errors():
    rng = np.random.default_rng(5)
    pts = rng.uniform(-3, 2, (300, 3))
    idx = rng.permutation(300)[:120]
    assert list(pa.distance_filter(pts, 0.4, idx)) == brute_filter(pts, 0.4, list(idx))
    assert list(pa.distance_filter(pts, 1.1e-19, idx)) == list(idx)
    with pytest.raises(ValueError):
        pa.distance_filter(pts, 0.0, idx)


def sample(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0, 1, (n, 3))
    v = np.sin(4 * p[:, 0]) * np.cos(3 * p[:, 1]) + p[:, 2] ** 2 + 0.3
    return p, v


@pytest.mark.gpu
def test_incremental_fit_meets_tolerance_everywhere():
    pts, vals = sample(10_000, 31)
    tol = 1e-2
    rbf = pa.Triharmonic3D([1.0])
    rbf.set_anisotropy(A)
    ip = pa.Interpolant(pa.Model(rbf, poly_degree=1))
    ip.fit_incrementally(pts, vals, tol, max_iter=100, accuracy=tol / 100)
    c = ip.centers()
    n = len(c)
    st = ip.fit_stats()
    print(f"incremental: {n} of {len(pts)} points as centres in {st['rounds']} rounds, "
          f"{st['iterations']} FGMRES iterations", flush=True)
    assert ip.weights().shape == (n + 4,)
    assert n < len(pts)
    # centres are distinct input points
    rows = {tuple(r) for r in pts}
    assert all(tuple(r) in rows for r in c) and len({tuple(r) for r in c}) == n
    err = np.abs(ip.evaluate(pts, accuracy=tol / 100) - vals).max()
    assert err < tol

    # The rounds before the polynomial is determined (fewer than 4 centres)
    # see the zero function: residual |v|, thinned at the bbox diagonal, then
    # half of it, and so on.
    dist = np.linalg.norm(pts.max(axis=0) - pts.min(axis=0))
    centers = []
    while len(centers) < 4:
        rest = [i for i in range(len(pts)) if i not in set(centers)]
        r = np.abs(vals[rest])
        order = [rest[k] for k in np.argsort(r, kind="stable")]
        need = [i for i in reversed(order) if abs(vals[i]) >= tol]
        kept = set(brute_filter(pts, dist, centers + list(reversed(order))))
        centers += [i for i in need if i in kept]
        dist *= 0.5
    assert np.array_equal(c[:len(centers)], pts[centers])


@pytest.mark.gpu
def test_incremental_fit_without_polynomial():
    # An exponential model: kernel without polynomial (degree -1 allowed,
    # cpd order 0); the first centre is the point of largest |value|.
    pts, vals = sample(4_000, 32)
    ip = pa.Interpolant(pa.Model(pa.CovExponential([1.0, 0.5]), poly_degree=-1))
    ip.fit_incrementally(pts, vals, 1e-3, accuracy=1e-5)
    c = ip.centers()
    assert np.array_equal(c[0], pts[np.argmax(np.abs(vals))])
    assert np.abs(ip.evaluate(pts, accuracy=1e-5) - vals).max() < 1e-3
