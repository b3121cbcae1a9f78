"""Incremental fitting (interpolation/incremental_fitter.hpp:25-183,
Interpolant::fit_incrementally, interpolant.hpp:115-147), inequality fitting
(interpolation/inequality_fitter.hpp:25-187, interpolant.hpp:149-190) and the
distance filter both use (point_cloud/distance_filter.hpp:16-99).

The filter is pinned against a brute-force greedy walk on small inputs. The
fitter is checked as the reference's own test checks it (every point within
the tolerance afterwards, test/interpolation/test_incremental_fitter.cpp), and
its first rounds, which pick centres before any solve (the fit is zero until
the polynomial is determined), are replayed with the brute-force filter. The
inequality fitter is checked on the reference test's two cases: the Kostov
(1986) problem with its data, and bounds only around a smooth function.
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
def test_incremental_fit_with_gradient_data(orc):
    # test/interpolation/test_incremental_fitter.cpp: 10,000 points and 10,000
    # gradient points, triharmonic3d with anisotropy, degree 1, tolerance 1e-2
    # on the values and 1e-1 on the gradients; the fit from the chosen centres
    # is checked at every point and gradient point by the oracle's direct
    # operator (the targets enter as centres of weight zero).
    rng = np.random.default_rng(41)
    pts = rng.uniform(-1, 1, (10_000, 3))
    gpts = rng.uniform(-1, 1, (10_000, 3))
    vals = np.sin(np.pi * pts @ A.T).sum(axis=1)
    grads = (np.pi * np.cos(np.pi * gpts @ A.T)) @ A
    data = np.concatenate([vals, grads.ravel()])
    tol, gtol = 1e-2, 1e-1
    rbf = pa.Triharmonic3D([1.0])
    rbf.set_anisotropy(A)
    ip = pa.Interpolant(pa.Model(rbf, poly_degree=1))
    ip.fit_incrementally_hermite(pts, gpts, data, tol, gtol, max_iter=100)
    c, g, w = ip.centers(), ip.grad_centers(), ip.weights()
    st = ip.fit_stats()
    print(f"incremental with gradients: {len(c)} points and {len(g)} gradient points as centres in "
          f"{st['rounds']} rounds, {st['iterations']} FGMRES iterations", flush=True)
    assert w.shape == (len(c) + 3 * len(g) + 4,)
    assert len(c) < len(pts) and len(g) < len(gpts)
    mu, sg = len(c), len(g)
    x = np.concatenate([w[:mu], np.zeros(len(pts)), w[mu:mu + 3 * sg], np.zeros(3 * len(gpts)), w[-4:]])
    r = orc.make_rbf(orc.TH3, [1.0, 0.0], A)
    y = orc.hermite_operator([r], 1, 0.0, np.vstack([c, pts]), np.vstack([g, gpts]), x)
    fv = y[mu:mu + len(pts)]
    fg = y[mu + len(pts) + 3 * sg:mu + len(pts) + 3 * sg + 3 * len(gpts)]
    assert np.abs(fv - vals).max() < tol
    assert np.abs(fg - grads.ravel()).max() < gtol
    # the interpolant's own evaluation agrees
    f = ip.evaluate_hermite(pts[:2000], gpts[:2000])
    assert np.abs(f[:2000] - fv[:2000]).max() <= 1e-6
    assert np.abs(f[2000:] - fg[:6000]).max() <= 1e-5


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


def kostov86():
    """The 1-D inequality problem of the reference's test
    (test/interpolation/test_inequality_fitter.cpp:60-109, after Kostov 1986),
    placed on the x axis of 3-D space (the exponential kernel sees distances only)."""
    nan = np.nan
    values = [1, nan, 4, nan, 3, nan, nan, nan, nan, nan, 3, nan, nan, 7, nan, 8, nan, nan, nan, 8, 3,
              nan, nan, 6, nan]
    lb = [nan, 6, nan, 2, nan, 2, 9, 4, 3, 3, nan, nan, 7, nan, 4, nan, nan, nan, 5, nan, nan, 9, 5, nan, nan]
    ub = [nan, nan, nan, 4, nan, 4, nan, nan, nan, nan, nan, 1, nan, nan, nan, nan, 4, 7, nan, nan, nan, nan,
          nan, nan, 3]
    pts = np.zeros((25, 3))
    pts[:, 0] = np.arange(25)
    return pts, np.array(values, float), np.array(lb, float), np.array(ub, float)


@pytest.mark.gpu
def test_inequality_fit_kostov86():
    pts, v, lb, ub = kostov86()
    tol = 1e-5
    ip = pa.Interpolant(pa.Model(pa.CovExponential([1.0, 3.0]), poly_degree=-1))
    ip.fit_inequality(pts, v, lb, ub, tol, max_iter=32, accuracy=tol / 100)
    fit = ip.evaluate(pts, accuracy=tol / 100)
    eq = ~np.isnan(v)
    assert np.abs(fit[eq] - v[eq]).max() < tol
    hl, hu = ~np.isnan(lb), ~np.isnan(ub)
    assert (fit[hl] > lb[hl] - tol).all() and (fit[hu] < ub[hu] + tol).all()
    c = ip.centers()
    # the equality points come first, then the active bounds in index order
    assert np.array_equal(c[:eq.sum()], pts[eq])
    act = c[eq.sum():, 0].astype(int)
    assert list(act) == sorted(act) and not eq[act].any()
    print(f"kostov86: {len(act)} active bounds at x = {[int(i) for i in act]}, "
          f"rounds {ip.fit_stats()['rounds']}", flush=True)
    # an active bound is fitted at one of its bounds
    for i in act:
        gaps = [abs(fit[i] - b) for b in (lb[i], ub[i]) if not np.isnan(b)]
        assert min(gaps) < tol


@pytest.mark.gpu
def test_inequality_fit_bounds_only():
    """The reference's inequality_only case (test_inequality_fitter.cpp:25-58):
    no values, bounds of +-0.001 around a smooth function."""
    pts, vals = sample(10_000, 33)
    tol = 1e-4
    rbf = pa.Biharmonic3D([1.0])
    rbf.set_anisotropy(A)
    ip = pa.Interpolant(pa.Model(rbf, poly_degree=0))
    lb, ub = vals - 0.001, vals + 0.001
    ip.fit_inequality(pts, np.full(len(pts), np.nan), lb, ub, tol, max_iter=32, accuracy=tol / 100)
    st = ip.fit_stats()
    n = len(ip.centers())
    print(f"bounds only: {n} of {len(pts)} points active in {st['rounds']} rounds", flush=True)
    assert ip.weights().shape == (n + 1,)
    fit = ip.evaluate(pts, accuracy=tol / 100)
    assert (fit > lb - tol).all() and (fit < ub + tol).all()


@pytest.mark.gpu
def test_inequality_fit_value_and_bound_on_one_point():
    """A point with a value and a bound that its value violates is fitted at
    the value and appears once among the centres."""
    pts, v, lb, ub = kostov86()
    lb[0] = 5.0  # v[0] = 1
    tol = 1e-5
    ip = pa.Interpolant(pa.Model(pa.CovExponential([1.0, 3.0]), poly_degree=-1))
    ip.fit_inequality(pts, v, lb, ub, tol, max_iter=32, accuracy=tol / 100)
    c = ip.centers()
    assert len(np.unique(c, axis=0)) == len(c)
    fit = ip.evaluate(pts, accuracy=tol / 100)
    assert abs(fit[0] - 1.0) < tol
