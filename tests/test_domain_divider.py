"""The RAS domain divider (host code of the library, no GPU): the invariants
of the reference's test/preconditioner/test_domain_divider.cpp, and the same
domains and coarse points as the oracle's restatement of
domain_divider.hpp:54-249, with and without gradient points."""
import numpy as np
import pytest

import polatory_amd as pa
from oracle import cpu_fit


def canonical(domains):
    return sorted((tuple(i), tuple(a), tuple(g), tuple(b)) for i, a, g, b in
                  ((np.asarray(d[0]).tolist(), np.asarray(d[1]).astype(bool).tolist(),
                    np.asarray(d[2]).tolist(), np.asarray(d[3]).astype(bool).tolist()) for d in domains))


@pytest.mark.parametrize("n_grad", [10000, 0])
def test_reference_invariants(n_grad):
    # test_domain_divider.cpp:20-89
    rng = np.random.default_rng(3)
    n, n_poly = 10000, 10
    pts = rng.uniform(0, 1, (n, 3))
    gpts = rng.uniform(0, 1, (n_grad, 3))
    poly = np.arange(n_poly)
    div = pa.DomainDivider(pts, gpts, np.arange(n), np.arange(n_grad), poly)
    inner, ginner = [], []
    for idx, inn, gidx, ginn in div.domains():
        assert np.array_equal(idx[:n_poly], poly)
        assert len(np.unique(idx)) == len(idx) and len(np.unique(gidx)) == len(gidx)
        assert len(idx) - n_poly + 3 * len(gidx) <= 1024
        inner += idx[inn].tolist()
        ginner += gidx[ginn].tolist()
    assert sorted(inner) == list(range(n))
    assert sorted(ginner) == list(range(n_grad))
    n_coarse = 1000
    cidx, cgidx = div.choose_coarse_points(n_coarse)
    rows = len(cidx) + 3 * len(cgidx)
    # A split trades one centre (1 or 3 rows) for two (up to 3 rows each), so
    # the loop that stops at n_coarse rows overshoots by at most 4 rows (the
    # reference's test data, uninitialized points, stays within 2).
    assert n_coarse + n_poly <= rows <= n_coarse + n_poly + (4 if n_grad else 1)
    assert np.array_equal(cidx[:n_poly], poly)


@pytest.mark.parametrize("n_pts,n_grad", [(3000, 2000), (1, 2500), (6000, 0), (300000, 0)])
def test_matches_oracle_restatement(n_pts, n_grad):
    rng = np.random.default_rng(5 + n_pts)
    pts = rng.uniform(-1, 1, (n_pts, 3)) * np.array([1.0, 0.7, 0.4])
    gpts = rng.uniform(-1, 1, (n_grad, 3)) * np.array([1.0, 0.7, 0.4])
    poly = np.arange(min(4, n_pts))
    div = pa.DomainDivider(pts, gpts, np.arange(n_pts), np.arange(n_grad), poly)
    if n_grad:
        ref = cpu_fit.divide_domains_mixed(pts, gpts, np.arange(n_pts), np.arange(n_grad), poly)
    else:
        ref = [(i, a, np.zeros(0, int), np.zeros(0, bool)) for i, a in cpu_fit.divide_domains(pts, np.arange(n_pts), poly)]
    assert canonical(div.domains()) == canonical(ref)
    n_coarse = 800
    cidx, cgidx = div.choose_coarse_points(n_coarse)
    if n_grad:
        ridx, rgidx = cpu_fit.choose_coarse_points_mixed(pts, gpts, np.arange(n_pts), np.arange(n_grad), poly, n_coarse)
    else:
        ridx, rgidx = cpu_fit.choose_coarse_points(pts, np.arange(n_pts), poly, n_coarse), np.zeros(0, int)
    assert np.array_equal(cidx[:len(poly)], poly)
    assert sorted(cidx.tolist()) == sorted(np.asarray(ridx).tolist())
    assert sorted(cgidx.tolist()) == sorted(np.asarray(rgidx).tolist())
