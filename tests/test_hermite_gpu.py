"""Fits with gradient data (Hermite-Birkhoff) on the GPU against the CPU
oracle: the operator of interpolation/operator.hpp:56-82 with the blocks of
preconditioner/mat_a.hpp:11-61 (the reference's test_operator.cpp sizes).
Calls go through the C ABI."""
import sys

import numpy as np
import pytest

import polatory_amd as pa

pytestmark = pytest.mark.gpu

ANISO = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.1, 0.0, 1.2]])


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def th3_model(nugget=0.01):
    r = pa.Triharmonic3D([1.0])
    r.set_anisotropy(ANISO)
    return pa.Model(r, poly_degree=1, nugget=nugget), r


def orc_rbf(orc, r):
    return orc.make_rbf(orc.TH3, r.parameters(), r.anisotropy())


def block_errors(y, ref, mu, sigma):
    n = mu + 3 * sigma
    out = []
    for a, b in ((0, mu), (mu, n), (n, len(y))):
        if b > a:
            out.append(np.abs(y[a:b] - ref[a:b]).max() / max(np.abs(ref[a:b]).max(), 1e-300))
    return out


@pytest.mark.parametrize("mode,mu,sigma,tols", [
    ("direct", 1024, 1024, (1e-12, 1e-12, 1e-12)),
    # FP64 FMM of order 10: gradients of the interpolated far field lose about
    # two digits against values (the reference's test_operator.cpp allows 1e-4
    # absolute on the gradient rows)
    ("fmm", 6000, 5000, (1e-8, 1e-6, 1e-12)),
    ("direct", 1, 3000, (1e-12, 1e-12, 1e-12)),
    ("direct", 2500, 0, (1e-12, 1e-12)),
])
def test_hermite_operator_matches_oracle(orc, mode, mu, sigma, tols):
    log("hermite operator", mode, mu, sigma)
    rng = np.random.default_rng(50 + mu + sigma)
    pts = rng.uniform(-1, 1, (mu, 3))
    gpts = rng.uniform(-1, 1, (sigma, 3))
    model, r = th3_model()
    n = mu + 3 * sigma + 4
    x = rng.normal(size=n)
    op = pa.Operator(model, pts, pa.FmmConfig("fp64", 10, -1, mode), grad_points=gpts)
    assert op.size() == n
    y = op(x)
    ref = orc.hermite_operator([orc_rbf(orc, r)], 1, 0.01, pts, gpts, x)
    errs = block_errors(y, ref, mu, sigma)
    log("  relative errors (values, gradients, polynomial):", ", ".join(f"{e:.2e}" for e in errs))
    for e, t in zip(errs, tols):
        assert e <= t


def hermite_matrix(orc, rbfs, nugget, pts, gpts, degree=1):
    n = len(pts) + 3 * len(gpts) + orc.poly_basis_size(degree)
    eye = np.eye(n)
    return np.stack([orc.hermite_operator(rbfs, degree, nugget, pts, gpts, eye[:, j]) for j in range(n)], axis=1)


def test_hermite_ras_single_level_is_the_exact_solve(orc):
    # Below the coarse-grid size the preconditioner is one coarse solve
    # (coarse_grid.hpp:40-131): the solution of the whole saddle-point system.
    log("hermite ras, one level")
    rng = np.random.default_rng(61)
    pts = rng.uniform(-1, 1, (300, 3))
    gpts = rng.uniform(-1, 1, (150, 3))
    model, r = th3_model()
    ras = pa.RasPreconditioner(model, pts, grad_points=gpts)
    assert ras.n_levels() == 1
    n = 300 + 3 * 150
    v = np.concatenate([rng.normal(size=n), np.zeros(4)])
    x = ras.apply(v)
    m = hermite_matrix(orc, [orc_rbf(orc, r)], 0.01, pts, gpts)
    ref = np.linalg.solve(m, v)
    e = np.abs(x - ref).max() / np.abs(ref).max()
    log(f"  relative difference to the dense solve {e:.2e}")
    assert e <= 1e-6


def reference_data(n_points, n_grad_points, seed):
    # test/utility.hpp:45-93: f(p) = sum_j sin(pi (A p)_j), gradients A^T (pi cos(pi (A p)_j))
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-1, 1, (n_points, 3))
    gpts = rng.uniform(-1, 1, (n_grad_points, 3))
    values = np.sin(np.pi * pts @ ANISO.T).sum(axis=1)
    grads = (np.pi * np.cos(np.pi * gpts @ ANISO.T)) @ ANISO
    return pts, gpts, np.concatenate([values, grads.ravel()])


@pytest.mark.parametrize("n_points,n_grad_points", [(10000, 10000), (1, 10000), (3000, 0)])
def test_hermite_fit_reference_cases(orc, n_points, n_grad_points):
    # test/interpolation/test_fitter.cpp:24-73: triharmonic3d with anisotropy,
    # degree cpd_order - 1 = 1, nugget 0.01, tolerance 1e-3 on values and on
    # gradients, checked here by the oracle's FP64 direct operator.
    log("hermite fit", n_points, n_grad_points)
    pts, gpts, data = reference_data(n_points, n_grad_points, 70 + n_points)
    model, r = th3_model()
    ip = pa.Interpolant(model)
    ip.fit_hermite(pts, gpts, data, 1e-3, 1e-3, max_iter=100)
    st = ip.fit_stats()
    w = ip.weights()
    mu, n = n_points, n_points + 3 * n_grad_points
    assert w.shape[0] == n + 4
    y = orc.hermite_operator([orc_rbf(orc, r)], 1, 0.01, pts, gpts, w)
    res = np.abs(y[:mu] - data[:mu]).max()
    gres = np.abs(y[mu:n] - data[mu:n]).max() if n_grad_points else 0.0
    log(f"  {st['iterations']} iterations, {st['n_levels']} levels; oracle residuals {res:.2e} (values), "
        f"{gres:.2e} (gradients); solver's {st['residual']:.2e}, {st['grad_residual']:.2e}")
    assert res <= 1e-3 and gres <= 1e-3
    assert np.abs(y[n:]).max() <= 1e-8 * max(1.0, np.abs(w[:n]).max())  # P^T [w; g] = 0
    # The fitted function at the data sites through the interpolant's evaluation
    # (test_fitter.cpp:51-63 adds nugget * w to the values).
    f = ip.evaluate_hermite(pts, gpts)
    f[:mu] += 0.01 * w[:mu]
    assert np.abs(f[:mu] - data[:mu]).max() <= 1e-3 + 1e-6
    if n_grad_points:
        assert np.abs(f[mu:] - data[mu:]).max() <= 1e-3 + 1e-5
