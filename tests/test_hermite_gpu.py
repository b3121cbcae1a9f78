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
