"""GPU tests of the request loop over resident models (continuous batching).

Each answer is checked against the oracle's FP64 direct sum (with polynomial
terms) at the stated relative accuracy, and every request's result must be
identical alone and in the concurrent mix (tolerance 0: the kernels sum in a
fixed order, so sharing the device must not change a single bit).
"""
import sys

import numpy as np
import pytest

import polatory_amd as pa

pytestmark = pytest.mark.gpu

REL_ACC = 1e-6
# Measured against the oracle on at most 400 targets per request, relative to
# the largest |value| among them: the serving contract itself.
REL_TOL = REL_ACC


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def orc_model_rbfs(orc, rbfs):
    out = []
    for r in rbfs:
        kind = {"bh3": orc.BH3, "th3": orc.TH3, "exp": orc.EXP}[r.short_name]
        out.append(orc.make_rbf(kind, r.parameters(), r.anisotropy()))
    return out


def sample_values(p):
    return np.sin(3 * p[:, 0]) * np.cos(2 * p[:, 1]) + 0.5 * p[:, 2] ** 2


@pytest.fixture(scope="module")
def served(orc):
    """Three fitted models and one large model with N(0,1) weights."""
    log("serve: fitting models")
    models = []
    specs = [(3000, pa.Biharmonic3D(), 1, 61), (5000, pa.Triharmonic3D(), 1, 62),
             (4000, pa.CovExponential([1.0, 0.5]), 0, 63)]
    for n, rbf, deg, seed in specs:
        pts = np.random.default_rng(seed).uniform(0, 1, (n, 3))
        ip = pa.Interpolant(pa.Model(rbf, poly_degree=deg, nugget=0.0))
        ip.fit(pts, sample_values(pts), 1e-4)
        models.append((ip, rbf, deg))
    n = 200_000
    pts = np.random.default_rng(64).uniform(0, 1, (n, 3))
    w = np.concatenate([np.random.default_rng(65).normal(size=n), [0.3, -0.2, 0.1, 0.5]])
    ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
    ip.set_fitted(pts, w)
    models.append((ip, pa.Biharmonic3D(), 1))
    srv = pa.Server(n_slots=4, rel_accuracy=REL_ACC)
    ids = []
    for ip, _, _ in models:
        ids.append(srv.add_model(ip))
        cfg, nbytes = srv.model_config(ids[-1])
        log(f"  model {ids[-1]}: {ip.num_centers()} centres, {cfg}, resident {nbytes / 2**20:.1f} MiB")
    yield srv, models, ids
    srv.close()


def reference(orc, model_tuple, q):
    ip, rbf, deg = model_tuple
    return orc.direct_evaluate(orc_model_rbfs(orc, [rbf]), deg, ip.centers(), ip.weights(), q)


def make_requests(n_models, seed, sizes):
    rng = np.random.default_rng(seed)
    reqs = []
    for s in sizes:
        reqs.append((int(rng.integers(n_models)), rng.uniform(0, 1, (s, 3))))
    return reqs


def test_mixed_equals_alone_and_meets_accuracy(orc, served):
    srv, models, ids = served
    sizes = [1, 7, 300, 2000, 40000, 150000, 3, 90000, 600, 250000, 12, 5000]
    reqs = make_requests(len(ids), 71, sizes)
    log(f"serve: {len(reqs)} requests in the mix")
    srv.reset_stats()
    rids = [srv.submit(ids[m], q) for m, q in reqs]
    mixed = [srv.wait(r) for r in rids]
    st = srv.stats()
    log(f"  mixed: completed {st['completed']}, max in flight {st['max_in_flight']}, "
        f"p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms")
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_in_flight"] >= 2, "requests never overlapped"
    for i, (m, q) in enumerate(reqs):
        alone = srv.wait(srv.submit(ids[m], q))
        assert np.array_equal(alone, mixed[i]), f"request {i} differs alone vs mixed"
        sub = np.arange(min(len(q), 400))
        ref = reference(orc, models[m], q[sub])
        err = np.abs(mixed[i][sub] - ref).max() / max(np.abs(ref).max(), 1e-300)
        log(f"  request {i}: model {m}, {len(q)} points, rel err {err:.2e}")
        assert err <= REL_TOL


def test_admission_by_budget_and_errors(orc, served):
    srv, models, ids = served
    with pytest.raises(ValueError):
        srv.submit(99, np.zeros((1, 3)))
    # queries outside the model box fail on their own, without stopping the loop
    bad = srv.submit(ids[3], np.full((5000, 3), 2.0))
    good = srv.submit(ids[3], np.full((5, 3), 0.5))
    with pytest.raises(RuntimeError):
        srv.wait(bad)
    out = srv.wait(good)
    ref = reference(orc, models[3], np.full((5, 3), 0.5))
    assert np.abs(out - ref).max() <= REL_TOL * np.abs(ref).max()
    # a budget that holds only the resident model: large requests are refused
    tiny = pa.Server(n_slots=2, hbm_budget=1 << 30, rel_accuracy=REL_ACC)
    try:
        mid = tiny.add_model(models[0][0])
        r1 = tiny.submit(mid, np.random.default_rng(5).uniform(0, 1, (8_000_000, 3)))
        r2 = tiny.submit(mid, np.random.default_rng(6).uniform(0, 1, (100, 3)))
        with pytest.raises(RuntimeError):
            tiny.wait(r1)
        out = tiny.wait(r2)
        ref = reference(orc, models[0], np.random.default_rng(6).uniform(0, 1, (100, 3)))
        assert np.abs(out - ref).max() <= REL_TOL * np.abs(ref).max()
    finally:
        tiny.close()


def test_empty_request(served):
    srv, _, ids = served
    out = srv.wait(srv.submit(ids[0], np.zeros((0, 3))))
    assert out.shape == (0,)


def test_small_requests_share_launches(orc, served):
    """Small requests (answered by the FP64 direct sum) queued behind large ones
    are served several models to a launch; each answer is the same bits as the
    same request alone and agrees with the oracle's FP64 direct sum."""
    srv, models, ids = served
    rng = np.random.default_rng(81)
    bigq = [rng.uniform(0, 1, (250_000, 3)) for _ in range(4)]
    small = []
    for k in range(48):
        m = k % len(ids)
        n = int(rng.integers(1, 300 if m == 3 else 3000))
        small.append((m, rng.uniform(0, 1, (n, 3))))
    srv.reset_stats()
    big = [srv.submit(ids[3], q) for q in bigq]
    rids = [srv.submit(ids[m], q) for m, q in small]
    mixed = [srv.wait(r) for r in rids]
    for r in big:
        srv.wait(r)
    st = srv.stats()
    log(f"serve: {len(small)} small requests in {st['batches']} shared launches "
        f"(largest {st['max_batch']})")
    assert st["batched_requests"] == len(small)
    assert st["max_batch"] >= 2, "no two small requests shared a launch"
    for i, (m, q) in enumerate(small):
        alone = srv.wait(srv.submit(ids[m], q))
        assert np.array_equal(alone, mixed[i]), f"request {i} differs alone vs batched"
        ref = reference(orc, models[m], q)
        err = np.abs(mixed[i] - ref).max() / max(np.abs(ref).max(), 1e-300)
        # FP64 sums in another order than the oracle's: rounding only
        assert err <= 1e-10, f"request {i}: rel err {err:.2e} vs the FP64 direct sum"


def test_fmm_requests_share_near_and_far_field_launches(orc, served):
    """FMM-sized requests of several models queued together are served in
    batches whose near fields and far-field L2Ps run in shared launches; every
    answer is the same bits as the same request served alone, and within the
    accuracy of the oracle's FP64 direct sum."""
    srv, models, ids = served
    rng = np.random.default_rng(91)
    reqs = []
    for k in range(12):
        m = k % len(ids)
        reqs.append((m, rng.uniform(0, 1, (int(rng.integers(30_000, 120_000)), 3))))
    srv.reset_stats()
    rids = [srv.submit(ids[m], q) for m, q in reqs]
    mixed = [srv.wait(r) for r in rids]
    st = srv.stats()
    log(f"serve: {len(reqs)} FMM requests in {st['fmm_batches']} batches (largest {st['max_fmm_batch']})")
    assert st["fmm_batched_requests"] == len(reqs)
    assert st["max_fmm_batch"] >= 2, "no two FMM requests shared a launch"
    for i, (m, q) in enumerate(reqs):
        alone = srv.wait(srv.submit(ids[m], q))
        assert np.array_equal(alone, mixed[i]), f"request {i} differs alone vs in an FMM batch"
        sub = np.arange(300)
        ref = reference(orc, models[m], q[sub])
        err = np.abs(mixed[i][sub] - ref).max() / np.abs(ref).max()
        assert err <= REL_TOL, f"request {i}: rel err {err:.2e}"
