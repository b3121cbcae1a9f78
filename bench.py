"""Benchmark: FMM evaluation of 1,000,000 biharmonic3d sources at 1,000,000
targets on one MI355X (mixed FP32/FP32-MFMA form vs the plain FP64 form).

Workload (seeded, synthetic): sources U[0,1]^3 (seed 3), weights N(0,1) (seed 4),
degree-1 polynomial coefficients N(0,1) (seed 5), targets U[0,1]^3 (seed 6);
error vs FP64 direct summation on a 10,000-target subsample (seed 7).
A step = one full evaluation at the FMM seam from device-resident points and
weights: point transform, source and target octrees, interaction lists,
P2M/M2M, M2L, L2L, L2P and P2P, output in target input order in HBM. The
degree-1 polynomial terms are not part of the seam (polatory's Evaluator adds
them after the FMM, interpolation/evaluator.hpp:95-118); they are added on the
host after the timed steps, for the error check only.

Prints one JSON line (see the contract in the repository documentation).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_SRC = 1_000_000
N_TRG = 1_000_000
N_SUB = 10_000
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def hbm_traffic(kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 counter summary in
    profiles/hbm_traffic.txt (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by
    the factors measured on a plain device copy); None when not profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and not line.startswith("#") and f[0] == kernel:
            fetch_kib, write_kib, fs, ws = map(float, f[1:])
            return (fetch_kib * fs + write_kib * ws) * 1024.0
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs():
    src = np.random.default_rng(3).uniform(0, 1, (N_SRC, 3))
    w = np.random.default_rng(4).normal(size=N_SRC)
    c = np.random.default_rng(5).normal(size=4)
    trg = np.random.default_rng(6).uniform(0, 1, (N_TRG, 3))
    sub = np.random.default_rng(7).permutation(N_TRG)[:N_SUB]
    return src, w, c, trg, sub


class DeviceEval:
    """Evaluator with device-resident inputs (torch owns the HBM buffers)."""

    def __init__(self, torch, src, w, trg, cfg, gradient=False):
        import polatory_amd as pa
        from polatory_amd import _capi
        self.pa, self.capi, self.torch = pa, _capi, torch
        dev = torch.device("cuda:0")
        self.d_src = torch.from_numpy(src).to(dev).contiguous()
        self.d_trg = torch.from_numpy(trg).to(dev).contiguous()
        self.d_w = torch.from_numpy(w).to(dev).contiguous()
        self.gradient = gradient
        self.d_out = torch.empty(trg.shape[0] * (3 if gradient else 1), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        self.ev = pa.make_fmm_evaluator(pa.Biharmonic3D(), ((0, 0, 0), (1, 1, 1)))
        self.ev.set_config(cfg)
        self.n_src, self.n_trg = src.shape[0], trg.shape[0]

    def step(self):
        P = ctypes.c_void_p
        h = self.ev._h
        self.capi.call("pam_fmm_evaluator_set_source_points_device", h, P(self.d_src.data_ptr()), self.n_src)
        self.capi.call("pam_fmm_evaluator_set_weights_device", h, P(self.d_w.data_ptr()), self.n_src)
        self.capi.call("pam_fmm_evaluator_set_target_points_device", h, P(self.d_trg.data_ptr()), self.n_trg)
        what = "pam_fmm_evaluator_evaluate_gradient_device" if self.gradient else "pam_fmm_evaluator_evaluate_device"
        self.capi.call(what, h, P(self.d_out.data_ptr()), self.n_trg)

    def sync(self):
        self.capi.call("pam_fmm_evaluator_synchronize", self.ev._h)

    def output(self):
        return self.d_out.cpu().numpy()


def search_config(pa, src, w, prec, acc):
    """The configuration the accuracy search picks for one precision at the
    absolute accuracy acc (polatory's accuracy is absolute: here 1e-6 x max |y|):
    orders 4..12 of that precision tried cheapest first on 10,000 source points
    used as targets against the FP64 direct sum (fmm_accuracy_estimator.hpp:74-121)."""
    ev = pa.make_fmm_evaluator(pa.Biharmonic3D(), ((0, 0, 0), (1, 1, 1)))
    ev.set_source_points(src)
    ev.set_weights(w)
    ev.set_target_points(src[:1])
    ev.set_search_precision(prec)
    ev.set_accuracy(acc)
    ev.evaluate()
    c = ev.config()
    return pa.FmmConfig(prec, c["order"], -1, "fmm")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_fmm_baseline(src, w, trg, sub, ref, acc):
    """The reference algorithm on the host cores (oracle/cpu_fmm.c): polatory's
    evaluate path with its tree height, box and separation-1 lists over a
    uniform-interpolation FMM with DFT-accelerated M2L, FP64, OpenMP. Orders as
    polatory's accuracy search tries them (8, then 10, ...); the first whose
    error on the subsample is within acc is the one timed."""
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    r = oracle.make_rbf(oracle.BH3)
    for order in (8, 10, 12):
        y, st = oracle.cpu_fmm_evaluate(r, src, w, trg, ((0, 0, 0), (1, 1, 1)), order)
        err = float(np.abs(y[sub] - ref).max())
        log(f"cpu fmm order {order}: {st['t_total']:.1f} s, abs err {err:.3e} (accuracy {acc:.3e}), "
            f"height {st['height']}, m2l pairs {st['m2l_pairs']}")
        if err <= acc:
            break
    return {"value": len(trg) / st["t_total"], "unit": "evals/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"the whole workload ({len(src)} sources -> {len(trg)} targets), one evaluation: "
                      f"reference-algorithm CPU FMM (height {st['height']}, uniform interpolation order "
                      f"{order}, DFT M2L, FP64, OpenMP over {cores} threads), {st['t_total']:.1f} s",
            "order": order, "height": st["height"], "rel_err": err / float(np.abs(ref).max()),
            "stage_s": {k: st[k] for k in ("t_tree", "t_upward", "t_m2l", "t_downward", "t_p2p")}}


def poly(c, pts):
    return c[0] + pts @ c[1:]


def sample_f(p):
    """Data of the fit workloads: f(x, y, z) = sin(3x) cos(2y) + 0.5 z^2."""
    return np.sin(3 * p[:, 0]) * np.cos(2 * p[:, 1]) + 0.5 * p[:, 2] ** 2


def config1_section(pa, no_cpu):
    """The plain fit + evaluate that polatory runs on the CPU today: 10,000
    biharmonic3d sites in the unit cube with a degree-1 polynomial, f as in
    the other fits, tolerance 1e-4; then 100,000 evaluation points. The GPU
    fit and evaluation (default options) beside the reference algorithm on the
    host cores (oracle/cpu_fit.py: FGMRES + multilevel RAS with FP64 direct
    sums; oracle/cpu_fmm.c at interpolation order 6, polatory's default
    evaluation accuracy), both checked by the FP64 direct sum."""
    import oracle
    from oracle import cpu_fit
    pts = np.random.default_rng(19).uniform(0, 1, (10_000, 3))
    vals = sample_f(pts)
    q = np.random.default_rng(20).uniform(0, 1, (100_000, 3))
    sub = np.random.default_rng(21).permutation(len(q))[:2_000]
    r = oracle.make_rbf(oracle.BH3)
    out = {"workload": "10,000 bh3 sites in [0,1]^3, degree 1, tol 1e-4; evaluate at 100,000 points"}
    ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
    ip.fit(pts, vals, 1e-4)  # warm-up (code objects, allocations)
    ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
    t0 = time.perf_counter()
    ip.fit(pts, vals, 1e-4)
    t_fit = time.perf_counter() - t0
    w = ip.weights()
    res = float(np.abs(oracle.direct_evaluate([r], 1, pts, w, pts) - vals).max())
    ip.evaluate(q[:10])
    t0 = time.perf_counter()
    y = ip.evaluate(q)
    t_eval = time.perf_counter() - t0
    ref = oracle.direct_evaluate([r], 1, pts, w, q[sub])
    out["gpu"] = {"fit_s": t_fit, "iterations": ip.fit_stats()["iterations"], "fp64_residual_oracle": res,
                  "eval_s": t_eval, "eval_rel_err_vs_fp64_direct": float(np.abs(y[sub] - ref).max() / np.abs(ref).max())}
    log("config1 gpu", json.dumps(out["gpu"]))
    if no_cpu:
        return out
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    wc, st = cpu_fit.cpu_fit([r], 1, 0.0, pts, vals, 1e-4, threads=cores)
    t_cfit = time.perf_counter() - t0
    t0 = time.perf_counter()
    yc, _ = oracle.cpu_fmm_evaluate(r, pts, wc[:len(pts)], q, ((0, 0, 0), (1, 1, 1)), 6)
    yc += wc[len(pts)] + q @ wc[len(pts) + 1:]
    t_ceval = time.perf_counter() - t0
    refc = oracle.direct_evaluate([r], 1, pts, wc, q[sub])
    out["cpu_baseline"] = {"kind": "port", "cores": cores, "cpu": cpu_model(), "fit_s": t_cfit,
                           "iterations": st["iterations"], "fp64_residual": st["residual"], "levels": st["n_levels"],
                           "eval_s": t_ceval,
                           "eval_rel_err_vs_fp64_direct": float(np.abs(yc[sub] - refc).max() / np.abs(refc).max()),
                           "sample": "the whole workload: oracle/cpu_fit.py (FGMRES + multilevel RAS, FP64 direct "
                                     "sums, LAPACK factors) and oracle/cpu_fmm.c (order 6), one run each"}
    out["fit_speedup_vs_cpu"] = t_cfit / t_fit
    log("config1 cpu", json.dumps(out["cpu_baseline"]))
    return out


def ellipsoid_points(n_surf, seed=8):
    """333,334 area-uniform points on the ellipsoid (1, 0.7, 0.5) (rejection on the
    sphere's image) with value 0, plus the same points offset by +-0.01 along the
    analytic normal with values +-0.01."""
    a = np.array([1.0, 0.7, 0.5])
    rng = np.random.default_rng(seed)
    gmax = max(a[1] * a[2], a[0] * a[2], a[0] * a[1])
    got, parts = 0, []
    while got < n_surf:
        u = rng.normal(size=(2 * n_surf, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        g = np.sqrt((a[1] * a[2] * u[:, 0]) ** 2 + (a[0] * a[2] * u[:, 1]) ** 2 + (a[0] * a[1] * u[:, 2]) ** 2)
        keep = u[rng.uniform(size=len(u)) * gmax < g] * a
        parts.append(keep)
        got += len(keep)
    p = np.concatenate(parts)[:n_surf]
    nrm = p / a ** 2
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    pts = np.concatenate([p, p + 0.01 * nrm, p - 0.01 * nrm])
    vals = np.concatenate([np.zeros(n_surf), np.full(n_surf, 0.01), np.full(n_surf, -0.01)])
    return pts, vals


def fit_section(pa):
    """1,000,002-point ellipsoid fit, biharmonic3d degree 1, absolute tolerance 1e-5:
    mixed-precision operator (FP32 P2P, FP32-MFMA M2L, FP64 refinement) vs FP64
    everywhere (operator and RAS level updates as FP64 FMM of the same order).
    The final residual is recomputed in FP64 at every site by the solver.
    Each form is fitted once untimed first, so that one-time costs of the
    process (code objects, M2L operators, cached device buffers and streams)
    fall on neither timed fit."""
    pts, vals = ellipsoid_points(333_334)
    out = {"workload": "ellipsoid surface fit, 1,000,002 points, bh3 degree 1, tol 1e-5 absolute"}
    sub = np.random.default_rng(18).permutation(len(pts))[:10_000]
    forms = (("mixed", pa.FitOptions()),
             ("fp64", pa.FitOptions(op=pa.FmmConfig("fp64", 8, -1, "auto"),
                                    level=pa.FmmConfig("fp64", 8, -1, "auto"),
                                    refine_rtol=0.0, ras_fp32_factors=False)))
    for name, opt in forms:
        warm = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
        warm.fit(pts, vals, 1e-5, max_iter=100, options=opt)
        del warm
    for name, opt in forms:
        ip = pa.Interpolant(pa.Model(pa.Biharmonic3D(), poly_degree=1))
        t0 = time.perf_counter()
        ip.fit(pts, vals, 1e-5, max_iter=100, options=opt)
        wall = time.perf_counter() - t0
        st = ip.fit_stats()
        # the final residual again, by the GPU FP64 direct sum on 10,000 sites
        w = ip.weights()
        n = len(pts)
        y = pa.direct_evaluate(pa.Biharmonic3D(), pts, w[:n], pts[sub]) + w[n] + pts[sub] @ w[n + 1:n + 4]
        res_direct = float(np.abs(vals[sub] - y).max())
        out[name] = {"wall_s": wall, "iterations": st["iterations"], "restarts": st["restarts"],
                     "fp64_residual": st["residual"], "fp64_direct_residual_10k_sites": res_direct,
                     "solve_s": st["seconds"],
                     "precond_setup_s": st["precond_setup_seconds"], "ras_levels": st["n_levels"],
                     "factor_gb": st["factor_bytes"] / 1e9}
        log("fit", name, json.dumps(out[name]))
        del ip
    out["speedup_vs_fp64"] = out["fp64"]["wall_s"] / out["mixed"]["wall_s"]
    return out


def gmm_sites(n, seed_sites=9, seed_field=10, seed_noise=11):
    """500,000 sites from a 20-component Gaussian mixture (means U[0,10]^3,
    isotropic std U[0.2,1.0], equal weights); values = 8 random sines plus
    N(0, 0.05^2) noise."""
    rng = np.random.default_rng(seed_sites)
    means = rng.uniform(0, 10, (20, 3))
    stds = rng.uniform(0.2, 1.0, 20)
    comp = rng.integers(0, 20, n)
    pts = means[comp] + stds[comp, None] * rng.normal(size=(n, 3))
    rf = np.random.default_rng(seed_field)
    amp = rf.uniform(0.5, 1.0, 8)
    freq = rf.normal(0.0, 0.5, (8, 3))
    phase = rf.uniform(0, 2 * np.pi, 8)

    def field(p):
        return (np.sin(p @ freq.T + phase) * amp).sum(axis=1) / 4.0

    vals = field(pts) + np.random.default_rng(seed_noise).normal(0.0, 0.05, n)
    return pts, vals


HB_ANISO = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.1, 0.0, 1.2]])


def hermite_section(pa):
    """Fits with gradient data: the reference's fitter case (test_fitter.cpp:24-73:
    triharmonic3d with anisotropy, degree 1, nugget 0.01, tolerance 1e-3 on the
    values and on the gradients, f = sum_j sin(pi (A p)_j) on [-1, 1]^3) at
    10,000 + 10,000 points and at 100,000 + 50,000."""
    out = {"workload": "th3 + anisotropy, degree 1, nugget 0.01, values and gradients of "
                       "sum_j sin(pi (A p)_j), tolerance 1e-3 on both"}
    for n, ng in ((10_000, 10_000), (100_000, 50_000)):
        rng = np.random.default_rng(n + ng)
        p = rng.uniform(-1, 1, (n, 3))
        g = rng.uniform(-1, 1, (ng, 3))
        data = np.concatenate([np.sin(np.pi * p @ HB_ANISO.T).sum(axis=1),
                               ((np.pi * np.cos(np.pi * g @ HB_ANISO.T)) @ HB_ANISO).ravel()])
        r = pa.Triharmonic3D([1.0])
        r.set_anisotropy(HB_ANISO)
        ip = pa.Interpolant(pa.Model(r, poly_degree=1, nugget=0.01))
        t0 = time.perf_counter()
        ip.fit_hermite(p, g, data, 1e-3, 1e-3, max_iter=100)
        wall = time.perf_counter() - t0
        st = ip.fit_stats()
        out[f"{n}+{ng}"] = {"wall_s": wall, "iterations": st["iterations"], "ras_levels": st["n_levels"],
                            "residual": st["residual"], "grad_residual": st["grad_residual"]}
        log("hermite", n, ng, json.dumps(out[f"{n}+{ng}"]))
        del ip
    return out


def kriging_section(pa):
    """Anisotropic kriging: exponential covariance (sill 1, range 2), anisotropy
    A = diag(1, 1, 1/0.3) (ranges 2, 2, 0.6), nugget 0.0025, degree 0, fit to
    1e-4 absolute; prediction on the 128^3 cell centres of [0,10]^3 with the
    mixed (FP32 near field + FP32-MFMA M2L), the mixed form with compensated
    near-field sums, and the FP64 evaluator."""
    pts, vals = gmm_sites(500_000)
    rbf = pa.CovExponential([1.0, 2.0])
    rbf.set_anisotropy(np.diag([1.0, 1.0, 1.0 / 0.3]))
    ip = pa.Interpolant(pa.Model(rbf, poly_degree=0, nugget=0.0025))
    t0 = time.perf_counter()
    ip.fit(pts, vals, 1e-4, max_iter=100)
    t_fit = time.perf_counter() - t0
    st = ip.fit_stats()
    g = (np.arange(128) + 0.5) * (10.0 / 128)
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    out = {"workload": "exp covariance kriging, 500,000 GMM sites, aniso diag(1,1,1/0.3), nugget 0.0025, "
                       "tol 1e-4; predict 128^3 grid",
           "fit_s": t_fit, "iterations": st["iterations"], "fp64_residual": st["residual"]}
    sub = np.random.default_rng(17).permutation(len(grid))[:2000]
    w = ip.weights()
    n = ip.num_centers()
    ref = pa.direct_evaluate(rbf, ip.centers(), w[:n], grid[sub]) + w[n]
    preds = {}
    for name, cfg in (("mixed", pa.FmmConfig("mixed", 8, -1, "auto")),
                      ("mixed_comp", pa.FmmConfig("mixed_comp", 8, -1, "auto")),
                      ("fp64", pa.FmmConfig("fp64", 10, -1, "auto"))):
        ip.set_eval_config(cfg)
        ip.evaluate(grid)  # builds the evaluator and trees
        t0 = time.perf_counter()
        y = ip.evaluate(grid)
        dt = time.perf_counter() - t0
        preds[name] = y
        out[name] = {"predict_s": dt, "points_per_s": len(grid) / dt,
                     "rel_err_vs_fp64_direct": float(np.abs(y[sub] - ref).max() / np.abs(ref).max())}
        log("kriging", name, json.dumps(out[name]))
    out["max_rel_diff_mixed_vs_fp64"] = float(np.abs(preds["mixed"] - preds["fp64"]).max()
                                              / np.abs(preds["fp64"]).max())
    log("kriging", json.dumps({k: v for k, v in out.items() if not isinstance(v, dict)}))
    return out


def serve_section(pa, n_models, n_requests, rate):
    """32 resident models fitted to f on N ~ log-uniform[1e4, 1e6] U[0,1]^3 sites
    (bh3 or th3, 50/50, tol 1e-4), Poisson arrivals at `rate` req/s, each
    request on a uniformly chosen model with log-uniform[1, 1e6] U[0,1]^3 query
    points, 1e-6 relative accuracy, admission by HBM budget, 4 slots."""
    rng_n = np.random.default_rng(12)
    rng_k = np.random.default_rng(13)
    sizes = np.exp(rng_n.uniform(np.log(1e4), np.log(1e6), n_models)).astype(int)
    kinds = rng_k.integers(0, 2, n_models)
    srv = pa.Server(n_slots=4, rel_accuracy=1e-6)
    ips, mids = [], []
    t0 = time.perf_counter()
    n_conv, worst_res = 0, 0.0
    for i, (n, k) in enumerate(zip(sizes, kinds)):
        p = np.random.default_rng(1000 + i).uniform(0, 1, (n, 3))
        rbf = pa.Biharmonic3D() if k == 0 else pa.Triharmonic3D()
        ip = pa.Interpolant(pa.Model(rbf, poly_degree=1))
        # keep_unconverged only so that a model short of 1e-4 after 100 steps
        # shows in models_fit_to_1e-4 instead of ending the run
        ip.fit(p, sample_f(p), 1e-4, max_iter=100, options=pa.FitOptions(keep_unconverged=True))
        st = ip.fit_stats()
        n_conv += st["converged"]
        worst_res = max(worst_res, st["residual"])
        mids.append(srv.add_model(ip))
        ips.append((ip, rbf))
    t_models = time.perf_counter() - t0
    resident = sum(srv.model_config(m)[1] for m in mids)
    log(f"serve: {n_models} models fitted and resident in {t_models:.1f} s ({resident / 2**30:.2f} GiB)")
    rng_a = np.random.default_rng(14)
    rng_q = np.random.default_rng(15)
    arrivals = np.cumsum(rng_a.exponential(1.0 / rate, n_requests))
    models = rng_q.integers(0, n_models, n_requests)
    counts = np.exp(rng_q.uniform(0.0, np.log(1e6), n_requests)).astype(int)
    queries = [rng_q.uniform(0, 1, (c, 3)) for c in counts]
    # warm up: every model once, and every slot once at the largest request size
    # (first-use device allocations), then measure
    for m in range(n_models):
        srv.wait(srv.submit(mids[m], queries[0][:1]))
    big = np.random.default_rng(16).uniform(0, 1, (1_000_000, 3))
    for r in [srv.submit(mids[int(np.argmax(sizes))], big) for _ in range(4)]:
        srv.wait(r)
    srv.reset_stats()
    rids = []
    t_start = time.perf_counter()
    for i in range(n_requests):
        delay = t_start + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        rids.append(srv.submit(mids[models[i]], queries[i]))
    outs = {}
    check = set(range(0, n_requests, max(1, n_requests // 20)))
    for i, r in enumerate(rids):
        o = srv.wait(r)
        if i in check:
            outs[i] = o
    st = srv.stats()
    # service time (admission to completion) by query-count decade
    by_decade = {}
    for i, r in enumerate(rids):
        q_ms, t_ms = srv.request_times(r)
        d = int(np.log10(max(counts[i], 1)))
        by_decade.setdefault(d, []).append(t_ms - q_ms)
    service = {f"1e{d}": {"n": len(v), "mean_ms": float(np.mean(v)), "max_ms": float(np.max(v))}
               for d, v in sorted(by_decade.items())}
    cfgs = {}
    for m in mids:
        c = srv.model_config(m)[0]
        key = f"{c.precision}{c.order}"
        cfgs[key] = cfgs.get(key, 0) + 1
    log("serve service times", json.dumps(service), "configs", json.dumps(cfgs))
    # accuracy of a sample of answers vs the FP64 direct sum on the GPU
    worst = 0.0
    for i, o in outs.items():
        ip, rbf = ips[models[i]]
        q = queries[i][:1000]
        w = ip.weights()
        n = ip.num_centers()
        ref = pa.direct_evaluate(rbf, ip.centers(), w[:n], q) + w[n] + q @ w[n + 1:n + 4]
        worst = max(worst, np.abs(o[:len(q)] - ref).max() / np.abs(ref).max())
    srv.close()
    return {"workload": f"{n_models} resident bh3/th3 models (N log-uniform 1e4..1e6), "
                        f"{n_requests} Poisson requests at {rate:g}/s, 1..1e6 query points each",
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "max_ms": st["max_ms"], "mean_queue_ms": st["mean_queue_ms"],
            "points_per_s": st["points_per_s"], "completed": st["completed"], "failed": st["failed"],
            "max_in_flight": st["max_in_flight"], "shared_direct_launches": st["batches"],
            "requests_in_shared_launches": st["batched_requests"], "max_requests_per_launch": st["max_batch"],
            "fmm_batches": st["fmm_batches"], "requests_in_fmm_batches": st["fmm_batched_requests"],
            "max_requests_per_fmm_batch": st["max_fmm_batch"],
            "resident_gib": resident / 2**30,
            "max_rel_err_sampled": worst, "models_ready_s": t_models,
            "models_fit_to_1e-4": n_conv, "worst_fit_residual": worst_res,
            "service_ms_by_size": service, "model_configs": cfgs}


def gradient_section(pa, torch, src, w, trg, sub, cfgs, steps, warmup):
    """Gradients at the 1M targets (gradient-transpose FMM) with the value
    path's configurations, against the FP64 direct gradient sum on 2,000 of
    the subsample's targets."""
    ref = pa.direct_evaluate_gradient(pa.Biharmonic3D(), src, w, trg[sub[:2000]])
    ref_inf = np.abs(ref).max()
    out = {"workload": "gradients of the biharmonic3d sum, 1,000,000 sources -> 1,000,000 targets"}
    for name, cfg in cfgs:
        de = DeviceEval(torch, src, w, trg, cfg, gradient=True)
        sec, stats = time_steps(torch, de, steps, warmup)
        g = de.output().reshape(-1, 3)
        err = float(np.abs(g[sub[:2000]] - ref).max() / ref_inf)
        keys = ("ms_tree", "ms_upward", "ms_m2l", "ms_l2l", "ms_l2p", "ms_p2p", "ms_total")
        out[name] = {"ms_per_step": sec * 1e3, "grads_per_s": N_TRG / sec, "rel_err": err,
                     "order": stats[-1]["order"],
                     "stage_ms": {k: float(np.mean([s_[k] for s_ in stats])) for k in keys}}
        log("gradient", name, json.dumps(out[name]))
        del de
        torch.cuda.synchronize()
    return out


def time_steps(torch, de, steps, warmup):
    for _ in range(warmup):
        de.step()
    de.sync()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        de.step()
        stats.append(de.ev.stats())  # HIP-event stage times of this step
    de.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--order", type=int, default=0,
                    help="mixed-form order (0: the accuracy search's choice at 1e-6 x max|y|)")
    ap.add_argument("--order64", type=int, default=0,
                    help="FP64-form order (0: the accuracy search's choice at 1e-6 x max|y|)")
    ap.add_argument("--leaf-level", type=int, default=-1)
    ap.add_argument("--m2l-dense", action="store_true",
                    help="dense M2L operators instead of the default compressed transfers (svd_tol 0)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fit", action="store_true", help="skip the 1M-point fit comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--no-kriging", action="store_true", help="skip the kriging fit and prediction")
    ap.add_argument("--no-gradient", action="store_true", help="skip the gradient evaluation")
    ap.add_argument("--no-config1", action="store_true", help="skip the 10k-site fit and evaluate")
    ap.add_argument("--no-hermite", action="store_true", help="skip the fits with gradient data")
    ap.add_argument("--serve-models", type=int, default=32)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--only", choices=("both", "mixed", "fp64", "fit"), default="both",
                    help="run one form only, or only the ellipsoid fit (profiling runs); the JSON line needs both")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized buffer 3 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB x 3")
        return
    import polatory_amd as pa

    if args.only == "fit":
        fit_section(pa)
        return
    src, w, c, trg, sub = make_inputs()
    log("inputs ready")

    # FP64 direct reference on the subsample (GPU) plus polynomial terms
    ref_rbf = pa.direct_evaluate(pa.Biharmonic3D(), src, w, trg[sub])
    ref = ref_rbf + poly(c, trg[sub])
    ref_inf = np.abs(ref).max()
    # Each form's order: the cheapest of its precision that the accuracy search
    # accepts at 1e-6 relative (absolute accuracy 1e-6 x max |y_direct|).
    acc = 1e-6 * float(np.abs(ref_rbf).max())
    cfg32 = (pa.FmmConfig("mixed", args.order, args.leaf_level, "fmm") if args.order > 0
             else search_config(pa, src, w, "mixed", acc))
    cfg64 = (pa.FmmConfig("fp64", args.order64, args.leaf_level, "fmm") if args.order64 > 0
             else search_config(pa, src, w, "fp64", acc))
    if args.m2l_dense:
        cfg32.svd_tol = cfg64.svd_tol = 0.0
    log(f"configurations: mixed {cfg32}, fp64 {cfg64} (accuracy {acc:.3e})")

    results = {}
    for name, cfg in (("mixed", cfg32), ("fp64", cfg64)):
        if args.only not in ("both", name):
            continue
        de = DeviceEval(torch, src, w, trg, cfg)
        sec, stats = time_steps(torch, de, args.steps, args.warmup)
        y = de.output() + poly(c, trg)  # polynomial terms on the host for the error check only
        err = np.abs(y[sub] - ref).max() / ref_inf
        p2p_pairs, m2l_pairs = de.ev.count_work()
        st = stats[-1]
        keys = ("ms_tree", "ms_upward", "ms_m2l", "ms_l2l", "ms_l2p", "ms_p2p", "ms_total")
        avg = {k: float(np.mean([s[k] for s in stats])) for k in keys}
        # The same steps with the stages in sequence: each stage's kernels
        # alone, for the rooflines (the timed steps above overlap the near
        # field with the far field).
        de.ev.set_overlap(False)
        sec_serial, stats_serial = time_steps(torch, de, args.steps, 1)
        de.ev.set_overlap(True)
        avg_serial = {k: float(np.mean([s[k] for s in stats_serial])) for k in keys}
        results[name] = dict(sec=sec, err=err, p2p_pairs=p2p_pairs, m2l_pairs=m2l_pairs,
                             order=st["order"], leaf_level=st["leaf_level"], m2l_rank=st["m2l_rank"],
                             n_trg_leaves=st["n_trg_leaves"], stage_ms=avg, sec_serial=sec_serial,
                             stage_ms_serial=avg_serial)
        log(name, json.dumps(results[name]))
        del de
        torch.cuda.synchronize()

    if args.only != "both":
        return
    m = results["mixed"]
    f = results["fp64"]
    p = m["order"]
    r = m["m2l_rank"]
    stage = m["stage_ms_serial"]
    # M2L algorithmic flops: rank-r transfers (2 r^2 per cell pair) when compressed,
    # dense-equivalent 2 p^6 per cell pair reported beside it.
    m2l_dense_flops = 2.0 * p**6 * m["m2l_pairs"]
    m2l_flops = 2.0 * r * r * m["m2l_pairs"] if r > 0 else m2l_dense_flops
    p2p_flops = 11.0 * m["p2p_pairs"]
    m2l_ach = m2l_flops / (stage["ms_m2l"] * 1e-3)
    p2p_ach = p2p_flops / (stage["ms_p2p"] * 1e-3)
    # L2P: FP64 locals of every target leaf + FP64 tree-order targets read, FP64 far field written
    l2p_bytes = m["n_trg_leaves"] * p**3 * 8 + N_TRG * (32 + 8)
    stage_roofs = {
        "m2l": {"kernel": ("m2l_batch_kernel<double,double> + " if r > 0 else "")
                + "m2l_pair_kernel<float,float> + m2l_pair_reduce_kernel<float>", "bound": "mfma",
                "achieved": m2l_ach / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                "frac": m2l_ach / PEAK_FP32, "flops": m2l_flops, "cell_pairs": m["m2l_pairs"]},
        "p2p": {"kernel": "p2p_jobs_kernel<float,0,false>", "bound": "valu (fp32)", "achieved": p2p_ach / 1e12,
                "peak": PEAK_FP32 / 1e12, "peak_unpacked": PEAK_FP32 / 2e12, "unit": "TFLOP/s",
                "frac": p2p_ach / PEAK_FP32, "pairs": m["p2p_pairs"], "flops_per_pair": 11},
        "l2p": {"kernel": f"l2p_jobs_kernel<double,{p}>", "bound": "hbm",
                "achieved": l2p_bytes / (stage["ms_l2p"] * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": l2p_bytes / (stage["ms_l2p"] * 1e-3) / PEAK_HBM, "bytes": l2p_bytes},
    }
    dominant = "m2l" if stage["ms_m2l"] >= stage["ms_p2p"] else "p2p"
    roof = {k: stage_roofs[dominant][k] for k in ("kernel", "bound", "achieved", "peak", "unit", "frac")}
    parts = [hbm_traffic(k) for k in stage_roofs[dominant]["kernel"].split(" + ")]
    roof["traffic"] = None if any(t is None for t in parts) else sum(parts)
    if dominant == "m2l":
        roof["dense_equivalent_tflops"] = m2l_dense_flops / (stage["ms_m2l"] * 1e-3) / 1e12

    cpu = None if args.no_cpu else cpu_fmm_baseline(src, w, trg, sub, ref_rbf, acc)

    gradient = None if args.no_gradient else gradient_section(
        pa, torch, src, w, trg, sub, (("mixed", cfg32), ("fp64", cfg64)), args.steps, args.warmup)
    fit = None if args.no_fit else fit_section(pa)
    config1 = None if args.no_config1 else config1_section(pa, args.no_cpu)
    kriging = None if args.no_kriging else kriging_section(pa)
    hermite = None if args.no_hermite else hermite_section(pa)
    serve = None if args.no_serve else serve_section(pa, args.serve_models, args.serve_requests,
                                                     args.serve_rate)

    out = {
        "metric": "FMM evals/s at 1M pts, <=1e-6 rel err (mixed vs FP64 GPU); fit s; serve p99 ms",
        "value": N_TRG / m["sec"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["sec"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP32 P2P, FP32-MFMA M2L, FP64 outputs)",
        "data": "synthetic (seeded U[0,1]^3 points, N(0,1) weights)",
        "config": {"workload": "biharmonic3d FMM evaluation, 1,000,000 sources -> 1,000,000 targets, "
                               "degree-1 polynomial",
                   "n_src": N_SRC, "n_trg": N_TRG, "order": m["order"], "leaf_level": m["leaf_level"],
                   "m2l_rank": m["m2l_rank"]},
        "rel_err": m["err"],
        "fp64_gpu": {"value": N_TRG / f["sec"], "ms_per_step": f["sec"] * 1e3, "rel_err": f["err"],
                     "order": f["order"]},
        "speedup_vs_fp64_gpu": f["sec"] / m["sec"],
        "stage_ms": m["stage_ms"],
        "stage_ms_serial": stage,
        "ms_per_step_serial": m["sec_serial"] * 1e3,
        "roofline": roof,
        "stage_rooflines": stage_roofs,
        "cpu_baseline": cpu,
        "fit": fit,
        "config1": config1,
        "kriging": kriging,
        "hermite": hermite,
        "gradient": gradient,
        "serve": serve,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
