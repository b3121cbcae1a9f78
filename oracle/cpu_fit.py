"""Reference-algorithm CPU fit -- BASELINE AND TEST INFRASTRUCTURE ONLY.

A numpy restatement of polatory's fit path (Interpolant::fit ->
Solver::solve: FGMRES right-preconditioned by the multilevel RAS
preconditioner), used to time the reference algorithm on the host cores
beside the GPU (bench.py's config-1 baseline) and checked by tests/. The
product package never imports it.

What it follows (paths relative to the polatory source tree):
  - Solver::solve (include/polatory/interpolation/solver.hpp:76-142):
    FGMRES from zero with the right preconditioner, the residual checked
    after every step, max_iter steps at most;
  - RasPreconditioner (include/polatory/preconditioner/ras_preconditioner.hpp):
    the level count and sizes (:56-73,130-140), the unisolvent polynomial
    points and Lagrange basis (:78-100), the points of every level from
    DomainDivider::choose_coarse_points (:136-141), the fine grids of each
    level and the coarse grid (:143-170), P orthonormalised and A P
    (:172-188); the application (:191-250), solve (:280-297),
    update_residuals (:299-328) and orthogonalize (:271-278);
  - DomainDivider (include/polatory/preconditioner/domain_divider.hpp):
    choose_coarse_points (:54-125) with initialize_cluster, divide_domain
    (:160-240) and divide_domains / merge_poly_points (:245-269);
  - FineGrid (include/polatory/preconditioner/fine_grid.hpp:59-142) and
    CoarseGrid (coarse_grid.hpp:40-131): Q^T A Q with Q = [-Lag^T; I],
    factored (LAPACK Cholesky in place of Eigen's LDLT), and the solves;
  - UnisolventPointSet (include/polatory/polynomial/unisolvent_point_set.hpp):
    100 random trials, the best reciprocal condition number; drawn here
    from numpy's generator, not std::mt19937 (the indices differ).
The operator and every evaluation between levels are FP64 direct sums of the
C oracle (liboracle.so, OpenMP); polatory applies its FMM at accuracy 0
there (solver.hpp:40,48), whose error is below the fit tolerance.
Only the value kernels of the oracle (bh3, th3, exp) and value data (no
gradient points) are covered.
"""
import heapq
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
from scipy.linalg import cho_factor, cho_solve
from scipy.spatial.distance import cdist
from threadpoolctl import threadpool_limits

from . import BH3, TH3, EXP, direct_evaluate, direct_operator, monomials, poly_basis_size

K_FINE_TO_COARSE = 10.0  # ras_preconditioner.hpp:52
K_COARSEST = 2048        # ras_preconditioner.hpp:53
K_OVERLAP_QUOTA = 0.5    # domain_divider.hpp:26
K_MAX_LEAF = 1024        # domain_divider.hpp:27


def _round_half_even(d):
    return math.ceil((d - 0.5) / 2.0) + math.floor((d + 0.5) / 2.0)


def kernel_matrix(rbf, pts_a, pts_b):
    """phi(A (a_i - b_j)) for one oracle rbf (closed forms of
    rbf/polyharmonic_odd.hpp:32-72 and rbf/cov_exponential.hpp:33-40)."""
    aniso = np.array(rbf.aniso[:]).reshape(3, 3)
    r = cdist(pts_a @ aniso.T, pts_b @ aniso.T)
    p0, p1 = rbf.params[0], rbf.params[1]
    if rbf.kind == BH3:
        return -p0 * np.sqrt(r * r + p1 * p1)
    if rbf.kind == TH3:
        rho = np.sqrt(r * r + p1 * p1)
        return p0 * rho ** 3
    if rbf.kind == EXP:
        return p0 * np.exp(-3.0 * r / p1)
    raise ValueError("unsupported kernel")


def mat_a(rbfs, nugget, pts):
    """preconditioner/mat_a.hpp:11-61: sum of the rbfs, nugget on the diagonal."""
    a = sum(kernel_matrix(r, pts, pts) for r in rbfs)
    a[np.diag_indices_from(a)] += nugget
    return a


def unisolvent_point_set(pts, degree, rng):
    l = poly_basis_size(degree)
    if l == 0:
        return []
    best, best_rc = None, 0.0
    for _ in range(100):
        s = set()
        while len(s) < l:
            s.add(int(rng.integers(0, len(pts))))
        idx = sorted(s)
        m = monomials(degree, pts[idx])
        if np.linalg.matrix_rank(m) < l:
            continue
        rc = 1.0 / np.linalg.cond(m, 1)
        if rc > best_rc:
            best, best_rc = idx, rc
    if best is None:
        raise RuntimeError("could not find a unisolvent set of points")
    return best


def _axes_order(pts):
    w = pts.max(axis=0) - pts.min(axis=0)
    return sorted(range(3), key=lambda i: -w[i])


def _sort_by_axes(pts, idx):
    ax = _axes_order(pts[idx])
    keys = pts[idx][:, ax]
    return np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))


def choose_coarse_points(pts, point_idcs, poly_idcs, n_coarse):
    """domain_divider.hpp:54-125 with initialize_cluster: bisect the cluster of
    the lowest level (ties: the largest bbox volume) until there are n_coarse
    clusters; a cluster's point is the one closest to its bbox centre."""
    poly = set(poly_idcs)
    root = np.array([i for i in point_idcs if i not in poly], dtype=np.int64)

    def init(points, level, seq):
        p = pts[points]
        lo, hi = p.min(axis=0), p.max(axis=0)
        c = 0.5 * (lo + hi)
        center = int(points[int(np.argmin(((p - c) ** 2).sum(axis=1)))])
        order = _sort_by_axes(pts, points)
        vol = float(np.prod(hi - lo))
        # heap key: lowest level first, then the largest volume
        return (level, -vol, seq, points[order], center)

    seq = 0
    heap = [init(root, 0, seq)]
    size = 1
    while size < n_coarse:
        level, _, _, points, _ = heapq.heappop(heap)
        n = len(points)
        # the split closest to the middle, ties to the even index
        mid = n // 2 if n % 2 == 0 else ((n - 1) // 2 if ((n - 1) // 2) % 2 == 0 else (n + 1) // 2)
        size -= 1
        for part in (points[:mid], points[mid:]):
            if len(part):
                seq += 1
                heapq.heappush(heap, init(part, level + 1, seq))
                size += 1
    out = list(poly_idcs)
    while heap:
        out.append(heapq.heappop(heap)[4])
    return out


def divide_domains(pts, point_idcs, poly_idcs):
    """domain_divider.hpp:160-269: split along the widest axes with an overlap
    of about half a leaf until every domain holds at most 1024 points, then put
    the polynomial points first in every domain."""
    work = [(np.asarray(point_idcs, dtype=np.int64), np.ones(len(point_idcs), dtype=bool))]
    out = []
    while work:
        idx, inner = work.pop(0)
        n = len(idx)
        if n <= K_MAX_LEAF:
            out.append((idx, inner))
            continue
        order = _sort_by_axes(pts, idx)
        idx, inner = idx[order], inner[order]
        q = K_OVERLAP_QUOTA * K_MAX_LEAF / n
        n_sub = int(_round_half_even((1.0 + q) / 2.0 * n))
        left_part, right_part = n - n_sub, n_sub
        mid = int(_round_half_even((left_part + right_part) / 2.0))
        pos = np.arange(n)
        work.append((idx[:right_part], inner[:right_part] & (pos[:right_part] < mid)))
        work.append((idx[left_part:], inner[left_part:] & (pos[left_part:] >= mid)))
    domains = []
    poly = list(poly_idcs)
    for idx, inner in out:
        order = np.argsort(idx, kind="stable")
        idx, inner = idx[order], inner[order]
        is_poly = np.isin(idx, poly)
        p_inner = np.zeros(len(poly), dtype=bool)
        for k, pi in enumerate(poly):
            hit = np.nonzero(idx == pi)[0]
            if len(hit):
                p_inner[k] = inner[hit[0]]
        domains.append((np.concatenate([np.array(poly, dtype=np.int64), idx[~is_poly]]),
                        np.concatenate([p_inner, inner[~is_poly]])))
    return domains


def _prefix_position(prefix, x):
    """Largest i with prefix[i] <= x (std::upper_bound - 1)."""
    return int(np.searchsorted(prefix, x, side="right")) - 1


def divide_domains_mixed(pts, gpts, point_idcs, grad_idcs, poly_idcs):
    """domain_divider.hpp:160-249 with gradient points: a gradient point counts
    3 rows in the sizes and split positions (prefix sums of the multiplicities);
    leaves hold at most 1024 rows. Returns [(idx, inner, gidx, ginner)], the
    points with the polynomial points first, the rest of each list in index
    order."""
    mu = len(pts)
    allp = np.vstack([np.asarray(pts, dtype=float).reshape(-1, 3), np.asarray(gpts, dtype=float).reshape(-1, 3)])
    rec = np.concatenate([np.asarray(point_idcs, dtype=np.int64), mu + np.asarray(grad_idcs, dtype=np.int64)])
    work = [(rec, np.ones(len(rec), dtype=bool))]
    out = []
    while work:
        r, inner = work.pop(0)
        mult = np.where(r >= mu, 3, 1)
        n_mult = int(mult.sum())
        if n_mult <= K_MAX_LEAF:
            out.append((r, inner))
            continue
        order = _sort_by_axes(allp, r)
        r, inner, mult = r[order], inner[order], mult[order]
        prefix = np.concatenate([[0], np.cumsum(mult)])
        q = K_OVERLAP_QUOTA * K_MAX_LEAF / n_mult
        n_sub = int(_round_half_even((1.0 + q) / 2.0 * n_mult))
        left_mult, right_mult = n_mult - n_sub, n_sub
        mid_mult = int(_round_half_even((left_mult + right_mult) / 2.0))
        left, right, mid = (_prefix_position(prefix, x) for x in (left_mult, right_mult, mid_mult))
        pos = np.arange(len(r))
        work.append((r[:right], inner[:right] & (pos[:right] < mid)))
        work.append((r[left:], inner[left:] & (pos[left:] >= mid)))
    domains = []
    poly = list(poly_idcs)
    for r, inner in out:
        is_g = r >= mu
        v_idx, v_in = r[~is_g], inner[~is_g]
        g_idx, g_in = r[is_g] - mu, inner[is_g]
        o = np.argsort(v_idx, kind="stable")
        v_idx, v_in = v_idx[o], v_in[o]
        o = np.argsort(g_idx, kind="stable")
        g_idx, g_in = g_idx[o], g_in[o]
        is_poly = np.isin(v_idx, poly)
        p_inner = np.zeros(len(poly), dtype=bool)
        for k, pi in enumerate(poly):
            hit = np.nonzero(v_idx == pi)[0]
            if len(hit):
                p_inner[k] = v_in[hit[0]]
        domains.append((np.concatenate([np.array(poly, dtype=np.int64), v_idx[~is_poly]]),
                        np.concatenate([p_inner, v_in[~is_poly]]), g_idx, g_in))
    return domains


def choose_coarse_points_mixed(pts, gpts, point_idcs, grad_idcs, poly_idcs, n_coarse):
    """domain_divider.hpp:54-125 with gradient points: bisect the cluster of
    the lowest level (ties: the largest bbox volume) at the middle of its rows
    until the clusters' centres count n_coarse rows (a gradient point counts
    3); a cluster's centre is its point nearest its bbox centre. Returns the
    point indices (polynomial points first) and gradient point indices."""
    mu = len(pts)
    allp = np.vstack([np.asarray(pts, dtype=float).reshape(-1, 3), np.asarray(gpts, dtype=float).reshape(-1, 3)])
    poly = set(int(i) for i in poly_idcs)
    root = np.array([i for i in point_idcs if int(i) not in poly] + [mu + int(k) for k in grad_idcs], dtype=np.int64)
    mult_of = lambda r: 3 if r >= mu else 1

    def init(rec, level, seq):
        p = allp[rec]
        lo, hi = p.min(axis=0), p.max(axis=0)
        c = 0.5 * (lo + hi)
        center = int(rec[int(np.argmin(((p - c) ** 2).sum(axis=1)))])
        vol = float(np.prod(hi - lo))
        return (level, -vol, seq, rec[_sort_by_axes(allp, rec)], center)

    total = sum(mult_of(int(r)) for r in root)
    n_coarse = min(n_coarse, total)
    seq = 0
    heap = [init(root, 0, seq)]
    size = mult_of(heap[0][4])
    while size < n_coarse:
        level, _, _, rec, center = heapq.heappop(heap)
        prefix = np.concatenate([[0], np.cumsum([mult_of(int(r)) for r in rec])])
        cs = prefix[-1]
        mid, best = 0, None
        for i in range(len(rec)):
            d = abs(2 * int(prefix[i]) - int(cs))
            if best is None or d < best or (d == best and i % 2 == 0):
                best, mid = d, i
        size -= mult_of(center)
        for part in (rec[:mid], rec[mid:]):
            if len(part):
                seq += 1
                cl = init(part, level + 1, seq)
                size += mult_of(cl[4])
                heapq.heappush(heap, cl)
    idx, gidx = list(poly_idcs), []
    for cl in heap:
        c = cl[4]
        if c >= mu:
            gidx.append(c - mu)
        else:
            idx.append(c)
    return np.array(idx, dtype=np.int64), np.array(gidx, dtype=np.int64)


class _Grid:
    """FineGrid / CoarseGrid on one domain (poly points first)."""

    def __init__(self, rbfs, nugget, degree, pts, lag, idx, inner, coarse):
        self.idx, self.inner, self.coarse, self.l = idx, inner, coarse, lag.shape[1]
        l, m = self.l, len(idx)
        a = mat_a(rbfs, nugget, pts[idx])
        if l > 0:
            self.q_top = -lag[idx[l:]].T  # l x (m - l)
            qtaq = (self.q_top.T @ a[:l, :l] @ self.q_top + self.q_top.T @ a[:l, l:] +
                    a[l:, :l] @ self.q_top + a[l:, l:])
            self.chol = cho_factor(qtaq, lower=True) if m > l else None
            if coarse:
                self.a_top = a[:l, :]
                self.p_top = monomials(degree, pts[idx[:l]])
        else:
            self.chol = cho_factor(a, lower=True)

    def _chol_solve(self, b):
        return cho_solve(self.chol, b)

    def solve(self, res, weights):
        l, idx = self.l, self.idx
        d = res[idx]
        if l > 0:
            lam = np.zeros(len(idx))
            if self.chol is not None:
                gamma = self._chol_solve(self.q_top.T @ d[:l] + d[l:])
                lam[:l] = self.q_top @ gamma
                lam[l:] = gamma
        else:
            lam = self._chol_solve(d)
        if self.coarse:
            weights[idx] = lam
            if l > 0:
                weights[-l:] = np.linalg.solve(self.p_top, d[:l] - self.a_top @ lam)
        else:
            weights[idx[self.inner]] = lam[self.inner]


class RasPreconditioner:
    def __init__(self, rbfs, degree, nugget, pts, rng, threads=None):
        self.rbfs, self.degree, self.nugget, self.pts = rbfs, degree, nugget, pts
        # the subdomains of a level are set up and solved in parallel, as the
        # reference's OpenMP loops do (ras_preconditioner.hpp:153-157,287-293);
        # LAPACK and the oracle release the interpreter lock
        self.threads = threads or os.cpu_count() or 1
        self.pool = ThreadPoolExecutor(max_workers=self.threads)
        mu = len(pts)
        self.l = l = poly_basis_size(degree)
        self.n_levels = max(int(math.ceil(math.log(mu / K_COARSEST) / math.log(K_FINE_TO_COARSE))), 0) + 1
        poly = unisolvent_point_set(pts, degree, rng) if l > 0 else []
        if l > 0:
            self.lag = monomials(degree, pts) @ np.linalg.inv(monomials(degree, pts[poly]))
        else:
            self.lag = np.zeros((mu, 0))
        top = list(poly) + [i for i in range(mu) if i not in set(poly)]
        self.point_idcs = [None] * self.n_levels
        self.point_idcs[-1] = np.array(top, dtype=np.int64)
        a_pts = pts
        if len(rbfs) == 1:
            aniso = np.array(rbfs[0].aniso[:]).reshape(3, 3)
            if not np.array_equal(aniso, np.eye(3)):
                a_pts = pts @ aniso.T
        self.fine = [[] for _ in range(self.n_levels)]
        finest = math.log(mu) / math.log(K_FINE_TO_COARSE)
        coarsest = math.log(K_COARSEST) / math.log(K_FINE_TO_COARSE)
        for level in range(self.n_levels - 1, 0, -1):
            lp = self.point_idcs[level]
            n_coarse = int(K_FINE_TO_COARSE ** (coarsest + (level - 1) * (finest - coarsest) / (self.n_levels - 1)))
            self.point_idcs[level - 1] = np.array(choose_coarse_points(a_pts, lp, poly, n_coarse), dtype=np.int64)
            doms = divide_domains(a_pts, lp, poly)
            self.fine[level] = list(self.pool.map(
                lambda d: _Grid(rbfs, nugget, degree, pts, self.lag, d[0], d[1], False), doms))
        c = self.point_idcs[0]
        self.coarse = _Grid(rbfs, nugget, degree, pts, self.lag, c, np.ones(len(c), dtype=bool), True)
        if self.n_levels > 1 and l > 0:
            p, _ = np.linalg.qr(monomials(degree, pts))
            self.p = p
            self.ap = np.stack([direct_evaluate(rbfs, degree, pts, np.concatenate([p[:, i], np.zeros(l)]), pts)
                                for i in range(l)], axis=1)

    def _solve(self, level, res):
        w = np.zeros(len(self.pts) + self.l)
        if level == 0:
            self.coarse.solve(res, w)
        else:
            # each point is inner to one domain: the writes do not overlap
            list(self.pool.map(lambda g: g.solve(res, w), self.fine[level]))
        return w

    def _update(self, src, trg, w, res):
        s, t = self.point_idcs[src], self.point_idcs[trg]
        sw = np.concatenate([w[s], w[len(self.pts):]])
        res[t] -= direct_evaluate(self.rbfs, self.degree, self.pts[s], sw, self.pts[t])

    def _orthogonalize(self, w, res):
        if self.l > 0:
            mu = len(self.pts)
            dot = self.p.T @ w[:mu]
            w[:mu] -= self.p @ dot
            res += self.ap @ dot

    def apply(self, v):
        mu, L = len(self.pts), self.n_levels
        res = v[:mu].copy()
        if L == 1:
            return self._solve(0, res)
        total = np.zeros(mu + self.l)
        w = self._solve(0, res)
        self._update(0, L - 1, w, res)
        total += w
        for level in range(1, L - 1):
            w = self._solve(level, res)
            self._update(level, L - 1, w, res)
            total += w
            self._orthogonalize(total, res)
            w = self._solve(0, res)
            self._update(0, L - 1, w, res)
            total += w
        for level in range(L - 1, 0, -1):
            w = self._solve(level, res)
            self._update(level, level - 1, w, res)
            total += w
            self._orthogonalize(total, res)
            w = self._solve(0, res)
            if level > 1:
                self._update(0, level - 1, w, res)
            total += w
        return total


def cpu_fit(rbfs, degree, nugget, pts, values, tolerance, max_iter=100, seed=0, threads=None):
    """Interpolant::fit on the CPU: returns (weights [w; c], stats)."""
    # one BLAS thread per subdomain task: the parallelism is over subdomains
    with threadpool_limits(limits=1, user_api="blas"):
        return _cpu_fit(rbfs, degree, nugget, pts, values, tolerance, max_iter, seed, threads)


def _cpu_fit(rbfs, degree, nugget, pts, values, tolerance, max_iter, seed, threads):
    t0 = time.perf_counter()
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    values = np.asarray(values, dtype=np.float64)
    mu, l = len(pts), poly_basis_size(degree)
    pc = RasPreconditioner(rbfs, degree, nugget, pts, np.random.default_rng(seed), threads)
    t_setup = time.perf_counter() - t0
    m = mu + l
    rhs = np.concatenate([values, np.zeros(l)])
    sample = np.arange(0, mu, max(1, mu // 1024))[:1024]

    def residual_max(x, idx=None):
        # residual_evaluator.hpp:56-108 (FP64 direct; the sample first)
        trg = pts if idx is None else pts[idx]
        y = direct_evaluate(rbfs, degree, pts, x, trg)
        y += nugget * (x[:mu] if idx is None else x[idx])
        return np.abs((values if idx is None else values[idx]) - y).max()

    beta = np.linalg.norm(rhs)
    v = [rhs / beta]
    z = []
    h = np.zeros((max_iter + 1, max_iter))
    cs, sn = np.zeros(max_iter), np.zeros(max_iter)
    g = np.zeros(max_iter + 1)
    g[0] = beta
    x = np.zeros(m)
    it = 0
    res = np.inf
    converged = False
    while True:
        if it > 0:
            y = np.linalg.solve(np.triu(h[:it, :it]), g[:it])
            x = np.stack(z, axis=1) @ y
        r = residual_max(x, sample)
        if r <= tolerance:
            r = residual_max(x)
        res = r
        if r <= tolerance:
            converged = True
            break
        if it == max_iter:
            break
        zj = pc.apply(v[it])
        z.append(zj)
        w = direct_operator(rbfs, degree, nugget, pts, zj)
        for i in range(it + 1):  # modified Gram-Schmidt (gmres_base.cpp)
            h[i, it] = w @ v[i]
            w = w - h[i, it] * v[i]
        h[it + 1, it] = np.linalg.norm(w)
        v.append(w / h[it + 1, it])
        for i in range(it):  # Givens rotations (gmres.cpp:32-49)
            t = cs[i] * h[i, it] + sn[i] * h[i + 1, it]
            h[i + 1, it] = -sn[i] * h[i, it] + cs[i] * h[i + 1, it]
            h[i, it] = t
        d = math.hypot(h[it, it], h[it + 1, it])
        cs[it], sn[it] = h[it, it] / d, h[it + 1, it] / d
        h[it, it] = d
        h[it + 1, it] = 0.0
        g[it + 1] = -sn[it] * g[it]
        g[it] = cs[it] * g[it]
        it += 1
    pc.pool.shutdown()
    return x, {"iterations": it, "residual": float(res), "converged": converged, "n_levels": pc.n_levels,
               "setup_s": t_setup, "seconds": time.perf_counter() - t0, "threads": pc.threads}
