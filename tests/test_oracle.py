"""Pins the CPU oracle: golden vectors from closed forms (tests/golden/), the
invariants the reference's own tests check, and brute force on tiny inputs."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "closed_forms.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


def _rbf(orc, name, aniso=None):
    table = {
        "bh3": (orc.BH3, [1.0, 0.0]), "bh3c": (orc.BH3, [2.5, 0.3]),
        "th3": (orc.TH3, [1.0, 0.0]), "th3c": (orc.TH3, [0.7, 0.2]),
        "exp": (orc.EXP, [1.0, 2.0]), "exp2": (orc.EXP, [0.4, 0.6]),
    }
    kind, params = table[name]
    return orc.make_rbf(kind, params, aniso)


@pytest.mark.parametrize("name", ["bh3", "bh3c", "th3", "th3c", "exp", "exp2"])
@pytest.mark.parametrize("variant", ["iso", "aniso"])
def test_kernel_values_match_closed_forms(orc, g, name, variant):
    aniso = None if variant == "iso" else g["aniso"]
    r = _rbf(orc, name, aniso)
    got = np.array([orc.rbf_evaluate(r, d) for d in g["d"]])
    np.testing.assert_allclose(got, g[f"{name}_{variant}"], rtol=1e-14, atol=1e-15)


def test_bh3_sign_is_negative_r(orc):
    # polyharmonic_odd.hpp:21: kSign = -1 for K = 1
    assert orc.rbf_evaluate(_rbf(orc, "bh3"), [3.0, 4.0, 0.0]) == -5.0
    assert orc.rbf_evaluate(_rbf(orc, "th3"), [3.0, 4.0, 0.0]) == 125.0


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_monomials(orc, g, deg):
    np.testing.assert_array_equal(orc.monomials(deg, g["mono_pts"]), g[f"mono_d{deg}"])
    assert orc.poly_basis_size(deg) == {0: 1, 1: 4, 2: 10}[deg]
    assert orc.poly_basis_size(-1) == 0


def test_direct_evaluate_golden(orc, g):
    w = np.concatenate([g["de_w"], g["de_c"]])
    y = orc.direct_evaluate(_rbf(orc, "bh3"), 1, g["de_src"], w, g["de_trg"])
    np.testing.assert_allclose(y, g["de_bh3_y"], rtol=1e-13, atol=1e-12)
    y = orc.direct_evaluate(_rbf(orc, "th3", g["aniso"]), 1, g["de_src"], w, g["de_trg"])
    np.testing.assert_allclose(y, g["de_th3_aniso_y"], rtol=1e-13, atol=1e-12)
    w0 = np.concatenate([g["de_w"], g["de_c"][:1]])
    y = orc.direct_evaluate(_rbf(orc, "exp"), 0, g["de_src"], w0, g["de_trg"])
    np.testing.assert_allclose(y, g["de_exp_y"], rtol=1e-13, atol=1e-12)


def test_direct_operator_golden(orc, g):
    y = orc.direct_operator(_rbf(orc, "bh3"), 1, 0.01, g["op_pts"], g["op_x"])
    np.testing.assert_allclose(y, g["op_y"], rtol=1e-13, atol=1e-12)


def test_make_box_golden(orc, g):
    c, w = orc.make_box(g["box_min"], g["box_max"], g["aniso"])
    np.testing.assert_array_equal(c, g["box_center"])
    assert w == g["box_width"][0]
    # degenerate bbox -> width 1 (src/fmm/utility.hpp:25-27)
    c, w = orc.make_box([1, 1, 1], [1, 1, 1])
    assert w == 1.0


def _kernel_matrix(orc, r, pts):
    n = pts.shape[0]
    return np.array([[orc.rbf_evaluate(r, pts[i] - pts[j]) for j in range(n)] for i in range(n)])


def test_conditional_positive_definiteness_fixes_the_signs(orc):
    # A wrong sign on bh3 or th3 turns these quadratic forms negative.
    rng = np.random.default_rng(1)
    pts = rng.uniform(0, 1, (40, 3))
    P1 = np.concatenate([np.ones((40, 1)), pts], 1)
    for name, P in (("bh3", P1[:, :1]), ("th3", P1)):
        A = _kernel_matrix(orc, _rbf(orc, name), pts)
        q, _ = np.linalg.qr(P, mode="complete")
        Z = q[:, P.shape[1]:]
        ev = np.linalg.eigvalsh(Z.T @ A @ Z)
        assert ev.min() > 0, name
    A = _kernel_matrix(orc, _rbf(orc, "exp"), pts)
    assert np.linalg.eigvalsh(A).min() > 0


def test_homogeneity_and_anisotropy(orc):
    rng = np.random.default_rng(2)
    a = np.array([[1.5, 0.2, 0.0], [0.1, 0.8, 0.0], [0.0, 0.3, 1.1]])
    for _ in range(10):
        d = rng.normal(size=3)
        lam = rng.uniform(0.1, 5)
        bh, th = _rbf(orc, "bh3"), _rbf(orc, "th3")
        assert np.isclose(orc.rbf_evaluate(bh, lam * d), lam * orc.rbf_evaluate(bh, d), rtol=1e-14)
        assert np.isclose(orc.rbf_evaluate(th, lam * d), lam**3 * orc.rbf_evaluate(th, d), rtol=1e-13)
        for name in ("bh3", "th3", "exp"):
            ra = _rbf(orc, name, a)
            assert np.isclose(orc.rbf_evaluate(ra, d), orc.rbf_evaluate(_rbf(orc, name), a @ d),
                              rtol=1e-14)


def test_direct_operator_is_symmetric(orc):
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (50, 3))
    x, y = rng.normal(size=54), rng.normal(size=54)
    r = _rbf(orc, "th3")
    ax = orc.direct_operator(r, 1, 0.1, pts, x)
    ay = orc.direct_operator(r, 1, 0.1, pts, y)
    assert np.isclose(y @ ax, x @ ay, rtol=1e-12)


def _interleave(ix, iy, iz, level):
    c = 0
    for b in range(level):
        c |= ((ix >> b) & 1) << (3 * b + 2) | ((iy >> b) & 1) << (3 * b + 1) | ((iz >> b) & 1) << (3 * b)
    return c


def test_morton_codes_brute_force(orc):
    rng = np.random.default_rng(4)
    level = 4
    pts = rng.uniform(0, 1, (300, 3))
    pts[0] = [0.0, 0.0, 0.0]
    pts[1] = [1.0, 1.0, 1.0]  # on the upper face: clamped to the last cell
    lo = np.array([0.0, 0.0, 0.0])
    inv_cell = 2.0**level / 1.0
    codes = orc.morton_codes(pts, lo, inv_cell, level)
    for p, c in zip(pts, codes):
        ijk = [min(int(np.floor((p[k] - lo[k]) * inv_cell)), 2**level - 1) for k in range(3)]
        assert c == _interleave(*ijk, level)
    perm = orc.stable_sort_codes(codes)
    np.testing.assert_array_equal(perm, np.argsort(codes, kind="stable"))


def _coords(code, level):
    x = y = z = 0
    for b in range(level):
        x |= ((code >> (3 * b + 2)) & 1) << b
        y |= ((code >> (3 * b + 1)) & 1) << b
        z |= ((code >> (3 * b)) & 1) << b
    return np.array([x, y, z])


def test_interaction_lists_brute_force(orc):
    rng = np.random.default_rng(5)
    level = 3
    for trial in range(3):
        # sparse random occupancy: tests missing neighbours too
        tc = np.unique(rng.integers(0, 8**level, 60).astype(np.uint64))
        sc = np.unique(rng.integers(0, 8**level, 80).astype(np.uint64))
        p2p = orc.p2p_lists(tc, sc, level)
        m2l = orc.m2l_lists(tc, sc, level)
        sco = [_coords(int(c), level) for c in sc]
        for t, code in enumerate(tc):
            ct = _coords(int(code), level)
            exp_p2p = {}
            exp_m2l = {}
            for si, cs in enumerate(sco):
                d = cs - ct
                if np.abs(d).max() <= 1:
                    exp_p2p[(d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)] = si
                elif np.abs((cs >> 1) - (ct >> 1)).max() <= 1:
                    exp_m2l[(d[0] + 3) * 49 + (d[1] + 3) * 7 + (d[2] + 3)] = si
            got_p2p = {s: int(v) for s, v in enumerate(p2p[t]) if v >= 0}
            got_m2l = {s: int(v) for s, v in enumerate(m2l[t]) if v >= 0}
            assert got_p2p == exp_p2p
            assert got_m2l == exp_m2l
            assert len(got_m2l) <= 189


def test_leaves(orc):
    codes = np.array([3, 3, 5, 9, 9, 9, 12], dtype=np.uint64)
    lc, ls = orc.leaves(codes)
    np.testing.assert_array_equal(lc, [3, 5, 9, 12])
    np.testing.assert_array_equal(ls, [0, 2, 3, 6, 7])


def test_fgmres_invariants(orc):
    # test/krylov/test_krylov.cpp:83-113: Givens residual == true residual,
    # strictly decreasing; exact at k = n.
    rng = np.random.default_rng(6)
    n = 100
    b = rng.normal(size=(n, n))
    a = b + b.T + n * np.eye(n) * 0.1
    pert = rng.normal(size=(n, n)) * 1e-2
    m = np.linalg.inv(a) + pert
    rhs = rng.normal(size=n)
    for k in (5, 20, 40):
        x, hist, it = orc.fgmres_dense(a, rhs, k, m=m)
        true = np.linalg.norm(rhs - a @ x)
        assert abs(hist[-1] - true) <= 1e-12 * np.linalg.norm(rhs) + 1e-12 * true
        assert np.all(np.diff(hist) < 0)
    x, hist, it = orc.fgmres_dense(a, rhs, n)
    np.testing.assert_allclose(x, np.linalg.solve(a, rhs), rtol=1e-8, atol=1e-10)


def test_fgmres_planted_sign_error_is_caught(orc):
    # A classical-vs-modified Gram-Schmidt slip or a sign error in the Givens
    # update breaks the residual identity checked above; verify the check bites.
    rng = np.random.default_rng(7)
    n = 30
    a = rng.normal(size=(n, n)) + n * np.eye(n)
    rhs = rng.normal(size=n)
    x, hist, _ = orc.fgmres_dense(a, rhs, 10)
    wrong = x.copy()
    wrong[0] += 1e-6
    assert abs(hist[-1] - np.linalg.norm(rhs - a @ wrong)) > 1e-12 * np.linalg.norm(rhs)


# ---- the reference-algorithm CPU FMM (oracle/cpu_fmm.c), the CPU baseline ----

@pytest.fixture(scope="module")
def fmm_case(orc):
    rng = np.random.default_rng(91)
    src = rng.uniform(0, 1, (12000, 3))
    w = rng.normal(size=12000)
    trg = rng.uniform(0, 1, (3000, 3))
    return src, w, trg


@pytest.mark.parametrize("kind", ["bh3", "th3", "exp"])
def test_cpu_fmm_converges_to_direct_sum(orc, fmm_case, kind):
    """Against the brute-force FP64 sum: the error falls by at least 10x per two
    orders and is within 1e-6 relative at order 8, the first order polatory's
    accuracy search tries (fmm_accuracy_estimator.hpp:105-108). A dropped
    transfer, a wrong DFT index or sign, or a misplaced node fails this."""
    src, w, trg = fmm_case
    aniso = np.array([[1.0, 0.1, 0.0], [0.0, 0.8, 0.2], [0.0, 0.0, 1.3]])
    r = orc.make_rbf({"bh3": orc.BH3, "th3": orc.TH3, "exp": orc.EXP}[kind], None, aniso)
    ref = orc.direct_rbf_sum(r, src, w, trg)
    errs = []
    for order in (4, 6, 8):
        y, st = orc.cpu_fmm_evaluate(r, src, w, trg, ((0, 0, 0), (1, 1, 1)), order)
        assert st["height"] == 5 and st["m2l_pairs"] > 0 and st["p2p_pairs"] > 0
        errs.append(np.abs(y - ref).max() / np.abs(ref).max())
    assert errs[1] < errs[0] / 10 and errs[2] < errs[1] / 10, errs
    assert errs[2] < 1e-6, errs


def test_cpu_fmm_height_and_direct_threshold(orc, fmm_case):
    """Height max(2, round(log8 max(N, M))) (src/fmm/utility.hpp:12-17); below
    1024^2 pairs the direct sum (fmm_evaluator.hpp:227-234), bit for bit; at
    height 2 every leaf neighbours every other, so the FMM is the near field
    alone and equals the direct sum up to the summation order."""
    src, w, trg = fmm_case
    r = orc.make_rbf(orc.BH3)
    y, st = orc.cpu_fmm_evaluate(r, src[:1000], w[:1000], trg[:1000], ((0, 0, 0), (1, 1, 1)), 8)
    assert st["height"] == 0
    np.testing.assert_array_equal(y, orc.direct_rbf_sum(r, src[:1000], w[:1000], trg[:1000]))
    y, st = orc.cpu_fmm_evaluate(r, src, w, trg, ((0, 0, 0), (1, 1, 1)), 8, height=2)
    ref = orc.direct_rbf_sum(r, src, w, trg)
    assert st["m2l_pairs"] == 0
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    for n, h in ((4097, 4), (300_000, 6), (1_000_000, 7)):
        assert max(2, round(np.log(n) / np.log(8))) == h


def test_cpu_fmm_floater_hormann(orc, fmm_case):
    """The fit operator's interpolant (order 12, Floater-Hormann degree 8,
    solver.hpp:40 with fmm_accuracy_estimator.hpp:80-82) is accurate, and degree
    order - 1 is the classic Lagrange interpolant (same bits)."""
    src, w, trg = fmm_case
    r = orc.make_rbf(orc.BH3)
    ref = orc.direct_rbf_sum(r, src, w, trg)
    y, _ = orc.cpu_fmm_evaluate(r, src, w, trg, ((0, 0, 0), (1, 1, 1)), 12, fh_degree=8)
    assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-8
    a, _ = orc.cpu_fmm_evaluate(r, src, w, trg, ((0, 0, 0), (1, 1, 1)), 6, fh_degree=5)
    b, _ = orc.cpu_fmm_evaluate(r, src, w, trg, ((0, 0, 0), (1, 1, 1)), 6, fh_degree=-1)
    np.testing.assert_array_equal(a, b)


# ---- the reference-algorithm CPU fit (oracle/cpu_fit.py), the config-1 CPU baseline ----

def test_cpu_fit_kernel_matrix_matches_closed_forms(orc):
    """The baseline's kernel matrices equal the oracle's pinned closed forms,
    anisotropy included (a transposed anisotropy or a sign slip fails)."""
    from oracle import cpu_fit
    rng = np.random.default_rng(93)
    a, b = rng.uniform(0, 1, (7, 3)), rng.uniform(0, 1, (5, 3))
    aniso = np.array([[1.0, 0.2, 0.0], [0.0, 0.7, 0.3], [0.1, 0.0, 1.4]])
    for kind, params in ((orc.BH3, [1.3, 0.0]), (orc.TH3, [0.7, 0.1]), (orc.EXP, [2.0, 0.8])):
        r = orc.make_rbf(kind, params, aniso)
        k = cpu_fit.kernel_matrix(r, a, b)
        ref = np.array([[orc.rbf_evaluate(r, a[i] - b[j]) for j in range(5)] for i in range(7)])
        np.testing.assert_allclose(k, ref, rtol=1e-13, atol=1e-15)


def test_cpu_fit_single_level_is_exact(orc):
    """Below 2048 points the preconditioner is the coarse grid alone, the exact
    solve (coarse_grid.hpp:103-131): its weights interpolate the data."""
    from oracle import cpu_fit
    rng = np.random.default_rng(94)
    pts = rng.uniform(0, 1, (1500, 3))
    v = np.sin(3 * pts[:, 0]) * np.cos(2 * pts[:, 1]) + 0.5 * pts[:, 2] ** 2
    r = orc.make_rbf(orc.BH3)
    pc = cpu_fit.RasPreconditioner([r], 1, 0.0, pts, np.random.default_rng(0))
    assert pc.n_levels == 1
    w = pc.apply(np.concatenate([v, np.zeros(4)]))
    pc.pool.shutdown()
    y = orc.direct_evaluate([r], 1, pts, w, pts)
    assert np.abs(y - v).max() <= 1e-9 * np.abs(v).max()
    # the polynomial part annihilates the weights: P^T w = 0
    assert np.abs(orc.monomials(1, pts).T @ w[:1500]).max() <= 1e-8 * np.abs(w).max()


def test_cpu_fit_two_levels_reaches_tolerance(orc):
    """3,000 points (two levels: fine domains of at most 1,024 points and the
    2,048-point coarse grid): the fit reaches 1e-4 on the oracle's FP64
    residual within 10 steps, and every point is the inner point of exactly
    one fine domain."""
    from oracle import cpu_fit
    rng = np.random.default_rng(95)
    pts = rng.uniform(0, 1, (3000, 3))
    v = np.sin(3 * pts[:, 0]) * np.cos(2 * pts[:, 1]) + 0.5 * pts[:, 2] ** 2
    r = orc.make_rbf(orc.BH3)
    w, st = cpu_fit.cpu_fit([r], 1, 0.0, pts, v, 1e-4, max_iter=30, threads=4)
    assert st["n_levels"] == 2 and st["converged"] and st["iterations"] <= 10
    y = orc.direct_evaluate([r], 1, pts, w, pts)
    assert np.abs(y - v).max() <= 1e-4
    doms = cpu_fit.divide_domains(pts, np.arange(3000), [0, 1, 2, 3])
    inner = np.zeros(3000, dtype=int)
    for idx, inn in doms:
        assert len(idx) <= 1024 + 4
        np.add.at(inner, idx[inn], 1)
    assert np.all(inner == 1)  # every point is the inner point of exactly one domain


# ---- The operator of fits with gradient data (interpolation/operator.hpp:56-82) ----

def _hermite_case(orc, seed=41, mu=12, sigma=9, degree=1):
    rng = np.random.default_rng(seed)
    aniso = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.1, 0.0, 1.2]])
    rbfs = [_rbf(orc, "th3", aniso), _rbf(orc, "bh3c")]
    pts = rng.uniform(-1, 1, (mu, 3))
    gpts = rng.uniform(-1, 1, (sigma, 3))
    return rng, rbfs, pts, gpts, degree


def _hermite_matrix(orc, rbfs, degree, nugget, pts, gpts):
    n = len(pts) + 3 * len(gpts) + orc.poly_basis_size(degree)
    eye = np.eye(n)
    return np.stack([orc.hermite_operator(rbfs, degree, nugget, pts, gpts, eye[:, j]) for j in range(n)], axis=1)


def test_hermite_operator_without_gradient_points_is_the_value_operator(orc):
    rng, rbfs, pts, _, deg = _hermite_case(orc)
    x = rng.normal(size=len(pts) + orc.poly_basis_size(deg))
    y = orc.hermite_operator(rbfs, deg, 0.01, pts, np.zeros((0, 3)), x)
    ref = orc.direct_operator(rbfs, deg, 0.01, pts, x)
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_hermite_operator_is_symmetric(orc):
    # The Gram matrix of the functionals {delta_p, d/dq_c delta_q} is symmetric;
    # a sign error in either mixed block, or in P_g against P_g^T, breaks it.
    _, rbfs, pts, gpts, deg = _hermite_case(orc)
    a = _hermite_matrix(orc, rbfs, deg, 0.01, pts, gpts)
    assert np.abs(a - a.T).max() <= 1e-12 * np.abs(a).max()


def test_hermite_operator_blocks_are_derivatives_of_the_value_field(orc):
    # Each block against central differences of the value kernel:
    #   gradient rows from value weights = gradient of sum_j w_j phi(q - p_j) at q_k;
    #   value rows from gradient weights = d/de sum_k phi(p_i - (q_k + e g_k)) at e = 0;
    #   gradient rows from gradient weights = gradient at q_k of the latter field;
    #   polynomial gradient rows = differences of the monomials.
    rng, rbfs, pts, gpts, deg = _hermite_case(orc, mu=10, sigma=7)
    mu, sigma, l = len(pts), len(gpts), orc.poly_basis_size(deg)
    n = mu + 3 * sigma
    h = 1e-5
    w = rng.normal(size=mu)
    g = rng.normal(size=(sigma, 3))
    c = rng.normal(size=l)

    def field(src, wts, trg):  # sum over rbfs of sum_j wts_j phi(trg - src_j)
        return sum(orc.direct_rbf_sum(r, src, wts, trg) for r in rbfs)

    y = orc.hermite_operator(rbfs, deg, 0.0, pts, gpts, np.concatenate([w, np.zeros(3 * sigma + l)]))
    fd = np.stack([(field(pts, w, gpts + h * e) - field(pts, w, gpts - h * e)) / (2 * h) for e in np.eye(3)], axis=1)
    np.testing.assert_allclose(y[mu:n].reshape(-1, 3), fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())

    y = orc.hermite_operator(rbfs, deg, 0.0, pts, gpts, np.concatenate([np.zeros(mu), g.ravel(), np.zeros(l)]))
    fd = (field(gpts + h * g, np.ones(sigma), pts) - field(gpts - h * g, np.ones(sigma), pts)) / (2 * h)
    np.testing.assert_allclose(y[:mu], fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())

    def vfield(trg):  # value rows from the gradient weights, at arbitrary points
        x = np.concatenate([np.zeros(len(trg)), g.ravel(), np.zeros(l)])
        return orc.hermite_operator(rbfs, deg, 0.0, trg, gpts, x)[:len(trg)]

    fd = np.stack([(vfield(gpts + h * e) - vfield(gpts - h * e)) / (2 * h) for e in np.eye(3)], axis=1)
    np.testing.assert_allclose(y[mu:n].reshape(-1, 3), fd, rtol=1e-4, atol=1e-4 * np.abs(fd).max())

    y = orc.hermite_operator(rbfs, deg, 0.0, pts, gpts, np.concatenate([np.zeros(n), c]))
    mono = lambda q: orc.monomials(deg, q) @ c
    fd = np.stack([(mono(gpts + h * e) - mono(gpts - h * e)) / (2 * h) for e in np.eye(3)], axis=1)
    np.testing.assert_allclose(y[mu:n].reshape(-1, 3), fd, rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(y[:mu], mono(pts), rtol=1e-12, atol=1e-12)


def test_hermite_operator_nugget_only_on_value_rows(orc):
    rng, rbfs, pts, gpts, deg = _hermite_case(orc)
    n = len(pts) + 3 * len(gpts) + orc.poly_basis_size(deg)
    x = rng.normal(size=n)
    d = orc.hermite_operator(rbfs, deg, 0.25, pts, gpts, x) - orc.hermite_operator(rbfs, deg, 0.0, pts, gpts, x)
    np.testing.assert_allclose(d[:len(pts)], 0.25 * x[:len(pts)], rtol=1e-12, atol=1e-12)
    assert np.all(d[len(pts):] == 0.0)
