/*
 * CPU oracle for the RBF fit/evaluate hot path — TEST INFRASTRUCTURE ONLY.
 * Never part of the product: only tests, the smoke check and the CPU-baseline leg
 * of the benchmark load it. Built with -ffp-contract=off so that every FP64
 * operation is one IEEE operation (needed for the bit-exact tree comparison).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- RBF kernels (rbf/polyharmonic_odd.hpp:32-43,120-136; rbf/cov_exponential.hpp:33-40) */

static double phi_iso(const orc_rbf* rbf, double r2) {
  switch (rbf->kind) {
    case ORC_BH3: { /* PolyharmonicOdd<3,1>: kSign = -1 */
      double rho = sqrt(r2 + rbf->params[1] * rbf->params[1]);
      return -rbf->params[0] * rho;
    }
    case ORC_TH3: { /* PolyharmonicOdd<3,3>: kSign = +1 */
      double rho = sqrt(r2 + rbf->params[1] * rbf->params[1]);
      return rbf->params[0] * rho * rho * rho;
    }
    case ORC_EXP: { /* psill * exp(-3 r / range) */
      double r = sqrt(r2);
      return rbf->params[0] * exp(-3.0 * (r / rbf->params[1]));
    }
  }
  return NAN;
}

static void apply_aniso(const double* a, double x, double y, double z, double* o) {
  o[0] = (a[0] * x + a[1] * y) + a[2] * z;
  o[1] = (a[3] * x + a[4] * y) + a[5] * z;
  o[2] = (a[6] * x + a[7] * y) + a[8] * z;
}

double orc_rbf_evaluate(const orc_rbf* rbf, double dx, double dy, double dz) {
  double t[3];
  apply_aniso(rbf->aniso, dx, dy, dz, t);
  return phi_iso(rbf, t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
}

/* Gradient and Hessian (rbf/polyharmonic_odd.hpp:45-72, rbf/cov_exponential.hpp:42-62,
 * rbf/rbf_base.hpp:45-53): g = A^T g_iso(A d), H = A^T H_iso(A d) A, with
 * g_iso = f u and H_iso = f (I + t u u^T) for the factor f and coefficient t
 * of each kernel. */
static double grad_factor(const orc_rbf* rbf, const double* u, double* t) {
  double q = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  switch (rbf->kind) {
    case ORC_BH3:
    case ORC_TH3: {
      int k = rbf->kind == ORC_BH3 ? 1 : 3;
      double sign = rbf->kind == ORC_BH3 ? -1.0 : 1.0;
      double rho2 = q + rbf->params[1] * rbf->params[1];
      double rho = sqrt(rho2);
      if (rho == 0.0) {
        *t = 0.0;
        return 0.0;
      }
      *t = (k - 2) / rho2;
      return sign * k * rbf->params[0] * (k == 1 ? 1.0 / rho : rho);
    }
    case ORC_EXP: {
      double r = sqrt(q), range = rbf->params[1];
      *t = -(1.0 / (r * r) + 3.0 / (range * r));
      return -3.0 * rbf->params[0] * exp(-3.0 * (r / range)) / (range * r);
    }
  }
  *t = NAN;
  return NAN;
}

void orc_rbf_gradient(const orc_rbf* rbf, double dx, double dy, double dz, double* g) {
  const double* a = rbf->aniso;
  double u[3], t;
  apply_aniso(a, dx, dy, dz, u);
  double f = grad_factor(rbf, u, &t);
  double gi[3] = {f * u[0], f * u[1], f * u[2]};
  for (int k = 0; k < 3; k++) g[k] = (a[k] * gi[0] + a[3 + k] * gi[1]) + a[6 + k] * gi[2];
}

void orc_rbf_hessian(const orc_rbf* rbf, double dx, double dy, double dz, double* h) {
  const double* a = rbf->aniso;
  double u[3], t, hi[9], ha[9];
  apply_aniso(a, dx, dy, dz, u);
  double f = grad_factor(rbf, u, &t);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) hi[3 * i + j] = f * ((i == j ? 1.0 : 0.0) + t * u[i] * u[j]);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) ha[3 * r + c] = (hi[3 * r] * a[c] + hi[3 * r + 1] * a[3 + c]) + hi[3 * r + 2] * a[6 + c];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) h[3 * r + c] = (a[r] * ha[c] + a[3 + r] * ha[3 + c]) + a[6 + r] * ha[6 + c];
}

/* Gradient of sum_j w_j phi(t - s_j) at each target (full_direct.hpp:7-52 with
 * the gradient kernel), summed over sources in index order. */
void orc_direct_gradient_sum(const orc_rbf* rbf, const double* src, int64_t n_src, const double* w,
                             const double* trg, int64_t n_trg, double* out) {
  for (int64_t i = 0; i < n_trg; i++) {
    double acc[3] = {0.0, 0.0, 0.0}, g[3];
    for (int64_t j = 0; j < n_src; j++) {
      orc_rbf_gradient(rbf, trg[3 * i] - src[3 * j], trg[3 * i + 1] - src[3 * j + 1], trg[3 * i + 2] - src[3 * j + 2], g);
      for (int k = 0; k < 3; k++) acc[k] += w[j] * g[k];
    }
    for (int k = 0; k < 3; k++) out[3 * i + k] = acc[k];
  }
}

void orc_transform_points(const double* aniso, const double* pts, int64_t n, double* out) {
  for (int64_t i = 0; i < n; i++) apply_aniso(aniso, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], out + 3 * i);
}

/* ---- Polynomials (polynomial/polynomial_basis_base.hpp:26-45, monomial_basis.hpp:198-247) */

int64_t orc_poly_basis_size(int degree) {
  if (degree < 0) return 0;
  int64_t k = degree + 1;
  return k * (k + 1) * (k + 2) / 6;
}

void orc_monomials(int degree, const double* pts, int64_t n, double* out) {
  int64_t l = orc_poly_basis_size(degree);
  for (int64_t i = 0; i < n; i++) {
    double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    double m[10] = {1.0, x, y, z, x * x, x * y, x * z, y * y, y * z, z * z};
    for (int64_t k = 0; k < l; k++) out[i * l + k] = m[k];
  }
}

/* ---- Direct summation (interpolation/direct_evaluator.hpp:45-79, src/fmm/full_direct.hpp) */

void orc_direct_rbf_sum(const orc_rbf* rbf, const double* src, int64_t n_src, const double* w,
                        const double* trg, int64_t n_trg, double* out) {
  double* ts = (double*)malloc(sizeof(double) * 3 * (size_t)(n_src > 0 ? n_src : 1));
  orc_transform_points(rbf->aniso, src, n_src, ts);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_trg; i++) {
    double t[3];
    apply_aniso(rbf->aniso, trg[3 * i], trg[3 * i + 1], trg[3 * i + 2], t);
    double acc = 0.0;
    for (int64_t j = 0; j < n_src; j++) {
      double dx = t[0] - ts[3 * j], dy = t[1] - ts[3 * j + 1], dz = t[2] - ts[3 * j + 2];
      acc += w[j] * phi_iso(rbf, dx * dx + dy * dy + dz * dz);
    }
    out[i] = acc;
  }
  free(ts);
}

void orc_direct_evaluate(const orc_rbf* rbfs, int n_rbfs, int poly_degree, const double* src,
                         int64_t n_src, const double* weights, const double* trg, int64_t n_trg,
                         double* out) {
  int64_t l = orc_poly_basis_size(poly_degree);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)(n_trg > 0 ? n_trg : 1));
  for (int64_t i = 0; i < n_trg; i++) out[i] = 0.0;
  for (int r = 0; r < n_rbfs; r++) {
    orc_direct_rbf_sum(&rbfs[r], src, n_src, weights, trg, n_trg, tmp);
    for (int64_t i = 0; i < n_trg; i++) out[i] += tmp[i];
  }
  if (l > 0) {
    double m[10];
    for (int64_t i = 0; i < n_trg; i++) {
      orc_monomials(poly_degree, trg + 3 * i, 1, m);
      double acc = 0.0;
      for (int64_t k = 0; k < l; k++) acc += m[k] * weights[n_src + k];
      out[i] += acc;
    }
  }
  free(tmp);
}

void orc_direct_operator(const orc_rbf* rbfs, int n_rbfs, int poly_degree, double nugget,
                         const double* pts, int64_t n, const double* x, double* y) {
  int64_t l = orc_poly_basis_size(poly_degree);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) y[i] = nugget * x[i];
  for (int r = 0; r < n_rbfs; r++) {
    orc_direct_rbf_sum(&rbfs[r], pts, n, x, pts, n, tmp);
    for (int64_t i = 0; i < n; i++) y[i] += tmp[i];
  }
  if (l > 0) {
    double m[10];
    for (int64_t k = 0; k < l; k++) y[n + k] = 0.0;
    for (int64_t i = 0; i < n; i++) {
      orc_monomials(poly_degree, pts + 3 * i, 1, m);
      double acc = 0.0;
      for (int64_t k = 0; k < l; k++) {
        acc += m[k] * x[n + k];
        y[n + k] += m[k] * x[i];
      }
      y[i] += acc;
    }
  }
  free(tmp);
}

/* Gradients of the monomials: out[(3 i + c) * l + k] = d m_k / d x_c at point i
 * (MonomialBasis::evaluate with gradient points, monomial_basis.hpp:198-247). */
void orc_monomial_gradients(int degree, const double* pts, int64_t n, double* out) {
  int64_t l = orc_poly_basis_size(degree);
  for (int64_t i = 0; i < n; i++) {
    double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    double gx[10] = {0.0, 1.0, 0.0, 0.0, 2.0 * x, y, z, 0.0, 0.0, 0.0};
    double gy[10] = {0.0, 0.0, 1.0, 0.0, 0.0, x, 0.0, 2.0 * y, z, 0.0};
    double gz[10] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, x, 0.0, y, 2.0 * z};
    for (int64_t k = 0; k < l; k++) {
      out[(3 * i) * l + k] = gx[k];
      out[(3 * i + 1) * l + k] = gy[k];
      out[(3 * i + 2) * l + k] = gz[k];
    }
  }
}

/* The operator of a fit with gradient data (interpolation/operator.hpp:56-82,
 * with the blocks of preconditioner/mat_a.hpp:11-61), by direct summation:
 * x = [w (mu); g (3 sigma, x y z per gradient point); c (l)],
 *   y_v[i] = nugget w_i + sum_j w_j phi(p_i - p_j) - sum_k g_k . grad phi(p_i - q_k) + P_v c,
 *   y_g[k] = sum_j w_j grad phi(q_k - p_j) - sum_m H(q_k - q_m) g_m + P_g c,
 *   y_p    = P_v^T w + P_g^T g,
 * summed over the rbfs; P_v holds the monomials at the points p, P_g their
 * gradients at the gradient points q. */
void orc_hermite_operator(const orc_rbf* rbfs, int n_rbfs, int poly_degree, double nugget,
                          const double* pts, int64_t mu, const double* gpts, int64_t sigma,
                          const double* x, double* y) {
  int64_t l = orc_poly_basis_size(poly_degree);
  int64_t n = mu + 3 * sigma;
  const double* g = x + mu;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < mu; i++) {
    double acc = nugget * x[i], gr[3];
    for (int r = 0; r < n_rbfs; r++) {
      for (int64_t j = 0; j < mu; j++)
        acc += x[j] * orc_rbf_evaluate(&rbfs[r], pts[3 * i] - pts[3 * j], pts[3 * i + 1] - pts[3 * j + 1],
                                       pts[3 * i + 2] - pts[3 * j + 2]);
      for (int64_t k = 0; k < sigma; k++) {
        orc_rbf_gradient(&rbfs[r], pts[3 * i] - gpts[3 * k], pts[3 * i + 1] - gpts[3 * k + 1],
                         pts[3 * i + 2] - gpts[3 * k + 2], gr);
        acc -= (g[3 * k] * gr[0] + g[3 * k + 1] * gr[1]) + g[3 * k + 2] * gr[2];
      }
    }
    y[i] = acc;
  }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < sigma; k++) {
    double acc[3] = {0.0, 0.0, 0.0}, gr[3], h[9];
    for (int r = 0; r < n_rbfs; r++) {
      for (int64_t j = 0; j < mu; j++) {
        orc_rbf_gradient(&rbfs[r], gpts[3 * k] - pts[3 * j], gpts[3 * k + 1] - pts[3 * j + 1],
                         gpts[3 * k + 2] - pts[3 * j + 2], gr);
        for (int c = 0; c < 3; c++) acc[c] += x[j] * gr[c];
      }
      for (int64_t m = 0; m < sigma; m++) {
        orc_rbf_hessian(&rbfs[r], gpts[3 * k] - gpts[3 * m], gpts[3 * k + 1] - gpts[3 * m + 1],
                        gpts[3 * k + 2] - gpts[3 * m + 2], h);
        for (int c = 0; c < 3; c++)
          acc[c] -= (h[3 * c] * g[3 * m] + h[3 * c + 1] * g[3 * m + 1]) + h[3 * c + 2] * g[3 * m + 2];
      }
    }
    for (int c = 0; c < 3; c++) y[mu + 3 * k + c] = acc[c];
  }
  if (l > 0) {
    double m[10], mg[30];
    for (int64_t a = 0; a < l; a++) y[n + a] = 0.0;
    for (int64_t i = 0; i < mu; i++) {
      orc_monomials(poly_degree, pts + 3 * i, 1, m);
      double acc = 0.0;
      for (int64_t a = 0; a < l; a++) {
        acc += m[a] * x[n + a];
        y[n + a] += m[a] * x[i];
      }
      y[i] += acc;
    }
    for (int64_t k = 0; k < sigma; k++) {
      orc_monomial_gradients(poly_degree, gpts + 3 * k, 1, mg);
      for (int c = 0; c < 3; c++) {
        double acc = 0.0;
        for (int64_t a = 0; a < l; a++) {
          acc += mg[c * l + a] * x[n + a];
          y[n + a] += mg[c * l + a] * g[3 * k + c];
        }
        y[mu + 3 * k + c] += acc;
      }
    }
  }
}

/* ---- Box and octree (src/fmm/utility.hpp:19-33 for the box) */

void orc_make_box(const double* bmin, const double* bmax, const double* aniso, double* center,
                  double* width) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; c++) {
    double p[3] = {(c & 1) ? bmax[0] : bmin[0], (c & 2) ? bmax[1] : bmin[1],
                   (c & 4) ? bmax[2] : bmin[2]};
    double t[3];
    apply_aniso(aniso, p[0], p[1], p[2], t);
    for (int k = 0; k < 3; k++) {
      if (t[k] < lo[k]) lo[k] = t[k];
      if (t[k] > hi[k]) hi[k] = t[k];
    }
  }
  double wmax = 0.0;
  for (int k = 0; k < 3; k++) {
    double wk = hi[k] - lo[k];
    if (wk > wmax) wmax = wk;
    center[k] = lo[k] + wk / 2.0;
  }
  double w = 1.01 * wmax;
  *width = (w == 0.0) ? 1.0 : w;
}

static uint64_t spread3(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 21; b++) r |= ((v >> b) & 1ull) << (3 * b);
  return r;
}

void orc_morton_codes(const double* pts, int64_t n, const double* lo, double inv_cell, int level,
                      uint64_t* codes) {
  int64_t hi = ((int64_t)1 << level) - 1;
  for (int64_t i = 0; i < n; i++) {
    int64_t c[3];
    for (int k = 0; k < 3; k++) {
      double f = floor((pts[3 * i + k] - lo[k]) * inv_cell);
      int64_t v = (f < 0.0) ? 0 : (f > (double)hi ? hi : (int64_t)f);
      c[k] = v;
    }
    codes[i] = (spread3((uint64_t)c[0]) << 2) | (spread3((uint64_t)c[1]) << 1) | spread3((uint64_t)c[2]);
  }
}

static const uint64_t* g_sort_codes;
static int cmp_perm(const void* a, const void* b) {
  int32_t i = *(const int32_t*)a, j = *(const int32_t*)b;
  uint64_t ci = g_sort_codes[i], cj = g_sort_codes[j];
  if (ci != cj) return ci < cj ? -1 : 1;
  return (i > j) - (i < j); /* ties by input position: a stable order */
}

void orc_stable_sort_codes(const uint64_t* codes, int64_t n, int32_t* perm) {
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  g_sort_codes = codes;
  qsort(perm, (size_t)n, sizeof(int32_t), cmp_perm);
}

int64_t orc_leaves(const uint64_t* sorted_codes, int64_t n, uint64_t* leaf_codes,
                   int32_t* leaf_start) {
  int64_t nl = 0;
  for (int64_t i = 0; i < n; i++) {
    if (i == 0 || sorted_codes[i] != sorted_codes[i - 1]) {
      leaf_codes[nl] = sorted_codes[i];
      leaf_start[nl] = (int32_t)i;
      nl++;
    }
  }
  leaf_start[nl] = (int32_t)n;
  return nl;
}

static void decode(uint64_t code, int64_t* c) {
  c[0] = c[1] = c[2] = 0;
  for (int b = 0; b < 21; b++) {
    c[0] |= (int64_t)((code >> (3 * b + 2)) & 1) << b;
    c[1] |= (int64_t)((code >> (3 * b + 1)) & 1) << b;
    c[2] |= (int64_t)((code >> (3 * b)) & 1) << b;
  }
}

static int64_t find_code(const uint64_t* codes, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (codes[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < n && codes[lo] == key) ? lo : -1;
}

static uint64_t encode(const int64_t* c) {
  return (spread3((uint64_t)c[0]) << 2) | (spread3((uint64_t)c[1]) << 1) | spread3((uint64_t)c[2]);
}

void orc_p2p_lists(const uint64_t* trg_codes, int64_t n_trg, const uint64_t* src_codes,
                   int64_t n_src, int level, int32_t* slots) {
  int64_t side = (int64_t)1 << level;
  for (int64_t t = 0; t < n_trg; t++) {
    int64_t c[3];
    decode(trg_codes[t], c);
    for (int s = 0; s < 27; s++) {
      int64_t d[3] = {s / 9 - 1, (s / 3) % 3 - 1, s % 3 - 1};
      int64_t q[3] = {c[0] + d[0], c[1] + d[1], c[2] + d[2]};
      int32_t found = -1;
      if (q[0] >= 0 && q[0] < side && q[1] >= 0 && q[1] < side && q[2] >= 0 && q[2] < side)
        found = (int32_t)find_code(src_codes, n_src, encode(q));
      slots[t * 27 + s] = found;
    }
  }
}

void orc_m2l_lists(const uint64_t* trg_codes, int64_t n_trg, const uint64_t* src_codes,
                   int64_t n_src, int level, int32_t* slots) {
  int64_t side = (int64_t)1 << level;
  for (int64_t t = 0; t < n_trg; t++) {
    int64_t c[3];
    decode(trg_codes[t], c);
    for (int s = 0; s < 343; s++) {
      int64_t d[3] = {s / 49 - 3, (s / 7) % 7 - 3, s % 7 - 3};
      int64_t q[3] = {c[0] + d[0], c[1] + d[1], c[2] + d[2]};
      int32_t found = -1;
      int adjacent = llabs(d[0]) <= 1 && llabs(d[1]) <= 1 && llabs(d[2]) <= 1;
      int parents_adjacent = 1;
      for (int k = 0; k < 3; k++) {
        if (q[k] < 0 || q[k] >= side) parents_adjacent = 0;
        else if (llabs((q[k] >> 1) - (c[k] >> 1)) > 1) parents_adjacent = 0;
      }
      if (!adjacent && parents_adjacent) found = (int32_t)find_code(src_codes, n_src, encode(q));
      slots[t * 343 + s] = found;
    }
  }
}

/* ---- FGMRES (src/krylov/gmres.cpp:9-50, fgmres.cpp:8-26, gmres_base.cpp:35-69) */

static void matvec(const double* a, int64_t n, const double* x, double* y) {
  for (int64_t i = 0; i < n; i++) {
    double acc = 0.0;
    for (int64_t j = 0; j < n; j++) acc += a[i * n + j] * x[j];
    y[i] = acc;
  }
}

int orc_fgmres_dense(const double* a, const double* m, int64_t n, const double* rhs,
                     const double* x0, int n_iter, double* x, double* resid_hist) {
  int kmax = n_iter;
  double* v = (double*)calloc((size_t)(kmax + 1) * (size_t)n, sizeof(double));
  double* z = (double*)calloc((size_t)kmax * (size_t)n, sizeof(double));
  double* r = (double*)calloc((size_t)(kmax + 1) * (size_t)kmax, sizeof(double));
  double* c = (double*)calloc((size_t)kmax + 1, sizeof(double));
  double* s = (double*)calloc((size_t)kmax + 1, sizeof(double));
  double* g = (double*)calloc((size_t)kmax + 1, sizeof(double));
  double* tmp = (double*)calloc((size_t)n, sizeof(double));
  /* r0 = rhs - op(x0) */
  matvec(a, n, x0, tmp);
  double nrm = 0.0;
  for (int64_t i = 0; i < n; i++) {
    v[i] = rhs[i] - tmp[i];
    nrm += v[i] * v[i];
  }
  nrm = sqrt(nrm);
  g[0] = nrm;
  for (int64_t i = 0; i < n; i++) v[i] /= nrm;
  resid_hist[0] = fabs(g[0]);
  int it = 0;
  for (int j = 0; j < kmax; j++) {
    double* vj = v + (size_t)j * n;
    double* zj = z + (size_t)j * n;
    double* w = v + (size_t)(j + 1) * n;
    if (m) matvec(m, n, vj, zj); else memcpy(zj, vj, sizeof(double) * (size_t)n);
    matvec(a, n, zj, w);
    /* classical Gram-Schmidt: all dots against the unmodified w, then updates */
    for (int i = 0; i <= j; i++) {
      double d = 0.0;
      for (int64_t k = 0; k < n; k++) d += v[(size_t)i * n + k] * w[k];
      r[i * kmax + j] = d;
    }
    for (int i = 0; i <= j; i++)
      for (int64_t k = 0; k < n; k++) w[k] -= r[i * kmax + j] * v[(size_t)i * n + k];
    double h = 0.0;
    for (int64_t k = 0; k < n; k++) h += w[k] * w[k];
    h = sqrt(h);
    r[(j + 1) * kmax + j] = h;
    if (h != 0.0)
      for (int64_t k = 0; k < n; k++) w[k] /= h;
    for (int i = 0; i < j; i++) {
      double xx = r[i * kmax + j], yy = r[(i + 1) * kmax + j];
      r[i * kmax + j] = c[i] * xx + s[i] * yy;
      r[(i + 1) * kmax + j] = -s[i] * xx + c[i] * yy;
    }
    double xx = r[j * kmax + j], yy = r[(j + 1) * kmax + j];
    double den = hypot(xx, yy);
    c[j] = xx / den;
    s[j] = yy / den;
    r[j * kmax + j] = c[j] * xx + s[j] * yy;
    g[j + 1] = -s[j] * g[j];
    g[j] = c[j] * g[j];
    it = j + 1;
    resid_hist[j + 1] = fabs(g[j + 1]);
    if (h == 0.0) break;
  }
  /* back substitution r y = g, x = x0 + sum y_i z_i */
  double* y = (double*)calloc((size_t)(it > 0 ? it : 1), sizeof(double));
  for (int j = it - 1; j >= 0; j--) {
    double acc = g[j];
    for (int i = j + 1; i < it; i++) acc -= r[j * kmax + i] * y[i];
    y[j] = acc / r[j * kmax + j];
  }
  for (int64_t k = 0; k < n; k++) {
    double acc = x0[k];
    for (int i = 0; i < it; i++) acc += y[i] * z[(size_t)i * n + k];
    x[k] = acc;
  }
  free(v); free(z); free(r); free(c); free(s); free(g); free(tmp); free(y);
  return it;
}
