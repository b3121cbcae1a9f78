/*
 * CPU oracle for the RBF fit/evaluate hot path — TEST INFRASTRUCTURE ONLY.
 *
 * Plain C11 + OpenMP, FP64 throughout. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker.
 * The product path (polatory_amd) never links, loads or calls it.
 *
 * Each function cites the reference file:line whose semantics it restates
 * (paths relative to the polatory source tree).
 */
#ifndef POLATORY_ORACLE_H
#define POLATORY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BH3 = 0, ORC_TH3 = 1, ORC_EXP = 2 };

typedef struct {
  int kind;           /* ORC_BH3 | ORC_TH3 | ORC_EXP */
  double params[2];   /* bh3/th3: {slope, c}; exp: {psill, range} */
  double aniso[9];    /* row-major 3x3 anisotropy A, applied as d -> d * A^T */
} orc_rbf;

/* phi(A d) for one difference vector d (rbf/rbf_base.hpp:40-43). */
double orc_rbf_evaluate(const orc_rbf* rbf, double dx, double dy, double dz);
void orc_rbf_gradient(const orc_rbf* rbf, double dx, double dy, double dz, double* g);
void orc_rbf_hessian(const orc_rbf* rbf, double dx, double dy, double dz, double* h);
void orc_direct_gradient_sum(const orc_rbf* rbf, const double* src, int64_t n_src, const double* w,
                             const double* trg, int64_t n_trg, double* out);

/* Points p -> p * A^T, evaluated as ((A0*x + A1*y) + A2*z) without contraction
 * (geometry/point3d.hpp:36-40). */
void orc_transform_points(const double* aniso, const double* pts, int64_t n, double* out);

/* Monomial basis [1, x, y, z, x^2, xy, xz, y^2, yz, z^2] truncated to degree
 * (polynomial/monomial_basis.hpp:198-247); out is n x l row-major. */
int64_t orc_poly_basis_size(int degree);
void orc_monomials(int degree, const double* pts, int64_t n, double* out);

/* DirectEvaluator::evaluate (interpolation/direct_evaluator.hpp:45-79) without
 * gradient points: y_i = sum_r sum_j w_j phi_r(x_i - s_j) + sum_k c_k m_k(x_i).
 * weights has n_src + l entries (poly coefficients last). */
void orc_direct_evaluate(const orc_rbf* rbfs, int n_rbfs, int poly_degree, const double* src,
                         int64_t n_src, const double* weights, const double* trg, int64_t n_trg,
                         double* out);

/* Same sum restricted to the RBF part (no polynomial), for the FMM seam
 * (src/fmm/full_direct.hpp:32-52). */
void orc_direct_rbf_sum(const orc_rbf* rbf, const double* src, int64_t n_src, const double* w,
                        const double* trg, int64_t n_trg, double* out);

/* DirectOperator::operator() (interpolation/direct_operator.hpp:42-76):
 * y = [(A + nugget I) w + P c ; P^T w] for x = [w ; c]. */
/* Gradients of the monomials: out[(3 i + c) * l + k] = d m_k / d x_c at point i. */
void orc_monomial_gradients(int degree, const double* pts, int64_t n, double* out);
/* Operator of a fit with gradient data: x, y have mu + 3 sigma + l entries
 * ([values; gradients x y z per gradient point; polynomial]). */
void orc_hermite_operator(const orc_rbf* rbfs, int n_rbfs, int poly_degree, double nugget,
                          const double* pts, int64_t mu, const double* gpts, int64_t sigma,
                          const double* x, double* y);
void orc_direct_operator(const orc_rbf* rbfs, int n_rbfs, int poly_degree, double nugget,
                         const double* pts, int64_t n, const double* x, double* y);

/* make_box (src/fmm/utility.hpp:19-33): the bbox corners transformed by A,
 * cube of width 1.01 x the largest extent around the transformed bbox centre. */
void orc_make_box(const double* bbox_min, const double* bbox_max, const double* aniso,
                  double* center, double* width);

/* Octree over already-transformed points (builder's definition, restated on
 * the CPU for a bit-exact check of the device tree):
 *   ix = clamp(floor((x - lo) * inv_cell), 0, 2^level - 1), likewise y, z;
 *   code = sum_b bit_b(ix) << (3b+2) | bit_b(iy) << (3b+1) | bit_b(iz) << 3b.
 * perm is the stable sort of codes (ties keep input order). */
void orc_morton_codes(const double* pts, int64_t n, const double* lo, double inv_cell, int level,
                      uint64_t* codes);
void orc_stable_sort_codes(const uint64_t* codes, int64_t n, int32_t* perm);

/* Unique leaf codes of sorted codes; returns the number of leaves, fills
 * leaf_codes[n_leaves] and leaf_start[n_leaves + 1] (point ranges). */
int64_t orc_leaves(const uint64_t* sorted_codes, int64_t n, uint64_t* leaf_codes,
                   int32_t* leaf_start);

/* P2P list: for each target leaf, 27 slots indexed (dx+1)*9+(dy+1)*3+(dz+1)
 * holding the source leaf index with that offset or -1 (separation criterion 1,
 * fmm/kernel.hpp:71). */
void orc_p2p_lists(const uint64_t* trg_codes, int64_t n_trg, const uint64_t* src_codes,
                   int64_t n_src, int level, int32_t* slots);

/* M2L list at one level: for each target cell, 343 slots indexed
 * (dx+3)*49+(dy+3)*7+(dz+3) holding the source cell index whose parent is a
 * neighbour of (or equal to) the target's parent while the cells themselves are
 * not neighbours, else -1. */
void orc_m2l_lists(const uint64_t* trg_codes, int64_t n_trg, const uint64_t* src_codes,
                   int64_t n_src, int level, int32_t* slots);

/* Right-preconditioned FGMRES without restart (src/krylov/gmres.cpp:9-50,
 * src/krylov/fgmres.cpp:8-26, gmres_base.cpp:35-69) on a dense row-major n x n
 * operator with a dense right preconditioner (NULL = identity). Runs exactly
 * n_iter iterations (or stops on a zero Krylov residual), writes the solution and
 * the Givens residual history |g_j|, j = 0..iters. Returns the iterations done. */
int orc_fgmres_dense(const double* a, const double* m, int64_t n, const double* rhs,
                     const double* x0, int n_iter, double* x, double* resid_hist);

/* Reference-algorithm CPU FMM (cpu_fmm.c): polatory's evaluate path over a
 * uniform-interpolation FMM with DFT-based M2L, FP64, OpenMP. order = nodes
 * per axis; fh_degree < 0: classic Lagrange, else Floater-Hormann degree;
 * height <= 0: the reference's height formula (and the direct sum below
 * 1024^2 pairs). Returns 0, or < 0 for an order or height out of range. */
typedef struct {
  double t_tree, t_upward, t_m2l, t_downward, t_p2p, t_total;
  int height, order;
  int64_t m2l_pairs, p2p_pairs, n_src_leaves, n_trg_leaves;
} orc_fmm_stats;
int orc_cpu_fmm_evaluate(const orc_rbf* rbf, const double* bbox_min, const double* bbox_max,
                         const double* src, int64_t n_src, const double* w, const double* trg,
                         int64_t n_trg, int order, int fh_degree, int height, double* out,
                         orc_fmm_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
