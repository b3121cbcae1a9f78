/*
 * polatory_amd — C ABI of the MI355X (gfx950) RBF fit/evaluate hot path.
 *
 * Plain C: opaque handles, host or device (HBM) pointers and int64 sizes.
 * Every entry point returns PAM_OK (0) or an error code; pam_last_error()
 * returns the message of the calling thread's last failure. Nothing throws
 * across this boundary. Points are row-major n x 3 FP64 (x, y, z interleaved),
 * the layout of polatory's geometry::Points (types.hpp:9-22); vectors are FP64.
 * Outputs are in the input order of the target points.
 *
 * Reference interfaces replaced (paths relative to the polatory tree):
 *   FMM seam     include/polatory/fmm/fmm_evaluator.hpp:17-41, 92-106
 *                include/polatory/fmm/fmm_symmetric_evaluator.hpp:15-37, 80-86
 *   evaluation   include/polatory/interpolation/evaluator.hpp:50-144
 *   fit          include/polatory/interpolation/fitter.hpp:21-26, solver.hpp:34-142
 *   direct sums  include/polatory/interpolation/direct_evaluator.hpp:38-131
 */
#ifndef POLATORY_AMD_H
#define POLATORY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define PAM_OK 0
#define PAM_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument in polatory */
#define PAM_ERR_RUNTIME 2          /* std::runtime_error in polatory */
#define PAM_ERR_HIP 3
#define PAM_ERR_OUT_OF_MEMORY 4
#define PAM_ERR_NOT_IMPLEMENTED 5

/* RBF families (rbf/polyharmonic_odd.hpp:120-136, rbf/cov_exponential.hpp:12-68). */
#define PAM_RBF_BH3 0 /* biharmonic3d  phi = -slope * sqrt(r^2 + c^2), params {slope, c} */
#define PAM_RBF_TH3 1 /* triharmonic3d phi = +slope * (r^2 + c^2)^1.5, params {slope, c} */
#define PAM_RBF_EXP 2 /* cov_exponential phi = psill * exp(-3 r / range), params {psill, range} */

typedef struct {
  int kind;           /* PAM_RBF_* */
  double params[2];   /* see above */
  double aniso[9];    /* row-major anisotropy A (det > 0); differences d -> d A^T (rbf_base.hpp:40-79) */
} pam_rbf;

/* Copies the calling thread's last error message (NUL-terminated, truncated to len). */
int pam_last_error(char* buf, size_t len);
/* Number of visible GPUs (0 when none). */
int pam_device_count(int* count);

/* ---------------------------------------------------------------------------
 * FMM seam: replaces fmm::FmmGenericEvaluatorBase<3> and
 * fmm::FmmGenericSymmetricEvaluatorBase<3> with make_fmm_evaluator /
 * make_fmm_symmetric_evaluator (fmm_evaluator.hpp:92-95, fmm_symmetric_evaluator.hpp:80-82).
 * The value kernel (fmm/kernel.hpp:14-75): out_i = sum_j w_j phi(A (t_i - s_j)).
 * ------------------------------------------------------------------------- */
typedef struct pam_fmm_evaluator pam_fmm_evaluator;

/* make_fmm_evaluator(rbf, bbox) — fmm_evaluator.hpp:92-93 */
int pam_fmm_evaluator_create(const pam_rbf* rbf, const double bbox_min[3], const double bbox_max[3],
                             pam_fmm_evaluator** out);
/* make_fmm_symmetric_evaluator(rbf, bbox) — fmm_symmetric_evaluator.hpp:80-82 */
int pam_fmm_symmetric_evaluator_create(const pam_rbf* rbf, const double bbox_min[3],
                                       const double bbox_max[3], pam_fmm_evaluator** out);
int pam_fmm_evaluator_destroy(pam_fmm_evaluator* h);
/* set_accuracy(double) — absolute L-inf accuracy; +inf = default configuration (fmm_evaluator.hpp:36) */
int pam_fmm_evaluator_set_accuracy(pam_fmm_evaluator* h, double accuracy);
/* set_source_points(const Points&) — fmm_evaluator.hpp:38 */
int pam_fmm_evaluator_set_source_points(pam_fmm_evaluator* h, const double* xyz, int64_t n);
/* set_target_points(const Points&) — fmm_evaluator.hpp:40 */
int pam_fmm_evaluator_set_target_points(pam_fmm_evaluator* h, const double* xyz, int64_t n);
/* set_points(const Points&) — fmm_symmetric_evaluator.hpp:35 (symmetric handles only) */
int pam_fmm_evaluator_set_points(pam_fmm_evaluator* h, const double* xyz, int64_t n);
/* set_weights(const VecX&) — fmm_evaluator.hpp:42 */
int pam_fmm_evaluator_set_weights(pam_fmm_evaluator* h, const double* w, int64_t n);
/* VecX evaluate() const — fmm_evaluator.hpp:34; out has n_targets entries */
int pam_fmm_evaluator_evaluate(pam_fmm_evaluator* h, double* out, int64_t n);

/* Same calls with pointers into device memory (HBM); asynchronous on the
 * handle's stream, ordered with the handle's other calls. */
int pam_fmm_evaluator_set_source_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n);
int pam_fmm_evaluator_set_target_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n);
int pam_fmm_evaluator_set_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n);
int pam_fmm_evaluator_set_weights_device(pam_fmm_evaluator* h, const double* d_w, int64_t n);
int pam_fmm_evaluator_evaluate_device(pam_fmm_evaluator* h, double* d_out, int64_t n);
int pam_fmm_evaluator_synchronize(pam_fmm_evaluator* h);

/* Explicit configuration (replaces the accuracy search; no reference counterpart:
 * polatory picks {tree height, order, d} internally, fmm_accuracy_estimator.hpp:74-121). */
#define PAM_PRECISION_FP64 0  /* every stage in FP64 */
#define PAM_PRECISION_MIXED 1 /* FP32 P2P/P2M/L2P and FP32-MFMA M2L, FP64 output */
#define PAM_MODE_AUTO 0
#define PAM_MODE_DIRECT 1
#define PAM_MODE_FMM 2
typedef struct {
  int precision;  /* PAM_PRECISION_* */
  int order;      /* Chebyshev points per axis, 3..12 */
  int leaf_level; /* octree leaf level 2..18, or -1 for automatic */
  int mode;       /* PAM_MODE_* */
} pam_fmm_config;
int pam_fmm_evaluator_set_config(pam_fmm_evaluator* h, const pam_fmm_config* cfg);
int pam_fmm_evaluator_get_config(pam_fmm_evaluator* h, pam_fmm_config* cfg);

/* Per-stage device times of the last evaluate (HIP events on the handle's stream). */
typedef struct {
  double ms_tree, ms_upward, ms_m2l, ms_l2l, ms_l2p, ms_p2p, ms_total;
  int precision, order, leaf_level, direct;
  int64_t n_src_leaves, n_trg_leaves;
} pam_fmm_stats;
int pam_fmm_evaluator_stats(pam_fmm_evaluator* h, pam_fmm_stats* out);
/* Near-field source-target pairs and M2L cell pairs of the current lists. */
int pam_fmm_evaluator_count_work(pam_fmm_evaluator* h, int64_t* p2p_pairs, int64_t* m2l_pairs);

/* Tree export for the bit-exact check against a CPU recomputation.
 * which: 0 = source tree, 1 = target tree. */
int pam_fmm_evaluator_tree_info(pam_fmm_evaluator* h, int which, int* level, double lo[3],
                                double* width, double* inv_cell, int64_t* n_points,
                                int64_t* n_leaves);
int pam_fmm_evaluator_export_tree(pam_fmm_evaluator* h, int which, uint64_t* point_codes,
                                  int32_t* perm, uint64_t* leaf_codes, int32_t* leaf_start);
/* Cell codes of one level (codes may be NULL to query the count). */
int pam_fmm_evaluator_export_level(pam_fmm_evaluator* h, int which, int level, int64_t* n,
                                   uint64_t* codes);
/* P2P slots [n_target_leaves][27]; M2L slots [n_target_cells(level)][343]. */
int pam_fmm_evaluator_export_p2p(pam_fmm_evaluator* h, int32_t* slots);
int pam_fmm_evaluator_export_m2l(pam_fmm_evaluator* h, int level, int32_t* slots);

/* FP64 direct O(N*M) summation on the GPU (DirectEvaluator's kernel sum,
 * direct_evaluator.hpp:45-79, without polynomial terms). Host pointers. */
int pam_direct_evaluate(const pam_rbf* rbf, const double* src, int64_t n_src, const double* w,
                        const double* trg, int64_t n_trg, double* out);

#ifdef __cplusplus
}
#endif

#endif /* POLATORY_AMD_H */
