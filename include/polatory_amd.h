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
/* Device memory cache (no counterpart in polatory, whose buffers are host
 * memory): device buffers the library releases are kept for reuse on their
 * GPU instead of being freed (hipFree synchronizes the device and costs about
 * 0.1 ms per buffer; a 1M-point fit releases about 900). The cache holds at
 * most a quarter of the device memory and is emptied by itself when an
 * allocation runs out of memory. PAM_DEVICE_POOL=0 in the environment turns
 * it off. pam_device_cache_release frees every cached buffer of the current
 * device (after waiting for the device); pam_device_cache_bytes reports the
 * bytes it holds. */
int pam_device_cache_release(void);
int pam_device_cache_bytes(size_t* bytes);

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
/* make_fmm_gradient_transpose_evaluator(rbf, bbox) — make_fmm_evaluator.cpp:119-148:
 * a source-target handle whose evaluate() returns 3 n_targets gradients (as
 * pam_fmm_evaluator_evaluate_gradient) and whose accuracy search measures them. */
int pam_fmm_gradient_transpose_evaluator_create(const pam_rbf* rbf, const double bbox_min[3],
                                                const double bbox_max[3], pam_fmm_evaluator** out);
/* make_fmm_gradient_evaluator(rbf, bbox) — make_fmm_evaluator.cpp:78-107,
 * fmm/gradient_kernel.hpp:46-62: sources carry gradient weights (set_weights takes
 * 3 n_sources entries, x y z per source) and evaluate() returns n_targets values,
 * out[i] = sum_j w_j . (-A^T grad phi(A (t_i - s_j))). Evaluated as three
 * gradient-transpose passes; the accuracy search measures those gradients. */
int pam_fmm_gradient_evaluator_create(const pam_rbf* rbf, const double bbox_min[3], const double bbox_max[3],
                                      pam_fmm_evaluator** out);
/* make_fmm_hessian_evaluator(rbf, bbox) — make_fmm_evaluator.cpp:160-188,
 * fmm/hessian_kernel.hpp:46-64: sources carry gradient weights (3 n_sources entries)
 * and evaluate() returns 3 n_targets entries, out[3 i + c] =
 * -(sum_j A^T H(A (t_i - s_j)) A w_j)_c. Evaluated as three Hessian-column passes; the
 * accuracy search measures gradient-transpose errors (Hessians at the same
 * configuration are less accurate). */
int pam_fmm_hessian_evaluator_create(const pam_rbf* rbf, const double bbox_min[3], const double bbox_max[3],
                                     pam_fmm_evaluator** out);
/* make_fmm_hessian_symmetric_evaluator (fmm_symmetric_evaluator.hpp:85-86,
 * make_fmm_evaluator.cpp:241-270): the Hessian kernel with the points as both
 * sources and targets. Use pam_fmm_evaluator_set_points; weights and results
 * are 3 per point (x y z). The zero-distance term is zero for bh3 and th3. */
int pam_fmm_hessian_symmetric_evaluator_create(const pam_rbf* rbf, const double bbox_min[3],
                                               const double bbox_max[3], pam_fmm_evaluator** out);
int pam_fmm_evaluator_destroy(pam_fmm_evaluator* h);
/* set_accuracy(double) — absolute L-inf accuracy; +inf = default configuration (fmm_evaluator.hpp:36) */
int pam_fmm_evaluator_set_accuracy(pam_fmm_evaluator* h, double accuracy);
/* No counterpart in the reference (its search has one candidate list,
 * fmm_accuracy_estimator.hpp:105-117): restricts the accuracy search to one
 * precision (PAM_PRECISION_*), trying orders 4..12 of it, cheapest first, and
 * undoes a configuration fixed by pam_fmm_evaluator_set_config; -1 restores
 * the default list (mixed orders 4-8, compensated mixed 8, FP64 8, 10, 12). */
int pam_fmm_evaluator_set_search_precision(pam_fmm_evaluator* h, int precision);
/* set_source_points(const Points&) — fmm_evaluator.hpp:38 */
int pam_fmm_evaluator_set_source_points(pam_fmm_evaluator* h, const double* xyz, int64_t n);
/* set_target_points(const Points&) — fmm_evaluator.hpp:40 */
int pam_fmm_evaluator_set_target_points(pam_fmm_evaluator* h, const double* xyz, int64_t n);
/* set_points(const Points&) — fmm_symmetric_evaluator.hpp:35 (symmetric handles only) */
int pam_fmm_evaluator_set_points(pam_fmm_evaluator* h, const double* xyz, int64_t n);
/* set_weights(const VecX&) — fmm_evaluator.hpp:42 */
int pam_fmm_evaluator_set_weights(pam_fmm_evaluator* h, const double* w, int64_t n);
/* VecX evaluate() const — fmm_evaluator.hpp:34; out has n_targets entries
 * (3 n_targets for a gradient-transpose handle); n is the number of targets */
int pam_fmm_evaluator_evaluate(pam_fmm_evaluator* h, double* out, int64_t n);
/* evaluate() of make_fmm_gradient_transpose_evaluator (make_fmm_evaluator.cpp:119-148,
 * fmm/gradient_transpose_kernel.hpp:47-62) on the same handle: out has 3 n_targets
 * entries, out[3 i + k] = (A^T sum_j w_j grad phi(A (t_i - s_j)))_k, using the
 * handle's configuration, trees and expansions. Source-target handles only. */
int pam_fmm_evaluator_evaluate_gradient(pam_fmm_evaluator* h, double* out, int64_t n);

/* Same calls with pointers into device memory (HBM), enqueued on the
 * handle's stream and ordered with the handle's other calls. The setters
 * return without waiting for the device, so the caller keeps its buffers
 * unchanged until the evaluation that reads them is complete (after
 * pam_fmm_evaluator_synchronize). The evaluations wait for the device once,
 * when the trees are rebuilt (their occupied-cell counts choose the leaf level
 * and size the lists), and return with the output still being written. */
int pam_fmm_evaluator_set_source_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n);
int pam_fmm_evaluator_set_target_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n);
int pam_fmm_evaluator_set_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n);
int pam_fmm_evaluator_set_weights_device(pam_fmm_evaluator* h, const double* d_w, int64_t n);
int pam_fmm_evaluator_evaluate_device(pam_fmm_evaluator* h, double* d_out, int64_t n);
int pam_fmm_evaluator_evaluate_gradient_device(pam_fmm_evaluator* h, double* d_out, int64_t n);
int pam_fmm_evaluator_synchronize(pam_fmm_evaluator* h);

/* Explicit configuration (replaces the accuracy search; no reference counterpart:
 * polatory picks {tree height, order, d} internally, fmm_accuracy_estimator.hpp:74-121). */
#define PAM_PRECISION_FP64 0  /* every stage in FP64 */
#define PAM_PRECISION_MIXED 1 /* FP32 P2P and FP32-MFMA M2L; FP64 expansions, P2M/M2M/L2L/L2P and output */
#define PAM_PRECISION_MIXED_COMPENSATED 2 /* MIXED with compensated (Kahan) near-field sums, for
                                             strongly cancelling weights (value evaluations) */
#define PAM_MODE_AUTO 0
#define PAM_MODE_DIRECT 1
#define PAM_MODE_FMM 2
typedef struct {
  int precision;  /* PAM_PRECISION_* */
  int order;      /* Chebyshev points per axis, 3..12 */
  int leaf_level; /* octree leaf level 2..18, or -1 for automatic */
  int mode;       /* PAM_MODE_* */
  double svd_tolerance; /* M2L compression: > 0 = compressed transfers whose
                           basis leaves an estimated relative residual within
                           this value (rank a multiple of 64, at most 256);
                           0 = dense M2L operators; negative = the default
                           (compressed at 2e-7 for biharmonic and triharmonic
                           kernels at orders 7 and 8, dense otherwise) */
} pam_fmm_config;
int pam_fmm_evaluator_set_config(pam_fmm_evaluator* h, const pam_fmm_config* cfg);
int pam_fmm_evaluator_get_config(pam_fmm_evaluator* h, pam_fmm_config* cfg);
/* No counterpart in the reference (its evaluator runs the stages in
 * sequence). overlap != 0 (the default): the near field runs on a second
 * stream beside the far field, the result is the same bit for bit; 0: every
 * stage runs in sequence on the handle's stream, so that the per-stage times
 * of pam_fmm_evaluator_stats are those of each stage's kernels alone. */
int pam_fmm_evaluator_set_overlap(pam_fmm_evaluator* h, int overlap);

/* Per-stage device times of the last evaluate (HIP events on the handle's
 * streams); ms_total is the whole evaluation, less than the sum of the stages
 * when the near field overlaps the far field. */
typedef struct {
  double ms_tree, ms_upward, ms_m2l, ms_l2l, ms_l2p, ms_p2p, ms_total;
  int precision, order, leaf_level, direct;
  int64_t n_src_leaves, n_trg_leaves;
  int m2l_rank; /* rank of the compressed M2L, 0 when dense */
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
/* Gradient of that sum at the targets, out[3 i + k] = sum_j w_j d/dx_k phi(t_i - s_j):
 * the full_direct form (src/fmm/full_direct.hpp:7-52) of the gradient kernel
 * (include/polatory/fmm/gradient_kernel.hpp). Host pointers. */
int pam_direct_evaluate_gradient(const pam_rbf* rbf, const double* src, int64_t n_src, const double* w,
                                 const double* trg, int64_t n_trg, double* out);

/* pam_direct_evaluate with device pointers that allocates no memory. No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The transformed points and the partial sums live in a fixed
 * device work area of PAM_NOALLOC_SCRATCH_BYTES per GPU, held from the first
 * call on that GPU until the process ends. Sizes whose work area
 * (3 n_trg + 3 n_src + the partial sums, in doubles) exceeds it are refused
 * with PAM_ERR_INVALID_ARGUMENT. Work is enqueued on `stream` (a hipStream_t; NULL for
 * the default stream) and the call returns without waiting; d_out is complete
 * when the stream reaches that point. Same kernels and summation order as
 * pam_direct_evaluate, so the same bits. All calls on one GPU share the work
 * area: calls that overlap on one GPU (other streams or threads) race. */
#define PAM_NOALLOC_SCRATCH_BYTES (64LL << 20)
int pam_direct_evaluate_device_noalloc(const pam_rbf* rbf, const double* d_src, int64_t n_src,
                                       const double* d_w, const double* d_trg, int64_t n_trg,
                                       double* d_out, void* stream);

/* Pointwise kernel values at n differences diff[3 i .. 3 i + 2], evaluated on
 * the GPU: RbfBase::evaluate (rbf/rbf_base.hpp:40-43) -> out[i];
 * evaluate_gradient (:45-48) -> out[3 i + k]; evaluate_hessian (:50-53) ->
 * out[9 i + 3 r + c] (row-major). The anisotropy applies as in the reference:
 * phi(A d), A^T grad(A d), A^T H(A d) A. Host pointers. */
int pam_rbf_evaluate(const pam_rbf* rbf, const double* diff, int64_t n, double* out);
int pam_rbf_evaluate_gradient(const pam_rbf* rbf, const double* diff, int64_t n, double* out);
int pam_rbf_evaluate_hessian(const pam_rbf* rbf, const double* diff, int64_t n, double* out);


/* ---------------------------------------------------------------------------
 * Model (include/polatory/model.hpp:32-183): rbfs, polynomial degree, nugget.
 * ------------------------------------------------------------------------- */
#define PAM_POLY_DEGREE_MIN (-2) /* Model::kMinRequiredPolyDegree: cpd_order - 1 */
typedef struct pam_model pam_model;
/* Model(std::vector<Rbf>, poly_degree) + set_nugget (model.hpp:36-51, 135-141) */
int pam_model_create(const pam_rbf* rbfs, int n_rbfs, int poly_degree, double nugget,
                     pam_model** out);
int pam_model_destroy(pam_model* m);
/* poly_degree() and poly_basis_size() (model.hpp:126-131) */
int pam_model_poly_degree(const pam_model* m, int* degree);
int pam_model_poly_basis_size(const pam_model* m, int* l);
/* rbfs() and nugget() (model.hpp:79, 131-133) */
int pam_model_num_rbfs(const pam_model* m, int* n);
int pam_model_rbf(const pam_model* m, int i, pam_rbf* out);
int pam_model_nugget(const pam_model* m, double* nugget);
/* Model::save / Model::load (model.hpp:169, POLATORY_IMPLEMENT_LOAD_SAVE in
 * common/io.hpp:147-187): polatory's binary model file, field order
 * model.hpp:304-320 and rbf/rbf_io.hpp:9-35. Loading an rbf that polatory knows
 * but this library does not implement returns PAM_ERR_NOT_IMPLEMENTED. */
int pam_model_save(const pam_model* m, const char* filename);
int pam_model_load(const char* filename, pam_model** out);

/* ---------------------------------------------------------------------------
 * Interpolant (include/polatory/interpolant.hpp:34-367): fit and evaluate.
 * Weights are [w (n) ; c (l)] as in Interpolant::weights() (interpolant.hpp:212-216).
 * ------------------------------------------------------------------------- */
typedef struct pam_interpolant pam_interpolant;
typedef struct {
  pam_fmm_config op;        /* FGMRES operator (default: mixed FMM, direct FP64 when small) */
  pam_fmm_config residual;  /* FP64 residual recomputation (default: FP64 FMM / direct) */
  pam_fmm_config level;     /* RAS inter-level updates when use_level_config != 0 */
  int use_level_config;
  int use_preconditioner;   /* 1: multilevel RAS (ras_preconditioner.hpp), 0: none */
  double refine_rtol;       /* inner relative residual that restarts FGMRES on the FP64 residual */
  int verbose;              /* print the residual per iteration (solver.hpp:111-128) */
  int use_graphs;           /* replay "preconditioner + matvec" of each FGMRES step from a HIP graph */
  int keep_unconverged;     /* at max_iter keep the last iterate instead of failing (default 0) */
  int tune_solves;          /* time the RAS batched-solve workgroup sizes per level at set-up (default 0:
                               a fixed rule that reproduces the timed choice on the 1M-point fit) */
  int ras_fp32_factors;     /* RAS fine-level factors in FP32 (FP32 MFMA factorisation, FP64 solve
                               arithmetic; the coarse grid stays FP64): 1 on, 0 off, -1 (default) on
                               when the operator is not FP64 and no rbf has CPD order 2 (bh3, exp) */
  int tune_factor;          /* time the RAS batched factorisation's forms (left-looking, right-looking
                               with 32- or 64-wide panels) on a sample of each fine level's domains
                               (default 0: a fixed rule from that timing) */
  int factor_panel;         /* RAS batched factorisation form: 0 left-looking, 32 or 64 right-looking
                               with that panel width, -1 (default) the rule or tune_factor */
  int cpd2_auto;            /* 1 (default): for kernels of CPD order 2 (triharmonic3d) a mixed-precision
                               operator is raised to FP64 of order >= 10, and the RAS coarse grid has
                               4096 points from 2e5 points and 6000 from 8e5 (the reference's 2048
                               stalls there); 0: the configurations as given, 2048 coarse points */
} pam_fit_options;
int pam_fit_options_default(pam_fit_options* opt);
/* Interpolant(const Model&) — interpolant.hpp:43 */
int pam_interpolant_create(const pam_model* model, pam_interpolant** out);
int pam_interpolant_destroy(pam_interpolant* h);
/* fit(points, values, tolerance, max_iter = 100, accuracy = inf, initial = nullptr)
 * — interpolant.hpp:77-113. Absolute L-inf tolerance, recomputed in FP64.
 * opt and initial may be NULL. accuracy must be positive, as in the reference,
 * where it sets the accuracy of the residual evaluation (solver.hpp:34-41);
 * here that evaluation is always FP64 (opt->residual, by default an order-10
 * FP64 FMM or the direct sum), at least as accurate as any accuracy asks, so
 * the value does not change the fit. */
int pam_interpolant_fit(pam_interpolant* h, const double* points, int64_t n, const double* values,
                        double tolerance, int max_iter, double accuracy, const pam_fit_options* opt,
                        const pam_interpolant* initial);
/* fit(points, grad_points, values, tolerance, grad_tolerance, max_iter, accuracy,
 * grad_accuracy, initial) — interpolant.hpp:83-113 with gradient data
 * (Hermite-Birkhoff): values has mu + 3 sigma entries, the values at the points
 * then the gradients (x y z) at the grad points; tolerance bounds the value rows
 * and grad_tolerance the gradient rows (absolute L-inf, recomputed in FP64).
 * The weights are then [w (mu); g (3 sigma); c (l)]. The operator, the RAS
 * level updates and the residual are FP64 FMM (order 10, or opt's FP64
 * configurations) or direct sums; the RAS factors are FP64. accuracy and
 * grad_accuracy must be positive (see pam_interpolant_fit). opt and initial
 * may be NULL. mu may be 1 with degree 1 and sigma >= 1 (the reference's
 * special case, ras_preconditioner.hpp:81-87). */
int pam_interpolant_fit_hermite(pam_interpolant* h, const double* points, int64_t mu, const double* grad_points,
                                int64_t sigma, const double* values, double tolerance, double grad_tolerance,
                                int max_iter, double accuracy, double grad_accuracy, const pam_fit_options* opt,
                                const pam_interpolant* initial);
/* evaluate(points, grad_points) — interpolant.hpp:58-67 for any interpolant:
 * out receives n values at points, then 3 ng gradient entries (x y z) at
 * grad_points. With gradient centres the evaluation is FP64 FMM of order 10
 * (or direct sums). */
int pam_interpolant_evaluate_hermite(pam_interpolant* h, const double* points, int64_t n,
                                     const double* grad_points, int64_t ng, double* out);
/* grad_centers() — interpolant.hpp:48 */
int pam_interpolant_num_grad_centers(const pam_interpolant* h, int64_t* n);
int pam_interpolant_grad_centers(const pam_interpolant* h, double* out);
/* fit_incrementally(points, values, tolerance, max_iter = 100, accuracy = inf)
 * — interpolant.hpp:115-147, interpolation/incremental_fitter.hpp:25-183: fits on a
 * growing subset of the points; centers() then holds the points chosen, in the
 * order they were added. opt may be NULL. */
int pam_interpolant_fit_incrementally(pam_interpolant* h, const double* points, int64_t n,
                                      const double* values, double tolerance, int max_iter,
                                      double accuracy, const pam_fit_options* opt);
/* fit_incrementally(points, grad_points, values, tolerance, grad_tolerance,
 * max_iter, accuracy, grad_accuracy) — interpolant.hpp:121-147,
 * incremental_fitter.hpp:36-183 with gradient data: values as in
 * pam_interpolant_fit_hermite; centers() and grad_centers() then hold the
 * points and gradient points chosen, in the order they were added. */
int pam_interpolant_fit_incrementally_hermite(pam_interpolant* h, const double* points, int64_t mu,
                                              const double* grad_points, int64_t sigma, const double* values,
                                              double tolerance, double grad_tolerance, int max_iter,
                                              double accuracy, double grad_accuracy, const pam_fit_options* opt);
/* fit_inequality(points, values, values_lb, values_ub, tolerance, max_iter = 100,
 * accuracy = inf, initial = nullptr) — interpolant.hpp:149-190,
 * interpolation/inequality_fitter.hpp:25-187. NaN in values / values_lb / values_ub
 * means no equality / lower / upper constraint at that point. centers() then
 * holds the equality points and the active bounds. opt and initial may be NULL. */
int pam_interpolant_fit_inequality(pam_interpolant* h, const double* points, int64_t n, const double* values,
                                   const double* values_lb, const double* values_ub, double tolerance,
                                   int max_iter, double accuracy, const pam_fit_options* opt,
                                   const pam_interpolant* initial);
/* kriging::cross_validate(model, points, values, set_ids, tolerance, max_iter, accuracy)
 * — kriging/cross_validate.hpp:14-47: for each set id, fits on the points of the
 * other sets and writes the predictions at the set's points into out (n). opt may
 * be NULL. */
int pam_cross_validate(const pam_model* model, const double* points, int64_t n, const double* values,
                       const int32_t* set_ids, double tolerance, int max_iter, double accuracy,
                       const pam_fit_options* opt, double* out);
/* DistanceFilter(points).filter(distance, indices).filtered_indices()
 * (point_cloud/distance_filter.hpp:21-56): walks indices in order and drops every
 * later point within distance of a kept one. out receives the kept indices (at
 * most n_idx), *n_out their count. Host only. */
int pam_distance_filter(const double* points, int64_t n, double distance, const int64_t* indices,
                        int64_t n_idx, int64_t* out, int64_t* n_out);
/* evaluate(points, accuracy = inf) — interpolant.hpp:54-56 */
int pam_interpolant_evaluate(pam_interpolant* h, const double* points, int64_t n, double accuracy,
                             double* out);
int pam_interpolant_evaluate_device(pam_interpolant* h, const double* d_points, int64_t n,
                                    double accuracy, double* d_out);
/* The grad_points half of evaluate(points, grad_points, accuracy, grad_accuracy)
 * (interpolant.hpp:58-67, interpolation/evaluator.hpp:68-74): gradients of the
 * fitted function at n points, out has 3 n entries (x y z per point).
 * `accuracy` is the grad_accuracy: the FMM configuration is searched on the
 * gradients. */
int pam_interpolant_evaluate_gradient(pam_interpolant* h, const double* points, int64_t n, double accuracy,
                                      double* out);
/* set_evaluation_bbox_impl(bbox, accuracy, grad_accuracy) — interpolant.hpp:201-210:
 * builds the evaluators for the union of [bbox_min, bbox_max] and the centres' box,
 * at these accuracies; evaluate_impl then reuses them. This pair is the seam
 * isosurface::RbfFieldFunction calls (rbf_field_function.hpp:21-27). */
int pam_interpolant_set_evaluation_bbox(pam_interpolant* h, const double* bbox_min, const double* bbox_max,
                                        double accuracy, double grad_accuracy);
/* evaluate_impl(points) — interpolant.hpp:69-75: evaluation at the accuracy and box
 * of the last set_evaluation_bbox (or evaluate) call. */
int pam_interpolant_evaluate_impl(pam_interpolant* h, const double* points, int64_t n, double* out);
int pam_interpolant_evaluate_impl_device(pam_interpolant* h, const double* d_points, int64_t n, double* d_out);
/* Fixes the evaluation FMM configuration (NULL restores the accuracy-driven choice). */
int pam_interpolant_set_eval_config(pam_interpolant* h, const pam_fmm_config* cfg);
int pam_interpolant_num_centers(const pam_interpolant* h, int64_t* n);
/* centers() — interpolant.hpp:48-52 (n x 3) */
int pam_interpolant_centers(const pam_interpolant* h, double* out);
/* weights() — interpolant.hpp:212-216 (n + l) */
int pam_interpolant_weights(const pam_interpolant* h, double* out);
/* Installs centres and weights (e.g. from a saved interpolant, interpolant.hpp:345-367). */
int pam_interpolant_set_fitted(pam_interpolant* h, const double* centers, int64_t n,
                               const double* weights);
/* Interpolant::save / Interpolant::load (interpolant.hpp:218; field order
 * interpolant.hpp:345-367): model, fitted, centres, gradient centres, bbox,
 * weights (the weights [w; g; c] when there are gradient centres). */
int pam_interpolant_save(const pam_interpolant* h, const char* filename);
int pam_interpolant_load(const char* filename, pam_interpolant** out);
/* model() (interpolant.hpp:199) as a new handle the caller destroys */
int pam_interpolant_model(const pam_interpolant* h, pam_model** out);
int pam_interpolant_fitted(const pam_interpolant* h, int* fitted);
typedef struct {
  int iterations, restarts, n_levels;
  double residual, seconds, setup_seconds, precond_setup_seconds;
  int64_t factor_bytes;
  int graph_replays;        /* FGMRES steps replayed from the captured graph */
  int escalations;          /* precision raises after FGMRES stagnated (FP64 level updates, then operator) */
  int converged;            /* 0 when keep_unconverged stopped the fit at max_iter */
  int rounds;               /* incremental and inequality fits: rounds of solve + residuals */
  double grad_residual;     /* fits with gradient data: L-inf residual of the gradient rows */
} pam_fit_stats;
int pam_interpolant_fit_stats(const pam_interpolant* h, pam_fit_stats* out);
/* Residual after each FGMRES iteration (out may be NULL to query the length). */
int pam_interpolant_residual_history(const pam_interpolant* h, double* out, int64_t* n);

/* ---------------------------------------------------------------------------
 * Krylov and preconditioner seams (krylov::LinearOperator users).
 * ------------------------------------------------------------------------- */
/* Right-preconditioned FGMRES on the device for a dense row-major operator
 * (src/krylov/fgmres.cpp:8-26, gmres.cpp:9-50); m may be NULL (no preconditioner).
 * Runs n_iter iterations; resid_hist receives |g_j| for j = 0..iters. */
int pam_fgmres_dense(const double* a, const double* m, int64_t n, const double* rhs,
                     const double* x0, int n_iter, double* x, double* resid_hist, int* iters);
/* RasPreconditioner(model, points) — ras_preconditioner.hpp:56-189 */
typedef struct pam_ras pam_ras;
int pam_ras_create(const pam_model* model, const double* points, int64_t n, pam_ras** out);
/* The preconditioner that a fit with these options builds for these points
 * (RasPreconditioner(model, points, ...) with the fit's precision, factor form
 * and coarse-grid choices); opt must not be NULL. */
int pam_ras_create_with_options(const pam_model* model, const double* points, int64_t n,
                                const pam_fit_options* opt, pam_ras** out);
/* RasPreconditioner(model, points, grad_points) — ras_preconditioner.hpp:56-189
 * with gradient points: vectors of mu + 3 sigma + l entries. The preconditioner
 * of pam_interpolant_fit_hermite (FP64 factors and level updates); opt may be
 * NULL. */
int pam_ras_create_hermite(const pam_model* model, const double* points, int64_t mu, const double* grad_points,
                           int64_t sigma, const pam_fit_options* opt, pam_ras** out);
int pam_ras_destroy(pam_ras* h);
/* DomainDivider(points, grad_points, point_indices, grad_point_indices,
 * poly_point_indices) — domain_divider.hpp:30-48, 160-269 (host only): the
 * overlapping domains of one RAS level, with every gradient point counting 3
 * rows. Indices are int32 into points (mu) and grad_points (sigma). */
typedef struct pam_domain_divider pam_domain_divider;
int pam_domain_divider_create(const double* points, int64_t mu, const double* grad_points, int64_t sigma,
                              const int32_t* point_idcs, int64_t n_idx, const int32_t* grad_idcs, int64_t n_gidx,
                              const int32_t* poly_idcs, int64_t n_poly, pam_domain_divider** out);
int pam_domain_divider_destroy(pam_domain_divider* h);
int pam_domain_divider_num_domains(const pam_domain_divider* h, int64_t* n);
/* Domain d: its point indices (polynomial points first) and inner flags (n),
 * its gradient point indices and inner flags (ng); any output may be NULL. */
int pam_domain_divider_domain(const pam_domain_divider* h, int64_t d, int32_t* idx, uint8_t* inner, int64_t* n,
                              int32_t* gidx, uint8_t* ginner, int64_t* ng);
/* choose_coarse_points(n_coarse) — domain_divider.hpp:54-125: the coarser
 * level's points (polynomial points first) and gradient points. With idx and
 * gidx NULL, *n and *ng receive the counts; otherwise they give the
 * capacities of idx and gidx and receive the counts. */
int pam_domain_divider_choose_coarse_points(const pam_domain_divider* h, int64_t n_coarse, int32_t* idx, int64_t* n,
                                            int32_t* gidx, int64_t* ng);
int pam_ras_num_levels(const pam_ras* h, int* n_levels);
/* point indices of a level (level n_levels-1 = all points; 0 = coarse grid) */
int pam_ras_level_points(const pam_ras* h, int level, int32_t* out, int64_t* n);
int pam_ras_num_domains(const pam_ras* h, int level, int* n);
int pam_ras_domain(const pam_ras* h, int level, int dom, int32_t* idx, uint8_t* inner, int64_t* n);
/* operator()(v) — ras_preconditioner.hpp:191-250 (host vectors of length n + l) */
int pam_ras_apply(pam_ras* h, const double* v, double* out);
/* The same as a device-resident krylov::LinearOperator (linear_operator.hpp:7-22):
 * d_v and d_out are device vectors of pam_ras_size entries; the work is
 * enqueued on the handle's stream and complete after pam_ras_synchronize. */
int pam_ras_apply_device(pam_ras* h, const double* d_v, double* d_out);
/* LinearOperator::size() — n + l */
int pam_ras_size(const pam_ras* h, int64_t* size);
int pam_ras_synchronize(pam_ras* h);

/* interpolation::Operator(model, points) — operator.hpp:38-111, the fit's
 * LinearOperator: y = [(A + nugget I) w + P c ; P^T w] for x = [w ; c] at the
 * points (n x 3, host, copied), with A evaluated by the symmetric FMM of cfg
 * (NULL: the fit's default inner operator, mixed precision, FP64 direct sum
 * below 2^26 pairs). operator.hpp builds it at accuracy 0 inside Solver
 * (solver.hpp:40); pass an FP64 configuration for that form. */
typedef struct pam_operator pam_operator;
int pam_operator_create(const pam_model* model, const double* points, int64_t n, const pam_fmm_config* cfg,
                        pam_operator** out);
/* Operator(model, points, grad_points) — operator.hpp:30-36, 56-82: the
 * operator of a fit with gradient data, on vectors of mu + 3 sigma + l entries
 * [w; g (x y z per gradient point); c]:
 *   [A_vv w + nugget w + A_vg g + P_v c; A_gv w + A_gg g + P_g c; P_v^T w + P_g^T g],
 * A_vv = phi(p_i - p_j), A_vg = -grad phi(p_i - q_k)^T, A_gv = grad phi(q_k - p_j),
 * A_gg = -Hessian phi(q_k - q_m) (mat_a.hpp:11-61); P_g holds the monomial
 * gradients. cfg == NULL: FP64 FMM of order 10. */
int pam_operator_create_hermite(const pam_model* model, const double* points, int64_t mu,
                                const double* grad_points, int64_t sigma, const pam_fmm_config* cfg,
                                pam_operator** out);
int pam_operator_destroy(pam_operator* h);
/* LinearOperator::size() — n + l */
int pam_operator_size(const pam_operator* h, int64_t* size);
/* operator()(x) with device vectors of pam_operator_size entries, enqueued on
 * the handle's stream (complete after pam_operator_synchronize) */
int pam_operator_apply_device(pam_operator* h, const double* d_x, double* d_y);
/* operator()(x) with host vectors (waits for the result) */
int pam_operator_apply(pam_operator* h, const double* x, double* y);
int pam_operator_synchronize(pam_operator* h);

/* ---------------------------------------------------------------------------
 * Serving: fitted models resident in HBM, requests admitted by HBM budget and
 * run concurrently (one stream per slot). No reference counterpart; replaces a
 * loop of Interpolant::evaluate calls (interpolant.hpp:54-67), each of which
 * rebuilds the evaluator.
 * ------------------------------------------------------------------------- */
typedef struct pam_server pam_server;
typedef struct {
  int n_slots;          /* requests on the device at once (default 4) */
  int64_t hbm_budget;   /* bytes for resident models + admitted requests; 0 = 90% of free HBM */
  double rel_accuracy;  /* target relative accuracy of every answer (default 1e-6);
                           0 = FP64 FMM of order 10 */
} pam_server_options;
int pam_server_options_default(pam_server_options* opt);
int pam_server_create(const pam_server_options* opt, pam_server** out);
/* Waits for the requests in flight, then frees every model. */
int pam_server_destroy(pam_server* h);
/* Makes a fitted interpolant's model resident. Query points must lie in
 * [lo, hi] (united with the centres' bbox). cfg may be NULL (accuracy search). */
int pam_server_add_model(pam_server* h, const pam_interpolant* fitted, const double lo[3],
                         const double hi[3], const pam_fmm_config* cfg, int* model_id);
int pam_server_model_config(pam_server* h, int model_id, pam_fmm_config* cfg, int64_t* resident_bytes);
/* Queues an evaluation of n points (host, n x 3) into out (host, n); both must
 * stay valid until the request completes. */
int pam_server_submit(pam_server* h, int model_id, const double* points, int64_t n, double* out,
                      int64_t* request_id);
/* Blocks until the request completes and returns its status. */
int pam_server_wait(pam_server* h, int64_t request_id);
int pam_server_test(pam_server* h, int64_t request_id, int* done);
int pam_server_drain(pam_server* h);
int pam_server_request_times(pam_server* h, int64_t request_id, double* queue_ms, double* total_ms);
typedef struct {
  int64_t submitted, completed, failed, points;
  double p50_ms, p99_ms, mean_ms, max_ms, mean_queue_ms;
  double points_per_s, wall_s;
  int max_in_flight;
  int64_t resident_bytes, budget, peak_admitted_bytes;
  /* small requests (answered by the FP64 direct sum) of any model are served in
     shared launches: the launches, the requests answered in them, the most
     requests in one */
  int64_t batches, batched_requests;
  int max_batch;
  int64_t fmm_batches;          /* FMM batches: near fields and far-field L2Ps in shared launches */
  int64_t fmm_batched_requests; /* requests served in FMM batches */
  int max_fmm_batch;            /* most requests in one FMM batch */
} pam_server_stats;
int pam_server_get_stats(pam_server* h, pam_server_stats* out);
/* Forgets completed requests and restarts the latency statistics. */
int pam_server_reset_stats(pam_server* h);

#ifdef __cplusplus
}
#endif

#endif /* POLATORY_AMD_H */
