// Device FMM evaluator: the gfx950 replacement of polatory's
// fmm::FmmGenericEvaluator / FmmGenericSymmetricEvaluator Impl classes
// (src/fmm/fmm_evaluator.hpp:32-292, src/fmm/fmm_symmetric_evaluator.hpp:30-270).
//
// Points, weights and outputs live in HBM; host entry points copy in and out.
// Source tree, multipole expansions and M2L operators stay resident across
// evaluate() calls until the points, weights or configuration change.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "fmm.hpp"
#include "tree.hpp"

namespace pam {

// kMixedComp: kMixed with compensated (Kahan) sums in the near field, for
// strongly cancelling weights.
enum Precision : int { kFp64 = 0, kMixed = 1, kMixedComp = 2 };
enum EvalMode : int { kModeAuto = 0, kModeDirect = 1, kModeFmm = 2 };

struct FmmConfig {
  int precision = kMixed;
  int order = 6;
  int leaf_level = -1;  // -1: chosen from the point counts
  int mode = kModeAuto;
  // M2L compression: > 0 selects compressed transfers whose basis leaves an
  // estimated relative residual within this value (rank a multiple of 64, at
  // most 256); 0 = dense operators; negative = the default (kDefaultSvdTol for
  // the homogeneous kernels at orders 7 and 8, dense otherwise).
  double svd_tol = -1.0;
};

constexpr double kDefaultSvdTol = 2e-7;

// Compression tolerance actually used for a configuration and kernel.
double effective_svd_tol(const FmmConfig& c, const RbfDesc& rbf);

struct FmmStats {
  double ms_tree = 0, ms_upward = 0, ms_m2l = 0, ms_l2l = 0, ms_l2p = 0, ms_p2p = 0, ms_total = 0;
  int precision = 0, order = 0, leaf_level = 0, direct = 0;
  int64_t n_src_leaves = 0, n_trg_leaves = 0;
  int m2l_rank = 0;  // 0: dense
};

// Expansions and the light O(N P^3) stages (P2M, M2M, L2L, L2P) are FP64 in
// both forms; the mixed form runs P2P on FP32 copies of the points and M2L on
// FP32 MFMA (operands rounded to FP32, FP64 accumulation).
struct FmmState {
  DevBuf<V4<double>> spts64, tpts64;
  DevBuf<V4<float>> spts32, tpts32;
  std::vector<DevBuf<double>> m, l;    // per level, P^3 node values
  std::vector<DevBuf<float>> m32;      // per level, FP32 copy of m (mixed-form M2L operand)
  std::vector<DevBuf<double>> mt, lt;  // per level, compressed coefficients
  std::vector<DevBuf<float>> mt32;     // per level, FP32 copy of mt (mixed-form transfer operand)
  DevBuf<double> y_far;
  DevBuf<double> m2l_partial;          // split-reduction partial sums of the dense M2L
  M2LSchedule m2l_sched;
  DevBuf<float> m2l_pairs32;           // per-pair transfers of the pair-major M2L (mixed form)
  DevBuf<double> m2l_pairs64;          // per-pair transfers of the pair-major M2L (FP64 form)
  M2LPairSchedule m2l_psched;
  uint64_t m2l_sched_gen = 0;
  int m2l_sched_rpad = 0;
  bool s64_valid = false, t64_valid = false, s32_valid = false, t32_valid = false;
  bool multipoles_valid = false, m32_valid = false;
  int multipole_order = 0;
  int m_level = -1;  // leaf level the multipoles were computed for
};

class FmmEvaluator {
 public:
  // stream == nullptr: the evaluator owns a stream; otherwise all work is
  // ordered on the caller's stream (which must outlive the evaluator).
  FmmEvaluator(const RbfDesc& rbf, const double* bbox_min, const double* bbox_max, bool symmetric,
               hipStream_t stream = nullptr);
  ~FmmEvaluator();
  FmmEvaluator(const FmmEvaluator&) = delete;
  FmmEvaluator& operator=(const FmmEvaluator&) = delete;

  void set_accuracy(double accuracy);
  // Restricts the accuracy search to one precision (orders 4..12 of it); -1
  // restores the default list (mixed orders first, then FP64).
  void set_search_precision(int precision);
  void set_config(const FmmConfig& c);  // fixes the configuration (no accuracy search)
  // Near field on a second stream beside the far field (default), or all
  // stages one after another on the evaluator's stream.
  void set_overlap(bool on) { overlap_ = on; }
  void set_source_points(const double* xyz, int64_t n, bool on_device);
  void set_target_points(const double* xyz, int64_t n, bool on_device);
  void set_weights(const double* w, int64_t n, bool on_device);
  // out has n_trg entries (input order of the targets).
  void evaluate(double* out, int64_t n, bool on_device);
  // evaluate() on device memory that stops before the near field and the
  // far-field L2P and records them in *job instead (job->valid), for launches
  // shared with other evaluations (launch_deferred); the output is complete
  // once they have run. When the evaluation goes to the direct sum, the
  // output is written and job->valid is false. Source-target value
  // evaluators only; the recorded buffers stay valid until this evaluator's
  // next call.
  void evaluate_deferred(double* d_out, int64_t n, DeferredFmm* job);
  // Gradients at the targets (make_fmm_gradient_transpose_evaluator,
  // fmm/gradient_transpose_kernel.hpp): out has 3 n_trg entries, x y z per
  // target in input order. Same configuration, trees and expansions as the
  // values; source-target evaluators only.
  void evaluate_gradient(double* out, int64_t n, bool on_device);
  // Column k of the Hessian kernel with the scalar weights (the building
  // block of make_fmm_hessian_evaluator): out has 3 n_trg entries,
  // A^T sum_j w_j H(A (t_i - s_j)) A e_k per target. Source-target only.
  void evaluate_hessian_column(int k, double* out, int64_t n, bool on_device);
  // Makes this a gradient-transpose evaluator (make_fmm_evaluator.cpp:119-148):
  // evaluate() returns the gradients and the accuracy search measures them.
  void set_gradient_kernel(bool on);
  bool gradient_kernel() const { return grad_kernel_; }

  hipStream_t stream() const { return stream_; }
  const RbfDesc& rbf() const { return rbf_; }
  const FmmStats& stats();
  const Tree& source_tree() const { return src_tree_; }
  const Tree& target_tree() const { return symmetric_ ? src_tree_ : trg_tree_; }
  const Lists& lists() const { return lists_; }
  // Builds trees/lists for the current configuration without evaluating.
  void prepare_trees();
  // Builds, on stream s, the dense M2L operator that evaluations with the
  // fixed configuration will take from the process-wide operator cache, so
  // that the build can overlap other work. Only for kernels homogeneous in r
  // (their operator does not depend on the tree) and point sets large enough
  // for the FMM; otherwise does nothing. Synchronizes s.
  void prebuild_operator(hipStream_t s);
  // Builds the source tree and the multipole expansions now, so that other
  // evaluators can share them (attach_source). Synchronizes the stream.
  void prepare_source();
  // Evaluates against the source side (points, weights, tree, expansions, box,
  // configuration) of a prepared evaluator instead of an own one; this
  // evaluator then owns only its targets. The owner must stay unchanged and
  // alive while attached. Targets must lie in the owner's box. nullptr detaches.
  void attach_source(const FmmEvaluator* owner);
  const FmmConfig& active_config() const { return cfg_; }
  int64_t n_src() const { return n_src_; }
  int64_t n_trg() const { return symmetric_ ? n_src_ : n_trg_; }
  const double* d_source_points() const { return src_pts_.get(); }
  const double* d_target_points() const { return symmetric_ ? src_pts_.get() : trg_pts_.get(); }
  const double* d_weights() const { return w_.get(); }
  // Counts of near-field pairs and far-field cell pairs for the current lists.
  void count_work(int64_t* p2p_pairs, int64_t* m2l_pairs);

 private:
  void choose_config();
  void ensure_box();
  void ensure_trees();
  bool use_direct();
  // The evaluator that holds the source side (this one unless attached).
  const FmmEvaluator& S() const { return so_ ? *so_ : *this; }
  template <class TP>
  void prepare_source_expansions();
  template <class TP>
  void prepare_source_points();
  template <class TP>
  void prepare_source_multipoles();
  template <class TP>
  void run_fmm(double* d_out);
  void evaluate_impl(double* out, int64_t n, bool on_device);
  void invalidate_expansions();
  void record(int idx, hipStream_t s = nullptr);

  RbfDesc rbf_;
  bool symmetric_;
  bool grad_ = false;  // the evaluation in progress computes gradients
  bool grad_kernel_ = false;
  int hess_col_ = -1;  // >= 0: the evaluation in progress computes this Hessian column
  double user_min_[3], user_max_[3];
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;

  double accuracy_;
  int search_precision_ = -1;  // accuracy search over one precision's orders; -1: the mixed list
  bool fixed_config_ = false;
  bool config_valid_ = false;
  FmmConfig cfg_;

  int64_t n_src_ = 0, n_trg_ = 0;
  DevBuf<double> src_pts_, trg_pts_, w_, staging_, bbox_scratch_, bbox_out_;
  bool src_tree_dirty_ = true, trg_tree_dirty_ = true;
  bool have_weights_ = false;
  Box box_{};
  bool box_valid_ = false;

  Tree src_tree_, trg_tree_;
  // per-level cell counts of the trees as built before the leaf level was
  // chosen (automatic leaf level only)
  std::vector<int64_t> src_counts_, trg_counts_;
  Lists lists_;
  // all cells of each level as M2L columns (compress / decompress transforms)
  std::vector<M2LLevel> src_cols_, trg_cols_;
  TreeWorkspace ws_;
  TreeWorkspace ws_trg_;  // the target tree's, when it is built on the second stream
  DevBuf<double> direct_scratch_;

  FmmState st_;
  const FmmEvaluator* so_ = nullptr;  // shared source side (attach_source)
  uint64_t lists_gen_ = 0;            // bumped whenever the interaction lists change

  hipEvent_t ev_[10] = {};
  // The near field runs on a second stream, concurrently with the far field
  // (VALU work beside MFMA work); fork_ev_ / join_ev_ order it.
  hipStream_t p2p_stream_ = nullptr;
  bool overlap_ = true;
  DeferredFmm* defer_ = nullptr;  // evaluation in progress records its near and far field here
  hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
  FmmStats stats_;
  bool stats_pending_ = false;
  bool timing_valid_ = true;  // false when the last evaluation was captured into a graph
};

// Leaf level for uniformly spread points: about points_per_leaf points per
// leaf, clamped to [2, 18]. The evaluator starts its occupancy search there.
int auto_leaf_level(int64_t n_src, int64_t n_trg, double points_per_leaf);
// Points per occupied leaf that balance the near field against the dense M2L
// at interpolation order `order`.
double leaf_occupancy_target(int order, int precision);

}  // namespace pam
