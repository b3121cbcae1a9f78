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

enum Precision : int { kFp64 = 0, kMixed = 1 };
enum EvalMode : int { kModeAuto = 0, kModeDirect = 1, kModeFmm = 2 };

struct FmmConfig {
  int precision = kMixed;
  int order = 6;
  int leaf_level = -1;  // -1: chosen from the point counts
  int mode = kModeAuto;
};

struct FmmStats {
  double ms_tree = 0, ms_upward = 0, ms_m2l = 0, ms_l2l = 0, ms_l2p = 0, ms_p2p = 0, ms_total = 0;
  int precision = 0, order = 0, leaf_level = 0, direct = 0;
  int64_t n_src_leaves = 0, n_trg_leaves = 0;
};

template <class T>
struct FmmState {
  DevBuf<V4<T>> spts, tpts;
  std::vector<DevBuf<T>> m, l;  // per level
  DevBuf<T> y_far;
  bool spts_valid = false, tpts_valid = false, multipoles_valid = false;
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
  void set_config(const FmmConfig& c);  // fixes the configuration (no accuracy search)
  void set_source_points(const double* xyz, int64_t n, bool on_device);
  void set_target_points(const double* xyz, int64_t n, bool on_device);
  void set_weights(const double* w, int64_t n, bool on_device);
  // out has n_trg entries (input order of the targets).
  void evaluate(double* out, int64_t n, bool on_device);

  hipStream_t stream() const { return stream_; }
  const FmmStats& stats();
  const Tree& source_tree() const { return src_tree_; }
  const Tree& target_tree() const { return symmetric_ ? src_tree_ : trg_tree_; }
  const Lists& lists() const { return lists_; }
  // Builds trees/lists for the current configuration without evaluating.
  void prepare_trees();
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
  bool use_direct() const;
  template <class T>
  void run_fmm(double* d_out);
  template <class T>
  FmmState<T>& state();
  void invalidate_expansions();
  void record(int idx);

  RbfDesc rbf_;
  bool symmetric_;
  double user_min_[3], user_max_[3];
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;

  double accuracy_;
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
  Lists lists_;
  TreeWorkspace ws_;
  DevBuf<double> direct_scratch_;

  FmmState<float> s32_;
  FmmState<double> s64_;

  hipEvent_t ev_[10] = {};
  FmmStats stats_;
  bool stats_pending_ = false;
};

// Tree height used when no leaf level is configured: target occupancy of
// about 24 points per leaf, clamped to [2, 18].
int auto_leaf_level(int64_t n_src, int64_t n_trg);

}  // namespace pam
