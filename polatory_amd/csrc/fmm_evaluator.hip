// Device FMM evaluator orchestration: box, trees, lists, operator cache, the
// upward/downward passes and the accuracy-driven configuration search.
#include "fmm_evaluator.hpp"

#include <array>

#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <tuple>

#include "device_api.hpp"
#include "m2l_compress.hpp"

namespace pam {

namespace {

// ---- M2L operator cache shared by all evaluators of the process ----
template <class T>
struct OpEntry {
  M2LOperator op;
  DevBuf<T> a;
};

using OpKey = std::tuple<int, int, int, int, double, double, double>;  // dev, sizeof, P, kind, p0, p1, h

template <class T>
struct OpCache {
  std::mutex mu;
  std::map<OpKey, std::pair<std::shared_ptr<OpEntry<T>>, uint64_t>> map;
  uint64_t tick = 0;
};

template <class T>
OpCache<T>& op_cache() {
  static OpCache<T> c;
  return c;
}

template <class T>
std::shared_ptr<OpEntry<T>> get_op(int p, const RbfDesc& rbf, double h, hipStream_t s) {
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  OpKey key{dev, static_cast<int>(sizeof(T)), p, rbf.kind, rbf.p0, rbf.p1, h};
  auto& c = op_cache<T>();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.map.find(key);
  if (it != c.map.end()) {
    it->second.second = ++c.tick;
    return it->second.first;
  }
  constexpr size_t kMaxEntries = 24;
  while (c.map.size() >= kMaxEntries) {
    auto victim = c.map.begin();
    for (auto j = c.map.begin(); j != c.map.end(); ++j)
      if (j->second.second < victim->second.second) victim = j;
    c.map.erase(victim);
  }
  auto e = std::make_shared<OpEntry<T>>();
  build_m2l_dense<T>(p, rbf, h, e->op, e->a, s);
  PAM_HIP(hipStreamSynchronize(s));
  c.map[key] = {e, ++c.tick};
  return e;
}

using CKey = std::tuple<int, int, int, int, double, double, double, double>;  // + tolerance

template <class T>
struct CCache {
  std::mutex mu;
  std::map<CKey, std::shared_ptr<CompressedM2L<T>>> map;
};

template <class T>
std::shared_ptr<CompressedM2L<T>> get_compressed(int p, const RbfDesc& rbf, double h, double tol,
                                                 hipStream_t s) {
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  CKey key{dev, static_cast<int>(sizeof(T)), p, rbf.kind, rbf.p0, rbf.p1, h, tol};
  static CCache<T> c;
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.map.find(key);
  if (it != c.map.end()) return it->second;
  if (c.map.size() >= 16) c.map.erase(c.map.begin());
  auto e = std::make_shared<CompressedM2L<T>>();
  build_m2l_compressed<T>(p, rbf, h, tol, *e, s);
  c.map[key] = e;
  return e;
}

// Degree of homogeneity phi(l r) = l^deg phi(r), or 0 when phi is not homogeneous.
int homogeneity(const RbfDesc& r) {
  if ((r.kind == kBh3 || r.kind == kTh3) && r.p1 == 0.0) return r.kind == kBh3 ? 1 : 3;
  return 0;
}

__global__ void identity_cols_kernel(int32_t* cols, int64_t n, int64_t n_cols) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n_cols) cols[i] = i < n ? static_cast<int32_t>(i) : -1;
}

void build_identity_cols(const Tree& t, std::vector<M2LLevel>& out, hipStream_t s) {
  if (out.size() < static_cast<size_t>(t.level) + 1) out.resize(t.level + 1);
  for (int l = 2; l <= t.level; l++) {
    int64_t n = t.levels[l].n;
    M2LLevel& m = out[l];
    m.n_cols = (n + kM2LCols - 1) / kM2LCols * kM2LCols;
    m.col_cell.resize(static_cast<size_t>(std::max<int64_t>(m.n_cols, 1)));
    if (m.n_cols > 0) {
      identity_cols_kernel<<<ceil_div(m.n_cols, 256), 256, 0, s>>>(m.col_cell.get(), n, m.n_cols);
      PAM_LAUNCH_CHECK();
    }
  }
}

}  // namespace

double effective_svd_tol(const FmmConfig& c, const RbfDesc& rbf) {
  // Default: compressed transfers for the homogeneous kernels (one operator
  // serves every level) at orders 7 and 8, where the rank-128 basis cuts the
  // transfer flops 7 to 16 times and, measured at 1M uniform sites, moves the
  // error against the FP64 direct sum from 7.4e-7 to 8.0e-7 (mixed) and from
  // 2.6e-7 to 5.6e-7 (FP64). Other orders, and the exponential kernel, keep
  // dense operators unless a tolerance is given.
  if (c.svd_tol >= 0.0) return c.svd_tol;
  if (homogeneity(rbf) > 0 && c.order >= 7 && c.order <= 8) return kDefaultSvdTol;
  return 0.0;
}

namespace {

void make_box_from_bbox(const double* bmin, const double* bmax, const double* aniso, Box& box) {
  // src/fmm/utility.hpp:19-33: corners through A, 1.01 x the largest extent.
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; c++) {
    double p[3] = {(c & 1) ? bmax[0] : bmin[0], (c & 2) ? bmax[1] : bmin[1], (c & 4) ? bmax[2] : bmin[2]};
    for (int k = 0; k < 3; k++) {
      double t = (aniso[3 * k] * p[0] + aniso[3 * k + 1] * p[1]) + aniso[3 * k + 2] * p[2];
      lo[k] = std::min(lo[k], t);
      hi[k] = std::max(hi[k], t);
    }
  }
  double wmax = 0.0, center[3];
  for (int k = 0; k < 3; k++) {
    double wk = hi[k] - lo[k];
    wmax = std::max(wmax, wk);
    center[k] = lo[k] + wk / 2.0;
  }
  double w = 1.01 * wmax;
  if (w == 0.0) w = 1.0;
  box.width = w;
  for (int k = 0; k < 3; k++) box.lo[k] = center[k] - w / 2.0;
}

}  // namespace

int auto_leaf_level(int64_t n_src, int64_t n_trg, double points_per_leaf) {
  double n = static_cast<double>(std::max<int64_t>(std::max(n_src, n_trg), 1));
  int l = static_cast<int>(std::lround(std::log(n / points_per_leaf) / std::log(8.0)));
  return std::clamp(l, 2, 18);
}

double leaf_occupancy_target(int order, int precision) {
  // Per leaf of n points the near field costs ~27 n^2 pair evaluations (rate
  // R_pair) and the far field ~1.14 x 189 dense transfers of 2 P^6 flops
  // (rate R_m2l); the cost per point is least at n = P^3 sqrt(16 R_pair/R_m2l):
  // 0.8 P^3 in the mixed form (1.9e12 pairs/s, 4.8e13 flop/s measured on
  // MI355X), 0.66 P^3 in FP64 (6.6e11 pairs/s, 2.4e13 flop/s).
  const double c = precision == kFp64 ? 0.66 : 0.8;
  return std::max(64.0, c * static_cast<double>(order) * order * order);
}

FmmEvaluator::FmmEvaluator(const RbfDesc& rbf, const double* bbox_min, const double* bbox_max,
                           bool symmetric, hipStream_t stream)
    : rbf_(rbf), symmetric_(symmetric), accuracy_(std::numeric_limits<double>::infinity()) {
  for (int k = 0; k < 3; k++) {
    user_min_[k] = bbox_min ? bbox_min[k] : INFINITY;
    user_max_[k] = bbox_max ? bbox_max[k] : -INFINITY;
  }
  if (stream) {
    stream_ = stream;
    own_stream_ = false;
  } else {
    stream_ = stream_acquire();
  }
  for (auto& e : ev_) PAM_HIP(hipEventCreate(&e));
  p2p_stream_ = stream_acquire();
  PAM_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  PAM_HIP(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
  bbox_out_.resize(12);
}

FmmEvaluator::~FmmEvaluator() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (p2p_stream_) (void)hipStreamSynchronize(p2p_stream_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (join_ev_) (void)hipEventDestroy(join_ev_);
  if (p2p_stream_) stream_release(p2p_stream_);
  if (stream_ && own_stream_) stream_release(stream_);
}

void FmmEvaluator::set_accuracy(double accuracy) {
  if (!(accuracy >= 0.0)) throw_invalid("accuracy must be nonnegative");
  accuracy_ = accuracy;
  if (!fixed_config_) config_valid_ = false;
}

void FmmEvaluator::set_search_precision(int precision) {
  if (precision != -1 && precision != kFp64 && precision != kMixed && precision != kMixedComp)
    throw_invalid("unknown precision");
  search_precision_ = precision;
  fixed_config_ = false;
  config_valid_ = false;
}

void FmmEvaluator::set_config(const FmmConfig& c) {
  if (c.precision != kFp64 && c.precision != kMixed && c.precision != kMixedComp) throw_invalid("unknown precision");
  if (!order_supported(c.order)) throw_invalid("unsupported interpolation order");
  if (c.leaf_level != -1 && (c.leaf_level < 2 || c.leaf_level > 18)) throw_invalid("leaf level out of range");
  FmmConfig old = cfg_;
  cfg_ = c;
  fixed_config_ = true;
  config_valid_ = true;
  if (old.leaf_level != c.leaf_level) src_tree_dirty_ = trg_tree_dirty_ = true;
  if (old.order != c.order || old.leaf_level != c.leaf_level) invalidate_expansions();
}

void FmmEvaluator::invalidate_expansions() {
  st_.multipoles_valid = false;
  st_.s64_valid = st_.s32_valid = false;
  st_.t64_valid = st_.t32_valid = false;
}

void FmmEvaluator::set_source_points(const double* xyz, int64_t n, bool on_device) {
  if (so_) throw_runtime("this evaluator shares another evaluator's sources");
  if (n < 0) throw_invalid("negative point count");
  n_src_ = n;
  src_pts_.resize(static_cast<size_t>(n) * 3);
  const double* d_in = xyz;
  if (!on_device) {
    staging_.upload(xyz, static_cast<size_t>(n) * 3, stream_);
    d_in = staging_.get();
  }
  launch_transform_points(rbf_.aniso, d_in, n, src_pts_.get(), stream_);
  if (n > 0) launch_bbox(src_pts_.get(), n, bbox_out_.get(), bbox_scratch_, stream_);
  src_tree_dirty_ = true;
  if (symmetric_) trg_tree_dirty_ = true;
  box_valid_ = false;
  have_weights_ = false;
  invalidate_expansions();
  if (!fixed_config_) config_valid_ = false;
  // A host input is staged through pinned memory that the next call reuses.
  if (!on_device) PAM_HIP(hipStreamSynchronize(stream_));
}

void FmmEvaluator::set_target_points(const double* xyz, int64_t n, bool on_device) {
  if (symmetric_) throw_runtime("set_target_points on a symmetric evaluator");
  if (n < 0) throw_invalid("negative point count");
  n_trg_ = n;
  trg_pts_.resize(static_cast<size_t>(n) * 3);
  const double* d_in = xyz;
  if (!on_device) {
    staging_.upload(xyz, static_cast<size_t>(n) * 3, stream_);
    d_in = staging_.get();
  }
  launch_transform_points(rbf_.aniso, d_in, n, trg_pts_.get(), stream_);
  if (n > 0) launch_bbox(trg_pts_.get(), n, bbox_out_.get() + 6, bbox_scratch_, stream_);
  trg_tree_dirty_ = true;
  box_valid_ = false;
  st_.t64_valid = st_.t32_valid = false;
  if (!on_device) PAM_HIP(hipStreamSynchronize(stream_));
}

void FmmEvaluator::set_weights(const double* w, int64_t n, bool on_device) {
  if (so_) throw_runtime("this evaluator shares another evaluator's sources");
  if (n != n_src_) throw_invalid("weights.rows() must equal the number of source points");
  w_.resize(static_cast<size_t>(n));
  if (n > 0) {
    PAM_HIP(hipMemcpyAsync(w_.get(), w, sizeof(double) * n,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  }
  have_weights_ = true;
  // Point coordinates stay; only the weights and the multipoles change.
  if (st_.s32_valid) launch_set_point_weights<float>(src_tree_, w_.get(), st_.spts32.get(), stream_);
  if (st_.s64_valid) launch_set_point_weights<double>(src_tree_, w_.get(), st_.spts64.get(), stream_);
  st_.multipoles_valid = false;
  if (!on_device) PAM_HIP(hipStreamSynchronize(stream_));
}

bool FmmEvaluator::use_direct() {
  if (cfg_.mode == kModeDirect) return true;
  if (cfg_.mode == kModeFmm) return false;
  const int64_t ns = S().n_src_;
  const int64_t nt = symmetric_ ? n_src_ : n_trg_;
  const double pairs = static_cast<double>(ns) * static_cast<double>(nt);
  // Below ~2^26 pairs the FP64 direct sum is both exact and cheaper on one GPU.
  if (pairs < 67108864.0) return true;
  if (symmetric_) return false;
  // Few or clustered targets against many sources: the far field costs one
  // dense transfer set (189 x 2 P^6 flops) per occupied target cell and level
  // whatever the number of targets in it, the FP64 direct sum n_src pair
  // evaluations per target. The trees are built to count the occupied cells
  // (they are needed for the FMM anyway). Rates measured on MI355X.
  ensure_trees();
  const FmmEvaluator& src = S();
  const Tree& tt = target_tree();
  const int P = cfg_.order;
  const int L = tt.level;
  double cells = 0.0;
  for (int l = 2; l <= L; l++) cells += static_cast<double>(tt.levels[l].n);
  const bool fp64 = cfg_.precision == kFp64;
  const double p6 = std::pow(static_cast<double>(P), 6);
  const double m2l_s = cells * kM2LSlots * 2.0 * p6 / (fp64 ? 2.4e13 : 4.8e13);
  const double src_leaves = std::max<double>(1.0, static_cast<double>(src.src_tree_.levels[L].n));
  const double p2p_s = static_cast<double>(nt) * 27.0 * (static_cast<double>(ns) / src_leaves) /
                       (fp64 ? 6.6e11 : 1.9e12);
  const double direct_s = pairs / 8e11;
  return direct_s < m2l_s + p2p_s;
}

void FmmEvaluator::ensure_box() {
  if (box_valid_) return;
  double pb[12];
  PAM_HIP(hipMemcpyAsync(pb, bbox_out_.get(), sizeof pb, hipMemcpyDeviceToHost, stream_));
  PAM_HIP(hipStreamSynchronize(stream_));
  if (so_) {
    const Box& b = so_->box_;
    if (n_trg_ > 0) {
      for (int k = 0; k < 3; k++)
        if (pb[6 + k] < b.lo[k] || pb[9 + k] > b.lo[k] + b.width)
          throw_runtime("target points lie outside the box of the shared source tree");
    }
    bool same = b.width == box_.width && b.lo[0] == box_.lo[0] && b.lo[1] == box_.lo[1] && b.lo[2] == box_.lo[2];
    if (!same) {
      box_ = b;
      trg_tree_dirty_ = true;
      st_.t32_valid = st_.t64_valid = false;
    }
    box_valid_ = true;
    return;
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  auto merge = [&](const double* b) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], b[k]);
      hi[k] = std::max(hi[k], b[3 + k]);
    }
  };
  if (n_src_ > 0) merge(pb);
  if (!symmetric_ && n_trg_ > 0) merge(pb + 6);
  Box box{};
  bool user_ok = user_min_[0] <= user_max_[0] && user_min_[1] <= user_max_[1] && user_min_[2] <= user_max_[2];
  if (user_ok) {
    make_box_from_bbox(user_min_, user_max_, rbf_.aniso, box);
    bool inside = true;
    for (int k = 0; k < 3; k++)
      if (lo[k] < box.lo[k] || hi[k] > box.lo[k] + box.width) inside = false;
    if (!inside) user_ok = false;
  }
  if (!user_ok) {
    // Points outside the given bbox (or none given): box of the transformed points.
    double ident[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (!(lo[0] <= hi[0])) {
      for (int k = 0; k < 3; k++) lo[k] = hi[k] = 0.0;
    }
    make_box_from_bbox(lo, hi, ident, box);
  }
  bool same = box.width == box_.width && box.lo[0] == box_.lo[0] && box.lo[1] == box_.lo[1] && box.lo[2] == box_.lo[2];
  if (!same) {
    box_ = box;
    src_tree_dirty_ = trg_tree_dirty_ = true;
    invalidate_expansions();
  }
  box_valid_ = true;
}

void FmmEvaluator::ensure_trees() {
  ensure_box();
  if (so_) {
    const Tree& stree = so_->src_tree_;
    const int L = stree.level;
    bool rebuilt = false;
    if (trg_tree_dirty_ || trg_tree_.level != L) {
      build_tree(trg_tree_, trg_pts_.get(), n_trg_, box_, L, ws_, stream_);
      build_identity_cols(trg_tree_, trg_cols_, stream_);
      trg_tree_dirty_ = false;
      rebuilt = true;
      st_.t32_valid = st_.t64_valid = false;
    }
    if (rebuilt || lists_.level != L) {
      build_lists(trg_tree_, stree, lists_, ws_, stream_, effective_svd_tol(cfg_, rbf_) > 0.0);
      lists_gen_++;
    }
    return;
  }
  bool rebuilt = false;
  // The two trees are independent: with the stage overlap on, the target
  // tree is built on the second stream with its own workspace while the
  // source tree is built here.
  const bool par = overlap_ && !symmetric_;
  hipStream_t ts = par ? p2p_stream_ : stream_;
  TreeWorkspace& tws = par ? ws_trg_ : ws_;
  auto fork = [&]() {
    if (!par) return;
    PAM_HIP(hipEventRecord(fork_ev_, stream_));
    PAM_HIP(hipStreamWaitEvent(ts, fork_ev_, 0));
  };
  auto join = [&]() {
    if (!par) return;
    PAM_HIP(hipEventRecord(join_ev_, ts));
    PAM_HIP(hipStreamWaitEvent(stream_, join_ev_, 0));
  };
  auto build_src = [&](int level) {
    build_tree(src_tree_, src_pts_.get(), n_src_, box_, level, ws_, stream_);
    src_tree_dirty_ = false;
    rebuilt = true;
    st_.s32_valid = st_.s64_valid = false;
    st_.multipoles_valid = false;
  };
  auto build_trg = [&](int level) {
    build_tree(trg_tree_, trg_pts_.get(), n_trg_, box_, level, tws, ts);
    trg_tree_dirty_ = false;
    rebuilt = true;
    st_.t32_valid = st_.t64_valid = false;
  };
  int L;
  if (cfg_.leaf_level >= 0) {
    L = cfg_.leaf_level;
    fork();
    if (!symmetric_ && (trg_tree_dirty_ || trg_tree_.level != L)) build_trg(L);
    if (src_tree_dirty_ || src_tree_.level != L) build_src(L);
    join();
    src_counts_.clear();
    trg_counts_.clear();
  } else {
    // Leaf level from the occupied cells: the points are sorted by their
    // codes kLevelsBeyond levels below the uniform-density choice, the cells
    // of every level are counted on the device (one read-back for both
    // trees), the level whose occupied leaves hold closest to kPointsPerLeaf
    // points (in ratio) is kept, and the trees are built at it. Clustered or
    // surface-like point sets get deeper trees than their point count alone
    // suggests.
    constexpr int kLevelsBeyond = 2;
    const double kPointsPerLeaf = leaf_occupancy_target(cfg_.order, cfg_.precision);
    const int L0 = auto_leaf_level(n_src_, n_trg(), kPointsPerLeaf);
    const int Lmax = std::min(L0 + kLevelsBeyond, 18);
    auto sort_both = [&](bool do_src, bool do_trg) {
      fork();
      if (do_trg) sort_tree_points(trg_tree_, trg_pts_.get(), n_trg_, box_, Lmax, tws, ts);
      if (do_src) sort_tree_points(src_tree_, src_pts_.get(), n_src_, box_, Lmax, ws_, stream_);
      std::vector<int64_t> hs(static_cast<size_t>(Lmax) + 1), ht(static_cast<size_t>(Lmax) + 1);
      const size_t bytes = hs.size() * sizeof(int64_t);
      if (do_trg) PAM_HIP(hipMemcpyAsync(ht.data(), trg_tree_.open_hist.get(), bytes, hipMemcpyDeviceToHost, ts));
      if (do_src) PAM_HIP(hipMemcpyAsync(hs.data(), src_tree_.open_hist.get(), bytes, hipMemcpyDeviceToHost, stream_));
      join();
      PAM_HIP(hipStreamSynchronize(stream_));
      if (do_src) {
        src_counts_ = level_counts(hs);
        src_tree_dirty_ = false;
        st_.s32_valid = st_.s64_valid = false;
        st_.multipoles_valid = false;
      }
      if (do_trg) {
        trg_counts_ = level_counts(ht);
        trg_tree_dirty_ = false;
        st_.t32_valid = st_.t64_valid = false;
      }
    };
    const bool sort_src = src_tree_dirty_ || static_cast<int>(src_counts_.size()) != Lmax + 1;
    const bool sort_trg =
        !symmetric_ && (trg_tree_dirty_ || static_cast<int>(trg_counts_.size()) != Lmax + 1);
    if (sort_src || sort_trg) sort_both(sort_src, sort_trg);
    L = L0;
    double best = INFINITY;
    for (int l = L0; l <= Lmax; l++) {
      double occ = 0.0;
      if (n_src_ > 0 && src_counts_[l] > 0) occ = std::max(occ, static_cast<double>(n_src_) / src_counts_[l]);
      if (!symmetric_ && n_trg_ > 0 && trg_counts_[l] > 0)
        occ = std::max(occ, static_cast<double>(n_trg_) / trg_counts_[l]);
      if (occ <= 0.0) continue;
      double d = std::abs(std::log(occ / kPointsPerLeaf));
      if (d < best) {
        best = d;
        L = l;
      }
    }
    // Points sorted at a level coarser than L (a tree finished earlier at a
    // coarser leaf level) are sorted again.
    const bool resort_src = src_tree_.sort_level < L;
    const bool resort_trg = !symmetric_ && trg_tree_.sort_level < L;
    if (resort_src || resort_trg) sort_both(resort_src, resort_trg);
    fork();
    if (!symmetric_ && (sort_trg || resort_trg || trg_tree_.level != L)) {
      finish_tree(trg_tree_, L, trg_counts_, tws, ts);
      rebuilt = true;
      st_.t32_valid = st_.t64_valid = false;
    }
    if (sort_src || resort_src || src_tree_.level != L) {
      finish_tree(src_tree_, L, src_counts_, ws_, stream_);
      rebuilt = true;
      st_.s32_valid = st_.s64_valid = false;
      st_.multipoles_valid = false;
    }
    join();
  }
  if (symmetric_) trg_tree_dirty_ = false;
  if (rebuilt) {
    build_identity_cols(src_tree_, src_cols_, stream_);
    if (!symmetric_) build_identity_cols(trg_tree_, trg_cols_, stream_);
  }
  if (rebuilt || lists_.level != L) {
    build_lists(target_tree(), src_tree_, lists_, ws_, stream_, effective_svd_tol(cfg_, rbf_) > 0.0);
    lists_gen_++;
  }
}

void FmmEvaluator::prepare_trees() {
  if (!config_valid_) choose_config();
  ensure_trees();
  PAM_HIP(hipStreamSynchronize(stream_));
}

void FmmEvaluator::prebuild_operator(hipStream_t s) {
  if (!fixed_config_ || cfg_.mode == kModeDirect || effective_svd_tol(cfg_, rbf_) > 0.0 || homogeneity(rbf_) == 0)
    return;
  // Below 2^16 points the tree may have no level with well-separated cells.
  const int64_t nt = symmetric_ ? n_src_ : n_trg_;
  if (std::min(n_src_, nt) < (int64_t{1} << 16)) return;
  if (cfg_.precision == kFp64) get_op<double>(cfg_.order, rbf_, 1.0, s);
  else get_op<float>(cfg_.order, rbf_, 1.0, s);
}

void FmmEvaluator::prepare_source() {
  if (so_) throw_runtime("prepare_source on an evaluator that shares another's sources");
  if (!have_weights_ && n_src_ > 0) throw_runtime("weights are not set");
  if (!config_valid_) choose_config();
  // Built unless the configuration forces the direct sum: whether a given
  // target set is evaluated directly is decided per evaluation.
  if (n_src_ > 0 && cfg_.mode != kModeDirect) {
    ensure_trees();
    if (cfg_.precision == kFp64) prepare_source_expansions<double>();
    else prepare_source_expansions<float>();
  }
  PAM_HIP(hipStreamSynchronize(stream_));
}

void FmmEvaluator::attach_source(const FmmEvaluator* owner) {
  if (owner == this) owner = nullptr;
  if (symmetric_) throw_runtime("a symmetric evaluator cannot share sources");
  if (owner && (owner->symmetric_ || owner->so_)) throw_runtime("the owner must be a non-symmetric, non-attached evaluator");
  if (owner == so_) return;
  so_ = owner;
  if (owner) {
    rbf_ = owner->rbf_;
    cfg_ = owner->cfg_;
    fixed_config_ = true;
    config_valid_ = true;
  }
  box_valid_ = false;
  trg_tree_dirty_ = true;
  lists_.level = -1;  // rebuilt on the next evaluation (buffers kept)
  lists_gen_++;
  st_.t32_valid = st_.t64_valid = false;
  // Targets are transformed with the rbf's anisotropy: they must be set again.
  n_trg_ = 0;
}


void FmmEvaluator::record(int idx, hipStream_t s) {
  // Stage timing is skipped while the stream is being captured into a graph.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  PAM_HIP(hipStreamIsCapturing(stream_, &cs));
  if (cs != hipStreamCaptureStatusNone) {
    timing_valid_ = false;
    return;
  }
  PAM_HIP(hipEventRecord(ev_[idx], s ? s : stream_));
}

template <class TP>
void FmmEvaluator::prepare_source_expansions() {
  prepare_source_points<TP>();
  prepare_source_multipoles<TP>();
}

template <class TP>
void FmmEvaluator::prepare_source_points() {
  FmmState& st = st_;
  const int P = cfg_.order;
  const int kp = kpad_of(P);
  const int L = src_tree_.level;
  if (st.m.size() < static_cast<size_t>(L) + 1) st.m.resize(L + 1);
  if (st.m_level != L || st.multipole_order != P) {
    st.m_level = L;
    st.multipoles_valid = false;
    st.multipole_order = P;
  }
  if (!st.s64_valid) {
    st.spts64.resize(n_src_);
    launch_prepare_points<double>(src_tree_, src_pts_.get(), w_.get(), st.spts64.get(), stream_);
    st.s64_valid = true;
    st.multipoles_valid = false;
  }
  if constexpr (sizeof(TP) == 4) {
    if (!st.s32_valid) {
      st.spts32.resize(n_src_);
      launch_prepare_points<float>(src_tree_, src_pts_.get(), w_.get(), st.spts32.get(), stream_);
      st.s32_valid = true;
    }
  }
}

template <class TP>
void FmmEvaluator::prepare_source_multipoles() {
  FmmState& st = st_;
  const int P = cfg_.order;
  const int kp = kpad_of(P);
  const int L = src_tree_.level;
  if (!st.multipoles_valid) {
    for (int l = 2; l <= L; l++) st.m[l].resize(static_cast<size_t>(src_tree_.levels[l].n) * kp);
    launch_p2m<double>(P, src_tree_, st.spts64.get(), st.m[L].get(), stream_);
    for (int l = L - 1; l >= 2; l--) launch_m2m<double>(P, src_tree_, l, st.m[l + 1].get(), st.m[l].get(), stream_);
    st.multipoles_valid = true;
    st.m32_valid = false;
  }
  if constexpr (sizeof(TP) == 4) {
    if (!st.m32_valid) {
      st.m32.resize(L + 1);
      for (int l = 2; l <= L; l++) {
        size_t n = static_cast<size_t>(src_tree_.levels[l].n) * kp;
        st.m32[l].resize(n);
        launch_convert_f64_f32(st.m[l].get(), st.m32[l].get(), static_cast<int64_t>(n), stream_);
      }
      st.m32_valid = true;
    }
  }
}

template <class TP>
void FmmEvaluator::run_fmm(double* d_out) {
  FmmState& st = st_;
  const int P = cfg_.order;
  const int kp = kpad_of(P);
  ensure_trees();
  const Tree& tt = target_tree();
  const Tree& stree = S().src_tree_;
  const FmmState& sst = S().st_;
  const int L = stree.level;
  record(1);
  if (!so_) {
    prepare_source_points<TP>();
  } else if (!sst.multipoles_valid || sst.multipole_order != P || !sst.s64_valid ||
             (sizeof(TP) == 4 && (!sst.s32_valid || !sst.m32_valid))) {
    throw_runtime("the shared source side is not prepared (call prepare_source on the owner)");
  }
  if (st.l.size() < static_cast<size_t>(L) + 1) st.l.resize(L + 1);
  if (!symmetric_ && !st.t64_valid) {
    st.tpts64.resize(n_trg_);
    launch_prepare_points<double>(trg_tree_, trg_pts_.get(), nullptr, st.tpts64.get(), stream_);
    st.t64_valid = true;
  }
  const V4<double>* tpts64 = symmetric_ ? sst.spts64.get() : st.tpts64.get();
  const V4<TP>* sp2p;
  const V4<TP>* tp2p;
  if constexpr (sizeof(TP) == 4) {
    if (!symmetric_ && !st.t32_valid) {
      st.tpts32.resize(n_trg_);
      launch_prepare_points<float>(trg_tree_, trg_pts_.get(), nullptr, st.tpts32.get(), stream_);
      st.t32_valid = true;
    }
    sp2p = sst.spts32.get();
    tp2p = symmetric_ ? sst.spts32.get() : st.tpts32.get();
  } else {
    sp2p = sst.spts64.get();
    tp2p = tpts64;
  }
  if (!so_) prepare_source_multipoles<TP>();
  record(2);
  // Fork: the near field (VALU work) writes the output on the second stream
  // while M2L (MFMA work) and L2L run here; L2P adds the far field after the
  // join (near + far, the same sum in the same order as one stream would do).
  // Without overlap the near field runs on this stream after L2L.
  auto near_field = [&](hipStream_t ps) {
    if (defer_) {
      // Recorded for a launch shared with other evaluations (value kernel).
      defer_->fp32 = sizeof(TP) == 4;
      defer_->compensated = cfg_.precision == kMixedComp;
      defer_->kind = rbf_.kind;
      if constexpr (sizeof(TP) == 4) defer_->p2p32 = make_p2p_job<float>(rbf_, tt, stree, lists_, tp2p, sp2p, nullptr, d_out);
      else defer_->p2p64 = make_p2p_job<double>(rbf_, tt, stree, lists_, tp2p, sp2p, nullptr, d_out);
      return;
    }
    PAM_HIP(hipEventRecord(fork_ev_, stream_));
    PAM_HIP(hipStreamWaitEvent(ps, fork_ev_, 0));
    record(7, ps);
    if (hess_col_ >= 0) launch_p2p_hess_col<TP>(rbf_, tt, stree, lists_, tp2p, sp2p, hess_col_, d_out, ps);
    else if (grad_) launch_p2p_grad<TP>(rbf_, tt, stree, lists_, tp2p, sp2p, d_out, ps);
    else launch_p2p<TP>(rbf_, tt, stree, lists_, tp2p, sp2p, nullptr, d_out, ps, cfg_.precision == kMixedComp);
    record(8, ps);
    PAM_HIP(hipEventRecord(join_ev_, ps));
  };
  const bool overlap = overlap_ && !defer_;
  if (overlap) near_field(p2p_stream_);
  const int deg = homogeneity(rbf_);
  const int32_t* vl = octant_vlist_device(stream_);
  const double tol = effective_svd_tol(cfg_, rbf_);
  stats_.m2l_rank = 0;
  for (int l = 2; l <= L; l++) st.l[l].resize(static_cast<size_t>(tt.levels[l].n) * kp);
  if (tol > 0.0) {
    // SVD-compressed M2L: compress multipoles, transfer in rank r, expand locals.
    if (!lists_.has_columns) {
      build_lists(tt, stree, lists_, ws_, stream_, true);
      lists_gen_++;
    }
    if (st.mt.size() < static_cast<size_t>(L) + 1) st.mt.resize(L + 1);
    if (st.lt.size() < static_cast<size_t>(L) + 1) st.lt.resize(L + 1);
    if (sizeof(TP) == 4 && st.mt32.size() < static_cast<size_t>(L) + 1) st.mt32.resize(L + 1);
    std::vector<std::shared_ptr<CompressedM2L<TP>>> ops(L + 1);
    std::shared_ptr<CompressedM2L<TP>> unit;
    if (deg > 0) unit = get_compressed<TP>(P, rbf_, 1.0, tol, stream_);
    // Compression and expansion in FP64 (per cell, cheap); the transfers (per
    // cell pair) in the form's operand type. An FP32 expansion of the summed
    // transfers measured 1.25e-6 against 7.4e-7 uncompressed at 1M sites.
    std::vector<M2LLevelSpec<double, double>> cmp, exp;
    std::vector<M2LDenseLevel<TP>> xfer;
    for (int l = 2; l <= L; l++) {
      double h = box_.width / std::ldexp(1.0, l) / 2.0;
      ops[l] = deg > 0 ? unit : get_compressed<TP>(P, rbf_, h, tol, stream_);
      const int kr = ops[l]->kpad_r;
      stats_.m2l_rank = ops[l]->r;
      size_t nmt = static_cast<size_t>(stree.levels[l].n) * kr;
      if (st.mt[l].size() != nmt) {
        st.mt[l].resize(nmt);
        st.mt[l].zero(stream_);
      }
      size_t nlt = static_cast<size_t>(tt.levels[l].n) * kr;
      if (st.lt[l].size() != nlt) {
        st.lt[l].resize(nlt);
        st.lt[l].zero(stream_);
      }
      double scale = deg > 0 ? std::pow(h, deg) : 1.0;
      cmp.push_back({&S().src_cols_[l], nullptr, sst.m[l].get(), st.mt[l].get(), ops[l]->ut.get(), 1.0});
      const TP* in;
      if constexpr (sizeof(TP) == 4) {
        st.mt32[l].resize(nmt);
        in = st.mt32[l].get();
      } else {
        in = st.mt[l].get();
      }
      xfer.push_back({&lists_.m2l[l], in, st.lt[l].get(), ops[l]->c.get(), scale});
      exp.push_back({symmetric_ ? &src_cols_[l] : &trg_cols_[l], nullptr, st.lt[l].get(), st.l[l].get(),
                     ops[l]->u.get(), 1.0});
    }
    if (L >= 2) {
      const auto& o = *ops[2];
      launch_m2l_batch<double, double>(o.ut_op, cmp, vl, false, kp, o.kpad_r, stream_);
      if constexpr (sizeof(TP) == 4)
        for (int l = 2; l <= L; l++)
          launch_convert_f64_f32(st.mt[l].get(), st.mt32[l].get(), static_cast<int64_t>(st.mt[l].size()), stream_);
      if (st.m2l_sched_gen != lists_gen_ || st.m2l_sched_rpad != o.c_op.rpad) {
        st.m2l_psched.valid = false;
        st.m2l_sched_gen = lists_gen_;
        st.m2l_sched_rpad = o.c_op.rpad;
      }
      if constexpr (sizeof(TP) == 4) launch_m2l_pairs<float, float>(o.c_op, xfer, st.m2l_psched, st.m2l_pairs32, stream_);
      else launch_m2l_pairs<double, double>(o.c_op, xfer, st.m2l_psched, st.m2l_pairs64, stream_);
      launch_m2l_batch<double, double>(o.u_op, exp, vl, false, o.kpad_r, kp, stream_);
    }
  } else {
    std::vector<std::shared_ptr<OpEntry<TP>>> ops(L + 1);
    std::shared_ptr<OpEntry<TP>> unit_op;
    if (deg > 0) unit_op = get_op<TP>(P, rbf_, 1.0, stream_);
    std::vector<M2LDenseLevel<TP>> xfer;
    for (int l = 2; l <= L; l++) {
      double h = box_.width / std::ldexp(1.0, l) / 2.0;
      ops[l] = deg > 0 ? unit_op : get_op<TP>(P, rbf_, h, stream_);
      double scale = deg > 0 ? std::pow(h, deg) : 1.0;
      const TP* in;
      if constexpr (sizeof(TP) == 4) in = sst.m32[l].get();
      else in = sst.m[l].get();
      xfer.push_back({&lists_.m2l[l], in, st.l[l].get(), ops[l]->a.get(), scale});
    }
    if (L >= 2) {
      // The work list follows the lists generation and the operator shape.
      if (st.m2l_sched_gen != lists_gen_ || st.m2l_sched_rpad != ops[2]->op.rpad) {
        st.m2l_psched.valid = false;
        st.m2l_sched_gen = lists_gen_;
        st.m2l_sched_rpad = ops[2]->op.rpad;
      }
      // The mixed form keeps the per-pair transfers in FP32: on the 500k-site
      // kriging prediction, FP64 per-pair storage left the error unchanged
      // (2.8e-4, set by the FP32 near field of strongly cancelling weights).
      if constexpr (sizeof(TP) == 4) {
        launch_m2l_pairs<float, float>(ops[2]->op, xfer, st.m2l_psched, st.m2l_pairs32, stream_);
      } else {
        launch_m2l_pairs<double, double>(ops[2]->op, xfer, st.m2l_psched, st.m2l_pairs64, stream_);
      }
    }
  }
  record(3);
  for (int l = 3; l <= L; l++) launch_l2l<double>(P, tt, l, st.l[l - 1].get(), st.l[l].get(), stream_);
  record(4);
  if (!overlap) near_field(stream_);
  if (defer_) {
    defer_->l2p = make_l2p_job<double>(tt, st.l[L].get(), tpts64, d_out, true, kp);
    defer_->order = P;
    defer_->valid = true;
    for (int i = 5; i <= 8; i++) record(i);
    return;
  }
  // Join the near field, then add the far field to it.
  PAM_HIP(hipStreamWaitEvent(stream_, join_ev_, 0));
  record(5);
  if (hess_col_ >= 0) launch_l2p_hess_col(P, tt, st.l[L].get(), tpts64, rbf_.aniso, hess_col_, d_out, stream_);
  else if (grad_) launch_l2p_grad(P, tt, st.l[L].get(), tpts64, rbf_.aniso, d_out, stream_);
  else launch_l2p<double>(P, tt, st.l[L].get(), tpts64, d_out, stream_, true);
  record(6);
}

void FmmEvaluator::choose_config() {
  FmmConfig c;
  c.leaf_level = -1;
  if (std::isinf(accuracy_)) {
    c.precision = kMixed;
    c.order = 6;
  } else if (accuracy_ == 0.0) {
    c.precision = kFp64;
    c.order = 10;
  } else {
    // src/fmm/fmm_accuracy_estimator.hpp:74-121: try candidate configurations on
    // up to 10,000 source points used as targets, keep the first whose L-inf
    // error against the direct sum is within the accuracy.
    if (!have_weights_) throw_runtime("weights must be set before a finite accuracy is tuned");
    int64_t m = std::min<int64_t>(n_src_, 10000);
    c.precision = kMixed;
    c.order = 6;
    if (m > 0) {
      // Evenly strided sample of the source points (deterministic).
      std::vector<double> all(static_cast<size_t>(n_src_) * 3);
      // The raw (untransformed) points are not kept; sample the transformed ones
      // and evaluate with an identity-anisotropy copy of the RBF.
      PAM_HIP(hipMemcpyAsync(all.data(), src_pts_.get(), all.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
      std::vector<double> wts(static_cast<size_t>(n_src_));
      PAM_HIP(hipMemcpyAsync(wts.data(), w_.get(), wts.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
      PAM_HIP(hipStreamSynchronize(stream_));
      std::vector<double> smp(static_cast<size_t>(m) * 3);
      for (int64_t i = 0; i < m; i++) {
        int64_t j = (i * n_src_) / m;
        if (grad_kernel_) {
          // Gradients are measured between the sources (at a source the
          // exponential's gradient is undefined): midpoints of two samples.
          int64_t j2 = (j + n_src_ / 2 + 1) % n_src_;
          for (int k = 0; k < 3; k++) smp[3 * i + k] = 0.5 * (all[3 * j + k] + all[3 * j2 + k]);
        } else {
          for (int k = 0; k < 3; k++) smp[3 * i + k] = all[3 * j + k];
        }
      }
      RbfDesc iso = rbf_;
      for (int k = 0; k < 9; k++) iso.aniso[k] = (k % 4 == 0) ? 1.0 : 0.0;
      ensure_box();
      double bmin[3], bmax[3];
      for (int k = 0; k < 3; k++) {
        bmin[k] = box_.lo[k];
        bmax[k] = box_.lo[k] + box_.width / 1.01;
      }
      FmmEvaluator probe(iso, bmin, bmax, false);
      probe.set_source_points(all.data(), n_src_, false);
      probe.set_weights(wts.data(), n_src_, false);
      probe.set_target_points(smp.data(), m, false);
      // A gradient-transpose evaluator measures the error of its gradients.
      probe.set_gradient_kernel(grad_kernel_);
      const int64_t per = grad_kernel_ ? 3 : 1;
      std::vector<double> exact(m * per), approx(m * per);
      FmmConfig dc;
      dc.mode = kModeDirect;
      probe.set_config(dc);
      probe.evaluate(exact.data(), m, false);
      std::vector<std::array<int, 2>> cand;
      if (search_precision_ < 0) {
        cand = {{kMixed, 4}, {kMixed, 5}, {kMixed, 6}, {kMixed, 7}, {kMixed, 8}, {kMixedComp, 8},
                {kFp64, 8}, {kFp64, 10}, {kFp64, 12}};
      } else {
        for (int o = 4; o <= 12; o++) cand.push_back({search_precision_, o});
      }
      bool found = false;
      for (auto& cd : cand) {
        FmmConfig t;
        t.precision = cd[0];
        t.order = cd[1];
        t.leaf_level = -1;
        t.mode = kModeFmm;
        probe.set_config(t);
        probe.evaluate(approx.data(), m, false);
        double err = 0.0;
        for (int64_t i = 0; i < m * per; i++) {
          const double d = std::abs(approx[i] - exact[i]);
          if (!(d <= err)) err = d;  // a NaN difference fails the candidate
        }
        if (err <= accuracy_) {
          c.precision = cd[0];
          c.order = cd[1];
          found = true;
          break;
        }
      }
      if (!found) throw_runtime("failed to construct an evaluator that meets the desired accuracy");
    }
  }
  FmmConfig old = cfg_;
  cfg_ = c;
  if (old.order != c.order) invalidate_expansions();
  config_valid_ = true;
}

void FmmEvaluator::set_gradient_kernel(bool on) {
  if (on && symmetric_) throw_runtime("gradients are evaluated at target points: use a source-target evaluator");
  if (on == grad_kernel_) return;
  grad_kernel_ = on;
  if (!fixed_config_) config_valid_ = false;
}

void FmmEvaluator::evaluate(double* out, int64_t n, bool on_device) {
  if (grad_kernel_) {
    evaluate_gradient(out, n, on_device);
    return;
  }
  grad_ = false;
  evaluate_impl(out, n, on_device);
}

void FmmEvaluator::evaluate_deferred(double* d_out, int64_t n, DeferredFmm* job) {
  if (!job) throw_invalid("null job");
  if (grad_kernel_ || symmetric_) throw_runtime("deferred evaluation is for source-target value evaluators");
  *job = DeferredFmm{};
  grad_ = false;
  defer_ = job;
  try {
    evaluate_impl(d_out, n, true);
  } catch (...) {
    defer_ = nullptr;
    throw;
  }
  defer_ = nullptr;
}

void FmmEvaluator::evaluate_gradient(double* out, int64_t n, bool on_device) {
  if (symmetric_) throw_runtime("gradients are evaluated at target points: use a source-target evaluator");
  if (!config_valid_) choose_config();
  grad_ = true;
  try {
    evaluate_impl(out, n, on_device);
  } catch (...) {
    grad_ = false;
    throw;
  }
  grad_ = false;
}

void FmmEvaluator::evaluate_hessian_column(int k, double* out, int64_t n, bool on_device) {
  if (symmetric_) throw_runtime("Hessian columns are evaluated at target points: use a source-target evaluator");
  if (k < 0 || k > 2) throw_invalid("Hessian column out of range");
  if (!config_valid_) choose_config();
  grad_ = false;
  hess_col_ = k;
  try {
    evaluate_impl(out, n, on_device);
  } catch (...) {
    hess_col_ = -1;
    throw;
  }
  hess_col_ = -1;
}

void FmmEvaluator::evaluate_impl(double* out, int64_t n, bool on_device) {
  int64_t nt = n_trg();
  if (n != nt) throw_invalid("output size must equal the number of target points");
  const FmmEvaluator& src = S();
  if (!src.have_weights_ && src.n_src_ > 0) throw_runtime("weights are not set");
  if (!config_valid_) choose_config();
  const int per = (grad_ || hess_col_ >= 0) ? 3 : 1;
  double* d_out = out;
  DevBuf<double> tmp;
  if (!on_device) {
    tmp.resize(static_cast<size_t>(nt) * per);
    d_out = tmp.get();
  }
  timing_valid_ = true;
  record(0);
  stats_ = FmmStats{};
  if (nt > 0) {
    if (src.n_src_ == 0) {
      PAM_HIP(hipMemsetAsync(d_out, 0, sizeof(double) * nt * per, stream_));
      for (int i = 1; i <= 8; i++) record(i);
      stats_.direct = 1;
    } else if (use_direct()) {
      for (int i = 1; i <= 5; i++) record(i);
      if (hess_col_ >= 0) {
        direct_hessian_column_f64(rbf_, src.src_pts_.get(), src.w_.get(), src.n_src_, d_target_points(), nt,
                                  hess_col_, d_out, direct_scratch_, stream_);
      } else if (grad_) {
        direct_gradient_f64(rbf_, src.src_pts_.get(), src.w_.get(), src.n_src_, d_target_points(), nt, d_out,
                            direct_scratch_, stream_);
      } else {
        direct_sum_f64(rbf_, src.src_pts_.get(), src.w_.get(), src.n_src_, d_target_points(), nt, d_out, false,
                       direct_scratch_, stream_);
      }
      record(6);
      record(7);
      record(8);
      stats_.direct = 1;
    } else if (cfg_.precision == kFp64) {
      run_fmm<double>(d_out);
    } else {
      run_fmm<float>(d_out);
    }
  } else {
    for (int i = 1; i <= 8; i++) record(i);
  }
  stats_.precision = cfg_.precision;
  stats_.order = cfg_.order;
  stats_.leaf_level = stats_.direct ? -1 : src.src_tree_.level;
  stats_.n_src_leaves = stats_.direct ? 0 : src.src_tree_.n_leaves();
  stats_.n_trg_leaves = stats_.direct ? 0 : target_tree().n_leaves();
  stats_pending_ = true;
  if (!on_device) {
    tmp.download(out, static_cast<size_t>(nt) * per, stream_);
    PAM_HIP(hipStreamSynchronize(stream_));
  }
}

const FmmStats& FmmEvaluator::stats() {
  if (stats_pending_ && !timing_valid_) stats_pending_ = false;  // last evaluation was captured
  if (stats_pending_) {
    // Events: 0 start, 1 trees, 2 upward, 3 M2L, 4 L2L, 5 near field joined,
    // 6 L2P (end); 7 / 8 around the near field on its own stream.
    PAM_HIP(hipEventSynchronize(ev_[6]));
    PAM_HIP(hipEventSynchronize(ev_[8]));
    auto el = [&](int a, int b) {
      float ms = 0.0f;
      PAM_HIP(hipEventElapsedTime(&ms, ev_[a], ev_[b]));
      return static_cast<double>(ms);
    };
    stats_.ms_tree = el(0, 1);
    stats_.ms_upward = el(1, 2);
    stats_.ms_m2l = el(2, 3);
    stats_.ms_l2l = el(3, 4);
    stats_.ms_l2p = el(5, 6);
    stats_.ms_p2p = el(7, 8);
    stats_.ms_total = el(0, 6);
    stats_pending_ = false;
  }
  return stats_;
}

void FmmEvaluator::count_work(int64_t* p2p_pairs, int64_t* m2l_pairs) {
  *p2p_pairs = 0;
  *m2l_pairs = 0;
  if (!config_valid_) choose_config();
  ensure_trees();
  PAM_HIP(hipStreamSynchronize(stream_));
  const Tree& tt = target_tree();
  const Tree& stree = S().src_tree_;
  int L = stree.level;
  int64_t nt = tt.n_leaves(), ns = stree.n_leaves();
  std::vector<int32_t> tls(nt + 1), sls(ns + 1), p2p(static_cast<size_t>(nt) * 27);
  PAM_HIP(hipMemcpy(tls.data(), tt.leaf_start.get(), tls.size() * 4, hipMemcpyDeviceToHost));
  PAM_HIP(hipMemcpy(sls.data(), stree.leaf_start.get(), sls.size() * 4, hipMemcpyDeviceToHost));
  PAM_HIP(hipMemcpy(p2p.data(), lists_.p2p.get(), p2p.size() * 4, hipMemcpyDeviceToHost));
  for (int64_t t = 0; t < nt; t++) {
    int64_t ssum = 0;
    for (int k = 0; k < 27; k++) {
      int32_t s = p2p[t * 27 + k];
      if (s >= 0) ssum += sls[s + 1] - sls[s];
    }
    *p2p_pairs += ssum * (tls[t + 1] - tls[t]);
  }
  for (int l = 2; l <= L; l++) {
    const auto& m = lists_.m2l[l];
    std::vector<int32_t> src(static_cast<size_t>(tt.levels[l].n) * kM2LSlots);
    if (src.empty()) continue;
    PAM_HIP(hipMemcpy(src.data(), m.src.get(), src.size() * 4, hipMemcpyDeviceToHost));
    for (int32_t s : src) *m2l_pairs += (s >= 0);
  }
}

}  // namespace pam
