// Multilevel restricted additive Schwarz preconditioner on the device.
// Restates preconditioner::RasPreconditioner (ras_preconditioner.hpp:56-328)
// with DomainDivider (domain_divider.hpp:30-269), FineGrid (fine_grid.hpp:59-170)
// and CoarseGrid (coarse_grid.hpp:40-131). Domain division runs on the host;
// the kernel matrices Q^T A Q of all subdomains are assembled, Cholesky-factored
// (FP64) and solved in batched device kernels, and the factors stay in HBM
// (polatory keeps them in a temporary file, binary_cache.hpp).
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "hermite.hpp"
#include "interpolation.hpp"
#include "krylov.hpp"

namespace pam {

struct Domain {
  std::vector<int32_t> idx;     // point indices (polynomial points first after merging)
  std::vector<uint8_t> inner;   // 1 where this domain owns the point's solution
  std::vector<int32_t> gidx;    // gradient point indices (fits with gradient data)
  std::vector<uint8_t> ginner;  // 1 where this domain owns the gradient point's solution
};

// DomainDivider (domain_divider.hpp:30-269), point multiplicity 1 (no gradient data).
class DomainDivider {
 public:
  DomainDivider(const std::vector<double>& pts, const std::vector<int32_t>& point_idcs,
                const std::vector<int32_t>& poly_point_idcs);
  std::vector<int32_t> choose_coarse_points(int64_t n_coarse) const;
  std::vector<Domain>& domains() { return domains_; }

 private:
  const std::vector<double>& pts_;
  std::vector<int32_t> point_idcs_, poly_idcs_;
  std::vector<Domain> domains_;
};

// DomainDivider::choose_coarse_points (domain_divider.hpp:54-125) as a free
// function of the level's points, so that it can run beside the division.
std::vector<int32_t> choose_coarse_points(const std::vector<double>& pts, const std::vector<int32_t>& point_idcs,
                                          const std::vector<int32_t>& poly_idcs, int64_t n_coarse);

// With gradient points (domain_divider.hpp:30-48, 160-269): every gradient
// point counts 3 (its three rows) in the sizes that drive the splits, and the
// points of a domain are sorted along its widest axes together.
std::vector<Domain> divide_mixed(const std::vector<double>& pts, const std::vector<double>& gpts,
                                 const std::vector<int32_t>& point_idcs, const std::vector<int32_t>& grad_idcs,
                                 const std::vector<int32_t>& poly_idcs);
// choose_coarse_points with gradient points (domain_divider.hpp:54-125):
// returns the coarse level's point and gradient point indices.
std::pair<std::vector<int32_t>, std::vector<int32_t>> choose_coarse_points_mixed(
    const std::vector<double>& pts, const std::vector<double>& gpts, const std::vector<int32_t>& point_idcs,
    const std::vector<int32_t>& grad_idcs, const std::vector<int32_t>& poly_idcs, int64_t n_coarse);

// Batched subdomain factors of one level (or the coarse grid as one domain).
// Rows are flat indices into the fit's vectors: a point's row is its index,
// row mu + 3 k + c is component c of gradient point k.
struct DomainBatch {
  int n_dom = 0;
  int max_m = 0;
  std::vector<int64_t> h_ptr;            // [n_dom+1] offsets into idx
  DevBuf<int32_t> idx;
  DevBuf<uint8_t> inner;
  DevBuf<int64_t> ptr, mat_ptr;          // [n_dom+1]
  DevBuf<double> mat;                    // lower Cholesky factors, column-major (FP64)
  DevBuf<float> mat32;                   // the same in FP32 (single-precision fine levels)
  bool fp32 = false;                     // the factors are mat32
  DevBuf<double> op;                     // coarse grid: its solve as a dense (m + l) x m operator
  bool has_op = false;
  int solve_threads = 0;                 // workgroup size of the batched solve (0: default)
};

struct RasOptions {
  FmmConfig level_cfg;      // inter-level residual updates (default accuracy)
  bool use_level_cfg = false;
  int verbose = 0;          // print the set-up phase times
  // Time each level's batched solve at 128, 256 and 512 threads per domain and
  // keep the fastest (off: the rule below, which reproduces what the timing
  // picked on the 1M-point fit without its cost).
  bool tune_solves = false;
  // Fine-level factors in FP32 (assembled in FP64, factored on FP32 MFMA,
  // solved with FP64 arithmetic on the FP32 factors); the coarse grid stays
  // FP64. A level whose FP32 factorisation meets a non-positive pivot is
  // factored again in FP64.
  bool fp32_factors = false;
  // Time the batched factorisation's forms (left-looking; right-looking with
  // 32- or 64-wide panels) on a sample of each fine level's domains and keep
  // the fastest (cached per precision and domain size for the process); off:
  // a fixed rule from such a timing.
  bool tune_factor = false;
  // Fixes the form instead: 0 left-looking, 32 or 64 right-looking; -1 none.
  int factor_panel = -1;
  // Points of the coarse grid (ras_preconditioner.hpp:53 uses 2048); at most
  // kMaxCoarsest.
  int64_t n_coarsest = 2048;
  static constexpr int64_t kMaxCoarsest = 6000;
};

class RasPreconditioner : public LinearOperator {
 public:
  // d_points: the same points already on the device (optional; copied on the
  // device instead of uploaded again).
  RasPreconditioner(const Model& model, const std::vector<double>& points, hipStream_t s,
                    const RasOptions& opt, const double* d_points = nullptr);
  // With gradient points (ras_preconditioner.hpp:56-189): vectors have
  // mu + 3 sigma + l entries ([values; gradients x y z per point; polynomial]).
  RasPreconditioner(const Model& model, const std::vector<double>& points, const std::vector<double>& grad_points,
                    hipStream_t s, const RasOptions& opt, const double* d_points = nullptr);
  void apply(const double* d_v, double* d_out) override;
  int64_t size() const override { return n_ + l_; }
  int64_t n_grad_points() const { return sigma_; }
  const std::vector<std::vector<int32_t>>& level_grad_points() const { return gpoint_idcs_; }
  int n_levels() const { return n_levels_; }
  const std::vector<std::vector<int32_t>>& level_points() const { return point_idcs_; }
  const std::vector<Domain>& level_domains(int level) const { return h_domains_.at(level); }
  double setup_seconds() const { return setup_s_; }
  // Changes the FMM configuration of the inter-level residual updates (and of
  // A P); the update evaluators are rebuilt on the next apply.
  void set_level_config(const FmmConfig& cfg);
  const RasOptions& options() const { return opt_; }
  int64_t factor_bytes() const;

 private:
  // Assembles and factors a batch asynchronously; its temporaries stay in
  // setup_tmp_ until the end of the constructor.
  void setup_batch(DomainBatch& b, const std::vector<Domain>& doms, bool coarse, hipStream_t ls, bool fp32,
                   int32_t* flag);
  // Set-up streams: each level's batch (assembly, factorisation) runs on its
  // own stream, forked from and joined back into s_.
  hipStream_t fork_stream();
  void join_streams();
  std::vector<hipStream_t> setup_streams_;
  struct SetupTmp {
    std::vector<int32_t> idx;
    std::vector<uint8_t> inner;
    std::vector<int64_t> mptr;
    DevBuf<double> u, atl, rows, inv;
  };
  std::vector<std::unique_ptr<SetupTmp>> setup_tmp_;
  void solve(int level, const double* d_res, double* d_w);
  void launch_solve(const DomainBatch& b, int level, int threads, const double* d_res, double* d_w);
  // Form of a fine level's batched factorisation (0 left-looking, else the
  // right-looking panel width; RasOptions::factor_panel, tune_factor).
  int factor_panel(const DomainBatch& b, bool fp32, hipStream_t ls, const std::vector<int64_t>& mptr);
  // Picks each level's solve workgroup size by timing (an autotuner over
  // launch shapes; the factorisation tiles stay fixed).
  void tune_solves();
  void update_residuals(int src, int trg, const double* d_w, double* d_res);
  void orthogonalize(double* d_w, double* d_res);
  // ap = A P with the level-update configuration, on stream s (synchronized).
  void compute_ap(hipStream_t s);
  Evaluator& evaluator(int src, int trg);
  HermiteEvaluator& hermite_evaluator(int src, int trg);
  // Level updates and A P of fits with gradient data: the level configuration
  // when set, else FP64 of order 10.
  FmmConfig hermite_level_cfg() const;
  void setup(const std::vector<double>& points, const std::vector<double>& grad_points);
  // Device copies of a level's points and gradient points.
  const double* level_points(int level);
  const double* level_grad_points(int level);

  Model model_;
  hipStream_t s_;
  RasOptions opt_;
  int64_t mu_;
  int64_t sigma_ = 0;  // gradient points
  int64_t n_ = 0;      // rows: mu + 3 sigma
  int l_;
  int n_levels_ = 1;
  std::vector<double> h_pts_, h_gpts_;
  DevBuf<double> d_pts_, d_gpts_, lag_, p_orth_, ap_, pinv_;
  std::vector<std::vector<int32_t>> point_idcs_, gpoint_idcs_;
  std::vector<DevBuf<int32_t>> d_point_idcs_, d_gpoint_idcs_;
  std::vector<DevBuf<int32_t>> d_level_rows_;  // per level: its rows (flat indices)
  std::vector<int64_t> level_rows_;
  std::vector<std::vector<Domain>> h_domains_;
  std::vector<DomainBatch> fine_;  // index = level (1..n-1)
  DomainBatch coarse_;
  std::map<std::pair<int, int>, std::unique_ptr<Evaluator>> evaluators_;
  std::map<std::pair<int, int>, std::unique_ptr<HermiteEvaluator>> hevaluators_;
  std::map<int, DevBuf<double>> level_pts_, level_gpts_;
  DevBuf<double> res_, w_, total_, tmp_, sw_, dot_;
  DevBuf<int32_t> flag_;
  Workspace1 ws_;
  double setup_s_ = 0.0;
  int verbose_applies_ = 0;  // applies whose phases have been timed and printed
};

}  // namespace pam
