// Fit path: the saddle-point operator, FP64 residual evaluation and the
// FGMRES+RAS solver with mixed-precision iterative refinement.
#pragma once

#include <future>
#include <memory>
#include <vector>

#include "interpolation.hpp"
#include "krylov.hpp"
#include "ras.hpp"

namespace pam {

// interpolation/operator.hpp:56-82: [[A + nugget I, P], [P^T, 0]] x.
class Operator : public LinearOperator {
 public:
  Operator(const Model& model, const double* d_pts, int64_t n, const FmmConfig& cfg, hipStream_t s);
  void apply(const double* d_x, double* d_y) override;
  int64_t size() const override { return mu_ + l_; }
  // Builds the M2L operators of the FMM evaluators on stream s (see
  // FmmEvaluator::prebuild_operator).
  void prebuild(hipStream_t s);

 private:
  Model model_;
  hipStream_t s_;
  int64_t mu_;
  int l_;
  const double* d_pts_;
  std::vector<std::unique_ptr<FmmEvaluator>> fmm_;
  DevBuf<double> tmp_;
  Workspace1 ws_;
};

struct Convergence {
  bool converged = false;
  double residual = 0.0;
  bool exact = false;
};

// interpolation/residual_evaluator.hpp:35-151: L-inf residual at the data
// sites, first on <= 1024 sampled sites with nonzero values (FP64 direct),
// then at all sites in FP64.
class ResidualEvaluator {
 public:
  ResidualEvaluator(const Model& model, const double* d_pts, int64_t n, const FmmConfig& full_cfg,
                    hipStream_t s);
  void set_values(const std::vector<double>& values);
  // allow_full = false: a sample within the tolerance is reported as not
  // converged (exact = false) instead of being confirmed at every site.
  Convergence converged(const double* d_weights, double tolerance, bool allow_full = true);
  // r = values - (A + nugget) w - P c at all sites (FP64), device output.
  void full_residual(const double* d_weights, double* d_r);
  // Builds the FP64 operator's M2L matrices on stream s.
  void prebuild(hipStream_t s) { full_.prebuild(s); }

 private:
  Model model_;
  hipStream_t s_;
  int64_t mu_;
  int l_;
  const double* d_pts_;
  Operator full_;
  std::vector<int32_t> sample_;
  DevBuf<int32_t> d_sample_;
  DevBuf<double> sample_pts_, sample_vals_, values_, tmp_, y_;
  std::vector<std::unique_ptr<FmmEvaluator>> direct_;
  Workspace1 ws_;
};

struct FitOptions {
  FmmConfig op_cfg;       // inner operator (mixed-precision FMM by default)
  FmmConfig res_cfg;      // FP64 residual recomputation
  FmmConfig level_cfg;    // RAS inter-level updates
  bool use_level_cfg = false;
  bool use_preconditioner = true;
  // Inner relative residual that triggers a refinement restart on the FP64
  // residual. The mixed operator is accurate to about 1e-6 relative, so the
  // inner solve is trusted down to 1e-5; a restart earlier throws away the
  // Krylov space (measured: 9 iterations and one restart at 1e-4 against 8
  // and none at 1e-5 for the 1M-point ellipsoid fit). Stalls below that are
  // caught by the stagnation check.
  double refine_rtol = 1e-5;
  int verbose = 0;
  // Capture "preconditioner + matvec" of an Arnoldi step in a HIP graph after
  // the first (eager) step and replay it for the following ones.
  bool use_graphs = true;
  // At max_iter keep the last iterate (FitStats::converged = false) instead of
  // throwing, as polatory does (solver.hpp:131-133).
  bool keep_unconverged = false;
  // Time the RAS solve workgroup sizes per level at set-up (RasOptions::tune_solves).
  bool tune_solves = false;
  // RAS fine-level factors in FP32 (RasOptions::fp32_factors): 1 on, 0 off,
  // -1 on when the operator is not FP64 and every rbf has CPD order <= 1.
  int ras_fp32_factors = -1;
  // Time the RAS factorisation's forms (RasOptions::tune_factor), or fix one
  // (RasOptions::factor_panel).
  bool tune_factor = false;
  int factor_panel = -1;
  // Kernels of CPD order 2 (triharmonic3d): raise a mixed-precision operator
  // to FP64 of order >= 10 and grow the RAS coarse grid with the point count.
  bool cpd2_auto = true;
};

struct FitStats {
  int iterations = 0;
  int restarts = 0;
  int graph_replays = 0;
  int escalations = 0;  // precision raises after stagnation
  bool converged = true;
  double residual = 0.0;
  double grad_residual = 0.0;   // fits with gradient data: L-inf residual of the gradient rows
  double seconds = 0.0;
  double setup_seconds = 0.0;
  double precond_setup_seconds = 0.0;
  int n_levels = 0;
  int64_t factor_bytes = 0;
  std::vector<double> history;  // residual per iteration (sampled or exact)
  int rounds = 0;               // incremental fit: rounds of solve + residual evaluation
};

class GraphStep;
class TimedOperator;

// interpolation/solver.hpp:34-142 + fitter.hpp:21-26.
// Applies the rules for kernels of CPD order 2 (FP64 operator with
// cpd2_auto, FP64 level updates unless configured) to a fit's options.
void resolve_fit_options(const Model& model, FitOptions& opt);
// The preconditioner options of a fit of n points with resolved options.
RasOptions fit_ras_options(const Model& model, const FitOptions& opt, int64_t n);

class Solver {
 public:
  Solver(const Model& model, const std::vector<double>& points, const FitOptions& opt, hipStream_t s);
  std::vector<double> solve(const std::vector<double>& values, double tolerance, int max_iter,
                            const std::vector<double>* initial, FitStats* stats);
  RasPreconditioner* preconditioner() { return pc_.get(); }

 private:
  bool can_escalate() const;
  void escalate(std::unique_ptr<GraphStep>& gstep, std::unique_ptr<TimedOperator>& t_op,
                std::unique_ptr<TimedOperator>& t_pc, LinearOperator*& op, LinearOperator*& pc,
                FitStats& st);
  Model model_;
  FitOptions opt_;
  hipStream_t s_;
  int64_t mu_;
  int l_;
  DevBuf<double> d_pts_, p_orth_;
  std::unique_ptr<Operator> op_;
  std::unique_ptr<ResidualEvaluator> res_;
  std::unique_ptr<RasPreconditioner> pc_;
  Workspace1 ws_;
  double setup_s_ = 0.0;
  // Build of the residual evaluator's M2L operator on a host thread and
  // stream of its own, overlapping the preconditioner set-up; joined before
  // the first residual check (and, being the last member, before the
  // evaluator is destroyed).
  std::future<void> prebuild_;
};

// Fits with gradient data (interpolant.hpp:77-113, residual_evaluator.hpp:35-151
// with gradient points): vectors [w (mu); g (3 sigma); c (l)] and data
// [values (mu); gradients (3 sigma)].
class HermiteResidualEvaluator {
 public:
  HermiteResidualEvaluator(const Model& model, const double* d_pts, int64_t mu, const double* d_gpts,
                           int64_t sigma, const FmmConfig& full_cfg, hipStream_t s);
  void set_values(const std::vector<double>& values);
  struct Result {
    bool converged = false;
    double residual = 0.0, grad_residual = 0.0;
    bool exact = false;
  };
  // First <= 1024 sampled sites with nonzero values and <= 1024 gradient
  // sites with nonzero gradients by the FP64 direct sum, then every row with
  // the FP64 operator (residual_evaluator.hpp:56-108).
  Result converged(const double* d_weights, double tolerance, double grad_tolerance);
  // r = data - (operator w)[0 : mu + 3 sigma] in FP64.
  void full_residual(const double* d_weights, double* d_r);

 private:
  Model model_;
  hipStream_t s_;
  int64_t mu_, sigma_;
  const double *d_pts_, *d_gpts_;
  HermiteOperator full_;
  std::unique_ptr<HermiteEvaluator> direct_;
  std::vector<int32_t> sample_, gsample_;
  DevBuf<int32_t> d_sample_, d_gsample_;
  DevBuf<double> sample_pts_, sample_gpts_, sample_vals_, values_, tmp_, y_;
  Workspace1 ws_;
};

class HermiteSolver {
 public:
  HermiteSolver(const Model& model, const std::vector<double>& points, const std::vector<double>& grad_points,
                const FitOptions& opt, hipStream_t s);
  std::vector<double> solve(const std::vector<double>& values, double tolerance, double grad_tolerance,
                            int max_iter, const std::vector<double>* initial, FitStats* stats);

 private:
  Model model_;
  FitOptions opt_;
  hipStream_t s_;
  int64_t mu_, sigma_, n_;
  int l_;
  DevBuf<double> d_pts_, d_gpts_, p_orth_;
  std::unique_ptr<HermiteOperator> op_;
  std::unique_ptr<HermiteResidualEvaluator> res_;
  std::unique_ptr<RasPreconditioner> pc_;
  Workspace1 ws_;
  double setup_s_ = 0.0;
};

}  // namespace pam
