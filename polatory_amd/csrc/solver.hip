// Fit operator, residual evaluator and the FGMRES solver (see solver.hpp).
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>

#include "device_api.hpp"

namespace pam {

Operator::Operator(const Model& model, const double* d_pts, int64_t n, const FmmConfig& cfg, hipStream_t s)
    : model_(model), s_(s), mu_(n), l_(model.l()), d_pts_(d_pts) {
  model_.validate();
  for (const auto& r : model_.rbfs) {
    fmm_.push_back(std::make_unique<FmmEvaluator>(r, nullptr, nullptr, true, s));
    fmm_.back()->set_config(cfg);
    fmm_.back()->set_source_points(d_pts, n, true);
  }
}

void Operator::prebuild(hipStream_t s) {
  for (auto& f : fmm_) f->prebuild_operator(s);
}

void Operator::apply(const double* d_x, double* d_y) {
  // y[0:mu] = nugget w + sum_rbf A w + P c; y[mu:] = P^T w   (operator.hpp:56-82)
  tmp_.resize(static_cast<size_t>(mu_));
  for (size_t i = 0; i < fmm_.size(); i++) {
    fmm_[i]->set_weights(d_x, mu_, true);
    fmm_[i]->evaluate(i == 0 ? d_y : tmp_.get(), mu_, true);
    if (i > 0) dev_axpy(1.0, tmp_.get(), d_y, mu_, s_);
  }
  if (model_.nugget != 0.0) dev_axpy(model_.nugget, d_x, d_y, mu_, s_);
  if (l_ > 0) {
    launch_poly_eval(model_.poly_degree, d_pts_, mu_, d_x + mu_, d_y, s_);
    dev_poly_transpose_apply(model_.poly_degree, d_pts_, mu_, d_x, d_y + mu_, ws_, s_);
  }
}

ResidualEvaluator::ResidualEvaluator(const Model& model, const double* d_pts, int64_t n,
                                     const FmmConfig& full_cfg, hipStream_t s)
    : model_(model), s_(s), mu_(n), l_(model.l()), d_pts_(d_pts), full_(model, d_pts, n, full_cfg, s) {}

void ResidualEvaluator::set_values(const std::vector<double>& values) {
  // residual_evaluator.hpp:121-151: up to 1024 sites with nonzero values.
  // polatory shuffles with a default-seeded mt19937; this uses an evenly
  // strided pick among the nonzero sites (deterministic, different sites).
  constexpr int64_t kDirect = 1024;
  values_.upload(values.data(), values.size(), s_);
  std::vector<int32_t> nz, zero;
  for (int64_t i = 0; i < mu_; i++) (values[i] != 0.0 ? nz : zero).push_back(static_cast<int32_t>(i));
  std::vector<int32_t> order = nz;
  order.insert(order.end(), zero.begin(), zero.end());
  int64_t k = std::min<int64_t>(mu_, kDirect);
  sample_.clear();
  int64_t nnz = static_cast<int64_t>(nz.size());
  if (nnz >= k) {
    for (int64_t i = 0; i < k; i++) sample_.push_back(nz[(i * nnz) / k]);
  } else {
    sample_ = nz;
    for (int64_t i = 0; static_cast<int64_t>(sample_.size()) < k; i++) sample_.push_back(zero[i]);
  }
  std::sort(sample_.begin(), sample_.end());
  d_sample_.upload(sample_.data(), sample_.size(), s_);
  int64_t ns = static_cast<int64_t>(sample_.size());
  sample_pts_.resize(static_cast<size_t>(ns) * 3);
  dev_gather_points(d_pts_, d_sample_.get(), ns, sample_pts_.get(), s_);
  sample_vals_.resize(static_cast<size_t>(ns));
  dev_gather(values_.get(), d_sample_.get(), ns, sample_vals_.get(), s_);
  direct_.clear();
  for (const auto& r : model_.rbfs) {
    direct_.push_back(std::make_unique<FmmEvaluator>(r, nullptr, nullptr, false, s_));
    FmmConfig c;
    c.mode = kModeDirect;
    direct_.back()->set_config(c);
    direct_.back()->set_source_points(d_pts_, mu_, true);
    direct_.back()->set_target_points(sample_pts_.get(), ns, true);
  }
  PAM_HIP(hipStreamSynchronize(s_));
}

void ResidualEvaluator::full_residual(const double* d_weights, double* d_r) {
  y_.resize(static_cast<size_t>(mu_ + l_));
  full_.apply(d_weights, y_.get());
  dev_copy(values_.get(), d_r, mu_, s_);
  dev_axpy(-1.0, y_.get(), d_r, mu_, s_);
}

Convergence ResidualEvaluator::converged(const double* d_weights, double tolerance, bool allow_full) {
  // residual_evaluator.hpp:56-108
  Convergence c;
  int64_t ns = static_cast<int64_t>(sample_.size());
  bool exact = ns == mu_;
  {
    y_.resize(static_cast<size_t>(std::max<int64_t>(ns, mu_ + l_)));
    tmp_.resize(static_cast<size_t>(ns));
    for (size_t i = 0; i < direct_.size(); i++) {
      direct_[i]->set_weights(d_weights, mu_, true);
      direct_[i]->evaluate(i == 0 ? y_.get() : tmp_.get(), ns, true);
      if (i > 0) dev_axpy(1.0, tmp_.get(), y_.get(), ns, s_);
    }
    if (l_ > 0) launch_poly_eval(model_.poly_degree, sample_pts_.get(), ns, d_weights + mu_, y_.get(), s_);
    if (model_.nugget != 0.0) {
      dev_gather(d_weights, d_sample_.get(), ns, tmp_.get(), s_);
      dev_axpy(model_.nugget, tmp_.get(), y_.get(), ns, s_);
    }
    double r = dev_max_abs_diff(y_.get(), sample_vals_.get(), ns, ws_, s_);
    c.residual = r;
    c.exact = exact;
    if (r > tolerance) return c;
    if (exact) {
      c.converged = true;
      return c;
    }
    if (!allow_full) return c;
  }
  tmp_.resize(static_cast<size_t>(mu_));
  full_residual(d_weights, tmp_.get());
  double r = dev_max_abs_diff(tmp_.get(), nullptr, mu_, ws_, s_);
  c.residual = r;
  c.exact = true;
  c.converged = r <= tolerance;
  return c;
}

void resolve_fit_options(const Model& model, FitOptions& opt) {
  if (model.cpd_order() < 2) return;
  if (opt.cpd2_auto && opt.op_cfg.precision != kFp64) {
    // Triharmonic systems are ill-conditioned: FGMRES on the mixed operator
    // stagnates and escalates to FP64 after wasted steps (measured from 3e4
    // to 1e6 points), so the fit starts with the FP64 operator.
    opt.op_cfg.precision = kFp64;
    opt.op_cfg.order = std::max(opt.op_cfg.order, 10);
  }
  if (!opt.use_level_cfg) {
    // Their multilevel updates need FP64 too: with the default mixed-precision
    // updates FGMRES stalls above 1e-3 relative (measured at 2e5 points), with
    // FP64 order-8 updates it converges.
    opt.level_cfg = FmmConfig{};
    opt.level_cfg.precision = kFp64;
    opt.level_cfg.order = 8;
    opt.use_level_cfg = true;
  }
}

RasOptions fit_ras_options(const Model& model, const FitOptions& opt, int64_t n) {
  RasOptions ro;
  ro.level_cfg = opt.level_cfg;
  ro.use_level_cfg = opt.use_level_cfg;
  ro.verbose = opt.verbose;
  ro.tune_solves = opt.tune_solves;
  ro.tune_factor = opt.tune_factor;
  ro.factor_panel = opt.factor_panel;
  // With the reference's 2048-point coarse grid, triharmonic fits stall above
  // about 5e5 points (at 1e-3 absolute at 1e6 after 100 steps); a coarse grid
  // of 4096 points from 2e5 and 6000 from 8e5 converges (1e6 points: 23 steps
  // to 1e-4).
  if (model.cpd_order() >= 2 && opt.cpd2_auto) ro.n_coarsest = n >= 800000 ? 6000 : n >= 200000 ? 4096 : 2048;
  // Triharmonic (CPD order 2) subdomain matrices are too ill-conditioned for
  // FP32 factors; an FP64 operator asks for the FP64 fit throughout.
  ro.fp32_factors = opt.ras_fp32_factors == 1 ||
                    (opt.ras_fp32_factors == -1 && model.cpd_order() <= 1 && opt.op_cfg.precision != kFp64);
  return ro;
}

Solver::Solver(const Model& model, const std::vector<double>& points, const FitOptions& opt, hipStream_t s)
    : model_(model), opt_(opt), s_(s), mu_(static_cast<int64_t>(points.size() / 3)), l_(model.l()) {
  auto t0 = std::chrono::steady_clock::now();
  model_.validate();
  if (mu_ < l_) throw_invalid("points.rows() must be greater than or equal to " + std::to_string(l_));
  d_pts_.upload(points.data(), points.size(), s_);
  resolve_fit_options(model_, opt_);
  op_ = std::make_unique<Operator>(model_, d_pts_.get(), mu_, opt_.op_cfg, s_);
  res_ = std::make_unique<ResidualEvaluator>(model_, d_pts_.get(), mu_, opt_.res_cfg, s_);
  {
    int dev = 0;
    PAM_HIP(hipGetDevice(&dev));
    prebuild_ = std::async(std::launch::async, [this, dev] {
      PAM_HIP(hipSetDevice(dev));
      hipStream_t ps = stream_acquire();
      try {
        res_->prebuild(ps);
        PAM_HIP(hipStreamSynchronize(ps));  // the operator is complete before solve() uses it
      } catch (...) {
        stream_release(ps);
        throw;
      }
      stream_release(ps);
    });
  }
  if (opt_.use_preconditioner)
    pc_ = std::make_unique<RasPreconditioner>(model_, points, s_, fit_ras_options(model_, opt_, mu_), d_pts_.get());
  if (l_ > 0) {
    p_orth_.resize(static_cast<size_t>(mu_) * l_);
    dev_poly_matrix(model_.poly_degree, d_pts_.get(), mu_, p_orth_.get(), s_);
    dev_orthonormalize_cols(p_orth_.get(), mu_, l_, ws_, s_);
  }
  PAM_HIP(hipStreamSynchronize(s_));
  setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Wall time of a linear operator's applications (verbose fits only: each
// application synchronizes the stream).
class TimedOperator : public LinearOperator {
 public:
  TimedOperator(LinearOperator& inner, hipStream_t s) : inner_(inner), s_(s) {}
  void apply(const double* d_x, double* d_y) override {
    PAM_HIP(hipStreamSynchronize(s_));
    auto t0 = std::chrono::steady_clock::now();
    inner_.apply(d_x, d_y);
    PAM_HIP(hipStreamSynchronize(s_));
    seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    calls++;
  }
  int64_t size() const override { return inner_.size(); }
  double seconds = 0.0;
  int calls = 0;

 private:
  LinearOperator& inner_;
  hipStream_t s_;
};

// "z = M v; w = A z" on fixed device buffers: the first call runs eagerly
// (it creates every lazily built evaluator and buffer), the second is
// captured into a HIP graph, and later calls replay the graph. Inputs and
// outputs are copied through the fixed buffers. If capture fails, the step
// stays eager (every evaluation re-uploads its weights first, so no state of
// the failed capture is reused).
class GraphStep {
 public:
  GraphStep(LinearOperator& op, LinearOperator* pc, int64_t m, hipStream_t s)
      : op_(op), pc_(pc), m_(m), s_(s), in_(m), z_(m), w_(m) {}
  ~GraphStep() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
  }
  void run(const double* vj, double* zj, double* w) {
    dev_copy(vj, in_.get(), m_, s_);
    if (exec_) {
      PAM_HIP(hipGraphLaunch(exec_, s_));
      replays++;
    } else if (calls_ == 0 || failed_) {
      eager();
    } else {
      capture();
    }
    calls_++;
    dev_copy(z_.get(), zj, m_, s_);
    dev_copy(w_.get(), w, m_, s_);
  }
  int replays = 0;

 private:
  void eager() {
    if (pc_) pc_->apply(in_.get(), z_.get());
    else dev_copy(in_.get(), z_.get(), m_, s_);
    op_.apply(z_.get(), w_.get());
  }
  void capture() {
    hipGraph_t g = nullptr;
    bool capturing = false;
    try {
      PAM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeRelaxed));
      capturing = true;
      {
        PoolCaptureScope no_sync;
        eager();
      }
      capturing = false;
      PAM_HIP(hipStreamEndCapture(s_, &g));
      PAM_HIP(hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      PAM_HIP(hipGraphLaunch(exec_, s_));
      replays++;
    } catch (const Error&) {
      if (capturing) {
        hipGraph_t bad = nullptr;
        (void)hipStreamEndCapture(s_, &bad);
        if (bad) (void)hipGraphDestroy(bad);
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
      if (exec_) {
        (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
      }
      (void)hipGetLastError();
      failed_ = true;
      eager();
    }
  }
  LinearOperator& op_;
  LinearOperator* pc_;
  int64_t m_;
  hipStream_t s_;
  DevBuf<double> in_, z_, w_;
  hipGraphExec_t exec_ = nullptr;
  int calls_ = 0;
  bool failed_ = false;
};

bool Solver::can_escalate() const {
  bool levels_fp64 = !pc_ || (pc_->options().use_level_cfg && pc_->options().level_cfg.precision == kFp64);
  return !(levels_fp64 && opt_.op_cfg.precision == kFp64 && opt_.op_cfg.order >= 12);
}

void Solver::escalate(std::unique_ptr<GraphStep>& gstep, std::unique_ptr<TimedOperator>& t_op,
                      std::unique_ptr<TimedOperator>& t_pc, LinearOperator*& op, LinearOperator*& pc,
                      FitStats& st) {
  PAM_HIP(hipStreamSynchronize(s_));
  const char* what = nullptr;
  if (pc_ && !(pc_->options().use_level_cfg && pc_->options().level_cfg.precision == kFp64)) {
    FmmConfig c;
    c.precision = kFp64;
    c.order = 8;
    pc_->set_level_config(c);
    what = "FP64 multilevel updates";
  } else if (opt_.op_cfg.precision != kFp64 || opt_.op_cfg.order < 10) {
    opt_.op_cfg.precision = kFp64;
    opt_.op_cfg.order = std::max(opt_.op_cfg.order, 10);
    op_ = std::make_unique<Operator>(model_, d_pts_.get(), mu_, opt_.op_cfg, s_);
    what = "FP64 operator of order 10";
  } else if (opt_.op_cfg.order < 12) {
    opt_.op_cfg.order = 12;
    op_ = std::make_unique<Operator>(model_, d_pts_.get(), mu_, opt_.op_cfg, s_);
    what = "FP64 operator of order 12";
  } else {
    return;  // nothing left to raise
  }
  st.escalations++;
  if (opt_.verbose) std::fprintf(stderr, "  stagnation: switching to %s\n", what);
  op = op_.get();
  pc = pc_.get();
  if (t_op) {
    t_op = std::make_unique<TimedOperator>(*op_, s_);
    op = t_op.get();
  }
  if (t_pc && pc_) {
    t_pc = std::make_unique<TimedOperator>(*pc_, s_);
    pc = t_pc.get();
  }
  if (gstep) gstep = std::make_unique<GraphStep>(*op_, pc_.get(), mu_ + l_, s_);
}

std::vector<double> Solver::solve(const std::vector<double>& values, double tolerance, int max_iter,
                                  const std::vector<double>* initial, FitStats* stats) {
  // solver.hpp:76-142, with the operator in mixed precision and FGMRES
  // restarted on the FP64 residual when the inner solve reaches refine_rtol.
  auto t0 = std::chrono::steady_clock::now();
  if (static_cast<int64_t>(values.size()) != mu_) throw_invalid("values.rows() must be equal to " + std::to_string(mu_));
  if (!(tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  const int64_t m = mu_ + l_;
  DevBuf<double> x(m), xb(m), rhs(m), dot(10);
  dev_fill(0.0, x.get(), m, s_);
  if (initial) {
    if (static_cast<int64_t>(initial->size()) != m) throw_invalid("initial weights have the wrong size");
    PAM_HIP(hipMemcpyAsync(x.get(), initial->data(), sizeof(double) * m, hipMemcpyHostToDevice, s_));
    if (l_ > 0) {  // orthogonalize the warm start against P (solver.hpp:85-93)
      double h[10];
      dev_multi_dot(p_orth_.get(), mu_, l_, x.get(), mu_, h, ws_, s_);
      PAM_HIP(hipMemcpyAsync(dot.get(), h, sizeof(double) * l_, hipMemcpyHostToDevice, s_));
      dev_gemv_update(x.get(), p_orth_.get(), mu_, l_, dot.get(), mu_, -1.0, s_);
    }
    PAM_HIP(hipStreamSynchronize(s_));
  }
  if (prebuild_.valid()) prebuild_.get();
  res_->set_values(values);
  FitStats st;
  st.setup_seconds = setup_s_;
  if (pc_) {
    st.precond_setup_seconds = pc_->setup_seconds();
    st.n_levels = pc_->n_levels();
    st.factor_bytes = pc_->factor_bytes();
  }
  int total_iter = 0;
  DevBuf<double> r(mu_);
  bool first = true;
  std::unique_ptr<TimedOperator> t_op, t_pc;
  LinearOperator* op = op_.get();
  LinearOperator* pc = pc_.get();
  if (opt_.verbose) {
    t_op = std::make_unique<TimedOperator>(*op_, s_);
    op = t_op.get();
    if (pc_) {
      t_pc = std::make_unique<TimedOperator>(*pc_, s_);
      pc = t_pc.get();
    }
  }
  double t_check = 0.0;
  // Graph replay of the Arnoldi step (timed runs synchronize inside the step,
  // which capture does not allow, so verbose fits stay eager).
  std::unique_ptr<GraphStep> gstep;
  if (opt_.use_graphs && !opt_.verbose) gstep = std::make_unique<GraphStep>(*op_, pc_.get(), m, s_);
  while (true) {
    // right-hand side of this refinement cycle
    if (first && !initial) {
      PAM_HIP(hipMemcpyAsync(rhs.get(), values.data(), sizeof(double) * mu_, hipMemcpyHostToDevice, s_));
      dev_fill(0.0, rhs.get() + mu_, l_, s_);
      dev_fill(0.0, xb.get(), m, s_);
    } else {
      dev_copy(x.get(), xb.get(), m, s_);
      res_->full_residual(xb.get(), rhs.get());
      if (l_ > 0) {  // second block: 0 - P^T w
        dev_poly_transpose_apply(model_.poly_degree, d_pts_.get(), mu_, xb.get(), rhs.get() + mu_, ws_, s_);
        dev_scal(-1.0, rhs.get() + mu_, l_, s_);
      }
    }
    first = false;
    Fgmres solver(*op, rhs.get(), m, std::max(max_iter - total_iter, 0), s_);
    solver.set_right_preconditioner(pc);
    if (gstep) solver.set_step([&](const double* vj, double* zj, double* w) { gstep->run(vj, zj, w); });
    solver.setup();
    if (solver.relative_residual() == 0.0) {
      dev_copy(xb.get(), x.get(), m, s_);
      break;
    }
    bool done = false, restart = false, stagnated = false;
    std::vector<double> inner_rel;
    // A check at every site costs an FP64 evaluation of the whole operator.
    // After one has failed in this cycle, the next is made when the inner
    // residual has fallen enough for the full residual, scaled by the same
    // factor, to be within twice the tolerance; every third step at the
    // latest; and at the last step allowed.
    double full_res = 0.0, full_inner = 0.0;
    int since_full = 0;
    while (true) {
      solver.solution_vector(x.get());
      dev_axpy(1.0, xb.get(), x.get(), m, s_);
      auto tc = std::chrono::steady_clock::now();
      bool allow_full = true;
      if (full_res > 0.0 && full_inner > 0.0) {
        const double predicted = full_res * solver.relative_residual() / full_inner;
        allow_full = predicted <= 2.0 * tolerance || since_full >= 2 ||
                     total_iter + solver.iteration_count() >= max_iter;
      }
      Convergence c = res_->converged(x.get(), tolerance, allow_full);
      if (c.exact && !c.converged) {
        full_res = c.residual;
        full_inner = solver.relative_residual();
        since_full = 0;
      } else {
        since_full++;
      }
      if (opt_.verbose) t_check += std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
      st.history.push_back(c.residual);
      st.residual = c.residual;
      if (opt_.verbose)
        std::fprintf(stderr, "%8d %s%12.6e  (inner rel %.3e)\n", total_iter + solver.iteration_count(),
                     c.exact ? " " : "~", c.residual, solver.relative_residual());
      if (c.converged) {
        done = true;
        break;
      }
      if (total_iter + solver.iteration_count() >= max_iter) {
        if (!opt_.keep_unconverged) throw_runtime("reached the maximum number of iterations");
        st.converged = false;
        done = true;
        break;
      }
      const int it = solver.iteration_count();
      inner_rel.push_back(solver.relative_residual());
      if (it > 0 && solver.relative_residual() < opt_.refine_rtol) {
        restart = true;
        break;
      }
      // Stagnation: less than a factor 2 over the last kWindow steps. The
      // inner operator (or the preconditioner) is too inexact for this
      // system: restart on the FP64 residual and raise the precision.
      constexpr int kWindow = 4;
      if (can_escalate() && it >= kWindow && inner_rel[it] > 0.5 * inner_rel[it - kWindow]) {
        restart = true;
        stagnated = true;
        break;
      }
      solver.iterate_process();
    }
    total_iter += solver.iteration_count();
    if (done) break;
    if (restart) st.restarts++;
    if (stagnated) escalate(gstep, t_op, t_pc, op, pc, st);
  }
  st.iterations = total_iter;
  st.graph_replays = gstep ? gstep->replays : 0;
  if (opt_.verbose) {
    std::fprintf(stderr, "  operator %d x %.4f s, preconditioner %d x %.4f s, residual checks %.3f s\n",
                 t_op->calls, t_op->calls ? t_op->seconds / t_op->calls : 0.0, t_pc ? t_pc->calls : 0,
                 t_pc && t_pc->calls ? t_pc->seconds / t_pc->calls : 0.0, t_check);
  }
  std::vector<double> out(m);
  PAM_HIP(hipMemcpyAsync(out.data(), x.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s_));
  PAM_HIP(hipStreamSynchronize(s_));
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return out;
}

}  // namespace pam

// ---- Fits with gradient data ----

namespace pam {

namespace {

// Up to k sites, evenly strided among those with nonzero data, then the
// others; sorted (the sample of ResidualEvaluator::set_values).
std::vector<int32_t> strided_sample(const std::vector<uint8_t>& nonzero, int64_t k) {
  std::vector<int32_t> nz, zero;
  for (size_t i = 0; i < nonzero.size(); i++) (nonzero[i] ? nz : zero).push_back(static_cast<int32_t>(i));
  std::vector<int32_t> out;
  const int64_t nnz = static_cast<int64_t>(nz.size());
  if (nnz >= k) {
    for (int64_t i = 0; i < k; i++) out.push_back(nz[(i * nnz) / k]);
  } else {
    out = nz;
    for (int64_t i = 0; static_cast<int64_t>(out.size()) < k; i++) out.push_back(zero[i]);
  }
  std::sort(out.begin(), out.end());
  return out;
}

FmmConfig fp64_at_least(const FmmConfig& c, int order) {
  FmmConfig r = c;
  if (r.precision != kFp64) {
    r.precision = kFp64;
    r.order = order;
  }
  r.order = std::max(r.order, order);
  return r;
}

}  // namespace

HermiteResidualEvaluator::HermiteResidualEvaluator(const Model& model, const double* d_pts, int64_t mu,
                                                   const double* d_gpts, int64_t sigma, const FmmConfig& full_cfg,
                                                   hipStream_t s)
    : model_(model),
      s_(s),
      mu_(mu),
      sigma_(sigma),
      d_pts_(d_pts),
      d_gpts_(d_gpts),
      full_(model, d_pts, mu, d_gpts, sigma, full_cfg, s) {}

void HermiteResidualEvaluator::set_values(const std::vector<double>& values) {
  // residual_evaluator.hpp:121-151 with gradient points
  constexpr int64_t kDirect = 1024;
  const int64_t n = mu_ + 3 * sigma_;
  if (static_cast<int64_t>(values.size()) != n) throw_invalid("values.rows() must be mu + 3 sigma");
  values_.upload(values.data(), values.size(), s_);
  std::vector<uint8_t> nzv(static_cast<size_t>(mu_)), nzg(static_cast<size_t>(sigma_));
  for (int64_t i = 0; i < mu_; i++) nzv[i] = values[i] != 0.0;
  for (int64_t k = 0; k < sigma_; k++)
    nzg[k] = values[mu_ + 3 * k] != 0.0 || values[mu_ + 3 * k + 1] != 0.0 || values[mu_ + 3 * k + 2] != 0.0;
  sample_ = strided_sample(nzv, std::min(mu_, kDirect));
  gsample_ = strided_sample(nzg, std::min(sigma_, kDirect));
  const int64_t ns = static_cast<int64_t>(sample_.size()), ng = static_cast<int64_t>(gsample_.size());
  d_sample_.upload(sample_.data(), sample_.size(), s_);
  d_gsample_.upload(gsample_.data(), gsample_.size(), s_);
  sample_pts_.resize(static_cast<size_t>(ns) * 3);
  sample_gpts_.resize(static_cast<size_t>(ng) * 3);
  if (ns > 0) dev_gather_points(d_pts_, d_sample_.get(), ns, sample_pts_.get(), s_);
  if (ng > 0) dev_gather_points(d_gpts_, d_gsample_.get(), ng, sample_gpts_.get(), s_);
  std::vector<double> sv;
  for (int32_t i : sample_) sv.push_back(values[i]);
  for (int32_t k : gsample_)
    for (int c = 0; c < 3; c++) sv.push_back(values[mu_ + 3 * static_cast<int64_t>(k) + c]);
  sample_vals_.upload(sv.data(), sv.size(), s_);
  FmmConfig direct;
  direct.precision = kFp64;
  direct.mode = kModeDirect;
  direct_ = std::make_unique<HermiteEvaluator>(model_, direct, s_);
  direct_->set_sources(d_pts_, mu_, d_gpts_, sigma_);
  direct_->set_targets(sample_pts_.get(), ns, sample_gpts_.get(), ng);
  PAM_HIP(hipStreamSynchronize(s_));  // sv is a host temporary
}

void HermiteResidualEvaluator::full_residual(const double* d_weights, double* d_r) {
  const int64_t n = mu_ + 3 * sigma_;
  y_.resize(static_cast<size_t>(full_.size()));
  full_.apply(d_weights, y_.get());
  dev_copy(values_.get(), d_r, n, s_);
  dev_axpy(-1.0, y_.get(), d_r, n, s_);
}

HermiteResidualEvaluator::Result HermiteResidualEvaluator::converged(const double* d_weights, double tolerance,
                                                                     double grad_tolerance) {
  // residual_evaluator.hpp:56-108
  Result c;
  const int64_t ns = static_cast<int64_t>(sample_.size()), ng = static_cast<int64_t>(gsample_.size());
  const int64_t n = mu_ + 3 * sigma_;
  const bool exact = ns == mu_ && ng == sigma_;
  auto max_diff = [&](const double* a, const double* b, int64_t k) {
    return k > 0 ? dev_max_abs_diff(a, b, k, ws_, s_) : 0.0;
  };
  {
    y_.resize(static_cast<size_t>(std::max(ns + 3 * ng, full_.size())));
    direct_->set_weights(d_weights);
    direct_->evaluate(y_.get());
    if (model_.nugget != 0.0 && ns > 0) {
      tmp_.resize(static_cast<size_t>(std::max(ns, n)));
      dev_gather(d_weights, d_sample_.get(), ns, tmp_.get(), s_);
      dev_axpy(model_.nugget, tmp_.get(), y_.get(), ns, s_);
    }
    c.residual = max_diff(y_.get(), sample_vals_.get(), ns);
    c.grad_residual = max_diff(y_.get() + ns, sample_vals_.get() + ns, 3 * ng);
    c.exact = exact;
    if (c.residual > tolerance || c.grad_residual > grad_tolerance) return c;
    if (exact) {
      c.converged = true;
      return c;
    }
  }
  tmp_.resize(static_cast<size_t>(std::max(ns, n)));
  full_residual(d_weights, tmp_.get());
  c.residual = max_diff(tmp_.get(), nullptr, mu_);
  c.grad_residual = max_diff(tmp_.get() + mu_, nullptr, 3 * sigma_);
  c.exact = true;
  c.converged = c.residual <= tolerance && c.grad_residual <= grad_tolerance;
  return c;
}

HermiteSolver::HermiteSolver(const Model& model, const std::vector<double>& points,
                             const std::vector<double>& grad_points, const FitOptions& opt, hipStream_t s)
    : model_(model),
      opt_(opt),
      s_(s),
      mu_(static_cast<int64_t>(points.size() / 3)),
      sigma_(static_cast<int64_t>(grad_points.size() / 3)),
      n_(mu_ + 3 * sigma_),
      l_(model.l()) {
  auto t0 = std::chrono::steady_clock::now();
  model_.validate();
  if (!(model_.poly_degree == 1 && mu_ == 1 && sigma_ >= 1) && mu_ < l_)
    throw_invalid("points.rows() must be greater than or equal to " + std::to_string(l_));
  resolve_fit_options(model_, opt_);
  d_pts_.upload(points.data(), points.size(), s_);
  d_gpts_.upload(grad_points.data(), grad_points.size(), s_);
  // The gradient rows are less accurate than the values in a mixed-precision
  // FMM (5.8e-5 relative at order 7): the operator of these fits is FP64.
  const FmmConfig op_cfg = fp64_at_least(opt_.op_cfg, 10);
  op_ = std::make_unique<HermiteOperator>(model_, d_pts_.get(), mu_, d_gpts_.get(), sigma_, op_cfg, s_);
  res_ = std::make_unique<HermiteResidualEvaluator>(model_, d_pts_.get(), mu_, d_gpts_.get(), sigma_,
                                                    fp64_at_least(opt_.res_cfg, 10), s_);
  if (opt_.use_preconditioner) {
    RasOptions ro = fit_ras_options(model_, opt_, n_);
    ro.fp32_factors = false;
    if (!ro.use_level_cfg) {
      ro.level_cfg = op_cfg;
      ro.use_level_cfg = true;
    }
    ro.level_cfg = fp64_at_least(ro.level_cfg, 8);
    pc_ = std::make_unique<RasPreconditioner>(model_, points, grad_points, s_, ro, d_pts_.get());
  }
  if (l_ > 0) {
    p_orth_.resize(static_cast<size_t>(n_) * l_);
    dev_poly_matrix_hermite(model_.poly_degree, d_pts_.get(), mu_, d_gpts_.get(), sigma_, p_orth_.get(), s_);
    dev_orthonormalize_cols(p_orth_.get(), n_, l_, ws_, s_);
  }
  PAM_HIP(hipStreamSynchronize(s_));
  setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::vector<double> HermiteSolver::solve(const std::vector<double>& values, double tolerance, double grad_tolerance,
                                         int max_iter, const std::vector<double>* initial, FitStats* stats) {
  // solver.hpp:76-142 with gradient data
  auto t0 = std::chrono::steady_clock::now();
  if (static_cast<int64_t>(values.size()) != n_)
    throw_invalid("values.rows() must be equal to points.rows() + 3 * grad_points.rows()");
  if (!(tolerance > 0.0) || !(grad_tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  const int64_t m = n_ + l_;
  DevBuf<double> x(m), xb(m), rhs(m), dot(10);
  dev_fill(0.0, x.get(), m, s_);
  if (initial) {
    if (static_cast<int64_t>(initial->size()) != m) throw_invalid("initial weights have the wrong size");
    PAM_HIP(hipMemcpyAsync(x.get(), initial->data(), sizeof(double) * m, hipMemcpyHostToDevice, s_));
    if (l_ > 0) {  // orthogonalize the warm start against P (solver.hpp:85-93)
      double h[10];
      dev_multi_dot(p_orth_.get(), n_, l_, x.get(), n_, h, ws_, s_);
      PAM_HIP(hipMemcpyAsync(dot.get(), h, sizeof(double) * l_, hipMemcpyHostToDevice, s_));
      dev_gemv_update(x.get(), p_orth_.get(), n_, l_, dot.get(), n_, -1.0, s_);
    }
    PAM_HIP(hipStreamSynchronize(s_));
  }
  res_->set_values(values);
  FitStats st;
  st.setup_seconds = setup_s_;
  if (pc_) {
    st.precond_setup_seconds = pc_->setup_seconds();
    st.n_levels = pc_->n_levels();
    st.factor_bytes = pc_->factor_bytes();
  }
  int total_iter = 0;
  bool first = true;
  while (true) {
    if (first && !initial) {
      PAM_HIP(hipMemcpyAsync(rhs.get(), values.data(), sizeof(double) * n_, hipMemcpyHostToDevice, s_));
      dev_fill(0.0, rhs.get() + n_, l_, s_);
      dev_fill(0.0, xb.get(), m, s_);
    } else {
      dev_copy(x.get(), xb.get(), m, s_);
      res_->full_residual(xb.get(), rhs.get());
      if (l_ > 0) {  // second block: 0 - P^T [w; g]
        dev_multi_dot_device(op_->poly_matrix(), n_, l_, xb.get(), n_, rhs.get() + n_, ws_, s_);
        dev_scal(-1.0, rhs.get() + n_, l_, s_);
      }
    }
    first = false;
    Fgmres solver(*op_, rhs.get(), m, std::max(max_iter - total_iter, 0), s_);
    solver.set_right_preconditioner(pc_.get());
    solver.setup();
    if (solver.relative_residual() == 0.0) {
      dev_copy(xb.get(), x.get(), m, s_);
      break;
    }
    bool done = false;
    while (true) {
      solver.solution_vector(x.get());
      dev_axpy(1.0, xb.get(), x.get(), m, s_);
      HermiteResidualEvaluator::Result c = res_->converged(x.get(), tolerance, grad_tolerance);
      st.history.push_back(std::max(c.residual, c.grad_residual));
      st.residual = c.residual;
      st.grad_residual = c.grad_residual;
      if (opt_.verbose)
        std::fprintf(stderr, "%8d %s%12.6e %12.6e  (inner rel %.3e)\n", total_iter + solver.iteration_count(),
                     c.exact ? " " : "~", c.residual, c.grad_residual, solver.relative_residual());
      if (c.converged) {
        done = true;
        break;
      }
      if (total_iter + solver.iteration_count() >= max_iter) {
        if (!opt_.keep_unconverged) throw_runtime("reached the maximum number of iterations");
        st.converged = false;
        done = true;
        break;
      }
      solver.iterate_process();
    }
    total_iter += solver.iteration_count();
    if (done) break;
    st.restarts++;
  }
  st.iterations = total_iter;
  std::vector<double> out(m);
  PAM_HIP(hipMemcpyAsync(out.data(), x.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s_));
  PAM_HIP(hipStreamSynchronize(s_));
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return out;
}

}  // namespace pam
