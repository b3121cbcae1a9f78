// Interpolant facade (see interpolant.hpp).
#include "interpolant.hpp"

#include "incremental_fitter.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <unordered_map>

namespace pam {

namespace {

struct PointKey {
  uint64_t b[3];
  bool operator==(const PointKey& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
};
struct PointHash {
  size_t operator()(const PointKey& k) const {
    uint64_t h = k.b[0] * 0x9e3779b97f4a7c15ull;
    h ^= k.b[1] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h ^= k.b[2] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return static_cast<size_t>(h);
  }
};
PointKey key_of(const double* p) {
  PointKey k;
  std::memcpy(k.b, p, sizeof k.b);
  return k;
}

bool same_model(const Model& a, const Model& b) {
  if (a.poly_degree != b.poly_degree || a.nugget != b.nugget || a.rbfs.size() != b.rbfs.size()) return false;
  for (size_t i = 0; i < a.rbfs.size(); i++) {
    const auto &x = a.rbfs[i], &y = b.rbfs[i];
    if (x.kind != y.kind || x.p0 != y.p0 || x.p1 != y.p1) return false;
    for (int k = 0; k < 9; k++)
      if (x.aniso[k] != y.aniso[k]) return false;
  }
  return true;
}

}  // namespace

Interpolant::Interpolant(const Model& model) : model_(model) {
  model_.validate();
  s_ = stream_acquire();
}

Interpolant::~Interpolant() {
  ev_.reset();
  gev_.reset();
  if (s_) {
    (void)hipStreamSynchronize(s_);
    stream_release(s_);
  }
}

void Interpolant::throw_if_not_fitted() const {
  if (!fitted_) throw_runtime("the interpolant is not fitted");
}

std::vector<double> Interpolant::build_initial_weights(const std::vector<double>& points,
                                                       const Interpolant& initial) const {
  // interpolant.hpp:232-273: weights of coincident centres, tail from the initial.
  int l = model_.l();
  int64_t mu = static_cast<int64_t>(points.size() / 3);
  std::vector<double> w(static_cast<size_t>(mu + l), 0.0);
  if (!same_model(model_, initial.model_)) return w;
  std::unordered_map<PointKey, int64_t, PointHash> ini;
  int64_t ini_mu = static_cast<int64_t>(initial.centers_.size() / 3);
  for (int64_t i = 0; i < ini_mu; i++) ini.emplace(key_of(&initial.centers_[3 * i]), i);
  for (int64_t i = 0; i < mu; i++) {
    auto it = ini.find(key_of(&points[3 * i]));
    if (it != ini.end()) w[i] = initial.weights_[it->second];
  }
  const int64_t ini_n = ini_mu + static_cast<int64_t>(initial.grad_centers_.size());  // + 3 sigma
  for (int k = 0; k < l; k++) w[mu + k] = initial.weights_[ini_n + k];
  return w;
}

std::vector<double> Interpolant::build_initial_weights(const std::vector<double>& points,
                                                       const std::vector<double>& grad_points,
                                                       const Interpolant& initial) const {
  // interpolant.hpp:232-273: weights of coincident centres and gradient
  // centres, polynomial coefficients from the initial interpolant.
  const int l = model_.l();
  const int64_t mu = static_cast<int64_t>(points.size() / 3), sigma = static_cast<int64_t>(grad_points.size() / 3);
  std::vector<double> w(static_cast<size_t>(mu + 3 * sigma + l), 0.0);
  if (!same_model(model_, initial.model_)) return w;
  const int64_t ini_mu = static_cast<int64_t>(initial.centers_.size() / 3);
  const int64_t ini_sigma = static_cast<int64_t>(initial.grad_centers_.size() / 3);
  std::unordered_map<PointKey, int64_t, PointHash> ini, gini;
  for (int64_t i = 0; i < ini_mu; i++) ini.emplace(key_of(&initial.centers_[3 * i]), i);
  for (int64_t i = 0; i < ini_sigma; i++) gini.emplace(key_of(&initial.grad_centers_[3 * i]), i);
  for (int64_t i = 0; i < mu; i++) {
    auto it = ini.find(key_of(&points[3 * i]));
    if (it != ini.end()) w[i] = initial.weights_[it->second];
  }
  for (int64_t i = 0; i < sigma; i++) {
    auto it = gini.find(key_of(&grad_points[3 * i]));
    if (it != gini.end())
      for (int c = 0; c < 3; c++) w[mu + 3 * i + c] = initial.weights_[ini_mu + 3 * it->second + c];
  }
  for (int k = 0; k < l; k++) w[mu + 3 * sigma + k] = initial.weights_[ini_mu + 3 * ini_sigma + k];
  return w;
}

void Interpolant::set_bbox() {
  for (int k = 0; k < 3; k++) {
    bbox_lo_[k] = INFINITY;
    bbox_hi_[k] = -INFINITY;
  }
  for (const std::vector<double>* v : {&centers_, &grad_centers_})
    for (size_t i = 0; i + 2 < v->size(); i += 3)
      for (int k = 0; k < 3; k++) {
        bbox_lo_[k] = std::min(bbox_lo_[k], (*v)[i + k]);
        bbox_hi_[k] = std::max(bbox_hi_[k], (*v)[i + k]);
      }
}

void Interpolant::fit_hermite(const std::vector<double>& points, const std::vector<double>& grad_points,
                              const std::vector<double>& values, double tolerance, double grad_tolerance,
                              int max_iter, const FitOptions& opt, const Interpolant* initial) {
  const int64_t mu = static_cast<int64_t>(points.size() / 3), sigma = static_cast<int64_t>(grad_points.size() / 3);
  if (static_cast<int64_t>(values.size()) != mu + 3 * sigma)
    throw_invalid("values.rows() must be equal to points.rows() + 3 * grad_points.rows()");
  if (!(tolerance > 0.0) || !(grad_tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  // The initial weights are built before clearing: `initial` may be this.
  std::vector<double> init;
  if (initial && initial->fitted_) init = build_initial_weights(points, grad_points, *initial);
  fitted_ = false;
  ev_.reset();
  gev_.reset();
  hev_.reset();
  {
    HermiteSolver solver(model_, points, grad_points, opt, s_);
    weights_ = solver.solve(values, tolerance, grad_tolerance, max_iter, init.empty() ? nullptr : &init, &stats_);
  }
  centers_ = points;
  grad_centers_ = grad_points;
  fitted_ = true;
  have_eval_bbox_ = false;
  set_bbox();
}

void Interpolant::evaluate_hermite(const double* points, int64_t n, const double* grad_points, int64_t ng,
                                   double* out, bool on_device) {
  throw_if_not_fitted();
  if (n < 0 || ng < 0) throw_invalid("negative point count");
  if (!hev_) {
    FmmConfig c;
    c.precision = kFp64;
    c.order = 10;
    if (have_eval_cfg_) {
      c = eval_cfg_;
      if (c.precision != kFp64) c.precision = kFp64, c.order = std::max(c.order, 10);
    }
    hev_ = std::make_unique<HermiteEvaluator>(model_, c, s_);
    d_centers_.upload(centers_.data(), centers_.size(), s_);
    d_gcenters_.upload(grad_centers_.data(), grad_centers_.size(), s_);
    d_weights_.upload(weights_.data(), weights_.size(), s_);
    hev_->set_sources(d_centers_.get(), static_cast<int64_t>(centers_.size() / 3), d_gcenters_.get(),
                      static_cast<int64_t>(grad_centers_.size() / 3));
  }
  const double *d_in = points, *d_gin = grad_points;
  if (!on_device) {
    d_pts_.upload(points, static_cast<size_t>(n) * 3, s_);
    d_gpts_.upload(grad_points, static_cast<size_t>(ng) * 3, s_);
    d_in = d_pts_.get();
    d_gin = d_gpts_.get();
  }
  double* d_out = out;
  if (!on_device) {
    d_out_.resize(static_cast<size_t>(n + 3 * ng));
    d_out = d_out_.get();
  }
  hev_->set_targets(d_in, n, d_gin, ng);
  hev_->set_weights(d_weights_.get());
  hev_->evaluate(d_out);
  if (!on_device) PAM_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * (n + 3 * ng), hipMemcpyDeviceToHost, s_));
  PAM_HIP(hipStreamSynchronize(s_));
}


void Interpolant::fit(const std::vector<double>& points, const std::vector<double>& values,
                      double tolerance, int max_iter, double accuracy, const FitOptions& opt,
                      const Interpolant* initial) {
  int64_t mu = static_cast<int64_t>(points.size() / 3);
  int l = model_.l();
  if (mu < l) throw_invalid("points.rows() must be greater than or equal to " + std::to_string(l));
  if (static_cast<int64_t>(values.size()) != mu) throw_invalid("values.rows() must be equal to " + std::to_string(mu));
  if (!(tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  if (!(accuracy > 0.0)) throw_invalid("accuracy must be positive");
  std::vector<double> init;
  if (initial) {
    if (initial->fitted_) init = build_initial_weights(points, *initial);
    else initial = nullptr;
  }
  fitted_ = false;
  ev_.reset();
  gev_.reset();
  hev_.reset();
  grad_centers_.clear();
  auto t0 = std::chrono::steady_clock::now();
  auto since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
  };
  {
    Solver solver(model_, points, opt, s_);
    const double t_setup = since(t0);
    weights_ = solver.solve(values, tolerance, max_iter, initial ? &init : nullptr, &stats_);
    const double t_solve = since(t0);
    if (opt.verbose) std::fprintf(stderr, "  fit: solver set-up %.3f s, solve to %.3f s\n", t_setup, t_solve);
    t0 = std::chrono::steady_clock::now();
  }
  if (opt.verbose) std::fprintf(stderr, "  fit: solver teardown %.3f s\n", since(t0));
  centers_ = points;
  fitted_ = true;
  have_eval_bbox_ = false;
  for (int k = 0; k < 3; k++) {
    bbox_lo_[k] = INFINITY;
    bbox_hi_[k] = -INFINITY;
  }
  for (int64_t i = 0; i < mu; i++)
    for (int k = 0; k < 3; k++) {
      bbox_lo_[k] = std::min(bbox_lo_[k], points[3 * i + k]);
      bbox_hi_[k] = std::max(bbox_hi_[k], points[3 * i + k]);
    }
}

void Interpolant::fit_incrementally(const std::vector<double>& points, const std::vector<double>& values,
                                    double tolerance, int max_iter, double accuracy, const FitOptions& opt) {
  fitted_ = false;
  ev_.reset();
  gev_.reset();
  IncrementalResult r = pam::fit_incrementally(model_, points, values, tolerance, max_iter, accuracy, opt, s_);
  std::vector<double> c(r.centers.size() * 3);
  for (size_t i = 0; i < r.centers.size(); i++)
    for (int k = 0; k < 3; k++) c[3 * i + k] = points[3 * r.centers[i] + k];
  set_fitted(c, r.weights);
  stats_ = FitStats{};
  stats_.iterations = r.iterations;
  stats_.rounds = r.rounds;
}

void Interpolant::fit_incrementally_hermite(const std::vector<double>& points, const std::vector<double>& grad_points,
                                            const std::vector<double>& values, double tolerance,
                                            double grad_tolerance, int max_iter, const FitOptions& opt) {
  fitted_ = false;
  ev_.reset();
  gev_.reset();
  hev_.reset();
  IncrementalHermiteResult r =
      pam::fit_incrementally_hermite(model_, points, grad_points, values, tolerance, grad_tolerance, max_iter, opt, s_);
  std::vector<double> c(r.centers.size() * 3), g(r.grad_centers.size() * 3);
  for (size_t i = 0; i < r.centers.size(); i++)
    for (int k = 0; k < 3; k++) c[3 * i + k] = points[3 * r.centers[i] + k];
  for (size_t i = 0; i < r.grad_centers.size(); i++)
    for (int k = 0; k < 3; k++) g[3 * i + k] = grad_points[3 * r.grad_centers[i] + k];
  set_fitted(c, r.weights, g);
  stats_ = FitStats{};
  stats_.iterations = r.iterations;
  stats_.rounds = r.rounds;
}

void Interpolant::fit_inequality(const std::vector<double>& points, const std::vector<double>& values,
                                 const std::vector<double>& lb, const std::vector<double>& ub, double tolerance,
                                 int max_iter, double accuracy, const FitOptions& opt, const Interpolant* initial) {
  // The initial weights are built before clearing: `initial` may be this.
  std::vector<double> init;
  if (initial && initial->fitted_) init = build_initial_weights(points, *initial);
  fitted_ = false;
  ev_.reset();
  gev_.reset();
  InequalityResult r = pam::fit_inequality(model_, points, values, lb, ub, tolerance, max_iter, accuracy, opt,
                                           init.empty() ? nullptr : &init, s_);
  std::vector<double> c(r.centers.size() * 3);
  for (size_t i = 0; i < r.centers.size(); i++)
    for (int k = 0; k < 3; k++) c[3 * i + k] = points[3 * r.centers[i] + k];
  set_fitted(c, r.weights);
  stats_ = FitStats{};
  stats_.iterations = r.iterations;
  stats_.rounds = r.rounds;
}

void Interpolant::set_fitted(const std::vector<double>& centers, const std::vector<double>& weights) {
  int64_t mu = static_cast<int64_t>(centers.size() / 3);
  if (static_cast<int64_t>(weights.size()) != mu + model_.l()) throw_invalid("weights.rows() must be centers.rows() + l");
  centers_ = centers;
  weights_ = weights;
  fitted_ = true;
  have_eval_bbox_ = false;
  ev_.reset();
  gev_.reset();
  hev_.reset();
  grad_centers_.clear();
  for (int k = 0; k < 3; k++) {
    bbox_lo_[k] = INFINITY;
    bbox_hi_[k] = -INFINITY;
  }
  for (int64_t i = 0; i < mu; i++)
    for (int k = 0; k < 3; k++) {
      bbox_lo_[k] = std::min(bbox_lo_[k], centers[3 * i + k]);
      bbox_hi_[k] = std::max(bbox_hi_[k], centers[3 * i + k]);
    }
}

void Interpolant::set_fitted(const std::vector<double>& centers, const std::vector<double>& weights,
                             const std::vector<double>& grad_centers) {
  if (grad_centers.empty()) {
    set_fitted(centers, weights);
    return;
  }
  const int64_t mu = static_cast<int64_t>(centers.size() / 3), sigma = static_cast<int64_t>(grad_centers.size() / 3);
  if (static_cast<int64_t>(weights.size()) != mu + 3 * sigma + model_.l())
    throw_invalid("weights.rows() must be centers.rows() + 3 grad_centers.rows() + l");
  centers_ = centers;
  grad_centers_ = grad_centers;
  weights_ = weights;
  fitted_ = true;
  have_eval_bbox_ = false;
  ev_.reset();
  gev_.reset();
  hev_.reset();
  set_bbox();
}

void Interpolant::set_eval_config(const FmmConfig* c) {
  have_eval_cfg_ = c != nullptr;
  if (c) eval_cfg_ = *c;
  ev_.reset();
  gev_.reset();
}

Evaluator& Interpolant::evaluator() {
  throw_if_not_fitted();
  if (!ev_) {
    int64_t mu = static_cast<int64_t>(centers_.size() / 3);
    ev_ = std::make_unique<Evaluator>(model_, have_eval_bbox_ ? eval_lo_ : bbox_lo_,
                                      have_eval_bbox_ ? eval_hi_ : bbox_hi_, ev_acc_ > 0 ? ev_acc_ : INFINITY, s_);
    if (have_eval_cfg_) ev_->set_fmm_config(eval_cfg_);
    d_centers_.upload(centers_.data(), centers_.size(), s_);
    d_weights_.upload(weights_.data(), weights_.size(), s_);
    ev_->set_source_points(d_centers_.get(), mu);
    ev_->set_weights(d_weights_.get());
    PAM_HIP(hipStreamSynchronize(s_));
  }
  return *ev_;
}

Evaluator& Interpolant::gradient_evaluator() {
  throw_if_not_fitted();
  if (!gev_) {
    int64_t mu = static_cast<int64_t>(centers_.size() / 3);
    gev_ = std::make_unique<Evaluator>(model_, have_eval_bbox_ ? eval_lo_ : bbox_lo_,
                                       have_eval_bbox_ ? eval_hi_ : bbox_hi_, gev_acc_ > 0 ? gev_acc_ : INFINITY,
                                       s_, true);
    if (have_eval_cfg_) gev_->set_fmm_config(eval_cfg_);
    d_centers_.upload(centers_.data(), centers_.size(), s_);
    d_weights_.upload(weights_.data(), weights_.size(), s_);
    gev_->set_source_points(d_centers_.get(), mu);
    gev_->set_weights(d_weights_.get());
    PAM_HIP(hipStreamSynchronize(s_));
  }
  return *gev_;
}

void Interpolant::set_evaluation_bbox(const double lo[3], const double hi[3], double accuracy,
                                      double grad_accuracy) {
  throw_if_not_fitted();
  if (!(accuracy > 0.0) || !(grad_accuracy > 0.0)) throw_invalid("accuracy must be positive");
  for (int k = 0; k < 3; k++)
    if (!(lo[k] <= hi[k])) throw_invalid("bbox min must not exceed bbox max");
  for (int k = 0; k < 3; k++) {
    eval_lo_[k] = std::min(lo[k], bbox_lo_[k]);
    eval_hi_[k] = std::max(hi[k], bbox_hi_[k]);
  }
  have_eval_bbox_ = true;
  ev_.reset();
  gev_.reset();
  ev_acc_ = accuracy;
  gev_acc_ = grad_accuracy;
}

void Interpolant::evaluate_impl(const double* points, int64_t n, double* out, bool on_device) {
  evaluate(points, n, ev_acc_ > 0.0 ? ev_acc_ : INFINITY, out, on_device);
}

void Interpolant::evaluate(const double* points, int64_t n, double accuracy, double* out, bool on_device) {
  throw_if_not_fitted();
  if (!(accuracy > 0.0)) throw_invalid("accuracy must be positive");
  if (!grad_centers_.empty()) {
    evaluate_hermite(points, n, nullptr, 0, out, on_device);
    return;
  }
  if (accuracy != ev_acc_) {
    // A box installed by set_evaluation_bbox belongs to its accuracy; at
    // another accuracy the evaluator is rebuilt for the centres' box, as the
    // reference rebuilds it per call (interpolant.hpp:58-67).
    ev_.reset();
    ev_acc_ = accuracy;
    have_eval_bbox_ = false;
  }
  Evaluator& ev = evaluator();
  const double* d_in = points;
  if (!on_device) {
    d_pts_.upload(points, static_cast<size_t>(n) * 3, s_);
    d_in = d_pts_.get();
  }
  double* d_out = out;
  if (!on_device) {
    d_out_.resize(static_cast<size_t>(n));
    d_out = d_out_.get();
  }
  ev.set_target_points(d_in, n);
  ev.evaluate(d_out);
  if (!on_device) {
    PAM_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, s_));
  }
  PAM_HIP(hipStreamSynchronize(s_));
}

void Interpolant::evaluate_gradient(const double* points, int64_t n, double accuracy, double* out3,
                                    bool on_device) {
  throw_if_not_fitted();
  if (!(accuracy > 0.0)) throw_invalid("accuracy must be positive");
  if (n < 0) throw_invalid("negative point count");
  if (!grad_centers_.empty()) {
    evaluate_hermite(nullptr, 0, points, n, out3, on_device);
    return;
  }
  if (accuracy != gev_acc_) {
    gev_.reset();
    gev_acc_ = accuracy;
    have_eval_bbox_ = false;
  }
  Evaluator& ev = gradient_evaluator();
  const double* d_in = points;
  if (!on_device) {
    d_pts_.upload(points, static_cast<size_t>(n) * 3, s_);
    d_in = d_pts_.get();
  }
  double* d_out = out3;
  if (!on_device) {
    d_out_.resize(static_cast<size_t>(n) * 3);
    d_out = d_out_.get();
  }
  ev.set_target_points(d_in, n);
  ev.evaluate_gradient(d_out);
  if (!on_device) {
    PAM_HIP(hipMemcpyAsync(out3, d_out, sizeof(double) * n * 3, hipMemcpyDeviceToHost, s_));
  }
  PAM_HIP(hipStreamSynchronize(s_));
}

int64_t Interpolant::resident_bytes() const {
  return static_cast<int64_t>(d_centers_.bytes() + d_weights_.bytes());
}

}  // namespace pam
