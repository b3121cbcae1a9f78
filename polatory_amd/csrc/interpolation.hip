// Model validation and the interpolation-level evaluators (device resident).
#include "interpolation.hpp"

#include <algorithm>

#include "device_api.hpp"

namespace pam {

int rbf_cpd_order(int kind) {
  switch (kind) {
    case kBh3: return 1;  // polyharmonic_odd.hpp:30, (K+1)/2 for K=1
    case kTh3: return 2;  // K=3
    case kExp: return 0;  // covariance_function_base.hpp:17
  }
  throw_invalid("unknown rbf kind");
}

int Model::cpd_order() const {
  int o = 0;
  for (const auto& r : rbfs) o = std::max(o, rbf_cpd_order(r.kind));
  return o;
}

void Model::validate() const {
  if (rbfs.empty()) throw_invalid("rbfs must not be empty");
  int lo = cpd_order() - 1;
  if (poly_degree < lo || poly_degree > 2)
    throw_invalid("poly_degree must be within " + std::to_string(lo) + " to 2");
  if (!(nugget >= 0.0)) throw_invalid("nugget must be non-negative");
}

Evaluator::Evaluator(const Model& model, const double* bbox_min, const double* bbox_max,
                     double accuracy, hipStream_t s, bool gradient)
    : model_(model), s_(s) {
  model_.validate();
  double acc = accuracy / static_cast<double>(model_.rbfs.size());
  for (const auto& r : model_.rbfs) {
    fmm_.push_back(std::make_unique<FmmEvaluator>(r, bbox_min, bbox_max, false, s));
    fmm_.back()->set_gradient_kernel(gradient);
    fmm_.back()->set_accuracy(acc);
  }
  coeffs_.resize(10);
}

void Evaluator::set_fmm_config(const FmmConfig& c) {
  for (auto& f : fmm_) f->set_config(c);
}

void Evaluator::set_source_points(const double* d_pts, int64_t n) {
  n_src_ = n;
  for (auto& f : fmm_) f->set_source_points(d_pts, n, true);
}

void Evaluator::set_target_points(const double* d_pts, int64_t n) {
  n_trg_ = n;
  trg_pts_.resize(static_cast<size_t>(n) * 3);
  dev_copy(d_pts, trg_pts_.get(), n * 3, s_);
  for (auto& f : fmm_) f->set_target_points(d_pts, n, true);
}

void Evaluator::set_weights(const double* d_w) {
  for (auto& f : fmm_) f->set_weights(d_w, n_src_, true);
  int l = model_.l();
  if (l > 0) dev_copy(d_w + n_src_, coeffs_.get(), l, s_);
}

void Evaluator::evaluate(double* d_out) {
  if (fmm_.size() == 1) {
    fmm_[0]->evaluate(d_out, n_trg_, true);
  } else {
    tmp_.resize(static_cast<size_t>(n_trg_));
    dev_fill(0.0, d_out, n_trg_, s_);
    for (auto& f : fmm_) {
      f->evaluate(tmp_.get(), n_trg_, true);
      dev_axpy(1.0, tmp_.get(), d_out, n_trg_, s_);
    }
  }
  if (model_.l() > 0) launch_poly_eval(model_.poly_degree, trg_pts_.get(), n_trg_, coeffs_.get(), d_out, s_);
}

void Evaluator::evaluate_gradient(double* d_out3) {
  const int64_t n3 = n_trg_ * 3;
  if (fmm_.size() == 1) {
    fmm_[0]->evaluate_gradient(d_out3, n_trg_, true);
  } else {
    tmp_.resize(static_cast<size_t>(n3));
    dev_fill(0.0, d_out3, n3, s_);
    for (auto& f : fmm_) {
      f->evaluate_gradient(tmp_.get(), n_trg_, true);
      dev_axpy(1.0, tmp_.get(), d_out3, n3, s_);
    }
  }
  if (model_.l() > 0) launch_poly_grad(model_.poly_degree, trg_pts_.get(), n_trg_, coeffs_.get(), d_out3, s_);
}

SymmetricEvaluator::SymmetricEvaluator(const Model& model, const double* bbox_min,
                                       const double* bbox_max, double accuracy, hipStream_t s)
    : model_(model), s_(s) {
  model_.validate();
  double acc = accuracy / static_cast<double>(model_.rbfs.size());
  for (const auto& r : model_.rbfs) {
    fmm_.push_back(std::make_unique<FmmEvaluator>(r, bbox_min, bbox_max, true, s));
    fmm_.back()->set_accuracy(acc);
  }
  coeffs_.resize(10);
}

void SymmetricEvaluator::set_fmm_config(const FmmConfig& c) {
  for (auto& f : fmm_) f->set_config(c);
}

void SymmetricEvaluator::set_points(const double* d_pts, int64_t n) {
  n_ = n;
  pts_.resize(static_cast<size_t>(n) * 3);
  dev_copy(d_pts, pts_.get(), n * 3, s_);
  for (auto& f : fmm_) f->set_source_points(d_pts, n, true);
}

void SymmetricEvaluator::set_weights(const double* d_w) {
  for (auto& f : fmm_) f->set_weights(d_w, n_, true);
  int l = model_.l();
  if (l > 0) dev_copy(d_w + n_, coeffs_.get(), l, s_);
}

void SymmetricEvaluator::evaluate(double* d_out) {
  if (fmm_.size() == 1) {
    fmm_[0]->evaluate(d_out, n_, true);
  } else {
    tmp_.resize(static_cast<size_t>(n_));
    dev_fill(0.0, d_out, n_, s_);
    for (auto& f : fmm_) {
      f->evaluate(tmp_.get(), n_, true);
      dev_axpy(1.0, tmp_.get(), d_out, n_, s_);
    }
  }
  if (model_.l() > 0) launch_poly_eval(model_.poly_degree, pts_.get(), n_, coeffs_.get(), d_out, s_);
}

}  // namespace pam
