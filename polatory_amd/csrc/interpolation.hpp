// Interpolation layer on the device: model, evaluators with polynomial terms,
// the fit operator, residual evaluation, FGMRES solver and the fit driver.
// Restates include/polatory/{model.hpp, interpolation/*.hpp} for the value
// kernels (no gradient data).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "fmm_evaluator.hpp"
#include "krylov.hpp"
#include "linalg.hpp"

namespace pam {

// model.hpp:32-183: rbfs, polynomial degree (cpd_order-1 .. 2), nugget >= 0.
struct Model {
  std::vector<RbfDesc> rbfs;
  int poly_degree = -1;
  double nugget = 0.0;
  int l() const { return poly_basis_size(poly_degree); }
  int cpd_order() const;
  void validate() const;
};

int rbf_cpd_order(int kind);

// interpolation/evaluator.hpp:50-144: sum over rbfs of FMM values + polynomial.
class Evaluator {
 public:
  // gradient: the FMM evaluators are gradient-transpose evaluators whose
  // accuracy search measures gradients (evaluate_gradient only).
  Evaluator(const Model& model, const double* bbox_min, const double* bbox_max, double accuracy,
            hipStream_t s, bool gradient = false);
  void set_fmm_config(const FmmConfig& c);
  void set_source_points(const double* d_pts, int64_t n);
  void set_target_points(const double* d_pts, int64_t n);
  void set_weights(const double* d_w);  // n_src + l entries
  void evaluate(double* d_out);         // n_trg entries
  // Gradients at the targets (the gradient-transpose terms and the polynomial
  // gradient of interpolation/evaluator.hpp:68-74): 3 n_trg entries.
  void evaluate_gradient(double* d_out3);
  int64_t n_src() const { return n_src_; }
  int64_t n_trg() const { return n_trg_; }
  FmmEvaluator& fmm(int i) { return *fmm_.at(i); }

 private:
  Model model_;
  hipStream_t s_;
  std::vector<std::unique_ptr<FmmEvaluator>> fmm_;
  DevBuf<double> trg_pts_, coeffs_, tmp_;
  int64_t n_src_ = 0, n_trg_ = 0;
};

// interpolation/symmetric_evaluator.hpp:40-114: values at the source points.
class SymmetricEvaluator {
 public:
  SymmetricEvaluator(const Model& model, const double* bbox_min, const double* bbox_max,
                     double accuracy, hipStream_t s);
  void set_fmm_config(const FmmConfig& c);
  void set_points(const double* d_pts, int64_t n);
  void set_weights(const double* d_w);  // n + l entries
  void evaluate(double* d_out);         // n entries

 private:
  Model model_;
  hipStream_t s_;
  std::vector<std::unique_ptr<FmmEvaluator>> fmm_;
  DevBuf<double> pts_, coeffs_, tmp_;
  int64_t n_ = 0;
};

}  // namespace pam
