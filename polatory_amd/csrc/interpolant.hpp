// Interpolant facade (include/polatory/interpolant.hpp:34-367) for the value
// kernels: fit with FGMRES + RAS on the device, evaluate with the device FMM.
#pragma once

#include <memory>
#include <vector>

#include "solver.hpp"

namespace pam {

class Interpolant {
 public:
  explicit Interpolant(const Model& model);
  ~Interpolant();
  Interpolant(const Interpolant&) = delete;
  Interpolant& operator=(const Interpolant&) = delete;

  // interpolant.hpp:77-113
  void fit(const std::vector<double>& points, const std::vector<double>& values, double tolerance,
           int max_iter, double accuracy, const FitOptions& opt, const Interpolant* initial);
  // interpolant.hpp:77-113 with gradient data: values has mu + 3 sigma entries
  // (values, then x y z per gradient point); the weights are [w; g; c].
  void fit_hermite(const std::vector<double>& points, const std::vector<double>& grad_points,
                   const std::vector<double>& values, double tolerance, double grad_tolerance, int max_iter,
                   const FitOptions& opt, const Interpolant* initial);
  // interpolant.hpp:58-67: values at points, then gradients (x y z) at
  // grad_points, of an interpolant with value and gradient centres (FP64 FMM
  // of order 10, or the evaluation configuration raised to FP64). Host or
  // device pointers; out has n + 3 ng entries.
  void evaluate_hermite(const double* points, int64_t n, const double* grad_points, int64_t ng, double* out,
                        bool on_device);
  const std::vector<double>& grad_centers() const { return grad_centers_; }
  // interpolant.hpp:54-67 (+ polynomial terms); pointers are host or device.
  void evaluate(const double* points, int64_t n, double accuracy, double* out, bool on_device);
  // Gradients at grad points (the grad_points half of interpolant.hpp:58-67):
  // out has 3 n entries, x y z per point; accuracy is the grad_accuracy.
  void evaluate_gradient(const double* points, int64_t n, double accuracy, double* out3, bool on_device);
  // interpolant.hpp:115-147: fits on a growing subset of the points; the
  // centres are the points chosen, in the order they were added.
  void fit_incrementally(const std::vector<double>& points, const std::vector<double>& values, double tolerance,
                         int max_iter, double accuracy, const FitOptions& opt);
  // interpolant.hpp:121-147 with gradient data.
  void fit_incrementally_hermite(const std::vector<double>& points, const std::vector<double>& grad_points,
                                 const std::vector<double>& values, double tolerance, double grad_tolerance,
                                 int max_iter, const FitOptions& opt);
  // interpolant.hpp:149-190: values where not NaN, bounds lb/ub where not NaN;
  // the centres are the equality points, then the active bounds.
  void fit_inequality(const std::vector<double>& points, const std::vector<double>& values,
                      const std::vector<double>& lb, const std::vector<double>& ub, double tolerance, int max_iter,
                      double accuracy, const FitOptions& opt, const Interpolant* initial);
  // interpolant.hpp:201-210: the evaluators that evaluate_impl uses are built
  // for the union of this box and the centres' box, at these accuracies
  // (the seam isosurface::RbfFieldFunction calls, rbf_field_function.hpp:21-27).
  void set_evaluation_bbox(const double lo[3], const double hi[3], double accuracy, double grad_accuracy);
  // interpolant.hpp:69-75: evaluation with the evaluator as it stands (the
  // accuracy and box of the last set_evaluation_bbox or evaluate call).
  void evaluate_impl(const double* points, int64_t n, double* out, bool on_device);
  // Installs centres and weights directly (a saved or externally fitted model).
  void set_fitted(const std::vector<double>& centers, const std::vector<double>& weights);
  // The same with gradient centres: weights = [w; g; c].
  void set_fitted(const std::vector<double>& centers, const std::vector<double>& weights,
                  const std::vector<double>& grad_centers);
  void set_eval_config(const FmmConfig* c);

  bool fitted() const { return fitted_; }
  const Model& model() const { return model_; }
  const std::vector<double>& centers() const { return centers_; }
  const std::vector<double>& weights() const { return weights_; }
  const FitStats& fit_stats() const { return stats_; }
  hipStream_t stream() const { return s_; }
  Evaluator& evaluator();
  Evaluator& gradient_evaluator();
  int64_t resident_bytes() const;

 private:
  std::vector<double> build_initial_weights(const std::vector<double>& points,
                                            const Interpolant& initial) const;
  std::vector<double> build_initial_weights(const std::vector<double>& points,
                                            const std::vector<double>& grad_points,
                                            const Interpolant& initial) const;
  void set_bbox();
  void throw_if_not_fitted() const;

  Model model_;
  hipStream_t s_ = nullptr;
  bool fitted_ = false;
  std::vector<double> centers_, weights_, grad_centers_;
  std::unique_ptr<HermiteEvaluator> hev_;
  DevBuf<double> d_gcenters_, d_gpts_;
  FitStats stats_;
  std::unique_ptr<Evaluator> ev_, gev_;
  double ev_acc_ = -1.0, gev_acc_ = -1.0;
  bool have_eval_cfg_ = false;
  FmmConfig eval_cfg_;
  DevBuf<double> d_centers_, d_weights_, d_pts_, d_out_;
  double bbox_lo_[3] = {0, 0, 0}, bbox_hi_[3] = {0, 0, 0};
  // The box of the last set_evaluation_bbox (centres' box included); cleared
  // by every fit.
  bool have_eval_bbox_ = false;
  double eval_lo_[3] = {0, 0, 0}, eval_hi_[3] = {0, 0, 0};
};

// kriging/cross_validate.hpp:14-47: prediction at every point from the fit
// on the points of the other sets (set_ids labels the folds).
std::vector<double> cross_validate(const Model& model, const std::vector<double>& points,
                                   const std::vector<double>& values, const std::vector<int32_t>& set_ids,
                                   double tolerance, int max_iter, double accuracy, const FitOptions& opt);

}  // namespace pam
