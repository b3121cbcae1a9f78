// Fitting on subsets of the points: the incremental fitter
// (interpolation/incremental_fitter.hpp:25-183), the inequality fitter
// (interpolation/inequality_fitter.hpp:25-187) and the distance filter both
// thin their candidates with (point_cloud/distance_filter.hpp:16-99).
#pragma once

#include <cstdint>
#include <vector>

#include "solver.hpp"

namespace pam {

// Greedy thinning: walks `indices` in order and drops every later point
// within `distance` of a kept one (radius search on a uniform grid of cell
// size `distance` instead of a k-d tree).
class DistanceFilter {
 public:
  static constexpr double kMinDistance = 1.0842021724855044e-19;  // sqrt(float min normal)
  explicit DistanceFilter(const std::vector<double>& xyz) : xyz_(xyz) {}
  std::vector<int64_t> filter(double distance, const std::vector<int64_t>& indices) const;

 private:
  const std::vector<double>& xyz_;
};

struct IncrementalResult {
  std::vector<int64_t> centers;  // indices into the points, in the order they were added
  std::vector<double> weights;   // centres' weights then the polynomial coefficients
  int rounds = 0;
  int iterations = 0;  // FGMRES iterations over all rounds
};

// Fits on a growing subset of the points: each round solves on the current
// centres, evaluates the residuals at the other points, and adds the points
// whose residual is at least `tolerance`, largest first, thinned by a
// distance that halves every round. Stops when no residual reaches the
// tolerance or every point is a centre.
IncrementalResult fit_incrementally(const Model& model, const std::vector<double>& points,
                                    const std::vector<double>& values, double tolerance, int max_iter,
                                    double accuracy, const FitOptions& opt, hipStream_t s);

struct IncrementalHermiteResult {
  std::vector<int64_t> centers, grad_centers;  // in the order they were added
  std::vector<double> weights;                 // [w; g; c]
  int rounds = 0;
  int iterations = 0;
};

// fit_incrementally with gradient points (incremental_fitter.hpp:36-183):
// values has n + 3 ng entries (values, then gradients x y z).
IncrementalHermiteResult fit_incrementally_hermite(const Model& model, const std::vector<double>& points,
                                                   const std::vector<double>& grad_points,
                                                   const std::vector<double>& values, double tolerance,
                                                   double grad_tolerance, int max_iter, const FitOptions& opt,
                                                   hipStream_t s);

struct InequalityResult {
  std::vector<int64_t> centers;  // equality points, then the active bounds in index order
  std::vector<double> weights;   // centres' weights then the polynomial coefficients
  int rounds = 0;
  int iterations = 0;
};

// Fits values where they are given (not NaN) and keeps the fit within
// [lb, ub] where bounds are given, by an active set: a bound joins the
// centres (fitted at the bound) when the fit violates it by more than
// `tolerance`, thinned by a halving distance, and leaves when its weight
// takes the sign that pulls the fit back inside. initial: n + l weights or
// nullptr.
InequalityResult fit_inequality(const Model& model, const std::vector<double>& points,
                                const std::vector<double>& values, const std::vector<double>& lb,
                                const std::vector<double>& ub, double tolerance, int max_iter, double accuracy,
                                const FitOptions& opt, const std::vector<double>* initial, hipStream_t s);

}  // namespace pam
