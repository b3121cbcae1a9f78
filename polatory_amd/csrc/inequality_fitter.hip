// Active-set fitting with bounds (host loop over the device solver and
// evaluator).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <set>

#include "incremental_fitter.hpp"
#include "interpolation.hpp"

namespace pam {

InequalityResult fit_inequality(const Model& model, const std::vector<double>& points,
                                const std::vector<double>& values, const std::vector<double>& lb,
                                const std::vector<double>& ub, double tolerance, int max_iter, double accuracy,
                                const FitOptions& opt, const std::vector<double>* initial, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(points.size() / 3);
  const int l = model.l();
  // interpolant.hpp:152-172
  if (model.nugget != 0.0) throw_runtime("Non-zero nugget is not supported");
  if (static_cast<int64_t>(values.size()) != n) throw_invalid("values.rows() must be equal to points.rows()");
  if (static_cast<int64_t>(lb.size()) != n) throw_invalid("values_lb.rows() must be equal to points.rows()");
  if (static_cast<int64_t>(ub.size()) != n) throw_invalid("values_ub.rows() must be equal to points.rows()");
  if (!(tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  if (!(accuracy > 0.0)) throw_invalid("accuracy must be positive");
  if (initial && static_cast<int64_t>(initial->size()) != n + l) throw_invalid("initial weights must have n + l entries");

  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], points[3 * i + k]);
      hi[k] = std::max(hi[k], points[3 * i + k]);
    }
  double dist = 0.0;
  if (n > 0)
    dist = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                     (hi[2] - lo[2]) * (hi[2] - lo[2]));
  dist = std::max(dist, DistanceFilter::kMinDistance);
  DistanceFilter filter(points);

  // inequality_fitter.hpp:52-61: equality points, and points with a bound
  std::vector<int64_t> eq, ineq;
  std::vector<int64_t> ineq_pos(static_cast<size_t>(n), -1);
  for (int64_t i = 0; i < n; i++) {
    if (!std::isnan(values[i])) eq.push_back(i);
    if (!std::isnan(lb[i]) || !std::isnan(ub[i])) {
      ineq_pos[i] = static_cast<int64_t>(ineq.size());
      ineq.push_back(i);
    }
  }
  const int64_t n_eq = static_cast<int64_t>(eq.size());
  const int64_t n_ineq = static_cast<int64_t>(ineq.size());
  std::vector<double> ineq_pts(static_cast<size_t>(n_ineq) * 3);
  for (int64_t i = 0; i < n_ineq; i++)
    for (int k = 0; k < 3; k++) ineq_pts[3 * i + k] = points[3 * ineq[i] + k];

  std::vector<double> weights = initial ? *initial : std::vector<double>(static_cast<size_t>(n + l), 0.0);
  std::set<int64_t> act_lb, act_ub;
  Evaluator res_eval(model, lo, hi, accuracy, s);
  DevBuf<double> d_centers, d_weights, d_ineq, d_fit;
  d_ineq.upload(ineq_pts.data(), ineq_pts.size(), s);
  d_fit.resize(static_cast<size_t>(n_ineq));

  InequalityResult res;
  std::vector<double> cw;
  // The reference loops until the active set stops changing; the bound on
  // rounds only turns a cycle into an error.
  const int64_t max_rounds = 4 * n + 64;
  for (;;) {
    if (res.rounds++ >= max_rounds) throw_runtime("the active set did not settle");
    std::vector<int64_t> centers(eq);
    std::set_union(act_lb.begin(), act_lb.end(), act_ub.begin(), act_ub.end(), std::back_inserter(centers));
    const int64_t nc = static_cast<int64_t>(centers.size());
    std::vector<double> fit(static_cast<size_t>(n_ineq), 0.0);
    if (nc > 0 && nc >= l) {
      std::vector<double> cpts(static_cast<size_t>(nc) * 3), cvals(static_cast<size_t>(nc));
      cw.assign(static_cast<size_t>(nc + l), 0.0);
      for (int64_t i = 0; i < nc; i++) {
        const int64_t j = centers[i];
        for (int k = 0; k < 3; k++) cpts[3 * i + k] = points[3 * j + k];
        cvals[i] = i < n_eq ? values[j] : (act_lb.count(j) ? lb[j] : ub[j]);
        cw[i] = weights[j];
      }
      for (int k = 0; k < l; k++) cw[nc + k] = weights[n + k];
      Solver solver(model, cpts, opt, s);
      FitStats st;
      cw = solver.solve(cvals, tolerance, max_iter, &cw, &st);
      res.iterations += st.iterations;
      for (int64_t i = 0; i < nc; i++) weights[centers[i]] = cw[i];
      for (int k = 0; k < l; k++) weights[n + k] = cw[nc + k];
      if (n_ineq > 0) {
        d_centers.upload(cpts.data(), cpts.size(), s);
        d_weights.upload(cw.data(), cw.size(), s);
        res_eval.set_source_points(d_centers.get(), nc);
        res_eval.set_weights(d_weights.get());
        res_eval.set_target_points(d_ineq.get(), n_ineq);
        res_eval.evaluate(d_fit.get());
        d_fit.download(fit.data(), fit.size(), s);
        PAM_HIP(hipStreamSynchronize(s));
      }
    } else {
      cw.assign(static_cast<size_t>(l), 0.0);
    }

    // Non-centre points by bound violation, largest first, thinned
    // (inequality_fitter.hpp:122-137). Each point's violation is taken at its
    // own fitted value.
    std::vector<char> is_center(static_cast<size_t>(n), 0);
    for (int64_t j : centers) is_center[j] = 1;
    std::vector<int64_t> idx;
    std::vector<double> viol;
    for (int64_t j = 0; j < n; j++) {
      if (is_center[j]) continue;
      double r = 0.0;
      const int64_t p = ineq_pos[j];
      if (p >= 0) {
        const double lr = std::isnan(lb[j]) ? 0.0 : std::max(lb[j] - fit[p], 0.0);
        const double ur = std::isnan(ub[j]) ? 0.0 : std::max(fit[p] - ub[j], 0.0);
        r = std::max(lr, ur);
      }
      idx.push_back(j);
      viol.push_back(r);
    }
    std::vector<int64_t> order(idx.size());
    std::iota(order.begin(), order.end(), int64_t{0});
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return viol[a] > viol[b]; });
    std::vector<int64_t> sorted(idx.size());
    for (size_t i = 0; i < order.size(); i++) sorted[i] = idx[order[i]];
    std::vector<char> kept(static_cast<size_t>(n), 0);
    for (int64_t j : filter.filter(dist, sorted)) kept[j] = 1;

    // Update the active set (inequality_fitter.hpp:139-181). A point that
    // also has a value is fitted at that value and never joins as a bound
    // (the reference would add it to the centres a second time).
    bool changed = false;
    for (int64_t i = 0; i < n_ineq; i++) {
      const int64_t j = ineq[i];
      if (!std::isnan(values[j])) continue;
      if (!std::isnan(lb[j])) {
        if (act_lb.count(j)) {
          if (weights[j] <= 0.0) {
            act_lb.erase(j);
            changed = true;
          }
        } else if (fit[i] < lb[j] - tolerance) {
          if (kept[j]) {
            act_lb.insert(j);
            weights[j] = 0.0;
          }
          changed = true;
        }
      }
      if (!std::isnan(ub[j])) {
        if (act_ub.count(j)) {
          if (weights[j] >= 0.0) {
            act_ub.erase(j);
            changed = true;
          }
        } else if (fit[i] > ub[j] + tolerance) {
          if (kept[j]) {
            act_ub.insert(j);
            weights[j] = 0.0;
          }
          changed = true;
        }
      }
    }
    if (!changed) {
      res.centers = std::move(centers);
      break;
    }
    dist = std::max(0.5 * dist, DistanceFilter::kMinDistance);
  }
  if (res.centers.size() < static_cast<size_t>(std::max(l, 1)))
    throw_runtime("the fit has fewer centres than the polynomial needs");
  res.weights = std::move(cw);
  return res;
}

}  // namespace pam
