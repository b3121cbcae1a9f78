// k-fold cross-validation of a model (kriging/cross_validate.hpp:14-47): one
// device fit per fold on the other points, evaluated at the fold's points.
#include <set>

#include "interpolant.hpp"

namespace pam {

std::vector<double> cross_validate(const Model& model, const std::vector<double>& points,
                                   const std::vector<double>& values, const std::vector<int32_t>& set_ids,
                                   double tolerance, int max_iter, double accuracy, const FitOptions& opt) {
  const int64_t n = static_cast<int64_t>(points.size() / 3);
  if (static_cast<int64_t>(values.size()) != n) throw_invalid("values.rows() must be equal to points.rows()");
  if (static_cast<int64_t>(set_ids.size()) != n) throw_invalid("set_ids.rows() must be equal to points.rows()");
  std::vector<double> pred(static_cast<size_t>(n), 0.0);
  const std::set<int32_t> ids(set_ids.begin(), set_ids.end());
  for (int32_t id : ids) {
    std::vector<double> train_pts, train_vals, test_pts;
    std::vector<int64_t> test;
    for (int64_t i = 0; i < n; i++) {
      if (set_ids[i] == id) {
        test.push_back(i);
        for (int k = 0; k < 3; k++) test_pts.push_back(points[3 * i + k]);
      } else {
        for (int k = 0; k < 3; k++) train_pts.push_back(points[3 * i + k]);
        train_vals.push_back(values[i]);
      }
    }
    Interpolant ip(model);
    ip.fit(train_pts, train_vals, tolerance, max_iter, accuracy, opt, nullptr);
    std::vector<double> fit(test.size());
    ip.evaluate(test_pts.data(), static_cast<int64_t>(test.size()), accuracy, fit.data(), false);
    for (size_t j = 0; j < test.size(); j++) pred[test[j]] = fit[j];
  }
  return pred;
}

}  // namespace pam
