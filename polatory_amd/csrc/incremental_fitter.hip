// Incremental fitting on growing subsets of the points (host loop over the
// device solver and evaluator).
#include "incremental_fitter.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "device_api.hpp"
#include "interpolation.hpp"

namespace pam {

std::vector<int64_t> DistanceFilter::filter(double distance, const std::vector<int64_t>& indices) const {
  if (!(distance >= kMinDistance)) throw_invalid("distance must be larger than or equal to 1.0842021724855044e-19");
  const int64_t n_pts = static_cast<int64_t>(xyz_.size() / 3);
  if (indices.empty()) return {};
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i : indices) {
    if (i < 0 || i >= n_pts) throw_invalid("index out of range");
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], xyz_[3 * i + k]);
      hi[k] = std::max(hi[k], xyz_[3 * i + k]);
    }
  }
  // Cells no smaller than the distance (so 27 cells hold every neighbour),
  // and at most 2^20 per axis (so a key packs into 63 bits).
  double extent = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  const double cell = std::max(distance, extent / 1048575.0);
  auto coord = [&](int64_t i, int k) {
    return static_cast<int64_t>(std::floor((xyz_[3 * i + k] - lo[k]) / cell));
  };
  auto key_of = [](int64_t x, int64_t y, int64_t z) {
    return static_cast<uint64_t>(x) | (static_cast<uint64_t>(y) << 21) | (static_cast<uint64_t>(z) << 42);
  };
  std::vector<std::pair<uint64_t, int64_t>> grid;
  grid.reserve(indices.size());
  for (int64_t i : indices) grid.emplace_back(key_of(coord(i, 0), coord(i, 1), coord(i, 2)), i);
  std::sort(grid.begin(), grid.end());
  const double r2 = distance * distance;
  std::vector<char> removed(static_cast<size_t>(n_pts), 0);
  for (int64_t i : indices) {
    if (removed[i]) continue;
    const int64_t c[3] = {coord(i, 0), coord(i, 1), coord(i, 2)};
    for (int64_t dx = -1; dx <= 1; dx++)
      for (int64_t dy = -1; dy <= 1; dy++)
        for (int64_t dz = -1; dz <= 1; dz++) {
          const int64_t x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
          if (x < 0 || y < 0 || z < 0) continue;
          const uint64_t key = key_of(x, y, z);
          auto it = std::lower_bound(grid.begin(), grid.end(), std::make_pair(key, INT64_MIN));
          for (; it != grid.end() && it->first == key; ++it) {
            const int64_t j = it->second;
            if (j == i) continue;
            double d2 = 0.0;
            for (int k = 0; k < 3; k++) {
              const double d = xyz_[3 * j + k] - xyz_[3 * i + k];
              d2 += d * d;
            }
            if (d2 <= r2) removed[j] = 1;
          }
        }
  }
  std::vector<int64_t> out;
  for (int64_t i : indices)
    if (!removed[i]) out.push_back(i);
  return out;
}

IncrementalResult fit_incrementally(const Model& model, const std::vector<double>& points,
                                    const std::vector<double>& values, double tolerance, int max_iter,
                                    double accuracy, const FitOptions& opt, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(points.size() / 3);
  const int l = model.l();
  if (static_cast<int64_t>(values.size()) != n) throw_invalid("values.rows() must be equal to " + std::to_string(n));
  if (!(tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  if (!(accuracy > 0.0)) throw_invalid("accuracy must be positive");
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], points[3 * i + k]);
      hi[k] = std::max(hi[k], points[3 * i + k]);
    }
  // incremental_fitter.hpp:53: the first thinning distance is the bbox diagonal
  double dist = 0.0;
  if (n > 0)
    dist = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                     (hi[2] - lo[2]) * (hi[2] - lo[2]));
  dist = std::max(dist, DistanceFilter::kMinDistance);
  DistanceFilter filter(points);

  IncrementalResult res;
  std::vector<double> weights(static_cast<size_t>(l), 0.0);  // always (centres + l) entries
  Evaluator res_eval(model, lo, hi, accuracy, s);
  DevBuf<double> d_centers, d_weights, d_cpts, d_fit;
  std::vector<char> is_center(static_cast<size_t>(n), 0);
  for (;;) {
    const int64_t mu = static_cast<int64_t>(res.centers.size());
    std::vector<double> cpts(static_cast<size_t>(mu) * 3), cvals(static_cast<size_t>(mu));
    for (int64_t i = 0; i < mu; i++) {
      const int64_t j = res.centers[i];
      for (int k = 0; k < 3; k++) cpts[3 * i + k] = points[3 * j + k];
      cvals[i] = values[j];
    }
    // incremental_fitter.hpp:83-91: solve once the polynomial is determined
    const bool solved = mu > 0 && mu >= l;
    if (solved) {
      Solver solver(model, cpts, opt, s);
      FitStats st;
      weights = solver.solve(cvals, tolerance, max_iter, &weights, &st);
      res.iterations += st.iterations;
    }
    res.rounds++;
    if (mu == n) break;

    // Residuals at the other points (ascending index order).
    std::vector<int64_t> c_idx;
    c_idx.reserve(static_cast<size_t>(n - mu));
    for (int64_t i = 0; i < n; i++)
      if (!is_center[i]) c_idx.push_back(i);
    const int64_t nc = static_cast<int64_t>(c_idx.size());
    std::vector<double> fit(static_cast<size_t>(nc), 0.0);
    if (solved) {
      std::vector<double> qpts(static_cast<size_t>(nc) * 3);
      for (int64_t i = 0; i < nc; i++)
        for (int k = 0; k < 3; k++) qpts[3 * i + k] = points[3 * c_idx[i] + k];
      d_centers.upload(cpts.data(), cpts.size(), s);
      d_weights.upload(weights.data(), weights.size(), s);
      d_cpts.upload(qpts.data(), qpts.size(), s);
      d_fit.resize(static_cast<size_t>(nc));
      res_eval.set_source_points(d_centers.get(), mu);
      res_eval.set_weights(d_weights.get());
      res_eval.set_target_points(d_cpts.get(), nc);
      res_eval.evaluate(d_fit.get());
      d_fit.download(fit.data(), fit.size(), s);
      PAM_HIP(hipStreamSynchronize(s));
    }
    // (no centres yet, or too few for the polynomial: the fit is zero)
    std::vector<double> r(static_cast<size_t>(nc));
    for (int64_t i = 0; i < nc; i++) r[i] = std::abs(fit[i] - values[c_idx[i]]);

    // Sort the other points by residual (incremental_fitter.hpp:114-118).
    std::vector<int64_t> order(static_cast<size_t>(nc));
    std::iota(order.begin(), order.end(), int64_t{0});
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return r[a] < r[b]; });
    int64_t n_need = 0;
    for (int64_t i = 0; i < nc; i++)
      if (r[i] >= tolerance) n_need++;
    if (n_need == 0) break;

    // Candidates, largest residual first, thinned together with the current
    // centres; those among the n_need largest that survive join.
    std::vector<int64_t> cand(res.centers);
    for (int64_t i = nc - 1; i >= 0; i--) cand.push_back(c_idx[order[i]]);
    std::vector<int64_t> kept = filter.filter(dist, cand);
    std::vector<char> keep(static_cast<size_t>(n), 0);
    for (int64_t i : kept) keep[i] = 1;
    const int64_t last_mu = mu;
    for (int64_t t = 0; t < n_need; t++) {
      const int64_t j = c_idx[order[nc - 1 - t]];
      if (keep[j]) {
        res.centers.push_back(j);
        is_center[j] = 1;
      }
    }
    const int64_t new_mu = static_cast<int64_t>(res.centers.size());
    // Warm start (incremental_fitter.hpp:173-177): the previous weights of the
    // previous centres and of the polynomial, zero for the new centres.
    std::vector<double> w0(static_cast<size_t>(new_mu + l), 0.0);
    for (int64_t i = 0; i < last_mu; i++) w0[i] = weights[i];
    for (int k = 0; k < l; k++) w0[new_mu + k] = weights[last_mu + k];
    weights = std::move(w0);
    dist = std::max(0.5 * dist, DistanceFilter::kMinDistance);
  }
  // Every residual below the tolerance before the polynomial was determined
  // (the zero function fits): this library keeps no interpolant without centres.
  if (res.centers.size() < static_cast<size_t>(std::max(l, 1)))
    throw_runtime("the data are fitted by zero: fewer centres than the polynomial needs");
  res.weights = std::move(weights);
  return res;
}

}  // namespace pam

namespace pam {

IncrementalHermiteResult fit_incrementally_hermite(const Model& model, const std::vector<double>& points,
                                                   const std::vector<double>& grad_points,
                                                   const std::vector<double>& values, double tolerance,
                                                   double grad_tolerance, int max_iter, const FitOptions& opt,
                                                   hipStream_t s) {
  // incremental_fitter.hpp:36-183 with gradient points: rounds add the points
  // and the gradient points whose residuals (a gradient point's largest
  // component) reach their tolerances, each set thinned by its own filter.
  const int64_t n = static_cast<int64_t>(points.size() / 3), ng = static_cast<int64_t>(grad_points.size() / 3);
  const int l = model.l();
  if (static_cast<int64_t>(values.size()) != n + 3 * ng)
    throw_invalid("values.rows() must be equal to points.rows() + 3 * grad_points.rows()");
  if (!(tolerance > 0.0) || !(grad_tolerance > 0.0)) throw_invalid("tolerance must be positive");
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const std::vector<double>* v : {&points, &grad_points})
    for (size_t i = 0; i + 2 < v->size(); i += 3)
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], (*v)[i + k]);
        hi[k] = std::max(hi[k], (*v)[i + k]);
      }
  double dist = 0.0;
  if (n + ng > 0)
    dist = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                     (hi[2] - lo[2]) * (hi[2] - lo[2]));
  dist = std::max(dist, DistanceFilter::kMinDistance);
  DistanceFilter filter(points), gfilter(grad_points);
  const bool special = model.poly_degree == 1 && n == 1 && ng >= 1;

  IncrementalHermiteResult res;
  std::vector<char> is_center(static_cast<size_t>(n), 0), is_gcenter(static_cast<size_t>(ng), 0);
  if (special) {  // incremental_fitter.hpp:58-61
    res.centers.push_back(0);
    is_center[0] = 1;
  }
  std::vector<double> weights(static_cast<size_t>(res.centers.size() + l), 0.0);
  FmmConfig ecfg;
  ecfg.precision = kFp64;
  ecfg.order = 10;
  HermiteEvaluator res_eval(model, ecfg, s);
  DevBuf<double> d_c, d_g, d_w, d_q, d_qg, d_fit;
  for (;;) {
    const int64_t mu = static_cast<int64_t>(res.centers.size()), sigma = static_cast<int64_t>(res.grad_centers.size());
    std::vector<double> cpts(static_cast<size_t>(mu) * 3), gpts(static_cast<size_t>(sigma) * 3),
        cvals(static_cast<size_t>(mu + 3 * sigma));
    for (int64_t i = 0; i < mu; i++) {
      const int64_t j = res.centers[i];
      for (int k = 0; k < 3; k++) cpts[3 * i + k] = points[3 * j + k];
      cvals[i] = values[j];
    }
    for (int64_t i = 0; i < sigma; i++) {
      const int64_t j = res.grad_centers[i];
      for (int k = 0; k < 3; k++) {
        gpts[3 * i + k] = grad_points[3 * j + k];
        cvals[mu + 3 * i + k] = values[n + 3 * j + k];
      }
    }
    // incremental_fitter.hpp:83-91
    const bool solved = (mu > 0 && mu >= l) || (model.poly_degree == 1 && mu == 1 && sigma >= 1);
    if (solved) {
      HermiteSolver solver(model, cpts, gpts, opt, s);
      FitStats st;
      weights = solver.solve(cvals, tolerance, grad_tolerance, max_iter, &weights, &st);
      res.iterations += st.iterations;
    }
    res.rounds++;
    if (mu == n && sigma == ng) break;

    // Residuals at the other points and gradient points (ascending index order).
    std::vector<int64_t> c_idx, g_idx;
    for (int64_t i = 0; i < n; i++)
      if (!is_center[i]) c_idx.push_back(i);
    for (int64_t i = 0; i < ng; i++)
      if (!is_gcenter[i]) g_idx.push_back(i);
    const int64_t nc = static_cast<int64_t>(c_idx.size()), ngc = static_cast<int64_t>(g_idx.size());
    std::vector<double> fit(static_cast<size_t>(nc + 3 * ngc), 0.0);
    if (solved) {
      std::vector<double> q(static_cast<size_t>(nc) * 3), qg(static_cast<size_t>(ngc) * 3);
      for (int64_t i = 0; i < nc; i++)
        for (int k = 0; k < 3; k++) q[3 * i + k] = points[3 * c_idx[i] + k];
      for (int64_t i = 0; i < ngc; i++)
        for (int k = 0; k < 3; k++) qg[3 * i + k] = grad_points[3 * g_idx[i] + k];
      d_c.upload(cpts.data(), cpts.size(), s);
      d_g.upload(gpts.data(), gpts.size(), s);
      d_w.upload(weights.data(), weights.size(), s);
      d_q.upload(q.data(), q.size(), s);
      d_qg.upload(qg.data(), qg.size(), s);
      d_fit.resize(static_cast<size_t>(nc + 3 * ngc));
      res_eval.set_sources(d_c.get(), mu, d_g.get(), sigma);
      res_eval.set_targets(d_q.get(), nc, d_qg.get(), ngc);
      res_eval.set_weights(d_w.get());
      res_eval.evaluate(d_fit.get());
      d_fit.download(fit.data(), fit.size(), s);
      PAM_HIP(hipStreamSynchronize(s));
    }
    std::vector<double> r(static_cast<size_t>(nc)), gr(static_cast<size_t>(ngc));
    for (int64_t i = 0; i < nc; i++) r[i] = std::abs(fit[i] - values[c_idx[i]]);
    for (int64_t i = 0; i < ngc; i++) {
      double m = 0.0;
      for (int k = 0; k < 3; k++) m = std::max(m, std::abs(fit[nc + 3 * i + k] - values[n + 3 * g_idx[i] + k]));
      gr[i] = m;
    }
    // Sort by residual (incremental_fitter.hpp:114-118) and count those at or
    // above the tolerance.
    auto sorted = [](const std::vector<double>& v) {
      std::vector<int64_t> o(v.size());
      std::iota(o.begin(), o.end(), int64_t{0});
      std::stable_sort(o.begin(), o.end(), [&](int64_t a, int64_t b) { return v[a] < v[b]; });
      return o;
    };
    const std::vector<int64_t> order = sorted(r), gorder = sorted(gr);
    int64_t n_need = 0, ng_need = 0;
    for (double v : r)
      if (v >= tolerance) n_need++;
    for (double v : gr)
      if (v >= grad_tolerance) ng_need++;
    if (n_need == 0 && ng_need == 0) break;

    const int64_t last_mu = mu, last_sigma = sigma;
    {
      std::vector<int64_t> cand(res.centers);
      for (int64_t i = nc - 1; i >= 0; i--) cand.push_back(c_idx[order[i]]);
      std::vector<char> keep(static_cast<size_t>(n), 0);
      for (int64_t i : filter.filter(dist, cand)) keep[i] = 1;
      for (int64_t t = 0; t < n_need; t++) {
        const int64_t j = c_idx[order[nc - 1 - t]];
        if (keep[j]) {
          res.centers.push_back(j);
          is_center[j] = 1;
        }
      }
    }
    {
      std::vector<int64_t> cand(res.grad_centers);
      for (int64_t i = ngc - 1; i >= 0; i--) cand.push_back(g_idx[gorder[i]]);
      std::vector<char> keep(static_cast<size_t>(ng), 0);
      for (int64_t i : gfilter.filter(dist, cand)) keep[i] = 1;
      for (int64_t t = 0; t < ng_need; t++) {
        const int64_t j = g_idx[gorder[ngc - 1 - t]];
        if (keep[j]) {
          res.grad_centers.push_back(j);
          is_gcenter[j] = 1;
        }
      }
    }
    // Warm start (incremental_fitter.hpp:173-177).
    const int64_t new_mu = static_cast<int64_t>(res.centers.size()),
                  new_sigma = static_cast<int64_t>(res.grad_centers.size());
    std::vector<double> w0(static_cast<size_t>(new_mu + 3 * new_sigma + l), 0.0);
    for (int64_t i = 0; i < last_mu; i++) w0[i] = weights[i];
    for (int64_t i = 0; i < 3 * last_sigma; i++) w0[new_mu + i] = weights[last_mu + i];
    for (int k = 0; k < l; k++) w0[new_mu + 3 * new_sigma + k] = weights[last_mu + 3 * last_sigma + k];
    weights = std::move(w0);
    dist = std::max(0.5 * dist, DistanceFilter::kMinDistance);
  }
  const bool determined = static_cast<int64_t>(res.centers.size()) >= std::max(l, 1) ||
                          (special && !res.grad_centers.empty());
  if (!determined) throw_runtime("the data are fitted by zero: fewer centres than the polynomial needs");
  res.weights = std::move(weights);
  return res;
}

}  // namespace pam
