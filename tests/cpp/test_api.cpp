// Exercises the C++ API mirror (polatory/polatory.hpp) the way polatory's own
// tests do: test/interpolation/test_fitter.cpp (fit, residual <= tolerance) and
// test/interpolation/test_evaluator.cpp (FMM vs direct within accuracy).
// Exit code 0 on success; prints one line per check.
#include <polatory/polatory.hpp>

#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>

using polatory::Index;
using polatory::Interpolant;
using polatory::Model;
using polatory::VecX;
using polatory::geometry::Bbox3;
using polatory::geometry::Points3;
using polatory::rbf::Biharmonic3D;
using polatory::rbf::Triharmonic3D;

static double value_at(double x, double y, double z) { return std::sin(3 * x) * std::cos(2 * y) + 0.5 * z * z; }

int main() {
  std::mt19937_64 gen(42);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  const Index n = 5000;
  Points3 pts(n, 3);
  VecX values(n);
  for (Index i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) pts(i, k) = u(gen);
    values(i) = value_at(pts(i, 0), pts(i, 1), pts(i, 2));
  }
  Model<3> model(Biharmonic3D<3>(), 1);
  Interpolant<3> interpolant(model);
  const double tol = 1e-4;
  interpolant.fit(pts, values, tol);
  VecX w = interpolant.weights();
  if (w.rows() != n + 4) return 1;

  // FP64 direct residual at the data sites (DirectEvaluator semantics, bh3 = -r).
  double res = 0.0;
  for (Index i = 0; i < n; i++) {
    double y = w(n) + w(n + 1) * pts(i, 0) + w(n + 2) * pts(i, 1) + w(n + 3) * pts(i, 2);
    for (Index j = 0; j < n; j++) {
      double dx = pts(i, 0) - pts(j, 0), dy = pts(i, 1) - pts(j, 1), dz = pts(i, 2) - pts(j, 2);
      y += -std::sqrt(dx * dx + dy * dy + dz * dz) * w(j);
    }
    res = std::max(res, std::abs(y - values(i)));
  }
  std::printf("fit residual %.3e (tolerance %.1e)\n", res, tol);
  if (!(res <= tol)) return 2;

  // FMM seam: evaluator vs direct sum on random targets
  const Index m = 2000;
  Points3 trg(m, 3);
  for (Index i = 0; i < m; i++)
    for (int k = 0; k < 3; k++) trg(i, k) = u(gen);
  auto ev = polatory::fmm::make_fmm_evaluator<3>(Biharmonic3D<3>(), Bbox3::from_points(pts).convex_hull(Bbox3::from_points(trg)));
  ev->set_source_points(pts);
  VecX wr(n);
  for (Index i = 0; i < n; i++) wr(i) = w(i);
  ev->set_weights(wr);
  ev->set_target_points(trg);
  VecX y = ev->evaluate();
  double err = 0.0, scale = 0.0;
  for (Index i = 0; i < m; i++) {
    double d = 0.0;
    for (Index j = 0; j < n; j++) {
      double dx = trg(i, 0) - pts(j, 0), dy = trg(i, 1) - pts(j, 1), dz = trg(i, 2) - pts(j, 2);
      d += -std::sqrt(dx * dx + dy * dy + dz * dz) * wr(j);
    }
    err = std::max(err, std::abs(y(i) - d));
    scale = std::max(scale, std::abs(d));
  }
  std::printf("evaluator rel err %.3e\n", err / scale);
  if (!(err <= 1e-8 * scale + 1e-12)) return 3;

  // interpolant evaluation reproduces the data
  VecX back = interpolant.evaluate(pts);
  double e2 = 0.0;
  for (Index i = 0; i < n; i++) e2 = std::max(e2, std::abs(back(i) - values(i)));
  std::printf("evaluate at data sites: max |f - s| %.3e\n", e2);
  if (!(e2 <= 10 * tol)) return 4;

  // Interpolant::save / load (interpolant.hpp:218): the loaded interpolant
  // evaluates to the same bits
  {
    const std::string file = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") +
                             "/pam_test_api_" + std::to_string(static_cast<long>(getpid())) + ".bin";
    interpolant.save(file);
    Interpolant<3> loaded = Interpolant<3>::load(file);
    std::remove(file.c_str());
    if (!loaded.fitted() || loaded.model().poly_degree() != 1) return 6;
    VecX a = interpolant.evaluate(trg), b = loaded.evaluate(trg);
    for (Index i = 0; i < m; i++)
      if (a(i) != b(i)) return 7;
    std::printf("save/load: identical evaluation\n");
  }

  // values then gradients (interpolant.hpp:58-67); the gradient of the fit
  // against central differences of its own values
  {
    Points3 q(3, 3), gq(2, 3);
    for (Index i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) q(i, k) = 0.3 + 0.1 * i + 0.05 * k;
    for (Index i = 0; i < 2; i++)
      for (int k = 0; k < 3; k++) gq(i, k) = 0.4 + 0.2 * i - 0.03 * k;
    // few points: the evaluators take the FP64 direct sums
    VecX vg = interpolant.evaluate(q, gq);
    if (vg.rows() != 3 + 6) return 8;
    VecX v = interpolant.evaluate(q);
    for (Index i = 0; i < 3; i++)
      if (vg(i) != v(i)) return 9;
    const double h = 1e-5;
    double gerr = 0.0;
    for (Index i = 0; i < 2; i++)
      for (int k = 0; k < 3; k++) {
        Points3 pp(2, 3);
        for (int c = 0; c < 3; c++) pp(0, c) = pp(1, c) = gq(i, c);
        pp(0, k) += h;
        pp(1, k) -= h;
        VecX f = interpolant.evaluate(pp);
        gerr = std::max(gerr, std::abs(vg(3 + 3 * i + k) - (f(0) - f(1)) / (2 * h)));
      }
    std::printf("gradients vs central differences: %.3e\n", gerr);
    if (!(gerr < 1e-4)) return 10;
  }

  // isosurface::RbfFieldFunction (rbf_field_function.hpp:21-27): the field on
  // a lattice that reaches past the centres' box, against the FP64 direct sum
  {
    polatory::isosurface::RbfFieldFunction field(interpolant, 1e-7);
    field.set_evaluation_bbox(Bbox3({-0.25, -0.25, -0.25}, {1.25, 1.25, 1.25}));
    const Index g = 16;
    Points3 lat(g * g * g, 3);
    for (Index i = 0; i < g * g * g; i++) {
      lat(i, 0) = -0.25 + 1.5 * (i % g) / (g - 1);
      lat(i, 1) = -0.25 + 1.5 * ((i / g) % g) / (g - 1);
      lat(i, 2) = -0.25 + 1.5 * (i / (g * g)) / (g - 1);
    }
    VecX f = field(lat);
    VecX f2 = field(lat);
    double ferr = 0.0, fscale = 0.0;
    for (Index i = 0; i < lat.rows(); i++) {
      if (f(i) != f2(i)) return 11;
      double d = w(n) + w(n + 1) * lat(i, 0) + w(n + 2) * lat(i, 1) + w(n + 3) * lat(i, 2);
      for (Index j = 0; j < n; j++) {
        double dx = lat(i, 0) - pts(j, 0), dy = lat(i, 1) - pts(j, 1), dz = lat(i, 2) - pts(j, 2);
        d += -std::sqrt(dx * dx + dy * dy + dz * dz) * w(j);
      }
      ferr = std::max(ferr, std::abs(f(i) - d));
      fscale = std::max(fscale, std::abs(d));
    }
    std::printf("RbfFieldFunction on a %ld^3 lattice: max error %.3e (scale %.3e)\n", static_cast<long>(g), ferr,
                fscale);
    if (!(ferr <= 1e-6 * std::max(1.0, fscale))) return 12;
  }

  // make_fmm_symmetric_evaluator and make_fmm_hessian_symmetric_evaluator
  // (fmm_symmetric_evaluator.hpp:80-86): the same bits as the source-target
  // evaluators given the points twice (Hessians), and values within the
  // mixed form's error (the value evaluator has its own one-tree path)
  {
    const Bbox3 box = Bbox3::from_points(pts);
    VecX w3(3 * n);
    for (Index i = 0; i < 3 * n; i++) w3(i) = std::sin(0.37 * static_cast<double>(i));
    auto sym = polatory::fmm::make_fmm_symmetric_evaluator<3>(Biharmonic3D<3>(), box);
    auto st = polatory::fmm::make_fmm_evaluator<3>(Biharmonic3D<3>(), box);
    auto hsym = polatory::fmm::make_fmm_hessian_symmetric_evaluator<3>(Biharmonic3D<3>(), box);
    auto hst = polatory::fmm::make_fmm_hessian_evaluator<3>(Biharmonic3D<3>(), box);
    VecX wr(n);
    for (Index i = 0; i < n; i++) wr(i) = w(i);
    sym->set_points(pts);
    sym->set_weights(wr);
    st->set_source_points(pts);
    st->set_weights(wr);
    st->set_target_points(pts);
    hsym->set_points(pts);
    hsym->set_weights(w3);
    hst->set_source_points(pts);
    hst->set_weights(w3);
    hst->set_target_points(pts);
    VecX a = sym->evaluate(), b = st->evaluate(), ha = hsym->evaluate(), hb = hst->evaluate();
    if (a.rows() != n || ha.rows() != 3 * n || hb.rows() != 3 * n) return 13;
    double dv = 0.0, dh = 0.0, sv = 0.0, sh = 0.0;
    for (Index i = 0; i < n; i++) {
      dv = std::max(dv, std::abs(a(i) - b(i)));
      sv = std::max(sv, std::abs(b(i)));
    }
    for (Index i = 0; i < 3 * n; i++) {
      dh = std::max(dh, std::abs(ha(i) - hb(i)));
      sh = std::max(sh, std::abs(hb(i)));
    }
    std::printf("symmetric vs source-target: values %.3e, Hessians %.3e (relative)\n", dv / sv, dh / sh);
    if (!(dh == 0.0) || !(dv <= 1e-5 * sv)) return 14;
  }

  // fit with gradient data (interpolant.hpp:83-113): f = x + 2 y - z + x y
  {
    const Index mu = 2000, sigma = 1000;
    Points3 p(mu, 3), g(sigma, 3);
    VecX rhs(mu + 3 * sigma);
    for (Index i = 0; i < mu; i++) {
      for (int k = 0; k < 3; k++) p(i, k) = u(gen);
      rhs(i) = p(i, 0) + 2.0 * p(i, 1) - p(i, 2) + p(i, 0) * p(i, 1);
    }
    for (Index i = 0; i < sigma; i++) {
      for (int k = 0; k < 3; k++) g(i, k) = u(gen);
      rhs(mu + 3 * i) = 1.0 + g(i, 1);
      rhs(mu + 3 * i + 1) = 2.0 + g(i, 0);
      rhs(mu + 3 * i + 2) = -1.0;
    }
    Model<3> hmodel(Triharmonic3D<3>(), 1);
    Interpolant<3> hip(hmodel);
    hip.fit(p, g, rhs, 1e-4, 1e-4);
    if (hip.weights().rows() != mu + 3 * sigma + 4) return 15;
    if (hip.grad_centers().rows() != sigma) return 16;
    VecX f = hip.evaluate(p, g);
    double e = 0.0;
    for (Index i = 0; i < mu + 3 * sigma; i++) e = std::max(e, std::abs(f(i) - rhs(i)));
    std::printf("fit with gradient data: max |f - data| %.3e\n", e);
    if (!(e <= 1e-4 + 1e-6)) return 17;
  }

  // error contract: invalid arguments throw std::invalid_argument
  try {
    interpolant.fit(pts, values, 0.0);
    return 5;
  } catch (const std::invalid_argument&) {
  }
  std::printf("ok\n");
  return 0;
}
