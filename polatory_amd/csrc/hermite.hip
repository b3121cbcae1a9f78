// Evaluator and operator of fits with gradient data (see hermite.hpp).
#include "hermite.hpp"

#include "device_api.hpp"

namespace pam {

namespace {

constexpr int kT = 256;

// Column k of P: monomial k at value centre i (row i), its gradient at
// gradient centre j (rows mu + 3 j + c). Monomials 1, x, y, z, x^2, xy, xz,
// y^2, yz, z^2 (monomial_basis.hpp:198-247).
__global__ void poly_matrix_hermite_kernel(int degree, int l, const double* __restrict__ pts, int64_t mu,
                                           const double* __restrict__ gpts, int64_t sigma, double* __restrict__ p) {
  const int64_t ld = mu + 3 * sigma;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
  if (r >= mu + sigma) return;
  if (r < mu) {
    const double x = pts[3 * r], y = pts[3 * r + 1], z = pts[3 * r + 2];
    const double m[10] = {1.0, x, y, z, x * x, x * y, x * z, y * y, y * z, z * z};
    for (int k = 0; k < l; k++) p[k * ld + r] = m[k];
    (void)degree;
    return;
  }
  const int64_t j = r - mu;
  const double x = gpts[3 * j], y = gpts[3 * j + 1], z = gpts[3 * j + 2];
  const double gx[10] = {0.0, 1.0, 0.0, 0.0, 2.0 * x, y, z, 0.0, 0.0, 0.0};
  const double gy[10] = {0.0, 0.0, 1.0, 0.0, 0.0, x, 0.0, 2.0 * y, z, 0.0};
  const double gz[10] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, x, 0.0, y, 2.0 * z};
  const int64_t row = mu + 3 * j;
  for (int k = 0; k < l; k++) {
    p[k * ld + row] = gx[k];
    p[k * ld + row + 1] = gy[k];
    p[k * ld + row + 2] = gz[k];
  }
}

}  // namespace

void dev_poly_matrix_hermite(int degree, const double* d_pts, int64_t mu, const double* d_gpts, int64_t sigma,
                             double* p, hipStream_t s) {
  const int l = poly_basis_size(degree);
  if (l == 0 || mu + sigma == 0) return;
  poly_matrix_hermite_kernel<<<ceil_div(mu + sigma, kT), kT, 0, s>>>(degree, l, d_pts, mu, d_gpts, sigma, p);
  PAM_LAUNCH_CHECK();
}

// ---- HermiteEvaluator ----

HermiteEvaluator::HermiteEvaluator(const Model& model, const FmmConfig& cfg, hipStream_t s)
    : model_(model), cfg_(cfg), s_(s) {
  model_.validate();
  coeffs_.resize(10);
}

void HermiteEvaluator::set_sources(const double* d_pts, int64_t mu, const double* d_gpts, int64_t sigma) {
  if (mu < 0 || sigma < 0) throw_invalid("negative point count");
  spts_ = d_pts;
  sgpts_ = d_gpts;
  mu_ = mu;
  sigma_ = sigma;
  built_ = false;
}

void HermiteEvaluator::set_targets(const double* d_pts, int64_t nt, const double* d_gpts, int64_t st) {
  if (nt < 0 || st < 0) throw_invalid("negative point count");
  tpts_ = d_pts;
  tgpts_ = d_gpts;
  nt_ = nt;
  st_ = st;
  built_ = false;
}

void HermiteEvaluator::build() {
  terms_.clear();
  for (const auto& r : model_.rbfs) {
    Terms t;
    auto make = [&](const double* src, int64_t ns, const double* trg, int64_t ntg, bool grad) {
      auto e = std::make_unique<FmmEvaluator>(r, nullptr, nullptr, false, s_);
      e->set_gradient_kernel(grad);
      e->set_config(cfg_);
      e->set_source_points(src, ns, true);
      e->set_target_points(trg, ntg, true);
      return e;
    };
    if (mu_ > 0 && nt_ > 0) t.vv = make(spts_, mu_, tpts_, nt_, false);
    if (sigma_ > 0 && nt_ > 0) {
      t.vg = make(sgpts_, sigma_, tpts_, nt_, true);
      t.vg3 = std::make_unique<DipoleSourceEvaluator>(*t.vg);
    }
    if (mu_ > 0 && st_ > 0) t.gv = make(spts_, mu_, tgpts_, st_, true);
    if (sigma_ > 0 && st_ > 0) {
      t.gg = make(sgpts_, sigma_, tgpts_, st_, true);
      t.gg3 = std::make_unique<DipoleHessianEvaluator>(*t.gg);
    }
    terms_.push_back(std::move(t));
  }
  built_ = true;
}

void HermiteEvaluator::set_weights(const double* d_w) {
  if (!built_) build();
  for (auto& t : terms_) {
    if (t.vv) t.vv->set_weights(d_w, mu_, true);
    if (t.gv) t.gv->set_weights(d_w, mu_, true);
    if (t.vg3) t.vg3->set_weights(d_w + mu_, 3 * sigma_, true);
    if (t.gg3) t.gg3->set_weights(d_w + mu_, 3 * sigma_, true);
  }
  const int l = model_.l();
  if (l > 0) dev_copy(d_w + mu_ + 3 * sigma_, coeffs_.get(), l, s_);
  d_w_ = d_w;
}

void HermiteEvaluator::evaluate(double* d_out) {
  if (!d_w_) throw_runtime("weights are not set");
  const int64_t n = nt_ + 3 * st_;
  dev_fill(0.0, d_out, n, s_);
  tmp_.resize(static_cast<size_t>(std::max<int64_t>(nt_, 3 * st_)));
  for (auto& t : terms_) {
    if (t.vv) {
      t.vv->evaluate(tmp_.get(), nt_, true);
      dev_axpy(1.0, tmp_.get(), d_out, nt_, s_);
    }
    if (t.vg3) {
      t.vg3->evaluate(tmp_.get(), nt_, true);
      dev_axpy(1.0, tmp_.get(), d_out, nt_, s_);
    }
    if (t.gv) {
      t.gv->evaluate(tmp_.get(), st_, true);
      dev_axpy(1.0, tmp_.get(), d_out + nt_, 3 * st_, s_);
    }
    if (t.gg3) {
      t.gg3->evaluate(tmp_.get(), st_, true);
      dev_axpy(1.0, tmp_.get(), d_out + nt_, 3 * st_, s_);
    }
  }
  if (model_.l() > 0) {
    if (nt_ > 0) launch_poly_eval(model_.poly_degree, tpts_, nt_, coeffs_.get(), d_out, s_);
    if (st_ > 0) launch_poly_grad(model_.poly_degree, tgpts_, st_, coeffs_.get(), d_out + nt_, s_);
  }
}

// ---- HermiteOperator ----

HermiteOperator::HermiteOperator(const Model& model, const double* d_pts, int64_t mu, const double* d_gpts,
                                 int64_t sigma, const FmmConfig& cfg, hipStream_t s)
    : model_(model), s_(s), mu_(mu), sigma_(sigma), l_(model.l()), ev_(model, cfg, s) {
  ev_.set_sources(d_pts, mu, d_gpts, sigma);
  ev_.set_targets(d_pts, mu, d_gpts, sigma);
  if (l_ > 0) {
    p_.resize(static_cast<size_t>(mu + 3 * sigma) * l_);
    dev_poly_matrix_hermite(model_.poly_degree, d_pts, mu, d_gpts, sigma, p_.get(), s_);
  }
}

void HermiteOperator::apply(const double* d_x, double* d_y) {
  const int64_t n = mu_ + 3 * sigma_;
  ev_.set_weights(d_x);
  ev_.evaluate(d_y);
  if (model_.nugget != 0.0) dev_axpy(model_.nugget, d_x, d_y, mu_, s_);
  if (l_ > 0) dev_multi_dot_device(p_.get(), n, l_, d_x, n, d_y + n, ws_, s_);
}

}  // namespace pam
