// Interpolation with gradient data (Hermite-Birkhoff): value centres p_i with
// weights w_i, gradient centres q_k with 3-vector weights g_k and polynomial
// coefficients c (interpolant.hpp:77-113, interpolation/operator.hpp:30-120).
// Vectors of such a fit are flat: [w (mu); g (3 sigma, x y z per gradient
// centre); c (l)], and the data [values at p (mu); gradients at q (3 sigma)].
#pragma once

#include <memory>
#include <vector>

#include "fmm_dipole.hpp"
#include "interpolation.hpp"
#include "krylov.hpp"
#include "linalg.hpp"

namespace pam {

// P of a fit with gradient data, l columns of length mu + 3 sigma
// (p[k * ld + r], ld = mu + 3 sigma): the monomials at the value centres,
// their gradients at the gradient centres (MonomialBasis::evaluate(points,
// grad_points), monomial_basis.hpp:198-247).
void dev_poly_matrix_hermite(int degree, const double* d_pts, int64_t mu, const double* d_gpts, int64_t sigma,
                             double* p, hipStream_t s);

// Values at target points and gradients at target gradient points of
//   f(x) = sum_i w_i phi(x - p_i) - sum_k g_k . grad phi(x - q_k) + P(x) c,
// summed over the model's rbfs (interpolation/evaluator.hpp:50-144 with
// gradient points; operator.hpp:56-82 for the four kernel blocks):
//   values from value weights      make_fmm_evaluator,
//   values from gradient weights   make_fmm_gradient_evaluator,
//   gradients from value weights   make_fmm_gradient_transpose_evaluator,
//   gradients from gradient weights make_fmm_hessian_evaluator.
// Every FMM evaluation uses one fixed configuration (direct sums where the
// configuration's mode or the point counts call for them).
class HermiteEvaluator {
 public:
  HermiteEvaluator(const Model& model, const FmmConfig& cfg, hipStream_t s);
  // Device n x 3 point arrays; they must stay valid while the evaluator is used.
  void set_sources(const double* d_pts, int64_t mu, const double* d_gpts, int64_t sigma);
  void set_targets(const double* d_pts, int64_t nt, const double* d_gpts, int64_t st);
  void set_weights(const double* d_w);  // mu + 3 sigma + l entries (device)
  void evaluate(double* d_out);         // nt + 3 st entries (device)
  int64_t n_sources() const { return mu_ + 3 * sigma_; }
  int64_t n_targets() const { return nt_ + 3 * st_; }

 private:
  struct Terms {
    std::unique_ptr<FmmEvaluator> vv, vg, gv, gg;
    std::unique_ptr<DipoleSourceEvaluator> vg3;
    std::unique_ptr<DipoleHessianEvaluator> gg3;
  };
  void build();
  Model model_;
  FmmConfig cfg_;
  hipStream_t s_;
  std::vector<Terms> terms_;
  const double *spts_ = nullptr, *sgpts_ = nullptr, *tpts_ = nullptr, *tgpts_ = nullptr;
  int64_t mu_ = 0, sigma_ = 0, nt_ = 0, st_ = 0;
  bool built_ = false;
  DevBuf<double> coeffs_, tmp_;
  const double* d_w_ = nullptr;
};

// The operator of the fit (interpolation/operator.hpp:56-82):
//   y = [A_vv w + nugget w + A_vg g + P_v c; A_gv w + A_gg g + P_g c; P^T [w; g]].
class HermiteOperator : public LinearOperator {
 public:
  HermiteOperator(const Model& model, const double* d_pts, int64_t mu, const double* d_gpts, int64_t sigma,
                  const FmmConfig& cfg, hipStream_t s);
  void apply(const double* d_x, double* d_y) override;
  int64_t size() const override { return mu_ + 3 * sigma_ + l_; }
  // P (see dev_poly_matrix_hermite), l columns of length mu + 3 sigma.
  const double* poly_matrix() const { return p_.get(); }

 private:
  Model model_;
  hipStream_t s_;
  int64_t mu_, sigma_;
  int l_;
  HermiteEvaluator ev_;
  DevBuf<double> p_;
  Workspace1 ws_;
};

}  // namespace pam
