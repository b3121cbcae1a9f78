// Device FGMRES (see krylov.hpp).
#include <cmath>

#include "krylov.hpp"

namespace pam {

Fgmres::Fgmres(LinearOperator& op, const double* d_rhs, int64_t m, int max_iter, hipStream_t s)
    : op_(op), m_(m), max_iter_(max_iter), s_(s) {
  if (max_iter < 0) throw_invalid("max_iter must be nonnegative");
  rhs_.resize(m);
  dev_copy(d_rhs, rhs_.get(), m, s);
  x0_.resize(m);
  dev_fill(0.0, x0_.get(), m, s);
  rhs_norm_ = std::sqrt(dev_dot(rhs_.get(), rhs_.get(), m, ws_, s));
}

void Fgmres::set_initial_solution(const double* d_x0) {
  dev_copy(d_x0, x0_.get(), m_, s_);
  x0_zero_ = dev_max_abs_diff(x0_.get(), nullptr, m_, ws_, s_) == 0.0;
}

void Fgmres::setup() {
  // gmres_base.cpp:35-47
  int k = max_iter_;
  c_.assign(k + 1, 0.0);
  sn_.assign(k + 1, 0.0);
  g_.assign(k + 2, 0.0);
  r_.assign(static_cast<size_t>(k + 1) * (k + 1), 0.0);
  h_.assign(k + 2, 0.0);
  v_.resize(static_cast<size_t>(k + 1) * m_);
  z_.resize(static_cast<size_t>(k > 0 ? k : 1) * m_);
  coef_.resize(static_cast<size_t>(k + 2));
  double* v0 = v_.get();
  if (x0_zero_) {
    dev_copy(rhs_.get(), v0, m_, s_);
  } else {
    op_.apply(x0_.get(), v0);
    dev_scal(-1.0, v0, m_, s_);
    dev_axpy(1.0, rhs_.get(), v0, m_, s_);
  }
  double nrm = std::sqrt(dev_dot(v0, v0, m_, ws_, s_));
  g_[0] = nrm;
  if (nrm > 0.0) dev_scal(1.0 / nrm, v0, m_, s_);
  iter_ = 0;
}

void Fgmres::iterate_process() {
  if (iter_ == max_iter_) return;
  const int j = iter_;
  const int k = max_iter_;
  double* vj = v_.get() + static_cast<int64_t>(j) * m_;
  double* zj = z_.get() + static_cast<int64_t>(j) * m_;
  double* w = v_.get() + static_cast<int64_t>(j + 1) * m_;
  // Arnoldi with a flexible right preconditioner
  if (step_) {
    step_(vj, zj, w);
  } else {
    if (pc_) pc_->apply(vj, zj); else dev_copy(vj, zj, m_, s_);
    op_.apply(zj, w);
  }
  // classical Gram-Schmidt: all projections against the unmodified w
  dev_multi_dot(v_.get(), m_, j + 1, w, m_, h_.data(), ws_, s_);
  PAM_HIP(hipMemcpyAsync(coef_.get(), h_.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s_));
  dev_gemv_update(w, v_.get(), m_, j + 1, coef_.get(), m_, -1.0, s_);
  double hn = std::sqrt(dev_dot(w, w, m_, ws_, s_));
  if (hn > 0.0) dev_scal(1.0 / hn, w, m_, s_);
  for (int i = 0; i <= j; i++) r_[i * (k + 1) + j] = h_[i];
  r_[(j + 1) * (k + 1) + j] = hn;
  // Givens rotations (gmres.cpp:32-49)
  for (int i = 0; i < j; i++) {
    double x = r_[i * (k + 1) + j], y = r_[(i + 1) * (k + 1) + j];
    r_[i * (k + 1) + j] = c_[i] * x + sn_[i] * y;
    r_[(i + 1) * (k + 1) + j] = -sn_[i] * x + c_[i] * y;
  }
  double x = r_[j * (k + 1) + j], y = r_[(j + 1) * (k + 1) + j];
  double den = std::hypot(x, y);
  c_[j] = x / den;
  sn_[j] = y / den;
  r_[j * (k + 1) + j] = c_[j] * x + sn_[j] * y;
  g_[j + 1] = -sn_[j] * g_[j];
  g_[j] = c_[j] * g_[j];
  iter_++;
}

void Fgmres::solution_vector(double* d_x) {
  const int k = max_iter_;
  std::vector<double> y(iter_ > 0 ? iter_ : 1, 0.0);
  for (int j = iter_ - 1; j >= 0; j--) {
    double acc = g_[j];
    for (int i = j + 1; i < iter_; i++) acc -= r_[j * (k + 1) + i] * y[i];
    y[j] = acc / r_[j * (k + 1) + j];
  }
  dev_copy(x0_.get(), d_x, m_, s_);
  if (iter_ > 0) {
    PAM_HIP(hipMemcpyAsync(coef_.get(), y.data(), sizeof(double) * iter_, hipMemcpyHostToDevice, s_));
    dev_gemv_update(d_x, z_.get(), m_, iter_, coef_.get(), m_, 1.0, s_);
  }
  PAM_HIP(hipStreamSynchronize(s_));  // y is a host temporary
}

}  // namespace pam
