// The gradient kernel of the FMM seam (fmm/gradient_kernel.hpp:46-62,
// make_fmm_gradient_evaluator, make_fmm_evaluator.cpp:78-107): sources carry
// 3-vector weights (gradient data) and the targets receive values,
//   out_i = sum_j w_j . (-A^T grad phi(A (t_i - s_j))) = -sum_k d/dx_k F_k(t_i),
// with F_k the scalar-weight sum of the k-th weight components. Evaluated as
// three gradient-transpose passes over one source tree, keeping the k-th
// component of pass k.
#pragma once

#include "fmm_evaluator.hpp"

namespace pam {

class DipoleSourceEvaluator {
 public:
  // ev: a source-target evaluator in gradient-kernel mode, owned by the caller.
  explicit DipoleSourceEvaluator(FmmEvaluator& ev);
  // w3: 3 n_src entries, x y z per source (input order).
  void set_weights(const double* w3, int64_t n, bool on_device);
  // out: n_trg values (input order of the targets).
  void evaluate(double* out, int64_t n, bool on_device);

 private:
  FmmEvaluator& ev_;
  DevBuf<double> w3_, wk_, g3_, tmp_;
  int64_t n_src_ = -1;
};

// The Hessian kernel (fmm/hessian_kernel.hpp:46-64, make_fmm_hessian_evaluator,
// make_fmm_evaluator.cpp:160-188): sources carry 3-vector weights and the
// targets receive 3-vectors,
//   out_i = -sum_j A^T H(A (t_i - s_j)) A w_j = -sum_k column k of the Hessian of F_k,
// as three Hessian-column passes over one source tree.
class DipoleHessianEvaluator {
 public:
  explicit DipoleHessianEvaluator(FmmEvaluator& ev);
  void set_weights(const double* w3, int64_t n, bool on_device);
  // out: 3 n_trg entries, x y z per target (input order).
  void evaluate(double* out, int64_t n, bool on_device);

 private:
  FmmEvaluator& ev_;
  DevBuf<double> w3_, wk_, g3_, tmp_;
  int64_t n_src_ = -1;
};

}  // namespace pam
