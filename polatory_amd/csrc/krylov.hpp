// Krylov solver on the device: right-preconditioned FGMRES without restart.
// Restates polatory's krylov::Fgmres (src/krylov/gmres.cpp:9-50,
// src/krylov/fgmres.cpp:8-26, src/krylov/gmres_base.cpp:7-87): classical
// Gram-Schmidt Arnoldi (one pass), Givens QR of the Hessenberg matrix.
// The Krylov bases V and Z live in HBM (FP64); the small Hessenberg/Givens
// state lives on the host.
#pragma once

#include <functional>
#include <vector>

#include "linalg.hpp"

namespace pam {

// krylov::LinearOperator (include/polatory/krylov/linear_operator.hpp:7-22):
// y = Op(x) for device vectors, ordered on the operator's stream.
class LinearOperator {
 public:
  virtual ~LinearOperator() = default;
  virtual void apply(const double* d_x, double* d_y) = 0;
  virtual int64_t size() const = 0;
};

class Fgmres {
 public:
  Fgmres(LinearOperator& op, const double* d_rhs, int64_t m, int max_iter, hipStream_t s);
  void set_initial_solution(const double* d_x0);
  void set_right_preconditioner(LinearOperator* pc) { pc_ = pc; }
  // Replaces "z_j = M v_j; w = A z_j" of an Arnoldi step (e.g. by a replay of
  // a captured graph); nullptr restores the operator calls.
  using Step = std::function<void(const double* vj, double* zj, double* w)>;
  void set_step(Step f) { step_ = std::move(f); }
  void setup();
  void iterate_process();
  // x = x0 + sum_i y_i z_i with R y = g (back substitution).
  void solution_vector(double* d_x);
  double absolute_residual() const { return std::abs(g_[iter_]); }
  double relative_residual() const { return std::abs(g_[iter_]) / rhs_norm_; }
  double rhs_norm() const { return rhs_norm_; }
  int iteration_count() const { return iter_; }
  int max_iterations() const { return max_iter_; }

 private:
  LinearOperator& op_;
  LinearOperator* pc_ = nullptr;
  Step step_;
  int64_t m_;
  int max_iter_;
  hipStream_t s_;
  int iter_ = 0;
  bool x0_zero_ = true;
  double rhs_norm_ = 0.0;
  DevBuf<double> rhs_, x0_, v_, z_, coef_;
  std::vector<double> r_, c_, sn_, g_, h_;
  Workspace1 ws_;
};

}  // namespace pam
