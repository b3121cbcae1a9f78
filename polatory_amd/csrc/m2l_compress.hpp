// SVD-compressed M2L (the bbFMM construction): with U_r the leading left
// singular vectors of [K_v]_v (from the eigenvectors of G = sum_v K_v K_v^T),
//   M~ = U_r^T M,   L~ = sum_v C_v M~_src,   C_v = U_r^T K_v U_r,   L = U_r L~.
// All three products run on the MFMA gather-GEMM kernel (launch_m2l).
#pragma once

#include "fmm.hpp"

namespace pam {

template <class T>
struct CompressedM2L {
  int p = 0, r = 0, kpad_r = 0;
  M2LOperator c_op, ut_op, u_op;   // C_v (r x r), U^T (r x P^3), U (P^3 x r)
  DevBuf<T> c, ut, u;
  std::vector<double> sv;          // singular values (descending) of the stacked operator
};

// Builds the compressed operator of one level (cells of half width h); the rank
// is the smallest multiple of 16 with sigma_r / sigma_0 below rel_tol.
template <class T>
void build_m2l_compressed(int p, const RbfDesc& rbf, double h, double rel_tol, CompressedM2L<T>& out,
                          hipStream_t s);

// Symmetric eigendecomposition (cyclic Jacobi, round-robin pair order, OpenMP over
// rotations of a round): a is n x n row-major (destroyed); evals descending,
// vecs column k = eigenvector k (row-major n x n).
void jacobi_eigen(std::vector<double>& a, int n, std::vector<double>& evals, std::vector<double>& vecs);

}  // namespace pam
