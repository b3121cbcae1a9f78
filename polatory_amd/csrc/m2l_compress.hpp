// SVD-compressed M2L (the bbFMM construction): with U_r the leading left
// singular vectors of [K_v]_v (from the eigenvectors of G = sum_v K_v K_v^T),
//   M~ = U_r^T M,   L~ = sum_v C_v M~_src,   C_v = U_r^T K_v U_r,   L = U_r L~.
// The compression and expansion run on the MFMA gather-GEMM kernel
// (launch_m2l_batch), the transfers on the pair-major one (launch_m2l_pairs).
#pragma once

#include "fmm.hpp"

namespace pam {

template <class T>
struct CompressedM2L {
  int p = 0, r = 0, kpad_r = 0;
  M2LOperator c_op, ut_op, u_op;   // C_v (r x r), U^T (r x P^3), U (P^3 x r)
  DevBuf<T> c;                     // transfers in the form's operand type
  DevBuf<double> ut, u;            // compression and expansion always in FP64
  std::vector<double> sv;          // estimated relative residual of each rank tried (64, 128, ...)
};

// Largest compressed rank.
constexpr int kMaxM2LRank = 256;

// Builds the compressed operator of one level (cells of half width h): U_r is
// an orthonormal basis of the range of [K_v]_v from randomized subspace
// iteration, and r the smallest multiple of 64 (at most kMaxM2LRank) whose
// estimated relative residual |(I - U_r U_r^T) [K_v]_v| is within rel_tol.
template <class T>
void build_m2l_compressed(int p, const RbfDesc& rbf, double h, double rel_tol, CompressedM2L<T>& out,
                          hipStream_t s);

}  // namespace pam
