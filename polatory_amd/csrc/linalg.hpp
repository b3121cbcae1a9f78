// Device BLAS-1/2 and polynomial-block helpers for the fit path (FP64).
// Reductions are two-stage with a fixed order, so results are reproducible.
#pragma once

#include <vector>

#include "common.hpp"

namespace pam {

struct Workspace1 {
  DevBuf<double> partial, small;
  std::vector<double> host;
};

// Returns x . y (synchronizes the stream).
double dev_dot(const double* x, const double* y, int64_t n, Workspace1& ws, hipStream_t s);
// h[i] = V_i . w for i < nvec, V_i = V + i * ldv (synchronizes, h is host memory).
void dev_multi_dot(const double* v, int64_t ldv, int nvec, const double* w, int64_t n, double* h,
                   Workspace1& ws, hipStream_t s);
// d_h[i] = V_i . w for i < nvec into device memory (no synchronization).
void dev_multi_dot_device(const double* v, int64_t ldv, int nvec, const double* w, int64_t n,
                          double* d_h, Workspace1& ws, hipStream_t s);
// w += alpha * sum_i coef[i] V_i (coef is device memory, nvec entries).
void dev_gemv_update(double* w, const double* v, int64_t ldv, int nvec, const double* d_coef,
                     int64_t n, double alpha, hipStream_t s);
void dev_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s);
void dev_scal(double a, double* x, int64_t n, hipStream_t s);
void dev_fill(double v, double* x, int64_t n, hipStream_t s);
void dev_copy(const double* x, double* y, int64_t n, hipStream_t s);
// dst[i] = src[idx[i]]
void dev_gather(const double* src, const int32_t* idx, int64_t n, double* dst, hipStream_t s);
// dst[idx[i]] += alpha * src[i]
void dev_scatter_add(const double* src, const int32_t* idx, int64_t n, double alpha, double* dst,
                     hipStream_t s);
// dst[idx[i]] = src[i]
void dev_scatter(const double* src, const int32_t* idx, int64_t n, double* dst, hipStream_t s);
// gather of 3-vectors: dst[i] = src[idx[i]] for row-major n x 3 points
void dev_gather_points(const double* src, const int32_t* idx, int64_t n, double* dst, hipStream_t s);
// max_i |a_i - b_i| (b may be null: max |a_i|); synchronizes.
double dev_max_abs_diff(const double* a, const double* b, int64_t n, Workspace1& ws, hipStream_t s);
// max_i |a_i - b_i| over the given indices of a (a[idx[i]] vs b[i]).
double dev_max_abs_diff_idx(const double* a, const int32_t* idx, const double* b, int64_t n,
                            Workspace1& ws, hipStream_t s);

// Monomial basis [1, x, y, z, x^2, xy, xz, y^2, yz, z^2] truncated to degree
// (polynomial/monomial_basis.hpp:198-247).
int poly_basis_size(int degree);
// P as l column vectors of length n: p[k * n + i] = m_k(x_i).
void dev_poly_matrix(int degree, const double* pts, int64_t n, double* p, hipStream_t s);
// out[k] = sum_i m_k(x_i) w_i (device out, l entries).
void dev_poly_transpose_apply(int degree, const double* pts, int64_t n, const double* w,
                              double* d_out, Workspace1& ws, hipStream_t s);
// Modified Gram-Schmidt of l columns of length n, in place
// (common/orthonormalize.hpp:8-18).
void dev_orthonormalize_cols(double* p, int64_t n, int l, Workspace1& ws, hipStream_t s);

}  // namespace pam
