// Host-side launchers of the device kernels (one translation unit per family).
#pragma once

#include <vector>

#include "common.hpp"

namespace pam {

// ---- util.hip ---------------------------------------------------------------
// out[i] = A * p_i for row-major n x 3 points, no contraction (bit-exact with a
// CPU evaluation of ((A0 x + A1 y) + A2 z)).
void launch_transform_points(const double* aniso, const double* in, int64_t n, double* out,
                             hipStream_t s);
// Componentwise min/max of n x 3 points; result lands in d_out[6] = {min xyz, max xyz}.
void launch_bbox(const double* pts, int64_t n, double* d_out, DevBuf<double>& scratch,
                 hipStream_t s);
// Exclusive prefix sum of n int32 values (in may alias out). Returns nothing;
// the total is out[n-1] + in[n-1] which callers read back when needed.
void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, DevBuf<int32_t>& scratch,
                        hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of key.
void radix_sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int bits, DevBuf<uint64_t>& tk,
                      DevBuf<int32_t>& tv, DevBuf<int32_t>& hist, DevBuf<int32_t>& scratch,
                      hipStream_t s);
// vals[i] = i
void launch_iota(int32_t* vals, int64_t n, hipStream_t s);
// out[i] = in[idx[i]] (gather) for doubles
void launch_gather_f64(const double* in, const int32_t* idx, int64_t n, double* out, hipStream_t s);
// FP64 direct sum: out[i] (+)= sum_j w[j] * phi(t_i - s_j) over transformed points.
void direct_sum_f64(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                    const double* trg, int64_t n_trg, double* out, bool accumulate,
                    DevBuf<double>& scratch, hipStream_t s);
// Doubles of partial sums the FP64 direct sum uses for n_trg targets and n_src
// sources (its split of the sources into chunks).
size_t direct_partial_size(int64_t n_trg, int64_t n_src);
// direct_sum_f64 with caller-provided partial sums (direct_partial_size
// doubles, device memory): allocates nothing, same kernels and split, so the
// same bits.
void direct_sum_f64_into(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                         const double* trg, int64_t n_trg, double* out, bool accumulate, double* partial,
                         hipStream_t s);
// FP64 direct gradient sum (gradient-transpose kernel): out3[3 i + k] =
// (A^T sum_j w[j] grad phi(t_i - s_j))_k over transformed points (rbf_eval.hip).
void direct_gradient_f64(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                         const double* trg, int64_t n_trg, double* out3, DevBuf<double>& scratch,
                         hipStream_t s);
// FP64 direct sum of column kcol of the Hessian kernel: out3[3 i + c] =
// (A^T sum_j w[j] H(t_i - s_j) A e_kcol)_c over transformed points (rbf_eval.hip).
void direct_hessian_column_f64(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                               const double* trg, int64_t n_trg, int kcol, double* out3, DevBuf<double>& scratch,
                               hipStream_t s);
// Batched FP64 direct sums over host query sets, for small evaluations of
// several resident models at once: one upload, one direct-sum launch, one
// reduction, one download. A term is one rbf of a model (its transformed
// sources and weights on the device); a query set sums its terms in order and
// adds its polynomial. Each term is split into blocks and source chunks as in
// direct_sum_f64, so a query set's result does not depend on the others in
// the batch.
struct DirectBatchTerm {
  RbfDesc rbf;        // kind, constants and the anisotropy applied to the queries
  const double* src;  // device, n_src x 3, transformed by rbf.aniso
  const double* w;    // device, n_src
  int64_t n_src;
};
struct DirectBatchQuery {
  const double* pts;  // host, n x 3, untransformed
  double* out;        // host, n
  int64_t n;
  int first_term, n_terms;
  int poly_degree;        // -1: no polynomial
  const double* coeffs;   // device, the polynomial coefficients
};
class DirectBatch {
 public:
  // Blocks until the outputs are written.
  void run(const std::vector<DirectBatchQuery>& queries, const std::vector<DirectBatchTerm>& terms,
           hipStream_t s);

 private:
  HostStage<double> up_, down_;
  DevBuf<double> d_in_, d_part_, d_out_;
};
// out3[3 i + k] += d/dx_k sum_m c_m m(p_i) (gradient of the polynomial part).
void launch_poly_grad(int degree, const double* pts, int64_t n, const double* d_coeffs, double* out3,
                      hipStream_t s);
// out[i] += sum_k c_k m_k(p_i) (monomials of degree 0..2, untransformed points).
void launch_poly_eval(int degree, const double* pts, int64_t n, const double* d_coeffs,
                      double* out, hipStream_t s);

}  // namespace pam
