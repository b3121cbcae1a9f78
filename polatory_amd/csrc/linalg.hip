// Device BLAS-1/2 kernels for the fit path: two-stage deterministic reductions,
// a fused multi-dot (classical Gram-Schmidt projections in one pass over the
// basis), the matching update, and polynomial-block helpers.
#include <algorithm>
#include <cmath>

#include "linalg.hpp"

namespace pam {

namespace {

constexpr int kT = 256;
constexpr int kRedBlocks = 512;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// partial[i * nb + b] = sum over block b's chunk of V_i . w
__global__ void multi_dot_kernel(const double* __restrict__ v, int64_t ldv, int nvec,
                                 const double* __restrict__ w, int64_t n, int64_t chunk,
                                 double* __restrict__ partial) {
  __shared__ double sh[kT];
  int64_t b0 = static_cast<int64_t>(blockIdx.x) * chunk;
  int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
  for (int i = 0; i < nvec; i++) {
    const double* vi = v + static_cast<int64_t>(i) * ldv;
    double acc = 0.0;
    for (int64_t k = b0 + threadIdx.x; k < b1; k += kT) acc += vi[k] * w[k];
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[static_cast<int64_t>(i) * gridDim.x + blockIdx.x] = t;
  }
}

__global__ void reduce_rows_kernel(const double* __restrict__ partial, int nb, int nrows,
                                   double* __restrict__ out) {
  __shared__ double sh[kT];
  int i = blockIdx.x;
  if (i >= nrows) return;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += kT) acc += partial[static_cast<int64_t>(i) * nb + b];
  double t = block_sum(acc, sh);
  if (threadIdx.x == 0) out[i] = t;
}

__global__ void gemv_update_kernel(double* __restrict__ w, const double* __restrict__ v, int64_t ldv,
                                   int nvec, const double* __restrict__ coef, int64_t n,
                                   double alpha) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double acc = 0.0;
  for (int i = 0; i < nvec; i++) acc += coef[i] * v[static_cast<int64_t>(i) * ldv + k];
  w[k] += alpha * acc;
}

__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) y[k] += a * x[k];
}

__global__ void scal_kernel(double a, double* __restrict__ x, int64_t n) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) x[k] *= a;
}

__global__ void fill_kernel(double v, double* __restrict__ x, int64_t n) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) x[k] = v;
}

__global__ void gather_kernel(const double* __restrict__ src, const int32_t* __restrict__ idx, int64_t n,
                              double* __restrict__ dst) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[idx[k]];
}

__global__ void gather_points_kernel(const double* __restrict__ src, const int32_t* __restrict__ idx,
                                     int64_t n, double* __restrict__ dst) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t j = idx[k];
  dst[3 * k] = src[3 * j];
  dst[3 * k + 1] = src[3 * j + 1];
  dst[3 * k + 2] = src[3 * j + 2];
}

__global__ void scatter_add_kernel(const double* __restrict__ src, const int32_t* __restrict__ idx,
                                   int64_t n, double alpha, double* __restrict__ dst) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) dst[idx[k]] += alpha * src[k];
}

__global__ void scatter_kernel(const double* __restrict__ src, const int32_t* __restrict__ idx, int64_t n,
                               double* __restrict__ dst) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) dst[idx[k]] = src[k];
}

__global__ void max_abs_diff_kernel(const double* __restrict__ a, const int32_t* __restrict__ idx,
                                    const double* __restrict__ b, int64_t n,
                                    double* __restrict__ partial) {
  __shared__ double sh[kT];
  double m = 0.0;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * kT) {
    double av = a[idx ? idx[k] : k];
    double d = b ? av - b[k] : av;
    m = fmax(m, fabs(d));
  }
  double t = block_max(m, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__device__ __forceinline__ void monomials(int degree, double x, double y, double z, double* m) {
  m[0] = 1.0;
  if (degree >= 1) {
    m[1] = x;
    m[2] = y;
    m[3] = z;
  }
  if (degree >= 2) {
    m[4] = x * x;
    m[5] = x * y;
    m[6] = x * z;
    m[7] = y * y;
    m[8] = y * z;
    m[9] = z * z;
  }
}

__global__ void poly_matrix_kernel(int degree, int l, const double* __restrict__ pts, int64_t n,
                                   double* __restrict__ p) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m[10];
  monomials(degree, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], m);
  for (int k = 0; k < l; k++) p[k * n + i] = m[k];
}

// partial[k * nb + b] = sum over block b of m_k(x_i) w_i
__global__ void poly_t_kernel(int degree, int l, const double* __restrict__ pts,
                              const double* __restrict__ w, int64_t n, int64_t chunk,
                              double* __restrict__ partial) {
  __shared__ double sh[kT];
  int64_t b0 = static_cast<int64_t>(blockIdx.x) * chunk;
  int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
    double m[10];
    monomials(degree, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], m);
    for (int k = 0; k < l; k++) acc[k] += m[k] * w[i];
  }
  for (int k = 0; k < l; k++) {
    double t = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partial[static_cast<int64_t>(k) * gridDim.x + blockIdx.x] = t;
  }
}

int red_blocks(int64_t n) {
  int64_t nb = (n + kT * 4 - 1) / (kT * 4);
  if (nb > kRedBlocks) nb = kRedBlocks;
  if (nb < 1) nb = 1;
  return static_cast<int>(nb);
}

}  // namespace

void dev_multi_dot(const double* v, int64_t ldv, int nvec, const double* w, int64_t n, double* h,
                   Workspace1& ws, hipStream_t s) {
  if (nvec <= 0) return;
  int nb = red_blocks(n);
  int64_t chunk = (n + nb - 1) / nb;
  ws.partial.resize(static_cast<size_t>(nvec) * nb);
  ws.small.resize(static_cast<size_t>(nvec));
  multi_dot_kernel<<<nb, kT, 0, s>>>(v, ldv, nvec, w, n, chunk, ws.partial.get());
  PAM_LAUNCH_CHECK();
  reduce_rows_kernel<<<nvec, kT, 0, s>>>(ws.partial.get(), nb, nvec, ws.small.get());
  PAM_LAUNCH_CHECK();
  PAM_HIP(hipMemcpyAsync(h, ws.small.get(), sizeof(double) * nvec, hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
}

void dev_multi_dot_device(const double* v, int64_t ldv, int nvec, const double* w, int64_t n,
                          double* d_h, Workspace1& ws, hipStream_t s) {
  if (nvec <= 0) return;
  int nb = red_blocks(n);
  int64_t chunk = (n + nb - 1) / nb;
  ws.partial.resize(static_cast<size_t>(nvec) * nb);
  multi_dot_kernel<<<nb, kT, 0, s>>>(v, ldv, nvec, w, n, chunk, ws.partial.get());
  PAM_LAUNCH_CHECK();
  reduce_rows_kernel<<<nvec, kT, 0, s>>>(ws.partial.get(), nb, nvec, d_h);
  PAM_LAUNCH_CHECK();
}

double dev_dot(const double* x, const double* y, int64_t n, Workspace1& ws, hipStream_t s) {
  double r = 0.0;
  dev_multi_dot(x, 0, 1, y, n, &r, ws, s);
  return r;
}

void dev_gemv_update(double* w, const double* v, int64_t ldv, int nvec, const double* d_coef,
                     int64_t n, double alpha, hipStream_t s) {
  if (nvec <= 0 || n == 0) return;
  gemv_update_kernel<<<ceil_div(n, kT), kT, 0, s>>>(w, v, ldv, nvec, d_coef, n, alpha);
  PAM_LAUNCH_CHECK();
}

void dev_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s) {
  if (n == 0) return;
  axpy_kernel<<<ceil_div(n, kT), kT, 0, s>>>(a, x, y, n);
  PAM_LAUNCH_CHECK();
}

void dev_scal(double a, double* x, int64_t n, hipStream_t s) {
  if (n == 0) return;
  scal_kernel<<<ceil_div(n, kT), kT, 0, s>>>(a, x, n);
  PAM_LAUNCH_CHECK();
}

void dev_fill(double v, double* x, int64_t n, hipStream_t s) {
  if (n == 0) return;
  fill_kernel<<<ceil_div(n, kT), kT, 0, s>>>(v, x, n);
  PAM_LAUNCH_CHECK();
}

void dev_copy(const double* x, double* y, int64_t n, hipStream_t s) {
  if (n == 0 || x == y) return;
  PAM_HIP(hipMemcpyAsync(y, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
}

void dev_gather(const double* src, const int32_t* idx, int64_t n, double* dst, hipStream_t s) {
  if (n == 0) return;
  gather_kernel<<<ceil_div(n, kT), kT, 0, s>>>(src, idx, n, dst);
  PAM_LAUNCH_CHECK();
}

void dev_gather_points(const double* src, const int32_t* idx, int64_t n, double* dst, hipStream_t s) {
  if (n == 0) return;
  gather_points_kernel<<<ceil_div(n, kT), kT, 0, s>>>(src, idx, n, dst);
  PAM_LAUNCH_CHECK();
}

void dev_scatter_add(const double* src, const int32_t* idx, int64_t n, double alpha, double* dst,
                     hipStream_t s) {
  if (n == 0) return;
  scatter_add_kernel<<<ceil_div(n, kT), kT, 0, s>>>(src, idx, n, alpha, dst);
  PAM_LAUNCH_CHECK();
}

void dev_scatter(const double* src, const int32_t* idx, int64_t n, double* dst, hipStream_t s) {
  if (n == 0) return;
  scatter_kernel<<<ceil_div(n, kT), kT, 0, s>>>(src, idx, n, dst);
  PAM_LAUNCH_CHECK();
}

double dev_max_abs_diff_idx(const double* a, const int32_t* idx, const double* b, int64_t n,
                            Workspace1& ws, hipStream_t s) {
  if (n == 0) return 0.0;
  int nb = red_blocks(n);
  ws.partial.resize(nb);
  max_abs_diff_kernel<<<nb, kT, 0, s>>>(a, idx, b, n, ws.partial.get());
  PAM_LAUNCH_CHECK();
  ws.host.resize(nb);
  PAM_HIP(hipMemcpyAsync(ws.host.data(), ws.partial.get(), sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  double m = 0.0;
  for (double v : ws.host) m = std::max(m, v);
  return m;
}

double dev_max_abs_diff(const double* a, const double* b, int64_t n, Workspace1& ws, hipStream_t s) {
  return dev_max_abs_diff_idx(a, nullptr, b, n, ws, s);
}

int poly_basis_size(int degree) {
  if (degree < 0) return 0;
  int k = degree + 1;
  return k * (k + 1) * (k + 2) / 6;
}

void dev_poly_matrix(int degree, const double* pts, int64_t n, double* p, hipStream_t s) {
  int l = poly_basis_size(degree);
  if (l == 0 || n == 0) return;
  poly_matrix_kernel<<<ceil_div(n, kT), kT, 0, s>>>(degree, l, pts, n, p);
  PAM_LAUNCH_CHECK();
}

void dev_poly_transpose_apply(int degree, const double* pts, int64_t n, const double* w,
                              double* d_out, Workspace1& ws, hipStream_t s) {
  int l = poly_basis_size(degree);
  if (l == 0) return;
  if (n == 0) {
    dev_fill(0.0, d_out, l, s);
    return;
  }
  int nb = red_blocks(n);
  int64_t chunk = (n + nb - 1) / nb;
  ws.partial.resize(static_cast<size_t>(l) * nb);
  poly_t_kernel<<<nb, kT, 0, s>>>(degree, l, pts, w, n, chunk, ws.partial.get());
  PAM_LAUNCH_CHECK();
  reduce_rows_kernel<<<l, kT, 0, s>>>(ws.partial.get(), nb, l, d_out);
  PAM_LAUNCH_CHECK();
}

void dev_orthonormalize_cols(double* p, int64_t n, int l, Workspace1& ws, hipStream_t s) {
  for (int i = 0; i < l; i++) {
    double* ci = p + static_cast<int64_t>(i) * n;
    double nrm = std::sqrt(dev_dot(ci, ci, n, ws, s));
    dev_scal(1.0 / nrm, ci, n, s);
    for (int j = i + 1; j < l; j++) {
      double* cj = p + static_cast<int64_t>(j) * n;
      double d = dev_dot(ci, cj, n, ws, s);
      dev_axpy(-d, ci, cj, n, s);
    }
  }
}

}  // namespace pam
