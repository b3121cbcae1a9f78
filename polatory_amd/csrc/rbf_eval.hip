// Pointwise RBF values, gradients and Hessians, and the FP64 direct gradient
// sum, on the GPU (rbf/rbf_base.hpp:40-53, rbf/polyharmonic_odd.hpp:32-72,
// rbf/cov_exponential.hpp:33-62; the gradient sum is the full_direct form of
// fmm/gradient_kernel.hpp, src/fmm/full_direct.hpp:7-52).
#include "capi_util.hpp"
#include "device_api.hpp"
#include "rbf_device.hpp"

namespace pam {
namespace {

struct Aniso {
  double a[9];  // row-major A; a difference d maps to A d
};

__device__ inline void apply(const Aniso& A, const double d[3], double u[3]) {
  for (int i = 0; i < 3; i++) u[i] = (A.a[3 * i] * d[0] + A.a[3 * i + 1] * d[1]) + A.a[3 * i + 2] * d[2];
}

// what: 0 value (n), 1 gradient A^T g(A d) (3n), 2 Hessian A^T H(A d) A (9n, row-major).
template <int KIND>
__global__ void rbf_point_kernel(double p0, double p1, Aniso A, const double* __restrict__ diff, int64_t n,
                                 int what, double* __restrict__ out) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d[3] = {diff[3 * i], diff[3 * i + 1], diff[3 * i + 2]};
  double u[3];
  apply(A, d, u);
  if (what == 0) {
    out[i] = rbf_value_iso<KIND>(p0, p1, u);
  } else if (what == 1) {
    double g[3];
    rbf_grad_iso<KIND>(p0, p1, u, g);
    for (int k = 0; k < 3; k++) out[3 * i + k] = (A.a[k] * g[0] + A.a[3 + k] * g[1]) + A.a[6 + k] * g[2];
  } else {
    double h[9], ha[9];
    rbf_hess_iso<KIND>(p0, p1, u, h);
    for (int r = 0; r < 3; r++)  // H A
      for (int c = 0; c < 3; c++) ha[3 * r + c] = (h[3 * r] * A.a[c] + h[3 * r + 1] * A.a[3 + c]) + h[3 * r + 2] * A.a[6 + c];
    for (int r = 0; r < 3; r++)  // A^T (H A)
      for (int c = 0; c < 3; c++)
        out[9 * i + 3 * r + c] = (A.a[r] * ha[c] + A.a[3 + r] * ha[3 + c]) + A.a[6 + r] * ha[6 + c];
  }
}

// Gradient sum over transformed points: partial[chunk][target][3] =
// sum over the chunk's sources of w_j g(u_t - u_s); sources tiled through LDS.
constexpr int kGradThreads = 256;

template <int KIND>
__global__ void __launch_bounds__(kGradThreads)
direct_grad_kernel(double p0, double p1, const double* __restrict__ src, const double* __restrict__ w,
                   int64_t n_src, const double* __restrict__ trg, int64_t n_trg, double* __restrict__ partial,
                   int64_t chunk) {
  __shared__ double4 tile[kGradThreads];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kGradThreads + threadIdx.x;
  const int64_t j0 = static_cast<int64_t>(blockIdx.y) * chunk;
  const int64_t j1 = j0 + chunk < n_src ? j0 + chunk : n_src;
  double t[3] = {0.0, 0.0, 0.0};
  if (i < n_trg) {
    t[0] = trg[3 * i];
    t[1] = trg[3 * i + 1];
    t[2] = trg[3 * i + 2];
  }
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t jt = j0; jt < j1; jt += kGradThreads) {
    __syncthreads();
    const int64_t j = jt + threadIdx.x;
    if (j < j1) tile[threadIdx.x] = make_double4(src[3 * j], src[3 * j + 1], src[3 * j + 2], w[j]);
    __syncthreads();
    const int cnt = static_cast<int>((j1 - jt) < kGradThreads ? (j1 - jt) : kGradThreads);
    for (int k = 0; k < cnt; k++) {
      const double4 q = tile[k];
      const double u[3] = {t[0] - q.x, t[1] - q.y, t[2] - q.z};
      double g[3];
      rbf_grad_iso<KIND>(p0, p1, u, g);
      for (int c = 0; c < 3; c++) acc[c] += q.w * g[c];
    }
  }
  if (i < n_trg)
    for (int c = 0; c < 3; c++) partial[(static_cast<int64_t>(blockIdx.y) * n_trg + i) * 3 + c] = acc[c];
}

// Column kcol of the Hessian kernel by the direct sum:
// partial = sum_j w_j H(t_i - s_j) a_k (transformed points; A^T applied in the reduction).
template <int KIND>
__global__ void __launch_bounds__(kGradThreads)
direct_hess_col_kernel(double p0, double p1, Aniso A, int kcol, const double* __restrict__ src,
                       const double* __restrict__ w, int64_t n_src, const double* __restrict__ trg, int64_t n_trg,
                       double* __restrict__ partial, int64_t chunk) {
  __shared__ double4 tile[kGradThreads];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kGradThreads + threadIdx.x;
  const int64_t j0 = static_cast<int64_t>(blockIdx.y) * chunk;
  const int64_t j1 = j0 + chunk < n_src ? j0 + chunk : n_src;
  const double ak[3] = {A.a[kcol], A.a[3 + kcol], A.a[6 + kcol]};
  double t[3] = {0.0, 0.0, 0.0};
  if (i < n_trg) {
    t[0] = trg[3 * i];
    t[1] = trg[3 * i + 1];
    t[2] = trg[3 * i + 2];
  }
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t jt = j0; jt < j1; jt += kGradThreads) {
    __syncthreads();
    const int64_t j = jt + threadIdx.x;
    if (j < j1) tile[threadIdx.x] = make_double4(src[3 * j], src[3 * j + 1], src[3 * j + 2], w[j]);
    __syncthreads();
    const int cnt = static_cast<int>((j1 - jt) < kGradThreads ? (j1 - jt) : kGradThreads);
    for (int k = 0; k < cnt; k++) {
      const double4 q = tile[k];
      const double u[3] = {t[0] - q.x, t[1] - q.y, t[2] - q.z};
      double tt;
      const double f = rbf_grad_factor<KIND>(p0, p1, u, &tt);
      const double ua = u[0] * ak[0] + u[1] * ak[1] + u[2] * ak[2];
      for (int c = 0; c < 3; c++) acc[c] += q.w * (f * (ak[c] + tt * u[c] * ua));
    }
  }
  if (i < n_trg)
    for (int c = 0; c < 3; c++) partial[(static_cast<int64_t>(blockIdx.y) * n_trg + i) * 3 + c] = acc[c];
}

// out[i] = A^T (sum of the chunks in order)
__global__ void grad_reduce_kernel(const double* __restrict__ partial, int nchunks, int64_t n_trg, Aniso A,
                                   double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_trg) return;
  double g[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < nchunks; c++)
    for (int k = 0; k < 3; k++) g[k] += partial[(static_cast<int64_t>(c) * n_trg + i) * 3 + k];
  for (int k = 0; k < 3; k++) out[3 * i + k] = (A.a[k] * g[0] + A.a[3 + k] * g[1]) + A.a[6 + k] * g[2];
}

}  // namespace

void direct_gradient_f64(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                         const double* trg, int64_t n_trg, double* out3, DevBuf<double>& scratch,
                         hipStream_t s) {
  if (n_trg == 0) return;
  Aniso A;
  for (int k = 0; k < 9; k++) A.a[k] = rbf.aniso[k];
  const int64_t chunk = 8192;
  const int nchunks = n_src > 0 ? ceil_div(n_src, chunk) : 0;
  scratch.resize(static_cast<size_t>(std::max(nchunks, 1)) * n_trg * 3);
  if (n_src > 0) {
    dim3 grid(ceil_div(n_trg, kGradThreads), nchunks);
    dispatch_kind(rbf.kind, [&](auto K) {
      direct_grad_kernel<decltype(K)::value><<<grid, kGradThreads, 0, s>>>(rbf.p0, rbf.p1, src, w, n_src, trg,
                                                                           n_trg, scratch.get(), chunk);
    });
    PAM_LAUNCH_CHECK();
  }
  grad_reduce_kernel<<<ceil_div(n_trg, 256), 256, 0, s>>>(scratch.get(), nchunks, n_trg, A, out3);
  PAM_LAUNCH_CHECK();
}

void direct_hessian_column_f64(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                               const double* trg, int64_t n_trg, int kcol, double* out3, DevBuf<double>& scratch,
                               hipStream_t s) {
  if (n_trg == 0) return;
  if (kcol < 0 || kcol > 2) throw_invalid("Hessian column out of range");
  Aniso A;
  for (int k = 0; k < 9; k++) A.a[k] = rbf.aniso[k];
  const int64_t chunk = 8192;
  const int nchunks = n_src > 0 ? ceil_div(n_src, chunk) : 0;
  scratch.resize(static_cast<size_t>(std::max(nchunks, 1)) * n_trg * 3);
  if (n_src > 0) {
    dim3 grid(ceil_div(n_trg, kGradThreads), nchunks);
    dispatch_kind(rbf.kind, [&](auto K) {
      direct_hess_col_kernel<decltype(K)::value><<<grid, kGradThreads, 0, s>>>(rbf.p0, rbf.p1, A, kcol, src, w,
                                                                               n_src, trg, n_trg, scratch.get(), chunk);
    });
    PAM_LAUNCH_CHECK();
  }
  grad_reduce_kernel<<<ceil_div(n_trg, 256), 256, 0, s>>>(scratch.get(), nchunks, n_trg, A, out3);
  PAM_LAUNCH_CHECK();
}

namespace {

struct Stream {
  hipStream_t s = nullptr;
  Stream() { s = stream_acquire(); }
  ~Stream() {
    if (s) {
      (void)hipStreamSynchronize(s);
      stream_release(s);
    }
  }
};

void point_values(const pam_rbf* rbf, const double* diff, int64_t n, int what, double* out) {
  if (!rbf || (n > 0 && (!diff || !out))) throw_invalid("null argument");
  if (n < 0) throw_invalid("negative count");
  const RbfDesc d = to_desc(rbf);
  if (n == 0) return;
  const int per = what == 0 ? 1 : (what == 1 ? 3 : 9);
  Stream st;
  DevBuf<double> dd, dout;
  dd.upload(diff, static_cast<size_t>(n) * 3, st.s);
  dout.resize(static_cast<size_t>(n) * per);
  Aniso A;
  for (int k = 0; k < 9; k++) A.a[k] = d.aniso[k];
  dispatch_kind(d.kind, [&](auto K) {
    rbf_point_kernel<decltype(K)::value><<<ceil_div(n, 256), 256, 0, st.s>>>(d.p0, d.p1, A, dd.get(), n, what,
                                                                              dout.get());
  });
  PAM_LAUNCH_CHECK();
  dout.download(out, static_cast<size_t>(n) * per, st.s);
  PAM_HIP(hipStreamSynchronize(st.s));
}

}  // namespace
}  // namespace pam

extern "C" {

int pam_rbf_evaluate(const pam_rbf* rbf, const double* diff, int64_t n, double* out) {
  return pam::guard([&] { pam::point_values(rbf, diff, n, 0, out); });
}

int pam_rbf_evaluate_gradient(const pam_rbf* rbf, const double* diff, int64_t n, double* out) {
  return pam::guard([&] { pam::point_values(rbf, diff, n, 1, out); });
}

int pam_rbf_evaluate_hessian(const pam_rbf* rbf, const double* diff, int64_t n, double* out) {
  return pam::guard([&] { pam::point_values(rbf, diff, n, 2, out); });
}

int pam_direct_evaluate_gradient(const pam_rbf* rbf, const double* src, int64_t n_src, const double* w,
                                 const double* trg, int64_t n_trg, double* out) {
  using namespace pam;
  return guard([&] {
    if (!rbf || (n_src > 0 && (!src || !w)) || (n_trg > 0 && (!trg || !out))) throw_invalid("null argument");
    if (n_src < 0 || n_trg < 0) throw_invalid("negative count");
    const RbfDesc d = to_desc(rbf);
    if (n_trg == 0) return;
    Stream st;
    DevBuf<double> ds, dw, dt, us, ut, partial, dout;
    dt.upload(trg, static_cast<size_t>(n_trg) * 3, st.s);
    ut.resize(static_cast<size_t>(n_trg) * 3);
    launch_transform_points(d.aniso, dt.get(), n_trg, ut.get(), st.s);
    dout.resize(static_cast<size_t>(n_trg) * 3);
    if (n_src > 0) {
      ds.upload(src, static_cast<size_t>(n_src) * 3, st.s);
      dw.upload(w, static_cast<size_t>(n_src), st.s);
      us.resize(static_cast<size_t>(n_src) * 3);
      launch_transform_points(d.aniso, ds.get(), n_src, us.get(), st.s);
    }
    direct_gradient_f64(d, us.get(), dw.get(), n_src, ut.get(), n_trg, dout.get(), partial, st.s);
    dout.download(out, static_cast<size_t>(n_trg) * 3, st.s);
    PAM_HIP(hipStreamSynchronize(st.s));
  });
}

}  // extern "C"
