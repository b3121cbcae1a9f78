// Gradient-kernel sources as three gradient-transpose passes.
#include "fmm_dipole.hpp"

#include "device_api.hpp"

namespace pam {

namespace {

__global__ void extract_component_kernel(const double* __restrict__ w3, int64_t n, int k, double* __restrict__ wk) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wk[i] = w3[3 * i + k];
}

// out = -g3[., k] for k == 0, out -= g3[., k] after
__global__ void minus_component_kernel(const double* __restrict__ g3, int64_t n, int k, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = k == 0 ? -g3[3 * i] : out[i] - g3[3 * i + k];
}

// out3 = -g3 for k == 0, out3 -= g3 after
__global__ void minus_all_kernel(const double* __restrict__ g3, int64_t n3, int k, double* __restrict__ out3) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  out3[i] = k == 0 ? -g3[i] : out3[i] - g3[i];
}

}  // namespace

DipoleSourceEvaluator::DipoleSourceEvaluator(FmmEvaluator& ev) : ev_(ev) {
  if (!ev.gradient_kernel()) throw_runtime("the evaluator must be in gradient-kernel mode");
}

void DipoleSourceEvaluator::set_weights(const double* w3, int64_t n, bool on_device) {
  if (n != 3 * ev_.n_src()) throw_invalid("weights.rows() must be 3 x the number of source points");
  w3_.resize(static_cast<size_t>(n));
  if (n > 0)
    PAM_HIP(hipMemcpyAsync(w3_.get(), w3, sizeof(double) * n,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ev_.stream()));
  if (!on_device) PAM_HIP(hipStreamSynchronize(ev_.stream()));
  n_src_ = ev_.n_src();
}

void DipoleSourceEvaluator::evaluate(double* out, int64_t n, bool on_device) {
  const int64_t nt = ev_.n_trg();
  if (n != nt) throw_invalid("output size must equal the number of target points");
  if (n_src_ != ev_.n_src()) throw_runtime("weights are not set");
  hipStream_t s = ev_.stream();
  double* d_out = out;
  if (!on_device) {
    tmp_.resize(static_cast<size_t>(nt));
    d_out = tmp_.get();
  }
  wk_.resize(static_cast<size_t>(n_src_));
  g3_.resize(static_cast<size_t>(nt) * 3);
  for (int k = 0; k < 3; k++) {
    if (n_src_ > 0) {
      extract_component_kernel<<<ceil_div(n_src_, 256), 256, 0, s>>>(w3_.get(), n_src_, k, wk_.get());
      PAM_LAUNCH_CHECK();
    }
    ev_.set_weights(wk_.get(), n_src_, true);
    ev_.evaluate_gradient(g3_.get(), nt, true);
    if (nt > 0) {
      minus_component_kernel<<<ceil_div(nt, 256), 256, 0, s>>>(g3_.get(), nt, k, d_out);
      PAM_LAUNCH_CHECK();
    }
  }
  if (!on_device) {
    tmp_.download(out, static_cast<size_t>(nt), s);
    PAM_HIP(hipStreamSynchronize(s));
  }
}

DipoleHessianEvaluator::DipoleHessianEvaluator(FmmEvaluator& ev) : ev_(ev) {
  if (!ev.gradient_kernel()) throw_runtime("the evaluator must be in gradient-kernel mode");
}

void DipoleHessianEvaluator::set_weights(const double* w3, int64_t n, bool on_device) {
  if (n != 3 * ev_.n_src()) throw_invalid("weights.rows() must be 3 x the number of source points");
  w3_.resize(static_cast<size_t>(n));
  if (n > 0)
    PAM_HIP(hipMemcpyAsync(w3_.get(), w3, sizeof(double) * n,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ev_.stream()));
  if (!on_device) PAM_HIP(hipStreamSynchronize(ev_.stream()));
  n_src_ = ev_.n_src();
}

void DipoleHessianEvaluator::evaluate(double* out, int64_t n, bool on_device) {
  const int64_t nt = ev_.n_trg();
  if (n != nt) throw_invalid("output size must equal the number of target points");
  if (n_src_ != ev_.n_src()) throw_runtime("weights are not set");
  hipStream_t s = ev_.stream();
  double* d_out = out;
  if (!on_device) {
    tmp_.resize(static_cast<size_t>(nt) * 3);
    d_out = tmp_.get();
  }
  wk_.resize(static_cast<size_t>(n_src_));
  g3_.resize(static_cast<size_t>(nt) * 3);
  for (int k = 0; k < 3; k++) {
    if (n_src_ > 0) {
      extract_component_kernel<<<ceil_div(n_src_, 256), 256, 0, s>>>(w3_.get(), n_src_, k, wk_.get());
      PAM_LAUNCH_CHECK();
    }
    ev_.set_weights(wk_.get(), n_src_, true);
    ev_.evaluate_hessian_column(k, g3_.get(), nt, true);
    if (nt > 0) {
      minus_all_kernel<<<ceil_div(3 * nt, 256), 256, 0, s>>>(g3_.get(), 3 * nt, k, d_out);
      PAM_LAUNCH_CHECK();
    }
  }
  if (!on_device) {
    tmp_.download(out, static_cast<size_t>(nt) * 3, s);
    PAM_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace pam
