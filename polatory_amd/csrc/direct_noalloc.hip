// FP64 direct sum with device inputs and outputs that allocates no memory:
// the transformed points and the partial sums live in a fixed-size device
// array of this translation unit, one copy per GPU for the life of the
// process. For callers that must not call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// The kernels, the source split and the summation order are those of
// pam_direct_evaluate, so for the same inputs the result is the same bits.
#include <array>
#include <mutex>

#include "../../include/polatory_amd.h"
#include "capi_util.hpp"
#include "device_api.hpp"

namespace {

constexpr int64_t kScratchDoubles = PAM_NOALLOC_SCRATCH_BYTES / 8;

// The shared work area. Host code never uses this name as an address: it
// stands for the host-side shadow, not device memory. The device address is
// looked up per device with hipGetSymbolAddress and kept in g_ptr.
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<double*, kMaxDevices> g_ptr{};

// Device address of the current device's copy of the work area, looked up on
// the first call on that device.
double* scratch_ptr() {
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) pam::throw_runtime("device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    PAM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) pam::throw_runtime("hipGetSymbolAddress returned a null address");
    g_ptr[dev] = static_cast<double*>(p);
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" {

int pam_direct_evaluate_device_noalloc(const pam_rbf* rbf, const double* d_src, int64_t n_src,
                                       const double* d_w, const double* d_trg, int64_t n_trg,
                                       double* d_out, void* stream) {
  return pam::guard([&] {
    auto d = pam::to_desc(rbf);
    if (n_src < 0 || n_trg < 0) pam::throw_invalid("negative point count");
    if (n_src > 0 && (!d_src || !d_w)) pam::throw_invalid("null source or weight pointer");
    if (n_trg > 0 && (!d_trg || !d_out)) pam::throw_invalid("null target or output pointer");
    // Work area: transformed targets, transformed sources, partial sums.
    const int64_t need = 3 * n_trg + 3 * n_src + static_cast<int64_t>(pam::direct_partial_size(n_trg, n_src));
    if (need > kScratchDoubles)
      pam::throw_invalid("pam_direct_evaluate_device_noalloc: 3 n_trg + 3 n_src + partial sums = " +
                         std::to_string(need) + " doubles exceed the fixed work area of " +
                         std::to_string(kScratchDoubles) + " doubles (PAM_NOALLOC_SCRATCH_BYTES)");
    if (n_trg == 0) return;
    double* base = scratch_ptr();
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* t_pts = base;
    double* s_pts = t_pts + 3 * n_trg;
    double* partial = s_pts + 3 * n_src;
    pam::launch_transform_points(d.aniso, d_trg, n_trg, t_pts, s);
    pam::launch_transform_points(d.aniso, d_src, n_src, s_pts, s);
    pam::direct_sum_f64_into(d, s_pts, d_w, n_src, t_pts, n_trg, d_out, false, partial, s);
  });
}

}  // extern "C"
