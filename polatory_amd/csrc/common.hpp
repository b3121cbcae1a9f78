// Shared host/device infrastructure for the gfx950 RBF library: error handling,
// device buffers, RBF kernel descriptors.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace pam {

// Errors cross the C ABI as status codes; internally they are exceptions.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,  // std::invalid_argument in the reference API
  kRuntimeError = 2,     // std::runtime_error in the reference API
  kHipError = 3,
  kOutOfMemory = 4,
  kNotImplemented = 5,
};

[[noreturn]] inline void throw_invalid(const std::string& m) { throw Error(kInvalidArgument, m); }
[[noreturn]] inline void throw_runtime(const std::string& m) { throw Error(kRuntimeError, m); }

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
    throw Error(e == hipErrorOutOfMemory ? kOutOfMemory : kHipError, buf);
  }
}
#define PAM_HIP(x) ::pam::hip_check((x), #x, __FILE__, __LINE__)
#define PAM_LAUNCH_CHECK() ::pam::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Device memory through a per-device cache of released blocks
// (device_pool.hip): device_alloc never returns a block that work enqueued
// before its release may still use.
void* device_alloc(size_t bytes);
void device_release(void* p);
void device_pool_trim();             // frees every cached block of the current device
size_t device_pool_cached_bytes();   // bytes held for reuse on the current device
// Non-blocking streams of the current device, reused across owners.
hipStream_t stream_acquire();
void stream_release(hipStream_t s);
// Marks the calling thread as capturing a graph: allocations then never
// synchronize the device.
struct PoolCaptureScope {
  PoolCaptureScope();
  ~PoolCaptureScope();
  PoolCaptureScope(const PoolCaptureScope&) = delete;
  PoolCaptureScope& operator=(const PoolCaptureScope&) = delete;
};

// Owning device allocation (from device_alloc, released on destruction).
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr; o.n_ = o.cap_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_; cap_ = o.cap_;
      o.p_ = nullptr; o.n_ = o.cap_ = 0;
    }
    return *this;
  }
  // Grows the capacity when needed; contents are not preserved on growth.
  void resize(size_t n) {
    if (n > cap_) {
      release();
      size_t bytes = (n == 0 ? 1 : n) * sizeof(T);
      p_ = static_cast<T*>(device_alloc(bytes));
      cap_ = n;
    }
    n_ = n;
  }
  void release() {
    if (p_) device_release(p_);
    p_ = nullptr;
    n_ = cap_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }
  void upload(const T* h, size_t n, hipStream_t s) {
    resize(n);
    if (n) PAM_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t n, hipStream_t s) const {
    if (n) PAM_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n_) PAM_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
  size_t cap_ = 0;
};

// Pinned, grow-only host staging for uploads that must not synchronize the
// stream: the buffer is refilled only after the event of its previous upload
// has completed.
template <class T>
class HostStage {
 public:
  HostStage() = default;
  ~HostStage() {
    if (ev_) (void)hipEventDestroy(ev_);
    if (p_) (void)hipHostFree(p_);
  }
  HostStage(const HostStage&) = delete;
  HostStage& operator=(const HostStage&) = delete;
  // Host pointer to n elements, safe to overwrite.
  T* acquire(size_t n) {
    if (!ev_) PAM_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
    if (pending_) PAM_HIP(hipEventSynchronize(ev_));
    pending_ = false;
    if (n > cap_) {
      if (p_) PAM_HIP(hipHostFree(p_));
      p_ = nullptr;
      PAM_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), (n == 0 ? 1 : n) * sizeof(T)));
      cap_ = n;
    }
    return p_;
  }
  void upload(DevBuf<T>& d, size_t n, hipStream_t s) {
    d.resize(n);
    if (n) PAM_HIP(hipMemcpyAsync(d.get(), p_, n * sizeof(T), hipMemcpyHostToDevice, s));
    PAM_HIP(hipEventRecord(ev_, s));
    pending_ = true;
  }

 private:
  T* p_ = nullptr;
  size_t cap_ = 0;
  hipEvent_t ev_ = nullptr;
  bool pending_ = false;
};

// RBF families on the path (rbf/polyharmonic_odd.hpp, rbf/cov_exponential.hpp).
enum RbfKind : int { kBh3 = 0, kTh3 = 1, kExp = 2 };

struct RbfDesc {
  int kind;
  double p0, p1;  // bh3/th3: slope, c; exp: psill, range
  double aniso[9];
};

inline bool is_identity(const double* a) {
  for (int i = 0; i < 9; i++)
    if (a[i] != ((i % 4 == 0) ? 1.0 : 0.0)) return false;
  return true;
}

inline int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace pam
