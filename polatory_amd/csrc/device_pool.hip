// Caching device allocator behind DevBuf.
//
// hipFree synchronizes the device and unmaps the block; a 1M-point fit frees
// about 900 buffers, and their hipFree calls alone took 61 ms after the solve
// and blocked the host inside set-up. Released blocks are kept per device and
// handed out again instead:
// - a released block first goes to a pending list: work already enqueued on
//   any stream may still use it;
// - pending blocks become reusable after one hipDeviceSynchronize, issued by
//   an allocation that finds no reusable block but a pending one that fits
//   (never inside a graph capture on the calling thread: there the
//   allocation falls through to hipMalloc, and a block released there is
//   kept out of reuse for good, since the graph may reference it);
// - an allocation that hipMalloc cannot satisfy frees every cached block of
//   the device and tries once more;
// - cached bytes are capped (a quarter of the device memory); releases past
//   the cap go straight to hipFree.
// PAM_DEVICE_POOL=0 in the environment turns the cache off (every DevBuf is
// then hipMalloc'd and hipFree'd). The blocks are never freed at process
// exit: the runtime may already be gone by then.
#include "common.hpp"

#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace pam {
namespace {

constexpr int kMaxDevices = 64;

struct Block {
  size_t bytes;
  int dev;
};

struct Pool {
  std::multimap<size_t, void*> reusable[kMaxDevices];
  std::vector<std::pair<void*, size_t>> pending[kMaxDevices];
  size_t cached[kMaxDevices] = {};
  size_t cap[kMaxDevices] = {};
  std::unordered_map<void*, Block> live;
};

std::mutex g_mu;
Pool* g_pool = nullptr;  // never deleted
thread_local int t_capture_depth = 0;

bool pool_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("PAM_DEVICE_POOL");
    return !(e && e[0] == '0');
  }();
  return on;
}

Pool& pool() {
  if (!g_pool) g_pool = new Pool();
  return *g_pool;
}

size_t round_bytes(size_t bytes) {
  constexpr size_t kSmall = 512, kLarge = size_t{2} << 20;
  if (bytes < (size_t{1} << 20)) return (bytes + kSmall - 1) / kSmall * kSmall;
  return (bytes + kLarge - 1) / kLarge * kLarge;
}

// A cached block of at least r bytes and at most r + r/4 (under g_mu).
void* take(std::multimap<size_t, void*>& m, size_t r, size_t* got) {
  auto it = m.lower_bound(r);
  if (it == m.end() || it->first - r > r / 4) return nullptr;
  void* p = it->second;
  *got = it->first;
  m.erase(it);
  return p;
}

bool pending_fits(const std::vector<std::pair<void*, size_t>>& v, size_t r) {
  for (const auto& b : v)
    if (b.second >= r && b.second - r <= r / 4) return true;
  return false;
}

// Frees every cached block of dev (under g_mu; the pending ones after a
// device synchronization).
void trim(Pool& P, int dev) {
  if (!P.pending[dev].empty()) {
    (void)hipDeviceSynchronize();
    for (auto& b : P.pending[dev]) (void)hipFree(b.first);
    P.pending[dev].clear();
  }
  for (auto& b : P.reusable[dev]) (void)hipFree(b.second);
  P.reusable[dev].clear();
  P.cached[dev] = 0;
}

}  // namespace

void* device_alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (!pool_enabled()) {
    void* p = nullptr;
    PAM_HIP(hipMalloc(&p, bytes));
    return p;
  }
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw_runtime("device index out of range for the allocator");
  const size_t r = round_bytes(bytes);
  std::lock_guard<std::mutex> lock(g_mu);
  Pool& P = pool();
  size_t got = 0;
  void* p = take(P.reusable[dev], r, &got);
  if (!p && t_capture_depth == 0 && pending_fits(P.pending[dev], r) &&
      hipDeviceSynchronize() == hipSuccess) {
    for (auto& b : P.pending[dev]) P.reusable[dev].emplace(b.second, b.first);
    P.pending[dev].clear();
    p = take(P.reusable[dev], r, &got);
  }
  if (p) {
    P.cached[dev] -= got;
    P.live.emplace(p, Block{got, dev});
    return p;
  }
  hipError_t e = hipMalloc(&p, r);
  if (e == hipErrorOutOfMemory && P.cached[dev] > 0 && t_capture_depth == 0) {
    (void)hipGetLastError();
    trim(P, dev);
    e = hipMalloc(&p, r);
  }
  PAM_HIP(e);
  P.live.emplace(p, Block{r, dev});
  return p;
}

void device_release(void* p) {
  if (!p) return;
  if (!pool_enabled()) {
    (void)hipFree(p);
    return;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  Pool& P = pool();
  auto it = P.live.find(p);
  if (it == P.live.end()) {  // not ours
    (void)hipFree(p);
    return;
  }
  const Block b = it->second;
  P.live.erase(it);
  // A block released while this thread captures a graph may be referenced by
  // that graph: it is neither reused nor freed.
  if (t_capture_depth > 0) return;
  if (P.cap[b.dev] == 0) {
    size_t free_b = 0, total_b = 0;
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur == b.dev && hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0) P.cap[b.dev] = total_b / 4;
    else P.cap[b.dev] = size_t{16} << 30;
  }
  if (P.cached[b.dev] + b.bytes > P.cap[b.dev]) {
    (void)hipFree(p);
    return;
  }
  P.pending[b.dev].emplace_back(p, b.bytes);
  P.cached[b.dev] += b.bytes;
}

void device_pool_trim() {
  if (!pool_enabled()) return;
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_mu);
  trim(pool(), dev);
}

size_t device_pool_cached_bytes() {
  int dev = 0;
  if (!pool_enabled() || hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 0;
  std::lock_guard<std::mutex> lock(g_mu);
  return pool().cached[dev];
}

// Streams: hipStreamCreateWithFlags and hipStreamDestroy cost 0.1-2 ms each
// and a 1M-point fit made about 19 of each. Released non-blocking streams
// are kept per device and handed out again; work still queued on a released
// stream simply runs before the next owner's.
namespace {
std::mutex g_stream_mu;
std::vector<hipStream_t>* g_streams = nullptr;  // [kMaxDevices], never deleted
}  // namespace

hipStream_t stream_acquire() {
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  if (dev >= 0 && dev < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_stream_mu);
    if (!g_streams) g_streams = new std::vector<hipStream_t>[kMaxDevices];
    auto& v = g_streams[dev];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  PAM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void stream_release(hipStream_t s) {
  if (!s) return;
  int dev = -1;
  if (hipStreamGetDevice(s, &dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(s);
    return;
  }
  std::lock_guard<std::mutex> lock(g_stream_mu);
  if (!g_streams) g_streams = new std::vector<hipStream_t>[kMaxDevices];
  if (g_streams[dev].size() >= 64) {
    (void)hipStreamDestroy(s);
    return;
  }
  g_streams[dev].push_back(s);
}

PoolCaptureScope::PoolCaptureScope() { t_capture_depth++; }
PoolCaptureScope::~PoolCaptureScope() { t_capture_depth--; }

}  // namespace pam
