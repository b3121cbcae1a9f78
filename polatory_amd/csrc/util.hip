// Generic device building blocks: point transforms, bbox reduction, prefix
// sums, a stable LSD radix sort (4-bit digits), gathers, polynomial terms and
// the FP64 direct O(N*M) sum.
#include <cstring>
#include <type_traits>
#include <vector>

#include "device_api.hpp"
#include "rbf_device.hpp"

namespace pam {

namespace {

struct Mat3 {
  double a[9];
};

__global__ void transform_kernel(Mat3 m, const double* __restrict__ in, int64_t n,
                                 double* __restrict__ out) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
  // Explicit round-to-nearest ops: no FMA contraction, so the result equals the
  // CPU evaluation of ((A0 x + A1 y) + A2 z) bit for bit.
  out[3 * i] = __dadd_rn(__dadd_rn(__dmul_rn(m.a[0], x), __dmul_rn(m.a[1], y)), __dmul_rn(m.a[2], z));
  out[3 * i + 1] = __dadd_rn(__dadd_rn(__dmul_rn(m.a[3], x), __dmul_rn(m.a[4], y)), __dmul_rn(m.a[5], z));
  out[3 * i + 2] = __dadd_rn(__dadd_rn(__dmul_rn(m.a[6], x), __dmul_rn(m.a[7], y)), __dmul_rn(m.a[8], z));
}

constexpr int kBboxThreads = 256;

__global__ void bbox_partial_kernel(const double* __restrict__ pts, int64_t n,
                                    double* __restrict__ partial) {
  __shared__ double s[6][kBboxThreads];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    for (int k = 0; k < 3; k++) {
      double v = pts[3 * i + k];
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  }
  for (int k = 0; k < 3; k++) {
    s[k][threadIdx.x] = lo[k];
    s[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int off = kBboxThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + off]);
        s[3 + k][threadIdx.x] = fmax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + off]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void bbox_final_kernel(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double s[6][kBboxThreads];
  double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += kBboxThreads) {
    for (int k = 0; k < 3; k++) {
      r[k] = fmin(r[k], partial[b * 6 + k]);
      r[3 + k] = fmax(r[3 + k], partial[b * 6 + 3 + k]);
    }
  }
  for (int k = 0; k < 6; k++) s[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int off = kBboxThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + off]);
        s[3 + k][threadIdx.x] = fmax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + off]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// ---- exclusive scan: 256 threads x 4 items per block ----
constexpr int kScanThreads = 256;
constexpr int kScanTile = 1024;

__global__ void scan_block_kernel(const int32_t* in, int32_t* out, int64_t n, int32_t* block_sums) {
  __shared__ int32_t s[kScanThreads];
  int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * 4;
  int32_t v[4];
  int32_t tsum = 0;
  for (int q = 0; q < 4; q++) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    tsum += v[q];
  }
  s[threadIdx.x] = tsum;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int32_t t = (static_cast<int>(threadIdx.x) >= off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int32_t run = s[threadIdx.x] - tsum;
  for (int q = 0; q < 4; q++) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == kScanThreads - 1 && block_sums) block_sums[blockIdx.x] = s[threadIdx.x];
}

__global__ void scan_add_kernel(int32_t* out, int64_t n, const int32_t* offs) {
  int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int32_t o = offs[blockIdx.x];
  for (int q = 0; q < 4; q++) {
    int64_t i = base + threadIdx.x + q * kScanThreads;
    if (i < n) out[i] += o;
  }
}

// Exclusive scan of a short array in one workgroup: chunks of 4096 entries
// (4 consecutive per thread), wave scans by shuffles and a running carry.
// Every thread reads and writes only its own entries, so in may equal out.
constexpr int kScan1Threads = 1024;
constexpr int64_t kScan1Max = 65536;

__global__ void __launch_bounds__(kScan1Threads) scan_single_kernel(const int32_t* in, int32_t* out,
                                                                    int64_t n) {
  __shared__ int32_t wsum[kScan1Threads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int32_t carry = 0;
  for (int64_t base = 0; base < n; base += 4 * kScan1Threads) {
    const int64_t i0 = base + 4 * static_cast<int64_t>(tid);
    int32_t v[4];
    int32_t tsum = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      v[q] = i0 + q < n ? in[i0 + q] : 0;
      tsum += v[q];
    }
    int32_t incl = tsum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      int32_t u = __shfl_up(incl, d, 64);
      if (lane >= d) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int32_t before = 0, total = 0;
    for (int w = 0; w < kScan1Threads / 64; w++) {
      if (w < wave) before += wsum[w];
      total += wsum[w];
    }
    int32_t run = carry + before + incl - tsum;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (i0 + q < n) out[i0 + q] = run;
      run += v[q];
    }
    carry += total;
    __syncthreads();
  }
}

void scan_rec(const int32_t* in, int32_t* out, int64_t n, int32_t* scratch, hipStream_t s) {
  if (n <= kScan1Max) {
    scan_single_kernel<<<1, kScan1Threads, 0, s>>>(in, out, n);
    PAM_LAUNCH_CHECK();
    return;
  }
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb <= 1) {
    scan_block_kernel<<<1, kScanThreads, 0, s>>>(in, out, n, nullptr);
    PAM_LAUNCH_CHECK();
    return;
  }
  scan_block_kernel<<<static_cast<unsigned>(nb), kScanThreads, 0, s>>>(in, out, n, scratch);
  PAM_LAUNCH_CHECK();
  scan_rec(scratch, scratch, nb, scratch + nb, s);
  scan_add_kernel<<<static_cast<unsigned>(nb), kScanThreads, 0, s>>>(out, n, scratch);
  PAM_LAUNCH_CHECK();
}

// ---- radix sort: 4-bit digits, 256 threads x 8 consecutive items per tile ----
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;

__global__ void radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                  int32_t* __restrict__ hist, int nblocks) {
  __shared__ int32_t cnt[16];
  if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile;
  for (int q = 0; q < kRsItems; q++) {
    int64_t i = base + threadIdx.x + q * kRsThreads;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 15u], 1);
  }
  __syncthreads();
  if (threadIdx.x < 16) hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Counters for 16 digits packed as 4 x uint64 with 16-bit fields (a tile holds
// at most 2048 items, so no field overflows).
__global__ void radix_scatter_kernel(const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                     uint64_t* __restrict__ kout, int32_t* __restrict__ vout,
                                     int64_t n, int shift, const int32_t* __restrict__ offs,
                                     int nblocks) {
  __shared__ uint64_t pk[4][kRsThreads];
  __shared__ int32_t boff[16];
  int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile + threadIdx.x * kRsItems;
  uint64_t key[kRsItems];
  int32_t val[kRsItems];
  uint64_t c[4] = {0, 0, 0, 0};
  for (int q = 0; q < kRsItems; q++) {
    if (base + q < n) {
      key[q] = kin[base + q];
      val[q] = vin[base + q];
      unsigned d = (key[q] >> shift) & 15u;
      c[d >> 2] += 1ull << (16 * (d & 3));
    }
  }
  for (int k = 0; k < 4; k++) pk[k][threadIdx.x] = c[k];
  if (threadIdx.x < 16) boff[threadIdx.x] = offs[threadIdx.x * nblocks + blockIdx.x];
  __syncthreads();
  for (int off = 1; off < kRsThreads; off <<= 1) {
    uint64_t t[4];
    for (int k = 0; k < 4; k++) t[k] = (static_cast<int>(threadIdx.x) >= off) ? pk[k][threadIdx.x - off] : 0;
    __syncthreads();
    for (int k = 0; k < 4; k++) pk[k][threadIdx.x] += t[k];
    __syncthreads();
  }
  uint64_t run[4];
  for (int k = 0; k < 4; k++) run[k] = pk[k][threadIdx.x] - c[k];
  for (int q = 0; q < kRsItems; q++) {
    if (base + q < n) {
      unsigned d = (key[q] >> shift) & 15u;
      int64_t pos = boff[d] + static_cast<int64_t>((run[d >> 2] >> (16 * (d & 3))) & 0xFFFFull);
      run[d >> 2] += 1ull << (16 * (d & 3));
      kout[pos] = key[q];
      vout[pos] = val[q];
    }
  }
}

__global__ void iota_kernel(int32_t* v, int64_t n) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = static_cast<int32_t>(i);
}

__global__ void gather_f64_kernel(const double* __restrict__ in, const int32_t* __restrict__ idx,
                                  int64_t n, double* __restrict__ out) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

// ---- FP64 direct sum: targets x source chunks, deterministic two-stage sum ----
constexpr int kDirThreads = 256;

// One workgroup's share of an FP64 direct sum: the thread's target (tx, ty,
// tz) against sources [j0, j1), staged through LDS kDirThreads at a time and
// summed in source order.
template <int KIND>
__device__ __forceinline__ double direct_f64_range(double4* tile, double pa, double pb,
                                                   const double* __restrict__ src,
                                                   const double* __restrict__ w, int64_t j0, int64_t j1,
                                                   double tx, double ty, double tz) {
  double acc = 0.0;
  for (int64_t jt = j0; jt < j1; jt += kDirThreads) {
    __syncthreads();
    int64_t j = jt + threadIdx.x;
    if (j < j1) tile[threadIdx.x] = make_double4(src[3 * j], src[3 * j + 1], src[3 * j + 2], w[j]);
    __syncthreads();
    int cnt = static_cast<int>((j1 - jt) < kDirThreads ? (j1 - jt) : kDirThreads);
    for (int k = 0; k < cnt; k++) {
      double4 q = tile[k];
      double dx = tx - q.x, dy = ty - q.y, dz = tz - q.z;
      acc += q.w * phi_f64<KIND>(dx * dx + dy * dy + dz * dz, pa, pb);
    }
  }
  return acc;
}

template <int KIND>
__global__ void __launch_bounds__(kDirThreads)
direct_f64_kernel(double pa, double pb, const double* __restrict__ src, const double* __restrict__ w,
                  int64_t n_src, const double* __restrict__ trg, int64_t n_trg,
                  double* __restrict__ partial, int64_t chunk) {
  __shared__ double4 tile[kDirThreads];
  int64_t i = static_cast<int64_t>(blockIdx.x) * kDirThreads + threadIdx.x;
  int64_t j0 = static_cast<int64_t>(blockIdx.y) * chunk;
  int64_t j1 = j0 + chunk < n_src ? j0 + chunk : n_src;
  double tx = 0, ty = 0, tz = 0;
  if (i < n_trg) {
    tx = trg[3 * i];
    ty = trg[3 * i + 1];
    tz = trg[3 * i + 2];
  }
  double acc = direct_f64_range<KIND>(tile, pa, pb, src, w, j0, j1, tx, ty, tz);
  if (i < n_trg) partial[static_cast<int64_t>(blockIdx.y) * n_trg + i] = acc;
}

__global__ void reduce_partial_kernel(const double* __restrict__ partial, int nchunks, int64_t n,
                                      double* __restrict__ out, int accumulate) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int c = 0; c < nchunks; c++) acc += partial[c * n + i];
  out[i] = accumulate ? out[i] + acc : acc;
}

__global__ void fill_f64_kernel(double* __restrict__ out, int64_t n, double v) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

// ---- batched FP64 direct sums: several query sets, each against its own
// models' sources, in one launch and one reduction ----
//
// A segment is one (query set, rbf) pair. It is split into kDirThreads-target
// blocks and source chunks by the same rule as a single direct sum, so its
// partial sums, and hence every query set's result, do not depend on which
// other query sets share the launch.
struct DirSeg {
  const double* src;  // transformed sources, n_src x 3
  const double* w;    // weights, n_src
  int64_t n_src, chunk;
  int64_t part_off;   // its partial sums: nchunks x q_n
  int64_t q_off, q_n; // rows of the concatenated query points
  int32_t kind, gx, nchunks, work_first;
  double pa, pb;
  double a[9];        // anisotropy applied to the query points on load
};
static_assert(sizeof(DirSeg) % sizeof(double) == 0, "DirSeg is staged as doubles");

struct DirQuery {
  int64_t off, n;
  int32_t seg_first, n_seg, degree, pad;
  const double* coeffs;
};
static_assert(sizeof(DirQuery) % sizeof(double) == 0, "DirQuery is staged as doubles");

__global__ void __launch_bounds__(kDirThreads)
direct_batch_kernel(const DirSeg* __restrict__ segs, int nseg, const double* __restrict__ pts,
                    double* __restrict__ partial) {
  __shared__ double4 tile[kDirThreads];
  const int b = static_cast<int>(blockIdx.x);
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (segs[mid].work_first <= b) lo = mid;
    else hi = mid - 1;
  }
  const DirSeg& g = segs[lo];
  const int local = b - g.work_first;
  const int bx = local % g.gx, by = local / g.gx;
  const int64_t i = static_cast<int64_t>(bx) * kDirThreads + threadIdx.x;
  const int64_t j0 = static_cast<int64_t>(by) * g.chunk;
  const int64_t j1 = j0 + g.chunk < g.n_src ? j0 + g.chunk : g.n_src;
  double tx = 0, ty = 0, tz = 0;
  if (i < g.q_n) {
    const double* p = pts + 3 * (g.q_off + i);
    const double x = p[0], y = p[1], z = p[2];
    // the rounding of transform_kernel
    tx = __dadd_rn(__dadd_rn(__dmul_rn(g.a[0], x), __dmul_rn(g.a[1], y)), __dmul_rn(g.a[2], z));
    ty = __dadd_rn(__dadd_rn(__dmul_rn(g.a[3], x), __dmul_rn(g.a[4], y)), __dmul_rn(g.a[5], z));
    tz = __dadd_rn(__dadd_rn(__dmul_rn(g.a[6], x), __dmul_rn(g.a[7], y)), __dmul_rn(g.a[8], z));
  }
  double acc;
  switch (g.kind) {
    case kBh3: acc = direct_f64_range<kBh3>(tile, g.pa, g.pb, g.src, g.w, j0, j1, tx, ty, tz); break;
    case kTh3: acc = direct_f64_range<kTh3>(tile, g.pa, g.pb, g.src, g.w, j0, j1, tx, ty, tz); break;
    default: acc = direct_f64_range<kExp>(tile, g.pa, g.pb, g.src, g.w, j0, j1, tx, ty, tz); break;
  }
  if (i < g.q_n) partial[g.part_off + static_cast<int64_t>(by) * g.q_n + i] = acc;
}

// sum_k c_k m_k(x, y, z) over the monomials of degree 0..2
// (polynomial/monomial_basis.hpp:198-209 order: 1, x, y, z, x^2, xy, xz, y^2, yz, z^2).
__device__ __forceinline__ double poly_value(int degree, const double* __restrict__ c, double x, double y,
                                             double z) {
  double acc = c[0];
  if (degree >= 1) acc += c[1] * x + c[2] * y + c[3] * z;
  if (degree >= 2) acc += c[4] * x * x + c[5] * x * y + c[6] * x * z + c[7] * y * y + c[8] * y * z + c[9] * z * z;
  return acc;
}

__global__ void poly_eval_kernel(int degree, const double* __restrict__ pts, int64_t n,
                                 const double* __restrict__ c, double* __restrict__ out) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] += poly_value(degree, c, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

// Gradient of the polynomial part at untransformed points:
// out3[3 i + k] += d/dx_k sum_m c_m m(p_i) (monomial_basis.hpp:198-209 order).
__global__ void poly_grad_kernel(int degree, const double* __restrict__ pts, int64_t n,
                                 const double* __restrict__ c, double* __restrict__ out3) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (degree < 1) return;
  const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
  double gx = c[1], gy = c[2], gz = c[3];
  if (degree >= 2) {
    gx += 2.0 * c[4] * x + c[5] * y + c[6] * z;
    gy += c[5] * x + 2.0 * c[7] * y + c[8] * z;
    gz += c[6] * x + c[8] * y + 2.0 * c[9] * z;
  }
  out3[3 * i] += gx;
  out3[3 * i + 1] += gy;
  out3[3 * i + 2] += gz;
}

// One thread per query row: each segment's chunks summed in chunk order, the
// segments of a query set added in rbf order, then its polynomial.
__global__ void direct_batch_reduce_kernel(const DirQuery* __restrict__ qs, int nq,
                                           const DirSeg* __restrict__ segs, const double* __restrict__ pts,
                                           const double* __restrict__ partial, double* __restrict__ out,
                                           int64_t rows) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= rows) return;
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (qs[mid].off <= t) lo = mid;
    else hi = mid - 1;
  }
  const DirQuery& q = qs[lo];
  const int64_t i = t - q.off;
  double v = 0.0;
  for (int k = 0; k < q.n_seg; k++) {
    const DirSeg& g = segs[q.seg_first + k];
    double acc = 0.0;
    for (int c = 0; c < g.nchunks; c++) acc += partial[g.part_off + static_cast<int64_t>(c) * g.q_n + i];
    v = k == 0 ? acc : v + acc;
  }
  if (q.degree >= 0) v += poly_value(q.degree, q.coeffs, pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]);
  out[t] = v;
}

}  // namespace

void launch_transform_points(const double* aniso, const double* in, int64_t n, double* out,
                             hipStream_t s) {
  if (n == 0) return;
  Mat3 m;
  for (int i = 0; i < 9; i++) m.a[i] = aniso[i];
  transform_kernel<<<ceil_div(n, 256), 256, 0, s>>>(m, in, n, out);
  PAM_LAUNCH_CHECK();
}

void launch_bbox(const double* pts, int64_t n, double* d_out, DevBuf<double>& scratch,
                 hipStream_t s) {
  int nb = ceil_div(n, kBboxThreads);
  if (nb > 512) nb = 512;
  if (nb < 1) nb = 1;
  scratch.resize(static_cast<size_t>(nb) * 6);
  bbox_partial_kernel<<<nb, kBboxThreads, 0, s>>>(pts, n, scratch.get());
  PAM_LAUNCH_CHECK();
  bbox_final_kernel<<<1, kBboxThreads, 0, s>>>(scratch.get(), nb, d_out);
  PAM_LAUNCH_CHECK();
}

void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, DevBuf<int32_t>& scratch,
                        hipStream_t s) {
  if (n <= 0) return;
  int64_t need = 0;
  for (int64_t m = n; m > 1;) {
    m = (m + kScanTile - 1) / kScanTile;
    need += m;
  }
  scratch.resize(static_cast<size_t>(need + 1));
  scan_rec(in, out, n, scratch.get(), s);
}

void radix_sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int bits, DevBuf<uint64_t>& tk,
                      DevBuf<int32_t>& tv, DevBuf<int32_t>& hist, DevBuf<int32_t>& scratch,
                      hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  int nb = ceil_div(n, kRsTile);
  tk.resize(static_cast<size_t>(n));
  tv.resize(static_cast<size_t>(n));
  hist.resize(static_cast<size_t>(16) * nb);
  uint64_t* ka = keys;
  int32_t* va = vals;
  uint64_t* kb = tk.get();
  int32_t* vb = tv.get();
  int passes = (bits + 3) / 4;
  for (int p = 0; p < passes; p++) {
    int shift = 4 * p;
    radix_hist_kernel<<<nb, kRsThreads, 0, s>>>(ka, n, shift, hist.get(), nb);
    PAM_LAUNCH_CHECK();
    exclusive_scan_i32(hist.get(), hist.get(), static_cast<int64_t>(16) * nb, scratch, s);
    radix_scatter_kernel<<<nb, kRsThreads, 0, s>>>(ka, va, kb, vb, n, shift, hist.get(), nb);
    PAM_LAUNCH_CHECK();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    PAM_HIP(hipMemcpyAsync(keys, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
    PAM_HIP(hipMemcpyAsync(vals, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  }
}

void launch_iota(int32_t* vals, int64_t n, hipStream_t s) {
  if (n == 0) return;
  iota_kernel<<<ceil_div(n, 256), 256, 0, s>>>(vals, n);
  PAM_LAUNCH_CHECK();
}

void launch_gather_f64(const double* in, const int32_t* idx, int64_t n, double* out, hipStream_t s) {
  if (n == 0) return;
  gather_f64_kernel<<<ceil_div(n, 256), 256, 0, s>>>(in, idx, n, out);
  PAM_LAUNCH_CHECK();
}

namespace {

// Targets in blocks of kDirThreads; sources split so that the grid has enough
// workgroups to fill 256 CUs, at least 1024 sources per chunk.
void direct_split(int64_t n_trg, int64_t n_src, int* gx, int64_t* chunk, int* nchunks) {
  *gx = ceil_div(n_trg, kDirThreads);
  int64_t want_chunks = (2048 + *gx - 1) / *gx;
  int64_t c = n_src / (want_chunks > 0 ? want_chunks : 1);
  if (c < 1024) c = 1024;
  *chunk = c;
  *nchunks = static_cast<int>(n_src > 0 ? (n_src + c - 1) / c : 1);
}

}  // namespace

size_t direct_partial_size(int64_t n_trg, int64_t n_src) {
  if (n_trg == 0) return 0;
  int gx, nchunks;
  int64_t chunk;
  direct_split(n_trg, n_src, &gx, &chunk, &nchunks);
  return static_cast<size_t>(nchunks) * n_trg;
}

void direct_sum_f64_into(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                         const double* trg, int64_t n_trg, double* out, bool accumulate, double* partial,
                         hipStream_t s) {
  if (n_trg == 0) return;
  PhiConsts pc = phi_consts(rbf, 1.0);
  int gx, nchunks;
  int64_t chunk;
  direct_split(n_trg, n_src, &gx, &chunk, &nchunks);
  if (n_src == 0) {
    fill_f64_kernel<<<ceil_div(n_trg, 256), 256, 0, s>>>(partial, n_trg, 0.0);
    PAM_LAUNCH_CHECK();
  } else {
    dim3 grid(gx, nchunks);
    dispatch_kind(rbf.kind, [&](auto k) {
      direct_f64_kernel<decltype(k)::value><<<grid, kDirThreads, 0, s>>>(pc.a, pc.b, src, w, n_src, trg,
                                                                        n_trg, partial, chunk);
    });
    PAM_LAUNCH_CHECK();
  }
  reduce_partial_kernel<<<ceil_div(n_trg, 256), 256, 0, s>>>(partial, nchunks, n_trg, out,
                                                             accumulate ? 1 : 0);
  PAM_LAUNCH_CHECK();
}

void direct_sum_f64(const RbfDesc& rbf, const double* src, const double* w, int64_t n_src,
                    const double* trg, int64_t n_trg, double* out, bool accumulate,
                    DevBuf<double>& scratch, hipStream_t s) {
  if (n_trg == 0) return;
  scratch.resize(direct_partial_size(n_trg, n_src));
  direct_sum_f64_into(rbf, src, w, n_src, trg, n_trg, out, accumulate, scratch.get(), s);
}

void launch_poly_grad(int degree, const double* pts, int64_t n, const double* d_coeffs, double* out3,
                      hipStream_t s) {
  if (n == 0 || degree < 1) return;
  poly_grad_kernel<<<ceil_div(n, 256), 256, 0, s>>>(degree, pts, n, d_coeffs, out3);
  PAM_LAUNCH_CHECK();
}

void launch_poly_eval(int degree, const double* pts, int64_t n, const double* d_coeffs,
                      double* out, hipStream_t s) {
  if (n == 0 || degree < 0) return;
  poly_eval_kernel<<<ceil_div(n, 256), 256, 0, s>>>(degree, pts, n, d_coeffs, out);
  PAM_LAUNCH_CHECK();
}

void DirectBatch::run(const std::vector<DirectBatchQuery>& queries,
                      const std::vector<DirectBatchTerm>& terms, hipStream_t s) {
  std::vector<DirSeg> segs;
  std::vector<DirQuery> qs;
  int64_t rows = 0, part = 0, work = 0;
  for (const DirectBatchQuery& q : queries) {
    if (q.n < 0) throw_invalid("negative query count");
    if (q.n == 0) continue;
    if (!q.pts || !q.out) throw_invalid("null query or output buffer");
    if (q.n_terms < 1 || q.first_term < 0 || q.first_term + q.n_terms > static_cast<int>(terms.size()))
      throw_invalid("query terms out of range");
    if (q.poly_degree > 2) throw_invalid("polynomial degree must be at most 2");
    if (q.poly_degree >= 0 && !q.coeffs) throw_invalid("null polynomial coefficients");
    DirQuery d{};
    d.off = rows;
    d.n = q.n;
    d.seg_first = static_cast<int32_t>(segs.size());
    d.n_seg = q.n_terms;
    d.degree = q.poly_degree;
    d.coeffs = q.coeffs;
    qs.push_back(d);
    for (int k = q.first_term; k < q.first_term + q.n_terms; k++) {
      const DirectBatchTerm& t = terms[k];
      if (t.n_src < 1 || !t.src || !t.w) throw_invalid("a batched term needs sources and weights");
      DirSeg g{};
      g.src = t.src;
      g.w = t.w;
      g.n_src = t.n_src;
      direct_split(q.n, t.n_src, &g.gx, &g.chunk, &g.nchunks);
      g.part_off = part;
      g.q_off = rows;
      g.q_n = q.n;
      g.kind = t.rbf.kind;
      g.work_first = static_cast<int32_t>(work);
      PhiConsts pc = phi_consts(t.rbf, 1.0);
      g.pa = pc.a;
      g.pb = pc.b;
      for (int e = 0; e < 9; e++) g.a[e] = t.rbf.aniso[e];
      segs.push_back(g);
      part += static_cast<int64_t>(g.nchunks) * q.n;
      work += static_cast<int64_t>(g.gx) * g.nchunks;
      if (work > (int64_t{1} << 30)) throw_invalid("batched direct sum too large for one launch");
    }
    rows += q.n;
  }
  if (rows == 0) return;
  // One upload: segment table, query table, then the query points.
  const size_t seg_d = segs.size() * sizeof(DirSeg) / sizeof(double);
  const size_t q_d = qs.size() * sizeof(DirQuery) / sizeof(double);
  const size_t total = seg_d + q_d + static_cast<size_t>(rows) * 3;
  double* h = up_.acquire(total);
  std::memcpy(h, segs.data(), segs.size() * sizeof(DirSeg));
  std::memcpy(h + seg_d, qs.data(), qs.size() * sizeof(DirQuery));
  double* hp = h + seg_d + q_d;
  for (const DirectBatchQuery& q : queries) {
    if (q.n == 0) continue;
    std::memcpy(hp, q.pts, sizeof(double) * 3 * static_cast<size_t>(q.n));
    hp += 3 * q.n;
  }
  up_.upload(d_in_, total, s);
  const DirSeg* d_segs = reinterpret_cast<const DirSeg*>(d_in_.get());
  const DirQuery* d_qs = reinterpret_cast<const DirQuery*>(d_in_.get() + seg_d);
  const double* d_pts = d_in_.get() + seg_d + q_d;
  d_part_.resize(static_cast<size_t>(part));
  d_out_.resize(static_cast<size_t>(rows));
  direct_batch_kernel<<<static_cast<unsigned>(work), kDirThreads, 0, s>>>(d_segs, static_cast<int>(segs.size()),
                                                                          d_pts, d_part_.get());
  PAM_LAUNCH_CHECK();
  direct_batch_reduce_kernel<<<ceil_div(rows, 256), 256, 0, s>>>(d_qs, static_cast<int>(qs.size()), d_segs, d_pts,
                                                                d_part_.get(), d_out_.get(), rows);
  PAM_LAUNCH_CHECK();
  double* ho = down_.acquire(static_cast<size_t>(rows));
  PAM_HIP(hipMemcpyAsync(ho, d_out_.get(), sizeof(double) * rows, hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  for (const DirectBatchQuery& q : queries) {
    if (q.n == 0) continue;
    std::memcpy(q.out, ho, sizeof(double) * static_cast<size_t>(q.n));
    ho += q.n;
  }
}

}  // namespace pam
