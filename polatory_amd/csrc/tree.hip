// Octree construction on the device: Morton codes, stable radix sort, leaf
// ranges, coarser levels by shift+unique, and the P2P/M2L interaction lists.
#include <algorithm>
#include <mutex>

#include "device_api.hpp"
#include "tree.hpp"

namespace pam {

namespace {

__device__ __forceinline__ uint64_t spread3(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__device__ __forceinline__ uint64_t compact3(uint64_t v) {
  v &= 0x1249249249249249ull;
  v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
  v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
  v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
  v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
  v = (v ^ (v >> 32)) & 0x1fffffull;
  return v;
}

__device__ __forceinline__ uint64_t encode3(int64_t x, int64_t y, int64_t z) {
  return (spread3(static_cast<uint64_t>(x)) << 2) | (spread3(static_cast<uint64_t>(y)) << 1) |
         spread3(static_cast<uint64_t>(z));
}

__device__ __forceinline__ void decode3(uint64_t c, int64_t* x, int64_t* y, int64_t* z) {
  *x = static_cast<int64_t>(compact3(c >> 2));
  *y = static_cast<int64_t>(compact3(c >> 1));
  *z = static_cast<int64_t>(compact3(c));
}

__device__ __forceinline__ int64_t find_code(const uint64_t* codes, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (codes[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < n && codes[lo] == key) ? lo : -1;
}

struct Lo3 {
  double v[3];
};

__global__ void morton_kernel(const double* __restrict__ pts, int64_t n, Lo3 lo, double inv_cell,
                              int64_t side_max, uint64_t* __restrict__ codes, int32_t* __restrict__ idx) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c[3];
  for (int k = 0; k < 3; k++) {
    double f = floor(__dmul_rn(__dsub_rn(pts[3 * i + k], lo.v[k]), inv_cell));
    c[k] = (f < 0.0) ? 0 : (f > static_cast<double>(side_max) ? side_max : static_cast<int64_t>(f));
  }
  codes[i] = encode3(c[0], c[1], c[2]);
  idx[i] = static_cast<int32_t>(i);
}

// Run detection over the first n entries of a buffer of n_bound entries; n is
// read from the device (d_n) so that coarser levels need no host round trip.
__device__ __forceinline__ int64_t count_of(const int64_t* d_n, int64_t n_bound) {
  return d_n ? *d_n : n_bound;
}

// flags[i] = 1 where (codes[i] >> shift) starts a new run (0 beyond n)
__global__ void run_flags_kernel(const uint64_t* __restrict__ codes, int64_t n_bound,
                                 const int64_t* __restrict__ d_n, int shift, int32_t* __restrict__ flags) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_bound) return;
  const int64_t n = count_of(d_n, n_bound);
  flags[i] = (i < n && (i == 0 || (codes[i] >> shift) != (codes[i - 1] >> shift))) ? 1 : 0;
}

// For each run start: out_codes[r] = code >> shift, out_start[r] = i; member[i] = run index.
__global__ void run_emit_kernel(const uint64_t* __restrict__ codes, int64_t n_bound,
                                const int64_t* __restrict__ d_n, int shift,
                                const int32_t* __restrict__ flags, const int32_t* __restrict__ excl,
                                uint64_t* __restrict__ out_codes, int32_t* __restrict__ out_start,
                                int32_t* __restrict__ member) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count_of(d_n, n_bound)) return;
  int32_t r = excl[i] + flags[i] - 1;
  if (flags[i]) {
    out_codes[r] = codes[i] >> shift;
    out_start[r] = static_cast<int32_t>(i);
  }
  member[i] = r;
}

// runs = number of run starts; out_start[runs] = n.
__global__ void run_count_kernel(const int32_t* __restrict__ flags, const int32_t* __restrict__ excl,
                                 int64_t n_bound, const int64_t* __restrict__ d_n,
                                 int32_t* __restrict__ out_start, int64_t* __restrict__ d_runs) {
  const int64_t n = count_of(d_n, n_bound);
  int64_t runs = n > 0 ? static_cast<int64_t>(excl[n - 1]) + flags[n - 1] : 0;
  *d_runs = runs;
  out_start[runs] = static_cast<int32_t>(n);
}

// One thread per (target leaf, neighbour slot).
__global__ void p2p_list_kernel(const uint64_t* __restrict__ tcodes, int64_t nt,
                                const uint64_t* __restrict__ scodes, int64_t ns, int64_t side,
                                int32_t* __restrict__ out) {
  int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t t = gid / 27;
  int s = static_cast<int>(gid % 27);
  if (t >= nt) return;
  int64_t x, y, z;
  decode3(tcodes[t], &x, &y, &z);
  int64_t qx = x + s / 9 - 1, qy = y + (s / 3) % 3 - 1, qz = z + s % 3 - 1;
  int32_t f = -1;
  if (qx >= 0 && qx < side && qy >= 0 && qy < side && qz >= 0 && qz < side)
    f = static_cast<int32_t>(find_code(scodes, ns, encode3(qx, qy, qz)));
  out[t * 27 + s] = f;
}

// Per-level inputs of the list kernels that cover all M2L levels in one launch
// (blockIdx.y = level - 2).
struct LevelLists {
  int n_levels;
  int64_t nt[kMaxTreeLevel + 1], ns[kMaxTreeLevel + 1], col_bound[kMaxTreeLevel + 1],
      blk_bound[kMaxTreeLevel + 1];
  const uint64_t* tcodes[kMaxTreeLevel + 1];
  const uint64_t* scodes[kMaxTreeLevel + 1];
  int32_t* src[kMaxTreeLevel + 1];
  int32_t* col_cell[kMaxTreeLevel + 1];
  int32_t* blk_oct[kMaxTreeLevel + 1];
  int64_t* lay[kMaxTreeLevel + 1];
  int32_t* v_count[kMaxTreeLevel + 1];
  const int32_t* v_first[kMaxTreeLevel + 1];
  int32_t* blk_src[kMaxTreeLevel + 1];
  int32_t* pair_pos[kMaxTreeLevel + 1];
};

// Pair kernels: a workgroup takes kPairCells target cells (all their slots)
// of one level (blockIdx.y = level - 2) and counts its pairs per transfer
// vector in LDS, so that only one global atomic per vector and workgroup is
// needed.
constexpr int kPairCells = 32;
constexpr int kPairIters = (kPairCells * kM2LSlots + 255) / 256;

// Pairs per transfer vector.
__global__ void __launch_bounds__(256) pair_count_kernel(LevelLists a, const int32_t* __restrict__ vlist) {
  __shared__ int32_t h[343];
  const int l = static_cast<int>(blockIdx.y) + 2;
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kPairCells;
  if (t0 >= a.nt[l]) return;
  for (int v = threadIdx.x; v < 343; v += 256) h[v] = 0;
  __syncthreads();
  for (int it = 0; it < kPairIters; it++) {
    const int e = it * 256 + static_cast<int>(threadIdx.x);
    const int64_t t = t0 + e / kM2LSlots;
    const int k = e % kM2LSlots;
    if (e >= kPairCells * kM2LSlots || t >= a.nt[l]) continue;
    if (a.src[l][t * kM2LSlots + k] < 0) continue;
    const int oct = static_cast<int>(a.tcodes[l][t] & 7u);
    atomicAdd(&h[vlist[oct * kM2LSlots + k]], 1);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < 343; v += 256)
    if (h[v] != 0) atomicAdd(&a.v_count[l][v], h[v]);
}

// Places every pair into its transfer vector's blocks and records its index
// (-1 for slots without a source): ranks inside the workgroup from LDS
// atomics, the workgroup's range from one global atomic per vector on the
// fill counters (v_count, zeroed). The order inside a group is immaterial.
__global__ void __launch_bounds__(256) pair_scatter_kernel(LevelLists a, const int32_t* __restrict__ vlist) {
  __shared__ int32_t h[343];
  __shared__ int32_t base[343];
  const int l = static_cast<int>(blockIdx.y) + 2;
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kPairCells;
  if (t0 >= a.nt[l]) return;
  for (int v = threadIdx.x; v < 343; v += 256) h[v] = 0;
  __syncthreads();
  int32_t vv[kPairIters], rank[kPairIters];
#pragma unroll
  for (int it = 0; it < kPairIters; it++) {
    const int e = it * 256 + static_cast<int>(threadIdx.x);
    const int64_t t = t0 + e / kM2LSlots;
    const int k = e % kM2LSlots;
    vv[it] = -2;  // outside the tile
    rank[it] = 0;
    if (e >= kPairCells * kM2LSlots || t >= a.nt[l]) continue;
    vv[it] = -1;  // no source
    if (a.src[l][t * kM2LSlots + k] < 0) continue;
    const int oct = static_cast<int>(a.tcodes[l][t] & 7u);
    vv[it] = vlist[oct * kM2LSlots + k];
    rank[it] = atomicAdd(&h[vv[it]], 1);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < 343; v += 256)
    base[v] = h[v] != 0 ? a.v_first[l][v] * kM2LCols + atomicAdd(&a.v_count[l][v], h[v]) : 0;
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kPairIters; it++) {
    if (vv[it] == -2) continue;
    const int e = it * 256 + static_cast<int>(threadIdx.x);
    const int64_t gid = t0 * kM2LSlots + e;
    int32_t pos = -1;
    if (vv[it] >= 0) {
      pos = base[vv[it]] + rank[it];
      a.blk_src[l][pos] = a.src[l][gid];
    }
    a.pair_pos[l][gid] = pos;
  }
}

__global__ void m2l_list_kernel(LevelLists a, const int32_t* __restrict__ vlist) {
  const int l = static_cast<int>(blockIdx.y) + 2;
  const uint64_t* __restrict__ tcodes = a.tcodes[l];
  const uint64_t* __restrict__ scodes = a.scodes[l];
  const int64_t nt = a.nt[l], ns = a.ns[l], side = int64_t{1} << l;
  int32_t* __restrict__ out = a.src[l];
  int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t t = gid / kM2LSlots;
  int k = static_cast<int>(gid % kM2LSlots);
  if (t >= nt) return;
  uint64_t code = tcodes[t];
  int64_t x, y, z;
  decode3(code, &x, &y, &z);
  int oct = static_cast<int>(code & 7u);
  int v = vlist[oct * kM2LSlots + k];
  int64_t qx = x + v / 49 - 3, qy = y + (v / 7) % 7 - 3, qz = z + v % 7 - 3;
  int32_t f = -1;
  if (qx >= 0 && qx < side && qy >= 0 && qy < side && qz >= 0 && qz < side)
    f = static_cast<int32_t>(find_code(scodes, ns, encode3(qx, qy, qz)));
  out[t * kM2LSlots + k] = f;
}

// Column layout of one level: octant groups in octant order, each padded to a
// multiple of kM2LCols; inside a group the cells keep their (Morton) order.
// lay[0..7] = number of cells of the octants before each group, lay[8..15] =
// its first column, lay[16] = number of columns. col_cell[c] = target cell of
// column c or -1; blk_oct[b] = octant of column block b (-1 beyond the
// layout). One workgroup per level.
constexpr int kLayout = 17;
constexpr int kColThreads = 1024;

__global__ void __launch_bounds__(kColThreads) octant_columns_kernel(LevelLists a) {
  __shared__ int32_t cnt[8];
  __shared__ int64_t first[8], col0[8], total_cols;
  __shared__ int32_t wcnt[kColThreads / 64][8];
  const int l = static_cast<int>(blockIdx.x) + 2;
  const int64_t n = a.nt[l];
  const uint64_t* __restrict__ codes = a.tcodes[l];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < 8) cnt[tid] = 0;
  __syncthreads();
  for (int64_t i = tid; i < n; i += kColThreads) atomicAdd(&cnt[codes[i] & 7u], 1);
  __syncthreads();
  if (tid == 0) {
    int64_t start = 0, col = 0;
    for (int k = 0; k < 8; k++) {
      first[k] = start;
      col0[k] = col;
      start += cnt[k];
      col += (cnt[k] + kM2LCols - 1) / kM2LCols * kM2LCols;
    }
    total_cols = col;
    for (int k = 0; k < 8; k++) {
      a.lay[l][k] = first[k];
      a.lay[l][8 + k] = col0[k];
    }
    a.lay[l][16] = col;
  }
  __syncthreads();
  // Padding columns (past each group's cells, and past the layout) are -1;
  // they are disjoint from the columns the scatter below writes.
  for (int64_t c = tid; c < a.col_bound[l]; c += kColThreads) {
    bool pad = c >= total_cols;
    for (int k = 0; k < 8; k++)
      if (c >= col0[k] + cnt[k] && c < (k < 7 ? col0[k + 1] : total_cols)) pad = true;
    if (pad) a.col_cell[l][c] = -1;
  }
  // Stable scatter: a cell's column is its octant's first column plus the
  // number of earlier cells of the same octant.
  int64_t run[8];
#pragma unroll
  for (int k = 0; k < 8; k++) run[k] = col0[k];
  const uint64_t lt = (uint64_t{1} << lane) - 1;
  for (int64_t base = 0; base < n; base += kColThreads) {
    const int64_t i = base + tid;
    const int oct = i < n ? static_cast<int>(codes[i] & 7u) : -1;
    int rank = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint64_t m = __ballot(oct == k);
      if (oct == k) rank = __popcll(m & lt);
      if (lane == 0) wcnt[wave][k] = __popcll(m);
    }
    __syncthreads();
    if (oct >= 0) {
      int64_t pos = 0;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (k != oct) continue;
        pos = run[k];
        for (int w = 0; w < wave; w++) pos += wcnt[w][k];
      }
      a.col_cell[l][pos + rank] = static_cast<int32_t>(i);
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      for (int w = 0; w < kColThreads / 64; w++) run[k] += wcnt[w][k];
    __syncthreads();
  }
  for (int64_t b = tid; b < a.blk_bound[l]; b += kColThreads) {
    const int64_t c = b * kM2LCols;
    int oct = -1;
    if (c < total_cols) {
      for (int k = 0; k < 8; k++)
        if (c >= col0[k]) oct = k;
    }
    a.blk_oct[l][b] = oct;
  }
}

__global__ void fill_i32_kernel(int32_t* a, int64_t n, int32_t v) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Unique runs of (codes >> shift) over the first *d_n of n_bound entries
// (d_n == nullptr: all n_bound); the run count goes to *d_runs on the device.
// Output buffers are sized by n_bound.
void unique_runs(const uint64_t* codes, int64_t n_bound, const int64_t* d_n, int shift,
                 DevBuf<uint64_t>& out_codes, DevBuf<int32_t>& out_start, DevBuf<int32_t>& member,
                 int64_t* d_runs, TreeWorkspace& ws, hipStream_t s) {
  ws.flags.resize(n_bound);
  ws.scan.resize(n_bound);
  member.resize(n_bound);
  out_codes.resize(n_bound);
  out_start.resize(n_bound + 1);
  run_flags_kernel<<<ceil_div(n_bound, 256), 256, 0, s>>>(codes, n_bound, d_n, shift, ws.flags.get());
  PAM_LAUNCH_CHECK();
  exclusive_scan_i32(ws.flags.get(), ws.scan.get(), n_bound, ws.scan_scratch, s);
  run_emit_kernel<<<ceil_div(n_bound, 256), 256, 0, s>>>(codes, n_bound, d_n, shift, ws.flags.get(),
                                                         ws.scan.get(), out_codes.get(), out_start.get(),
                                                         member.get());
  PAM_LAUNCH_CHECK();
  run_count_kernel<<<1, 1, 0, s>>>(ws.flags.get(), ws.scan.get(), n_bound, d_n, out_start.get(), d_runs);
  PAM_LAUNCH_CHECK();
}

std::vector<int32_t> make_octant_vlist() {
  std::vector<int32_t> t;
  t.reserve(8 * kM2LSlots);
  for (int oct = 0; oct < 8; oct++) {
    int bx = (oct >> 2) & 1, by = (oct >> 1) & 1, bz = oct & 1;
    for (int v = 0; v < 343; v++) {
      int dx = v / 49 - 3, dy = (v / 7) % 7 - 3, dz = v % 7 - 3;
      bool in = dx >= -2 - bx && dx <= 3 - bx && dy >= -2 - by && dy <= 3 - by && dz >= -2 - bz &&
                dz <= 3 - bz;
      bool adj = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
      if (in && !adj) t.push_back(v);
    }
  }
  return t;
}

}  // namespace

const std::vector<int32_t>& octant_vlist_host() {
  static const std::vector<int32_t> t = make_octant_vlist();
  return t;
}

const int32_t* octant_vlist_device(hipStream_t s) {
  static std::mutex mu;
  static int32_t* per_dev[64] = {};
  int dev = 0;
  PAM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if (!per_dev[dev]) {
    const auto& h = octant_vlist_host();
    PAM_HIP(hipMalloc(reinterpret_cast<void**>(&per_dev[dev]), h.size() * sizeof(int32_t)));
    PAM_HIP(hipMemcpyAsync(per_dev[dev], h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PAM_HIP(hipStreamSynchronize(s));
  }
  return per_dev[dev];
}

namespace {

// Level at which sorted point i first opens a cell: 0 for the first point,
// else sort_level - floor(msb(c[i] ^ c[i-1]) / 3) (the codes of level l are
// the top 3l bits); points equal to their predecessor open none.
__global__ void level_open_kernel(const uint64_t* __restrict__ codes, int64_t n, int sort_level,
                                  int64_t* __restrict__ hist) {
  __shared__ unsigned int h[kMaxTreeLevel + 1];
  if (threadIdx.x <= kMaxTreeLevel) h[threadIdx.x] = 0u;
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int lv;
    if (i == 0) {
      lv = 0;
    } else {
      uint64_t x = codes[i] ^ codes[i - 1];
      if (x == 0) continue;
      lv = sort_level - (63 - __clzll(static_cast<long long>(x))) / 3;
    }
    atomicAdd(&h[lv], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= static_cast<unsigned>(sort_level) && h[threadIdx.x] != 0u)
    atomicAdd(reinterpret_cast<unsigned long long*>(hist + threadIdx.x),
              static_cast<unsigned long long>(h[threadIdx.x]));
}

__global__ void zero_i64_kernel(int64_t* a, int n) {
  if (static_cast<int>(threadIdx.x) < n) a[threadIdx.x] = 0;
}

__global__ void shift_codes_kernel(uint64_t* codes, int64_t n, int shift) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) codes[i] >>= shift;
}

// Levels level-1 .. 0 from the leaf codes in one workgroup: for each level,
// run starts of (child code >> 3) give the cells (codes, child_start) and
// every child's parent index. Counts are known, so every write is bounded.
struct CoarseArgs {
  int level;
  int64_t n[kMaxTreeLevel + 1];
  uint64_t* codes[kMaxTreeLevel + 1];
  int32_t* parent[kMaxTreeLevel + 1];
  int32_t* child_start[kMaxTreeLevel + 1];
};

constexpr int kCoarseThreads = 1024;
constexpr int64_t kCoarseMaxLeaves = 65536;

__global__ void __launch_bounds__(kCoarseThreads) coarse_levels_kernel(CoarseArgs a) {
  __shared__ int32_t wsum[kCoarseThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int l = a.level - 1; l >= 0; l--) {
    const uint64_t* cc = a.codes[l + 1];
    const int64_t nc = a.n[l + 1], np = a.n[l];
    int32_t carry = 0;
    for (int64_t base = 0; base < nc; base += kCoarseThreads) {
      const int64_t i = base + tid;
      int32_t f = 0;
      uint64_t c = 0;
      if (i < nc) {
        c = cc[i] >> 3;
        f = (i == 0 || c != (cc[i - 1] >> 3)) ? 1 : 0;
      }
      int32_t v = f;  // inclusive scan within the wave
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        int32_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
      }
      if (lane == 63) wsum[wave] = v;
      __syncthreads();
      int32_t before = 0, total = 0;
      for (int w = 0; w < kCoarseThreads / 64; w++) {
        if (w < wave) before += wsum[w];
        total += wsum[w];
      }
      const int32_t r = carry + before + v - 1;  // cell index of child i
      if (i < nc && r >= 0 && r < np) {
        if (f) {
          a.codes[l][r] = c;
          a.child_start[l][r] = static_cast<int32_t>(i);
        }
        a.parent[l + 1][i] = r;
      }
      carry += total;
      __syncthreads();
    }
    if (tid == 0) a.child_start[l][np] = static_cast<int32_t>(nc);
    // This level's codes are the next level's input.
    __threadfence();
    __syncthreads();
  }
}

}  // namespace

void sort_tree_points(Tree& t, const double* d_pts, int64_t n, const Box& box, int sort_level,
                      TreeWorkspace& ws, hipStream_t s) {
  if (sort_level < 0 || sort_level > kMaxTreeLevel) throw_invalid("tree level out of range");
  t.box = box;
  t.n_points = n;
  t.sort_level = sort_level;
  t.level = -1;
  t.open_hist.resize(static_cast<size_t>(sort_level) + 1);
  zero_i64_kernel<<<1, 64, 0, s>>>(t.open_hist.get(), sort_level + 1);
  PAM_LAUNCH_CHECK();
  if (n == 0) return;
  t.point_codes.resize(n);
  t.perm.resize(n);
  Lo3 lo{{box.lo[0], box.lo[1], box.lo[2]}};
  const double inv_cell = std::ldexp(1.0, sort_level) / box.width;
  morton_kernel<<<ceil_div(n, 256), 256, 0, s>>>(d_pts, n, lo, inv_cell,
                                                 (int64_t{1} << sort_level) - 1, t.point_codes.get(),
                                                 t.perm.get());
  PAM_LAUNCH_CHECK();
  radix_sort_pairs(t.point_codes.get(), t.perm.get(), n, 3 * sort_level, ws.tk, ws.tv, ws.hist,
                   ws.scan_scratch, s);
  const int nb = std::min(ceil_div(n, 256 * 16), 1024);
  level_open_kernel<<<nb, 256, 0, s>>>(t.point_codes.get(), n, sort_level, t.open_hist.get());
  PAM_LAUNCH_CHECK();
}

std::vector<int64_t> level_counts(const std::vector<int64_t>& open_hist) {
  std::vector<int64_t> c(open_hist.size());
  int64_t acc = 0;
  for (size_t l = 0; l < c.size(); l++) c[l] = acc += open_hist[l];
  return c;
}

void finish_tree(Tree& t, int level, const std::vector<int64_t>& counts, TreeWorkspace& ws,
                 hipStream_t s) {
  if (level < 0 || level > t.sort_level) throw_invalid("finish_tree: level beyond the sort level");
  if (static_cast<int>(counts.size()) <= level) throw_invalid("finish_tree: missing level counts");
  const int64_t n = t.n_points;
  t.level = level;
  t.inv_cell = std::ldexp(1.0, level) / t.box.width;
  // Level buffers are kept and only grown: freeing device memory would
  // synchronize the whole device (and every stream of other evaluators).
  if (t.levels.size() < static_cast<size_t>(level) + 1) t.levels.resize(level + 1);
  if (n == 0) {
    for (int l = 0; l <= level; l++) {
      TreeLevel& lv = t.levels[l];
      lv.n = 0;
      lv.child_start.resize(1);
      lv.child_start.zero(s);
    }
    t.leaf_start.resize(1);
    t.leaf_start.zero(s);
    return;
  }
  // Morton codes of level l are the level-L codes shifted right by 3(L - l):
  // the cell index floor((x - lo) * 2^L / w) scales by an exact power of two.
  if (t.sort_level > level) {
    shift_codes_kernel<<<ceil_div(n, 256), 256, 0, s>>>(t.point_codes.get(), n, 3 * (t.sort_level - level));
    PAM_LAUNCH_CHECK();
    t.sort_level = level;
  }
  ws.counts.resize(static_cast<size_t>(level) + 1);
  for (int l = 0; l <= level; l++) t.levels[l].n = counts[l];
  unique_runs(t.point_codes.get(), n, nullptr, 0, t.levels[level].codes, t.leaf_start, t.point_leaf,
              ws.counts.get() + level, ws, s);
  if (level == 0) return;
  if (counts[level] <= kCoarseMaxLeaves) {
    CoarseArgs a{};
    a.level = level;
    for (int l = 0; l <= level; l++) {
      TreeLevel& lv = t.levels[l];
      a.n[l] = counts[l];
      if (l < level) {
        lv.codes.resize(static_cast<size_t>(counts[l]));
        lv.child_start.resize(static_cast<size_t>(counts[l]) + 1);
        a.child_start[l] = lv.child_start.get();
      }
      if (l >= 1) {
        lv.parent.resize(static_cast<size_t>(counts[l]));
        a.parent[l] = lv.parent.get();
      }
      a.codes[l] = lv.codes.get();
    }
    coarse_levels_kernel<<<1, kCoarseThreads, 0, s>>>(a);
    PAM_LAUNCH_CHECK();
  } else {
    for (int l = level - 1; l >= 0; l--) {
      TreeLevel& child = t.levels[l + 1];
      TreeLevel& cur = t.levels[l];
      unique_runs(child.codes.get(), counts[l + 1], nullptr, 3, cur.codes, cur.child_start, child.parent,
                  ws.counts.get() + l, ws, s);
    }
  }
}

void build_tree(Tree& t, const double* d_pts, int64_t n, const Box& box, int level,
                TreeWorkspace& ws, hipStream_t s) {
  sort_tree_points(t, d_pts, n, box, level, ws, s);
  std::vector<int64_t> hist(static_cast<size_t>(level) + 1);
  PAM_HIP(hipMemcpyAsync(hist.data(), t.open_hist.get(), hist.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  finish_tree(t, level, level_counts(hist), ws, s);
}

void build_lists(const Tree& trg, const Tree& src, Lists& lists, TreeWorkspace& ws, hipStream_t s,
                 bool with_columns) {
  if (trg.level != src.level) throw_runtime("source and target trees must share the leaf level");
  int L = trg.level;
  int64_t nt = trg.n_leaves();
  lists.p2p.resize(static_cast<size_t>(nt) * 27);
  if (nt > 0) {
    p2p_list_kernel<<<ceil_div(nt * 27, 256), 256, 0, s>>>(trg.levels[L].codes.get(), nt,
                                                          src.levels[L].codes.get(), src.levels[L].n,
                                                          int64_t{1} << L, lists.p2p.get());
    PAM_LAUNCH_CHECK();
  }
  const int32_t* vl = octant_vlist_device(s);
  if (lists.m2l.size() < static_cast<size_t>(L) + 1) lists.m2l.resize(L + 1);
  lists.level = L;
  lists.has_columns = with_columns;
  // Per level: the M2L slots, then the columns grouped by octant (stable, Morton
  // order inside a group), each group padded to a multiple of kM2LCols so a
  // column block has a single octant. All levels go in one launch of each
  // kernel; the layouts come back to the host in one copy at the end.
  ws.layout.resize(static_cast<size_t>(L + 1) * kLayout);
  LevelLists a{};
  a.n_levels = L >= 2 ? L - 1 : 0;
  int64_t max_work = 0, max_cells = 0;
  for (int l = 2; l <= L; l++) {
    const TreeLevel& tl = trg.levels[l];
    const TreeLevel& sl = src.levels[l];
    M2LLevel& m = lists.m2l[l];
    const int64_t n = tl.n;
    m.src.resize(static_cast<size_t>(n) * kM2LSlots);
    m.n_cells = n;
    m.n_cols = 0;
    m.h_blk_oct.clear();
    const int64_t col_bound = n + 8 * (kM2LCols - 1);
    const int64_t blk_bound = col_bound / kM2LCols + 1;
    m.col_cell.resize(col_bound);
    m.blk_oct.resize(blk_bound);
    a.nt[l] = n;
    a.ns[l] = sl.n;
    a.col_bound[l] = col_bound;
    a.blk_bound[l] = blk_bound;
    a.tcodes[l] = tl.codes.get();
    a.scodes[l] = sl.codes.get();
    a.src[l] = m.src.get();
    a.col_cell[l] = m.col_cell.get();
    a.blk_oct[l] = m.blk_oct.get();
    a.lay[l] = ws.layout.get() + static_cast<int64_t>(l) * kLayout;
    max_work = std::max(max_work, n * kM2LSlots);
    max_cells = std::max(max_cells, n);
  }
  constexpr int kV = 343;
  lists.v_count.resize(static_cast<size_t>(L + 1) * kV);
  lists.v_first.resize(static_cast<size_t>(L + 1) * kV);
  for (int l = 2; l <= L; l++) {
    a.v_count[l] = lists.v_count.get() + static_cast<int64_t>(l) * kV;
    a.v_first[l] = lists.v_first.get() + static_cast<int64_t>(l) * kV;
  }
  if (a.n_levels > 0) {
    fill_i32_kernel<<<ceil_div(static_cast<int64_t>(L + 1) * kV, 256), 256, 0, s>>>(
        lists.v_count.get(), static_cast<int64_t>(L + 1) * kV, 0);
    PAM_LAUNCH_CHECK();
    if (max_work > 0) {
      m2l_list_kernel<<<dim3(ceil_div(max_work, 256), a.n_levels), 256, 0, s>>>(a, vl);
      PAM_LAUNCH_CHECK();
      pair_count_kernel<<<dim3(ceil_div(max_cells, kPairCells), a.n_levels), 256, 0, s>>>(a, vl);
      PAM_LAUNCH_CHECK();
    }
    if (with_columns) {
      octant_columns_kernel<<<a.n_levels, kColThreads, 0, s>>>(a);
      PAM_LAUNCH_CHECK();
    }
  }
  std::vector<int64_t> lay(static_cast<size_t>(L + 1) * kLayout, 0);
  std::vector<int32_t> vc(static_cast<size_t>(L + 1) * kV, 0);
  if (with_columns)
    PAM_HIP(hipMemcpyAsync(lay.data(), ws.layout.get(), lay.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  if (a.n_levels > 0)
    PAM_HIP(hipMemcpyAsync(vc.data(), lists.v_count.get(), vc.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  for (int l = 2; l <= L; l++) {
    M2LLevel& m = lists.m2l[l];
    m.h_v_blocks.assign(kV, 0);
    m.h_v_first.assign(kV, 0);
    m.n_blk = 0;
    if (m.n_cells == 0) continue;
    if (with_columns) {
      const int64_t* ly = lay.data() + static_cast<int64_t>(l) * kLayout;
      m.n_cols = ly[16];
      for (int k = 0; k < 8; k++) {
        int64_t end = k < 7 ? ly[8 + k + 1] : ly[16];
        for (int64_t c = ly[8 + k]; c < end; c += kM2LCols) m.h_blk_oct.push_back(k);
      }
    }
    int64_t nb = 0;
    for (int v = 0; v < kV; v++) {
      const int32_t c = vc[static_cast<size_t>(l) * kV + v];
      m.h_v_first[v] = static_cast<int32_t>(nb);
      m.h_v_blocks[v] = (c + kM2LCols - 1) / kM2LCols;
      nb += m.h_v_blocks[v];
    }
    m.n_blk = nb;
    m.blk_src.resize(static_cast<size_t>(std::max<int64_t>(nb, 1)) * kM2LCols);
    m.pair_pos.resize(static_cast<size_t>(m.n_cells) * kM2LSlots);
    a.blk_src[l] = m.blk_src.get();
    a.pair_pos[l] = m.pair_pos.get();
  }
  if (a.n_levels > 0 && max_work > 0) {
    int32_t* h = lists.v_stage.acquire(static_cast<size_t>(L + 1) * kV);
    for (int l = 0; l <= L; l++)
      for (int v = 0; v < kV; v++)
        h[static_cast<size_t>(l) * kV + v] = l >= 2 ? lists.m2l[l].h_v_first[v] : 0;
    lists.v_stage.upload(lists.v_first, static_cast<size_t>(L + 1) * kV, s);
    fill_i32_kernel<<<ceil_div(static_cast<int64_t>(L + 1) * kV, 256), 256, 0, s>>>(
        lists.v_count.get(), static_cast<int64_t>(L + 1) * kV, 0);
    PAM_LAUNCH_CHECK();
    for (int l = 2; l <= L; l++) {
      const M2LLevel& m = lists.m2l[l];
      if (m.n_blk == 0) continue;
      fill_i32_kernel<<<ceil_div(m.n_blk * kM2LCols, 256), 256, 0, s>>>(m.blk_src.get(), m.n_blk * kM2LCols, -1);
      PAM_LAUNCH_CHECK();
    }
    pair_scatter_kernel<<<dim3(ceil_div(max_cells, kPairCells), a.n_levels), 256, 0, s>>>(a, vl);
    PAM_LAUNCH_CHECK();
  }
}

}  // namespace pam
