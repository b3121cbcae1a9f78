// Octree construction on the device: Morton codes, stable radix sort, leaf
// ranges, coarser levels by shift+unique, and the P2P/M2L interaction lists.
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

// flags[i] = 1 where (codes[i] >> shift) starts a new run
__global__ void run_flags_kernel(const uint64_t* __restrict__ codes, int64_t n, int shift,
                                 int32_t* __restrict__ flags) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || (codes[i] >> shift) != (codes[i - 1] >> shift)) ? 1 : 0;
}

// For each run start: out_codes[r] = code >> shift, out_start[r] = i; member[i] = run index.
__global__ void run_emit_kernel(const uint64_t* __restrict__ codes, int64_t n, int shift,
                                const int32_t* __restrict__ flags, const int32_t* __restrict__ excl,
                                uint64_t* __restrict__ out_codes, int32_t* __restrict__ out_start,
                                int32_t* __restrict__ member) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t r = excl[i] + flags[i] - 1;
  if (flags[i]) {
    out_codes[r] = codes[i] >> shift;
    out_start[r] = static_cast<int32_t>(i);
  }
  member[i] = r;
}

__global__ void set_last_kernel(int32_t* a, int64_t idx, int32_t v) { a[idx] = v; }

__global__ void p2p_list_kernel(const uint64_t* __restrict__ tcodes, int64_t nt,
                                const uint64_t* __restrict__ scodes, int64_t ns, int64_t side,
                                int32_t* __restrict__ out) {
  int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int64_t x, y, z;
  decode3(tcodes[t], &x, &y, &z);
  for (int s = 0; s < 27; s++) {
    int64_t qx = x + s / 9 - 1, qy = y + (s / 3) % 3 - 1, qz = z + s % 3 - 1;
    int32_t f = -1;
    if (qx >= 0 && qx < side && qy >= 0 && qy < side && qz >= 0 && qz < side)
      f = static_cast<int32_t>(find_code(scodes, ns, encode3(qx, qy, qz)));
    out[t * 27 + s] = f;
  }
}

__global__ void m2l_list_kernel(const uint64_t* __restrict__ tcodes, int64_t nt,
                                const uint64_t* __restrict__ scodes, int64_t ns, int64_t side,
                                const int32_t* __restrict__ vlist, int32_t* __restrict__ out) {
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

__global__ void octant_key_kernel(const uint64_t* __restrict__ codes, int64_t n,
                                  uint64_t* __restrict__ keys, int32_t* __restrict__ idx) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = codes[i] & 7u;
  idx[i] = static_cast<int32_t>(i);
}

struct Oct8 {
  int32_t start[8];   // start of octant group in the sorted list
  int32_t cstart[8];  // start of octant group in the padded column list
};

__global__ void fill_cols_kernel(const int32_t* __restrict__ sorted_idx,
                                 const uint64_t* __restrict__ sorted_keys, int64_t n, Oct8 o,
                                 int32_t* __restrict__ col_cell) {
  int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int oct = static_cast<int>(sorted_keys[j]);
  col_cell[o.cstart[oct] + (j - o.start[oct])] = sorted_idx[j];
}

__global__ void fill_i32_kernel(int32_t* a, int64_t n, int32_t v) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Unique runs of (codes >> shift). Returns the run count.
int64_t unique_runs(const uint64_t* codes, int64_t n, int shift, DevBuf<uint64_t>& out_codes,
                    DevBuf<int32_t>& out_start, DevBuf<int32_t>& member, TreeWorkspace& ws,
                    hipStream_t s) {
  ws.flags.resize(n);
  ws.scan.resize(n);
  member.resize(n);
  run_flags_kernel<<<ceil_div(n, 256), 256, 0, s>>>(codes, n, shift, ws.flags.get());
  PAM_LAUNCH_CHECK();
  exclusive_scan_i32(ws.flags.get(), ws.scan.get(), n, ws.scan_scratch, s);
  int32_t pair[2];
  PAM_HIP(hipMemcpyAsync(&pair[0], ws.scan.get() + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipMemcpyAsync(&pair[1], ws.flags.get() + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  int64_t runs = static_cast<int64_t>(pair[0]) + pair[1];
  out_codes.resize(runs);
  out_start.resize(runs + 1);
  run_emit_kernel<<<ceil_div(n, 256), 256, 0, s>>>(codes, n, shift, ws.flags.get(), ws.scan.get(),
                                                   out_codes.get(), out_start.get(), member.get());
  PAM_LAUNCH_CHECK();
  set_last_kernel<<<1, 1, 0, s>>>(out_start.get(), runs, static_cast<int32_t>(n));
  PAM_LAUNCH_CHECK();
  return runs;
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

void build_tree(Tree& t, const double* d_pts, int64_t n, const Box& box, int level,
                TreeWorkspace& ws, hipStream_t s) {
  if (level < 0 || level > 20) throw_invalid("tree level out of range");
  t.level = level;
  t.box = box;
  t.inv_cell = std::ldexp(1.0, level) / box.width;
  t.n_points = n;
  t.levels.clear();
  t.levels.resize(level + 1);
  if (n == 0) {
    for (auto& lv : t.levels) {
      lv.n = 0;
      lv.child_start.resize(1);
      lv.child_start.zero(s);
    }
    t.leaf_start.resize(1);
    t.leaf_start.zero(s);
    return;
  }
  t.point_codes.resize(n);
  t.perm.resize(n);
  Lo3 lo{{box.lo[0], box.lo[1], box.lo[2]}};
  morton_kernel<<<ceil_div(n, 256), 256, 0, s>>>(d_pts, n, lo, t.inv_cell,
                                                 (int64_t{1} << level) - 1, t.point_codes.get(),
                                                 t.perm.get());
  PAM_LAUNCH_CHECK();
  radix_sort_pairs(t.point_codes.get(), t.perm.get(), n, 3 * level, ws.tk, ws.tv, ws.hist,
                   ws.scan_scratch, s);
  // leaves
  TreeLevel& leaves = t.levels[level];
  leaves.n = unique_runs(t.point_codes.get(), n, 0, leaves.codes, t.leaf_start, t.point_leaf, ws, s);
  // coarser levels: unique runs of the child-level codes >> 3
  for (int l = level - 1; l >= 0; l--) {
    TreeLevel& child = t.levels[l + 1];
    TreeLevel& cur = t.levels[l];
    cur.n = unique_runs(child.codes.get(), child.n, 3, cur.codes, cur.child_start, child.parent, ws, s);
  }
}

void build_lists(const Tree& trg, const Tree& src, Lists& lists, TreeWorkspace& ws, hipStream_t s) {
  if (trg.level != src.level) throw_runtime("source and target trees must share the leaf level");
  int L = trg.level;
  int64_t nt = trg.n_leaves();
  lists.p2p.resize(static_cast<size_t>(nt) * 27);
  if (nt > 0) {
    p2p_list_kernel<<<ceil_div(nt, 128), 128, 0, s>>>(trg.levels[L].codes.get(), nt,
                                                      src.levels[L].codes.get(), src.levels[L].n,
                                                      int64_t{1} << L, lists.p2p.get());
    PAM_LAUNCH_CHECK();
  }
  const int32_t* vl = octant_vlist_device(s);
  lists.m2l.clear();
  lists.m2l.resize(L + 1);
  for (int l = 2; l <= L; l++) {
    const TreeLevel& tl = trg.levels[l];
    const TreeLevel& sl = src.levels[l];
    M2LLevel& m = lists.m2l[l];
    int64_t n = tl.n;
    m.src.resize(static_cast<size_t>(n) * kM2LSlots);
    m.n_cells = n;
    m.h_blk_oct.clear();
    if (n == 0) {
      m.n_cols = 0;
      continue;
    }
    int64_t work = n * kM2LSlots;
    m2l_list_kernel<<<ceil_div(work, 256), 256, 0, s>>>(tl.codes.get(), n, sl.codes.get(), sl.n,
                                                        int64_t{1} << l, vl, m.src.get());
    PAM_LAUNCH_CHECK();
    // Columns grouped by octant (stable, Morton order inside a group), each group
    // padded to a multiple of kM2LCols so a column block has a single octant.
    ws.tmp_codes.resize(n);
    ws.tmp_idx.resize(n);
    octant_key_kernel<<<ceil_div(n, 256), 256, 0, s>>>(tl.codes.get(), n, ws.tmp_codes.get(),
                                                       ws.tmp_idx.get());
    PAM_LAUNCH_CHECK();
    radix_sort_pairs(ws.tmp_codes.get(), ws.tmp_idx.get(), n, 3, ws.tk, ws.tv, ws.hist,
                     ws.scan_scratch, s);
    std::vector<uint64_t> keys(n);
    PAM_HIP(hipMemcpyAsync(keys.data(), ws.tmp_codes.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    PAM_HIP(hipStreamSynchronize(s));
    Oct8 o{};
    int64_t cnt[8] = {};
    for (int64_t j = 0; j < n; j++) cnt[keys[j]]++;
    int64_t a = 0, c = 0;
    std::vector<int32_t> blk;
    for (int k = 0; k < 8; k++) {
      o.start[k] = static_cast<int32_t>(a);
      o.cstart[k] = static_cast<int32_t>(c);
      int64_t padded = (cnt[k] + kM2LCols - 1) / kM2LCols * kM2LCols;
      for (int64_t b = 0; b < padded / kM2LCols; b++) blk.push_back(k);
      a += cnt[k];
      c += padded;
    }
    m.n_cols = c;
    m.col_cell.resize(c);
    fill_i32_kernel<<<ceil_div(c, 256), 256, 0, s>>>(m.col_cell.get(), c, -1);
    PAM_LAUNCH_CHECK();
    fill_cols_kernel<<<ceil_div(n, 256), 256, 0, s>>>(ws.tmp_idx.get(), ws.tmp_codes.get(), n, o,
                                                      m.col_cell.get());
    PAM_LAUNCH_CHECK();
    m.h_blk_oct = blk;
    m.blk_oct.upload(m.h_blk_oct.data(), m.h_blk_oct.size(), s);
    PAM_HIP(hipStreamSynchronize(s));  // the upload reads host memory
  }
}

}  // namespace pam
