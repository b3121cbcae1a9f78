// FMM stage kernels for gfx950.
//   P2M / L2P : one workgroup per leaf, Lagrange weights staged in LDS, streaming
//   M2M / L2L : separable 1-D transfers (3 passes of P^4) through LDS
//   M2L       : gather-GEMM on MFMA (16x16x4 f32 or f64): 32 target cells of
//               one octant x 64 operator rows per workgroup, the 189 transfer
//               vectors of that octant reduced in registers
//   P2P       : one workgroup per target leaf, the 27 neighbour leaves staged
//               through LDS, targets x source-splits mapped onto 256 lanes
#include <array>
#include <cmath>
#include <algorithm>
#include <type_traits>
#include <vector>

#include "fmm.hpp"

namespace pam {

namespace {

constexpr int kThreads = 256;

// T_k(x_a) for the Chebyshev nodes x_a = cos((2a+1) pi / 2P): tk[a*P + k].
template <class T, int P>
__device__ __forceinline__ void init_tk(T* tk) {
  for (int idx = threadIdx.x; idx < P * P; idx += blockDim.x) {
    int a = idx / P, k = idx % P;
    tk[idx] = static_cast<T>(cospi(static_cast<double>(k * (2 * a + 1)) / (2.0 * P)));
  }
}

// Lagrange basis at the Chebyshev nodes: S_a(x) = 1/P + 2/P sum_{k>=1} T_k(x_a) T_k(x).
template <class T, int P>
__device__ __forceinline__ void lagrange(T xi, const T* __restrict__ tk, T* S) {
  T t[P];
  t[0] = T(1);
  if constexpr (P > 1) t[1] = xi;
#pragma unroll
  for (int k = 2; k < P; k++) t[k] = T(2) * xi * t[k - 1] - t[k - 2];
  const T c0 = T(1) / T(P), c1 = T(2) / T(P);
#pragma unroll
  for (int a = 0; a < P; a++) {
    T sum = T(0);
#pragma unroll
    for (int k = 1; k < P; k++) sum += tk[a * P + k] * t[k];
    S[a] = c0 + c1 * sum;
  }
}

// Same values written to column `col` of s[P][ncols] with a rolled loop over
// a, so that the compiler does not keep the P^2 table tk in registers.
template <class T, int P, int NC>
__device__ __forceinline__ void lagrange_lds(T xi, const T* __restrict__ tk, T (*s)[NC], int col) {
  T t[P];
  t[0] = T(1);
  if constexpr (P > 1) t[1] = xi;
#pragma unroll
  for (int k = 2; k < P; k++) t[k] = T(2) * xi * t[k - 1] - t[k - 2];
  const T c0 = T(1) / T(P), c1 = T(2) / T(P);
#pragma unroll 1
  for (int a = 0; a < P; a++) {
    T sum = T(0);
#pragma unroll
    for (int k = 1; k < P; k++) sum += tk[a * P + k] * t[k];
    s[a][col] = c0 + c1 * sum;
  }
}

// tr[o][a][a2] = S_a((x_a2 + (2o - 1)) / 2): parent-node weight a of child node a2
// for the child on side o of an axis. Computed in FP64.
template <class T, int P>
__device__ __forceinline__ void init_transfer(T* tr) {
  for (int idx = threadIdx.x; idx < 2 * P * P; idx += blockDim.x) {
    int o = idx / (P * P), a = (idx / P) % P, a2 = idx % P;
    double x = (cospi((2.0 * a2 + 1.0) / (2.0 * P)) + (2.0 * o - 1.0)) / 2.0;
    double t0 = 1.0, t1 = x, sum = 0.0;
    for (int k = 1; k < P; k++) {
      double tk = (k == 1) ? t1 : 2.0 * x * t1 - t0;
      if (k >= 2) {
        t0 = t1;
        t1 = tk;
      }
      sum += cospi(static_cast<double>(k * (2 * a + 1)) / (2.0 * P)) * tk;
    }
    tr[idx] = static_cast<T>(1.0 / P + 2.0 / P * sum);
  }
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

struct BoxArgs {
  double lo[3];
  double cell;   // leaf cell width
};

template <class T>
__global__ void prepare_points_kernel(const double* __restrict__ pts, const double* __restrict__ w,
                                      const int32_t* __restrict__ perm,
                                      const int32_t* __restrict__ point_leaf,
                                      const uint64_t* __restrict__ leaf_codes, int64_t n, BoxArgs b,
                                      V4<T>* __restrict__ out) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t j = perm[i];
  uint64_t code = leaf_codes[point_leaf[i]];
  double c[3] = {static_cast<double>(compact3(code >> 2)), static_cast<double>(compact3(code >> 1)),
                 static_cast<double>(compact3(code))};
  double inv_h = 2.0 / b.cell;
  double xi[3];
  for (int k = 0; k < 3; k++) {
    double center = b.lo[k] + (c[k] + 0.5) * b.cell;
    xi[k] = (pts[3 * j + k] - center) * inv_h;
  }
  V4<T> v;
  v.x = static_cast<T>(xi[0]);
  v.y = static_cast<T>(xi[1]);
  v.z = static_cast<T>(xi[2]);
  v.w = w ? static_cast<T>(w[j]) : T(0);
  out[i] = v;
}

template <class T>
__global__ void set_weights_kernel(const double* __restrict__ w, const int32_t* __restrict__ perm,
                                   int64_t n, V4<T>* __restrict__ pts) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pts[i].w = static_cast<T>(w[perm[i]]);
}

constexpr int kP2mChunk = 64;

template <class T, int P>
__global__ void __launch_bounds__(kThreads)
p2m_kernel(const int32_t* __restrict__ leaf_start, const V4<T>* __restrict__ pts, T* __restrict__ m,
           int kpad) {
  constexpr int P2 = P * P, P3 = P * P * P;
  // Thread (g, bc): the P coefficients (a, b, c) of one (b, c) pair, summed
  // over the points j = g (mod NG) of each chunk; the NG partial sums are
  // added in group order at the end. Per point a thread reads Y_b, Z_c and
  // the P weighted X_a from LDS for P multiply-adds.
  constexpr int NG = kThreads / P2 > 0 ? kThreads / P2 : 1;
  __shared__ T tk[P * P];
  __shared__ T sv[3][kP2mChunk][P];
  __shared__ T red[NG > 1 ? NG - 1 : 1][P3];
  const int leaf = blockIdx.x;
  const int ps = leaf_start[leaf], pe = leaf_start[leaf + 1];
  init_tk<T, P>(tk);
  __syncthreads();
  const int bc = threadIdx.x % P2, g = threadIdx.x / P2;
  const bool active = g < NG;
  const int b = bc / P, c = bc % P;
  T acc[P];
#pragma unroll
  for (int a = 0; a < P; a++) acc[a] = T(0);
  for (int c0 = ps; c0 < pe; c0 += kP2mChunk) {
    int cnt = min(kP2mChunk, pe - c0);
    if (threadIdx.x < 3 * kP2mChunk) {
      int j = threadIdx.x % kP2mChunk, d = threadIdx.x / kP2mChunk;
      T S[P];
      if (j < cnt) {
        V4<T> q = pts[c0 + j];
        T xi = d == 0 ? q.x : (d == 1 ? q.y : q.z);
        lagrange<T, P>(xi, tk, S);
        if (d == 0) {
#pragma unroll
          for (int a = 0; a < P; a++) S[a] *= q.w;
        }
      } else {
#pragma unroll
        for (int a = 0; a < P; a++) S[a] = T(0);
      }
#pragma unroll
      for (int a = 0; a < P; a++) sv[d][j][a] = S[a];
    }
    __syncthreads();
    if (active) {
      for (int j = g; j < cnt; j += NG) {
        const T yz = sv[1][j][b] * sv[2][j][c];
#pragma unroll
        for (int a = 0; a < P; a++) acc[a] += sv[0][j][a] * yz;
      }
    }
    __syncthreads();
  }
  if (NG > 1) {
    if (active && g > 0) {
#pragma unroll
      for (int a = 0; a < P; a++) red[g - 1][a * P2 + bc] = acc[a];
    }
    __syncthreads();
    if (g == 0) {
      for (int k = 1; k < NG; k++) {
#pragma unroll
        for (int a = 0; a < P; a++) acc[a] += red[k - 1][a * P2 + bc];
      }
    }
  }
  T* dst = m + static_cast<int64_t>(leaf) * kpad;
  if (g == 0) {
#pragma unroll
    for (int a = 0; a < P; a++) dst[a * P2 + bc] = acc[a];
  }
  for (int o = P3 + threadIdx.x; o < kpad; o += kThreads) dst[o] = T(0);
}

template <class T, int P>
__global__ void __launch_bounds__(kThreads)
m2m_kernel(const int32_t* __restrict__ child_start, const uint64_t* __restrict__ child_codes,
           const T* __restrict__ mc, T* __restrict__ mp, int kpad) {
  constexpr int P2 = P * P, P3 = P * P * P;
  constexpr int NO = (P3 + kThreads - 1) / kThreads;
  __shared__ T tr[2 * P2];
  __shared__ T b0[P3], b1[P3];
  const int parent = blockIdx.x;
  init_transfer<T, P>(tr);
  T acc[NO];
#pragma unroll
  for (int q = 0; q < NO; q++) acc[q] = T(0);
  const int cs = child_start[parent], ce = child_start[parent + 1];
  for (int c = cs; c < ce; c++) {
    int oct = static_cast<int>(child_codes[c] & 7u);
    const T* tx = tr + ((oct >> 2) & 1) * P2;
    const T* ty = tr + ((oct >> 1) & 1) * P2;
    const T* tz = tr + (oct & 1) * P2;
    __syncthreads();
    for (int idx = threadIdx.x; idx < P3; idx += kThreads) b0[idx] = mc[static_cast<int64_t>(c) * kpad + idx];
    __syncthreads();
    for (int idx = threadIdx.x; idx < P3; idx += kThreads) {  // z
      int ab = idx / P, g = idx % P;
      T s = T(0);
#pragma unroll
      for (int g2 = 0; g2 < P; g2++) s += tz[g * P + g2] * b0[ab * P + g2];
      b1[idx] = s;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < P3; idx += kThreads) {  // y
      int a2 = idx / P2, bb = (idx / P) % P, g = idx % P;
      T s = T(0);
#pragma unroll
      for (int b2 = 0; b2 < P; b2++) s += ty[bb * P + b2] * b1[(a2 * P + b2) * P + g];
      b0[idx] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NO; q++) {  // x
      int o = threadIdx.x + q * kThreads;
      if (o < P3) {
        int a = o / P2, bg = o % P2;
        T s = T(0);
#pragma unroll
        for (int a2 = 0; a2 < P; a2++) s += tx[a * P + a2] * b0[a2 * P2 + bg];
        acc[q] += s;
      }
    }
  }
  T* dst = mp + static_cast<int64_t>(parent) * kpad;
#pragma unroll
  for (int q = 0; q < NO; q++) {
    int o = threadIdx.x + q * kThreads;
    if (o < P3) dst[o] = acc[q];
  }
  for (int o = P3 + threadIdx.x; o < kpad; o += kThreads) dst[o] = T(0);
}

template <class T, int P>
__global__ void __launch_bounds__(kThreads)
l2l_kernel(const int32_t* __restrict__ parent_of, const uint64_t* __restrict__ codes,
           const T* __restrict__ lp, T* __restrict__ lc, int kpad) {
  constexpr int P2 = P * P, P3 = P * P * P;
  __shared__ T tr[2 * P2];
  __shared__ T b0[P3], b1[P3];
  const int c = blockIdx.x;
  const int parent = parent_of[c];
  const int oct = static_cast<int>(codes[c] & 7u);
  init_transfer<T, P>(tr);
  for (int idx = threadIdx.x; idx < P3; idx += kThreads) b0[idx] = lp[static_cast<int64_t>(parent) * kpad + idx];
  __syncthreads();
  const T* tx = tr + ((oct >> 2) & 1) * P2;
  const T* ty = tr + ((oct >> 1) & 1) * P2;
  const T* tz = tr + (oct & 1) * P2;
  for (int idx = threadIdx.x; idx < P3; idx += kThreads) {  // x: b1[a2][b][g] = sum_a tx[a][a2] b0[a][b][g]
    int a2 = idx / P2, bg = idx % P2;
    T s = T(0);
#pragma unroll
    for (int a = 0; a < P; a++) s += tx[a * P + a2] * b0[a * P2 + bg];
    b1[idx] = s;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < P3; idx += kThreads) {  // y
    int a2 = idx / P2, b2 = (idx / P) % P, g = idx % P;
    T s = T(0);
#pragma unroll
    for (int bb = 0; bb < P; bb++) s += ty[bb * P + b2] * b1[(a2 * P + bb) * P + g];
    b0[idx] = s;
  }
  __syncthreads();
  T* dst = lc + static_cast<int64_t>(c) * kpad;
  for (int idx = threadIdx.x; idx < P3; idx += kThreads) {  // z
    int ab = idx / P, g2 = idx % P;
    T s = T(0);
#pragma unroll
    for (int g = 0; g < P; g++) s += tz[g * P + g2] * b0[ab * P + g];
    dst[idx] += s;
  }
}

template <class T, int P>
__device__ __forceinline__ void l2p_leaf(int leaf, const int32_t* __restrict__ leaf_start,
                                         const V4<T>* __restrict__ pts, const T* __restrict__ l, int kpad,
                                         T* __restrict__ y_far, const int32_t* __restrict__ perm) {
  constexpr int P2 = P * P, P3 = P * P * P;
  __shared__ T tk[P2];
  __shared__ T ls[P3];
  // Per-thread Lagrange factors in LDS: the loops over them stay rolled (or
  // partly unrolled), otherwise the compiler keeps the P^2 Chebyshev table and
  // the P^3 local coefficients in registers across points and spills.
  __shared__ T sxs[P][kThreads];
  __shared__ T sys[P][kThreads];
  __shared__ T szs[P][kThreads];
  constexpr int kUnrollB = P <= 8 ? P : 2;
  init_tk<T, P>(tk);
  for (int idx = threadIdx.x; idx < P3; idx += kThreads) ls[idx] = l[static_cast<int64_t>(leaf) * kpad + idx];
  __syncthreads();
  const int ps = leaf_start[leaf], pe = leaf_start[leaf + 1];
  const int c = threadIdx.x;
  for (int i = ps + c; i < pe; i += kThreads) {
    V4<T> q = pts[i];
    lagrange_lds<T, P, kThreads>(q.x, tk, sxs, c);
    lagrange_lds<T, P, kThreads>(q.y, tk, sys, c);
    lagrange_lds<T, P, kThreads>(q.z, tk, szs, c);
    T sz[P];
#pragma unroll
    for (int g = 0; g < P; g++) sz[g] = szs[g][c];
    T y = T(0);
#pragma unroll 1
    for (int a = 0; a < P; a++) {
      const T* la = ls + a * P2;
      T ya = T(0);
#pragma unroll kUnrollB
      for (int bb = 0; bb < P; bb++) {
        T yb = T(0);
#pragma unroll
        for (int g = 0; g < P; g++) yb += la[bb * P + g] * sz[g];
        ya += yb * sys[bb][c];
      }
      y += ya * sxs[a][c];
    }
    // perm: add the far field into the output in input order (the near field
    // is already there); otherwise store it in tree order.
    if (perm) y_far[perm[i]] += y;
    else y_far[i] = y;
  }
}

// The far fields (L2P) of up to kMaxFmmJobs evaluations of one order in one
// launch, job lookup as in p2p_jobs_kernel.
template <class T, int P>
__global__ void __launch_bounds__(kThreads) l2p_jobs_kernel(L2PJobs<T> jobs) {
  const int g = blockIdx.x;
  int k = 0;
  while (k + 1 < jobs.n && jobs.j[k + 1].first_leaf <= g) k++;
  const L2PJob<T>& j = jobs.j[k];
  l2p_leaf<T, P>(g - j.first_leaf, j.leaf_start, j.pts, j.l, j.kpad, j.y_far, j.perm);
}

// Targets per thread of the near-field kernel (each source read from LDS
// serves R pairs): chosen per leaf in [4, kP2PMaxR].
constexpr int kP2PMinR = 4, kP2PMaxR = 8;

// The R and tp (target slots, a power of two) for a leaf of nt targets: the
// pair with the fewest idle target slots R tp - nt (ties: the smaller R). A
// fixed R = 4 rounds 257 targets up to 512 slots, and every idle slot costs
// a full pass over the leaf's sources; on 1M uniform points (about 244 per
// leaf) that was a quarter of the pairs computed.
__device__ __forceinline__ int p2p_shape(int nt, int* tp_out) {
  int best_r = kP2PMinR, best_tp = kThreads, best_cap = 1 << 30;
  for (int r = kP2PMinR; r <= kP2PMaxR; r++) {
    const int per = (nt + r - 1) / r;
    int tp = 1;
    while (tp < per && tp < kThreads) tp <<= 1;
    const int passes = (per + tp - 1) / tp;
    const int cap = r * tp * passes;
    if (cap < best_cap) best_cap = cap, best_r = r, best_tp = tp;
  }
  *tp_out = best_tp;
  return best_r;
}

// kComp: compensated (Kahan) sums per target and FP64 sums over the source
// splits, for weights that cancel strongly (kriging).
// The passes of one target leaf with R targets per thread and tp target
// slots (tp x nsplit threads; the sources staged through sp in chunks).
template <class T, int KIND, bool kComp, int R>
__device__ __forceinline__ void p2p_passes(int nt, int ts, int tp, int total, int nnb, const int* nb_start,
                                           const int* nb_src, const T (*nb_off)[3], V4<T>* sp, int cap,
                                           T (*red)[kThreads], const V4<T>* __restrict__ tpts,
                                           const V4<T>* __restrict__ spts, const double* __restrict__ y_far,
                                           const int32_t* __restrict__ tperm, T pa, T pb,
                                           double* __restrict__ out) {
  // Every thread carries R targets (ti + r tp), so each source read from LDS
  // serves R pairs; tp target slots x nsplit source splits fill the 256 lanes.
  const int nsplit = kThreads / tp;
  const int ti = threadIdx.x & (tp - 1);
  const int si = threadIdx.x / tp;
  for (int c0 = 0; c0 < nt; c0 += R * tp) {
    V4<T> tq[R];
    T acc[R], cmp[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = c0 + r * tp + ti;
      tq[r] = i < nt ? tpts[ts + i] : V4<T>{T(0), T(0), T(0), T(0)};
      acc[r] = T(0);
      cmp[r] = T(0);
    }
    for (int cb = 0; cb < total; cb += cap) {
      const int cnt = min(cap, total - cb);
      __syncthreads();
      for (int e = threadIdx.x; e < cnt; e += kThreads) {
        int g = cb + e;
        int k = 0;
        while (k + 1 < nnb && nb_start[k + 1] <= g) k++;
        V4<T> q = spts[nb_src[k] + (g - nb_start[k])];
        q.x += nb_off[k][0];
        q.y += nb_off[k][1];
        q.z += nb_off[k][2];
        q.w *= pa;  // the kernel's factor a folded into the weight
        sp[e] = q;
      }
      __syncthreads();
      // Per pair: three differences, three FMAs seeded with the kernel's
      // offset, the shape function and one FMA. Four sources per trip keep
      // their LDS reads ahead of the arithmetic; the sum order over sources
      // is the plain sequential one.
      const T seed = phi_seed<T, KIND>(pb);
      auto pair = [&](const V4<T>& q, int r) {
        T dx = tq[r].x - q.x, dy = tq[r].y - q.y, dz = tq[r].z - q.z;
        T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, seed)));
        if constexpr (kComp) {
          const T y = q.w * phi_shape<T, KIND>(r2, pb) - cmp[r];
          const T t = acc[r] + y;
          cmp[r] = (t - acc[r]) - y;
          acc[r] = t;
        } else {
          acc[r] = fma(q.w, phi_shape<T, KIND>(r2, pb), acc[r]);
        }
      };
      int j = si;
      for (; j + 3 * nsplit < cnt; j += 4 * nsplit) {
        const V4<T> q0 = sp[j], q1 = sp[j + nsplit], q2 = sp[j + 2 * nsplit], q3 = sp[j + 3 * nsplit];
#pragma unroll
        for (int r = 0; r < R; r++) {
          pair(q0, r);
          pair(q1, r);
          pair(q2, r);
          pair(q3, r);
        }
      }
      for (; j < cnt; j += nsplit) {
        const V4<T> q = sp[j];
#pragma unroll
        for (int r = 0; r < R; r++) pair(q, r);
      }
    }
    __syncthreads();  // red overlays the staged sources
#pragma unroll
    for (int r = 0; r < R; r++) red[r][threadIdx.x] = kComp ? acc[r] - cmp[r] : acc[r];
    __syncthreads();
    if (si == 0) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = c0 + r * tp + ti;
        if (i >= nt) continue;
        using S = std::conditional_t<kComp, double, T>;
        S sum = S(0);
        for (int k = 0; k < nsplit; k++) sum += static_cast<S>(red[r][ti + k * tp]);
        const int gi = ts + i;
        double far = y_far ? y_far[gi] : 0.0;
        out[tperm[gi]] = far + static_cast<double>(sum);
      }
    }
    __syncthreads();
  }
}

// One target leaf of one evaluation's near field.
template <class T, int KIND, bool kComp>
__device__ __forceinline__ void p2p_leaf(int leaf, const int32_t* __restrict__ tls, const int32_t* __restrict__ sls,
                                         const int32_t* __restrict__ p2p, const V4<T>* __restrict__ tpts,
                                         const V4<T>* __restrict__ spts, const double* __restrict__ y_far,
                                         const int32_t* __restrict__ tperm, T pa, T pb, double* __restrict__ out) {
  constexpr int CAP = 32768 / static_cast<int>(sizeof(V4<T>));
  __shared__ V4<T> sp[CAP];
  __shared__ int nb_start[28];
  __shared__ int nb_src[27];
  __shared__ T nb_off[27][3];
  __shared__ int nb_count;
  // The split partials of a pass, over the staged sources (read after them).
  static_assert(kP2PMaxR * kThreads * sizeof(T) <= sizeof(V4<T>) * CAP, "partials exceed the staging buffer");
  T (*red)[kThreads] = reinterpret_cast<T (*)[kThreads]>(sp);
  const int ts = tls[leaf], te = tls[leaf + 1];
  const int nt = te - ts;
  if (threadIdx.x == 0) {
    int k = 0, total = 0;
    for (int s = 0; s < 27; s++) {
      int sl = p2p[static_cast<int64_t>(leaf) * 27 + s];
      if (sl < 0) continue;
      int a = sls[sl], b = sls[sl + 1];
      nb_src[k] = a;
      nb_start[k] = total;
      nb_off[k][0] = T(2 * (s / 9 - 1));
      nb_off[k][1] = T(2 * ((s / 3) % 3 - 1));
      nb_off[k][2] = T(2 * (s % 3 - 1));
      total += b - a;
      k++;
    }
    nb_start[k] = total;
    nb_count = k;
  }
  __syncthreads();
  const int nnb = nb_count;
  const int total = nb_start[nnb];
  int tp = kThreads;
  const int r = p2p_shape(nt, &tp);
#define PAM_P2P_PASSES(RR)                                                                                     \
  p2p_passes<T, KIND, kComp, RR>(nt, ts, tp, total, nnb, nb_start, nb_src, nb_off, sp, CAP, red, tpts, spts, \
                                 y_far, tperm, pa, pb, out)
  switch (r) {
    case 4: PAM_P2P_PASSES(4); break;
    case 5: PAM_P2P_PASSES(5); break;
    case 6: PAM_P2P_PASSES(6); break;
    case 7: PAM_P2P_PASSES(7); break;
    default: PAM_P2P_PASSES(8); break;
  }
#undef PAM_P2P_PASSES
}

// The near fields of up to kMaxFmmJobs evaluations (same precision, kernel
// kind and summation form) in one launch: workgroup g takes target leaf
// g - first_leaf of the last job whose first_leaf <= g. A single evaluation
// is the one-job case of the same kernel, so its result does not depend on
// what shares the launch.
template <class T, int KIND, bool kComp>
__global__ void __launch_bounds__(kThreads) p2p_jobs_kernel(P2PJobs<T> jobs) {
  const int g = blockIdx.x;
  int k = 0;
  while (k + 1 < jobs.n && jobs.j[k + 1].first_leaf <= g) k++;
  const P2PJob<T>& j = jobs.j[k];
  p2p_leaf<T, KIND, kComp>(g - j.first_leaf, j.tls, j.sls, j.p2p, j.tpts, j.spts, j.y_far, j.tperm, j.pa, j.pb,
                           j.out);
}

// ---- gradients at the targets (gradient-transpose kernel,
// fmm/gradient_transpose_kernel.hpp:47-62): out_i = A^T sum_j w_j grad phi(A (t_i - s_j)) ----

// Lagrange basis and its derivative at the Chebyshev nodes:
// S_a'(x) = 2/P sum_{k>=1} T_k(x_a) k U_{k-1}(x), written to column `col`.
template <class T, int P, int NC>
__device__ __forceinline__ void lagrange_d_lds(T xi, const T* __restrict__ tk, T (*s)[NC], T (*d)[NC],
                                               int col) {
  T t[P], du[P];
  t[0] = T(1);
  du[0] = T(0);
  if constexpr (P > 1) {
    t[1] = xi;
    du[1] = T(1);  // 1 * U_0
  }
  T u0 = T(1), u1 = T(2) * xi;  // U_0, U_1
#pragma unroll
  for (int k = 2; k < P; k++) {
    t[k] = T(2) * xi * t[k - 1] - t[k - 2];
    du[k] = T(k) * u1;  // k U_{k-1}
    const T u2 = T(2) * xi * u1 - u0;
    u0 = u1;
    u1 = u2;
  }
  const T c0 = T(1) / T(P), c1 = T(2) / T(P);
#pragma unroll 1
  for (int a = 0; a < P; a++) {
    T sum = T(0), sd = T(0);
#pragma unroll
    for (int k = 1; k < P; k++) {
      sum += tk[a * P + k] * t[k];
      sd += tk[a * P + k] * du[k];
    }
    s[a][col] = c0 + c1 * sum;
    d[a][col] = c1 * sd;
  }
}

struct Aniso3 {
  double a[9];  // row-major A
};

constexpr int kL2PGThreads = 128;

// Far-field gradient: the derivative of each leaf's local interpolant,
// divided by the leaf half width, turned back by A^T and added to the output
// in input order (the near field is already there).
template <int P>
__global__ void __launch_bounds__(kL2PGThreads)
l2p_grad_kernel(const int32_t* __restrict__ leaf_start, const V4<double>* __restrict__ pts,
                const double* __restrict__ l, int kpad, double inv_h, Aniso3 A, const int32_t* __restrict__ perm,
                double* __restrict__ out) {
  constexpr int P2 = P * P, P3 = P * P * P;
  __shared__ double tk[P2];
  __shared__ double ls[P3];
  __shared__ double sxs[P][kL2PGThreads], sys[P][kL2PGThreads], szs[P][kL2PGThreads];
  __shared__ double dxs[P][kL2PGThreads], dys[P][kL2PGThreads], dzs[P][kL2PGThreads];
  const int leaf = blockIdx.x;
  init_tk<double, P>(tk);
  for (int idx = threadIdx.x; idx < P3; idx += kL2PGThreads) ls[idx] = l[static_cast<int64_t>(leaf) * kpad + idx];
  __syncthreads();
  const int ps = leaf_start[leaf], pe = leaf_start[leaf + 1];
  const int c = threadIdx.x;
  for (int i = ps + c; i < pe; i += kL2PGThreads) {
    const V4<double> q = pts[i];
    lagrange_d_lds<double, P, kL2PGThreads>(q.x, tk, sxs, dxs, c);
    lagrange_d_lds<double, P, kL2PGThreads>(q.y, tk, sys, dys, c);
    lagrange_d_lds<double, P, kL2PGThreads>(q.z, tk, szs, dzs, c);
    double sz[P], dz[P];
#pragma unroll
    for (int g = 0; g < P; g++) {
      sz[g] = szs[g][c];
      dz[g] = dzs[g][c];
    }
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll 1
    for (int a = 0; a < P; a++) {
      const double* la = ls + a * P2;
      double ya = 0.0, yay = 0.0, yaz = 0.0;
#pragma unroll 1
      for (int bb = 0; bb < P; bb++) {
        double yb = 0.0, ybz = 0.0;
#pragma unroll
        for (int g = 0; g < P; g++) {
          yb += la[bb * P + g] * sz[g];
          ybz += la[bb * P + g] * dz[g];
        }
        ya += yb * sys[bb][c];
        yay += yb * dys[bb][c];
        yaz += ybz * sys[bb][c];
      }
      gx += ya * dxs[a][c];
      gy += yay * sxs[a][c];
      gz += yaz * sxs[a][c];
    }
    gx *= inv_h;
    gy *= inv_h;
    gz *= inv_h;
    double* o = out + 3 * static_cast<int64_t>(perm[i]);
#pragma unroll
    for (int k = 0; k < 3; k++) o[k] += (A.a[k] * gx + A.a[3 + k] * gy) + A.a[6 + k] * gz;
  }
}

// Near-field gradient: the layout of p2p_kernel with three sums per target;
// writes A^T (sum over the 27 neighbour leaves) to the output in input order.
template <class T, int KIND>
__global__ void __launch_bounds__(kThreads)
p2p_grad_kernel(const int32_t* __restrict__ tls, const int32_t* __restrict__ sls,
                const int32_t* __restrict__ p2p, const V4<T>* __restrict__ tpts,
                const V4<T>* __restrict__ spts, const int32_t* __restrict__ tperm, T ag, T pb, Aniso3 A,
                double* __restrict__ out) {
  constexpr int CAP = 32768 / static_cast<int>(sizeof(V4<T>));
  constexpr int R = 2;
  __shared__ V4<T> sp[CAP];
  __shared__ int nb_start[28];
  __shared__ int nb_src[27];
  __shared__ T nb_off[27][3];
  __shared__ int nb_count;
  __shared__ T red[R][3][kThreads];
  const int leaf = blockIdx.x;
  const int ts = tls[leaf], te = tls[leaf + 1];
  const int nt = te - ts;
  if (threadIdx.x == 0) {
    int k = 0, total = 0;
    for (int s = 0; s < 27; s++) {
      int sl = p2p[static_cast<int64_t>(leaf) * 27 + s];
      if (sl < 0) continue;
      int a = sls[sl], b = sls[sl + 1];
      nb_src[k] = a;
      nb_start[k] = total;
      nb_off[k][0] = T(2 * (s / 9 - 1));
      nb_off[k][1] = T(2 * ((s / 3) % 3 - 1));
      nb_off[k][2] = T(2 * (s % 3 - 1));
      total += b - a;
      k++;
    }
    nb_start[k] = total;
    nb_count = k;
  }
  __syncthreads();
  const int nnb = nb_count;
  const int total = nb_start[nnb];
  const int per = (nt + R - 1) / R;
  int tp = 1;
  while (tp < per && tp < kThreads) tp <<= 1;
  const int nsplit = kThreads / tp;
  const int ti = threadIdx.x & (tp - 1);
  const int si = threadIdx.x / tp;
  const T seed = phi_seed<T, KIND>(pb);
  for (int c0 = 0; c0 < nt; c0 += R * tp) {
    V4<T> tq[R];
    T ax[R], ay[R], az[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = c0 + r * tp + ti;
      tq[r] = i < nt ? tpts[ts + i] : V4<T>{T(0), T(0), T(0), T(0)};
      ax[r] = ay[r] = az[r] = T(0);
    }
    for (int cb = 0; cb < total; cb += CAP) {
      const int cnt = min(CAP, total - cb);
      __syncthreads();
      for (int e = threadIdx.x; e < cnt; e += kThreads) {
        int g = cb + e;
        int k = 0;
        while (k + 1 < nnb && nb_start[k + 1] <= g) k++;
        V4<T> q = spts[nb_src[k] + (g - nb_start[k])];
        q.x += nb_off[k][0];
        q.y += nb_off[k][1];
        q.z += nb_off[k][2];
        q.w *= ag;  // the gradient's factor folded into the weight
        sp[e] = q;
      }
      __syncthreads();
      for (int j = si; j < cnt; j += nsplit) {
        const V4<T> q = sp[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
          const T dx = tq[r].x - q.x, dy = tq[r].y - q.y, dz = tq[r].z - q.z;
          const T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, seed)));
          const T f = q.w * grad_shape<T, KIND>(r2, pb);
          ax[r] = fma(f, dx, ax[r]);
          ay[r] = fma(f, dy, ay[r]);
          az[r] = fma(f, dz, az[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      red[r][0][threadIdx.x] = ax[r];
      red[r][1][threadIdx.x] = ay[r];
      red[r][2][threadIdx.x] = az[r];
    }
    __syncthreads();
    if (si == 0) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = c0 + r * tp + ti;
        if (i >= nt) continue;
        double g[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          T sum = T(0);
          for (int m = 0; m < nsplit; m++) sum += red[r][k][ti + m * tp];
          g[k] = static_cast<double>(sum);
        }
        double* o = out + 3 * static_cast<int64_t>(tperm[ts + i]);
#pragma unroll
        for (int k = 0; k < 3; k++) o[k] = (A.a[k] * g[0] + A.a[3 + k] * g[1]) + A.a[6 + k] * g[2];
      }
    }
    __syncthreads();
  }
}

// ---- one column of the Hessian at the targets (the Hessian kernel,
// fmm/hessian_kernel.hpp:46-64, applied to one weight component):
// out_i = A^T sum_j w_j H(A (t_i - s_j)) a_k, a_k the k-th column of A ----

// S_a, S_a' and S_a'' at x: T_k'' from T''_{k+1} = 4 T'_k + 2x T''_k - T''_{k-1}.
template <class T, int P, int NC>
__device__ __forceinline__ void lagrange_dd_lds(T xi, const T* __restrict__ tk, T (*s)[NC], T (*d)[NC],
                                                T (*dd)[NC], int col) {
  T t[P], d1[P], d2[P];
  t[0] = T(1);
  d1[0] = T(0);
  d2[0] = T(0);
  if constexpr (P > 1) {
    t[1] = xi;
    d1[1] = T(1);
    d2[1] = T(0);
  }
#pragma unroll
  for (int k = 2; k < P; k++) {
    t[k] = T(2) * xi * t[k - 1] - t[k - 2];
    d1[k] = T(2) * t[k - 1] + T(2) * xi * d1[k - 1] - d1[k - 2];
    d2[k] = T(4) * d1[k - 1] + T(2) * xi * d2[k - 1] - d2[k - 2];
  }
  const T c0 = T(1) / T(P), c1 = T(2) / T(P);
#pragma unroll 1
  for (int a = 0; a < P; a++) {
    T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
    for (int k = 1; k < P; k++) {
      s0 += tk[a * P + k] * t[k];
      s1 += tk[a * P + k] * d1[k];
      s2 += tk[a * P + k] * d2[k];
    }
    s[a][col] = c0 + c1 * s0;
    d[a][col] = c1 * s1;
    dd[a][col] = c1 * s2;
  }
}

constexpr int kL2PHThreads = 64;

template <int P>
__global__ void __launch_bounds__(kL2PHThreads)
l2p_hess_col_kernel(const int32_t* __restrict__ leaf_start, const V4<double>* __restrict__ pts,
                    const double* __restrict__ l, int kpad, double inv_h2, Aniso3 A, int kcol,
                    const int32_t* __restrict__ perm, double* __restrict__ out) {
  constexpr int P2 = P * P, P3 = P * P * P;
  constexpr int NC = kL2PHThreads;
  __shared__ double tk[P2];
  __shared__ double ls[P3];
  __shared__ double x0[P][NC], x1[P][NC], x2[P][NC];
  __shared__ double y0[P][NC], y1[P][NC], y2[P][NC];
  __shared__ double z0s[P][NC], z1s[P][NC], z2s[P][NC];
  const int leaf = blockIdx.x;
  init_tk<double, P>(tk);
  for (int idx = threadIdx.x; idx < P3; idx += NC) ls[idx] = l[static_cast<int64_t>(leaf) * kpad + idx];
  __syncthreads();
  const int ps = leaf_start[leaf], pe = leaf_start[leaf + 1];
  const int c = threadIdx.x;
  const double ak0 = A.a[kcol], ak1 = A.a[3 + kcol], ak2 = A.a[6 + kcol];
  for (int i = ps + c; i < pe; i += NC) {
    const V4<double> q = pts[i];
    lagrange_dd_lds<double, P, NC>(q.x, tk, x0, x1, x2, c);
    lagrange_dd_lds<double, P, NC>(q.y, tk, y0, y1, y2, c);
    lagrange_dd_lds<double, P, NC>(q.z, tk, z0s, z1s, z2s, c);
    double hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
#pragma unroll 1
    for (int a = 0; a < P; a++) {
      const double* la = ls + a * P2;
      double b00 = 0, b20 = 0, b02 = 0, b10 = 0, b01 = 0, b11 = 0;
#pragma unroll 1
      for (int bb = 0; bb < P; bb++) {
        double v0 = 0, v1 = 0, v2 = 0;
#pragma unroll
        for (int g = 0; g < P; g++) {
          const double lv = la[bb * P + g];
          v0 += lv * z0s[g][c];
          v1 += lv * z1s[g][c];
          v2 += lv * z2s[g][c];
        }
        const double Y0 = y0[bb][c], Y1 = y1[bb][c], Y2 = y2[bb][c];
        b00 += Y0 * v0;
        b20 += Y2 * v0;
        b10 += Y1 * v0;
        b02 += Y0 * v2;
        b01 += Y0 * v1;
        b11 += Y1 * v1;
      }
      const double X0 = x0[a][c], X1 = x1[a][c], X2 = x2[a][c];
      hxx += X2 * b00;
      hyy += X0 * b20;
      hzz += X0 * b02;
      hxy += X1 * b10;
      hxz += X1 * b01;
      hyz += X0 * b11;
    }
    const double v[3] = {(hxx * ak0 + hxy * ak1 + hxz * ak2) * inv_h2, (hxy * ak0 + hyy * ak1 + hyz * ak2) * inv_h2,
                         (hxz * ak0 + hyz * ak1 + hzz * ak2) * inv_h2};
    double* o = out + 3 * static_cast<int64_t>(perm[i]);
#pragma unroll
    for (int k = 0; k < 3; k++) o[k] += (A.a[k] * v[0] + A.a[3 + k] * v[1]) + A.a[6 + k] * v[2];
  }
}

template <class T, int KIND>
__global__ void __launch_bounds__(kThreads)
p2p_hess_col_kernel(const int32_t* __restrict__ tls, const int32_t* __restrict__ sls,
                    const int32_t* __restrict__ p2p, const V4<T>* __restrict__ tpts,
                    const V4<T>* __restrict__ spts, const int32_t* __restrict__ tperm, T hc, T pb, Aniso3 A,
                    int kcol, double* __restrict__ out) {
  constexpr int CAP = 32768 / static_cast<int>(sizeof(V4<T>));
  constexpr int R = 2;
  __shared__ V4<T> sp[CAP];
  __shared__ int nb_start[28];
  __shared__ int nb_src[27];
  __shared__ T nb_off[27][3];
  __shared__ int nb_count;
  __shared__ T red[R][3][kThreads];
  const int leaf = blockIdx.x;
  const int ts = tls[leaf], te = tls[leaf + 1];
  const int nt = te - ts;
  if (threadIdx.x == 0) {
    int k = 0, total = 0;
    for (int s = 0; s < 27; s++) {
      int sl = p2p[static_cast<int64_t>(leaf) * 27 + s];
      if (sl < 0) continue;
      int a = sls[sl], b = sls[sl + 1];
      nb_src[k] = a;
      nb_start[k] = total;
      nb_off[k][0] = T(2 * (s / 9 - 1));
      nb_off[k][1] = T(2 * ((s / 3) % 3 - 1));
      nb_off[k][2] = T(2 * (s % 3 - 1));
      total += b - a;
      k++;
    }
    nb_start[k] = total;
    nb_count = k;
  }
  __syncthreads();
  const int nnb = nb_count;
  const int total = nb_start[nnb];
  const int per = (nt + R - 1) / R;
  int tp = 1;
  while (tp < per && tp < kThreads) tp <<= 1;
  const int nsplit = kThreads / tp;
  const int ti = threadIdx.x & (tp - 1);
  const int si = threadIdx.x / tp;
  const T seed = phi_seed<T, KIND>(pb);
  const T ak0 = static_cast<T>(A.a[kcol]), ak1 = static_cast<T>(A.a[3 + kcol]), ak2 = static_cast<T>(A.a[6 + kcol]);
  for (int c0 = 0; c0 < nt; c0 += R * tp) {
    V4<T> tq[R];
    T ax[R], ay[R], az[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = c0 + r * tp + ti;
      tq[r] = i < nt ? tpts[ts + i] : V4<T>{T(0), T(0), T(0), T(0)};
      ax[r] = ay[r] = az[r] = T(0);
    }
    for (int cb = 0; cb < total; cb += CAP) {
      const int cnt = min(CAP, total - cb);
      __syncthreads();
      for (int e = threadIdx.x; e < cnt; e += kThreads) {
        int g = cb + e;
        int k = 0;
        while (k + 1 < nnb && nb_start[k + 1] <= g) k++;
        V4<T> q = spts[nb_src[k] + (g - nb_start[k])];
        q.x += nb_off[k][0];
        q.y += nb_off[k][1];
        q.z += nb_off[k][2];
        q.w *= hc;
        sp[e] = q;
      }
      __syncthreads();
      for (int j = si; j < cnt; j += nsplit) {
        const V4<T> q = sp[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
          const T dx = tq[r].x - q.x, dy = tq[r].y - q.y, dz = tq[r].z - q.z;
          const T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, seed)));
          const T f = q.w * grad_shape<T, KIND>(r2, pb);
          const T tu = hess_t<T, KIND>(r2, pb) * fma(dz, ak2, fma(dy, ak1, dx * ak0));
          ax[r] = fma(f, fma(tu, dx, ak0), ax[r]);
          ay[r] = fma(f, fma(tu, dy, ak1), ay[r]);
          az[r] = fma(f, fma(tu, dz, ak2), az[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      red[r][0][threadIdx.x] = ax[r];
      red[r][1][threadIdx.x] = ay[r];
      red[r][2][threadIdx.x] = az[r];
    }
    __syncthreads();
    if (si == 0) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = c0 + r * tp + ti;
        if (i >= nt) continue;
        double g[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          T sum = T(0);
          for (int m = 0; m < nsplit; m++) sum += red[r][k][ti + m * tp];
          g[k] = static_cast<double>(sum);
        }
        double* o = out + 3 * static_cast<int64_t>(tperm[ts + i]);
#pragma unroll
        for (int k = 0; k < 3; k++) o[k] = (A.a[k] * g[0] + A.a[3 + k] * g[1]) + A.a[6 + k] * g[2];
      }
    }
    __syncthreads();
  }
}

// ---- M2L on MFMA ----
template <class T>
struct Mfma;
template <>
struct Mfma<float> {
  using acc_t = float __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<double> {
  using acc_t = double __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

constexpr int kM2LKc = 256;  // reduction chunk staged in LDS
constexpr int kM2LBCols = 32;  // columns per workgroup of m2l_batch_kernel

// All levels of one M2L pass in one launch: column blocks of every level are
// concatenated along grid.x, so the few blocks of the coarse levels run
// beside the many blocks of the leaf level instead of after them.
template <class TM, class TS>
__global__ void __launch_bounds__(kThreads)
m2l_batch_kernel(M2LBatch<TM, TS> b) {
  using acc_t = typename Mfma<TM>::acc_t;
  constexpr int LDB = kM2LKc + 4;
  constexpr int NA = kM2LKc / 16;  // operator fragments per lane and chunk
  constexpr int NB = kM2LKc / 32;  // staged multipole vectors per thread and chunk
  __shared__ V4<TM> bs4[kM2LBCols * LDB / 4];
  __shared__ int cells[kM2LBCols];
  __shared__ int srcs[kM2LBCols];
  TM* bs = reinterpret_cast<TM*>(bs4);
  int lv = 0;
  while (lv + 1 < b.n && static_cast<int>(blockIdx.x) >= b.blk_begin[lv + 1]) lv++;
  const int blk = static_cast<int>(blockIdx.x) - b.blk_begin[lv];
  const int32_t* __restrict__ src_slots = b.src_slots[lv];
  const TS* __restrict__ in = b.in[lv];
  const int rpad = b.rpad, kpad = b.kpad;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int row0 = blockIdx.y * 64 + wave * 16;
  const int oct = b.blk_oct[lv] ? b.blk_oct[lv][blk * kM2LBCols / kM2LCols] : 0;
  const int nv = src_slots ? kM2LSlots : 1;
  if (threadIdx.x < kM2LBCols) cells[threadIdx.x] = b.col_cell[lv][blk * kM2LBCols + threadIdx.x];
  acc_t acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  // The MFMA accumulators are flushed into FP64 every 64 reduction steps (FP32)
  // or once at the end (FP64): one long FP32 running sum over 189 x P^3 terms
  // would otherwise dominate the error of the far field.
  double d0[4] = {0, 0, 0, 0}, d1[4] = {0, 0, 0, 0};
  auto flush = [&]() {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      d0[i] += static_cast<double>(acc0[i]);
      d1[i] += static_cast<double>(acc1[i]);
      acc0[i] = TM(0);
      acc1[i] = TM(0);
    }
  };
  const int col = threadIdx.x / 8, part = threadIdx.x % 8;  // staging map: 8 lanes per column
  for (int sidx = 0; sidx < nv; sidx++) {
    const int v = src_slots ? b.vlist[oct * kM2LSlots + sidx] : 0;
    __syncthreads();
    if (threadIdx.x < kM2LBCols) {
      int c = cells[threadIdx.x];
      srcs[threadIdx.x] = c < 0 ? -1 : (src_slots ? src_slots[static_cast<int64_t>(c) * kM2LSlots + sidx] : c);
    }
    __syncthreads();
    bool any = false;  // skip transfer vectors with no source in this block
    for (int k = 0; k < kM2LBCols; k++) any |= srcs[k] >= 0;
    if (!any) continue;
    const TM* av = b.a[lv] + static_cast<int64_t>(v) * b.per_v;
    const int src = srcs[col];
    for (int k0 = 0; k0 < kpad; k0 += kM2LKc) {
      const int kc = min(kM2LKc, kpad - k0);
      // Issue every global load of the chunk before waiting on any of them.
      V4<TM> breg[NB];
#pragma unroll
      for (int i = 0; i < NB; i++) {
        int k4 = part + 8 * i;
        V4<TM> val{TM(0), TM(0), TM(0), TM(0)};
        if (k4 < kc / 4 && src >= 0) {
          V4<TS> t = *reinterpret_cast<const V4<TS>*>(in + static_cast<int64_t>(src) * b.in_stride + k0 + 4 * k4);
          val.x = static_cast<TM>(t.x);
          val.y = static_cast<TM>(t.y);
          val.z = static_cast<TM>(t.z);
          val.w = static_cast<TM>(t.w);
        }
        breg[i] = val;
      }
      V4<TM> areg[NA];
#pragma unroll
      for (int kb = 0; kb < NA; kb++) {
        if (kb < kc / 16)
          areg[kb] = *reinterpret_cast<const V4<TM>*>(
              av + ((static_cast<int64_t>((k0 >> 4) + kb) * rpad + row0 + (lane & 15)) << 4) + ((lane >> 4) << 2));
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NB; i++) {
        int k4 = part + 8 * i;
        if (k4 < kc / 4) *reinterpret_cast<V4<TM>*>(bs + col * LDB + 4 * k4) = breg[i];
      }
      __syncthreads();
#pragma unroll
      for (int kb = 0; kb < NA; kb++) {
        if (kb < kc / 16) {
          const V4<TM> a4 = areg[kb];
          const V4<TM> b0 = *reinterpret_cast<const V4<TM>*>(bs + (lane & 15) * LDB + kb * 16 + ((lane >> 4) << 2));
          const V4<TM> b1 = *reinterpret_cast<const V4<TM>*>(bs + (16 + (lane & 15)) * LDB + kb * 16 + ((lane >> 4) << 2));
          acc0 = Mfma<TM>::op(a4.x, b0.x, acc0);
          acc1 = Mfma<TM>::op(a4.x, b1.x, acc1);
          acc0 = Mfma<TM>::op(a4.y, b0.y, acc0);
          acc1 = Mfma<TM>::op(a4.y, b1.y, acc1);
          acc0 = Mfma<TM>::op(a4.z, b0.z, acc0);
          acc1 = Mfma<TM>::op(a4.z, b1.z, acc1);
          acc0 = Mfma<TM>::op(a4.w, b0.w, acc0);
          acc1 = Mfma<TM>::op(a4.w, b1.w, acc1);
          if constexpr (sizeof(TM) == 4) {
            if ((kb & 3) == 3) flush();
          }
        }
      }
      if constexpr (sizeof(TM) == 4) flush();
    }
  }
  if constexpr (sizeof(TM) == 8) flush();
  // D layout (16x16), column lane%16; register i holds row 4*(lane/16) + i for
  // the f32 instruction and row lane/16 + 4*i for the f64 one (measured on gfx950).
  TS* __restrict__ out = b.out[lv];
  const double scale = b.scale[lv];
#pragma unroll
  for (int ct = 0; ct < 2; ct++) {
    const int c = cells[ct * 16 + (lane & 15)];
    if (c < 0) continue;
    const double* acc = ct == 0 ? d0 : d1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int row = row0 + (sizeof(TM) == 4 ? 4 * (lane >> 4) + i : (lane >> 4) + 4 * i);
      if (row < b.r) {
        TS val = static_cast<TS>(acc[i] * scale);
        TS* dst = out + static_cast<int64_t>(c) * b.out_stride + row;
        *dst = b.accumulate ? *dst + val : val;
      }
    }
  }
}

// ---- dense M2L on MFMA: 128 operator rows x 64 target cells per workgroup ----
//
// Workgroup = 4 waves in a 2 x 2 grid, each wave 64 rows x 32 cells (4 x 2
// MFMA tiles of 16 x 16). The reduction runs over the active slots of the
// workgroup's split (transfer vectors with at least one source cell among its
// 64 cells) and, inside a slot, over chunks of kK operator columns. Operator
// and multipole chunks are double-buffered in LDS: the global loads of chunk
// t+1 are in flight while chunk t is multiplied.
constexpr int kM2LR = 128;
constexpr int kM2LC = kM2LCols;
static_assert(kM2LC == 64, "the dense M2L kernel maps 64 columns onto 2 x 32-column waves");
// Slots one split may cover; the schedule uses at least
// ceil(kM2LSlots / kM2LSplitSlots) splits so their source table fits in LDS.
constexpr int kM2LSplitSlots = 48;
constexpr int kM2LMinSplits = (kM2LSlots + kM2LSplitSlots - 1) / kM2LSplitSlots;

template <class TM>
struct M2LTile {
  static constexpr int kK = sizeof(TM) == 4 ? 32 : 16;  // reduction chunk
  static constexpr int kKB = kK / 16;                    // 16-wide fragment blocks per chunk
  // Row / column pitches padded by 4 elements: keeps V4 fragment reads aligned
  // and spreads the 16 rows of a fragment over distinct LDS banks.
  // FP32: no padding, 16-byte units XOR-swizzled by row (A) / column (B)
  // instead, so that a workgroup's LDS stays under 53 KB (three per CU).
  static constexpr bool kSwz = sizeof(TM) == 4;
  static constexpr int kAP = kSwz ? 16 : 16 + 4;
  static constexpr int kBP = kSwz ? kK : kK + 4;
  static constexpr int kUpr = static_cast<int>(sizeof(TM));  // 16-byte units per 16-element row
  static constexpr int kAUnits = kKB * kM2LR * kUpr;         // 1024
  static constexpr int kBUnitsPerCol = kK * static_cast<int>(sizeof(TM)) / 16;  // 8
  static_assert(kAUnits == 4 * kThreads, "4 operator loads per thread and chunk");
  static_assert(kBUnitsPerCol == 8, "8 threads per column, 2 columns per thread");
};

template <class TM>
struct M2LDenseArgs {
  int n;
  int r, rpad, kpad, nsplit;
  int64_t per_v;
  int64_t cells_total;
  const int32_t* vlist;
  const int4* sched;
  double* partial;
  const int32_t* col_cell[kMaxM2LLevels];
  const int32_t* blk_oct[kMaxM2LLevels];
  const int32_t* src_slots[kMaxM2LLevels];
  const TM* in[kMaxM2LLevels];
  double* out[kMaxM2LLevels];
  const TM* a[kMaxM2LLevels];
  double scale[kMaxM2LLevels];
  int64_t cell_off[kMaxM2LLevels];
};

// Shared-memory tiles of the M2L main loop (double-buffered operator and
// multipole chunks).
template <class TM>
struct M2LSmem {
  TM as[2][M2LTile<TM>::kKB][kM2LR][M2LTile<TM>::kAP];
  TM bs[2][kM2LC][M2LTile<TM>::kBP];
};

// Main loop shared by the M2L kernels: for the workgroup's 128 operator rows
// (row block rb) and 64 columns, dacc = sum over the nact active transfer
// vectors act_v[ia] of A_v[rows][0:kpad] x in[src_tab[ia][col]][0:kpad]
// (columns with src_tab -1 contribute zero). Chunks of kK reduction indices
// are multiplied on MFMA in the operand type, FP32 partial sums being folded
// into FP64 every 64 indices. With kEmit, the sum of each ia is handed to
// emit(ia, dacc) as soon as it is complete and dacc restarts from zero.
// Ends with a barrier.
struct NoEmit {
  template <class DT>
  __device__ void operator()(int, const DT (&)[4][2][4]) const {}
};

template <class TM, bool kEmit = false, class Emit = NoEmit, class DT = double>
__device__ __forceinline__ void m2l_mainloop(M2LSmem<TM>& sm, const int (*src_tab)[kM2LC],
                                             const int* act_v, int nact, const TM* __restrict__ a,
                                             int64_t per_v, const TM* __restrict__ in, int kpad, int rpad,
                                             int rb, int r, DT (&dacc)[4][2][4], Emit emit = Emit{}) {
  using Tl = M2LTile<TM>;
  using acc_t = typename Mfma<TM>::acc_t;
  constexpr int kK = Tl::kK, kKB = Tl::kKB, kUpr = Tl::kUpr;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int nkc = kpad / kK;
  const int total = nact * nkc;

  // Loader roles: operator units tid + 256 q (q < 4); multipole columns
  // bc0 = tid / 8 and bc1 = 32 + tid / 8, 16-byte piece bq = tid % 8.
  // Every global load of the main loop is issued unconditionally (an absent
  // source reads cell 0 and is zeroed when stored), so the compiler can wait
  // for exactly the register set being stored while the next one stays in
  // flight.
  const int bc0 = tid / 8, bc1 = 32 + tid / 8, bq = tid % 8;
  constexpr int kAU = kM2LR * kUpr;  // operator units per 16-wide block
  static_assert(kAU % kThreads == 0, "operator units of one thread never straddle a block");
  // Two register sets: the chunk two steps ahead is in flight while the next
  // one waits to be stored into the free LDS buffer.
  // Native 4 x 32-bit vectors (not HIP's uint4 wrapper) so the sets stay in
  // registers.
  using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
  struct RegSet {
    u32x4 a[4], b[2];
    uint32_t m0, m1;  // all ones where the loader's column has a source in this chunk's slot
  };
  RegSet r0, r1;

  // Chunk t (clamped to the last one past the end, where its multipole
  // columns are masked to zero so the extra step adds nothing).
  auto load_chunk = [&](int t_req, RegSet& rs) {
    const int t = min(t_req, total - 1);
    const uint32_t live = t_req < total ? ~0u : 0u;
    const int ia = t / nkc, kc = t - ia * nkc;
    const int k0 = kc * kK;
    const TM* av = a + static_cast<int64_t>(act_v[ia]) * per_v;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int seg = (q * kThreads) / kAU;
      const int rem = (q * kThreads) % kAU + tid;
      const int row = rem / kUpr, pc = rem % kUpr;
      const int kb = (k0 >> 4) + seg;
      const TM* g = av + ((static_cast<int64_t>(kb) * rpad + rb * kM2LR + row) << 4);
      rs.a[q] = reinterpret_cast<const u32x4*>(g)[pc];
    }
    const int kel = k0 + bq * (16 / static_cast<int>(sizeof(TM)));
    const int s0 = src_tab[ia][bc0], s1 = src_tab[ia][bc1];
    rs.m0 = s0 >= 0 ? live : 0u;
    rs.m1 = s1 >= 0 ? live : 0u;
    rs.b[0] = *reinterpret_cast<const u32x4*>(in + static_cast<int64_t>(max(s0, 0)) * kpad + kel);
    rs.b[1] = *reinterpret_cast<const u32x4*>(in + static_cast<int64_t>(max(s1, 0)) * kpad + kel);
  };
  auto store_chunk = [&](int buf, const RegSet& rs) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int seg = (q * kThreads) / kAU;
      const int rem = (q * kThreads) % kAU + tid;
      const int row = rem / kUpr, pc = rem % kUpr;
      const int pcs = Tl::kSwz ? (pc ^ ((row >> 2) & 3)) : pc;
      *reinterpret_cast<u32x4*>(&sm.as[buf][seg][row][pcs * (16 / static_cast<int>(sizeof(TM)))]) = rs.a[q];
    }
    // Masked with AND rather than a select: a select of two loaded values
    // becomes a select of addresses, which pins the sets in scratch.
    const int bq0 = Tl::kSwz ? (bq ^ (bc0 & 7)) : bq, bq1 = Tl::kSwz ? (bq ^ (bc1 & 7)) : bq;
    *reinterpret_cast<u32x4*>(&sm.bs[buf][bc0][bq0 * (16 / static_cast<int>(sizeof(TM)))]) = rs.b[0] & rs.m0;
    *reinterpret_cast<u32x4*>(&sm.bs[buf][bc1][bq1 * (16 / static_cast<int>(sizeof(TM)))]) = rs.b[1] & rs.m1;
  };

  acc_t acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      acc[i][j] = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 4; q++) dacc[i][j][q] = 0.0;
    }
  auto flush = [&]() {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          dacc[i][j][q] += static_cast<DT>(acc[i][j][q]);
          acc[i][j][q] = TM(0);
        }
  };

  const int wr = wave >> 1, wc = wave & 1;
  // 16-row tiles of this wave that hold operator rows (< r); tiles wholly in
  // the padding rows are not multiplied (their accumulators stay zero).
  const int nvt = __builtin_amdgcn_readfirstlane(min(4, max(0, (r - rb * kM2LR - wr * 64 + 15) / 16)));
  auto compute = [&](int t) {
    const int buf = t & 1;
    const int kc = t % nkc;
#pragma unroll
    for (int kb = 0; kb < kKB; kb++) {
      using V4T = V4<TM>;
      V4T af[4], bf[2];
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (i < nvt)
          af[i] = *reinterpret_cast<const V4T*>(
              &sm.as[buf][kb][wr * 64 + i * 16 + (lane & 15)]
                    [(Tl::kSwz ? ((lane >> 4) ^ (((lane & 15) >> 2) & 3)) : (lane >> 4)) * 4]);
#pragma unroll
      for (int j = 0; j < 2; j++)
        bf[j] = *reinterpret_cast<const V4T*>(
            &sm.bs[buf][wc * 32 + j * 16 + (lane & 15)]
                  [(Tl::kSwz ? ((kb * 4 + (lane >> 4)) ^ (lane & 7)) : (kb * 4 + (lane >> 4))) * 4]);
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          if (i >= nvt) continue;
          acc[i][j] = Mfma<TM>::op(af[i].x, bf[j].x, acc[i][j]);
          acc[i][j] = Mfma<TM>::op(af[i].y, bf[j].y, acc[i][j]);
          acc[i][j] = Mfma<TM>::op(af[i].z, bf[j].z, acc[i][j]);
          acc[i][j] = Mfma<TM>::op(af[i].w, bf[j].w, acc[i][j]);
        }
    }
    // FP32 partial sums over at most 64 reduction steps are flushed into FP64.
    if constexpr (sizeof(TM) == 4) {
      if ((kc & 1) == 1 || kc == nkc - 1) flush();
    } else if constexpr (kEmit) {
      if (kc == nkc - 1) flush();
    }
    if constexpr (kEmit) {
      if (kc == nkc - 1) {
        emit(t / nkc, dacc);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) dacc[i][j][q] = 0.0;
      }
    }
  };
  // Step t: chunk t+1 sits in registers (cur), chunk t+2 is loaded into the
  // other set (nxt) while chunk t is multiplied from LDS buffer t&1; then
  // chunk t+1 goes to LDS buffer (t+1)&1. Past the end the last chunk is
  // loaded and stored again into the buffer no later step reads, so no load
  // or store sits behind a branch.
  auto step = [&](int t, const RegSet& cur, RegSet& nxt) {
    load_chunk(t + 2, nxt);
    compute(t);
    store_chunk((t + 1) & 1, cur);
    __syncthreads();
  };
  if (total > 0) {
    load_chunk(0, r1);
    store_chunk(0, r1);
    load_chunk(1, r0);
    __syncthreads();
    for (int t = 0; t < total; t += 2) {
      step(t, r0, r1);
      step(t + 1, r1, r0);
    }
  }
  if constexpr (sizeof(TM) == 8) flush();

}

template <class TM>
__global__ void __launch_bounds__(kThreads)
m2l_dense_kernel(M2LDenseArgs<TM> b) {
  __shared__ __align__(16) M2LSmem<TM> sm;
  __shared__ int cells[kM2LC];
  __shared__ uint32_t mask[6];
  __shared__ int act_v[kM2LSplitSlots];
  __shared__ int act_s[kM2LSplitSlots];
  __shared__ int src_tab[kM2LSplitSlots][kM2LC];  // source cell per active slot and column, -1 if none
  __shared__ int nact_s;

  const int4 w = b.sched[blockIdx.x];
  if (w.x < 0) return;
  const int lv = w.x, cb = w.y, rb = w.z, sp = w.w;
  const int32_t* __restrict__ src_slots = b.src_slots[lv];
  const TM* __restrict__ in = b.in[lv];
  const int kpad = b.kpad, rpad = b.rpad;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int oct = b.blk_oct[lv][cb];
  const int s_begin = sp * kM2LSlots / b.nsplit, s_end = (sp + 1) * kM2LSlots / b.nsplit;

  if (tid < kM2LC) cells[tid] = b.col_cell[lv][static_cast<int64_t>(cb) * kM2LC + tid];
  if (tid < 6) mask[tid] = 0u;
  __syncthreads();
  // Active slots of this split: at least one of the 64 cells has a source there.
  for (int idx = tid; idx < kM2LC * (s_end - s_begin); idx += kThreads) {
    int c = idx / (s_end - s_begin), sl = s_begin + idx % (s_end - s_begin);
    int cell = cells[c];
    if (cell >= 0 && src_slots[static_cast<int64_t>(cell) * kM2LSlots + sl] >= 0)
      atomicOr(&mask[sl >> 5], 1u << (sl & 31));
  }
  __syncthreads();
  if (tid == 0) {
    int na = 0;
    for (int sl = s_begin; sl < s_end; sl++) {
      if (mask[sl >> 5] & (1u << (sl & 31))) {
        act_s[na] = sl;
        act_v[na] = b.vlist[oct * kM2LSlots + sl];
        na++;
      }
    }
    nact_s = na;
  }
  __syncthreads();
  const int nact = nact_s;
  for (int idx = tid; idx < nact * kM2LC; idx += kThreads) {
    int ia = idx / kM2LC, c = idx % kM2LC;
    int cell = cells[c];
    src_tab[ia][c] = cell >= 0 ? src_slots[static_cast<int64_t>(cell) * kM2LSlots + act_s[ia]] : -1;
  }
  __syncthreads();
  double dacc[4][2][4];
  m2l_mainloop<TM>(sm, src_tab, act_v, nact, b.a[lv], b.per_v, in, kpad, rpad, rb, b.r, dacc);
  const int wr = wave >> 1, wc = wave & 1;

  // D layout (16 x 16), column lane % 16; register q holds row 4 (lane / 16) + q
  // for the f32 instruction and row lane / 16 + 4 q for the f64 one.
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int cell = cells[wc * 32 + j * 16 + (lane & 15)];
    if (cell < 0) continue;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int row = rb * kM2LR + wr * 64 + i * 16 + (sizeof(TM) == 4 ? 4 * (lane >> 4) + q : (lane >> 4) + 4 * q);
        if (row >= b.r) continue;
        if (b.nsplit == 1) {
          b.out[lv][static_cast<int64_t>(cell) * kpad + row] = dacc[i][j][q] * b.scale[lv];
        } else {
          b.partial[(static_cast<int64_t>(sp) * b.cells_total + b.cell_off[lv] + cell) * rpad + row] = dacc[i][j][q];
        }
      }
  }
}

// Pair-major M2L: a workgroup takes up to kPairBlocks consecutive blocks of
// 64 (target cell, source cell) pairs that share the transfer vector v, for
// one row block, and writes each pair's transfer A_v x in[src] to its own
// partial row (stride rpad, operand type; rows >= r and padding pairs hold
// zeros). The pairs of a block are independent columns, so a pair's result
// does not depend on which block it sits in.
constexpr int kPairBlocks = 4;

template <class TM, class PT>
struct M2LPairArgs {
  int r, rpad, kpad;
  int64_t per_v;
  // (level | (blocks << 5), first block, row block, v) per workgroup; x < 0: idle
  const int4* sched;
  const TM* in[kMaxM2LLevels];
  const TM* a[kMaxM2LLevels];
  const int32_t* blk_src[kMaxM2LLevels];
  PT* partial[kMaxM2LLevels];
};

template <class TM, class PT>
__global__ void __launch_bounds__(kThreads)
m2l_pair_kernel(M2LPairArgs<TM, PT> b) {
  __shared__ __align__(16) M2LSmem<TM> sm;
  __shared__ int src_tab[kPairBlocks][kM2LC];
  __shared__ int act_v[kPairBlocks];
  const int4 w = b.sched[blockIdx.x];
  if (w.x < 0) return;
  const int lv = w.x & 31, nb = w.x >> 5, blk0 = w.y, rb = w.z;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  for (int idx = tid; idx < nb * kM2LC; idx += kThreads)
    src_tab[idx / kM2LC][idx % kM2LC] = b.blk_src[lv][static_cast<int64_t>(blk0) * kM2LC + idx];
  if (tid < nb) act_v[tid] = w.w;
  __syncthreads();
  const int wr = wave >> 1, wc = wave & 1;
  const int rpad = b.rpad;
  PT* __restrict__ out = b.partial[lv];
  // D layout (16 x 16), column lane % 16; register q holds row 4 (lane / 16) + q
  // for the f32 instruction and row lane / 16 + 4 q for the f64 one.
  // FP32 operands: the partial sums of 64 reduction steps are added in FP32
  // (the per-pair transfer is stored in FP32 anyway), which keeps the
  // workgroup within three per CU.
  using DT = std::conditional_t<sizeof(TM) == 4 && sizeof(PT) == 4, float, double>;
  auto emit = [&](int ia, const DT (&d)[4][2][4]) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int col = wc * 32 + j * 16 + (lane & 15);
      PT* __restrict__ po = out + (static_cast<int64_t>(blk0 + ia) * kM2LC + col) * rpad;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r0 = rb * kM2LR + wr * 64 + i * 16;
        if (r0 >= rpad) continue;  // the half block past a 64-row padding
        if constexpr (sizeof(TM) == 4 && sizeof(PT) == 4) {
          using f32x4 = float __attribute__((ext_vector_type(4)));
          f32x4 v = {static_cast<float>(d[i][j][0]), static_cast<float>(d[i][j][1]),
                     static_cast<float>(d[i][j][2]), static_cast<float>(d[i][j][3])};
          *reinterpret_cast<f32x4*>(po + r0 + 4 * (lane >> 4)) = v;
        } else if constexpr (sizeof(TM) == 4) {
          using f64x2 = double __attribute__((ext_vector_type(2)));
          *reinterpret_cast<f64x2*>(po + r0 + 4 * (lane >> 4)) = f64x2{d[i][j][0], d[i][j][1]};
          *reinterpret_cast<f64x2*>(po + r0 + 4 * (lane >> 4) + 2) = f64x2{d[i][j][2], d[i][j][3]};
        } else {
#pragma unroll
          for (int q = 0; q < 4; q++) po[r0 + (lane >> 4) + 4 * q] = d[i][j][q];
        }
      }
    }
  };
  DT dacc[4][2][4];
  m2l_mainloop<TM, true, decltype(emit), DT>(sm, src_tab, act_v, nb, b.a[lv], b.per_v, b.in[lv], b.kpad, rpad, rb,
                                             b.r, dacc, emit);
}

// out[cell][row] = scale * sum over the cell's slots, in slot order, of the
// partial row of its (cell, slot) pair. One workgroup per target cell.
template <class TM>
struct M2LPairReduceArgs {  // TM: partial type
  int n, r, kpad, pstride;
  int64_t cell_off[kMaxM2LLevels + 1];
  const int32_t* pair_pos[kMaxM2LLevels];
  const TM* partial[kMaxM2LLevels];
  double* out[kMaxM2LLevels];
  double scale[kMaxM2LLevels];
};

constexpr int kPairReduceThreads = 128;

template <class TM, int kRows>
__global__ void __launch_bounds__(kPairReduceThreads) m2l_pair_reduce_kernel(M2LPairReduceArgs<TM> b) {
  __shared__ int32_t pos[kM2LSlots + 7];
  const int64_t g = blockIdx.x;
  int k = 0;
  while (k + 1 < b.n && b.cell_off[k + 1] <= g) k++;
  const int64_t cell = g - b.cell_off[k];
  for (int sl = threadIdx.x; sl < kM2LSlots + 7; sl += kPairReduceThreads)
    pos[sl] = sl < kM2LSlots ? b.pair_pos[k][cell * kM2LSlots + sl] : -1;
  __syncthreads();
  const TM* __restrict__ partial = b.partial[k];
  // kRows rows per thread (three for dense operators, one for compressed
  // ones of at most 128 rows) and eight slots per trip keep up to 24 loads in
  // flight; the additions still run in slot order.
  constexpr int kSl = 8;
  for (int row0 = 0; row0 < b.r; row0 += kRows * kPairReduceThreads) {
    double acc[kRows];
#pragma unroll
    for (int q = 0; q < kRows; q++) acc[q] = 0.0;
    int row[kRows];
#pragma unroll
    for (int q = 0; q < kRows; q++) row[q] = min(row0 + q * kPairReduceThreads + static_cast<int>(threadIdx.x), b.r - 1);
    for (int sl = 0; sl < kM2LSlots; sl += kSl) {
      // Loads are unconditional (slots without a pair read row 0) and masked
      // afterwards: loads behind branches would each be waited for on their own.
      using UT = std::conditional_t<sizeof(TM) == 4, uint32_t, uint64_t>;  // TM: partial type
      UT x[kSl][kRows], m[kSl];
#pragma unroll
      for (int u = 0; u < kSl; u++) {
        const int32_t p = pos[sl + u];
        m[u] = p >= 0 ? ~UT(0) : UT(0);
        const UT* pr = reinterpret_cast<const UT*>(partial) + static_cast<int64_t>(max(p, 0)) * b.pstride;
#pragma unroll
        for (int q = 0; q < kRows; q++) x[u][q] = pr[row[q]];
      }
#pragma unroll
      for (int u = 0; u < kSl; u++)
#pragma unroll
        for (int q = 0; q < kRows; q++)
          acc[q] += static_cast<double>(__builtin_bit_cast(TM, static_cast<UT>(x[u][q] & m[u])));
    }
#pragma unroll
    for (int q = 0; q < kRows; q++) {
      const int rr = row0 + q * kPairReduceThreads + static_cast<int>(threadIdx.x);
      if (rr < b.r) b.out[k][cell * b.kpad + rr] = acc[q] * b.scale[k];
    }
  }
}

// out[cell][row] = scale * sum_s partial[s][cell][row], splits summed in order.
struct M2LReduceArgs {
  int n, r, rpad, kpad, nsplit;
  int64_t cells_total;
  const double* partial;
  double* out[kMaxM2LLevels];
  double scale[kMaxM2LLevels];
  int64_t cell_off[kMaxM2LLevels + 1];
};

__global__ void m2l_reduce_kernel(M2LReduceArgs b) {
  int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t gcell = idx / b.rpad;
  int row = static_cast<int>(idx % b.rpad);
  if (gcell >= b.cells_total || row >= b.r) return;
  int lv = 0;
  while (lv + 1 < b.n && gcell >= b.cell_off[lv + 1]) lv++;
  double sum = 0.0;
  for (int s = 0; s < b.nsplit; s++) sum += b.partial[(s * b.cells_total + gcell) * b.rpad + row];
  b.out[lv][(gcell - b.cell_off[lv]) * b.kpad + row] = sum * b.scale[lv];
}

__global__ void convert_f64_f32_kernel(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<float>(in[i]);
}

}  // namespace

void launch_convert_f64_f32(const double* in, float* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  convert_f64_f32_kernel<<<ceil_div(n, 256), 256, 0, s>>>(in, out, n);
  PAM_LAUNCH_CHECK();
}

namespace {

// Work items grouped by (octant, row block, split); groups dealt round-robin
// onto the 8 XCDs (workgroup i runs on XCD i % 8), so the workgroups that
// stream the same operator rows share that XCD's L2.
void build_m2l_schedule(const std::vector<const M2LLevel*>& lvls, int nrowblk, M2LSchedule& sc,
                        hipStream_t s) {
  constexpr int kXcd = 8;
  int64_t items = 0;
  for (const M2LLevel* l : lvls) items += (l->n_cols / kM2LC) * nrowblk;
  // Enough workgroups to fill 256 CUs about sixteen times over: with the
  // active-slot mask the work per workgroup varies, and many small splits
  // balance the tail better than a few large ones (each split still covers
  // at least three of the 189 slots).
  int nsplit = items > 0 ? static_cast<int>((4096 + items - 1) / items) : 1;
  sc.nsplit = std::clamp(nsplit, kM2LMinSplits, 48);
  // Column blocks per octant in (level, block) order; a group (octant, row
  // block, split) holds all of its octant's blocks. Groups go round-robin to
  // the XCDs in (octant, row block, split) order.
  std::array<std::vector<std::array<int32_t, 2>>, 8> per_oct;
  for (size_t li = 0; li < lvls.size(); li++) {
    const M2LLevel& l = *lvls[li];
    for (int64_t cb = 0; cb < l.n_cols / kM2LC; cb++)
      per_oct[l.h_blk_oct[cb]].push_back({static_cast<int32_t>(li), static_cast<int32_t>(cb)});
  }
  std::array<size_t, kXcd> len_x{};
  int g = 0;
  for (int oct = 0; oct < 8; oct++) {
    if (per_oct[oct].empty()) continue;
    for (int rb = 0; rb < nrowblk; rb++)
      for (int sp = 0; sp < sc.nsplit; sp++) len_x[g++ % kXcd] += per_oct[oct].size();
  }
  size_t len = *std::max_element(len_x.begin(), len_x.end());
  int32_t* h = sc.h.acquire(len * kXcd * 4);
  std::fill(h, h + len * kXcd * 4, -1);
  std::array<size_t, kXcd> pos{};
  g = 0;
  for (int oct = 0; oct < 8; oct++) {
    if (per_oct[oct].empty()) continue;
    for (int rb = 0; rb < nrowblk; rb++)
      for (int sp = 0; sp < sc.nsplit; sp++) {
        const int x = g++ % kXcd;
        for (const auto& it : per_oct[oct]) {
          int32_t* e = h + (pos[x]++ * kXcd + x) * 4;
          e[0] = it[0];
          e[1] = it[1];
          e[2] = rb;
          e[3] = sp;
        }
      }
  }
  sc.n_wg = static_cast<int>(len * kXcd);
  sc.cells_total = 0;
  for (const M2LLevel* l : lvls) sc.cells_total += l->n_cells;
  sc.h.upload(sc.d, len * kXcd * 4, s);
  sc.valid = true;
}

}  // namespace

template <class TM>
void launch_m2l_dense(const M2LOperator& op, const std::vector<M2LDenseLevel<TM>>& levels,
                      const int32_t* vlist, M2LSchedule& sched, DevBuf<double>& partial, hipStream_t s) {
  // Rows padded to 64; a final half block of 64 rows reads into the operator's
  // slack (64 zero rows after its last transfer vector) and writes no partials.
  if (op.rpad % 64 != 0) throw_runtime("M2L operator rows are not padded to 64");
  if (op.kpad % M2LTile<TM>::kK != 0) throw_runtime("M2L reduction length is not a multiple of the chunk");
  M2LDenseArgs<TM> b{};
  b.r = op.r;
  b.rpad = op.rpad;
  b.kpad = op.kpad;
  b.per_v = static_cast<int64_t>(op.per_v);
  b.vlist = vlist;
  std::vector<const M2LLevel*> lvls;
  int64_t off = 0;
  for (const auto& l : levels) {
    if (!l.lv || l.lv->n_cols == 0) continue;
    if (static_cast<int64_t>(l.lv->h_blk_oct.size()) != l.lv->n_cols / kM2LC)
      throw_runtime("M2L column blocks without octants");
    if (b.n >= kMaxM2LLevels) throw_runtime("too many M2L levels");
    int k = b.n++;
    b.col_cell[k] = l.lv->col_cell.get();
    b.blk_oct[k] = l.lv->blk_oct.get();
    b.src_slots[k] = l.lv->src.get();
    b.in[k] = l.in;
    b.out[k] = l.out;
    b.a[k] = l.a;
    b.scale[k] = l.scale;
    b.cell_off[k] = off;
    off += l.lv->n_cells;
    lvls.push_back(l.lv);
  }
  if (b.n == 0) return;
  if (!sched.valid) build_m2l_schedule(lvls, op.rpad / kM2LR, sched, s);
  if (sched.cells_total != off) throw_runtime("stale M2L schedule");
  b.nsplit = sched.nsplit;
  b.cells_total = sched.cells_total;
  b.sched = reinterpret_cast<const int4*>(sched.d.get());
  if (sched.nsplit > 1) {
    partial.resize(static_cast<size_t>(sched.nsplit) * sched.cells_total * op.rpad);
    b.partial = partial.get();
  }
  m2l_dense_kernel<TM><<<static_cast<unsigned>(sched.n_wg), kThreads, 0, s>>>(b);
  PAM_LAUNCH_CHECK();
  if (sched.nsplit > 1) {
    M2LReduceArgs rb{};
    rb.n = b.n;
    rb.r = op.r;
    rb.rpad = op.rpad;
    rb.kpad = op.kpad;
    rb.nsplit = sched.nsplit;
    rb.cells_total = sched.cells_total;
    rb.partial = partial.get();
    for (int k = 0; k < b.n; k++) {
      rb.out[k] = b.out[k];
      rb.scale[k] = b.scale[k];
      rb.cell_off[k] = b.cell_off[k];
    }
    rb.cell_off[b.n] = sched.cells_total;
    int64_t work = sched.cells_total * op.rpad;
    m2l_reduce_kernel<<<ceil_div(work, 256), 256, 0, s>>>(rb);
    PAM_LAUNCH_CHECK();
  }
}

namespace {

// Pair-major work list: workgroups grouped by (level, v, row block), groups
// dealt round-robin onto the 8 XCDs so that the workgroups streaming the same
// operator rows share an L2.
void build_m2l_pair_schedule(const std::vector<const M2LLevel*>& lvls, int nrowblk,
                             M2LPairSchedule& sc, hipStream_t s) {
  constexpr int kXcd = 8;
  auto items_of = [](int blocks) { return (blocks + kPairBlocks - 1) / kPairBlocks; };
  std::array<size_t, kXcd> len_x{};
  int g = 0;
  for (const M2LLevel* l : lvls)
    for (int v = 0; v < 343; v++) {
      if (l->h_v_blocks[v] == 0) continue;
      for (int rb = 0; rb < nrowblk; rb++) len_x[g++ % kXcd] += items_of(l->h_v_blocks[v]);
    }
  const size_t len = *std::max_element(len_x.begin(), len_x.end());
  int32_t* h = sc.h.acquire(len * kXcd * 4);
  std::fill(h, h + len * kXcd * 4, -1);
  std::array<size_t, kXcd> pos{};
  g = 0;
  for (size_t li = 0; li < lvls.size(); li++) {
    const M2LLevel& l = *lvls[li];
    for (int v = 0; v < 343; v++) {
      const int nblk = l.h_v_blocks[v];
      if (nblk == 0) continue;
      for (int rb = 0; rb < nrowblk; rb++) {
        const int x = g++ % kXcd;
        for (int bk = 0; bk < nblk; bk += kPairBlocks) {
          int32_t* e = h + (pos[x]++ * kXcd + x) * 4;
          e[0] = static_cast<int32_t>(li) | (std::min(kPairBlocks, nblk - bk) << 5);
          e[1] = l.h_v_first[v] + bk;
          e[2] = rb;
          e[3] = v;
        }
      }
    }
  }
  sc.n_wg = static_cast<int>(len * kXcd);
  sc.h.upload(sc.d, len * kXcd * 4, s);
  sc.valid = true;
}

}  // namespace

template <class TM, class PT>
void launch_m2l_pairs(const M2LOperator& op, const std::vector<M2LDenseLevel<TM>>& levels,
                      M2LPairSchedule& sched, DevBuf<PT>& partial, hipStream_t s) {
  // Rows padded to 64; a final half block of 64 rows reads into the operator's
  // slack (64 zero rows after its last transfer vector) and writes no partials.
  if (op.rpad % 64 != 0) throw_runtime("M2L operator rows are not padded to 64");
  if (op.kpad % M2LTile<TM>::kK != 0) throw_runtime("M2L reduction length is not a multiple of the chunk");
  M2LPairArgs<TM, PT> b{};
  M2LPairReduceArgs<PT> rb{};
  b.r = rb.r = op.r;
  b.rpad = rb.pstride = op.rpad;
  b.kpad = rb.kpad = op.kpad;
  b.per_v = static_cast<int64_t>(op.per_v);
  std::vector<const M2LLevel*> lvls;
  int64_t pairs = 0, cells = 0;
  for (const auto& l : levels) {
    if (!l.lv || l.lv->n_cells == 0) continue;
    if (static_cast<int>(l.lv->h_v_blocks.size()) != 343) throw_runtime("M2L pair lists missing");
    pairs += l.lv->n_blk * kM2LC;
  }
  partial.resize(static_cast<size_t>(std::max<int64_t>(pairs, 1)) * op.rpad);
  pairs = 0;
  for (const auto& l : levels) {
    if (!l.lv || l.lv->n_cells == 0) continue;
    if (lvls.size() >= static_cast<size_t>(std::min(kMaxM2LLevels, 32))) throw_runtime("too many M2L levels");
    const int k = static_cast<int>(lvls.size());
    b.in[k] = l.in;
    b.a[k] = l.a;
    b.blk_src[k] = l.lv->blk_src.get();
    b.partial[k] = partial.get() + pairs * op.rpad;
    rb.pair_pos[k] = l.lv->pair_pos.get();
    rb.partial[k] = b.partial[k];
    rb.out[k] = l.out;
    rb.scale[k] = l.scale;
    rb.cell_off[k] = cells;
    pairs += l.lv->n_blk * kM2LC;
    cells += l.lv->n_cells;
    lvls.push_back(l.lv);
  }
  if (lvls.empty()) return;
  rb.n = static_cast<int>(lvls.size());
  rb.cell_off[rb.n] = cells;
  if (!sched.valid) build_m2l_pair_schedule(lvls, (op.rpad + kM2LR - 1) / kM2LR, sched, s);
  b.sched = reinterpret_cast<const int4*>(sched.d.get());
  if (sched.n_wg > 0) {
    m2l_pair_kernel<TM, PT><<<static_cast<unsigned>(sched.n_wg), kThreads, 0, s>>>(b);
    PAM_LAUNCH_CHECK();
  }
  if (op.r <= kPairReduceThreads)
    m2l_pair_reduce_kernel<PT, 1><<<static_cast<unsigned>(cells), kPairReduceThreads, 0, s>>>(rb);
  else
    m2l_pair_reduce_kernel<PT, 3><<<static_cast<unsigned>(cells), kPairReduceThreads, 0, s>>>(rb);
  PAM_LAUNCH_CHECK();
}

template void launch_m2l_pairs<float, float>(const M2LOperator&, const std::vector<M2LDenseLevel<float>>&,
                                             M2LPairSchedule&, DevBuf<float>&, hipStream_t);
template void launch_m2l_pairs<double, double>(const M2LOperator&, const std::vector<M2LDenseLevel<double>>&,
                                               M2LPairSchedule&, DevBuf<double>&, hipStream_t);

namespace {

template <class T>
__global__ void build_m2l_dense_kernel(int p, int kind, double pa, double pb, int rpad, int kpad,
                                       T* __restrict__ a) {
  int64_t per_v = static_cast<int64_t>(rpad) * kpad;
  int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= 343 * per_v) return;
  int v = static_cast<int>(idx / per_v);
  int64_t rem = idx % per_v;
  int kb = static_cast<int>(rem / (static_cast<int64_t>(rpad) * 16));
  int row = static_cast<int>((rem / 16) % rpad);
  int k = kb * 16 + static_cast<int>(rem % 16);
  int p3 = p * p * p;
  int dx = v / 49 - 3, dy = (v / 7) % 7 - 3, dz = v % 7 - 3;
  bool adj = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
  T val = T(0);
  if (!adj && row < p3 && k < p3) {
    int ta = row / (p * p), tb = (row / p) % p, tc = row % p;
    int sa = k / (p * p), sb = (k / p) % p, sc = k % p;
    // source cell centre = target centre + 2 d (local units of the cell half width)
    double ex = cospi((2.0 * ta + 1) / (2.0 * p)) - 2.0 * dx - cospi((2.0 * sa + 1) / (2.0 * p));
    double ey = cospi((2.0 * tb + 1) / (2.0 * p)) - 2.0 * dy - cospi((2.0 * sb + 1) / (2.0 * p));
    double ez = cospi((2.0 * tc + 1) / (2.0 * p)) - 2.0 * dz - cospi((2.0 * sc + 1) / (2.0 * p));
    double r2 = ex * ex + ey * ey + ez * ez;
    double f;
    if (kind == kBh3) f = pa * sqrt(r2 + pb);
    else if (kind == kTh3) { double rho = sqrt(r2 + pb); f = pa * rho * rho * rho; }
    else f = pa * exp(pb * sqrt(r2));
    val = static_cast<T>(f);
  }
  a[idx] = val;
}

template <class F>
void dispatch_order(int p, F&& f) {
  switch (p) {
    case 3: f(std::integral_constant<int, 3>{}); return;
    case 4: f(std::integral_constant<int, 4>{}); return;
    case 5: f(std::integral_constant<int, 5>{}); return;
    case 6: f(std::integral_constant<int, 6>{}); return;
    case 7: f(std::integral_constant<int, 7>{}); return;
    case 8: f(std::integral_constant<int, 8>{}); return;
    case 9: f(std::integral_constant<int, 9>{}); return;
    case 10: f(std::integral_constant<int, 10>{}); return;
    case 11: f(std::integral_constant<int, 11>{}); return;
    case 12: f(std::integral_constant<int, 12>{}); return;
  }
  throw_invalid("unsupported interpolation order");
}

}  // namespace

template <class T>
void launch_prepare_points(const Tree& t, const double* d_pts, const double* d_w, V4<T>* out,
                           hipStream_t s) {
  int64_t n = t.n_points;
  if (n == 0) return;
  BoxArgs b{{t.box.lo[0], t.box.lo[1], t.box.lo[2]}, t.box.width / std::ldexp(1.0, t.level)};
  prepare_points_kernel<T><<<ceil_div(n, 256), 256, 0, s>>>(d_pts, d_w, t.perm.get(), t.point_leaf.get(),
                                                            t.levels[t.level].codes.get(), n, b, out);
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_set_point_weights(const Tree& t, const double* d_w, V4<T>* pts, hipStream_t s) {
  int64_t n = t.n_points;
  if (n == 0) return;
  set_weights_kernel<T><<<ceil_div(n, 256), 256, 0, s>>>(d_w, t.perm.get(), n, pts);
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_p2m(int p, const Tree& t, const V4<T>* pts, T* m_leaf, hipStream_t s) {
  int64_t nl = t.n_leaves();
  if (nl == 0) return;
  dispatch_order(p, [&](auto P) {
    p2m_kernel<T, decltype(P)::value><<<static_cast<unsigned>(nl), kThreads, 0, s>>>(
        t.leaf_start.get(), pts, m_leaf, kpad_of(P));
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_m2m(int p, const Tree& t, int l, const T* m_child, T* m_parent, hipStream_t s) {
  int64_t n = t.levels[l].n;
  if (n == 0) return;
  dispatch_order(p, [&](auto P) {
    m2m_kernel<T, decltype(P)::value><<<static_cast<unsigned>(n), kThreads, 0, s>>>(
        t.levels[l].child_start.get(), t.levels[l + 1].codes.get(), m_child, m_parent, kpad_of(P));
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_l2l(int p, const Tree& t, int l, const T* l_parent, T* l_child, hipStream_t s) {
  int64_t n = t.levels[l].n;
  if (n == 0) return;
  dispatch_order(p, [&](auto P) {
    l2l_kernel<T, decltype(P)::value><<<static_cast<unsigned>(n), kThreads, 0, s>>>(
        t.levels[l].parent.get(), t.levels[l].codes.get(), l_parent, l_child, kpad_of(P));
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
L2PJob<T> make_l2p_job(const Tree& t, const T* l_leaf, const V4<T>* pts, T* y_far, bool add_to_output, int kpad) {
  L2PJob<T> j{};
  j.leaf_start = t.leaf_start.get();
  j.pts = pts;
  j.l = l_leaf;
  j.kpad = kpad;
  j.y_far = y_far;
  j.perm = add_to_output ? t.perm.get() : nullptr;
  j.n_leaves = static_cast<int>(t.n_leaves());
  return j;
}

template <class T>
P2PJob<T> make_p2p_job(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists, const V4<T>* tpts,
                       const V4<T>* spts, const double* y_far, double* out) {
  P2PJob<T> j{};
  const double h = trg.box.width / std::ldexp(1.0, trg.level) / 2.0;
  const PhiConsts pc = phi_consts(rbf, h);
  j.tls = trg.leaf_start.get();
  j.sls = src.leaf_start.get();
  j.p2p = lists.p2p.get();
  j.tpts = tpts;
  j.spts = spts;
  j.y_far = y_far;
  j.tperm = trg.perm.get();
  j.pa = static_cast<T>(pc.a);
  j.pb = static_cast<T>(pc.b);
  j.out = out;
  j.n_leaves = static_cast<int>(trg.n_leaves());
  return j;
}

template <class T>
void launch_l2p_jobs(int p, const L2PJob<T>* jobs, int n, hipStream_t s) {
  if (n < 1 || n > kMaxFmmJobs) throw_runtime("L2P job count out of range");
  L2PJobs<T> t{};
  int64_t total = 0;
  for (int k = 0; k < n; k++) {
    t.j[t.n] = jobs[k];
    t.j[t.n].first_leaf = static_cast<int>(total);
    total += jobs[k].n_leaves;
    if (jobs[k].n_leaves > 0) t.n++;
  }
  if (total == 0) return;
  if (total > INT32_MAX) throw_runtime("too many leaves in one launch");
  dispatch_order(p, [&](auto P) {
    l2p_jobs_kernel<T, decltype(P)::value><<<static_cast<unsigned>(total), kThreads, 0, s>>>(t);
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_p2p_jobs(int kind, bool compensated, const P2PJob<T>* jobs, int n, hipStream_t s) {
  if (n < 1 || n > kMaxFmmJobs) throw_runtime("P2P job count out of range");
  P2PJobs<T> t{};
  int64_t total = 0;
  for (int k = 0; k < n; k++) {
    t.j[t.n] = jobs[k];
    t.j[t.n].first_leaf = static_cast<int>(total);
    total += jobs[k].n_leaves;
    if (jobs[k].n_leaves > 0) t.n++;
  }
  if (total == 0) return;
  if (total > INT32_MAX) throw_runtime("too many leaves in one launch");
  dispatch_kind(kind, [&](auto K) {
    if constexpr (sizeof(T) == 4) {
      if (compensated) {
        p2p_jobs_kernel<T, decltype(K)::value, true><<<static_cast<unsigned>(total), kThreads, 0, s>>>(t);
        return;
      }
    }
    p2p_jobs_kernel<T, decltype(K)::value, false><<<static_cast<unsigned>(total), kThreads, 0, s>>>(t);
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_l2p(int p, const Tree& t, const T* l_leaf, const V4<T>* pts, T* y_far, hipStream_t s,
                bool add_to_output) {
  if (t.n_leaves() == 0) return;
  const L2PJob<T> j = make_l2p_job<T>(t, l_leaf, pts, y_far, add_to_output, kpad_of(p));
  launch_l2p_jobs<T>(p, &j, 1, s);
}

template <class T>
void launch_p2p(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists,
                const V4<T>* tpts, const V4<T>* spts, const double* y_far, double* out, hipStream_t s,
                bool compensated) {
  if (trg.n_leaves() == 0) return;
  const P2PJob<T> j = make_p2p_job<T>(rbf, trg, src, lists, tpts, spts, y_far, out);
  launch_p2p_jobs<T>(rbf.kind, compensated, &j, 1, s);
}

void launch_deferred(const std::vector<const DeferredFmm*>& jobs, hipStream_t s) {
  // Near fields first (each writes its output), grouped by precision, kernel
  // kind and summation form; then the far fields (added to the outputs),
  // grouped by order. Each group goes out in launches of kMaxFmmJobs.
  auto flush_p2p = [&](auto tag, int kind, bool comp) {
    using T = decltype(tag);
    std::vector<P2PJob<T>> g;
    for (const DeferredFmm* d : jobs) {
      if (!d || !d->valid || d->kind != kind || d->compensated != comp || (sizeof(T) == 4) != d->fp32) continue;
      if constexpr (sizeof(T) == 4) g.push_back(d->p2p32);
      else g.push_back(d->p2p64);
      if (static_cast<int>(g.size()) == kMaxFmmJobs) {
        launch_p2p_jobs<T>(kind, comp, g.data(), kMaxFmmJobs, s);
        g.clear();
      }
    }
    if (!g.empty()) launch_p2p_jobs<T>(kind, comp, g.data(), static_cast<int>(g.size()), s);
  };
  for (int kind : {kBh3, kTh3, kExp}) {
    flush_p2p(0.0, kind, false);
    flush_p2p(0.0f, kind, false);
    flush_p2p(0.0f, kind, true);
  }
  for (int p = kMinOrder; p <= kMaxOrder; p++) {
    std::vector<L2PJob<double>> g;
    for (const DeferredFmm* d : jobs) {
      if (!d || !d->valid || d->order != p) continue;
      g.push_back(d->l2p);
      if (static_cast<int>(g.size()) == kMaxFmmJobs) {
        launch_l2p_jobs<double>(p, g.data(), kMaxFmmJobs, s);
        g.clear();
      }
    }
    if (!g.empty()) launch_l2p_jobs<double>(p, g.data(), static_cast<int>(g.size()), s);
  }
}

void launch_l2p_grad(int p, const Tree& t, const double* l_leaf, const V4<double>* pts, const double* aniso,
                     double* out3, hipStream_t s) {
  int64_t nl = t.n_leaves();
  if (nl == 0) return;
  const double h = t.box.width / std::ldexp(1.0, t.level) / 2.0;
  Aniso3 A;
  for (int k = 0; k < 9; k++) A.a[k] = aniso[k];
  dispatch_order(p, [&](auto P) {
    l2p_grad_kernel<decltype(P)::value><<<static_cast<unsigned>(nl), kL2PGThreads, 0, s>>>(
        t.leaf_start.get(), pts, l_leaf, kpad_of(P), 1.0 / h, A, t.perm.get(), out3);
  });
  PAM_LAUNCH_CHECK();
}

void launch_l2p_hess_col(int p, const Tree& t, const double* l_leaf, const V4<double>* pts, const double* aniso,
                         int kcol, double* out3, hipStream_t s) {
  int64_t nl = t.n_leaves();
  if (nl == 0) return;
  if (kcol < 0 || kcol > 2) throw_invalid("Hessian column out of range");
  const double h = t.box.width / std::ldexp(1.0, t.level) / 2.0;
  Aniso3 A;
  for (int k = 0; k < 9; k++) A.a[k] = aniso[k];
  dispatch_order(p, [&](auto P) {
    l2p_hess_col_kernel<decltype(P)::value><<<static_cast<unsigned>(nl), kL2PHThreads, 0, s>>>(
        t.leaf_start.get(), pts, l_leaf, kpad_of(P), 1.0 / (h * h), A, kcol, t.perm.get(), out3);
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_p2p_hess_col(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists,
                         const V4<T>* tpts, const V4<T>* spts, int kcol, double* out3, hipStream_t s) {
  int64_t nl = trg.n_leaves();
  if (nl == 0) return;
  if (kcol < 0 || kcol > 2) throw_invalid("Hessian column out of range");
  const double h = trg.box.width / std::ldexp(1.0, trg.level) / 2.0;
  const PhiConsts gc = grad_consts(rbf, h);
  Aniso3 A;
  for (int k = 0; k < 9; k++) A.a[k] = rbf.aniso[k];
  dispatch_kind(rbf.kind, [&](auto K) {
    p2p_hess_col_kernel<T, decltype(K)::value><<<static_cast<unsigned>(nl), kThreads, 0, s>>>(
        trg.leaf_start.get(), src.leaf_start.get(), lists.p2p.get(), tpts, spts, trg.perm.get(),
        static_cast<T>(gc.a / h), static_cast<T>(gc.b), A, kcol, out3);
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void launch_p2p_grad(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists,
                     const V4<T>* tpts, const V4<T>* spts, double* out3, hipStream_t s) {
  int64_t nl = trg.n_leaves();
  if (nl == 0) return;
  const double h = trg.box.width / std::ldexp(1.0, trg.level) / 2.0;
  const PhiConsts gc = grad_consts(rbf, h);
  Aniso3 A;
  for (int k = 0; k < 9; k++) A.a[k] = rbf.aniso[k];
  dispatch_kind(rbf.kind, [&](auto K) {
    p2p_grad_kernel<T, decltype(K)::value><<<static_cast<unsigned>(nl), kThreads, 0, s>>>(
        trg.leaf_start.get(), src.leaf_start.get(), lists.p2p.get(), tpts, spts, trg.perm.get(),
        static_cast<T>(gc.a), static_cast<T>(gc.b), A, out3);
  });
  PAM_LAUNCH_CHECK();
}

template <class T>
void build_m2l_dense(int p, const RbfDesc& rbf, double h, M2LOperator& op, DevBuf<T>& a,
                     hipStream_t s) {
  op.r = p * p * p;
  op.rpad = rpad_of(op.r);
  op.kpad = kpad_of(p);
  op.per_v = static_cast<size_t>(op.rpad) * op.kpad;
  a.resize(343 * op.per_v);
  PhiConsts pc = phi_consts(rbf, h);
  int64_t total = static_cast<int64_t>(343) * op.per_v;
  build_m2l_dense_kernel<T><<<ceil_div(total, 256), 256, 0, s>>>(p, rbf.kind, pc.a, pc.b, op.rpad,
                                                                 op.kpad, a.get());
  PAM_LAUNCH_CHECK();
}

template <class TM, class TS>
void launch_m2l_batch(const M2LOperator& op, const std::vector<M2LLevelSpec<TM, TS>>& levels,
                      const int32_t* vlist, bool accumulate, int in_stride, int out_stride, hipStream_t s) {
  if (op.kpad % 16 != 0 || in_stride % 4 != 0) throw_runtime("m2l: bad padding");
  if (levels.size() > static_cast<size_t>(kMaxM2LLevels)) throw_runtime("m2l: too many levels");
  M2LBatch<TM, TS> b{};
  b.r = op.r;
  b.rpad = op.rpad;
  b.kpad = op.kpad;
  b.per_v = static_cast<int64_t>(op.per_v);
  b.vlist = vlist;
  b.accumulate = accumulate ? 1 : 0;
  b.in_stride = in_stride;
  b.out_stride = out_stride;
  int blocks = 0;
  for (const auto& l : levels) {
    if (!l.lv || l.lv->n_cols == 0) continue;
    int k = b.n++;
    b.blk_begin[k] = blocks;
    b.col_cell[k] = l.lv->col_cell.get();
    b.blk_oct[k] = l.src_slots ? l.lv->blk_oct.get() : nullptr;
    b.src_slots[k] = l.src_slots;
    b.in[k] = l.in;
    b.out[k] = l.out;
    b.a[k] = l.a;
    b.scale[k] = l.scale;
    blocks += static_cast<int>(l.lv->n_cols / kM2LBCols);
  }
  b.blk_begin[b.n] = blocks;
  if (blocks == 0) return;
  dim3 grid(static_cast<unsigned>(blocks), static_cast<unsigned>(op.rpad / 64));
  m2l_batch_kernel<TM, TS><<<grid, kThreads, 0, s>>>(b);
  PAM_LAUNCH_CHECK();
}

#define PAM_INSTANTIATE(T)                                                                          \
  template void launch_prepare_points<T>(const Tree&, const double*, const double*, V4<T>*,        \
                                         hipStream_t);                                               \
  template void launch_set_point_weights<T>(const Tree&, const double*, V4<T>*, hipStream_t);       \
  template void launch_p2m<T>(int, const Tree&, const V4<T>*, T*, hipStream_t);                     \
  template void launch_m2m<T>(int, const Tree&, int, const T*, T*, hipStream_t);                    \
  template void launch_l2l<T>(int, const Tree&, int, const T*, T*, hipStream_t);                    \
  template void launch_l2p<T>(int, const Tree&, const T*, const V4<T>*, T*, hipStream_t, bool);           \
  template L2PJob<T> make_l2p_job<T>(const Tree&, const T*, const V4<T>*, T*, bool, int);                   \
  template P2PJob<T> make_p2p_job<T>(const RbfDesc&, const Tree&, const Tree&, const Lists&, const V4<T>*,  \
                                     const V4<T>*, const double*, double*);                                 \
  template void launch_p2p<T>(const RbfDesc&, const Tree&, const Tree&, const Lists&, const V4<T>*, \
                              const V4<T>*, const double*, double*, hipStream_t, bool);                 \
  template void launch_p2p_grad<T>(const RbfDesc&, const Tree&, const Tree&, const Lists&, const V4<T>*, \
                                   const V4<T>*, double*, hipStream_t);                                  \
  template void launch_p2p_hess_col<T>(const RbfDesc&, const Tree&, const Tree&, const Lists&, const V4<T>*, \
                                       const V4<T>*, int, double*, hipStream_t);                           \
  template void build_m2l_dense<T>(int, const RbfDesc&, double, M2LOperator&, DevBuf<T>&,          \
                                   hipStream_t);                                                    \
  template void launch_m2l_batch<T, double>(const M2LOperator&,                                        \
                                            const std::vector<M2LLevelSpec<T, double>>&, const int32_t*, \
                                            bool, int, int, hipStream_t);                                 \
  template void launch_m2l_dense<T>(const M2LOperator&, const std::vector<M2LDenseLevel<T>>&,       \
                                    const int32_t*, M2LSchedule&, DevBuf<double>&, hipStream_t);

PAM_INSTANTIATE(float)
PAM_INSTANTIATE(double)

}  // namespace pam
