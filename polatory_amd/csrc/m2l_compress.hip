// Setup of the SVD-compressed M2L operators (see m2l_compress.hpp). Runs once
// per (order, kernel, cell size) and is cached by the evaluator.
#include "m2l_compress.hpp"

#include <algorithm>
#include <cmath>

namespace pam {

namespace {

// K[v][i][j] = phi(x_i - 2 d_v - y_j) for the 316 well-separated v (0 otherwise).
__global__ void dense_kv_kernel(int p, int kind, double pa, double pb, double* __restrict__ k) {
  const int p3 = p * p * p;
  int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t per_v = static_cast<int64_t>(p3) * p3;
  if (idx >= 343 * per_v) return;
  int v = static_cast<int>(idx / per_v);
  int64_t rem = idx % per_v;
  int row = static_cast<int>(rem / p3), col = static_cast<int>(rem % p3);
  int dx = v / 49 - 3, dy = (v / 7) % 7 - 3, dz = v % 7 - 3;
  bool adj = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
  double val = 0.0;
  if (!adj) {
    int ta = row / (p * p), tb = (row / p) % p, tc = row % p;
    int sa = col / (p * p), sb = (col / p) % p, sc = col % p;
    double ex = cospi((2.0 * ta + 1) / (2.0 * p)) - 2.0 * dx - cospi((2.0 * sa + 1) / (2.0 * p));
    double ey = cospi((2.0 * tb + 1) / (2.0 * p)) - 2.0 * dy - cospi((2.0 * sb + 1) / (2.0 * p));
    double ez = cospi((2.0 * tc + 1) / (2.0 * p)) - 2.0 * dz - cospi((2.0 * sc + 1) / (2.0 * p));
    double r2 = ex * ex + ey * ey + ez * ez;
    if (kind == kBh3) val = pa * sqrt(r2 + pb);
    else if (kind == kTh3) { double rho = sqrt(r2 + pb); val = pa * rho * rho * rho; }
    else val = pa * exp(pb * sqrt(r2));
  }
  k[idx] = val;
}

// Batched strided FP64 GEMM, 16x16 tiles: C[b](i,j) = sum_s sum_k A(b,s,i,k) B(b,s,k,j)
struct GemmArgs {
  int m, n, k, nsum;
  int64_t a_b, a_s, a_i, a_k;
  int64_t b_b, b_s, b_k, b_j;
  int64_t c_b, c_i, c_j;
};

__global__ void gemm_f64_kernel(GemmArgs g, const double* __restrict__ a, const double* __restrict__ b,
                                double* __restrict__ c) {
  __shared__ double as[16][17];
  __shared__ double bs[16][17];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int i = blockIdx.y * 16 + ty, j = blockIdx.x * 16 + tx;
  const int bt = blockIdx.z;
  double acc = 0.0;
  for (int s = 0; s < g.nsum; s++) {
    const double* ab = a + bt * g.a_b + s * g.a_s;
    const double* bb = b + bt * g.b_b + s * g.b_s;
    for (int k0 = 0; k0 < g.k; k0 += 16) {
      int ai = blockIdx.y * 16 + ty, ak = k0 + tx;
      as[ty][tx] = (ai < g.m && ak < g.k) ? ab[ai * g.a_i + ak * g.a_k] : 0.0;
      int bk = k0 + ty, bj = blockIdx.x * 16 + tx;
      bs[ty][tx] = (bk < g.k && bj < g.n) ? bb[bk * g.b_k + bj * g.b_j] : 0.0;
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 16; t++) acc += as[ty][t] * bs[t][tx];
      __syncthreads();
    }
  }
  if (i < g.m && j < g.n) c[bt * g.c_b + i * g.c_i + j * g.c_j] = acc;
}

void gemm(const GemmArgs& g, int batch, const double* a, const double* b, double* c, hipStream_t s) {
  dim3 grid((g.n + 15) / 16, (g.m + 15) / 16, batch);
  gemm_f64_kernel<<<grid, 256, 0, s>>>(g, a, b, c);
  PAM_LAUNCH_CHECK();
}

// Packs row-major [batch][rows][cols] FP64 into the MFMA fragment layout of
// launch_m2l: ((b * (kpad/16) + k/16) * rpad + row) * 16 + k % 16, zero padded.
template <class T>
__global__ void pack_fragment_kernel(const double* __restrict__ src, int batch, int rows, int cols,
                                     int rpad, int kpad, int64_t slack, T* __restrict__ dst) {
  int64_t per = static_cast<int64_t>(rpad) * kpad;
  int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= batch * per) {
    if (idx < batch * per + slack) dst[idx] = T(0);
    return;
  }
  int b = static_cast<int>(idx / per);
  int64_t rem = idx % per;
  int kb = static_cast<int>(rem / (static_cast<int64_t>(rpad) * 16));
  int row = static_cast<int>((rem / 16) % rpad);
  int k = kb * 16 + static_cast<int>(rem % 16);
  double v = (row < rows && k < cols) ? src[(static_cast<int64_t>(b) * rows + row) * cols + k] : 0.0;
  dst[idx] = static_cast<T>(v);
}

// row_pad: row padding of the layout (128 for launch_m2l; 64 for the pair-major
// transfer, whose workgroups read 128 rows: `slack` zero rows past the last
// batch keep the reads of a final half block in bounds).
template <class T>
void pack(const double* src, int batch, int rows, int cols, M2LOperator& op, DevBuf<T>& dst, hipStream_t s,
          int row_pad = 128, int64_t slack = 0) {
  op.r = rows;
  op.rpad = (rows + row_pad - 1) / row_pad * row_pad;
  op.kpad = (cols + 15) / 16 * 16;
  op.per_v = static_cast<size_t>(op.rpad) * op.kpad;
  int64_t total = static_cast<int64_t>(batch) * op.per_v;
  dst.resize(static_cast<size_t>(total + slack));
  pack_fragment_kernel<T><<<ceil_div(total + slack, 256), 256, 0, s>>>(src, batch, rows, cols, op.rpad, op.kpad,
                                                                       slack, dst.get());
  PAM_LAUNCH_CHECK();
}

// Standard normal samples from a counter hash (splitmix64 finaliser, Box-Muller),
// reproducible for a given seed.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

__global__ void gaussian_kernel(uint64_t seed, int64_t n, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = mix64(seed + 2 * static_cast<uint64_t>(i)), b = mix64(seed + 2 * static_cast<uint64_t>(i) + 1);
  const double u1 = (static_cast<double>(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  out[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

void gaussian(uint64_t seed, int64_t n, double* out, hipStream_t s) {
  gaussian_kernel<<<ceil_div(n, 256), 256, 0, s>>>(seed, n, out);
  PAM_LAUNCH_CHECK();
}

// Orthonormal columns of the row-major n x k matrix y, in place: modified
// Gram-Schmidt applied twice per column ("twice is enough", Giraud et al. 2005).
void orthonormalize_columns(std::vector<double>& y, int n, int k) {
  std::vector<double> col(n);
  for (int j = 0; j < k; j++) {
    for (int i = 0; i < n; i++) col[i] = y[static_cast<size_t>(i) * k + j];
    for (int pass = 0; pass < 2; pass++)
      for (int q = 0; q < j; q++) {
        double d = 0.0;
        for (int i = 0; i < n; i++) d += y[static_cast<size_t>(i) * k + q] * col[i];
        for (int i = 0; i < n; i++) col[i] -= d * y[static_cast<size_t>(i) * k + q];
      }
    double nrm = 0.0;
    for (int i = 0; i < n; i++) nrm += col[i] * col[i];
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0)) throw_runtime("M2L compression: sample matrix lost rank");
    for (int i = 0; i < n; i++) y[static_cast<size_t>(i) * k + j] = col[i] / nrm;
  }
}

}  // namespace

template <class T>
void build_m2l_compressed(int p, const RbfDesc& rbf, double h, double rel_tol, CompressedM2L<T>& out,
                          hipStream_t s) {
  const int p3 = p * p * p;
  out.p = p;
  const int kmax = std::min(kMaxM2LRank, p3 / 64 * 64);
  if (kmax < 64) throw_invalid("M2L compression needs an order of at least 4");
  PhiConsts pc = phi_consts(rbf, h);
  const int64_t per_v = static_cast<int64_t>(p3) * p3;
  const int64_t n_wide = 343 * static_cast<int64_t>(p3);  // columns of the stacked [K_v]_v
  DevBuf<double> kv(static_cast<size_t>(343 * per_v));
  dense_kv_kernel<<<ceil_div(343 * per_v, 256), 256, 0, s>>>(p, rbf.kind, pc.a, pc.b, kv.get());
  PAM_LAUNCH_CHECK();
  // Range of K = [K_v]_v by randomized subspace iteration (Halko, Martinsson
  // and Tropp 2011, algorithms 4.1 and 4.4): Q1 = orth(K W), Q = orth(K K^T Q1)
  // with W an n_wide x kmax Gaussian matrix. Unlike the eigenvectors of
  // sum_v K_v K_v^T, the basis resolves directions down to rounding of K itself.
  DevBuf<double> w(static_cast<size_t>(n_wide) * kmax), y(static_cast<size_t>(p3) * kmax),
      z(static_cast<size_t>(n_wide) * kmax);
  gaussian(0x5eed0001ULL, n_wide * kmax, w.get(), s);
  // Y = sum_v K_v W_v (W_v: rows v p3 .. v p3 + p3 of W)
  const GemmArgs stack{p3, kmax, p3, 343, 0, per_v, p3, 1, 0, static_cast<int64_t>(p3) * kmax, kmax, 1, 0, kmax, 1};
  gemm(stack, 1, kv.get(), w.get(), y.get(), s);
  std::vector<double> hq(static_cast<size_t>(p3) * kmax);
  PAM_HIP(hipMemcpyAsync(hq.data(), y.get(), hq.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  orthonormalize_columns(hq, p3, kmax);
  DevBuf<double> dq;
  dq.upload(hq.data(), hq.size(), s);
  // Z_v = K_v^T Q1, then Y = sum_v K_v Z_v
  const GemmArgs kt{p3, kmax, p3, 1, per_v, 0, 1, p3, 0, 0, kmax, 1, static_cast<int64_t>(p3) * kmax, kmax, 1};
  gemm(kt, 343, kv.get(), dq.get(), z.get(), s);
  gemm(stack, 1, kv.get(), z.get(), y.get(), s);
  PAM_HIP(hipMemcpyAsync(hq.data(), y.get(), hq.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  orthonormalize_columns(hq, p3, kmax);
  // Rank: the first multiple of 64 whose leading columns Q_r leave
  // max |E - Q_r Q_r^T E| <= rel_tol max |E| for a fresh sample E = K W'
  // of 8 columns (the a-posteriori estimate of Halko et al., section 4.3).
  constexpr int kProbe = 8;
  gaussian(0x5eed0002ULL, n_wide * kProbe, w.get(), s);
  const GemmArgs probe{p3, kProbe, p3, 343, 0, per_v, p3, 1, 0, static_cast<int64_t>(p3) * kProbe, kProbe, 1, 0,
                       kProbe, 1};
  gemm(probe, 1, kv.get(), w.get(), y.get(), s);
  std::vector<double> he(static_cast<size_t>(p3) * kProbe);
  PAM_HIP(hipMemcpyAsync(he.data(), y.get(), he.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  double emax = 0.0;
  for (double e : he) emax = std::max(emax, std::abs(e));
  out.sv.clear();
  int r = kmax;
  for (int rr = 64; rr <= kmax; rr += 64) {
    double res = 0.0;
    for (int c = 0; c < kProbe; c++) {
      std::vector<double> proj(rr, 0.0);
      for (int i = 0; i < p3; i++)
        for (int q = 0; q < rr; q++) proj[q] += hq[static_cast<size_t>(i) * kmax + q] * he[static_cast<size_t>(i) * kProbe + c];
      for (int i = 0; i < p3; i++) {
        double v = he[static_cast<size_t>(i) * kProbe + c];
        for (int q = 0; q < rr; q++) v -= hq[static_cast<size_t>(i) * kmax + q] * proj[q];
        res = std::max(res, std::abs(v));
      }
    }
    out.sv.push_back(emax > 0.0 ? res / emax : 0.0);
    if (emax == 0.0 || res <= rel_tol * emax) {
      r = rr;
      break;
    }
  }
  out.r = r;
  out.kpad_r = r;
  std::vector<double> hu(static_cast<size_t>(p3) * r), hut(static_cast<size_t>(r) * p3);
  for (int i = 0; i < p3; i++)
    for (int k = 0; k < r; k++) {
      hu[static_cast<size_t>(i) * r + k] = hq[static_cast<size_t>(i) * kmax + k];
      hut[static_cast<size_t>(k) * p3 + i] = hq[static_cast<size_t>(i) * kmax + k];
    }
  DevBuf<double> du, dut, tv(static_cast<size_t>(343) * p3 * r), cv(static_cast<size_t>(343) * r * r);
  du.upload(hu.data(), hu.size(), s);
  dut.upload(hut.data(), hut.size(), s);
  // T_v = K_v U (p3 x r), C_v = U^T T_v (r x r)
  GemmArgs g1{p3, r, p3, 1, per_v, 0, p3, 1, 0, 0, r, 1, static_cast<int64_t>(p3) * r, r, 1};
  gemm(g1, 343, kv.get(), du.get(), tv.get(), s);
  GemmArgs g2{r, r, p3, 1, 0, 0, p3, 1, static_cast<int64_t>(p3) * r, 0, r, 1, static_cast<int64_t>(r) * r, r, 1};
  gemm(g2, 343, dut.get(), tv.get(), cv.get(), s);
  // C_v in the pair-major layout (rows padded to 64, 64 zero rows of slack x 16).
  pack<T>(cv.get(), 343, r, r, out.c_op, out.c, s, 64, 64 * 16);
  pack<double>(dut.get(), 1, r, p3, out.ut_op, out.ut, s);
  pack<double>(du.get(), 1, p3, r, out.u_op, out.u, s);
  PAM_HIP(hipStreamSynchronize(s));  // temporaries go out of scope
}

template void build_m2l_compressed<float>(int, const RbfDesc&, double, double, CompressedM2L<float>&, hipStream_t);
template void build_m2l_compressed<double>(int, const RbfDesc&, double, double, CompressedM2L<double>&, hipStream_t);

}  // namespace pam
