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
                                     int rpad, int kpad, T* __restrict__ dst) {
  int64_t per = static_cast<int64_t>(rpad) * kpad;
  int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= batch * per) return;
  int b = static_cast<int>(idx / per);
  int64_t rem = idx % per;
  int kb = static_cast<int>(rem / (static_cast<int64_t>(rpad) * 16));
  int row = static_cast<int>((rem / 16) % rpad);
  int k = kb * 16 + static_cast<int>(rem % 16);
  double v = (row < rows && k < cols) ? src[(static_cast<int64_t>(b) * rows + row) * cols + k] : 0.0;
  dst[idx] = static_cast<T>(v);
}

template <class T>
void pack(const double* src, int batch, int rows, int cols, M2LOperator& op, DevBuf<T>& dst, hipStream_t s) {
  op.r = rows;
  op.rpad = rpad_of(rows);
  op.kpad = (cols + 15) / 16 * 16;
  op.per_v = static_cast<size_t>(op.rpad) * op.kpad;
  dst.resize(static_cast<size_t>(batch) * op.per_v);
  int64_t total = static_cast<int64_t>(batch) * op.per_v;
  pack_fragment_kernel<T><<<ceil_div(total, 256), 256, 0, s>>>(src, batch, rows, cols, op.rpad, op.kpad, dst.get());
  PAM_LAUNCH_CHECK();
}

}  // namespace

void jacobi_eigen(std::vector<double>& a, int n, std::vector<double>& evals, std::vector<double>& vecs) {
  // Pad to an even size with a zero row/column so that every round pairs all indices.
  int m = n + (n & 1);
  std::vector<double> A(static_cast<size_t>(m) * m, 0.0), V(static_cast<size_t>(m) * m, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A[static_cast<size_t>(i) * m + j] = a[static_cast<size_t>(i) * n + j];
  for (int i = 0; i < m; i++) V[static_cast<size_t>(i) * m + i] = 1.0;
  std::vector<int> pos(m);
  for (int i = 0; i < m; i++) pos[i] = i;
  std::vector<double> cs(m / 2), sn(m / 2);
  std::vector<int> pp(m / 2), qq(m / 2);
  double total = 0.0;
  for (double x : A) total += x * x;
  for (int sweep = 0; sweep < 40; sweep++) {
    double off = 0.0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++)
        if (i != j) off += A[static_cast<size_t>(i) * m + j] * A[static_cast<size_t>(i) * m + j];
    if (off <= 1e-30 * total) break;
    for (int round = 0; round < m - 1; round++) {
      for (int k = 0; k < m / 2; k++) {
        int p = std::min(pos[k], pos[m - 1 - k]), q = std::max(pos[k], pos[m - 1 - k]);
        pp[k] = p;
        qq[k] = q;
        double apq = A[static_cast<size_t>(p) * m + q];
        if (std::abs(apq) < 1e-300) {
          cs[k] = 1.0;
          sn[k] = 0.0;
          continue;
        }
        double tau = (A[static_cast<size_t>(q) * m + q] - A[static_cast<size_t>(p) * m + p]) / (2.0 * apq);
        double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        cs[k] = 1.0 / std::sqrt(1.0 + t * t);
        sn[k] = t * cs[k];
      }
      // columns of A and V (all rotations of the round touch disjoint columns)
#pragma omp parallel for schedule(static)
      for (int row = 0; row < m; row++) {
        double* ar = &A[static_cast<size_t>(row) * m];
        double* vr = &V[static_cast<size_t>(row) * m];
        for (int k = 0; k < m / 2; k++) {
          double c = cs[k], s = sn[k];
          if (s == 0.0) continue;
          int p = pp[k], q = qq[k];
          double x = ar[p], y = ar[q];
          ar[p] = c * x - s * y;
          ar[q] = s * x + c * y;
          x = vr[p];
          y = vr[q];
          vr[p] = c * x - s * y;
          vr[q] = s * x + c * y;
        }
      }
      // rows of A
#pragma omp parallel for schedule(static)
      for (int k = 0; k < m / 2; k++) {
        double c = cs[k], s = sn[k];
        if (s == 0.0) continue;
        double* ap = &A[static_cast<size_t>(pp[k]) * m];
        double* aq = &A[static_cast<size_t>(qq[k]) * m];
        for (int col = 0; col < m; col++) {
          double x = ap[col], y = aq[col];
          ap[col] = c * x - s * y;
          aq[col] = s * x + c * y;
        }
      }
      int last = pos[m - 1];
      for (int i = m - 1; i > 1; i--) pos[i] = pos[i - 1];
      pos[1] = last;
    }
  }
  std::vector<int> order(m);
  for (int i = 0; i < m; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    return A[static_cast<size_t>(x) * m + x] > A[static_cast<size_t>(y) * m + y];
  });
  evals.assign(n, 0.0);
  vecs.assign(static_cast<size_t>(n) * n, 0.0);
  int k = 0;
  for (int oi = 0; oi < m && k < n; oi++) {
    int c = order[oi];
    // skip the padding direction (a unit vector on the added index)
    if (m != n && std::abs(V[static_cast<size_t>(n) * m + c]) > 0.5) continue;
    evals[k] = A[static_cast<size_t>(c) * m + c];
    for (int i = 0; i < n; i++) vecs[static_cast<size_t>(i) * n + k] = V[static_cast<size_t>(i) * m + c];
    k++;
  }
}

template <class T>
void build_m2l_compressed(int p, const RbfDesc& rbf, double h, double rel_tol, CompressedM2L<T>& out,
                          hipStream_t s) {
  const int p3 = p * p * p;
  out.p = p;
  PhiConsts pc = phi_consts(rbf, h);
  int64_t per_v = static_cast<int64_t>(p3) * p3;
  DevBuf<double> kv(static_cast<size_t>(343 * per_v));
  dense_kv_kernel<<<ceil_div(343 * per_v, 256), 256, 0, s>>>(p, rbf.kind, pc.a, pc.b, kv.get());
  PAM_LAUNCH_CHECK();
  // G = sum_v K_v K_v^T
  DevBuf<double> g(static_cast<size_t>(p3) * p3);
  GemmArgs ga{p3, p3, p3, 343, 0, per_v, p3, 1, 0, per_v, 1, p3, 0, p3, 1};
  gemm(ga, 1, kv.get(), kv.get(), g.get(), s);
  std::vector<double> hg(static_cast<size_t>(p3) * p3);
  PAM_HIP(hipMemcpyAsync(hg.data(), g.get(), hg.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PAM_HIP(hipStreamSynchronize(s));
  std::vector<double> evals, vecs;
  jacobi_eigen(hg, p3, evals, vecs);
  out.sv.resize(p3);
  for (int i = 0; i < p3; i++) out.sv[i] = std::sqrt(std::max(evals[i], 0.0));
  int r = p3;
  for (int k = 16; k < p3; k += 16) {
    if (out.sv[k] <= rel_tol * out.sv[0]) {
      r = k;
      break;
    }
  }
  r = std::min(r, p3);
  out.r = r;
  out.kpad_r = (r + 15) / 16 * 16;
  std::vector<double> hu(static_cast<size_t>(p3) * r), hut(static_cast<size_t>(r) * p3);
  for (int i = 0; i < p3; i++)
    for (int k = 0; k < r; k++) {
      hu[static_cast<size_t>(i) * r + k] = vecs[static_cast<size_t>(i) * p3 + k];
      hut[static_cast<size_t>(k) * p3 + i] = vecs[static_cast<size_t>(i) * p3 + k];
    }
  DevBuf<double> du, dut, tv(static_cast<size_t>(343) * p3 * r), cv(static_cast<size_t>(343) * r * r);
  du.upload(hu.data(), hu.size(), s);
  dut.upload(hut.data(), hut.size(), s);
  // T_v = K_v U (p3 x r), C_v = U^T T_v (r x r)
  GemmArgs g1{p3, r, p3, 1, per_v, 0, p3, 1, 0, 0, r, 1, static_cast<int64_t>(p3) * r, r, 1};
  gemm(g1, 343, kv.get(), du.get(), tv.get(), s);
  GemmArgs g2{r, r, p3, 1, 0, 0, p3, 1, static_cast<int64_t>(p3) * r, 0, r, 1, static_cast<int64_t>(r) * r, r, 1};
  gemm(g2, 343, dut.get(), tv.get(), cv.get(), s);
  pack<T>(cv.get(), 343, r, r, out.c_op, out.c, s);
  pack<T>(dut.get(), 1, r, p3, out.ut_op, out.ut, s);
  pack<T>(du.get(), 1, p3, r, out.u_op, out.u, s);
  PAM_HIP(hipStreamSynchronize(s));  // temporaries go out of scope
}

template void build_m2l_compressed<float>(int, const RbfDesc&, double, double, CompressedM2L<float>&, hipStream_t);
template void build_m2l_compressed<double>(int, const RbfDesc&, double, double, CompressedM2L<double>&, hipStream_t);

}  // namespace pam
