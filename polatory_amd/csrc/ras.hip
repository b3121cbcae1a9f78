// Multilevel RAS preconditioner: host domain division, batched device assembly,
// batched FP64 Cholesky of Q^T A Q and batched subdomain solves.
#include "ras.hpp"

#include <future>
#include <mutex>

#include <algorithm>
#include <parallel/algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <list>
#include <numeric>
#include <queue>

#include "device_api.hpp"
#include "rbf_device.hpp"

namespace pam {

namespace {

constexpr int kT = 256;
constexpr int kNb = 32;     // Cholesky panel width
constexpr int kTile = 64;   // trailing-update tile

struct RbfSet {
  int n;
  int kind[4];
  double p0[4], p1[4];
  double a[4][9];
  double nugget;
};

RbfSet make_rbf_set(const Model& m) {
  if (m.rbfs.size() > 4) throw_invalid("at most 4 rbfs per model are supported");
  RbfSet r{};
  r.n = static_cast<int>(m.rbfs.size());
  for (int k = 0; k < r.n; k++) {
    r.kind[k] = m.rbfs[k].kind;
    r.p0[k] = m.rbfs[k].p0;
    r.p1[k] = m.rbfs[k].p1;
    for (int i = 0; i < 9; i++) r.a[k][i] = m.rbfs[k].aniso[i];
  }
  r.nugget = m.nugget;
  return r;
}

// Sum of the rbfs at the separation (dx, dy, dz) (preconditioner/mat_a.hpp:11-61).
__device__ __forceinline__ double kernel_value(const RbfSet& r, double dx, double dy, double dz) {
  double s = 0.0;
  for (int k = 0; k < r.n; k++) {
    const double* a = r.a[k];
    double tx = a[0] * dx + a[1] * dy + a[2] * dz;
    double ty = a[3] * dx + a[4] * dy + a[5] * dz;
    double tz = a[6] * dx + a[7] * dy + a[8] * dz;
    double r2 = tx * tx + ty * ty + tz * tz;
    if (r.kind[k] == kBh3) {
      s += -r.p0[k] * sqrt(r2 + r.p1[k] * r.p1[k]);
    } else if (r.kind[k] == kTh3) {
      double rho = sqrt(r2 + r.p1[k] * r.p1[k]);
      s += r.p0[k] * rho * rho * rho;
    } else {
      s += r.p0[k] * exp(-3.0 * sqrt(r2) / r.p1[k]);
    }
  }
  return s;
}

// mat_a entry (preconditioner/mat_a.hpp:11-61): sum of rbfs + nugget on the diagonal.
__device__ __forceinline__ double kernel_entry(const RbfSet& r, const double* __restrict__ pts, int u, int v) {
  double s = kernel_value(r, pts[3 * u] - pts[3 * v], pts[3 * u + 1] - pts[3 * v + 1], pts[3 * u + 2] - pts[3 * v + 2]);
  if (u == v) s += r.nugget;
  return s;
}

// Entry of mat_a between two rows of a fit with gradient data
// (preconditioner/mat_a.hpp:25-56), summed over the rbfs; ti, tj = -1 for a
// value row, else the gradient component; d = position_i - position_j:
//   value, value:       phi(d)
//   value, gradient e:  -d phi / d x_e (d)
//   gradient c, value:  +d phi / d x_c (d)
//   gradient c, gradient e: -d^2 phi / d x_c d x_e (d)
// With u = A d and the isotropic g = f u, H = f (I + t u u^T) (rbf_grad_factor),
// grad phi = f A^T u and Hess phi = f (A^T A + t (A^T u)(A^T u)^T).
__device__ double kernel_hb(const RbfSet& r, int ti, int tj, double dx, double dy, double dz) {
  if (ti < 0 && tj < 0) return kernel_value(r, dx, dy, dz);
  double s = 0.0;
  for (int k = 0; k < r.n; k++) {
    const double* a = r.a[k];
    const double u[3] = {a[0] * dx + a[1] * dy + a[2] * dz, a[3] * dx + a[4] * dy + a[5] * dz,
                         a[6] * dx + a[7] * dy + a[8] * dz};
    double t = 0.0, f = 0.0;
    if (r.kind[k] == kBh3) f = rbf_grad_factor<kBh3>(r.p0[k], r.p1[k], u, &t);
    else if (r.kind[k] == kTh3) f = rbf_grad_factor<kTh3>(r.p0[k], r.p1[k], u, &t);
    else f = rbf_grad_factor<kExp>(r.p0[k], r.p1[k], u, &t);
    double v[3];
    for (int c = 0; c < 3; c++) v[c] = a[c] * u[0] + a[3 + c] * u[1] + a[6 + c] * u[2];
    if (ti < 0) {
      s -= f * v[tj];
    } else if (tj < 0) {
      s += f * v[ti];
    } else {
      const double g = a[ti] * a[tj] + a[3 + ti] * a[3 + tj] + a[6 + ti] * a[6 + tj];
      s -= f * (g + t * v[ti] * v[tj]);
    }
  }
  return s;
}

// Position and type (-1 value, else gradient component) of flat row f.
__device__ __forceinline__ const double* flat_row(const double* __restrict__ pts, const double* __restrict__ gpts,
                                                  int64_t mu, int64_t f, int* type) {
  if (f < mu) {
    *type = -1;
    return pts + 3 * f;
  }
  const int64_t k = (f - mu) / 3;
  *type = static_cast<int>(f - mu - 3 * k);
  return gpts + 3 * k;
}

// mat_a entry between flat rows u and v (nugget on the diagonal of value rows).
__device__ __forceinline__ double flat_entry(const RbfSet& r, const double* __restrict__ pts,
                                             const double* __restrict__ gpts, int64_t mu, int64_t u, int64_t v) {
  if (u < mu && v < mu) return kernel_entry(r, pts, static_cast<int>(u), static_cast<int>(v));
  int tu, tv;
  const double* xu = flat_row(pts, gpts, mu, u, &tu);
  const double* xv = flat_row(pts, gpts, mu, v, &tv);
  return kernel_hb(r, tu, tv, xu[0] - xv[0], xu[1] - xv[1], xu[2] - xv[2]);
}

__device__ __forceinline__ void monomials(int degree, double x, double y, double z, double* m) {
  m[0] = 1.0;
  if (degree >= 1) { m[1] = x; m[2] = y; m[3] = z; }
  if (degree >= 2) { m[4] = x * x; m[5] = x * y; m[6] = x * z; m[7] = y * y; m[8] = y * z; m[9] = z * z; }
}

// lag[i*l + k] = sum_j m_j(x_i) pinv[j*l + k]   (LagrangeBasis::evaluate, lagrange_basis.hpp:50-58)
__global__ void lagrange_kernel(int degree, int l, const double* __restrict__ pts, int64_t n,
                                const double* __restrict__ pinv, double* __restrict__ lag) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m[10];
  monomials(degree, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], m);
  for (int k = 0; k < l; k++) {
    double s = 0.0;
    for (int j = 0; j < l; j++) s += m[j] * pinv[j * l + k];
    lag[i * l + k] = s;
  }
}

// Gradient rows of the Lagrange basis: lag[(3 k + c) * l + a] =
// sum_j d m_j / d x_c (q_k) pinv[j * l + a] (LagrangeBasis::evaluate with
// gradient points, lagrange_basis.hpp:50-58); lag points at the first
// gradient row.
__global__ void lagrange_grad_kernel(int degree, int l, const double* __restrict__ gpts, int64_t n,
                                     const double* __restrict__ pinv, double* __restrict__ lag) {
  int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double x = gpts[3 * k], y = gpts[3 * k + 1], z = gpts[3 * k + 2];
  const double gm[3][10] = {{0.0, 1.0, 0.0, 0.0, 2.0 * x, y, z, 0.0, 0.0, 0.0},
                            {0.0, 0.0, 1.0, 0.0, 0.0, x, 0.0, 2.0 * y, z, 0.0},
                            {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, x, 0.0, y, 2.0 * z}};
  (void)degree;
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < l; a++) {
      double s = 0.0;
      for (int j = 0; j < l; j++) s += gm[c][j] * pinv[j * l + a];
      lag[(3 * k + c) * l + a] = s;
    }
}

// Fields of the inner rows of every domain of a fit with gradient data, for
// M = Q^T A Q as in assemble_rows_kernel: x, y, z, type, q_0.., g_0.. with
// g_i = U[:, i] + Atl q_i / 2, U[a][i] = A(top_a, row_i), Atl = A(top, top)
// (the first l rows of a domain are the polynomial rows). Field-major:
// rows[(4 + 2l) o_d + f n_d + i].
__global__ void assemble_hb_rows_kernel(RbfSet r, int l, const double* __restrict__ pts,
                                        const double* __restrict__ gpts, int64_t mu,
                                        const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                        const double* __restrict__ lag, double* __restrict__ rows) {
  __shared__ double atl[10][10];
  const int d = blockIdx.y;
  const int64_t p0 = ptr[d];
  const int n = static_cast<int>(ptr[d + 1] - p0) - l;
  const int64_t o = p0 - static_cast<int64_t>(d) * l;
  for (int e = threadIdx.x; e < l * l; e += kT)
    atl[e / l][e % l] = flat_entry(r, pts, gpts, mu, idx[p0 + e / l], idx[p0 + e % l]);
  __syncthreads();
  double* rd = rows + static_cast<int64_t>(4 + 2 * l) * o;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const int64_t f = idx[p0 + l + i];
    int type;
    const double* x = flat_row(pts, gpts, mu, f, &type);
    rd[i] = x[0];
    rd[n + i] = x[1];
    rd[2 * n + i] = x[2];
    rd[3 * n + i] = static_cast<double>(type);
    double q[10];
    for (int a = 0; a < l; a++) {
      q[a] = -lag[f * l + a];
      rd[static_cast<int64_t>(4 + a) * n + i] = q[a];
    }
    for (int a = 0; a < l; a++) {
      double t = 0.0;
      for (int b = 0; b < l; b++) t += atl[a][b] * q[b];
      rd[static_cast<int64_t>(4 + l + a) * n + i] = flat_entry(r, pts, gpts, mu, idx[p0 + a], f) + 0.5 * t;
    }
  }
}

// U[a][j] = A(p_a, q_j) (l x n_d), Atl[a][b] = A(p_a, p_b)
__global__ void assemble_u_kernel(RbfSet r, int l, const double* __restrict__ pts,
                                  const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                  double* __restrict__ u, double* __restrict__ atl) {
  int d = blockIdx.y;
  int64_t p0 = ptr[d];
  int m = static_cast<int>(ptr[d + 1] - p0);
  int n = m - l;
  int64_t uoff = static_cast<int64_t>(l) * (p0 - static_cast<int64_t>(d) * l);
  int total = l * n + l * l;
  for (int e = blockIdx.x * kT + threadIdx.x; e < total; e += gridDim.x * kT) {
    if (e < l * n) {
      int a = e / n, j = e % n;
      u[uoff + e] = kernel_entry(r, pts, idx[p0 + a], idx[p0 + l + j]);
    } else {
      int f = e - l * n;
      int a = f / l, b = f % l;
      atl[static_cast<int64_t>(d) * l * l + f] = kernel_entry(r, pts, idx[p0 + a], idx[p0 + b]);
    }
  }
}

// Rows of the inner points of every domain for the assembly of
// M = Q^T A Q (fine_grid.hpp:79-86), Q = [-Lag_inner^T ; I]. Writing
// q_i = -lag(x_i) and g_i = U[:, i] + Atl q_i / 2 (U, Atl from
// assemble_u_kernel, Atl symmetric),
//   M_ij = A(x_i, x_j) + q_i . U[:, j] + U[:, i] . q_j + q_i^T Atl q_j
//        = A(x_i, x_j) + q_i . g_j + g_i . q_j.
// Per domain d (inner offset o_d = ptr[d] - d l, n_d inner points) the
// 3 + 2l fields are stored field-major: rows[(3 + 2l) o_d + f n_d + i] with
// f = x, y, z, q_0..q_{l-1}, g_0..g_{l-1}.
__global__ void assemble_rows_kernel(int l, const double* __restrict__ pts, const int64_t* __restrict__ ptr,
                                     const int32_t* __restrict__ idx, const double* __restrict__ lag,
                                     const double* __restrict__ u, const double* __restrict__ atl,
                                     double* __restrict__ rows) {
  const int d = blockIdx.y;
  const int64_t p0 = ptr[d];
  const int n = static_cast<int>(ptr[d + 1] - p0) - l;
  const int64_t o = p0 - static_cast<int64_t>(d) * l;
  double* rd = rows + static_cast<int64_t>(3 + 2 * l) * o;
  const double* ud = u + static_cast<int64_t>(l) * o;
  const double* ad = atl + static_cast<int64_t>(d) * l * l;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const int qi = idx[p0 + l + i];
    rd[i] = pts[3 * qi];
    rd[n + i] = pts[3 * qi + 1];
    rd[2 * n + i] = pts[3 * qi + 2];
    double q[10];
    for (int a = 0; a < l; a++) {
      q[a] = -lag[static_cast<int64_t>(qi) * l + a];
      rd[static_cast<int64_t>(3 + a) * n + i] = q[a];
    }
    for (int a = 0; a < l; a++) {
      double t = 0.0;
      for (int b = 0; b < l; b++) t += ad[a * l + b] * q[b];
      rd[static_cast<int64_t>(3 + l + a) * n + i] = ud[static_cast<int64_t>(a) * n + i] + 0.5 * t;
    }
  }
}

// Lower triangle of M (column-major, n_d x n_d per domain) in 64 x 64 tiles:
// workgroup tile t of domain d covers rows 64 I.., columns 64 J.., I >= J.
// The tile's column fields are staged in LDS; each thread keeps one row's
// fields in registers and computes 16 columns, so the stores of a wavefront
// go to 64 consecutive rows of one column. Entries are computed in FP64 and
// stored in T (FP32 for single-precision factors). The strict upper triangle
// is never written: the factorisation and the solves read only the lower one.
template <class T, int L>
__global__ void __launch_bounds__(kT)
assemble_m_kernel(RbfSet r, const int64_t* __restrict__ ptr, const double* __restrict__ rows,
                  const int64_t* __restrict__ mat_ptr, T* __restrict__ mat) {
  constexpr int F = 3 + 2 * L;
  __shared__ double cj[F][kTile];
  const int d = blockIdx.y;
  const int64_t p0 = ptr[d];
  const int n = static_cast<int>(ptr[d + 1] - p0) - L;
  if (n <= 0) return;
  const int nt = (n + kTile - 1) / kTile;
  const int ntiles = nt * (nt + 1) / 2;
  const double* rd = rows + static_cast<int64_t>(F) * (p0 - static_cast<int64_t>(d) * L);
  T* md = mat + mat_ptr[d];
  const int tid = threadIdx.x, li = tid & 63, cg = tid >> 6;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int I = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    while (I * (I + 1) / 2 > t) I--;
    const int J = t - I * (I + 1) / 2;
    __syncthreads();
    for (int e = tid; e < F * kTile; e += kT) {
      const int f = e / kTile, c = e % kTile, j = J * kTile + c;
      cj[f][c] = j < n ? rd[static_cast<int64_t>(f) * n + j] : 0.0;
    }
    const int i = I * kTile + li;
    double ri[F];
#pragma unroll
    for (int f = 0; f < F; f++) ri[f] = i < n ? rd[static_cast<int64_t>(f) * n + i] : 0.0;
    __syncthreads();
    if (i >= n) continue;
#pragma unroll 4
    for (int c = 0; c < 16; c++) {
      const int jc = cg * 16 + c, j = J * kTile + jc;
      if (j > i || j >= n) break;
      double v = kernel_value(r, ri[0] - cj[0][jc], ri[1] - cj[1][jc], ri[2] - cj[2][jc]);
      if (j == i) v += r.nugget;
#pragma unroll
      for (int a = 0; a < L; a++) v += ri[3 + a] * cj[3 + L + a][jc] + ri[3 + L + a] * cj[3 + a][jc];
      md[static_cast<int64_t>(j) * n + i] = static_cast<T>(v);
    }
  }
}

// assemble_m_kernel for fits with gradient data: the rows carry their type
// (field 3) and entries come from kernel_hb.
template <class T, int L>
__global__ void __launch_bounds__(kT)
assemble_m_hb_kernel(RbfSet r, const int64_t* __restrict__ ptr, const double* __restrict__ rows,
                     const int64_t* __restrict__ mat_ptr, T* __restrict__ mat) {
  constexpr int F = 4 + 2 * L;
  __shared__ double cj[F][kTile];
  const int d = blockIdx.y;
  const int64_t p0 = ptr[d];
  const int n = static_cast<int>(ptr[d + 1] - p0) - L;
  if (n <= 0) return;
  const int nt = (n + kTile - 1) / kTile;
  const int ntiles = nt * (nt + 1) / 2;
  const double* rd = rows + static_cast<int64_t>(F) * (p0 - static_cast<int64_t>(d) * L);
  T* md = mat + mat_ptr[d];
  const int tid = threadIdx.x, li = tid & 63, cg = tid >> 6;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int I = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    while (I * (I + 1) / 2 > t) I--;
    const int J = t - I * (I + 1) / 2;
    __syncthreads();
    for (int e = tid; e < F * kTile; e += kT) {
      const int f = e / kTile, c = e % kTile, j = J * kTile + c;
      cj[f][c] = j < n ? rd[static_cast<int64_t>(f) * n + j] : 0.0;
    }
    const int i = I * kTile + li;
    double ri[F];
    for (int f = 0; f < F; f++) ri[f] = i < n ? rd[static_cast<int64_t>(f) * n + i] : 0.0;
    __syncthreads();
    if (i >= n) continue;
    const int ti = static_cast<int>(ri[3]);
    for (int c = 0; c < 16; c++) {
      const int jc = cg * 16 + c, j = J * kTile + jc;
      if (j > i || j >= n) break;
      const int tj = static_cast<int>(cj[3][jc]);
      double v = kernel_hb(r, ti, tj, ri[0] - cj[0][jc], ri[1] - cj[1][jc], ri[2] - cj[2][jc]);
      if (j == i && ti < 0) v += r.nugget;
      for (int a = 0; a < L; a++) v += ri[4 + a] * cj[4 + L + a][jc] + ri[4 + L + a] * cj[4 + a][jc];
      md[static_cast<int64_t>(j) * n + i] = static_cast<T>(v);
    }
  }
}

template <class T>
void launch_assemble_m_hb(int l, int max_n, int n_dom, const RbfSet& rs, const int64_t* ptr, const double* rows,
                          const int64_t* mat_ptr, T* mat, hipStream_t s) {
  if (n_dom <= 0 || max_n <= 0) return;
  const int nt = (max_n + kTile - 1) / kTile;
  const dim3 grid(static_cast<unsigned>(std::min(nt * (nt + 1) / 2, 2048)), static_cast<unsigned>(n_dom));
  switch (l) {
    case 0: assemble_m_hb_kernel<T, 0><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    case 1: assemble_m_hb_kernel<T, 1><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    case 4: assemble_m_hb_kernel<T, 4><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    case 10: assemble_m_hb_kernel<T, 10><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    default: throw_invalid("unsupported polynomial basis size");
  }
  PAM_LAUNCH_CHECK();
}

// Launch of assemble_m_kernel for the polynomial basis size l.
template <class T>
void launch_assemble_m(int l, int max_n, int n_dom, const RbfSet& rs, const int64_t* ptr, const double* rows,
                       const int64_t* mat_ptr, T* mat, hipStream_t s) {
  if (n_dom <= 0 || max_n <= 0) return;
  const int nt = (max_n + kTile - 1) / kTile;
  const dim3 grid(static_cast<unsigned>(std::min(nt * (nt + 1) / 2, 2048)), static_cast<unsigned>(n_dom));
  switch (l) {
    case 0: assemble_m_kernel<T, 0><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    case 1: assemble_m_kernel<T, 1><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    case 4: assemble_m_kernel<T, 4><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    case 10: assemble_m_kernel<T, 10><<<grid, kT, 0, s>>>(rs, ptr, rows, mat_ptr, mat); break;
    default: throw_invalid("unsupported polynomial basis size");
  }
  PAM_LAUNCH_CHECK();
}

// One workgroup factors one domain: right-looking Cholesky with 32-wide
// panels. Per panel: one wavefront factors the 32 x 32 diagonal block in LDS
// (wave-level synchronisation only) and inverts it; the rows below are the
// panel times that inverse (independent dot products, no sequential solve);
// the trailing update A22 -= L21 L21^T runs over 64 x 64 tiles of the lower
// triangle on FP64 MFMA (4 waves as 2 x 2, 32 x 32 each), the tile result
// staged in LDS so that the read-modify-write of A22 is coalesced.
// T = double (FP64 MFMA) or float (FP32 MFMA, single-precision fine-level
// factors).
template <class T, int NB>
struct CholSmem {
  T dg[NB][NB + 1];     // diagonal block (L11 after the factorisation)
  T pi[kTile][NB + 1];   // panel rows of tile I
  T pj[kTile][NB + 1];   // panel rows of tile J
  T dt[kTile][kTile + 1]; // MFMA result of a tile; the inverse of L11 during the panel phase
};

template <class T>
struct CholMfma;
template <>
struct CholMfma<double> {
  using acc_t = double __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D layout: register q holds row lane / 16 + 4 q, column lane % 16
  static __device__ __forceinline__ int row(int lane, int q) { return (lane >> 4) + 4 * q; }
  static constexpr double kTiny = 1e-300;
};
template <>
struct CholMfma<float> {
  using acc_t = float __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D layout: register q holds row 4 (lane / 16) + q, column lane % 16
  static __device__ __forceinline__ int row(int lane, int q) { return 4 * (lane >> 4) + q; }
  static constexpr float kTiny = 1e-30f;
};

template <class T, int NB>
__device__ void chol_rows(T* __restrict__ a, int n, int k0, const T (*inv)[NB + 1], int first, int stride);

// Diagonal block of panel step k0 of the factorisation of the n x n matrix a
// (column-major, lower triangle): L11 by one wavefront, stored, and its
// inverse left in sm.dt. chol_panel adds L21 = A21 L11^-T and ends with a
// barrier.
template <class T, int NB>
__device__ void chol_diag(CholSmem<T, NB>& sm, T* __restrict__ a, int n, int k0, int32_t* __restrict__ flag) {
  const int lda = n;
  const int kb = min(NB, n - k0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  T (*inv)[NB + 1] = reinterpret_cast<T (*)[NB + 1]>(&sm.dt[0][0]);
  for (int e = tid; e < kb * kb; e += kT) {
    int rr = e % kb, cc = e / kb;
    sm.dg[rr][cc] = a[static_cast<int64_t>(k0 + cc) * lda + k0 + rr];
  }
  __syncthreads();
  if (wave == 0) {
    for (int j = 0; j < kb; j++) {
      if (lane == 0) {
        T v = sm.dg[j][j];
        if (!(v > T(0))) {
          atomicExch(flag, 1);
          v = fabs(v) + CholMfma<T>::kTiny;
        }
        sm.dg[j][j] = sqrt(v);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      if (lane > j && lane < kb) sm.dg[lane][j] /= sm.dg[j][j];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      const int w = kb - j - 1;
      for (int e = lane; e < w * w; e += 64) {
        int rr = j + 1 + e / w, cc = j + 1 + e % w;
        if (cc <= rr) sm.dg[rr][cc] -= sm.dg[rr][j] * sm.dg[cc][j];
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    // inverse of L11, one column per lane: inv[r][c] for r >= c
    if (lane < kb) {
      const int c = lane;
      for (int r = 0; r < kb; r++) {
        if (r < c) {
          inv[r][c] = T(0);
          continue;
        }
        T v = r == c ? T(1) : T(0);
        for (int t = c; t < r; t++) v -= sm.dg[r][t] * inv[t][c];
        inv[r][c] = v / sm.dg[r][r];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < kb * kb; e += kT) {
    int rr = e % kb, cc = e / kb;
    if (rr >= cc) a[static_cast<int64_t>(k0 + cc) * lda + k0 + rr] = sm.dg[rr][cc];
  }
}

template <class T, int NB>
__device__ void chol_panel(CholSmem<T, NB>& sm, T* __restrict__ a, int n, int k0, int32_t* __restrict__ flag) {
  chol_diag(sm, a, n, k0, flag);
  const int kb = min(NB, n - k0);
  chol_rows<T, NB>(a, n, k0, reinterpret_cast<const T (*)[NB + 1]>(&sm.dt[0][0]), k0 + kb + static_cast<int>(threadIdx.x),
                   kT);
  __syncthreads();
}

// Rows i = first, first + stride, ... < n of L21 = A21 L11^-T, given the
// inverse of the panel's diagonal block.
template <class T, int NB>
__device__ void chol_rows(T* __restrict__ a, int n, int k0, const T (*inv)[NB + 1], int first, int stride) {
  const int lda = n;
  const int kb = min(NB, n - k0);
  for (int i = first; i < n; i += stride) {
    T x[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) x[t] = t < kb ? a[static_cast<int64_t>(k0 + t) * lda + i] : T(0);
#pragma unroll 4
    for (int c = 0; c < kb; c++) {
      T v = T(0);
#pragma unroll
      for (int t = 0; t < NB; t++)
        if (t <= c) v += x[t] * inv[c][t];
      a[static_cast<int64_t>(k0 + c) * lda + i] = v;
    }
  }
}

// Trailing tile (I, J), I >= J, of panel step k0: A22 -= L21 L21^T on FP64
// MFMA (4 waves as 2 x 2, 32 x 32 each). load_i: stage tile I's panel rows
// (they can stay staged across the J of one I). Ends with a barrier.
template <class T, int NB>
__device__ void chol_tile(CholSmem<T, NB>& sm, T* __restrict__ a, int n, int k0, int I, int J, bool load_i) {
  using acc_t = typename CholMfma<T>::acc_t;
  const int lda = n;
  const int kb = min(NB, n - k0);
  const int base = k0 + kb;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wi = wave >> 1, wj = wave & 1;
  const int r0 = base + I * kTile, c0 = base + J * kTile;
  if (load_i) {
    for (int e = tid; e < kTile * NB; e += kT) {
      int rr = e % kTile, cc = e / kTile;
      sm.pi[rr][cc] = (r0 + rr < n && cc < kb) ? a[static_cast<int64_t>(k0 + cc) * lda + r0 + rr] : T(0);
    }
  }
  for (int e = tid; e < kTile * NB; e += kT) {
    int rr = e % kTile, cc = e / kTile;
    sm.pj[rr][cc] = (c0 + rr < n && cc < kb) ? a[static_cast<int64_t>(k0 + cc) * lda + c0 + rr] : T(0);
  }
  __syncthreads();
  acc_t acc[2][2];
#pragma unroll
  for (int ii = 0; ii < 2; ii++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++) acc[ii][jj] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int ks = 0; ks < NB / 4; ks++) {
    T af[2], bf[2];
#pragma unroll
    for (int ii = 0; ii < 2; ii++) af[ii] = sm.pi[wi * 32 + ii * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
    for (int jj = 0; jj < 2; jj++) bf[jj] = sm.pj[wj * 32 + jj * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
    for (int ii = 0; ii < 2; ii++)
#pragma unroll
      for (int jj = 0; jj < 2; jj++)
        acc[ii][jj] = CholMfma<T>::op(af[ii], bf[jj], acc[ii][jj]);
  }
#pragma unroll
  for (int ii = 0; ii < 2; ii++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        sm.dt[wi * 32 + ii * 16 + CholMfma<T>::row(lane, q)][wj * 32 + jj * 16 + (lane & 15)] = acc[ii][jj][q];
  __syncthreads();
  for (int e = tid; e < kTile * kTile; e += kT) {
    int rr = e % kTile, cc = e / kTile;
    int row = r0 + rr, col = c0 + cc;
    if (row < n && col < n && row >= col) a[static_cast<int64_t>(col) * lda + row] -= sm.dt[rr][cc];
  }
  __syncthreads();
}

// One workgroup per domain, all panel steps (many domains at once).
template <class T, int NB>
__global__ void __launch_bounds__(kT)
cholesky_kernel(const int64_t* __restrict__ ptr, int l, const int64_t* __restrict__ mat_ptr,
                T* __restrict__ mat, int32_t* __restrict__ flag) {
  __shared__ CholSmem<T, NB> sm;
  const int d = blockIdx.x;
  const int n = static_cast<int>(ptr[d + 1] - ptr[d]) - l;
  if (n <= 0) return;
  T* a = mat + mat_ptr[d];
  for (int k0 = 0; k0 < n; k0 += NB) {
    chol_panel(sm, a, n, k0, flag);
    const int rem = n - min(NB, n - k0) - k0;
    if (rem <= 0) break;
    const int nt = (rem + kTile - 1) / kTile;
    for (int I = 0; I < nt; I++)
      for (int J = 0; J <= I; J++) chol_tile(sm, a, n, k0, I, J, J == 0);
  }
}

// Left-looking batched factorisation: one workgroup per domain, block column
// by block column (64 wide). Tile (I, J), I >= J, is
//   A_IJ - sum_{K<J} L_IK L_JK^T
// accumulated on MFMA in registers (4 waves as 2 x 2, 32 x 32 each) from
// 32-column chunks of the two row panels, staged through LDS with the next
// chunk's loads in flight; the diagonal tile is then factored in LDS by all
// threads (one barrier per column) and inverted (Gauss-Jordan, one barrier per
// column), and each tile below it is multiplied by that inverse's transpose on
// MFMA. Every entry of L is written once and A is never read back, so the
// trailing matrix is not re-read and re-written per panel as in the
// right-looking form.
constexpr int kKc = 32;  // reduction chunk of the tile updates

template <class T>
struct LlSmem {
  union {
    struct {
      T pi[kTile][kKc + 1];  // chunk of row panel I (negated)
      T pj[kTile][kKc + 1];  // chunk of row panel J
    } g;
    T dt[kTile][kTile + 1];  // a tile in transit
  } u;
  T dg[kTile][kTile + 1];   // diagonal tile; L_JJ after the factorisation
  T inv[kTile][kTile + 1];  // L_JJ^-1 (lower)
};

// acc = A_IJ - sum_{K<J} L_IK L_JK^T for the calling wave's 32 x 32 part.
template <class T>
__device__ void ll_tile_update(LlSmem<T>& sm, const T* __restrict__ a, int n, int I, int J,
                               typename CholMfma<T>::acc_t (&acc)[2][2]) {
  using M = CholMfma<T>;
  const int64_t lda = n;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wi = wave >> 1, wj = wave & 1;
  const int r0 = I * kTile, c0 = J * kTile;
#pragma unroll
  for (int ii = 0; ii < 2; ii++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int r = r0 + wi * 32 + ii * 16 + M::row(lane, q), c = c0 + wj * 32 + jj * 16 + (lane & 15);
        acc[ii][jj][q] = (r < n && c < n) ? a[static_cast<int64_t>(c) * lda + r] : T(0);
      }
  const int nk = c0 / kKc;
  if (nk == 0) return;
  constexpr int kPer = kTile * kKc / kT;  // 8 entries of each panel chunk per thread
  const int rr = tid & 63, kk0 = tid >> 6;
  const bool in_i = r0 + rr < n, in_j = c0 + rr < n;
  T ri[kPer], rj[kPer];
  auto load = [&](int k) {
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int64_t col = static_cast<int64_t>(k * kKc + kk0 + 4 * q) * lda;
      ri[q] = in_i ? a[col + r0 + rr] : T(0);
      rj[q] = in_j ? a[col + c0 + rr] : T(0);
    }
  };
  load(0);
  for (int k = 0; k < nk; k++) {
    __syncthreads();  // the previous chunk's reads are done
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      sm.u.g.pi[rr][kk0 + 4 * q] = -ri[q];
      sm.u.g.pj[rr][kk0 + 4 * q] = rj[q];
    }
    __syncthreads();
    if (k + 1 < nk) load(k + 1);
#pragma unroll
    for (int ks = 0; ks < kKc / 4; ks++) {
      T af[2], bf[2];
#pragma unroll
      for (int ii = 0; ii < 2; ii++) af[ii] = sm.u.g.pi[wi * 32 + ii * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
      for (int jj = 0; jj < 2; jj++) bf[jj] = sm.u.g.pj[wj * 32 + jj * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
      for (int ii = 0; ii < 2; ii++)
#pragma unroll
        for (int jj = 0; jj < 2; jj++) acc[ii][jj] = M::op(af[ii], bf[jj], acc[ii][jj]);
    }
  }
}

template <class T>
__device__ __forceinline__ void ll_store_acc(T (*dst)[kTile + 1], const typename CholMfma<T>::acc_t (&acc)[2][2]) {
  using M = CholMfma<T>;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, wi = wave >> 1, wj = wave & 1;
#pragma unroll
  for (int ii = 0; ii < 2; ii++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        dst[wi * 32 + ii * 16 + M::row(lane, q)][wj * 32 + jj * 16 + (lane & 15)] = acc[ii][jj][q];
}

template <class T>
__global__ void __launch_bounds__(kT)
cholesky_ll_kernel(const int64_t* __restrict__ ptr, int l, const int64_t* __restrict__ mat_ptr,
                   T* __restrict__ mat, int32_t* __restrict__ flag) {
  using M = CholMfma<T>;
  using acc_t = typename M::acc_t;
  __shared__ LlSmem<T> sm;
  const int d = blockIdx.x;
  const int n = static_cast<int>(ptr[d + 1] - ptr[d]) - l;
  if (n <= 0) return;
  T* a = mat + mat_ptr[d];
  const int64_t lda = n;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wi = wave >> 1, wj = wave & 1;
  const int nt = (n + kTile - 1) / kTile;
  const int rr = tid & 63, cg = tid >> 6;  // row, and column group (of four) of the LDS passes
  for (int J = 0; J < nt; J++) {
    const int c0 = J * kTile, kb = min(kTile, n - c0);
    acc_t acc[2][2];
    ll_tile_update(sm, a, n, J, J, acc);
    // The diagonal tile, padded with the identity past the matrix.
#pragma unroll
    for (int ii = 0; ii < 2; ii++)
#pragma unroll
      for (int jj = 0; jj < 2; jj++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int r = wi * 32 + ii * 16 + M::row(lane, q), c = wj * 32 + jj * 16 + (lane & 15);
          sm.dg[r][c] = (r < kb && c < kb) ? acc[ii][jj][q] : (r == c ? T(1) : T(0));
        }
    __syncthreads();
    // Cholesky of dg. Step j updates columns c > j with column j as it was
    // before its scaling (a_rj a_cj / a_jj); column j is scaled in step j + 1,
    // which reads columns > j + 1 only, so one barrier per step suffices.
    for (int j = 0; j < kTile; j++) {
      T piv = sm.dg[j][j];
      if (!(piv > T(0))) {
        if (tid == 0) atomicExch(flag, 1);
        piv = fabs(piv) + M::kTiny;
      }
      const T rp = T(1) / piv;
      if (j > 0 && tid < kTile) {  // scale column j - 1
        const T pj1 = sm.dg[j - 1][j - 1];
        const T p1 = pj1 > T(0) ? pj1 : fabs(pj1) + M::kTiny;
        const T s = sqrt(p1);
        if (rr > j - 1) sm.dg[rr][j - 1] /= s;
      }
      const T lrj = sm.dg[rr][j];
      for (int c = j + 1 + cg; c < kTile; c += 4)
        if (rr >= c) sm.dg[rr][c] -= lrj * (sm.dg[c][j] * rp);
      __syncthreads();
      if (j > 0 && tid == 0) {  // the diagonal entry of column j - 1, after every reader of it
        const T pj1 = sm.dg[j - 1][j - 1];
        sm.dg[j - 1][j - 1] = sqrt(pj1 > T(0) ? pj1 : fabs(pj1) + M::kTiny);
      }
    }
    if (tid < kTile) {
      const T p = sm.dg[kTile - 1][kTile - 1];
      const T s = sqrt(p > T(0) ? p : fabs(p) + M::kTiny);
      if (rr == kTile - 1) sm.dg[rr][rr] = s;
    }
    __syncthreads();
    for (int e = tid; e < kb * kb; e += kT) {
      const int r = e % kb, c = e / kb;
      if (r >= c) a[static_cast<int64_t>(c0 + c) * lda + c0 + r] = sm.dg[r][c];
    }
    if (J + 1 == nt) break;
    // inv = L_JJ^-1 by forward elimination on the identity. Step j
    // subtracts multiples of row j (as it was before its division) from the
    // rows below; row j is divided in step j + 1, which does not read it.
    for (int e = tid; e < kTile * kTile; e += kT) {
      const int r = e % kTile, c = e / kTile;
      sm.inv[r][c] = r == c ? T(1) : T(0);
    }
    __syncthreads();
    for (int j = 0; j < kTile; j++) {
      const T rd = T(1) / sm.dg[j][j];
      if (j > 0 && tid < kTile && rr <= j - 1) sm.inv[j - 1][rr] /= sm.dg[j - 1][j - 1];
      if (rr > j) {
        const T f = sm.dg[rr][j] * rd;
        for (int c = cg; c <= j; c += 4) sm.inv[rr][c] -= f * sm.inv[j][c];
      }
      __syncthreads();
    }
    if (tid < kTile) sm.inv[kTile - 1][rr] /= sm.dg[kTile - 1][kTile - 1];
    __syncthreads();
    for (int I = J + 1; I < nt; I++) {
      ll_tile_update(sm, a, n, I, J, acc);
      __syncthreads();  // dt overlays the panel chunks
      ll_store_acc(sm.u.dt, acc);
      __syncthreads();
      acc_t o[2][2];
#pragma unroll
      for (int ii = 0; ii < 2; ii++)
#pragma unroll
        for (int jj = 0; jj < 2; jj++) o[ii][jj] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int ks = 0; ks < kTile / 4; ks++) {
        T af[2], bf[2];
#pragma unroll
        for (int ii = 0; ii < 2; ii++) af[ii] = sm.u.dt[wi * 32 + ii * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
        for (int jj = 0; jj < 2; jj++) bf[jj] = sm.inv[wj * 32 + jj * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
#pragma unroll
        for (int ii = 0; ii < 2; ii++)
#pragma unroll
          for (int jj = 0; jj < 2; jj++) o[ii][jj] = M::op(af[ii], bf[jj], o[ii][jj]);
      }
      __syncthreads();
      ll_store_acc(sm.u.dt, o);
      __syncthreads();
      const int r0 = I * kTile;
      for (int e = tid; e < kTile * kTile; e += kT) {
        const int r = e % kTile, c = e / kTile;
        if (r0 + r < n) a[static_cast<int64_t>(c0 + c) * lda + r0 + r] = sm.u.dt[r][c];
      }
      __syncthreads();
    }
  }
}

// The batched factorisation: right-looking with panel width nb (32 or 64), or
// left-looking (nb = 0).
template <class T>
void launch_cholesky_batch(int nb, int n_dom, const int64_t* ptr, int l, const int64_t* mat_ptr, T* mat,
                           int32_t* flag, hipStream_t s) {
  if (n_dom <= 0) return;
  if (nb == 0) {
    cholesky_ll_kernel<T><<<n_dom, kT, 0, s>>>(ptr, l, mat_ptr, mat, flag);
  } else if (nb == 64) {
    cholesky_kernel<T, 64><<<n_dom, kT, 0, s>>>(ptr, l, mat_ptr, mat, flag);
  } else if (nb == 32) {
    cholesky_kernel<T, 32><<<n_dom, kT, 0, s>>>(ptr, l, mat_ptr, mat, flag);
  } else {
    throw_invalid("unsupported Cholesky panel width");
  }
  PAM_LAUNCH_CHECK();
}

// A single large matrix, per panel step: the diagonal block by one
// workgroup (its inverse written out), the rows below it by one thread per
// row, the trailing tiles by one workgroup per tile.
__global__ void __launch_bounds__(kT)
cholesky_panel_kernel(double* __restrict__ a, int n, int k0, int32_t* __restrict__ flag, double* __restrict__ inv_out) {
  __shared__ CholSmem<double, kNb> sm;
  chol_diag(sm, a, n, k0, flag);
  const double (*inv)[kNb + 1] = reinterpret_cast<const double (*)[kNb + 1]>(&sm.dt[0][0]);
  for (int e = threadIdx.x; e < kNb * kNb; e += kT) inv_out[e] = inv[e / kNb][e % kNb];
}

// The rows below the panel's diagonal block, spread over workgroups.
__global__ void __launch_bounds__(kT)
cholesky_rows_kernel(double* __restrict__ a, int n, int k0, const double* __restrict__ inv_in) {
  __shared__ double inv[kNb][kNb + 1];
  for (int e = threadIdx.x; e < kNb * kNb; e += kT) inv[e / kNb][e % kNb] = inv_in[e];
  __syncthreads();
  const int kb = min(kNb, n - k0);
  chol_rows<double, kNb>(a, n, k0, inv, k0 + kb + static_cast<int>(blockIdx.x * kT + threadIdx.x),
                         static_cast<int>(gridDim.x * kT));
}

__global__ void __launch_bounds__(kT)
cholesky_trailing_kernel(double* __restrict__ a, int n, int k0) {
  __shared__ CholSmem<double, kNb> sm;
  // blockIdx.x -> (I, J) with I >= J
  const int t = blockIdx.x;
  int I = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) / 2.0);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  const int J = t - I * (I + 1) / 2;
  chol_tile(sm, a, n, k0, I, J, true);
}


// Sum of the partials of the first 64 threads, added in lane order by one
// thread: the result does not depend on the workgroup size.
__device__ __forceinline__ double lane_ordered_sum(double v, double* sh) {
  if (threadIdx.x < 64) sh[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 64; k++) t += sh[k];
    sh[64] = t;
  }
  __syncthreads();
  const double r = sh[64];
  __syncthreads();
  return r;
}

// One workgroup solves one domain (fine_grid.hpp:97-142, coarse_grid.hpp:105-131):
// Q^T d, L L^T gamma = Q^T d, lambda = Q gamma; the coarse grid also solves
// P_top c = d_top - A_top lambda. vals (LDS, m entries) holds d on entry;
// emit(u, v) receives entry u < m of the solution (domain order) and
// emit(m + a, c_a) the polynomial coefficients of the coarse grid.
// KT threads (a multiple of 64, at least 128); every sum is taken in an order
// that does not depend on KT, so the solution is the same bits for any KT.
template <int KT, class Emit, class TF = double>
__device__ void solve_domain(RbfSet r, int l, const double* __restrict__ pts, const double* __restrict__ gpts,
                             int64_t mu_v, const int32_t* __restrict__ didx,
                             int m, const double* __restrict__ lag, const TF* __restrict__ L,
                             double* vals, double* y, double* sh, int coarse, const double* __restrict__ pinv,
                             Emit emit) {
  __shared__ double lam_top[10];
  __shared__ double rhs_top[10];
  const int n = m - l;
  if (n > 0) {
    for (int i = threadIdx.x; i < n; i += KT) {
      double v = vals[l + i];
      int qi = didx[l + i];
      for (int a = 0; a < l; a++) v -= lag[static_cast<int64_t>(qi) * l + a] * vals[a];
      y[i] = v;
    }
    __syncthreads();
    const int lda = n;
    // Diagonal blocks are staged in LDS before their sequential solve, and
    // full panels read their 32 factor entries per row with independent
    // loads: both keep the sequential parts off global-memory latency.
    __shared__ double dgl[kNb][kNb + 1];
    auto stage_diag = [&](int k0, int kb) {
      for (int e = threadIdx.x; e < kb * kb; e += KT) {
        const int rr = e % kb, cc = e / kb;
        dgl[rr][cc] = static_cast<double>(L[static_cast<int64_t>(k0 + cc) * lda + k0 + rr]);
      }
      __syncthreads();
    };
    // forward: L y = qtd, 32-row blocks
    for (int k0 = 0; k0 < n; k0 += kNb) {
      int kb = min(kNb, n - k0);
      stage_diag(k0, kb);
      if (threadIdx.x < 64) {  // one wavefront solves the diagonal block
        for (int t = 0; t < kb; t++) {
          if (static_cast<int>(threadIdx.x) == t) y[k0 + t] /= dgl[t][t];
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          int c = threadIdx.x;
          if (c > t && c < kb) y[k0 + c] -= dgl[c][t] * y[k0 + t];
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
      }
      __syncthreads();
      if (kb == kNb) {
        for (int i = k0 + kNb + threadIdx.x; i < n; i += KT) {
          double lv[kNb];
#pragma unroll
          for (int t = 0; t < kNb; t++) lv[t] = static_cast<double>(L[static_cast<int64_t>(k0 + t) * lda + i]);
          double v = y[i];
#pragma unroll
          for (int t = 0; t < kNb; t++) v -= lv[t] * y[k0 + t];
          y[i] = v;
        }
      } else {
        for (int i = k0 + kb + threadIdx.x; i < n; i += KT) {
          double v = y[i];
          for (int t = 0; t < kb; t++) v -= static_cast<double>(L[static_cast<int64_t>(k0 + t) * lda + i]) * y[k0 + t];
          y[i] = v;
        }
      }
      __syncthreads();
    }
    // backward: L^T x = y
    for (int k0 = ((n - 1) / kNb) * kNb; k0 >= 0; k0 -= kNb) {
      int kb = min(kNb, n - k0);
      stage_diag(k0, kb);
      if (threadIdx.x < 64) {
        for (int t = kb - 1; t >= 0; t--) {
          if (static_cast<int>(threadIdx.x) == t) y[k0 + t] /= dgl[t][t];
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          int c = threadIdx.x;
          if (c < t) y[k0 + c] -= dgl[t][c] * y[k0 + t];
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
      }
      __syncthreads();
      if (kb == kNb) {
        for (int i = threadIdx.x; i < k0; i += KT) {
          double lv[kNb];
#pragma unroll
          for (int t = 0; t < kNb; t++) lv[t] = static_cast<double>(L[static_cast<int64_t>(i) * lda + k0 + t]);
          double v = y[i];
#pragma unroll
          for (int t = 0; t < kNb; t++) v -= lv[t] * y[k0 + t];
          y[i] = v;
        }
      } else {
        for (int i = threadIdx.x; i < k0; i += KT) {
          double v = y[i];
          for (int t = 0; t < kb; t++) v -= static_cast<double>(L[static_cast<int64_t>(i) * lda + k0 + t]) * y[k0 + t];
          y[i] = v;
        }
      }
      __syncthreads();
    }
  }
  // lambda_top[a] = sum_i Q[a][i] gamma_i = -sum_i lag[q_i][a] gamma_i
  for (int a = 0; a < l; a++) {
    double acc = 0.0;
    if (threadIdx.x < 64)
      for (int i = threadIdx.x; i < n; i += 64) acc -= lag[static_cast<int64_t>(didx[l + i]) * l + a] * y[i];
    double t = lane_ordered_sum(acc, sh);
    if (threadIdx.x == 0) lam_top[a] = t;
  }
  __syncthreads();
  for (int u = threadIdx.x; u < m; u += KT) emit(u, u < l ? lam_top[u] : y[u - l]);
  if (coarse && l > 0) {
    // polynomial part: P_top c = d_top - A_top lambda
    for (int a = 0; a < l; a++) {
      double acc = 0.0;
      int pa = didx[a];
      if (threadIdx.x < 64)
        for (int u = threadIdx.x; u < m; u += 64) {
          double lu = u < l ? lam_top[u] : y[u - l];
          acc += flat_entry(r, pts, gpts, mu_v, pa, didx[u]) * lu;
        }
      double t = lane_ordered_sum(acc, sh);
      if (threadIdx.x == 0) rhs_top[a] = vals[a] - t;
    }
    __syncthreads();
    if (threadIdx.x < static_cast<unsigned>(l)) {
      double c = 0.0;
      for (int b = 0; b < l; b++) c += pinv[threadIdx.x * l + b] * rhs_top[b];
      emit(m + static_cast<int>(threadIdx.x), c);
    }
  }
}

// Domain solves of one batch: gather d from res, scatter the owned entries
// of the solution (every entry for the coarse grid) and the coarse grid's
// polynomial coefficients into out.
template <int KT, class TF>
__global__ void __launch_bounds__(KT)
ras_solve_kernel(RbfSet r, int l, const double* __restrict__ pts, const double* __restrict__ gpts, int64_t mu_v,
                 const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                 const uint8_t* __restrict__ inner, const double* __restrict__ lag,
                 const int64_t* __restrict__ mat_ptr, const TF* __restrict__ mat,
                 const double* __restrict__ res, double* __restrict__ out, int64_t mu, int coarse,
                 const double* __restrict__ pinv) {
  extern __shared__ double sm[];
  __shared__ double sh[KT];
  const int d = blockIdx.x;
  const int64_t p0 = ptr[d];
  const int m = static_cast<int>(ptr[d + 1] - p0);
  double* vals = sm;      // [m]
  double* y = sm + m;     // [m - l]
  for (int u = threadIdx.x; u < m; u += KT) vals[u] = res[idx[p0 + u]];
  __syncthreads();
  const int32_t* didx = idx + p0;
  const uint8_t* din = inner + p0;
  solve_domain<KT>(r, l, pts, gpts, mu_v, didx, m, lag, mat + mat_ptr[d], vals, y, sh, coarse, pinv, [&](int u, double v) {
    if (u < m) {
      if (coarse || din[u]) out[didx[u]] = v;
    } else {
      out[mu + (u - m)] = v;
    }
  });
}

// The coarse grid's solve as a dense operator: workgroup u0 solves for the
// unit vector e_u0 and writes column u0 of op ((m + l) x m, row-major).
__global__ void __launch_bounds__(kT)
coarse_operator_kernel(RbfSet r, int l, const double* __restrict__ pts, const double* __restrict__ gpts, int64_t mu_v,
                       const int32_t* __restrict__ idx, int m,
                       const double* __restrict__ lag, const double* __restrict__ mat,
                       const double* __restrict__ pinv, double* __restrict__ op) {
  extern __shared__ double sm[];
  __shared__ double sh[kT];
  const int u0 = blockIdx.x;
  double* vals = sm;
  double* y = sm + m;
  for (int u = threadIdx.x; u < m; u += kT) vals[u] = u == u0 ? 1.0 : 0.0;
  __syncthreads();
  solve_domain<kT>(r, l, pts, gpts, mu_v, idx, m, lag, mat, vals, y, sh, 1, pinv,
               [&](int u, double v) { op[static_cast<int64_t>(u) * m + u0] = v; });
}

// Coarse solve by the dense operator: out[idx[u]] = (op d)[u] for u < m and
// out[mu + a] = (op d)[m + a], d = res[idx[.]]. One wavefront per row, the
// row's products summed in a fixed order.
constexpr int kCoarseRowsPerBlock = kT / 64;
constexpr int kCoarseOpMax = 8192;  // largest coarse grid given a dense operator (512 MiB)

__global__ void __launch_bounds__(kT)
coarse_apply_kernel(const int32_t* __restrict__ idx, int m, int l, const double* __restrict__ op,
                    const double* __restrict__ res, double* __restrict__ out, int64_t mu) {
  extern __shared__ double x[];
  for (int u = threadIdx.x; u < m; u += kT) x[u] = res[idx[u]];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kCoarseRowsPerBlock + static_cast<int>(threadIdx.x >> 6);
  if (row >= m + l) return;
  const double* __restrict__ orow = op + static_cast<int64_t>(row) * m;
  double acc = 0.0;
  for (int u = lane; u < m; u += 64) acc += orow[u] * x[u];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) {
    if (row < m) out[idx[row]] = acc;
    else out[mu + (row - m)] = acc;
  }
}

// ---- host helpers ----

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

void host_monomials(int degree, const double* p, double* m) {
  m[0] = 1.0;
  if (degree >= 1) { m[1] = p[0]; m[2] = p[1]; m[3] = p[2]; }
  if (degree >= 2) {
    m[4] = p[0] * p[0]; m[5] = p[0] * p[1]; m[6] = p[0] * p[2];
    m[7] = p[1] * p[1]; m[8] = p[1] * p[2]; m[9] = p[2] * p[2];
  }
}

// Inverse by Gauss-Jordan with partial pivoting; returns the reciprocal
// condition estimate 1 / (|A|_1 |A^-1|_1), 0 when singular.
double invert(std::vector<double> a, int n, std::vector<double>& inv) {
  inv.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; i++) inv[i * n + i] = 1.0;
  double norm_a = 0.0;
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += std::abs(a[i * n + j]);
    norm_a = std::max(norm_a, s);
  }
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int i = c + 1; i < n; i++)
      if (std::abs(a[i * n + c]) > std::abs(a[p * n + c])) p = i;
    if (std::abs(a[p * n + c]) <= 1e-14 * norm_a) return 0.0;
    if (p != c) {
      for (int j = 0; j < n; j++) {
        std::swap(a[p * n + j], a[c * n + j]);
        std::swap(inv[p * n + j], inv[c * n + j]);
      }
    }
    double piv = a[c * n + c];
    for (int j = 0; j < n; j++) {
      a[c * n + j] /= piv;
      inv[c * n + j] /= piv;
    }
    for (int i = 0; i < n; i++) {
      if (i == c) continue;
      double f = a[i * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        a[i * n + j] -= f * a[c * n + j];
        inv[i * n + j] -= f * inv[c * n + j];
      }
    }
  }
  double norm_inv = 0.0;
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += std::abs(inv[i * n + j]);
    norm_inv = std::max(norm_inv, s);
  }
  return 1.0 / (norm_a * norm_inv);
}

// UnisolventPointSet (polynomial/unisolvent_point_set.hpp:25-66): 100 random
// trials of l distinct points, keep the best-conditioned. The random stream is
// this library's own (splitmix64), so the chosen indices differ from polatory's.
std::vector<int32_t> unisolvent_point_set(const std::vector<double>& pts, int degree) {
  int l = poly_basis_size(degree);
  int64_t n = static_cast<int64_t>(pts.size() / 3);
  if (l == 0) return {};
  if (n < l) throw_invalid("points.rows() must be greater than or equal to " + std::to_string(l));
  uint64_t seed = 0x5eedu;
  std::vector<int32_t> best;
  double best_rcond = 0.0;
  for (int trial = 0; trial < 100; trial++) {
    std::vector<int32_t> set;
    while (static_cast<int>(set.size()) < l) {
      int32_t c = static_cast<int32_t>(splitmix64(seed) % static_cast<uint64_t>(n));
      if (std::find(set.begin(), set.end(), c) == set.end()) set.push_back(c);
    }
    std::sort(set.begin(), set.end());
    std::vector<double> p(static_cast<size_t>(l) * l), inv;
    for (int i = 0; i < l; i++) {
      double m[10];
      host_monomials(degree, &pts[3 * static_cast<size_t>(set[i])], m);
      for (int k = 0; k < l; k++) p[i * l + k] = m[k];
    }
    double rc = invert(p, l, inv);
    if (rc > best_rcond) {
      best_rcond = rc;
      best = set;
    }
  }
  if (best.empty()) throw_runtime("could not find a unisolvent set of points");
  return best;
}

double round_half_to_even(double d) { return std::ceil((d - 0.5) / 2.0) + std::floor((d + 0.5) / 2.0); }

struct Bbox3 {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void add(const double* p) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  std::array<int, 3> axes_by_width() const {
    std::array<int, 3> ax{0, 1, 2};
    std::stable_sort(ax.begin(), ax.end(), [&](int i, int j) { return hi[i] - lo[i] > hi[j] - lo[j]; });
    return ax;
  }
};

// Positions of ids in lexicographic order of their coordinates along ax, ties
// by position: the order a stable sort by coordinates gives. Keys are copied
// into one contiguous array; large inputs use the OpenMP parallel sort
// (deterministic, since the order is total).
std::vector<int32_t> order_by_axes(const int32_t* ids, int64_t n, const std::vector<double>& pts,
                                   const std::array<int, 3>& ax, bool parallel) {
  struct Key {
    double k[3];
    int32_t pos;
  };
  std::vector<Key> keys(static_cast<size_t>(n));
#pragma omp parallel for schedule(static) if (parallel && n >= (int64_t{1} << 16))
  for (int64_t i = 0; i < n; i++) {
    const double* p = &pts[3 * static_cast<size_t>(ids[i])];
    keys[i] = Key{{p[ax[0]], p[ax[1]], p[ax[2]]}, static_cast<int32_t>(i)};
  }
  auto less = [](const Key& a, const Key& b) {
    if (a.k[0] != b.k[0]) return a.k[0] < b.k[0];
    if (a.k[1] != b.k[1]) return a.k[1] < b.k[1];
    if (a.k[2] != b.k[2]) return a.k[2] < b.k[2];
    return a.pos < b.pos;
  };
  if (parallel && n >= (int64_t{1} << 16)) __gnu_parallel::sort(keys.begin(), keys.end(), less);
  else std::sort(keys.begin(), keys.end(), less);
  std::vector<int32_t> order(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; i++) order[i] = keys[i].pos;
  return order;
}

// Like order_by_axes, but only as far as a split needs it: positions
// [lo, hi) hold exactly the entries of those ranks, in order; the entries
// below lo and at or above hi are the right sets in any order.
std::vector<int32_t> order_for_split(const int32_t* ids, int64_t n, const std::vector<double>& pts,
                                     const std::array<int, 3>& ax, int64_t lo, int64_t hi, bool parallel) {
  if (hi - lo > n / 4) return order_by_axes(ids, n, pts, ax, parallel);
  struct Key {
    double k[3];
    int32_t pos;
  };
  std::vector<Key> keys(static_cast<size_t>(n));
#pragma omp parallel for schedule(static) if (parallel && n >= (int64_t{1} << 16))
  for (int64_t i = 0; i < n; i++) {
    const double* p = &pts[3 * static_cast<size_t>(ids[i])];
    keys[i] = Key{{p[ax[0]], p[ax[1]], p[ax[2]]}, static_cast<int32_t>(i)};
  }
  auto less = [](const Key& a, const Key& b) {
    if (a.k[0] != b.k[0]) return a.k[0] < b.k[0];
    if (a.k[1] != b.k[1]) return a.k[1] < b.k[1];
    if (a.k[2] != b.k[2]) return a.k[2] < b.k[2];
    return a.pos < b.pos;
  };
  std::vector<int32_t> order(static_cast<size_t>(n));
  if (parallel && n >= (int64_t{1} << 17)) {
    // Parallel selection: pivots from an evenly strided sample bracket the
    // ranks [lo, hi) with a margin; one pass sorts the keys into the three
    // ranges below, between and above the pivots (fixed chunks, so the
    // result does not depend on the thread count); the middle range then
    // holds ranks lo..hi and is finished sequentially. If the pivots miss
    // (the sample is off by more than the margin), the sequential selection
    // below runs instead.
    constexpr int64_t kSample = 8192, kMargin = 192, kChunk = int64_t{1} << 15;
    std::vector<Key> sample(static_cast<size_t>(kSample));
    for (int64_t i = 0; i < kSample; i++) sample[i] = keys[(i * n) / kSample];
    std::sort(sample.begin(), sample.end(), less);
    const int64_t s_lo = (lo * kSample) / n - kMargin, s_hi = (hi * kSample + n - 1) / n + kMargin;
    if (s_lo >= 0 && s_hi < kSample) {
      const Key p_lo = sample[s_lo], p_hi = sample[s_hi];
      const int64_t n_chunks = (n + kChunk - 1) / kChunk;
      std::vector<int64_t> cnt(static_cast<size_t>(3 * n_chunks), 0);
      auto cls = [&](const Key& k) { return less(k, p_lo) ? 0 : (less(p_hi, k) ? 2 : 1); };
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < n_chunks; c++)
        for (int64_t i = c * kChunk; i < std::min(n, (c + 1) * kChunk); i++) cnt[3 * c + cls(keys[i])]++;
      int64_t tot[3] = {0, 0, 0};
      for (int64_t c = 0; c < n_chunks; c++)
        for (int b = 0; b < 3; b++) tot[b] += cnt[3 * c + b];
      if (tot[0] <= lo && tot[0] + tot[1] >= hi) {
        std::vector<int64_t> off(static_cast<size_t>(3 * n_chunks));
        int64_t base[3] = {0, tot[0], tot[0] + tot[1]};
        for (int64_t c = 0; c < n_chunks; c++)
          for (int b = 0; b < 3; b++) {
            off[3 * c + b] = base[b];
            base[b] += cnt[3 * c + b];
          }
        std::vector<Key> out(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n_chunks; c++) {
          int64_t o[3] = {off[3 * c], off[3 * c + 1], off[3 * c + 2]};
          for (int64_t i = c * kChunk; i < std::min(n, (c + 1) * kChunk); i++) out[o[cls(keys[i])]++] = keys[i];
        }
        auto mid_b = out.begin() + tot[0], mid_e = mid_b + tot[1];
        std::nth_element(mid_b, out.begin() + lo, mid_e, less);
        std::nth_element(out.begin() + lo, out.begin() + hi, mid_e, less);
        std::sort(out.begin() + lo, out.begin() + hi, less);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) order[i] = out[i].pos;
        return order;
      }
    }
  }
  std::nth_element(keys.begin(), keys.begin() + lo, keys.end(), less);
  std::nth_element(keys.begin() + lo, keys.begin() + hi, keys.end(), less);
  std::sort(keys.begin() + lo, keys.begin() + hi, less);
  for (int64_t i = 0; i < n; i++) order[i] = keys[i].pos;
  return order;
}

}  // namespace

// ---- DomainDivider ----

DomainDivider::DomainDivider(const std::vector<double>& pts, const std::vector<int32_t>& point_idcs,
                             const std::vector<int32_t>& poly_point_idcs)
    : pts_(pts), point_idcs_(point_idcs), poly_idcs_(poly_point_idcs) {
  constexpr double kOverlapQuota = 0.5;   // domain_divider.hpp:26
  constexpr int64_t kMaxLeafSize = 1024;  // domain_divider.hpp:27
  // divide_domain (domain_divider.hpp:160-240): sort along the widest axes and
  // split with an overlap of about half a leaf, until every domain holds at
  // most kMaxLeafSize points. Domains of one depth are independent: few large
  // ones use the parallel sort, many small ones are split in parallel.
  auto divide = [&](const Domain& d, Domain& left, Domain& right, bool parallel_sort) {
    int64_t n = static_cast<int64_t>(d.idx.size());
    const bool par = parallel_sort && n >= (int64_t{1} << 17);
    Bbox3 bb;
    if (par) {
      double lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY, hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY;
#pragma omp parallel for schedule(static) reduction(min : lo0, lo1, lo2) reduction(max : hi0, hi1, hi2)
      for (int64_t k = 0; k < n; k++) {
        const double* p = &pts_[3 * static_cast<size_t>(d.idx[k])];
        lo0 = std::min(lo0, p[0]);
        lo1 = std::min(lo1, p[1]);
        lo2 = std::min(lo2, p[2]);
        hi0 = std::max(hi0, p[0]);
        hi1 = std::max(hi1, p[1]);
        hi2 = std::max(hi2, p[2]);
      }
      bb.lo[0] = lo0, bb.lo[1] = lo1, bb.lo[2] = lo2, bb.hi[0] = hi0, bb.hi[1] = hi1, bb.hi[2] = hi2;
    } else {
      for (int32_t i : d.idx) bb.add(&pts_[3 * static_cast<size_t>(i)]);
    }
    double q = kOverlapQuota * static_cast<double>(kMaxLeafSize) / static_cast<double>(n);
    int64_t n_sub = static_cast<int64_t>(round_half_to_even((1.0 + q) / 2.0 * static_cast<double>(n)));
    int64_t left_part = n - n_sub, right_part = n_sub;
    auto order = order_for_split(d.idx.data(), n, pts_, bb.axes_by_width(), left_part, right_part,
                                 parallel_sort);
    int64_t mid = static_cast<int64_t>(round_half_to_even(static_cast<double>(left_part + right_part) / 2.0));
    left.idx.resize(right_part);
    left.inner.resize(right_part);
#pragma omp parallel for schedule(static) if (par)
    for (int64_t i = 0; i < right_part; i++) {
      int32_t o = order[i];
      left.idx[i] = d.idx[o];
      left.inner[i] = d.inner[o] && i < mid;
    }
    right.idx.resize(n - left_part);
    right.inner.resize(n - left_part);
#pragma omp parallel for schedule(static) if (par)
    for (int64_t i = left_part; i < n; i++) {
      int32_t o = order[i];
      right.idx[i - left_part] = d.idx[o];
      right.inner[i - left_part] = d.inner[o] && i >= mid;
    }
  };
  std::vector<Domain> work, leaves;
  {
    Domain root;
    root.idx = point_idcs;
    root.inner.assign(point_idcs.size(), 1);
    work.push_back(std::move(root));
  }
  while (!work.empty()) {
    const int64_t nw = static_cast<int64_t>(work.size());
    std::vector<Domain> next(static_cast<size_t>(2 * nw));
    std::vector<uint8_t> split(static_cast<size_t>(nw), 0);
    for (int64_t i = 0; i < nw; i++) split[i] = static_cast<int64_t>(work[i].idx.size()) > kMaxLeafSize;
    if (nw == 1) {
      if (split[0]) divide(work[0], next[0], next[1], true);
    } else if (nw < 8 && static_cast<int64_t>(work[0].idx.size()) >= (int64_t{1} << 18)) {
      // the few large domains of the first levels, one after another, each
      // with every thread
      for (int64_t i = 0; i < nw; i++)
        if (split[i]) divide(work[i], next[2 * i], next[2 * i + 1], true);
    } else if (nw < 8) {
      // a few domains, one thread each
#pragma omp parallel for schedule(static, 1) num_threads(static_cast<int>(nw))
      for (int64_t i = 0; i < nw; i++)
        if (split[i]) divide(work[i], next[2 * i], next[2 * i + 1], false);
    } else {
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t i = 0; i < nw; i++)
        if (split[i]) divide(work[i], next[2 * i], next[2 * i + 1], false);
    }
    std::vector<Domain> cont;
    for (int64_t i = 0; i < nw; i++) {
      if (split[i]) {
        cont.push_back(std::move(next[2 * i]));
        cont.push_back(std::move(next[2 * i + 1]));
      } else {
        leaves.push_back(std::move(work[i]));
      }
    }
    work.swap(cont);
  }
  auto& work_out = leaves;
  // merge_poly_points (domain.hpp:30-51)
  domains_.resize(work_out.size());
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t di = 0; di < work_out.size(); di++) {
    Domain& d = work_out[di];
    std::vector<int32_t> order(d.idx.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return d.idx[a] < d.idx[b]; });
    Domain md;
    std::vector<int32_t> sorted_idx;
    std::vector<uint8_t> sorted_inner;
    for (int32_t o : order) {
      sorted_idx.push_back(d.idx[o]);
      sorted_inner.push_back(d.inner[o]);
    }
    md.idx = poly_idcs_;
    md.inner.assign(poly_idcs_.size(), 0);
    for (size_t i = 0; i < poly_idcs_.size(); i++) {
      auto jt = std::lower_bound(sorted_idx.begin(), sorted_idx.end(), poly_idcs_[i]);
      if (jt != sorted_idx.end() && *jt == poly_idcs_[i]) {
        size_t k = static_cast<size_t>(jt - sorted_idx.begin());
        md.inner[i] = sorted_inner[k];
        sorted_idx.erase(jt);
        sorted_inner.erase(sorted_inner.begin() + static_cast<std::ptrdiff_t>(k));
      }
    }
    md.idx.insert(md.idx.end(), sorted_idx.begin(), sorted_idx.end());
    md.inner.insert(md.inner.end(), sorted_inner.begin(), sorted_inner.end());
    domains_[di] = std::move(md);
  }
}

std::vector<int32_t> DomainDivider::choose_coarse_points(int64_t n_coarse) const {
  return pam::choose_coarse_points(pts_, point_idcs_, poly_idcs_, n_coarse);
}

std::vector<int32_t> choose_coarse_points(const std::vector<double>& pts, const std::vector<int32_t>& point_idcs,
                                          const std::vector<int32_t>& poly_idcs, int64_t n_coarse) {
  // domain_divider.hpp:54-125: split the cluster of lowest level (largest bbox
  // volume first) at its median until n_coarse clusters exist; take the point
  // nearest each cluster's bbox centre.
  //
  // The reference pops clusters from a priority queue (lowest level first,
  // then largest volume): every cluster of one level is split before any of
  // the next, so the split runs level by level; at the last level only the
  // largest clusters are split, up to n_coarse clusters in total. Clusters of
  // one level are independent and handled in parallel. All points live in
  // one array that the splits partition in place: a cluster is a range of it,
  // its points' coordinates stored with them so the passes read memory in
  // order.
  struct Rec {
    double x[3];
    int32_t id, pos;
  };
  struct Cluster {
    int64_t b = 0, e = 0;
    double volume = 0.0;
    int32_t center = -1;
    std::array<int, 3> axes{0, 1, 2};
  };
  std::vector<Rec> all;
  all.reserve(point_idcs.size());
  {
    std::vector<int32_t> poly_sorted(poly_idcs);
    std::sort(poly_sorted.begin(), poly_sorted.end());
    for (int32_t i : point_idcs) {
      if (std::binary_search(poly_sorted.begin(), poly_sorted.end(), i)) continue;
      const double* p = &pts[3 * static_cast<size_t>(i)];
      all.push_back(Rec{{p[0], p[1], p[2]}, i, 0});
    }
  }
  // Bbox, volume, axes and the point nearest the bbox centre (the first in
  // the range's order among equally near ones); large ranges in parallel with
  // the same result.
  auto init = [&](Cluster& c) {
    const int64_t n = c.e - c.b;
    const Rec* r = all.data() + c.b;
    Bbox3 bb;
    if (n >= (int64_t{1} << 16)) {
      double lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY, hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY;
#pragma omp parallel for schedule(static) reduction(min : lo0, lo1, lo2) reduction(max : hi0, hi1, hi2)
      for (int64_t i = 0; i < n; i++) {
        lo0 = std::min(lo0, r[i].x[0]);
        lo1 = std::min(lo1, r[i].x[1]);
        lo2 = std::min(lo2, r[i].x[2]);
        hi0 = std::max(hi0, r[i].x[0]);
        hi1 = std::max(hi1, r[i].x[1]);
        hi2 = std::max(hi2, r[i].x[2]);
      }
      bb.lo[0] = lo0, bb.lo[1] = lo1, bb.lo[2] = lo2, bb.hi[0] = hi0, bb.hi[1] = hi1, bb.hi[2] = hi2;
    } else {
      for (int64_t i = 0; i < n; i++) bb.add(r[i].x);
    }
    double cen[3];
    c.volume = 1.0;
    for (int k = 0; k < 3; k++) {
      cen[k] = bb.lo[k] + (bb.hi[k] - bb.lo[k]) / 2.0;
      c.volume *= bb.hi[k] - bb.lo[k];
    }
    auto d2 = [&](const Rec& q) {
      return (q.x[0] - cen[0]) * (q.x[0] - cen[0]) + (q.x[1] - cen[1]) * (q.x[1] - cen[1]) +
             (q.x[2] - cen[2]) * (q.x[2] - cen[2]);
    };
    int64_t best_i = 0;
    if (n >= (int64_t{1} << 16)) {
      double best = INFINITY;
#pragma omp parallel
      {
        double tb = INFINITY;
        int64_t ti = n;
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < n; i++) {
          const double d = d2(r[i]);
          if (d < tb) tb = d, ti = i;
        }
#pragma omp critical
        if (tb < best || (tb == best && ti < best_i)) best = tb, best_i = ti;
      }
    } else {
      double best = INFINITY;
      for (int64_t i = 0; i < n; i++) {
        const double d = d2(r[i]);
        if (d < best) best = d, best_i = i;
      }
    }
    c.center = r[best_i].id;
    c.axes = bb.axes_by_width();
  };
  auto split_mid = [](int64_t sz) {
    // median split (ties broken towards an even index, domain_divider.hpp:82-88):
    // the index i minimising |2 i - sz|, the even one of the two when sz is odd
    if (sz % 2 == 0) return sz / 2;
    return ((sz - 1) / 2) % 2 == 0 ? (sz - 1) / 2 : (sz + 1) / 2;
  };
  std::vector<int32_t> out(poly_idcs);
  if (all.empty()) return out;
  std::vector<Cluster> cur(1);
  cur[0].b = 0;
  cur[0].e = static_cast<int64_t>(all.size());
  init(cur[0]);
  std::vector<int32_t> done;  // centres of clusters that stay unsplit
  int64_t current = 1;
  while (current < n_coarse && !cur.empty()) {
    std::stable_sort(cur.begin(), cur.end(), [](const Cluster& x, const Cluster& y) { return x.volume > y.volume; });
    int64_t splittable = 0;
    for (const auto& c : cur) splittable += c.e - c.b > 1;
    int64_t n_split = std::min<int64_t>(splittable, n_coarse - current);
    std::vector<int64_t> which;
    for (int64_t i = 0; i < static_cast<int64_t>(cur.size()) && static_cast<int64_t>(which.size()) < n_split; i++)
      if (cur[i].e - cur[i].b > 1) which.push_back(i);
    const int64_t nw = static_cast<int64_t>(which.size());
    std::vector<Cluster> next(static_cast<size_t>(2 * nw));
    auto do_split = [&](int64_t w) {
      const Cluster& c = cur[which[w]];
      const int64_t sz = c.e - c.b;
      const int64_t mid = split_mid(sz);
      // The first mid points in the order of coordinates along the cluster's
      // axes (ties by position) form the left half; only the membership of the
      // halves matters, so a selection replaces the full sort.
      Rec* r = all.data() + c.b;
      for (int64_t i = 0; i < sz; i++) r[i].pos = static_cast<int32_t>(i);
      const auto ax = c.axes;
      auto less = [&ax](const Rec& p, const Rec& q) {
        for (int k : ax)
          if (p.x[k] != q.x[k]) return p.x[k] < q.x[k];
        return p.pos < q.pos;
      };
      std::nth_element(r, r + mid, r + sz, less);
      // (the order inside a half only breaks exact ties of the nearest-to-centre search)
      Cluster& l = next[2 * w];
      Cluster& rt = next[2 * w + 1];
      l.b = c.b;
      l.e = rt.b = c.b + mid;
      rt.e = c.e;
      if (l.e > l.b) init(l);
      if (rt.e > rt.b) init(rt);
    };
    if (nw == 1) {
      do_split(0);  // its passes over the points run in parallel
    } else if (nw < 8) {
      // the few large clusters of the first levels, one thread each
#pragma omp parallel for schedule(static, 1) num_threads(static_cast<int>(nw))
      for (int64_t w = 0; w < nw; w++) do_split(w);
    } else {
#pragma omp parallel for schedule(guided)
      for (int64_t w = 0; w < nw; w++) do_split(w);
    }
    std::vector<uint8_t> was_split(cur.size(), 0);
    for (int64_t w : which) was_split[w] = 1;
    for (size_t i = 0; i < cur.size(); i++)
      if (!was_split[i]) done.push_back(cur[i].center);
    std::vector<Cluster> lvl;
    lvl.reserve(next.size());
    for (auto& c : next)
      if (c.e > c.b) lvl.push_back(c);
    current += nw;
    cur.swap(lvl);
    if (nw == 0) break;
  }
  out.insert(out.end(), done.begin(), done.end());
  for (const auto& c : cur) out.push_back(c.center);
  return out;
}

// ---- Division with gradient points ----

namespace {

// A point or a gradient point of a domain or cluster (domain_divider.hpp:127-139).
struct MixedPoint {
  int32_t index;
  uint8_t inner;
  uint8_t grad;
  int mult() const { return grad ? 3 : 1; }
};

struct MixedSet {
  const std::vector<double>& pts;
  const std::vector<double>& gpts;
  const double* at(const MixedPoint& p) const { return (p.grad ? gpts : pts).data() + 3 * static_cast<size_t>(p.index); }
  Bbox3 bbox(const std::vector<MixedPoint>& v) const {
    Bbox3 bb;
    for (const auto& p : v) bb.add(at(p));
    return bb;
  }
  // Sort along the axes (widest first); ties keep their order.
  void sort(std::vector<MixedPoint>& v, const std::array<int, 3>& ax) const {
    std::stable_sort(v.begin(), v.end(), [&](const MixedPoint& a, const MixedPoint& b) {
      const double *p = at(a), *q = at(b);
      for (int k : ax)
        if (p[k] != q[k]) return p[k] < q[k];
      return false;
    });
  }
};

// Largest i with prefix[i] <= x (std::upper_bound(...) - 1 in domain_divider.hpp:199-212).
int64_t prefix_position(const std::vector<int64_t>& prefix, int64_t x) {
  return static_cast<int64_t>(std::upper_bound(prefix.begin(), prefix.end(), x) - prefix.begin()) - 1;
}

}  // namespace

std::vector<Domain> divide_mixed(const std::vector<double>& pts, const std::vector<double>& gpts,
                                 const std::vector<int32_t>& point_idcs, const std::vector<int32_t>& grad_idcs,
                                 const std::vector<int32_t>& poly_idcs) {
  // divide_domain (domain_divider.hpp:160-249) with the multiplicities of
  // gradient points, until every domain holds at most 1024 rows.
  constexpr double kOverlapQuota = 0.5;
  constexpr int64_t kMaxLeafSize = 1024;
  const MixedSet ms{pts, gpts};
  std::vector<std::vector<MixedPoint>> work(1), leaves;
  for (int32_t i : point_idcs) work[0].push_back({i, 1, 0});
  for (int32_t i : grad_idcs) work[0].push_back({i, 1, 1});
  auto rows = [](const std::vector<MixedPoint>& v) {
    int64_t n = 0;
    for (const auto& p : v) n += p.mult();
    return n;
  };
  while (!work.empty()) {
    const int64_t nw = static_cast<int64_t>(work.size());
    std::vector<std::vector<MixedPoint>> next(static_cast<size_t>(2 * nw));
    std::vector<uint8_t> split(static_cast<size_t>(nw));
    for (int64_t w = 0; w < nw; w++) split[w] = rows(work[w]) > kMaxLeafSize;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t w = 0; w < nw; w++) {
      if (!split[w]) continue;
      std::vector<MixedPoint> v = work[w];
      ms.sort(v, ms.bbox(v).axes_by_width());
      std::vector<int64_t> prefix{0};
      for (const auto& p : v) prefix.push_back(prefix.back() + p.mult());
      const int64_t n_mult = prefix.back();
      const double q = kOverlapQuota * static_cast<double>(kMaxLeafSize) / static_cast<double>(n_mult);
      const int64_t n_sub = static_cast<int64_t>(round_half_to_even((1.0 + q) / 2.0 * static_cast<double>(n_mult)));
      const int64_t left_mult = n_mult - n_sub, right_mult = n_sub;
      const int64_t mid_mult =
          static_cast<int64_t>(round_half_to_even(static_cast<double>(left_mult + right_mult) / 2.0));
      const int64_t left = prefix_position(prefix, left_mult);
      const int64_t right = prefix_position(prefix, right_mult);
      const int64_t mid = prefix_position(prefix, mid_mult);
      const int64_t n = static_cast<int64_t>(v.size());
      auto& lo = next[2 * w];
      auto& hi = next[2 * w + 1];
      for (int64_t i = 0; i < right; i++) lo.push_back({v[i].index, static_cast<uint8_t>(v[i].inner && i < mid), v[i].grad});
      for (int64_t i = left; i < n; i++) hi.push_back({v[i].index, static_cast<uint8_t>(v[i].inner && i >= mid), v[i].grad});
    }
    std::vector<std::vector<MixedPoint>> cont;
    for (int64_t w = 0; w < nw; w++) {
      if (split[w]) {
        cont.push_back(std::move(next[2 * w]));
        cont.push_back(std::move(next[2 * w + 1]));
      } else {
        leaves.push_back(std::move(work[w]));
      }
    }
    work.swap(cont);
  }
  // merge_poly_points (domain.hpp:30-51) on the points; the gradient points
  // are kept in index order as well.
  std::vector<Domain> out(leaves.size());
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t di = 0; di < leaves.size(); di++) {
    std::vector<std::pair<int32_t, uint8_t>> vp, gp;
    for (const auto& p : leaves[di]) (p.grad ? gp : vp).push_back({p.index, p.inner});
    std::sort(vp.begin(), vp.end());
    std::sort(gp.begin(), gp.end());
    Domain& d = out[di];
    d.idx = poly_idcs;
    d.inner.assign(poly_idcs.size(), 0);
    for (size_t a = 0; a < poly_idcs.size(); a++) {
      auto it = std::lower_bound(vp.begin(), vp.end(), std::make_pair(poly_idcs[a], uint8_t{0}));
      if (it != vp.end() && it->first == poly_idcs[a]) {
        d.inner[a] = it->second;
        vp.erase(it);
      }
    }
    for (const auto& p : vp) {
      d.idx.push_back(p.first);
      d.inner.push_back(p.second);
    }
    for (const auto& p : gp) {
      d.gidx.push_back(p.first);
      d.ginner.push_back(p.second);
    }
  }
  return out;
}

std::pair<std::vector<int32_t>, std::vector<int32_t>> choose_coarse_points_mixed(
    const std::vector<double>& pts, const std::vector<double>& gpts, const std::vector<int32_t>& point_idcs,
    const std::vector<int32_t>& grad_idcs, const std::vector<int32_t>& poly_idcs, int64_t n_coarse) {
  // domain_divider.hpp:54-125: split the cluster of lowest level (largest
  // bbox volume first) at the middle of its rows until the clusters' centres
  // count n_coarse rows; the centre of a cluster is its point (or gradient
  // point) nearest its bbox centre.
  const MixedSet ms{pts, gpts};
  struct Cluster {
    std::vector<MixedPoint> p;
    int level = 0;
    double volume = 0.0;
    int64_t seq = 0;
    MixedPoint center{};
  };
  int64_t seq = 0;
  auto init = [&](Cluster& c) {
    const Bbox3 bb = ms.bbox(c.p);
    double cen[3];
    c.volume = 1.0;
    for (int k = 0; k < 3; k++) {
      cen[k] = bb.lo[k] + (bb.hi[k] - bb.lo[k]) / 2.0;
      c.volume *= bb.hi[k] - bb.lo[k];
    }
    double best = INFINITY;
    for (const auto& p : c.p) {
      const double* x = ms.at(p);
      const double d = (x[0] - cen[0]) * (x[0] - cen[0]) + (x[1] - cen[1]) * (x[1] - cen[1]) +
                       (x[2] - cen[2]) * (x[2] - cen[2]);
      if (d < best) best = d, c.center = p;
    }
    ms.sort(c.p, bb.axes_by_width());
    c.seq = seq++;
  };
  auto lower = [](const Cluster& a, const Cluster& b) {  // a pops after b
    if (a.level != b.level) return a.level > b.level;
    if (a.volume != b.volume) return a.volume < b.volume;
    return a.seq > b.seq;
  };
  std::priority_queue<Cluster, std::vector<Cluster>, decltype(lower)> q(lower);
  std::vector<int32_t> sorted_poly(poly_idcs);
  std::sort(sorted_poly.begin(), sorted_poly.end());
  Cluster root;
  for (int32_t i : point_idcs)
    if (!std::binary_search(sorted_poly.begin(), sorted_poly.end(), i)) root.p.push_back({i, 1, 0});
  for (int32_t i : grad_idcs) root.p.push_back({i, 1, 1});
  std::vector<int32_t> idcs(poly_idcs), gidcs;
  if (root.p.empty()) return {idcs, gidcs};
  int64_t total = 0;  // all clusters single points: their centres count every row
  for (const auto& p : root.p) total += p.mult();
  n_coarse = std::min(n_coarse, total);
  init(root);
  int64_t current = root.center.mult();
  q.push(std::move(root));
  while (current < n_coarse) {
    Cluster c = q.top();
    q.pop();
    std::vector<int64_t> prefix{0};
    for (const auto& p : c.p) prefix.push_back(prefix.back() + p.mult());
    const int64_t size = prefix.back();
    int64_t mid = 0, bestd = INT64_MAX;
    for (int64_t i = 0; i < static_cast<int64_t>(c.p.size()); i++) {
      const int64_t dd = std::llabs(2 * prefix[i] - size);
      if (dd < bestd || (dd == bestd && i % 2 == 0)) bestd = dd, mid = i;
    }
    current -= c.center.mult();
    Cluster l, r;
    l.p.assign(c.p.begin(), c.p.begin() + mid);
    r.p.assign(c.p.begin() + mid, c.p.end());
    l.level = r.level = c.level + 1;
    if (!l.p.empty()) {
      init(l);
      current += l.center.mult();
      q.push(std::move(l));
    }
    if (!r.p.empty()) {
      init(r);
      current += r.center.mult();
      q.push(std::move(r));
    }
    if (q.empty()) break;
  }
  while (!q.empty()) {
    const MixedPoint& p = q.top().center;
    (p.grad ? gidcs : idcs).push_back(p.index);
    q.pop();
  }
  return {idcs, gidcs};
}

// ---- RasPreconditioner ----

RasPreconditioner::RasPreconditioner(const Model& model, const std::vector<double>& points,
                                     hipStream_t s, const RasOptions& opt, const double* d_points)
    : RasPreconditioner(model, points, std::vector<double>{}, s, opt, d_points) {}

RasPreconditioner::RasPreconditioner(const Model& model, const std::vector<double>& points,
                                     const std::vector<double>& grad_points, hipStream_t s, const RasOptions& opt,
                                     const double* d_points)
    : model_(model),
      s_(s),
      opt_(opt),
      mu_(static_cast<int64_t>(points.size() / 3)),
      sigma_(static_cast<int64_t>(grad_points.size() / 3)),
      n_(mu_ + 3 * sigma_),
      l_(model.l()) {
  auto t0 = std::chrono::steady_clock::now();
  model_.validate();
  constexpr double kFineToCoarseRatio = 10.0;  // ras_preconditioner.hpp:52
  const int64_t kNCoarsestPoints = opt_.n_coarsest;  // 2048 in ras_preconditioner.hpp:53
  if (kNCoarsestPoints < 1 || kNCoarsestPoints > RasOptions::kMaxCoarsest)
    throw_invalid("the coarse grid size must be in [1, " + std::to_string(RasOptions::kMaxCoarsest) + "]");
  h_pts_ = points;
  h_gpts_ = grad_points;
  if (d_points) {
    d_pts_.resize(h_pts_.size());
    dev_copy(d_points, d_pts_.get(), static_cast<int64_t>(h_pts_.size()), s_);
  } else {
    d_pts_.upload(h_pts_.data(), h_pts_.size(), s_);
  }
  d_gpts_.upload(h_gpts_.data(), h_gpts_.size(), s_);
  n_levels_ = std::max(static_cast<int>(std::ceil(std::log(static_cast<double>(n_) / kNCoarsestPoints) /
                                                  std::log(kFineToCoarseRatio))), 0) + 1;
  point_idcs_.resize(n_levels_);
  gpoint_idcs_.resize(n_levels_);
  std::vector<int32_t> poly_idcs;
  if (l_ > 0) {
    // The polynomial rows: a unisolvent set of points; in the special case of
    // one point with gradient points and degree 1, that point and the first
    // gradient point (ras_preconditioner.hpp:81-92). Degree-1 Lagrange
    // polynomials have constant gradients, so any gradient point of a domain
    // serves as its top rows.
    const bool special = model_.poly_degree == 1 && mu_ == 1 && sigma_ >= 1;
    std::vector<double> p(static_cast<size_t>(l_) * l_), inv;
    if (special) {
      poly_idcs = {0};
      double m[10];
      host_monomials(1, &h_pts_[0], m);
      for (int k = 0; k < l_; k++) p[k] = m[k];
      const double gm[3][4] = {{0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
      for (int c = 0; c < 3; c++)
        for (int k = 0; k < l_; k++) p[(1 + c) * l_ + k] = gm[c][k];
    } else {
      if (mu_ < l_) throw_invalid("points.rows() must be greater than or equal to " + std::to_string(l_));
      poly_idcs = unisolvent_point_set(h_pts_, model_.poly_degree);
      for (int i = 0; i < l_; i++) {
        double m[10];
        host_monomials(model_.poly_degree, &h_pts_[3 * static_cast<size_t>(poly_idcs[i])], m);
        for (int k = 0; k < l_; k++) p[i * l_ + k] = m[k];
      }
    }
    if (invert(p, l_, inv) == 0.0) throw_invalid("the set of points is not unisolvent");
    pinv_.upload(inv.data(), inv.size(), s_);
    lag_.resize(static_cast<size_t>(n_) * l_);
    if (mu_ > 0) {
      lagrange_kernel<<<ceil_div(mu_, kT), kT, 0, s_>>>(model_.poly_degree, l_, d_pts_.get(), mu_, pinv_.get(),
                                                        lag_.get());
      PAM_LAUNCH_CHECK();
    }
    if (sigma_ > 0) {
      lagrange_grad_kernel<<<ceil_div(sigma_, kT), kT, 0, s_>>>(model_.poly_degree, l_, d_gpts_.get(), sigma_,
                                                                pinv_.get(), lag_.get() + mu_ * l_);
      PAM_LAUNCH_CHECK();
    }
    auto& top = point_idcs_[n_levels_ - 1];
    top = poly_idcs;
    std::vector<uint8_t> is_poly(mu_, 0);
    for (int32_t i : poly_idcs) is_poly[i] = 1;
    for (int64_t i = 0; i < mu_; i++)
      if (!is_poly[i]) top.push_back(static_cast<int32_t>(i));
  } else {
    auto& top = point_idcs_[n_levels_ - 1];
    top.resize(mu_);
    std::iota(top.begin(), top.end(), 0);
    lag_.resize(1);
    pinv_.resize(1);
  }
  gpoint_idcs_[n_levels_ - 1].resize(sigma_);
  std::iota(gpoint_idcs_[n_levels_ - 1].begin(), gpoint_idcs_[n_levels_ - 1].end(), 0);
  // Domain geometry in the anisotropic frame when there is a single rbf
  // (ras_preconditioner.hpp:105-113).
  const bool transform = model_.rbfs.size() == 1 && !is_identity(model_.rbfs[0].aniso);
  std::vector<double> a_pts_t, a_gpts_t;
  if (transform) {
    a_pts_t = h_pts_;
    a_gpts_t = h_gpts_;
  }
  std::vector<double>& a_pts = transform ? a_pts_t : h_pts_;
  std::vector<double>& a_gpts = transform ? a_gpts_t : h_gpts_;
  if (transform) {
    const double* a = model_.rbfs[0].aniso;
    for (int64_t i = 0; i < mu_; i++) {
      const double* p = &h_pts_[3 * i];
      for (int k = 0; k < 3; k++) a_pts[3 * i + k] = a[3 * k] * p[0] + a[3 * k + 1] * p[1] + a[3 * k + 2] * p[2];
    }
    for (int64_t i = 0; i < sigma_; i++) {
      const double* p = &h_gpts_[3 * i];
      for (int k = 0; k < 3; k++) a_gpts[3 * i + k] = a[3 * k] * p[0] + a[3 * k + 1] * p[1] + a[3 * k + 2] * p[2];
    }
  }
  // A P depends on the points only: it is computed on a host thread and a
  // stream of its own (ras_preconditioner.hpp:173-188) while this thread
  // divides the levels and the device factors them.
  std::future<void> ap_done;
  if (n_levels_ > 1 && l_ > 0) {
    p_orth_.resize(static_cast<size_t>(n_) * l_);
    PAM_HIP(hipStreamSynchronize(s_));  // the points are on the device
    int dev = 0;
    PAM_HIP(hipGetDevice(&dev));
    ap_done = std::async(std::launch::async, [this, dev] {
      PAM_HIP(hipSetDevice(dev));
      hipStream_t as = stream_acquire();
      try {
        // P, orthonormalized (ras_preconditioner.hpp:168-171), then A P
        Workspace1 ws;
        if (sigma_ == 0) dev_poly_matrix(model_.poly_degree, d_pts_.get(), mu_, p_orth_.get(), as);
        else dev_poly_matrix_hermite(model_.poly_degree, d_pts_.get(), mu_, d_gpts_.get(), sigma_, p_orth_.get(), as);
        dev_orthonormalize_cols(p_orth_.get(), n_, l_, ws, as);
        compute_ap(as);
      } catch (...) {
        stream_release(as);
        throw;
      }
      stream_release(as);
    });
  }
  // one positive-definiteness flag per level (0 = the coarse grid)
  flag_.resize(static_cast<size_t>(n_levels_));
  flag_.zero(s_);
  fine_.clear();
  fine_.resize(n_levels_);
  h_domains_.assign(n_levels_, {});
  auto tick = [&](const char* what) {
    if (!opt_.verbose) return;
    double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  ras setup %-28s %8.3f s\n", what, t);
    std::fflush(stdout);
  };
  tick("unisolvent set, transforms");
  // Each level's batch is launched on its own stream as soon as its domains
  // exist; the host then picks the coarser level's points and divides it
  // while the device assembles and factors.
  for (int level = n_levels_ - 1; level >= 1; level--) {
    double finest = std::log(static_cast<double>(n_)) / std::log(kFineToCoarseRatio);
    double coarsest = std::log(static_cast<double>(kNCoarsestPoints)) / std::log(kFineToCoarseRatio);
    auto n_coarse = static_cast<int64_t>(std::pow(
        kFineToCoarseRatio, coarsest + (level - 1) * (finest - coarsest) / (n_levels_ - 1)));
    // The coarser level's points depend on this level's points only, not on
    // its domains: they are chosen on a second host thread while this one
    // divides the level and launches its batch.
    const std::vector<int32_t>& lp = point_idcs_[level];
    const std::vector<int32_t>& lg = gpoint_idcs_[level];
    if (sigma_ > 0) {
      // with gradient points (the sizes of fits with gradient data are small:
      // one host thread)
      auto cp = choose_coarse_points_mixed(a_pts, a_gpts, lp, lg, poly_idcs, n_coarse);
      std::vector<Domain> doms = divide_mixed(a_pts, a_gpts, lp, lg, poly_idcs);
      tick("divide");
      setup_batch(fine_[level], doms, false, fork_stream(), opt_.fp32_factors, flag_.get() + level);
      point_idcs_[level - 1] = std::move(cp.first);
      gpoint_idcs_[level - 1] = std::move(cp.second);
      h_domains_[level] = std::move(doms);
      continue;
    }
    auto coarse = std::async(std::launch::async,
                             [&] { return choose_coarse_points(a_pts, lp, poly_idcs, n_coarse); });
    std::vector<Domain> doms;
    try {
      DomainDivider div(a_pts, lp, poly_idcs);
      doms = std::move(div.domains());
    } catch (...) {
      coarse.wait();
      throw;
    }
    tick("divide");
    setup_batch(fine_[level], doms, false, fork_stream(), opt_.fp32_factors, flag_.get() + level);
    tick("batch launched");
    point_idcs_[level - 1] = coarse.get();
    tick("coarse points");
    h_domains_[level] = std::move(doms);
  }
  {
    Domain cd;
    cd.idx = point_idcs_[0];
    cd.inner.assign(cd.idx.size(), 1);
    cd.gidx = gpoint_idcs_[0];
    cd.ginner.assign(cd.gidx.size(), 1);
    h_domains_[0] = {cd};
    hipStream_t cs = fork_stream();
    setup_batch(coarse_, h_domains_[0], true, cs, false, flag_.get());
    // The coarse solve is one domain: solved by a single workgroup it would
    // serialize on one CU, so it is turned into a dense operator once (one
    // workgroup per unit vector) and applied as a matrix-vector product.
    const int m = static_cast<int>(coarse_.h_ptr[1] - coarse_.h_ptr[0]);
    const size_t shmem = sizeof(double) * 2 * static_cast<size_t>(m);
    if (coarse_.n_dom == 1 && m > 0 && m <= kCoarseOpMax && shmem <= 96 * 1024) {
      coarse_.op.resize(static_cast<size_t>(m + l_) * m);
      RbfSet rs = make_rbf_set(model_);
      coarse_operator_kernel<<<m, kT, shmem, cs>>>(rs, l_, d_pts_.get(), d_gpts_.get(), mu_, coarse_.idx.get(), m,
                                                   lag_.get(), coarse_.mat.get(), pinv_.get(), coarse_.op.get());
      PAM_LAUNCH_CHECK();
      coarse_.has_op = true;
    }
  }
  join_streams();
  std::vector<int32_t> flags(static_cast<size_t>(n_levels_), 0);
  PAM_HIP(hipMemcpyAsync(flags.data(), flag_.get(), sizeof(int32_t) * flags.size(), hipMemcpyDeviceToHost, s_));
  PAM_HIP(hipStreamSynchronize(s_));
  setup_tmp_.clear();
  // A single-precision level that met a non-positive pivot is factored again in FP64.
  bool refactored = false;
  for (int level = 1; level < n_levels_; level++) {
    if (!flags[level] || !fine_[level].fp32) continue;
    if (opt_.verbose) std::printf("  ras setup level %d: FP32 factorisation failed, refactoring in FP64\n", level);
    flag_.zero(s_);
    setup_batch(fine_[level], h_domains_[level], false, s_, false, flag_.get() + level);
    int32_t f = 0;
    PAM_HIP(hipMemcpyAsync(&f, flag_.get() + level, sizeof f, hipMemcpyDeviceToHost, s_));
    PAM_HIP(hipStreamSynchronize(s_));
    flags[level] = f;
    refactored = true;
  }
  if (refactored) setup_tmp_.clear();
  for (hipStream_t ls : setup_streams_) stream_release(ls);
  setup_streams_.clear();
  tick("factors done");
  for (int32_t f : flags)
    if (f) throw_runtime("a subdomain matrix Q^T A Q is not positive definite");
  d_point_idcs_.resize(n_levels_);
  d_gpoint_idcs_.resize(n_levels_);
  d_level_rows_.resize(n_levels_);
  level_rows_.assign(n_levels_, 0);
  for (int level = 0; level < n_levels_; level++) {
    d_point_idcs_[level].upload(point_idcs_[level].data(), point_idcs_[level].size(), s_);
    d_gpoint_idcs_[level].upload(gpoint_idcs_[level].data(), gpoint_idcs_[level].size(), s_);
    std::vector<int32_t> rows(point_idcs_[level]);
    for (int32_t k : gpoint_idcs_[level])
      for (int c = 0; c < 3; c++) rows.push_back(static_cast<int32_t>(mu_ + 3 * static_cast<int64_t>(k) + c));
    level_rows_[level] = static_cast<int64_t>(rows.size());
    d_level_rows_[level].upload(rows.data(), rows.size(), s_);
    PAM_HIP(hipStreamSynchronize(s_));  // rows is a host temporary
  }
  res_.resize(n_);
  w_.resize(n_ + l_);
  total_.resize(n_ + l_);
  dot_.resize(10);
  if (ap_done.valid()) ap_done.get();
  PAM_HIP(hipStreamSynchronize(s_));
  tick("A P done");
  if (opt_.tune_solves) {
    tune_solves();
    tick("solve tuning");
  } else {
    // Few large batches: more threads per domain; many domains: 256 (the
    // timing on the 1M-point fit: 512 for 32 and 256 domains, 256 for 2048).
    for (auto& b : fine_) b.solve_threads = b.n_dom < 1024 ? 512 : 256;
    coarse_.solve_threads = 512;
  }
  setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

hipStream_t RasPreconditioner::fork_stream() {
  hipStream_t ls = stream_acquire();
  setup_streams_.push_back(ls);
  hipEvent_t ev = nullptr;
  PAM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  PAM_HIP(hipEventRecord(ev, s_));
  PAM_HIP(hipStreamWaitEvent(ls, ev, 0));
  PAM_HIP(hipEventDestroy(ev));
  return ls;
}

void RasPreconditioner::join_streams() {
  for (hipStream_t ls : setup_streams_) {
    hipEvent_t ev = nullptr;
    PAM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    PAM_HIP(hipEventRecord(ev, ls));
    PAM_HIP(hipStreamWaitEvent(s_, ev, 0));
    PAM_HIP(hipEventDestroy(ev));
  }
}

void RasPreconditioner::setup_batch(DomainBatch& b, const std::vector<Domain>& doms, bool coarse, hipStream_t ls,
                                    bool fp32, int32_t* flag) {
  (void)coarse;
  b.n_dom = static_cast<int>(doms.size());
  b.h_ptr.assign(b.n_dom + 1, 0);
  std::vector<int64_t> mptr(b.n_dom + 1, 0);
  b.max_m = 0;
  for (int d = 0; d < b.n_dom; d++) {
    int64_t m = static_cast<int64_t>(doms[d].idx.size() + 3 * doms[d].gidx.size());
    if (l_ > 0 && m < l_) throw_runtime("a domain holds fewer points than the polynomial basis");
    b.h_ptr[d + 1] = b.h_ptr[d] + m;
    int64_t n = m - l_;
    mptr[d + 1] = mptr[d] + n * n;
    b.max_m = std::max<int>(b.max_m, static_cast<int>(m));
  }
  auto tmp = std::make_unique<SetupTmp>();
  tmp->mptr = std::move(mptr);
  tmp->idx.resize(b.h_ptr.back());
  tmp->inner.resize(b.h_ptr.back());
  for (int d = 0; d < b.n_dom; d++) {
    std::copy(doms[d].idx.begin(), doms[d].idx.end(), tmp->idx.begin() + b.h_ptr[d]);
    std::copy(doms[d].inner.begin(), doms[d].inner.end(), tmp->inner.begin() + b.h_ptr[d]);
    int64_t r = b.h_ptr[d] + static_cast<int64_t>(doms[d].idx.size());
    for (size_t g = 0; g < doms[d].gidx.size(); g++)
      for (int c = 0; c < 3; c++, r++) {
        tmp->idx[r] = static_cast<int32_t>(mu_ + 3 * static_cast<int64_t>(doms[d].gidx[g]) + c);
        tmp->inner[r] = doms[d].ginner[g];
      }
  }
  b.idx.upload(tmp->idx.data(), tmp->idx.size(), ls);
  b.inner.upload(tmp->inner.data(), tmp->inner.size(), ls);
  b.ptr.upload(b.h_ptr.data(), b.h_ptr.size(), ls);
  b.mat_ptr.upload(tmp->mptr.data(), tmp->mptr.size(), ls);
  b.fp32 = fp32;
  if (fp32) {
    b.mat32.resize(static_cast<size_t>(std::max<int64_t>(tmp->mptr.back(), 1)));
    b.mat.release();
  } else {
    b.mat.resize(static_cast<size_t>(std::max<int64_t>(tmp->mptr.back(), 1)));
    b.mat32.release();
  }
  RbfSet rs = make_rbf_set(model_);
  if (sigma_ > 0) {
    // rows of mixed type: fields with the row type, entries by kernel_hb
    const int max_n = b.max_m - l_;
    tmp->rows.resize(static_cast<size_t>(
        std::max<int64_t>((4 + 2 * l_) * (b.h_ptr.back() - static_cast<int64_t>(b.n_dom) * l_), 1)));
    if (max_n > 0) {
      assemble_hb_rows_kernel<<<dim3(static_cast<unsigned>(ceil_div(max_n, kT)), b.n_dom), kT, 0, ls>>>(
          rs, l_, d_pts_.get(), d_gpts_.get(), mu_, b.ptr.get(), b.idx.get(), lag_.get(), tmp->rows.get());
      PAM_LAUNCH_CHECK();
    }
    if (fp32) launch_assemble_m_hb<float>(l_, max_n, b.n_dom, rs, b.ptr.get(), tmp->rows.get(), b.mat_ptr.get(),
                                          b.mat32.get(), ls);
    else launch_assemble_m_hb<double>(l_, max_n, b.n_dom, rs, b.ptr.get(), tmp->rows.get(), b.mat_ptr.get(),
                                      b.mat.get(), ls);
  } else {
    tmp->u.resize(static_cast<size_t>(
        std::max<int64_t>(l_ * (b.h_ptr.back() - static_cast<int64_t>(b.n_dom) * l_), 1)));
    tmp->atl.resize(static_cast<size_t>(std::max(b.n_dom * l_ * l_, 1)));
    if (l_ > 0) {
      assemble_u_kernel<<<dim3(16, b.n_dom), kT, 0, ls>>>(rs, l_, d_pts_.get(), b.ptr.get(), b.idx.get(),
                                                          tmp->u.get(), tmp->atl.get());
      PAM_LAUNCH_CHECK();
    }
    tmp->rows.resize(static_cast<size_t>(
        std::max<int64_t>((3 + 2 * l_) * (b.h_ptr.back() - static_cast<int64_t>(b.n_dom) * l_), 1)));
    const int max_n = b.max_m - l_;
    if (max_n > 0) {
      assemble_rows_kernel<<<dim3(static_cast<unsigned>(ceil_div(max_n, kT)), b.n_dom), kT, 0, ls>>>(
          l_, d_pts_.get(), b.ptr.get(), b.idx.get(), lag_.get(), tmp->u.get(), tmp->atl.get(), tmp->rows.get());
      PAM_LAUNCH_CHECK();
    }
    if (fp32) launch_assemble_m<float>(l_, max_n, b.n_dom, rs, b.ptr.get(), tmp->rows.get(), b.mat_ptr.get(),
                                       b.mat32.get(), ls);
    else launch_assemble_m<double>(l_, max_n, b.n_dom, rs, b.ptr.get(), tmp->rows.get(), b.mat_ptr.get(),
                                   b.mat.get(), ls);
  }
  if (fp32) {
    launch_cholesky_batch<float>(factor_panel(b, true, ls, tmp->mptr), b.n_dom, b.ptr.get(), l_, b.mat_ptr.get(),
                                 b.mat32.get(), flag, ls);
    setup_tmp_.push_back(std::move(tmp));
    return;
  }
  const int n0 = static_cast<int>(b.h_ptr.back() - l_);
  if (b.n_dom == 1 && n0 > 4 * kTile) {
    // One large matrix (the coarse grid): many workgroups per panel step
    // instead of one workgroup for the whole factorisation.
    tmp->inv.resize(static_cast<size_t>(kNb) * kNb);
    for (int k0 = 0; k0 < n0; k0 += kNb) {
      cholesky_panel_kernel<<<1, kT, 0, ls>>>(b.mat.get(), n0, k0, flag, tmp->inv.get());
      PAM_LAUNCH_CHECK();
      const int rem = n0 - std::min(kNb, n0 - k0) - k0;
      if (rem <= 0) break;
      cholesky_rows_kernel<<<ceil_div(rem, kT), kT, 0, ls>>>(b.mat.get(), n0, k0, tmp->inv.get());
      PAM_LAUNCH_CHECK();
      const int nt = (rem + kTile - 1) / kTile;
      cholesky_trailing_kernel<<<nt * (nt + 1) / 2, kT, 0, ls>>>(b.mat.get(), n0, k0);
      PAM_LAUNCH_CHECK();
    }
  } else {
    launch_cholesky_batch<double>(coarse ? kNb : factor_panel(b, false, ls, tmp->mptr), b.n_dom, b.ptr.get(), l_,
                                  b.mat_ptr.get(), b.mat.get(), flag, ls);
  }
  // b.h_ptr is a member and outlives the upload; the rest stays until the
  // constructor synchronizes.
  setup_tmp_.push_back(std::move(tmp));
}

int RasPreconditioner::factor_panel(const DomainBatch& b, bool fp32, hipStream_t ls,
                                    const std::vector<int64_t>& mptr) {
  if (opt_.factor_panel >= 0) return opt_.factor_panel;
  // Without tuning, the rule that the timing below found on MI355X for the
  // reference's subdomain size (about 1,000 points, 1,500 with the overlap):
  // the left-looking form, 3.7x faster than the right-looking one in FP32
  // (3.6 against 13.5 ms for 32 domains of the 1M-point ellipsoid fit) and
  // 3.3x in FP64 (4.7 against 15.3 ms).
  if (!opt_.tune_factor) return 0;
  static std::mutex mu;
  static std::map<std::pair<bool, int>, int> chosen;  // (fp32, size class) -> panel width
  const std::pair<bool, int> key{fp32, b.max_m / 256};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = chosen.find(key);
    if (it != chosen.end()) return it->second;
  }
  // Time each width on copies of the first domains' assembled matrices.
  const int k = std::min(b.n_dom, 32);
  const int64_t elems = mptr[static_cast<size_t>(k)];
  const size_t es = fp32 ? sizeof(float) : sizeof(double);
  DevBuf<char> orig(static_cast<size_t>(elems) * es), work(static_cast<size_t>(elems) * es);
  DevBuf<int32_t> dummy(1);
  dummy.zero(ls);
  const void* src = fp32 ? static_cast<const void*>(b.mat32.get()) : static_cast<const void*>(b.mat.get());
  PAM_HIP(hipMemcpyAsync(orig.get(), src, static_cast<size_t>(elems) * es, hipMemcpyDeviceToDevice, ls));
  hipEvent_t e0, e1;
  PAM_HIP(hipEventCreate(&e0));
  PAM_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = INFINITY;
  for (int nb : {0, 32, 64}) {
    float ms = 0.0f;
    for (int rep = 0; rep < 2; rep++) {  // the first run warms the code object and caches
      PAM_HIP(hipMemcpyAsync(work.get(), orig.get(), static_cast<size_t>(elems) * es, hipMemcpyDeviceToDevice, ls));
      PAM_HIP(hipEventRecord(e0, ls));
      if (fp32) launch_cholesky_batch<float>(nb, k, b.ptr.get(), l_, b.mat_ptr.get(),
                                            reinterpret_cast<float*>(work.get()), dummy.get(), ls);
      else launch_cholesky_batch<double>(nb, k, b.ptr.get(), l_, b.mat_ptr.get(),
                                        reinterpret_cast<double*>(work.get()), dummy.get(), ls);
      PAM_HIP(hipEventRecord(e1, ls));
      PAM_HIP(hipEventSynchronize(e1));
      PAM_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    if (opt_.verbose) std::printf("  ras factor tuning (%s): %s %d, %d domains, %.3f ms\n", fp32 ? "FP32" : "FP64",
                                  nb ? "right-looking, panel" : "left-looking, tile", nb ? nb : kTile, k, ms);
    if (ms < best_ms) best_ms = ms, best = nb;
  }
  PAM_HIP(hipEventDestroy(e0));
  PAM_HIP(hipEventDestroy(e1));
  std::lock_guard<std::mutex> lk(mu);
  chosen[key] = best;
  return best;
}

int64_t RasPreconditioner::factor_bytes() const {
  int64_t b = static_cast<int64_t>(coarse_.mat.bytes());
  for (const auto& f : fine_) b += static_cast<int64_t>(f.mat.bytes() + f.mat32.bytes());
  return b;
}

void RasPreconditioner::solve(int level, const double* d_res, double* d_w) {
  // ras_preconditioner.hpp:280-297
  dev_fill(0.0, d_w, n_ + l_, s_);
  DomainBatch& b = level == 0 ? coarse_ : fine_[level];
  if (b.n_dom == 0) return;
  if (level == 0 && b.has_op) {
    const int m = static_cast<int>(b.h_ptr[1] - b.h_ptr[0]);
    coarse_apply_kernel<<<ceil_div(m + l_, kCoarseRowsPerBlock), kT, sizeof(double) * m, s_>>>(
        b.idx.get(), m, l_, b.op.get(), d_res, d_w, n_);
    PAM_LAUNCH_CHECK();
    return;
  }
  launch_solve(b, level, b.solve_threads > 0 ? b.solve_threads : kT, d_res, d_w);
}

void RasPreconditioner::launch_solve(const DomainBatch& b, int level, int threads, const double* d_res,
                                     double* d_w) {
  size_t shmem = sizeof(double) * 2 * static_cast<size_t>(b.max_m);
  if (shmem > 96 * 1024) throw_runtime("subdomain too large for the solve kernel");
  RbfSet rs = make_rbf_set(model_);
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    if (b.fp32) {
      ras_solve_kernel<KT, float><<<b.n_dom, KT, shmem, s_>>>(rs, l_, d_pts_.get(), d_gpts_.get(), mu_, b.ptr.get(),
                                                              b.idx.get(), b.inner.get(), lag_.get(), b.mat_ptr.get(),
                                                              b.mat32.get(), d_res, d_w, n_, level == 0 ? 1 : 0,
                                                              pinv_.get());
    } else {
      ras_solve_kernel<KT, double><<<b.n_dom, KT, shmem, s_>>>(rs, l_, d_pts_.get(), d_gpts_.get(), mu_, b.ptr.get(),
                                                               b.idx.get(), b.inner.get(), lag_.get(), b.mat_ptr.get(),
                                                               b.mat.get(), d_res, d_w, n_, level == 0 ? 1 : 0,
                                                               pinv_.get());
    }
  };
  switch (threads) {
    case 128: go(std::integral_constant<int, 128>{}); break;
    case 256: go(std::integral_constant<int, 256>{}); break;
    case 512: go(std::integral_constant<int, 512>{}); break;
    default: throw_invalid("unsupported solve workgroup size");
  }
  PAM_LAUNCH_CHECK();
}

void RasPreconditioner::tune_solves() {
  // Workgroup size of each level's batched solve, timed on the batch itself
  // (a zero right-hand side: the work does not depend on the values). Every
  // size gives the same bits, so the choice only changes the time.
  DevBuf<double> res(static_cast<size_t>(n_ + l_)), w(static_cast<size_t>(n_ + l_));
  dev_fill(0.0, res.get(), n_ + l_, s_);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PAM_HIP(hipEventCreate(&e0));
  PAM_HIP(hipEventCreate(&e1));
  for (int level = 0; level < n_levels_; level++) {
    DomainBatch& b = level == 0 ? coarse_ : fine_[level];
    if (b.n_dom == 0 || (level == 0 && b.has_op)) continue;
    float best = 0.0f;
    for (int threads : {128, 256, 512}) {
      launch_solve(b, level, threads, res.get(), w.get());  // warm-up
      PAM_HIP(hipEventRecord(e0, s_));
      for (int rep = 0; rep < 2; rep++) launch_solve(b, level, threads, res.get(), w.get());
      PAM_HIP(hipEventRecord(e1, s_));
      PAM_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      PAM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (b.solve_threads == 0 || ms < best) {
        best = ms;
        b.solve_threads = threads;
      }
    }
    if (opt_.verbose)
      std::printf("RAS level %d: %d domains, solve workgroup %d threads (%.3f ms)\n", level, b.n_dom,
                  b.solve_threads, best / 2.0f);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void RasPreconditioner::compute_ap(hipStream_t s) {
  // ap = A P (ras_preconditioner.hpp:173-188), with the level-update configuration
  ap_.resize(static_cast<size_t>(n_) * l_);
  if (sigma_ > 0) {
    HermiteEvaluator ev(model_, hermite_level_cfg(), s);
    ev.set_sources(d_pts_.get(), mu_, d_gpts_.get(), sigma_);
    ev.set_targets(d_pts_.get(), mu_, d_gpts_.get(), sigma_);
    DevBuf<double> wcol(static_cast<size_t>(n_ + l_));
    for (int k = 0; k < l_; k++) {
      dev_fill(0.0, wcol.get(), n_ + l_, s);
      dev_copy(p_orth_.get() + static_cast<int64_t>(k) * n_, wcol.get(), n_, s);
      ev.set_weights(wcol.get());
      ev.evaluate(ap_.get() + static_cast<int64_t>(k) * n_);
    }
    PAM_HIP(hipStreamSynchronize(s));
    return;
  }
  Bbox3 bb;
  for (int64_t i = 0; i < mu_; i++) bb.add(&h_pts_[3 * i]);
  SymmetricEvaluator ev(model_, bb.lo, bb.hi, INFINITY, s);
  if (opt_.use_level_cfg) ev.set_fmm_config(opt_.level_cfg);
  ev.set_points(d_pts_.get(), mu_);
  DevBuf<double> wcol(static_cast<size_t>(mu_ + l_));
  for (int k = 0; k < l_; k++) {
    dev_fill(0.0, wcol.get(), mu_ + l_, s);
    dev_copy(p_orth_.get() + static_cast<int64_t>(k) * mu_, wcol.get(), mu_, s);
    ev.set_weights(wcol.get());
    ev.evaluate(ap_.get() + static_cast<int64_t>(k) * mu_);
  }
  PAM_HIP(hipStreamSynchronize(s));
}

void RasPreconditioner::set_level_config(const FmmConfig& cfg) {
  PAM_HIP(hipStreamSynchronize(s_));
  opt_.level_cfg = cfg;
  opt_.use_level_cfg = true;
  evaluators_.clear();
  hevaluators_.clear();
  if (n_levels_ > 1 && l_ > 0) compute_ap(s_);
}

FmmConfig RasPreconditioner::hermite_level_cfg() const {
  if (opt_.use_level_cfg) return opt_.level_cfg;
  FmmConfig c;
  c.precision = kFp64;
  c.order = 10;
  return c;
}

const double* RasPreconditioner::level_points(int level) {
  auto jt = level_pts_.find(level);
  if (jt != level_pts_.end()) return jt->second.get();
  DevBuf<double>& buf = level_pts_[level];
  buf.resize(point_idcs_[level].size() * 3);
  dev_gather_points(d_pts_.get(), d_point_idcs_[level].get(), static_cast<int64_t>(point_idcs_[level].size()),
                    buf.get(), s_);
  return buf.get();
}

const double* RasPreconditioner::level_grad_points(int level) {
  auto jt = level_gpts_.find(level);
  if (jt != level_gpts_.end()) return jt->second.get();
  DevBuf<double>& buf = level_gpts_[level];
  buf.resize(gpoint_idcs_[level].size() * 3);
  dev_gather_points(d_gpts_.get(), d_gpoint_idcs_[level].get(), static_cast<int64_t>(gpoint_idcs_[level].size()),
                    buf.get(), s_);
  return buf.get();
}

HermiteEvaluator& RasPreconditioner::hermite_evaluator(int src, int trg) {
  auto key = std::make_pair(src, trg);
  auto it = hevaluators_.find(key);
  if (it != hevaluators_.end()) return *it->second;
  auto ev = std::make_unique<HermiteEvaluator>(model_, hermite_level_cfg(), s_);
  ev->set_sources(level_points(src), static_cast<int64_t>(point_idcs_[src].size()), level_grad_points(src),
                  static_cast<int64_t>(gpoint_idcs_[src].size()));
  ev->set_targets(level_points(trg), static_cast<int64_t>(point_idcs_[trg].size()), level_grad_points(trg),
                  static_cast<int64_t>(gpoint_idcs_[trg].size()));
  HermiteEvaluator& ref = *ev;
  hevaluators_[key] = std::move(ev);
  return ref;
}

Evaluator& RasPreconditioner::evaluator(int src, int trg) {
  auto key = std::make_pair(src, trg);
  auto it = evaluators_.find(key);
  if (it != evaluators_.end()) return *it->second;
  auto lp = [&](int level) -> const double* {
    auto jt = level_pts_.find(level);
    if (jt == level_pts_.end()) {
      DevBuf<double>& buf = level_pts_[level];
      buf.resize(point_idcs_[level].size() * 3);
      dev_gather_points(d_pts_.get(), d_point_idcs_[level].get(), static_cast<int64_t>(point_idcs_[level].size()),
                        buf.get(), s_);
      return buf.get();
    }
    return jt->second.get();
  };
  Bbox3 bb;
  for (int64_t i = 0; i < mu_; i++) bb.add(&h_pts_[3 * i]);
  auto ev = std::make_unique<Evaluator>(model_, bb.lo, bb.hi, INFINITY, s_);
  if (opt_.use_level_cfg) ev->set_fmm_config(opt_.level_cfg);
  ev->set_source_points(lp(src), static_cast<int64_t>(point_idcs_[src].size()));
  ev->set_target_points(lp(trg), static_cast<int64_t>(point_idcs_[trg].size()));
  Evaluator& ref = *ev;
  evaluators_[key] = std::move(ev);
  return ref;
}

void RasPreconditioner::update_residuals(int src, int trg, const double* d_w, double* d_res) {
  // ras_preconditioner.hpp:299-328
  if (sigma_ > 0) {
    const int64_t ns = level_rows_[src], nt = level_rows_[trg];
    sw_.resize(static_cast<size_t>(ns + l_));
    dev_gather(d_w, d_level_rows_[src].get(), ns, sw_.get(), s_);
    if (l_ > 0) dev_copy(d_w + n_, sw_.get() + ns, l_, s_);
    HermiteEvaluator& ev = hermite_evaluator(src, trg);
    ev.set_weights(sw_.get());
    tmp_.resize(static_cast<size_t>(nt));
    ev.evaluate(tmp_.get());
    dev_scatter_add(tmp_.get(), d_level_rows_[trg].get(), nt, -1.0, d_res, s_);
    return;
  }
  int64_t ns = static_cast<int64_t>(point_idcs_[src].size());
  int64_t nt = static_cast<int64_t>(point_idcs_[trg].size());
  sw_.resize(static_cast<size_t>(ns + l_));
  dev_gather(d_w, d_point_idcs_[src].get(), ns, sw_.get(), s_);
  if (l_ > 0) dev_copy(d_w + mu_, sw_.get() + ns, l_, s_);
  Evaluator& ev = evaluator(src, trg);
  ev.set_weights(sw_.get());
  tmp_.resize(static_cast<size_t>(nt));
  ev.evaluate(tmp_.get());
  dev_scatter_add(tmp_.get(), d_point_idcs_[trg].get(), nt, -1.0, d_res, s_);
}

void RasPreconditioner::orthogonalize(double* d_w, double* d_res) {
  // ras_preconditioner.hpp:271-278
  if (l_ == 0) return;
  // The projections stay on the device: no host round trip inside apply().
  dev_multi_dot_device(p_orth_.get(), n_, l_, d_w, n_, dot_.get(), ws_, s_);
  dev_gemv_update(d_w, p_orth_.get(), n_, l_, dot_.get(), n_, -1.0, s_);
  dev_gemv_update(d_res, ap_.get(), n_, l_, dot_.get(), n_, 1.0, s_);
}

void RasPreconditioner::apply(const double* d_v, double* d_out) {
  // ras_preconditioner.hpp:191-250
  // Verbose: the phases of the first two applies are timed (synchronizing
  // after each) and printed.
  const bool timed = opt_.verbose && verbose_applies_ < 2;
  std::string report;
  auto t_last = std::chrono::steady_clock::now();
  auto phase = [&](const std::string& what) {
    if (!timed) return;
    PAM_HIP(hipStreamSynchronize(s_));
    auto now = std::chrono::steady_clock::now();
    char buf[128];
    std::snprintf(buf, sizeof buf, " %s %.2f", what.c_str(), std::chrono::duration<double, std::milli>(now - t_last).count());
    report += buf;
    t_last = now;
  };
  struct Report {
    const bool& timed;
    std::string& text;
    int& count;
    ~Report() {
      if (timed) {
        std::printf("  ras apply (ms):%s\n", text.c_str());
        std::fflush(stdout);
        count++;
      }
    }
  } rep{timed, report, verbose_applies_};
  auto solve_t = [&](int level, const double* r, double* w) {
    solve(level, r, w);
    phase("solve" + std::to_string(level));
  };
  auto update_t = [&](int src, int trg, const double* w, double* r) {
    update_residuals(src, trg, w, r);
    phase("upd" + std::to_string(src) + ">" + std::to_string(trg));
  };
  dev_copy(d_v, res_.get(), n_, s_);
  if (n_levels_ == 1) {
    solve(0, res_.get(), d_out);
    return;
  }
  const int n = n_levels_;
  double* tot = total_.get();
  dev_fill(0.0, tot, n_ + l_, s_);
  solve_t(0, res_.get(), w_.get());
  update_t(0, n - 1, w_.get(), res_.get());
  dev_axpy(1.0, w_.get(), tot, n_ + l_, s_);
  for (int level = 1; level < n - 1; level++) {
    solve_t(level, res_.get(), w_.get());
    update_t(level, n - 1, w_.get(), res_.get());
    dev_axpy(1.0, w_.get(), tot, n_ + l_, s_);
    orthogonalize(tot, res_.get());
    solve_t(0, res_.get(), w_.get());
    update_t(0, n - 1, w_.get(), res_.get());
    dev_axpy(1.0, w_.get(), tot, n_ + l_, s_);
  }
  for (int level = n - 1; level >= 1; level--) {
    solve_t(level, res_.get(), w_.get());
    update_t(level, level - 1, w_.get(), res_.get());
    dev_axpy(1.0, w_.get(), tot, n_ + l_, s_);
    orthogonalize(tot, res_.get());
    solve_t(0, res_.get(), w_.get());
    if (level > 1) update_t(0, level - 1, w_.get(), res_.get());
    dev_axpy(1.0, w_.get(), tot, n_ + l_, s_);
  }
  dev_copy(tot, d_out, n_ + l_, s_);
}

}  // namespace pam
