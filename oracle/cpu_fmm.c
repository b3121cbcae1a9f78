/*
 * Reference-algorithm CPU FMM: BASELINE AND TEST INFRASTRUCTURE ONLY.
 *
 * An OpenMP FP64 restatement of polatory's FMM evaluate path, used to time the
 * reference algorithm on the host cores beside the GPU (bench.py's
 * cpu_baseline leg) and checked against the oracle's direct sum by the tests.
 * The product path never links, loads or calls it.
 *
 * What it follows (paths relative to the polatory source tree):
 *   - prepare() (src/fmm/fmm_evaluator.hpp:226-271): direct sum when
 *     n_src * n_trg < 1024^2; otherwise tree height
 *     H = max(2, round(ln max(n_src, n_trg) / ln 8)) (src/fmm/utility.hpp:12-17),
 *     leaves at level H - 1, the box of make_box (utility.hpp:19-33);
 *   - evaluate() (fmm_evaluator.hpp:77-112): P2M, M2M, then M2L, L2L, L2P, P2P
 *     with separation criterion 1 (fmm/kernel.hpp:71);
 *   - the interpolator: ScalFMM3's uniform interpolation FMM (external, the
 *     polatory fork at GIT_TAG polatory, src/CMakeLists.txt:155-164, not
 *     pinned to a commit), restated from its published algorithm (Blanchard,
 *     Coulaud and Darve, Inria research report, 2015: Lagrange interpolation
 *     on p equispaced nodes per axis, x_a = -1 + 2a/(p - 1); M2L between two
 *     cells of one level is a Toeplitz product, applied as a Hadamard product
 *     of discrete Fourier transforms of size 2p - 1 per axis). With d >= 0 the
 *     interpolant is Floater and Hormann's barycentric rational interpolant of
 *     degree d on the same nodes (Numer. Math. 107, 2007), the "modified"
 *     option polatory passes (fmm_accuracy_estimator.hpp:108-117); d < 0 is
 *     the classic Lagrange interpolant.
 * Points are pre-transformed by the anisotropy (fmm_evaluator.hpp:125-133), so
 * the kernel is evaluated isotropically.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAXP 24

static double now_s(void) { return omp_get_wtime(); }

static double phi_iso(const orc_rbf* r, double r2) {
  switch (r->kind) {
    case ORC_BH3: return -r->params[0] * sqrt(r2 + r->params[1] * r->params[1]);
    case ORC_TH3: {
      double rho = sqrt(r2 + r->params[1] * r->params[1]);
      return r->params[0] * rho * rho * rho;
    }
    default: return r->params[0] * exp(-3.0 * sqrt(r2) / r->params[1]);
  }
}

/* ---- interpolation on p equispaced nodes ---- */

typedef struct {
  int p, d;
  double x[MAXP];   /* nodes */
  double bw[MAXP];  /* barycentric weights */
} Interp;

static double binom(int n, int k) {
  double b = 1.0;
  for (int i = 1; i <= k; i++) b = b * (n - k + i) / i;
  return b;
}

static void interp_init(Interp* in, int p, int d) {
  in->p = p;
  in->d = d;
  for (int a = 0; a < p; a++) in->x[a] = -1.0 + 2.0 * a / (p - 1);
  const int n = p - 1;
  for (int k = 0; k <= n; k++) {
    if (d < 0 || d >= n) {
      /* classic Lagrange on equispaced nodes: (-1)^k binom(n, k) */
      in->bw[k] = ((k & 1) ? -1.0 : 1.0) * binom(n, k);
    } else {
      /* Floater-Hormann: (-1)^(k-d) sum over i in [max(0, k-d), min(k, n-d)] of binom(d, k-i) */
      double s = 0.0;
      int lo = k - d > 0 ? k - d : 0, hi = k < n - d ? k : n - d;
      for (int i = lo; i <= hi; i++) s += binom(d, k - i);
      in->bw[k] = (((k - d) & 1) ? -1.0 : 1.0) * s;
    }
  }
}

/* S_a(t) for all a: barycentric form, exact node hits handled. */
static void interp_basis(const Interp* in, double t, double* s) {
  const int p = in->p;
  for (int a = 0; a < p; a++) {
    if (t == in->x[a]) {
      for (int b = 0; b < p; b++) s[b] = b == a ? 1.0 : 0.0;
      return;
    }
  }
  double den = 0.0;
  for (int a = 0; a < p; a++) {
    s[a] = in->bw[a] / (t - in->x[a]);
    den += s[a];
  }
  for (int a = 0; a < p; a++) s[a] /= den;
}

/* ---- trees over pre-transformed points ---- */

static uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

static uint64_t morton(int ix, int iy, int iz) {
  return spread3((uint32_t)ix) << 2 | spread3((uint32_t)iy) << 1 | spread3((uint32_t)iz);
}

typedef struct {
  int nlev;            /* levels 0 .. nlev - 1, leaves at nlev - 1 */
  int64_t n;           /* points */
  int32_t* perm;       /* tree order -> input index */
  double* pts;         /* points in tree order */
  double* w;           /* weights in tree order (sources) */
  int64_t* ncell;      /* occupied cells per level */
  int32_t** cix;       /* [lev][cell*3]: integer cell coordinates */
  int32_t** lookup;    /* [lev][morton code]: cell index or -1 */
  int64_t* leaf_start; /* points of leaf cell c: [leaf_start[c], leaf_start[c + 1]) */
  int32_t** parent;    /* [lev][cell]: index at lev - 1 */
} Tree;

static void tree_free(Tree* t) {
  free(t->perm);
  free(t->pts);
  free(t->w);
  for (int l = 0; l < t->nlev; l++) {
    free(t->cix[l]);
    free(t->lookup[l]);
    free(t->parent[l]);
  }
  free(t->cix);
  free(t->lookup);
  free(t->parent);
  free(t->ncell);
  free(t->leaf_start);
}

static void tree_build(Tree* t, const double* pts, const double* w, int64_t n, const double* lo,
                       double width, int nlev) {
  const int L = nlev - 1;
  const int side = 1 << L;
  const int64_t ncode = (int64_t)1 << (3 * L);
  t->nlev = nlev;
  t->n = n;
  uint64_t* code = malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
  const double inv = side / width;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    int c[3];
    for (int k = 0; k < 3; k++) {
      double f = floor((pts[3 * i + k] - lo[k]) * inv);
      c[k] = f < 0 ? 0 : (f > side - 1 ? side - 1 : (int)f);
    }
    code[i] = morton(c[0], c[1], c[2]);
  }
  /* stable counting sort by leaf code */
  int64_t* cnt = calloc(ncode + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) cnt[code[i] + 1]++;
  for (int64_t c = 0; c < ncode; c++) cnt[c + 1] += cnt[c];
  t->perm = malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  {
    int64_t* pos = malloc(sizeof(int64_t) * (ncode + 1));
    memcpy(pos, cnt, sizeof(int64_t) * (ncode + 1));
    for (int64_t i = 0; i < n; i++) t->perm[pos[code[i]]++] = (int32_t)i;
    free(pos);
  }
  t->pts = malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
  t->w = malloc(sizeof(double) * (n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    int64_t j = t->perm[i];
    for (int k = 0; k < 3; k++) t->pts[3 * i + k] = pts[3 * j + k];
    t->w[i] = w ? w[j] : 0.0;
  }
  t->ncell = calloc(nlev, sizeof(int64_t));
  t->cix = calloc(nlev, sizeof(int32_t*));
  t->lookup = calloc(nlev, sizeof(int32_t*));
  t->parent = calloc(nlev, sizeof(int32_t*));
  /* leaves: occupied codes in Morton order */
  int64_t nl = 0;
  for (int64_t c = 0; c < ncode; c++)
    if (cnt[c + 1] > cnt[c]) nl++;
  t->leaf_start = malloc(sizeof(int64_t) * (nl + 1));
  uint64_t* lcode = malloc(sizeof(uint64_t) * (nl > 0 ? nl : 1));
  nl = 0;
  for (int64_t c = 0; c < ncode; c++)
    if (cnt[c + 1] > cnt[c]) {
      lcode[nl] = (uint64_t)c;
      t->leaf_start[nl++] = cnt[c];
    }
  t->leaf_start[nl] = n;
  free(cnt);
  free(code);
  /* every level: unique parent codes (Morton order is kept by the shift) */
  uint64_t* codes = lcode;
  int64_t nc = nl;
  for (int l = L; l >= 0; l--) {
    t->ncell[l] = nc;
    t->cix[l] = malloc(sizeof(int32_t) * 3 * (nc > 0 ? nc : 1));
    int64_t nco = (int64_t)1 << (3 * l);
    t->lookup[l] = malloc(sizeof(int32_t) * nco);
    for (int64_t c = 0; c < nco; c++) t->lookup[l][c] = -1;
    for (int64_t c = 0; c < nc; c++) {
      uint64_t m = codes[c];
      int ix = 0, iy = 0, iz = 0;
      for (int b = 0; b < l; b++) {
        ix |= (int)((m >> (3 * b + 2)) & 1) << b;
        iy |= (int)((m >> (3 * b + 1)) & 1) << b;
        iz |= (int)((m >> (3 * b)) & 1) << b;
      }
      t->cix[l][3 * c] = ix;
      t->cix[l][3 * c + 1] = iy;
      t->cix[l][3 * c + 2] = iz;
      t->lookup[l][m] = (int32_t)c;
    }
    if (l == 0) break;
    uint64_t* up = malloc(sizeof(uint64_t) * (nc > 0 ? nc : 1));
    int64_t nu = 0;
    t->parent[l] = malloc(sizeof(int32_t) * (nc > 0 ? nc : 1));
    for (int64_t c = 0; c < nc; c++) {
      uint64_t pc = codes[c] >> 3;
      if (nu == 0 || up[nu - 1] != pc) up[nu++] = pc;
      t->parent[l][c] = (int32_t)(nu - 1);
    }
    free(codes);
    codes = up;
    nc = nu;
  }
  free(codes);
}

/* ---- discrete Fourier transforms of size N = 2p - 1 (real data, half spectrum) ---- */

typedef struct {
  int p, N, nh;     /* nh = p: kz in [0, p) */
  double *cr, *ci;  /* cos / -sin tables [N][N]: e^{-2 pi i k n / N} */
} Dft;

static void dft_init(Dft* f, int p) {
  f->p = p;
  f->N = 2 * p - 1;
  f->nh = p;
  const int N = f->N;
  f->cr = malloc(sizeof(double) * N * N);
  f->ci = malloc(sizeof(double) * N * N);
  for (int k = 0; k < N; k++)
    for (int n = 0; n < N; n++) {
      double a = -2.0 * M_PI * (double)((int64_t)k * n % N) / N;
      f->cr[k * N + n] = cos(a);
      f->ci[k * N + n] = sin(a);
    }
}

static void dft_free(Dft* f) {
  free(f->cr);
  free(f->ci);
}

/* Half spectrum of a real N^3 array given on [0, m)^3 (zero elsewhere):
 * out[(kx * N + ky) * nh + kz], kz < nh. in is [m][m][m] with stride m. */
static void dft_forward(const Dft* f, const double* in, int m, double* outr, double* outi, double* wr,
                        double* wi) {
  const int N = f->N, nh = f->nh;
  /* z: in[x][y][z] -> A[x][y][kz] for x, y < m */
  double* ar = wr;
  double* ai = wi;
  for (int x = 0; x < m; x++)
    for (int y = 0; y < m; y++) {
      const double* v = in + (x * m + y) * m;
      for (int kz = 0; kz < nh; kz++) {
        double sr = 0.0, si = 0.0;
        for (int z = 0; z < m; z++) {
          sr += v[z] * f->cr[kz * N + z];
          si += v[z] * f->ci[kz * N + z];
        }
        ar[(x * m + y) * nh + kz] = sr;
        ai[(x * m + y) * nh + kz] = si;
      }
    }
  /* y: A[x][y][kz] -> B[x][ky][kz] */
  double* br = wr + m * m * nh;
  double* bi = wi + m * m * nh;
  for (int x = 0; x < m; x++)
    for (int ky = 0; ky < N; ky++) {
      double* orr = br + (x * N + ky) * nh;
      double* oii = bi + (x * N + ky) * nh;
      for (int kz = 0; kz < nh; kz++) orr[kz] = oii[kz] = 0.0;
      for (int y = 0; y < m; y++) {
        const double c = f->cr[ky * N + y], s = f->ci[ky * N + y];
        const double* xr = ar + (x * m + y) * nh;
        const double* xi = ai + (x * m + y) * nh;
        for (int kz = 0; kz < nh; kz++) {
          orr[kz] += xr[kz] * c - xi[kz] * s;
          oii[kz] += xr[kz] * s + xi[kz] * c;
        }
      }
    }
  /* x: B[x][ky][kz] -> out[kx][ky][kz] */
  for (int kx = 0; kx < N; kx++) {
    double* orr = outr + (int64_t)kx * N * nh;
    double* oii = outi + (int64_t)kx * N * nh;
    for (int q = 0; q < N * nh; q++) orr[q] = oii[q] = 0.0;
    for (int x = 0; x < m; x++) {
      const double c = f->cr[kx * N + x], s = f->ci[kx * N + x];
      const double* xr = br + (int64_t)x * N * nh;
      const double* xi = bi + (int64_t)x * N * nh;
      for (int q = 0; q < N * nh; q++) {
        orr[q] += xr[q] * c - xi[q] * s;
        oii[q] += xr[q] * s + xi[q] * c;
      }
    }
  }
}

/* Inverse of a half spectrum of a real N^3 array, the result on [0, p)^3 only,
 * scaled by 1 / N^3 and added to out[(x * p + y) * p + z]. */
static void dft_inverse_add(const Dft* f, const double* inr, const double* ini, double* out, double* wr,
                            double* wi) {
  const int N = f->N, nh = f->nh, p = f->p;
  /* x: in[kx][ky][kz] -> A[x][ky][kz], x < p (conjugate twiddles) */
  double* ar = wr;
  double* ai = wi;
  for (int x = 0; x < p; x++) {
    double* orr = ar + (int64_t)x * N * nh;
    double* oii = ai + (int64_t)x * N * nh;
    for (int q = 0; q < N * nh; q++) orr[q] = oii[q] = 0.0;
    for (int kx = 0; kx < N; kx++) {
      const double c = f->cr[kx * N + x], s = -f->ci[kx * N + x];
      const double* xr = inr + (int64_t)kx * N * nh;
      const double* xi = ini + (int64_t)kx * N * nh;
      for (int q = 0; q < N * nh; q++) {
        orr[q] += xr[q] * c - xi[q] * s;
        oii[q] += xr[q] * s + xi[q] * c;
      }
    }
  }
  /* y: A[x][ky][kz] -> B[x][y][kz], y < p */
  double* br = wr + (int64_t)p * N * nh;
  double* bi = wi + (int64_t)p * N * nh;
  for (int x = 0; x < p; x++)
    for (int y = 0; y < p; y++) {
      double* orr = br + (x * p + y) * nh;
      double* oii = bi + (x * p + y) * nh;
      for (int kz = 0; kz < nh; kz++) orr[kz] = oii[kz] = 0.0;
      for (int ky = 0; ky < N; ky++) {
        const double c = f->cr[ky * N + y], s = -f->ci[ky * N + y];
        const double* xr = ar + ((int64_t)x * N + ky) * nh;
        const double* xi = ai + ((int64_t)x * N + ky) * nh;
        for (int kz = 0; kz < nh; kz++) {
          orr[kz] += xr[kz] * c - xi[kz] * s;
          oii[kz] += xr[kz] * s + xi[kz] * c;
        }
      }
    }
  /* z: real line from its half spectrum (N odd: X[N - k] = conj X[k]) */
  const double scale = 1.0 / ((double)N * N * N);
  for (int x = 0; x < p; x++)
    for (int y = 0; y < p; y++) {
      const double* xr = br + (x * p + y) * nh;
      const double* xi = bi + (x * p + y) * nh;
      for (int z = 0; z < p; z++) {
        double s = xr[0];
        for (int kz = 1; kz < nh; kz++) {
          /* Re(X[kz] e^{+2 pi i kz z / N}) */
          s += 2.0 * (xr[kz] * f->cr[kz * N + z] + xi[kz] * f->ci[kz * N + z]);
        }
        out[(x * p + y) * p + z] += s * scale;
      }
    }
}

/* ---- the FMM ---- */

static int fmm_height(int64_t n) {
  int h = (int)llround(log((double)n) / log(8.0));
  return h > 2 ? h : 2;
}

int orc_cpu_fmm_evaluate(const orc_rbf* rbf, const double* bbox_min, const double* bbox_max,
                         const double* src_raw, int64_t n_src, const double* w, const double* trg_raw,
                         int64_t n_trg, int order, int fh_degree, int height, double* out,
                         orc_fmm_stats* st) {
  memset(st, 0, sizeof *st);
  double t0 = now_s();
  if (order < 2 || order > MAXP) return -1;
  for (int64_t i = 0; i < n_trg; i++) out[i] = 0.0;
  double* src = malloc(sizeof(double) * 3 * (n_src > 0 ? n_src : 1));
  double* trg = malloc(sizeof(double) * 3 * (n_trg > 0 ? n_trg : 1));
  orc_transform_points(rbf->aniso, src_raw, n_src, src);
  orc_transform_points(rbf->aniso, trg_raw, n_trg, trg);
  orc_rbf iso = *rbf;
  for (int k = 0; k < 9; k++) iso.aniso[k] = (k % 4 == 0) ? 1.0 : 0.0;
  if (n_src * n_trg < 1024 * 1024 && height <= 0) {
    /* fmm_evaluator.hpp:227-234: full_direct below 1024^2 pairs */
    orc_direct_rbf_sum(&iso, src, n_src, w, trg, n_trg, out);
    st->height = 0;
    st->t_total = now_s() - t0;
    free(src);
    free(trg);
    return 0;
  }
  const int H = height > 0 ? height : fmm_height(n_src > n_trg ? n_src : n_trg);
  if (H > 9) {
    free(src);
    free(trg);
    return -2;
  }
  const int L = H - 1, p = order, p3 = p * p * p;
  double center[3], width;
  orc_make_box(bbox_min, bbox_max, rbf->aniso, center, &width);
  double lo[3];
  for (int k = 0; k < 3; k++) lo[k] = center[k] - 0.5 * width;
  Tree ts, tt;
  tree_build(&ts, src, w, n_src, lo, width, H);
  tree_build(&tt, trg, NULL, n_trg, lo, width, H);
  free(src);
  free(trg);
  st->height = H;
  st->order = p;
  double t1 = now_s();
  st->t_tree = t1 - t0;

  Interp ip;
  interp_init(&ip, p, fh_degree);
  /* M2M / L2L transfer: T[s][alpha][a] = S_alpha((x_a + (2 s - 1)) / 2) */
  double T[2][MAXP][MAXP];
  for (int s = 0; s < 2; s++)
    for (int a = 0; a < p; a++) {
      double sv[MAXP];
      interp_basis(&ip, 0.5 * (ip.x[a] + (2 * s - 1)), sv);
      for (int al = 0; al < p; al++) T[s][al][a] = sv[al];
    }

  /* P2M at the leaves, M2M up to level 2 */
  double** mul = calloc(H, sizeof(double*));
  for (int l = 2; l <= L; l++) mul[l] = calloc((size_t)ts.ncell[l] * p3, sizeof(double));
  if (L >= 2) {
    const double cw = width / (1 << L);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < ts.ncell[L]; c++) {
      double* m = mul[L] + (size_t)c * p3;
      double ctr[3];
      for (int k = 0; k < 3; k++) ctr[k] = lo[k] + (ts.cix[L][3 * c + k] + 0.5) * cw;
      for (int64_t i = ts.leaf_start[c]; i < ts.leaf_start[c + 1]; i++) {
        double sx[MAXP], sy[MAXP], sz[MAXP];
        interp_basis(&ip, 2.0 * (ts.pts[3 * i] - ctr[0]) / cw, sx);
        interp_basis(&ip, 2.0 * (ts.pts[3 * i + 1] - ctr[1]) / cw, sy);
        interp_basis(&ip, 2.0 * (ts.pts[3 * i + 2] - ctr[2]) / cw, sz);
        const double wi = ts.w[i];
        for (int a = 0; a < p; a++)
          for (int b = 0; b < p; b++) {
            const double f = wi * sx[a] * sy[b];
            double* row = m + (a * p + b) * p;
            for (int cc = 0; cc < p; cc++) row[cc] += f * sz[cc];
          }
      }
    }
  }
  for (int l = L; l > 2; l--) {
    /* children at l -> parents at l - 1, gathered per parent */
    const int64_t nc = ts.ncell[l];
#pragma omp parallel
    {
      double* t1b = malloc(sizeof(double) * p3);
      double* t2b = malloc(sizeof(double) * p3);
#pragma omp for schedule(dynamic, 64)
      for (int64_t c = 0; c < nc; c++) {
        /* process each parent once: only from its first child */
        int32_t par = ts.parent[l][c];
        if (c > 0 && ts.parent[l][c - 1] == par) continue;
        double* pm = mul[l - 1] + (size_t)par * p3;
        for (int64_t ch = c; ch < nc && ts.parent[l][ch] == par; ch++) {
          const int sx = ts.cix[l][3 * ch] & 1, sy = ts.cix[l][3 * ch + 1] & 1, sz = ts.cix[l][3 * ch + 2] & 1;
          const double* cm = mul[l] + (size_t)ch * p3;
          /* z, then y, then x */
          for (int a = 0; a < p; a++)
            for (int b = 0; b < p; b++)
              for (int g = 0; g < p; g++) {
                double s = 0.0;
                for (int cc = 0; cc < p; cc++) s += T[sz][g][cc] * cm[(a * p + b) * p + cc];
                t1b[(a * p + b) * p + g] = s;
              }
          for (int a = 0; a < p; a++)
            for (int be = 0; be < p; be++)
              for (int g = 0; g < p; g++) {
                double s = 0.0;
                for (int b = 0; b < p; b++) s += T[sy][be][b] * t1b[(a * p + b) * p + g];
                t2b[(a * p + be) * p + g] = s;
              }
          for (int al = 0; al < p; al++)
            for (int be = 0; be < p; be++)
              for (int g = 0; g < p; g++) {
                double s = 0.0;
                for (int a = 0; a < p; a++) s += T[sx][al][a] * t2b[(a * p + be) * p + g];
                pm[(al * p + be) * p + g] += s;
              }
        }
      }
      free(t1b);
      free(t2b);
    }
  }
  double t2 = now_s();
  st->t_upward = t2 - t1;

  /* M2L per level through DFTs of size N = 2p - 1 */
  Dft f;
  dft_init(&f, p);
  const int N = f.N, nh = f.nh;
  const int64_t ns = (int64_t)N * N * nh;
  double** loc = calloc(H, sizeof(double*));
  for (int l = 2; l <= L; l++) loc[l] = calloc((size_t)tt.ncell[l] * p3, sizeof(double));
  int64_t m2l_pairs = 0;
  for (int l = 2; l <= L; l++) {
    const double cw = width / (1 << l);
    const double delta = cw / (p - 1);
    /* kernel spectra for u = t - s in [-3, 3]^3, |u|_inf >= 2 */
    double* kr = calloc((size_t)343 * ns, sizeof(double));
    double* ki = calloc((size_t)343 * ns, sizeof(double));
#pragma omp parallel
    {
      double* grid = malloc(sizeof(double) * N * N * N);
      double* wr = malloc(sizeof(double) * 2 * N * N * N);
      double* wi = malloc(sizeof(double) * 2 * N * N * N);
#pragma omp for schedule(dynamic, 1)
      for (int u = 0; u < 343; u++) {
        int ux = u / 49 - 3, uy = (u / 7) % 7 - 3, uz = u % 7 - 3;
        int ax = abs(ux), ay = abs(uy), az = abs(uz);
        int mx = ax > ay ? ax : ay;
        mx = mx > az ? mx : az;
        if (mx < 2) continue;
        /* C[k] = K(delta ((p - 1) u + d)), d = k (k < p) or k - N */
        for (int kx = 0; kx < N; kx++)
          for (int ky = 0; ky < N; ky++)
            for (int kz = 0; kz < N; kz++) {
              int dx = kx < p ? kx : kx - N, dy = ky < p ? ky : ky - N, dz = kz < p ? kz : kz - N;
              double rx = delta * ((p - 1) * ux + dx), ry = delta * ((p - 1) * uy + dy),
                     rz = delta * ((p - 1) * uz + dz);
              grid[(kx * N + ky) * N + kz] = phi_iso(&iso, rx * rx + ry * ry + rz * rz);
            }
        dft_forward(&f, grid, N, kr + (size_t)u * ns, ki + (size_t)u * ns, wr, wi);
      }
      free(grid);
      free(wr);
      free(wi);
    }
    /* multipole spectra of every source cell at level l */
    const int64_t nsc = ts.ncell[l];
    double* mr = malloc(sizeof(double) * (size_t)nsc * ns);
    double* mi = malloc(sizeof(double) * (size_t)nsc * ns);
#pragma omp parallel
    {
      double* wr = malloc(sizeof(double) * 2 * N * N * N);
      double* wi = malloc(sizeof(double) * 2 * N * N * N);
#pragma omp for schedule(dynamic, 64)
      for (int64_t c = 0; c < nsc; c++)
        dft_forward(&f, mul[l] + (size_t)c * p3, p, mr + (size_t)c * ns, mi + (size_t)c * ns, wr, wi);
      free(wr);
      free(wi);
    }
    /* targets: Hadamard products summed over the interaction list, one inverse each */
    const int side = 1 << l;
    int64_t pairs = 0;
#pragma omp parallel reduction(+ : pairs)
    {
      double* accr = malloc(sizeof(double) * ns);
      double* acci = malloc(sizeof(double) * ns);
      double* wr = malloc(sizeof(double) * 2 * N * N * N);
      double* wi = malloc(sizeof(double) * 2 * N * N * N);
#pragma omp for schedule(dynamic, 16)
      for (int64_t c = 0; c < tt.ncell[l]; c++) {
        const int tx = tt.cix[l][3 * c], ty = tt.cix[l][3 * c + 1], tz = tt.cix[l][3 * c + 2];
        int any = 0;
        for (int64_t q = 0; q < ns; q++) accr[q] = acci[q] = 0.0;
        for (int px = (tx >> 1) - 1; px <= (tx >> 1) + 1; px++)
          for (int py = (ty >> 1) - 1; py <= (ty >> 1) + 1; py++)
            for (int pz = (tz >> 1) - 1; pz <= (tz >> 1) + 1; pz++)
              for (int o = 0; o < 8; o++) {
                int sx = 2 * px + (o >> 2), sy = 2 * py + ((o >> 1) & 1), sz = 2 * pz + (o & 1);
                if (sx < 0 || sy < 0 || sz < 0 || sx >= side || sy >= side || sz >= side) continue;
                int ux = tx - sx, uy = ty - sy, uz = tz - sz;
                if (abs(ux) <= 1 && abs(uy) <= 1 && abs(uz) <= 1) continue;
                int32_t s = ts.lookup[l][morton(sx, sy, sz)];
                if (s < 0) continue;
                const int u = (ux + 3) * 49 + (uy + 3) * 7 + (uz + 3);
                const double* Kr = kr + (size_t)u * ns;
                const double* Ki = ki + (size_t)u * ns;
                const double* Mr = mr + (size_t)s * ns;
                const double* Mi = mi + (size_t)s * ns;
                for (int64_t q = 0; q < ns; q++) {
                  accr[q] += Kr[q] * Mr[q] - Ki[q] * Mi[q];
                  acci[q] += Kr[q] * Mi[q] + Ki[q] * Mr[q];
                }
                any = 1;
                pairs++;
              }
        if (any) dft_inverse_add(&f, accr, acci, loc[l] + (size_t)c * p3, wr, wi);
      }
      free(accr);
      free(acci);
      free(wr);
      free(wi);
    }
    m2l_pairs += pairs;
    free(mr);
    free(mi);
    free(kr);
    free(ki);
  }
  dft_free(&f);
  double t3 = now_s();
  st->t_m2l = t3 - t2;

  /* L2L down to the leaves */
  for (int l = 3; l <= L; l++) {
#pragma omp parallel
    {
      double* t1b = malloc(sizeof(double) * p3);
      double* t2b = malloc(sizeof(double) * p3);
#pragma omp for schedule(dynamic, 64)
      for (int64_t c = 0; c < tt.ncell[l]; c++) {
        const double* pl = loc[l - 1] + (size_t)tt.parent[l][c] * p3;
        double* cl = loc[l] + (size_t)c * p3;
        const int sx = tt.cix[l][3 * c] & 1, sy = tt.cix[l][3 * c + 1] & 1, sz = tt.cix[l][3 * c + 2] & 1;
        /* child[a][b][c] += sum T[sx][al][a] T[sy][be][b] T[sz][g][c] parent[al][be][g] */
        for (int a = 0; a < p; a++)
          for (int be = 0; be < p; be++)
            for (int g = 0; g < p; g++) {
              double s = 0.0;
              for (int al = 0; al < p; al++) s += T[sx][al][a] * pl[(al * p + be) * p + g];
              t1b[(a * p + be) * p + g] = s;
            }
        for (int a = 0; a < p; a++)
          for (int b = 0; b < p; b++)
            for (int g = 0; g < p; g++) {
              double s = 0.0;
              for (int be = 0; be < p; be++) s += T[sy][be][b] * t1b[(a * p + be) * p + g];
              t2b[(a * p + b) * p + g] = s;
            }
        for (int a = 0; a < p; a++)
          for (int b = 0; b < p; b++)
            for (int cc = 0; cc < p; cc++) {
              double s = 0.0;
              for (int g = 0; g < p; g++) s += T[sz][g][cc] * t2b[(a * p + b) * p + g];
              cl[(a * p + b) * p + cc] += s;
            }
      }
      free(t1b);
      free(t2b);
    }
  }
  /* L2P */
  double* yt = calloc(n_trg > 0 ? n_trg : 1, sizeof(double));
  if (L >= 2) {
    const double cw = width / (1 << L);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < tt.ncell[L]; c++) {
      const double* lc = loc[L] + (size_t)c * p3;
      double ctr[3];
      for (int k = 0; k < 3; k++) ctr[k] = lo[k] + (tt.cix[L][3 * c + k] + 0.5) * cw;
      for (int64_t i = tt.leaf_start[c]; i < tt.leaf_start[c + 1]; i++) {
        double sx[MAXP], sy[MAXP], sz[MAXP];
        interp_basis(&ip, 2.0 * (tt.pts[3 * i] - ctr[0]) / cw, sx);
        interp_basis(&ip, 2.0 * (tt.pts[3 * i + 1] - ctr[1]) / cw, sy);
        interp_basis(&ip, 2.0 * (tt.pts[3 * i + 2] - ctr[2]) / cw, sz);
        double s = 0.0;
        for (int a = 0; a < p; a++)
          for (int b = 0; b < p; b++) {
            const double* row = lc + (a * p + b) * p;
            double r = 0.0;
            for (int cc = 0; cc < p; cc++) r += row[cc] * sz[cc];
            s += sx[a] * sy[b] * r;
          }
        yt[i] = s;
      }
    }
  }
  double t4 = now_s();
  st->t_downward = t4 - t3;

  /* P2P over the 27 neighbour leaves */
  int64_t p2p_pairs = 0;
  {
    const int side = 1 << L;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : p2p_pairs)
    for (int64_t c = 0; c < tt.ncell[L]; c++) {
      const int tx = tt.cix[L][3 * c], ty = tt.cix[L][3 * c + 1], tz = tt.cix[L][3 * c + 2];
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dz = -1; dz <= 1; dz++) {
            int sx = tx + dx, sy = ty + dy, sz = tz + dz;
            if (sx < 0 || sy < 0 || sz < 0 || sx >= side || sy >= side || sz >= side) continue;
            int32_t s = ts.lookup[L][morton(sx, sy, sz)];
            if (s < 0) continue;
            for (int64_t i = tt.leaf_start[c]; i < tt.leaf_start[c + 1]; i++) {
              const double x = tt.pts[3 * i], y = tt.pts[3 * i + 1], z = tt.pts[3 * i + 2];
              double acc = 0.0;
              for (int64_t j = ts.leaf_start[s]; j < ts.leaf_start[s + 1]; j++) {
                double ex = x - ts.pts[3 * j], ey = y - ts.pts[3 * j + 1], ez = z - ts.pts[3 * j + 2];
                acc += ts.w[j] * phi_iso(&iso, ex * ex + ey * ey + ez * ez);
              }
              yt[i] += acc;
            }
            p2p_pairs += (tt.leaf_start[c + 1] - tt.leaf_start[c]) * (ts.leaf_start[s + 1] - ts.leaf_start[s]);
          }
    }
  }
  for (int64_t i = 0; i < n_trg; i++) out[tt.perm[i]] = yt[i];
  double t5 = now_s();
  st->t_p2p = t5 - t4;
  st->t_total = t5 - t0;
  st->m2l_pairs = m2l_pairs;
  st->p2p_pairs = p2p_pairs;
  st->n_src_leaves = ts.ncell[L];
  st->n_trg_leaves = tt.ncell[L];
  free(yt);
  for (int l = 0; l < H; l++) {
    free(mul[l]);
    free(loc[l]);
  }
  free(mul);
  free(loc);
  tree_free(&ts);
  tree_free(&tt);
  return 0;
}
