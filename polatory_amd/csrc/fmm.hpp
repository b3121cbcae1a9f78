// Interpolation FMM stage launchers (Chebyshev tensor-product interpolation of
// order P per axis, P^3 nodes per cell). T = float (mixed form) or double.
//
// Storage:
//   points in tree order as V4<T> = (xi, eta, zeta, w): coordinates local to the
//   point's leaf, scaled so the leaf is [-1,1]^3;
//   expansions [n_cells][kpad(P)] (zero padded), node index (a*P + b)*P + c,
//   a along x; nodes x_a = cos((2a+1) pi / (2P)).
#pragma once

#include <vector>

#include "common.hpp"
#include "rbf_device.hpp"
#include "tree.hpp"

namespace pam {

template <class T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

constexpr int kMinOrder = 3;
constexpr int kMaxOrder = 12;
inline bool order_supported(int p) { return p >= kMinOrder && p <= kMaxOrder; }

// Expansion stride: P^3 padded to 32 (the dense M2L reduces in chunks of 32
// FP32 or 16 FP64 columns, so every chunk is full and needs no bounds test).
inline int kpad_of(int p) { return (p * p * p + 31) / 32 * 32; }
// Row padding of the M2L operators: 128 rows per workgroup.
inline int rpad_of(int r) { return (r + 127) / 128 * 128; }

// M2L operator for one level in MFMA fragment order: A[v][kb][row][16] with
// v in [0, 343), kb in [0, kpad/16), row in [0, rpad); element (row, k) of
// transfer vector v sits at ((v * (kpad/16) + k/16) * rpad + row) * 16 + k % 16.
struct M2LOperator {
  int r = 0, rpad = 0, kpad = 0;
  size_t per_v = 0;  // elements per transfer vector
};

template <class T>
void launch_prepare_points(const Tree& t, const double* d_pts, const double* d_w, V4<T>* out,
                           hipStream_t s);
template <class T>
void launch_set_point_weights(const Tree& t, const double* d_w, V4<T>* pts, hipStream_t s);

template <class T>
void launch_p2m(int p, const Tree& t, const V4<T>* pts, T* m_leaf, hipStream_t s);
// parent level l from child level l+1
template <class T>
void launch_m2m(int p, const Tree& t, int l, const T* m_child, T* m_parent, hipStream_t s);
// child level l from parent level l-1 (accumulates into l_child)
template <class T>
void launch_l2l(int p, const Tree& t, int l, const T* l_parent, T* l_child, hipStream_t s);
// Near-field (P2P) and far-field (L2P) work of one evaluation as kernel
// arguments, so that several evaluations, of any model, can share one launch
// of each (serving). A lone evaluation runs the same kernels with one job.
constexpr int kMaxFmmJobs = 8;
template <class T>
struct P2PJob {
  const int32_t* tls;    // target leaf starts
  const int32_t* sls;    // source leaf starts
  const int32_t* p2p;    // [target leaves][27] source leaves
  const V4<T>* tpts;     // targets in tree order, leaf-local
  const V4<T>* spts;     // sources in tree order, leaf-local, with weights
  const double* y_far;   // far field in tree order, or null
  const int32_t* tperm;  // tree order -> input order of the targets
  double* out;           // output, input order
  T pa, pb;              // kernel constants in the leaf unit
  int n_leaves, first_leaf;
};
template <class T>
struct P2PJobs {
  int n;
  P2PJob<T> j[kMaxFmmJobs];
};
template <class T>
struct L2PJob {
  const int32_t* leaf_start;
  const V4<T>* pts;
  const T* l;            // leaf local expansions
  T* y_far;              // output (input order when perm is set)
  const int32_t* perm;
  int kpad, n_leaves, first_leaf;
};
template <class T>
struct L2PJobs {
  int n;
  L2PJob<T> j[kMaxFmmJobs];
};
template <class T>
P2PJob<T> make_p2p_job(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists, const V4<T>* tpts,
                       const V4<T>* spts, const double* y_far, double* out);
template <class T>
L2PJob<T> make_l2p_job(const Tree& t, const T* l_leaf, const V4<T>* pts, T* y_far, bool add_to_output, int kpad);
// The near field and far field of one evaluation, recorded instead of
// launched (FmmEvaluator::evaluate_deferred).
struct DeferredFmm {
  bool valid = false;
  bool fp32 = false, compensated = false;
  int kind = 0, order = 0;
  P2PJob<float> p2p32{};
  P2PJob<double> p2p64{};
  L2PJob<double> l2p{};
};
// Launches the recorded work of several evaluations: all near fields (shared
// launches per precision, kind and summation form), then all far fields
// (shared launches per order). Each output gets the bits a lone evaluation
// would give it.
void launch_deferred(const std::vector<const DeferredFmm*>& jobs, hipStream_t s);

// add_to_output: y_far is the output in input order, the far field is added
// to it through the tree permutation.
template <class T>
void launch_l2p(int p, const Tree& t, const T* l_leaf, const V4<T>* pts, T* y_far, hipStream_t s,
                bool add_to_output = false);
// out[perm[i]] = y_far[i] + near field, FP64 output in input order.
template <class T>
void launch_p2p(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists,
                const V4<T>* tpts, const V4<T>* spts, const double* y_far, double* out, hipStream_t s,
                bool compensated = false);  // compensated: Kahan sums (FP32 only)
// Gradients at the targets (fmm/gradient_transpose_kernel.hpp:47-62): the
// near field writes A^T sum_j w_j grad phi into out3 (n x 3, input order);
// the far field then adds A^T grad of the leaf local expansions.
template <class T>
void launch_p2p_grad(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists,
                     const V4<T>* tpts, const V4<T>* spts, double* out3, hipStream_t s);
void launch_l2p_grad(int p, const Tree& t, const double* l_leaf, const V4<double>* pts, const double* aniso,
                     double* out3, hipStream_t s);
// Column kcol of the Hessian kernel at the targets (fmm/hessian_kernel.hpp:46-64
// up to its sign): out3 = A^T sum_j w_j H(A (t - s_j)) a_k, a_k = A e_k; the
// near field writes, the far field adds.
template <class T>
void launch_p2p_hess_col(const RbfDesc& rbf, const Tree& trg, const Tree& src, const Lists& lists,
                         const V4<T>* tpts, const V4<T>* spts, int kcol, double* out3, hipStream_t s);
void launch_l2p_hess_col(int p, const Tree& t, const double* l_leaf, const V4<double>* pts, const double* aniso,
                         int kcol, double* out3, hipStream_t s);

// Builds the dense M2L operator of one level: entries phi(h * |x_a + 2v - x_b|)
// for cells of half width h (kernel constants folded by phi_consts).
template <class T>
void build_m2l_dense(int p, const RbfDesc& rbf, double h, M2LOperator& op, DevBuf<T>& a,
                     hipStream_t s);

// One level of an M2L pass: out[col_cell][row] (+)= scale * sum_slots A_v * in[src]
// over the octant's 189 transfer vectors (src_slots != null), or
// out[c] = scale * A_0 * in[c] for every column cell (src_slots == null).
// TM: MFMA arithmetic (float or double); TS: storage of in/out (expansions).
template <class TM, class TS>
struct M2LLevelSpec {
  const M2LLevel* lv = nullptr;
  const int32_t* src_slots = nullptr;
  const TS* in = nullptr;
  TS* out = nullptr;
  const TM* a = nullptr;
  double scale = 1.0;
};

constexpr int kMaxM2LLevels = 20;

template <class TM, class TS>
struct M2LBatch {
  int n;
  int r, rpad, kpad;
  int64_t per_v;
  const int32_t* vlist;
  int accumulate;
  int in_stride, out_stride;
  int blk_begin[kMaxM2LLevels + 1];
  const int32_t* col_cell[kMaxM2LLevels];
  const int32_t* blk_oct[kMaxM2LLevels];
  const int32_t* src_slots[kMaxM2LLevels];
  const TS* in[kMaxM2LLevels];
  TS* out[kMaxM2LLevels];
  const TM* a[kMaxM2LLevels];
  double scale[kMaxM2LLevels];
};

// All levels in one launch (they share the operator shape).
template <class TM, class TS>
void launch_m2l_batch(const M2LOperator& op, const std::vector<M2LLevelSpec<TM, TS>>& levels,
                      const int32_t* vlist, bool accumulate, int in_stride, int out_stride, hipStream_t s);

// Dense M2L, all levels in one launch: out[cell][row] = scale * sum over the
// cell's 189 slots of A_v * in[src], for the rows < r (expansion stride kpad).
// in holds the multipoles in the MFMA type (an FP32 copy for the mixed form).
template <class TM>
struct M2LDenseLevel {
  const M2LLevel* lv = nullptr;
  const TM* in = nullptr;
  double* out = nullptr;
  const TM* a = nullptr;
  double scale = 1.0;
};

// Work list of the dense M2L: one int4 (level, column block, row block, split)
// per workgroup, ordered so that the workgroups that read the same operator
// rows (same octant, row block and split of the slots) land on the same XCD.
struct M2LSchedule {
  HostStage<int32_t> h;
  DevBuf<int32_t> d;
  int n_wg = 0;
  int nsplit = 1;
  int64_t cells_total = 0;
  bool valid = false;
};

template <class TM>
void launch_m2l_dense(const M2LOperator& op, const std::vector<M2LDenseLevel<TM>>& levels,
                      const int32_t* vlist, M2LSchedule& sched, DevBuf<double>& partial, hipStream_t s);

// Pair-major dense M2L (same result as launch_m2l_dense, from the pair lists
// of each level): per-pair transfers into `partial` (type PT), then
// per-target sums in slot order in FP64.
struct M2LPairSchedule {
  HostStage<int32_t> h;
  DevBuf<int32_t> d;
  int n_wg = 0;
  bool valid = false;
};

template <class TM, class PT>
void launch_m2l_pairs(const M2LOperator& op, const std::vector<M2LDenseLevel<TM>>& levels,
                      M2LPairSchedule& sched, DevBuf<PT>& partial, hipStream_t s);

void launch_convert_f64_f32(const double* in, float* out, int64_t n, hipStream_t s);

}  // namespace pam
