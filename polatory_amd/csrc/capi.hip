// extern "C" entry points: exceptions become status codes, messages are kept
// per thread for pam_last_error().
#include <cstring>
#include <new>
#include <string>

#include "../../include/polatory_amd.h"
#include "capi_util.hpp"
#include "device_api.hpp"
#include "fmm_dipole.hpp"
#include "fmm_evaluator.hpp"

struct pam_fmm_evaluator {
  std::unique_ptr<pam::FmmEvaluator> impl;
  bool symmetric;
  std::unique_ptr<pam::DipoleSourceEvaluator> dipole;   // gradient-kernel handles
  std::unique_ptr<pam::DipoleHessianEvaluator> hessian;  // Hessian-kernel handles
  // Symmetric at the interface over a source-target evaluator: set_points
  // installs the points as both sources and targets.
  bool same_points = false;
};

namespace pam {

std::string& last_error_ref() {
  thread_local std::string e;
  return e;
}

RbfDesc to_desc(const pam_rbf* r) {
  if (!r) throw_invalid("rbf is null");
  if (r->kind != PAM_RBF_BH3 && r->kind != PAM_RBF_TH3 && r->kind != PAM_RBF_EXP)
    throw_invalid("unknown rbf kind");
  RbfDesc d{};
  d.kind = r->kind;
  d.p0 = r->params[0];
  d.p1 = r->params[1];
  for (int i = 0; i < 9; i++) d.aniso[i] = r->aniso[i];
  const double* a = d.aniso;
  double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
               a[2] * (a[3] * a[7] - a[4] * a[6]);
  if (!(det > 0.0)) throw_invalid("aniso must have a positive determinant");
  if (d.kind == PAM_RBF_EXP && !(d.p1 > 0.0)) throw_invalid("range must be positive");
  if (!(d.p0 >= 0.0) || !(d.p1 >= 0.0)) throw_invalid("rbf parameters must be nonnegative");
  return d;
}

}  // namespace pam

namespace {

using pam::guard;
using pam::to_desc;

pam::FmmEvaluator& ev(pam_fmm_evaluator* h) {
  if (!h || !h->impl) pam::throw_invalid("null evaluator handle");
  return *h->impl;
}

int create(const pam_rbf* rbf, const double* bmin, const double* bmax, pam_fmm_evaluator** out,
           bool symmetric) {
  return guard([&] {
    if (!out) pam::throw_invalid("out is null");
    auto d = to_desc(rbf);
    auto* h = new pam_fmm_evaluator{std::make_unique<pam::FmmEvaluator>(d, bmin, bmax, symmetric), symmetric};
    *out = h;
  });
}

const pam::Tree& which_tree(pam_fmm_evaluator* h, int which) {
  return which == 0 ? ev(h).source_tree() : ev(h).target_tree();
}

}  // namespace

extern "C" {

int pam_last_error(char* buf, size_t len) {
  if (buf && len > 0) {
    std::strncpy(buf, pam::last_error_ref().c_str(), len - 1);
    buf[len - 1] = '\0';
  }
  return PAM_OK;
}

int pam_device_count(int* count) {
  return guard([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int pam_device_cache_release(void) {
  return guard([&] { pam::device_pool_trim(); });
}

int pam_device_cache_bytes(size_t* bytes) {
  return guard([&] {
    if (!bytes) pam::throw_invalid("null argument");
    *bytes = pam::device_pool_cached_bytes();
  });
}

int pam_fmm_evaluator_create(const pam_rbf* rbf, const double bbox_min[3], const double bbox_max[3],
                             pam_fmm_evaluator** out) {
  return create(rbf, bbox_min, bbox_max, out, false);
}

int pam_fmm_symmetric_evaluator_create(const pam_rbf* rbf, const double bbox_min[3],
                                       const double bbox_max[3], pam_fmm_evaluator** out) {
  return create(rbf, bbox_min, bbox_max, out, true);
}

int pam_fmm_gradient_transpose_evaluator_create(const pam_rbf* rbf, const double bbox_min[3],
                                                const double bbox_max[3], pam_fmm_evaluator** out) {
  int rc = create(rbf, bbox_min, bbox_max, out, false);
  if (rc != 0) return rc;
  return guard([&] { ev(*out).set_gradient_kernel(true); });
}

int pam_fmm_gradient_evaluator_create(const pam_rbf* rbf, const double bbox_min[3], const double bbox_max[3],
                                      pam_fmm_evaluator** out) {
  int rc = create(rbf, bbox_min, bbox_max, out, false);
  if (rc != 0) return rc;
  return guard([&] {
    ev(*out).set_gradient_kernel(true);
    (*out)->dipole = std::make_unique<pam::DipoleSourceEvaluator>(ev(*out));
  });
}

int pam_fmm_hessian_evaluator_create(const pam_rbf* rbf, const double bbox_min[3], const double bbox_max[3],
                                     pam_fmm_evaluator** out) {
  int rc = create(rbf, bbox_min, bbox_max, out, false);
  if (rc != 0) return rc;
  return guard([&] {
    ev(*out).set_gradient_kernel(true);
    (*out)->hessian = std::make_unique<pam::DipoleHessianEvaluator>(ev(*out));
  });
}

int pam_fmm_hessian_symmetric_evaluator_create(const pam_rbf* rbf, const double bbox_min[3],
                                               const double bbox_max[3], pam_fmm_evaluator** out) {
  int rc = pam_fmm_hessian_evaluator_create(rbf, bbox_min, bbox_max, out);
  if (rc != 0) return rc;
  (*out)->symmetric = true;
  (*out)->same_points = true;
  return 0;
}

int pam_fmm_evaluator_destroy(pam_fmm_evaluator* h) {
  return guard([&] { delete h; });
}

int pam_fmm_evaluator_set_accuracy(pam_fmm_evaluator* h, double accuracy) {
  return guard([&] { ev(h).set_accuracy(accuracy); });
}

int pam_fmm_evaluator_set_search_precision(pam_fmm_evaluator* h, int precision) {
  return guard([&] { ev(h).set_search_precision(precision); });
}

int pam_fmm_evaluator_set_source_points(pam_fmm_evaluator* h, const double* xyz, int64_t n) {
  return guard([&] {
    if (h && h->symmetric) pam::throw_runtime("use set_points on a symmetric evaluator");
    ev(h).set_source_points(xyz, n, false);
  });
}

int pam_fmm_evaluator_set_target_points(pam_fmm_evaluator* h, const double* xyz, int64_t n) {
  return guard([&] {
    if (h && h->same_points) pam::throw_runtime("use set_points on a symmetric evaluator");
    ev(h).set_target_points(xyz, n, false);
  });
}

int pam_fmm_evaluator_set_points(pam_fmm_evaluator* h, const double* xyz, int64_t n) {
  return guard([&] {
    if (h && !h->symmetric) pam::throw_runtime("set_points needs a symmetric evaluator");
    ev(h).set_source_points(xyz, n, false);
    if (h->same_points) ev(h).set_target_points(xyz, n, false);
  });
}

int pam_fmm_evaluator_set_weights(pam_fmm_evaluator* h, const double* w, int64_t n) {
  return guard([&] {
    if (h && h->dipole) h->dipole->set_weights(w, n, false);
    else if (h && h->hessian) h->hessian->set_weights(w, n, false);
    else ev(h).set_weights(w, n, false);
  });
}

int pam_fmm_evaluator_evaluate(pam_fmm_evaluator* h, double* out, int64_t n) {
  return guard([&] {
    if (h && h->dipole) h->dipole->evaluate(out, n, false);
    else if (h && h->hessian) h->hessian->evaluate(out, n, false);
    else ev(h).evaluate(out, n, false);
  });
}

int pam_fmm_evaluator_set_source_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n) {
  return guard([&] {
    if (h && h->symmetric) pam::throw_runtime("use set_points on a symmetric evaluator");
    ev(h).set_source_points(d_xyz, n, true);
  });
}

int pam_fmm_evaluator_set_target_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n) {
  return guard([&] {
    if (h && h->same_points) pam::throw_runtime("use set_points on a symmetric evaluator");
    ev(h).set_target_points(d_xyz, n, true);
  });
}

int pam_fmm_evaluator_set_points_device(pam_fmm_evaluator* h, const double* d_xyz, int64_t n) {
  return guard([&] {
    if (h && !h->symmetric) pam::throw_runtime("set_points needs a symmetric evaluator");
    ev(h).set_source_points(d_xyz, n, true);
    if (h->same_points) ev(h).set_target_points(d_xyz, n, true);
  });
}

int pam_fmm_evaluator_set_weights_device(pam_fmm_evaluator* h, const double* d_w, int64_t n) {
  return guard([&] {
    if (h && h->dipole) h->dipole->set_weights(d_w, n, true);
    else if (h && h->hessian) h->hessian->set_weights(d_w, n, true);
    else ev(h).set_weights(d_w, n, true);
  });
}

int pam_fmm_evaluator_evaluate_device(pam_fmm_evaluator* h, double* d_out, int64_t n) {
  return guard([&] {
    if (h && h->dipole) h->dipole->evaluate(d_out, n, true);
    else if (h && h->hessian) h->hessian->evaluate(d_out, n, true);
    else ev(h).evaluate(d_out, n, true);
  });
}

int pam_fmm_evaluator_evaluate_gradient(pam_fmm_evaluator* h, double* out, int64_t n) {
  return guard([&] {
    if (h && (h->dipole || h->hessian)) pam::throw_runtime("use pam_fmm_evaluator_evaluate on this handle");
    ev(h).evaluate_gradient(out, n, false);
  });
}

int pam_fmm_evaluator_evaluate_gradient_device(pam_fmm_evaluator* h, double* d_out, int64_t n) {
  return guard([&] {
    if (h && (h->dipole || h->hessian)) pam::throw_runtime("use pam_fmm_evaluator_evaluate on this handle");
    ev(h).evaluate_gradient(d_out, n, true);
  });
}

int pam_fmm_evaluator_synchronize(pam_fmm_evaluator* h) {
  return guard([&] { PAM_HIP(hipStreamSynchronize(ev(h).stream())); });
}

int pam_fmm_evaluator_set_config(pam_fmm_evaluator* h, const pam_fmm_config* cfg) {
  return guard([&] {
    if (!cfg) pam::throw_invalid("cfg is null");
    pam::FmmConfig c;
    c.precision = cfg->precision;
    c.order = cfg->order;
    c.leaf_level = cfg->leaf_level;
    c.mode = cfg->mode;
    c.svd_tol = cfg->svd_tolerance;
    ev(h).set_config(c);
  });
}

int pam_fmm_evaluator_set_overlap(pam_fmm_evaluator* h, int overlap) {
  return guard([&] { ev(h).set_overlap(overlap != 0); });
}

int pam_fmm_evaluator_get_config(pam_fmm_evaluator* h, pam_fmm_config* cfg) {
  return guard([&] {
    const auto& c = ev(h).active_config();
    cfg->precision = c.precision;
    cfg->order = c.order;
    cfg->leaf_level = c.leaf_level;
    cfg->mode = c.mode;
    cfg->svd_tolerance = c.svd_tol;
  });
}

int pam_fmm_evaluator_stats(pam_fmm_evaluator* h, pam_fmm_stats* out) {
  return guard([&] {
    const auto& s = ev(h).stats();
    out->ms_tree = s.ms_tree;
    out->ms_upward = s.ms_upward;
    out->ms_m2l = s.ms_m2l;
    out->ms_l2l = s.ms_l2l;
    out->ms_l2p = s.ms_l2p;
    out->ms_p2p = s.ms_p2p;
    out->ms_total = s.ms_total;
    out->precision = s.precision;
    out->order = s.order;
    out->leaf_level = s.leaf_level;
    out->direct = s.direct;
    out->n_src_leaves = s.n_src_leaves;
    out->n_trg_leaves = s.n_trg_leaves;
    out->m2l_rank = s.m2l_rank;
  });
}

int pam_fmm_evaluator_count_work(pam_fmm_evaluator* h, int64_t* p2p_pairs, int64_t* m2l_pairs) {
  return guard([&] { ev(h).count_work(p2p_pairs, m2l_pairs); });
}

int pam_fmm_evaluator_tree_info(pam_fmm_evaluator* h, int which, int* level, double lo[3],
                                double* width, double* inv_cell, int64_t* n_points,
                                int64_t* n_leaves) {
  return guard([&] {
    ev(h).prepare_trees();
    const auto& t = which_tree(h, which);
    *level = t.level;
    for (int k = 0; k < 3; k++) lo[k] = t.box.lo[k];
    *width = t.box.width;
    *inv_cell = t.inv_cell;
    *n_points = t.n_points;
    *n_leaves = t.n_leaves();
  });
}

int pam_fmm_evaluator_export_tree(pam_fmm_evaluator* h, int which, uint64_t* point_codes,
                                  int32_t* perm, uint64_t* leaf_codes, int32_t* leaf_start) {
  return guard([&] {
    ev(h).prepare_trees();
    const auto& t = which_tree(h, which);
    int64_t n = t.n_points, nl = t.n_leaves();
    if (n > 0) {
      PAM_HIP(hipMemcpy(point_codes, t.point_codes.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost));
      PAM_HIP(hipMemcpy(perm, t.perm.get(), n * sizeof(int32_t), hipMemcpyDeviceToHost));
      PAM_HIP(hipMemcpy(leaf_codes, t.levels[t.level].codes.get(), nl * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    PAM_HIP(hipMemcpy(leaf_start, t.leaf_start.get(), (nl + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  });
}

int pam_fmm_evaluator_export_level(pam_fmm_evaluator* h, int which, int level, int64_t* n,
                                   uint64_t* codes) {
  return guard([&] {
    ev(h).prepare_trees();
    const auto& t = which_tree(h, which);
    if (level < 0 || level > t.level) pam::throw_invalid("level out of range");
    *n = t.levels[level].n;
    if (codes && *n > 0)
      PAM_HIP(hipMemcpy(codes, t.levels[level].codes.get(), *n * sizeof(uint64_t), hipMemcpyDeviceToHost));
  });
}

int pam_fmm_evaluator_export_p2p(pam_fmm_evaluator* h, int32_t* slots) {
  return guard([&] {
    ev(h).prepare_trees();
    int64_t nl = ev(h).target_tree().n_leaves();
    if (nl > 0)
      PAM_HIP(hipMemcpy(slots, ev(h).lists().p2p.get(), nl * 27 * sizeof(int32_t), hipMemcpyDeviceToHost));
  });
}

int pam_fmm_evaluator_export_m2l(pam_fmm_evaluator* h, int level, int32_t* slots) {
  return guard([&] {
    ev(h).prepare_trees();
    const auto& t = ev(h).target_tree();
    if (level < 2 || level > t.level) pam::throw_invalid("level out of range");
    int64_t n = t.levels[level].n;
    std::vector<int32_t> s(static_cast<size_t>(n) * pam::kM2LSlots);
    if (n > 0)
      PAM_HIP(hipMemcpy(s.data(), ev(h).lists().m2l[level].src.get(), s.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<uint64_t> codes(n);
    if (n > 0)
      PAM_HIP(hipMemcpy(codes.data(), t.levels[level].codes.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    const auto& vl = pam::octant_vlist_host();
    for (int64_t c = 0; c < n; c++) {
      for (int v = 0; v < 343; v++) slots[c * 343 + v] = -1;
      int oct = static_cast<int>(codes[c] & 7u);
      for (int k = 0; k < pam::kM2LSlots; k++) slots[c * 343 + vl[oct * pam::kM2LSlots + k]] = s[c * pam::kM2LSlots + k];
    }
  });
}

int pam_direct_evaluate(const pam_rbf* rbf, const double* src, int64_t n_src, const double* w,
                        const double* trg, int64_t n_trg, double* out) {
  return guard([&] {
    auto d = to_desc(rbf);
    pam::FmmEvaluator e(d, nullptr, nullptr, false);
    pam::FmmConfig c;
    c.mode = pam::kModeDirect;
    e.set_config(c);
    e.set_source_points(src, n_src, false);
    e.set_weights(w, n_src, false);
    e.set_target_points(trg, n_trg, false);
    e.evaluate(out, n_trg, false);
  });
}

}  // extern "C"
