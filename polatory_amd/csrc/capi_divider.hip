// C ABI of the domain divider (host only): DomainDivider(points, grad_points,
// point_indices, grad_point_indices, poly_point_indices) and its
// choose_coarse_points (domain_divider.hpp:30-125).
#include <memory>
#include <vector>

#include "capi_util.hpp"
#include "ras.hpp"

using pam::guard;

struct pam_domain_divider {
  std::vector<pam::Domain> domains;
  std::vector<double> pts, gpts;
  std::vector<int32_t> idx, gidx, poly;
};

namespace {

std::vector<int32_t> copy_idx(const int32_t* p, int64_t n, int64_t limit, const char* what) {
  if (n < 0) pam::throw_invalid("negative index count");
  if (n > 0 && !p) pam::throw_invalid("null argument");
  std::vector<int32_t> v(p, p + n);
  for (int32_t i : v)
    if (i < 0 || i >= limit) pam::throw_invalid(std::string(what) + " index out of range");
  return v;
}

}  // namespace

int pam_domain_divider_create(const double* points, int64_t mu, const double* grad_points, int64_t sigma,
                              const int32_t* point_idcs, int64_t n_idx, const int32_t* grad_idcs, int64_t n_gidx,
                              const int32_t* poly_idcs, int64_t n_poly, pam_domain_divider** out) {
  return guard([&] {
    if (!out) pam::throw_invalid("null argument");
    if (mu < 0 || sigma < 0) pam::throw_invalid("negative point count");
    if ((mu > 0 && !points) || (sigma > 0 && !grad_points)) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_domain_divider>();
    h->pts.assign(points, points + 3 * mu);
    h->gpts.assign(grad_points, grad_points + 3 * sigma);
    h->idx = copy_idx(point_idcs, n_idx, mu, "point");
    h->gidx = copy_idx(grad_idcs, n_gidx, sigma, "gradient point");
    h->poly = copy_idx(poly_idcs, n_poly, mu, "polynomial point");
    if (h->gidx.empty()) {
      pam::DomainDivider div(h->pts, h->idx, h->poly);
      h->domains = std::move(div.domains());
    } else {
      h->domains = pam::divide_mixed(h->pts, h->gpts, h->idx, h->gidx, h->poly);
    }
    *out = h.release();
  });
}

int pam_domain_divider_destroy(pam_domain_divider* h) {
  return guard([&] { delete h; });
}

int pam_domain_divider_num_domains(const pam_domain_divider* h, int64_t* n) {
  return guard([&] {
    if (!h || !n) pam::throw_invalid("null argument");
    *n = static_cast<int64_t>(h->domains.size());
  });
}

int pam_domain_divider_domain(const pam_domain_divider* h, int64_t d, int32_t* idx, uint8_t* inner, int64_t* n,
                              int32_t* gidx, uint8_t* ginner, int64_t* ng) {
  return guard([&] {
    if (!h || !n || !ng) pam::throw_invalid("null argument");
    if (d < 0 || d >= static_cast<int64_t>(h->domains.size())) pam::throw_invalid("domain index out of range");
    const pam::Domain& dm = h->domains[d];
    *n = static_cast<int64_t>(dm.idx.size());
    *ng = static_cast<int64_t>(dm.gidx.size());
    if (idx) std::copy(dm.idx.begin(), dm.idx.end(), idx);
    if (inner) std::copy(dm.inner.begin(), dm.inner.end(), inner);
    if (gidx) std::copy(dm.gidx.begin(), dm.gidx.end(), gidx);
    if (ginner) std::copy(dm.ginner.begin(), dm.ginner.end(), ginner);
  });
}

int pam_domain_divider_choose_coarse_points(const pam_domain_divider* h, int64_t n_coarse, int32_t* idx, int64_t* n,
                                            int32_t* gidx, int64_t* ng) {
  return guard([&] {
    if (!h || !n || !ng) pam::throw_invalid("null argument");
    if (n_coarse < 0) pam::throw_invalid("negative coarse size");
    std::vector<int32_t> a, b;
    if (h->gidx.empty()) {
      a = pam::choose_coarse_points(h->pts, h->idx, h->poly, n_coarse);
    } else {
      auto r = pam::choose_coarse_points_mixed(h->pts, h->gpts, h->idx, h->gidx, h->poly, n_coarse);
      a = std::move(r.first);
      b = std::move(r.second);
    }
    const bool query = !idx && !gidx;
    if (!query && (static_cast<int64_t>(a.size()) > *n || static_cast<int64_t>(b.size()) > *ng))
      pam::throw_invalid("output arrays too small");
    *n = static_cast<int64_t>(a.size());
    *ng = static_cast<int64_t>(b.size());
    if (idx) std::copy(a.begin(), a.end(), idx);
    if (gidx) std::copy(b.begin(), b.end(), gidx);
  });
}
