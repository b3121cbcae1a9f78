// extern "C" entry points of the request loop (serve.hpp).
#include <memory>

#include "capi_handles.hpp"
#include "serve.hpp"

struct pam_server {
  std::unique_ptr<pam::Server> impl;
};

namespace {

using pam::guard;

pam::Server& srv(pam_server* h) {
  if (!h || !h->impl) pam::throw_invalid("null server handle");
  return *h->impl;
}

}  // namespace

extern "C" {

int pam_server_options_default(pam_server_options* opt) {
  return guard([&] {
    if (!opt) pam::throw_invalid("null options");
    pam::ServeOptions d;
    opt->n_slots = d.n_slots;
    opt->hbm_budget = d.hbm_budget;
    opt->rel_accuracy = d.rel_accuracy;
  });
}

int pam_server_create(const pam_server_options* opt, pam_server** out) {
  return guard([&] {
    if (!out) pam::throw_invalid("null output handle");
    *out = nullptr;
    pam::ServeOptions o;
    if (opt) {
      o.n_slots = opt->n_slots;
      o.hbm_budget = opt->hbm_budget;
      o.rel_accuracy = opt->rel_accuracy;
    }
    auto h = std::make_unique<pam_server>();
    h->impl = std::make_unique<pam::Server>(o);
    *out = h.release();
  });
}

int pam_server_destroy(pam_server* h) {
  return guard([&] {
    if (!h) return;
    if (h->impl) h->impl->drain();
    delete h;
  });
}

int pam_server_add_model(pam_server* h, const pam_interpolant* fitted, const double lo[3],
                         const double hi[3], const pam_fmm_config* cfg, int* model_id) {
  return guard([&] {
    if (!fitted || !fitted->impl) pam::throw_invalid("null interpolant handle");
    if (!model_id) pam::throw_invalid("null model id");
    const pam::Interpolant& ip = *fitted->impl;
    if (!ip.fitted()) pam::throw_runtime("the interpolant is not fitted");
    if (!ip.grad_centers().empty())
      throw pam::Error(pam::kNotImplemented, "serving interpolants with gradient centres is not implemented");
    pam::FmmConfig c;
    if (cfg) c = pam::to_cfg(*cfg);
    int64_t n = static_cast<int64_t>(ip.centers().size() / 3);
    *model_id = srv(h).add_model(ip.model(), ip.centers().data(), n, ip.weights().data(), lo, hi,
                                 cfg ? &c : nullptr);
  });
}

int pam_server_model_config(pam_server* h, int model_id, pam_fmm_config* cfg, int64_t* resident_bytes) {
  return guard([&] {
    auto& s = srv(h);
    if (cfg) *cfg = pam::from_cfg(s.model_config(model_id));
    if (resident_bytes) *resident_bytes = s.model_bytes(model_id);
  });
}

int pam_server_submit(pam_server* h, int model_id, const double* points, int64_t n, double* out,
                      int64_t* request_id) {
  return guard([&] {
    if (!request_id) pam::throw_invalid("null request id");
    *request_id = srv(h).submit(model_id, points, n, out);
  });
}

int pam_server_wait(pam_server* h, int64_t request_id) {
  return guard([&] { srv(h).wait(request_id); });
}

int pam_server_test(pam_server* h, int64_t request_id, int* done) {
  return guard([&] {
    if (!done) pam::throw_invalid("null output");
    *done = srv(h).done(request_id) ? 1 : 0;
  });
}

int pam_server_drain(pam_server* h) {
  return guard([&] { srv(h).drain(); });
}

int pam_server_request_times(pam_server* h, int64_t request_id, double* queue_ms, double* total_ms) {
  return guard([&] {
    if (!queue_ms || !total_ms) pam::throw_invalid("null output");
    srv(h).request_times(request_id, queue_ms, total_ms);
  });
}

int pam_server_get_stats(pam_server* h, pam_server_stats* out) {
  return guard([&] {
    if (!out) pam::throw_invalid("null output");
    pam::ServeStats s = srv(h).stats();
    out->submitted = s.submitted;
    out->completed = s.completed;
    out->failed = s.failed;
    out->points = s.points;
    out->p50_ms = s.p50_ms;
    out->p99_ms = s.p99_ms;
    out->mean_ms = s.mean_ms;
    out->max_ms = s.max_ms;
    out->mean_queue_ms = s.mean_queue_ms;
    out->points_per_s = s.points_per_s;
    out->wall_s = s.wall_s;
    out->max_in_flight = s.max_in_flight;
    out->resident_bytes = s.resident_bytes;
    out->budget = s.budget;
    out->peak_admitted_bytes = s.peak_admitted_bytes;
    out->batches = s.batches;
    out->batched_requests = s.batched_requests;
    out->max_batch = s.max_batch;
    out->fmm_batches = s.fmm_batches;
    out->fmm_batched_requests = s.fmm_batched_requests;
    out->max_fmm_batch = s.max_fmm_batch;
  });
}

int pam_server_reset_stats(pam_server* h) {
  return guard([&] { srv(h).reset_stats(); });
}

}  // extern "C"
