// extern "C" entry points of the interpolation layer: model, interpolant fit /
// evaluate, and the Krylov / RAS seams.
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "capi_handles.hpp"
#include "hermite.hpp"
#include "incremental_fitter.hpp"
#include "model_io.hpp"
#include "ras.hpp"

struct pam_operator {
  hipStream_t s = nullptr;
  pam::Model model;
  pam::DevBuf<double> pts, gpts;
  std::unique_ptr<pam::LinearOperator> impl;
  pam::DevBuf<double> x, y;  // host-pointer applies
  ~pam_operator() {
    impl.reset();
    if (s) {
      (void)hipStreamSynchronize(s);
      pam::stream_release(s);
    }
  }
};

struct pam_ras {
  hipStream_t s = nullptr;
  pam::Model model;
  std::unique_ptr<pam::RasPreconditioner> impl;
  ~pam_ras() {
    impl.reset();
    if (s) {
      (void)hipStreamSynchronize(s);
      pam::stream_release(s);
    }
  }
};

namespace {

using pam::guard;

using pam::from_cfg;
using pam::to_cfg;

pam::FitOptions default_fit_options() {
  pam::FitOptions o;
  // Inner operator: mixed-precision FMM (FP32 P2P, FP32-MFMA M2L); the FP64
  // direct sum below ~2^26 pairs.
  o.op_cfg.precision = pam::kMixed;
  o.op_cfg.order = 8;
  o.op_cfg.mode = pam::kModeAuto;
  // Final residual recomputed in FP64: FP64 FMM of order 10, direct when small.
  o.res_cfg.precision = pam::kFp64;
  o.res_cfg.order = 10;
  o.res_cfg.mode = pam::kModeAuto;
  o.level_cfg = pam::FmmConfig{};
  return o;
}

pam::FitOptions to_opts(const pam_fit_options* o) {
  if (!o) return default_fit_options();
  pam::FitOptions r;
  r.op_cfg = to_cfg(o->op);
  r.res_cfg = to_cfg(o->residual);
  r.level_cfg = to_cfg(o->level);
  r.use_level_cfg = o->use_level_config != 0;
  r.use_preconditioner = o->use_preconditioner != 0;
  r.refine_rtol = o->refine_rtol;
  r.use_graphs = o->use_graphs != 0;
  r.keep_unconverged = o->keep_unconverged != 0;
  r.verbose = o->verbose;
  r.tune_solves = o->tune_solves != 0;
  if (o->ras_fp32_factors < -1 || o->ras_fp32_factors > 1) pam::throw_invalid("ras_fp32_factors must be -1, 0 or 1");
  r.ras_fp32_factors = o->ras_fp32_factors;
  r.tune_factor = o->tune_factor != 0;
  if (o->factor_panel != -1 && o->factor_panel != 0 && o->factor_panel != 32 && o->factor_panel != 64)
    pam::throw_invalid("factor_panel must be -1, 0, 32 or 64");
  r.factor_panel = o->factor_panel;
  r.cpd2_auto = o->cpd2_auto != 0;
  return r;
}

pam::Interpolant& ip(const pam_interpolant* h) {
  if (!h || !h->impl) pam::throw_invalid("null interpolant handle");
  return *h->impl;
}

// Dense row-major operator on the device (FGMRES tests).
__global__ void dense_gemv_kernel(const double* __restrict__ a, const double* __restrict__ x, int64_t n,
                                  double* __restrict__ y) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t j = 0; j < n; j++) acc += a[i * n + j] * x[j];
  y[i] = acc;
}

class DenseOperator : public pam::LinearOperator {
 public:
  DenseOperator(const double* h_a, int64_t n, hipStream_t s) : n_(n), s_(s) {
    a_.upload(h_a, static_cast<size_t>(n) * n, s);
  }
  void apply(const double* d_x, double* d_y) override {
    dense_gemv_kernel<<<pam::ceil_div(n_, 128), 128, 0, s_>>>(a_.get(), d_x, n_, d_y);
    PAM_LAUNCH_CHECK();
  }
  int64_t size() const override { return n_; }

 private:
  int64_t n_;
  hipStream_t s_;
  pam::DevBuf<double> a_;
};

}  // namespace

extern "C" {

int pam_model_create(const pam_rbf* rbfs, int n_rbfs, int poly_degree, double nugget, pam_model** out) {
  return guard([&] {
    if (!out) pam::throw_invalid("out is null");
    if (n_rbfs <= 0) pam::throw_invalid("rbfs must not be empty");
    auto m = std::make_unique<pam_model>();
    for (int i = 0; i < n_rbfs; i++) m->m.rbfs.push_back(pam::to_desc(&rbfs[i]));
    int lo = m->m.cpd_order() - 1;
    m->m.poly_degree = poly_degree == PAM_POLY_DEGREE_MIN ? lo : poly_degree;
    m->m.nugget = nugget;
    m->m.validate();
    *out = m.release();
  });
}

int pam_model_destroy(pam_model* m) {
  return guard([&] { delete m; });
}

int pam_model_poly_degree(const pam_model* m, int* degree) {
  return guard([&] {
    if (!m) pam::throw_invalid("null model");
    *degree = m->m.poly_degree;
  });
}

int pam_model_poly_basis_size(const pam_model* m, int* l) {
  return guard([&] {
    if (!m) pam::throw_invalid("null model");
    *l = m->m.l();
  });
}

int pam_model_num_rbfs(const pam_model* m, int* n) {
  return guard([&] {
    if (!m || !n) pam::throw_invalid("null argument");
    *n = static_cast<int>(m->m.rbfs.size());
  });
}

int pam_model_rbf(const pam_model* m, int i, pam_rbf* out) {
  return guard([&] {
    if (!m || !out) pam::throw_invalid("null argument");
    if (i < 0 || i >= static_cast<int>(m->m.rbfs.size())) pam::throw_invalid("rbf index out of range");
    const pam::RbfDesc& r = m->m.rbfs[i];
    out->kind = r.kind;
    out->params[0] = r.p0;
    out->params[1] = r.p1;
    for (int k = 0; k < 9; k++) out->aniso[k] = r.aniso[k];
  });
}

int pam_model_nugget(const pam_model* m, double* nugget) {
  return guard([&] {
    if (!m || !nugget) pam::throw_invalid("null argument");
    *nugget = m->m.nugget;
  });
}

int pam_model_save(const pam_model* m, const char* filename) {
  return guard([&] {
    if (!m || !filename) pam::throw_invalid("null argument");
    pam::save_model(filename, m->m);
  });
}

int pam_model_load(const char* filename, pam_model** out) {
  return guard([&] {
    if (!filename || !out) pam::throw_invalid("null argument");
    auto m = std::make_unique<pam_model>();
    m->m = pam::load_model(filename);
    *out = m.release();
  });
}

int pam_fit_options_default(pam_fit_options* opt) {
  return guard([&] {
    auto o = default_fit_options();
    opt->op = from_cfg(o.op_cfg);
    opt->residual = from_cfg(o.res_cfg);
    opt->level = from_cfg(o.level_cfg);
    opt->use_level_config = 0;
    opt->use_preconditioner = 1;
    opt->refine_rtol = o.refine_rtol;
    opt->use_graphs = o.use_graphs ? 1 : 0;
    opt->keep_unconverged = o.keep_unconverged ? 1 : 0;
    opt->verbose = 0;
    opt->tune_solves = o.tune_solves ? 1 : 0;
    opt->ras_fp32_factors = o.ras_fp32_factors;
    opt->tune_factor = o.tune_factor ? 1 : 0;
    opt->factor_panel = o.factor_panel;
    opt->cpd2_auto = o.cpd2_auto ? 1 : 0;
  });
}

int pam_interpolant_create(const pam_model* model, pam_interpolant** out) {
  return guard([&] {
    if (!model || !out) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_interpolant>();
    h->impl = std::make_unique<pam::Interpolant>(model->m);
    *out = h.release();
  });
}

int pam_interpolant_destroy(pam_interpolant* h) {
  return guard([&] { delete h; });
}

int pam_interpolant_fit(pam_interpolant* h, const double* points, int64_t n, const double* values,
                        double tolerance, int max_iter, double accuracy, const pam_fit_options* opt,
                        const pam_interpolant* initial) {
  return guard([&] {
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && (!points || !values)) pam::throw_invalid("null argument");
    std::vector<double> p(points, points + 3 * n), v(values, values + n);
    ip(h).fit(p, v, tolerance, max_iter, accuracy, to_opts(opt), initial ? &ip(initial) : nullptr);
  });
}

int pam_interpolant_fit_hermite(pam_interpolant* h, const double* points, int64_t mu, const double* grad_points,
                                int64_t sigma, const double* values, double tolerance, double grad_tolerance,
                                int max_iter, double accuracy, double grad_accuracy, const pam_fit_options* opt,
                                const pam_interpolant* initial) {
  return guard([&] {
    if (mu < 0 || sigma < 0) pam::throw_invalid("negative point count");
    if ((mu > 0 && !points) || (sigma > 0 && !grad_points) || (mu + sigma > 0 && !values))
      pam::throw_invalid("null argument");
    if (!(accuracy > 0.0) || !(grad_accuracy > 0.0)) pam::throw_invalid("accuracy must be positive");
    std::vector<double> p(points, points + 3 * mu), g(grad_points, grad_points + 3 * sigma),
        v(values, values + mu + 3 * sigma);
    ip(h).fit_hermite(p, g, v, tolerance, grad_tolerance, max_iter, to_opts(opt), initial ? &ip(initial) : nullptr);
  });
}

int pam_interpolant_fit_incrementally_hermite(pam_interpolant* h, const double* points, int64_t mu,
                                              const double* grad_points, int64_t sigma, const double* values,
                                              double tolerance, double grad_tolerance, int max_iter,
                                              double accuracy, double grad_accuracy, const pam_fit_options* opt) {
  return guard([&] {
    if (mu < 0 || sigma < 0) pam::throw_invalid("negative point count");
    if ((mu > 0 && !points) || (sigma > 0 && !grad_points) || (mu + sigma > 0 && !values))
      pam::throw_invalid("null argument");
    if (!(accuracy > 0.0) || !(grad_accuracy > 0.0)) pam::throw_invalid("accuracy must be positive");
    std::vector<double> p(points, points + 3 * mu), g(grad_points, grad_points + 3 * sigma),
        v(values, values + mu + 3 * sigma);
    ip(h).fit_incrementally_hermite(p, g, v, tolerance, grad_tolerance, max_iter, to_opts(opt));
  });
}

int pam_interpolant_evaluate_hermite(pam_interpolant* h, const double* points, int64_t n,
                                     const double* grad_points, int64_t ng, double* out) {
  return guard([&] {
    if (n < 0 || ng < 0) pam::throw_invalid("negative point count");
    if ((n > 0 && !points) || (ng > 0 && !grad_points) || (n + ng > 0 && !out)) pam::throw_invalid("null argument");
    ip(h).evaluate_hermite(points, n, grad_points, ng, out, false);
  });
}

int pam_interpolant_num_grad_centers(const pam_interpolant* h, int64_t* n) {
  return guard([&] { *n = static_cast<int64_t>(ip(h).grad_centers().size() / 3); });
}

int pam_interpolant_grad_centers(const pam_interpolant* h, double* out) {
  return guard([&] {
    const auto& c = ip(h).grad_centers();
    std::copy(c.begin(), c.end(), out);
  });
}

int pam_interpolant_fit_incrementally(pam_interpolant* h, const double* points, int64_t n,
                                      const double* values, double tolerance, int max_iter,
                                      double accuracy, const pam_fit_options* opt) {
  return guard([&] {
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && (!points || !values)) pam::throw_invalid("null argument");
    std::vector<double> p(points, points + 3 * n), v(values, values + n);
    ip(h).fit_incrementally(p, v, tolerance, max_iter, accuracy, to_opts(opt));
  });
}

int pam_interpolant_fit_inequality(pam_interpolant* h, const double* points, int64_t n, const double* values,
                                   const double* values_lb, const double* values_ub, double tolerance,
                                   int max_iter, double accuracy, const pam_fit_options* opt,
                                   const pam_interpolant* initial) {
  return guard([&] {
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && (!points || !values || !values_lb || !values_ub)) pam::throw_invalid("null argument");
    std::vector<double> p(points, points + 3 * n), v(values, values + n), lo(values_lb, values_lb + n),
        hi(values_ub, values_ub + n);
    ip(h).fit_inequality(p, v, lo, hi, tolerance, max_iter, accuracy, to_opts(opt), initial ? &ip(initial) : nullptr);
  });
}

int pam_cross_validate(const pam_model* model, const double* points, int64_t n, const double* values,
                       const int32_t* set_ids, double tolerance, int max_iter, double accuracy,
                       const pam_fit_options* opt, double* out) {
  return guard([&] {
    if (!model) pam::throw_invalid("null model");
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && (!points || !values || !set_ids || !out)) pam::throw_invalid("null argument");
    std::vector<double> p(points, points + 3 * n), v(values, values + n);
    std::vector<int32_t> ids(set_ids, set_ids + n);
    std::vector<double> pred = pam::cross_validate(model->m, p, v, ids, tolerance, max_iter, accuracy, to_opts(opt));
    std::copy(pred.begin(), pred.end(), out);
  });
}

int pam_distance_filter(const double* points, int64_t n, double distance, const int64_t* indices,
                        int64_t n_idx, int64_t* out, int64_t* n_out) {
  return guard([&] {
    if (n < 0 || n_idx < 0) pam::throw_invalid("negative count");
    if ((n > 0 && !points) || (n_idx > 0 && (!indices || !out)) || !n_out) pam::throw_invalid("null argument");
    std::vector<double> p(points, points + 3 * n);
    std::vector<int64_t> idx(indices, indices + n_idx);
    std::vector<int64_t> kept = pam::DistanceFilter(p).filter(distance, idx);
    std::copy(kept.begin(), kept.end(), out);
    *n_out = static_cast<int64_t>(kept.size());
  });
}

int pam_interpolant_evaluate(pam_interpolant* h, const double* points, int64_t n, double accuracy,
                             double* out) {
  return guard([&] { ip(h).evaluate(points, n, accuracy, out, false); });
}

int pam_interpolant_evaluate_device(pam_interpolant* h, const double* d_points, int64_t n,
                                    double accuracy, double* d_out) {
  return guard([&] { ip(h).evaluate(d_points, n, accuracy, d_out, true); });
}

int pam_interpolant_evaluate_gradient(pam_interpolant* h, const double* points, int64_t n, double accuracy,
                                      double* out) {
  return guard([&] {
    if (n > 0 && (!points || !out)) pam::throw_invalid("null argument");
    ip(h).evaluate_gradient(points, n, accuracy, out, false);
  });
}

int pam_interpolant_set_evaluation_bbox(pam_interpolant* h, const double* bbox_min, const double* bbox_max,
                                        double accuracy, double grad_accuracy) {
  return guard([&] {
    if (!bbox_min || !bbox_max) pam::throw_invalid("null argument");
    ip(h).set_evaluation_bbox(bbox_min, bbox_max, accuracy, grad_accuracy);
  });
}

int pam_interpolant_evaluate_impl(pam_interpolant* h, const double* points, int64_t n, double* out) {
  return guard([&] {
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && (!points || !out)) pam::throw_invalid("null argument");
    ip(h).evaluate_impl(points, n, out, false);
  });
}

int pam_interpolant_evaluate_impl_device(pam_interpolant* h, const double* d_points, int64_t n, double* d_out) {
  return guard([&] {
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && (!d_points || !d_out)) pam::throw_invalid("null argument");
    ip(h).evaluate_impl(d_points, n, d_out, true);
  });
}

int pam_interpolant_set_eval_config(pam_interpolant* h, const pam_fmm_config* cfg) {
  return guard([&] {
    if (cfg) {
      pam::FmmConfig c = to_cfg(*cfg);
      ip(h).set_eval_config(&c);
    } else {
      ip(h).set_eval_config(nullptr);
    }
  });
}

int pam_interpolant_num_centers(const pam_interpolant* h, int64_t* n) {
  return guard([&] { *n = static_cast<int64_t>(ip(h).centers().size() / 3); });
}

int pam_interpolant_centers(const pam_interpolant* h, double* out) {
  return guard([&] {
    const auto& c = ip(h).centers();
    std::copy(c.begin(), c.end(), out);
  });
}

int pam_interpolant_weights(const pam_interpolant* h, double* out) {
  return guard([&] {
    if (!ip(h).fitted()) pam::throw_runtime("the interpolant is not fitted");
    const auto& w = ip(h).weights();
    std::copy(w.begin(), w.end(), out);
  });
}

int pam_interpolant_set_fitted(pam_interpolant* h, const double* centers, int64_t n, const double* weights) {
  return guard([&] {
    int l = ip(h).model().l();
    std::vector<double> c(centers, centers + 3 * n), w(weights, weights + n + l);
    ip(h).set_fitted(c, w);
  });
}

int pam_interpolant_save(const pam_interpolant* h, const char* filename) {
  return guard([&] {
    if (!h || !filename) pam::throw_invalid("null argument");
    pam::save_interpolant(filename, *h->impl);
  });
}

int pam_interpolant_load(const char* filename, pam_interpolant** out) {
  return guard([&] {
    if (!filename || !out) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_interpolant>();
    h->impl = pam::load_interpolant(filename);
    *out = h.release();
  });
}

int pam_interpolant_model(const pam_interpolant* h, pam_model** out) {
  return guard([&] {
    if (!h || !out) pam::throw_invalid("null argument");
    auto m = std::make_unique<pam_model>();
    m->m = h->impl->model();
    *out = m.release();
  });
}

int pam_interpolant_fitted(const pam_interpolant* h, int* fitted) {
  return guard([&] {
    if (!h || !fitted) pam::throw_invalid("null argument");
    *fitted = h->impl->fitted() ? 1 : 0;
  });
}

int pam_interpolant_fit_stats(const pam_interpolant* h, pam_fit_stats* out) {
  return guard([&] {
    const auto& s = ip(h).fit_stats();
    out->iterations = s.iterations;
    out->restarts = s.restarts;
    out->n_levels = s.n_levels;
    out->residual = s.residual;
    out->seconds = s.seconds;
    out->setup_seconds = s.setup_seconds;
    out->precond_setup_seconds = s.precond_setup_seconds;
    out->factor_bytes = s.factor_bytes;
    out->graph_replays = s.graph_replays;
    out->escalations = s.escalations;
    out->converged = s.converged ? 1 : 0;
    out->rounds = s.rounds;
    out->grad_residual = s.grad_residual;
  });
}

int pam_interpolant_residual_history(const pam_interpolant* h, double* out, int64_t* n) {
  return guard([&] {
    const auto& hist = ip(h).fit_stats().history;
    *n = static_cast<int64_t>(hist.size());
    if (out) std::copy(hist.begin(), hist.end(), out);
  });
}

int pam_fgmres_dense(const double* a, const double* m, int64_t n, const double* rhs, const double* x0,
                     int n_iter, double* x, double* resid_hist, int* iters) {
  return guard([&] {
    hipStream_t s = pam::stream_acquire();
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() {
        (void)hipStreamSynchronize(s);
        pam::stream_release(s);
      }
    } sg{s};
    DenseOperator op(a, n, s);
    std::unique_ptr<DenseOperator> pc;
    if (m) pc = std::make_unique<DenseOperator>(m, n, s);
    pam::DevBuf<double> d_rhs, d_x0, d_x(n);
    d_rhs.upload(rhs, n, s);
    pam::Fgmres solver(op, d_rhs.get(), n, n_iter, s);
    if (x0) {
      d_x0.upload(x0, n, s);
      solver.set_initial_solution(d_x0.get());
    }
    solver.set_right_preconditioner(pc.get());
    solver.setup();
    resid_hist[0] = solver.absolute_residual();
    for (int i = 0; i < n_iter; i++) {
      solver.iterate_process();
      resid_hist[i + 1] = solver.absolute_residual();
    }
    solver.solution_vector(d_x.get());
    d_x.download(x, n, s);
    PAM_HIP(hipStreamSynchronize(s));
    *iters = solver.iteration_count();
  });
}

int pam_ras_create(const pam_model* model, const double* points, int64_t n, pam_ras** out) {
  return guard([&] {
    if (!model || !out) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_ras>();
    h->s = pam::stream_acquire();
    h->model = model->m;
    std::vector<double> p(points, points + 3 * n);
    h->impl = std::make_unique<pam::RasPreconditioner>(h->model, p, h->s, pam::RasOptions{});
    *out = h.release();
  });
}

int pam_ras_create_with_options(const pam_model* model, const double* points, int64_t n, const pam_fit_options* opt,
                                pam_ras** out) {
  return guard([&] {
    if (!model || !out || !opt) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_ras>();
    h->s = pam::stream_acquire();
    h->model = model->m;
    pam::FitOptions fo = to_opts(opt);
    pam::resolve_fit_options(h->model, fo);
    std::vector<double> p(points, points + 3 * n);
    h->impl = std::make_unique<pam::RasPreconditioner>(h->model, p, h->s, pam::fit_ras_options(h->model, fo, n));
    *out = h.release();
  });
}

int pam_ras_create_hermite(const pam_model* model, const double* points, int64_t mu, const double* grad_points,
                           int64_t sigma, const pam_fit_options* opt, pam_ras** out) {
  return guard([&] {
    if (!model || !out) pam::throw_invalid("null argument");
    if (mu < 0 || sigma < 0) pam::throw_invalid("negative point count");
    if ((mu > 0 && !points) || (sigma > 0 && !grad_points)) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_ras>();
    h->s = pam::stream_acquire();
    h->model = model->m;
    pam::FitOptions fo = to_opts(opt);
    pam::resolve_fit_options(h->model, fo);
    pam::RasOptions ro = pam::fit_ras_options(h->model, fo, mu + 3 * sigma);
    ro.fp32_factors = false;
    if (!ro.use_level_cfg || ro.level_cfg.precision != pam::kFp64) {
      ro.level_cfg = pam::FmmConfig{};
      ro.level_cfg.precision = pam::kFp64;
      ro.level_cfg.order = 10;
      ro.use_level_cfg = true;
    }
    std::vector<double> p(points, points + 3 * mu), g(grad_points, grad_points + 3 * sigma);
    h->impl = std::make_unique<pam::RasPreconditioner>(h->model, p, g, h->s, ro);
    *out = h.release();
  });
}

int pam_ras_destroy(pam_ras* h) {
  return guard([&] { delete h; });
}

int pam_ras_num_levels(const pam_ras* h, int* n_levels) {
  return guard([&] { *n_levels = h->impl->n_levels(); });
}

int pam_ras_level_points(const pam_ras* h, int level, int32_t* out, int64_t* n) {
  return guard([&] {
    const auto& lp = h->impl->level_points();
    if (level < 0 || level >= static_cast<int>(lp.size())) pam::throw_invalid("level out of range");
    *n = static_cast<int64_t>(lp[level].size());
    if (out) std::copy(lp[level].begin(), lp[level].end(), out);
  });
}

int pam_ras_num_domains(const pam_ras* h, int level, int* n) {
  return guard([&] { *n = static_cast<int>(h->impl->level_domains(level).size()); });
}

int pam_ras_domain(const pam_ras* h, int level, int dom, int32_t* idx, uint8_t* inner, int64_t* n) {
  return guard([&] {
    const auto& d = h->impl->level_domains(level).at(dom);
    *n = static_cast<int64_t>(d.idx.size());
    if (idx) std::copy(d.idx.begin(), d.idx.end(), idx);
    if (inner) std::copy(d.inner.begin(), d.inner.end(), inner);
  });
}

int pam_ras_apply_device(pam_ras* h, const double* d_v, double* d_out) {
  return guard([&] {
    if (!h || !h->impl) pam::throw_invalid("null preconditioner handle");
    if (h->impl->size() > 0 && (!d_v || !d_out)) pam::throw_invalid("null argument");
    h->impl->apply(d_v, d_out);
  });
}

int pam_ras_size(const pam_ras* h, int64_t* size) {
  return guard([&] {
    if (!h || !h->impl || !size) pam::throw_invalid("null argument");
    *size = h->impl->size();
  });
}

int pam_ras_synchronize(pam_ras* h) {
  return guard([&] {
    if (!h) pam::throw_invalid("null preconditioner handle");
    PAM_HIP(hipStreamSynchronize(h->s));
  });
}

int pam_operator_create(const pam_model* model, const double* points, int64_t n, const pam_fmm_config* cfg,
                        pam_operator** out) {
  return guard([&] {
    if (!model || !out) pam::throw_invalid("null argument");
    if (n < 0) pam::throw_invalid("negative point count");
    if (n > 0 && !points) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_operator>();
    h->s = pam::stream_acquire();
    h->model = model->m;
    h->pts.upload(points, static_cast<size_t>(3 * n), h->s);
    pam::FmmConfig c = cfg ? to_cfg(*cfg) : default_fit_options().op_cfg;
    h->impl = std::make_unique<pam::Operator>(h->model, h->pts.get(), n, c, h->s);
    PAM_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
  });
}

int pam_operator_create_hermite(const pam_model* model, const double* points, int64_t mu,
                                const double* grad_points, int64_t sigma, const pam_fmm_config* cfg,
                                pam_operator** out) {
  return guard([&] {
    if (!model || !out) pam::throw_invalid("null argument");
    if (mu < 0 || sigma < 0) pam::throw_invalid("negative point count");
    if ((mu > 0 && !points) || (sigma > 0 && !grad_points)) pam::throw_invalid("null argument");
    auto h = std::make_unique<pam_operator>();
    h->s = pam::stream_acquire();
    h->model = model->m;
    h->pts.upload(points, static_cast<size_t>(3 * mu), h->s);
    h->gpts.upload(grad_points, static_cast<size_t>(3 * sigma), h->s);
    pam::FmmConfig c;
    if (cfg) {
      c = to_cfg(*cfg);
    } else {
      c.precision = pam::kFp64;
      c.order = 10;
    }
    h->impl = std::make_unique<pam::HermiteOperator>(h->model, h->pts.get(), mu, h->gpts.get(), sigma, c, h->s);
    PAM_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
  });
}

int pam_operator_destroy(pam_operator* h) {
  return guard([&] { delete h; });
}

int pam_operator_size(const pam_operator* h, int64_t* size) {
  return guard([&] {
    if (!h || !h->impl || !size) pam::throw_invalid("null argument");
    *size = h->impl->size();
  });
}

int pam_operator_apply_device(pam_operator* h, const double* d_x, double* d_y) {
  return guard([&] {
    if (!h || !h->impl) pam::throw_invalid("null operator handle");
    if (h->impl->size() > 0 && (!d_x || !d_y)) pam::throw_invalid("null argument");
    h->impl->apply(d_x, d_y);
  });
}

int pam_operator_apply(pam_operator* h, const double* x, double* y) {
  return guard([&] {
    if (!h || !h->impl) pam::throw_invalid("null operator handle");
    const int64_t m = h->impl->size();
    if (m > 0 && (!x || !y)) pam::throw_invalid("null argument");
    h->x.upload(x, static_cast<size_t>(m), h->s);
    h->y.resize(static_cast<size_t>(m));
    h->impl->apply(h->x.get(), h->y.get());
    h->y.download(y, static_cast<size_t>(m), h->s);
    PAM_HIP(hipStreamSynchronize(h->s));
  });
}

int pam_operator_synchronize(pam_operator* h) {
  return guard([&] {
    if (!h) pam::throw_invalid("null operator handle");
    PAM_HIP(hipStreamSynchronize(h->s));
  });
}

int pam_ras_apply(pam_ras* h, const double* v, double* out) {
  return guard([&] {
    int64_t m = h->impl->size();
    pam::DevBuf<double> dv, dout(m);
    dv.upload(v, m, h->s);
    h->impl->apply(dv.get(), dout.get());
    dout.download(out, m, h->s);
    PAM_HIP(hipStreamSynchronize(h->s));
  });
}

}  // extern "C"
