// Continuous-batching request loop over resident fitted models.
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

#include "device_api.hpp"
#include "fmm.hpp"

namespace pam {

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  // nearest rank
  size_t k = static_cast<size_t>(std::ceil(q * static_cast<double>(v.size())));
  k = std::clamp<size_t>(k, 1, v.size());
  return v[k - 1];
}

}  // namespace

struct Server::ModelEntry {
  Model model;
  int64_t n = 0;
  int level = 0;
  FmmConfig cfg;
  int64_t bytes = 0;
  hipStream_t s = nullptr;
  DevBuf<double> d_centers, d_weights, d_coeffs;
  std::vector<std::unique_ptr<FmmEvaluator>> owners;  // one per rbf, source side prepared
  ~ModelEntry() {
    owners.clear();
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

struct Server::Request {
  int64_t id = 0;
  ModelEntry* model = nullptr;
  const double* pts = nullptr;
  int64_t n = 0;
  double* out = nullptr;
  int64_t est = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  bool done = false;
  int code = kOk;
  std::string err;
};

struct Server::Slot {
  hipStream_t s = nullptr;
  // Resources of one request of a shared FMM batch (position in the batch).
  struct Pos {
    std::vector<std::unique_ptr<FmmEvaluator>> ev;  // one per rbf, attached to the model's source side
    std::vector<DeferredFmm> jobs;                   // their recorded near and far fields
    DevBuf<double> d_pts, d_out;
    std::vector<DevBuf<double>> d_tmp;               // per rbf after the first
  };
  std::vector<std::unique_ptr<Pos>> pos;
  DirectBatch direct;
  ~Slot() {
    pos.clear();
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

Server::Server(const ServeOptions& opt) : opt_(opt) {
  if (opt_.n_slots < 1 || opt_.n_slots > 64) throw_invalid("n_slots must be within 1 to 64");
  if (!(opt_.rel_accuracy >= 0.0)) throw_invalid("rel_accuracy must be nonnegative");
  if (opt_.hbm_budget < 0) throw_invalid("hbm_budget must be nonnegative");
  if (opt_.hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    PAM_HIP(hipMemGetInfo(&free_b, &total_b));
    opt_.hbm_budget = static_cast<int64_t>(0.9 * static_cast<double>(free_b));
  }
  for (int i = 0; i < opt_.n_slots; i++) {
    auto sl = std::make_unique<Slot>();
    PAM_HIP(hipStreamCreateWithFlags(&sl->s, hipStreamNonBlocking));
    slots_.push_back(std::move(sl));
  }
  for (int i = 0; i < opt_.n_slots; i++) threads_.emplace_back([this, i] { worker(i); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& t : threads_) t.join();
  slots_.clear();
  models_.clear();
}

int Server::add_model(const Model& model, const double* centers, int64_t n, const double* weights,
                      const double* lo, const double* hi, const FmmConfig* cfg) {
  model.validate();
  if (n < 1) throw_invalid("a served model needs at least one centre");
  if (!centers || !weights) throw_invalid("null centres or weights");
  const int l = model.l();
  // Box: the query box united with the centres' bbox.
  double blo[3], bhi[3];
  for (int k = 0; k < 3; k++) {
    blo[k] = lo ? lo[k] : INFINITY;
    bhi[k] = hi ? hi[k] : -INFINITY;
  }
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      blo[k] = std::min(blo[k], centers[3 * i + k]);
      bhi[k] = std::max(bhi[k], centers[3 * i + k]);
    }
  PAM_HIP(hipDeviceSynchronize());
  size_t free0 = 0, total = 0;
  PAM_HIP(hipMemGetInfo(&free0, &total));

  auto m = std::make_unique<ModelEntry>();
  m->model = model;
  m->n = n;
  PAM_HIP(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
  m->d_centers.upload(centers, static_cast<size_t>(n) * 3, m->s);
  m->d_weights.upload(weights, static_cast<size_t>(n + l), m->s);
  m->d_coeffs.resize(10);
  if (l > 0) dev_copy(m->d_weights.get() + n, m->d_coeffs.get(), l, m->s);
  for (const auto& r : model.rbfs) {
    auto e = std::make_unique<FmmEvaluator>(r, blo, bhi, false, m->s);
    e->set_source_points(m->d_centers.get(), n, true);
    e->set_weights(m->d_weights.get(), n, true);
    m->owners.push_back(std::move(e));
  }
  if (cfg) {
    for (auto& e : m->owners) e->set_config(*cfg);
  } else if (opt_.rel_accuracy == 0.0) {
    FmmConfig c;
    c.precision = kFp64;
    c.order = 10;
    for (auto& e : m->owners) e->set_config(c);
  } else {
    // Absolute accuracy = rel_accuracy x max |value| over a sample of the
    // centres (FP64 direct sum + polynomial), split over the rbfs as in
    // interpolation/evaluator.hpp:95-97.
    int64_t ns = std::min<int64_t>(n, 2000);
    std::vector<double> smp(static_cast<size_t>(ns) * 3);
    for (int64_t i = 0; i < ns; i++) {
      int64_t j = (i * n) / ns;
      for (int k = 0; k < 3; k++) smp[3 * i + k] = centers[3 * j + k];
    }
    DevBuf<double> d_smp, d_val, d_tmp;
    d_smp.upload(smp.data(), smp.size(), m->s);
    d_val.resize(static_cast<size_t>(ns));
    d_tmp.resize(static_cast<size_t>(ns));
    dev_fill(0.0, d_val.get(), ns, m->s);
    DevBuf<double> scratch;
    for (const auto& r : model.rbfs) {
      FmmEvaluator probe(r, blo, bhi, false, m->s);
      probe.set_source_points(m->d_centers.get(), n, true);
      probe.set_weights(m->d_weights.get(), n, true);
      FmmConfig dc;
      dc.mode = kModeDirect;
      probe.set_config(dc);
      probe.set_target_points(d_smp.get(), ns, true);
      probe.evaluate(d_tmp.get(), ns, true);
      dev_axpy(1.0, d_tmp.get(), d_val.get(), ns, m->s);
    }
    if (l > 0) launch_poly_eval(model.poly_degree, d_smp.get(), ns, m->d_coeffs.get(), d_val.get(), m->s);
    std::vector<double> val(static_cast<size_t>(ns));
    d_val.download(val.data(), val.size(), m->s);
    PAM_HIP(hipStreamSynchronize(m->s));
    double vmax = 0.0;
    for (double v : val) vmax = std::max(vmax, std::abs(v));
    if (vmax == 0.0) vmax = 1.0;
    double acc = opt_.rel_accuracy * vmax / static_cast<double>(model.rbfs.size());
    for (auto& e : m->owners) e->set_accuracy(acc);
  }
  for (auto& e : m->owners) e->prepare_source();
  m->cfg = m->owners[0]->active_config();
  m->level = m->owners[0]->source_tree().level;
  PAM_HIP(hipStreamSynchronize(m->s));
  PAM_HIP(hipDeviceSynchronize());
  size_t free1 = 0;
  PAM_HIP(hipMemGetInfo(&free1, &total));
  m->bytes = free0 > free1 ? static_cast<int64_t>(free0 - free1) : 0;

  std::lock_guard<std::mutex> lk(mu_);
  resident_ += m->bytes;
  models_.push_back(std::move(m));
  return static_cast<int>(models_.size() - 1);
}

const FmmConfig& Server::model_config(int id) const {
  std::lock_guard<std::mutex> lk(mu_);
  if (id < 0 || id >= static_cast<int>(models_.size())) throw_invalid("unknown model id");
  return models_[id]->cfg;
}

int64_t Server::model_bytes(int id) const {
  std::lock_guard<std::mutex> lk(mu_);
  if (id < 0 || id >= static_cast<int>(models_.size())) throw_invalid("unknown model id");
  return models_[id]->bytes;
}

int64_t Server::estimate_bytes(const ModelEntry& m, int64_t n) const {
  // Working set of one request: the query points and outputs, and per rbf the
  // transformed targets, their octree and sort buffers, the FP64/FP32
  // tree-order copies, the far-field values, the local expansions and lists.
  const int kp = kpad_of(m.cfg.order);
  const int64_t nr = static_cast<int64_t>(m.model.rbfs.size());
  int64_t cells = 0;
  for (int l = 2; l <= m.level; l++) cells += std::min<int64_t>(n, int64_t{1} << (3 * l));
  int64_t leaves = std::min<int64_t>(n, int64_t{1} << (3 * m.level));
  int64_t per_pt = 24 + 8 + (nr > 1 ? 8 : 0) + nr * (24 + 40 + 32 + 16 + 8);
  int64_t per_cell = nr * (static_cast<int64_t>(kp) * 8 + kM2LSlots * 4 + 64);
  return n * per_pt + cells * per_cell + leaves * nr * 27 * 4 + (int64_t{1} << 20);
}

int64_t Server::submit(int model, const double* pts, int64_t n, double* out) {
  if (n < 0) throw_invalid("negative point count");
  if (n > 0 && (!pts || !out)) throw_invalid("null query or output buffer");
  auto r = std::make_unique<Request>();
  r->pts = pts;
  r->n = n;
  r->out = out;
  r->t_submit = now_s();
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (model < 0 || model >= static_cast<int>(models_.size())) throw_invalid("unknown model id");
    r->model = models_[model].get();
    r->est = estimate_bytes(*r->model, n);
    id = next_id_++;
    r->id = id;
    if (first_submit_ < 0) first_submit_ = r->t_submit;
    submitted_++;
    queue_.push_back(r.get());
    requests_[id] = std::move(r);
  }
  cv_work_.notify_one();
  return id;
}

bool Server::batchable(const Request& r) {
  // Requests that the evaluator would answer with the FP64 direct sum (below
  // 2^26 source-target pairs, or a model fixed to direct mode).
  if (r.n <= 0) return false;
  const ModelEntry& m = *r.model;
  if (m.cfg.mode == kModeDirect) return true;
  if (m.cfg.mode != kModeAuto) return false;
  return static_cast<double>(r.n) * static_cast<double>(m.n) < 67108864.0;
}

void Server::worker(int slot) {
  Slot& sl = *slots_[slot];
  std::vector<Request*> batch;
  for (;;) {
    Request* r = nullptr;
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        if (stop_) return;
        if (!queue_.empty()) {
          Request* head = queue_.front();
          int64_t room = opt_.hbm_budget - resident_;
          if (head->est > room) {
            // can never fit: fail it instead of blocking the queue
            queue_.pop_front();
            head->code = kOutOfMemory;
            head->err = "request working set exceeds the HBM budget left by the resident models";
            head->done = true;
            head->t_admit = head->t_done = now_s();
            failed_++;
            cv_done_.notify_all();
            continue;
          }
          if (in_flight_ == 0 || admitted_ + head->est <= room) {
            queue_.pop_front();
            r = head;
            break;
          }
        }
        cv_work_.wait(lk);
      }
      admitted_ += r->est;
      batch.push_back(r);
      if (!batchable(*r) && r->n > 0) {
        // Requests for the FMM: queued ones of any model join this batch, in
        // queue order, up to kMaxFmmBatch requests and kMaxFmmBatchPoints
        // query points, as far as the HBM budget allows; their near fields and
        // far-field L2Ps then run in shared launches.
        const int64_t room = opt_.hbm_budget - resident_;
        int64_t pts = r->n;
        for (auto it = queue_.begin(); it != queue_.end() && batch.size() < kMaxFmmBatch;) {
          Request* q = *it;
          if (!batchable(*q) && q->n > 0 && pts + q->n <= kMaxFmmBatchPoints && admitted_ + q->est <= room) {
            admitted_ += q->est;
            pts += q->n;
            batch.push_back(q);
            it = queue_.erase(it);
          } else {
            ++it;
          }
        }
      }
      if (batchable(*r)) {
        // Continuous batching: the queued requests that are also answered by
        // the direct sum join this launch, whatever their model, in queue
        // order, as far as the HBM budget and the batch limits allow.
        const int64_t room = opt_.hbm_budget - resident_;
        int64_t rows = r->n;
        for (auto it = queue_.begin(); it != queue_.end() && batch.size() < kMaxBatch;) {
          Request* q = *it;
          if (batchable(*q) && rows + q->n <= kMaxBatchRows && admitted_ + q->est <= room) {
            admitted_ += q->est;
            rows += q->n;
            batch.push_back(q);
            it = queue_.erase(it);
          } else {
            ++it;
          }
        }
      }
      peak_admitted_ = std::max(peak_admitted_, admitted_);
      in_flight_ += static_cast<int>(batch.size());
      max_in_flight_ = std::max(max_in_flight_, in_flight_);
      const double t = now_s();
      for (Request* q : batch) q->t_admit = t;
    }
    int code = kOk;
    std::string err;
    try {
      if (batchable(*r)) run_batch(sl, batch);
      else run_fmm_batch(sl, batch);
    } catch (const Error& e) {
      code = e.code;
      err = e.what();
    } catch (const std::exception& e) {
      code = kRuntimeError;
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      const double t = now_s();
      if (batchable(*r)) {
        batches_++;
        batched_requests_ += static_cast<int64_t>(batch.size());
        max_batch_ = std::max(max_batch_, static_cast<int>(batch.size()));
      } else if (r->n > 0) {
        fmm_batches_++;
        fmm_batched_requests_ += static_cast<int64_t>(batch.size());
        max_fmm_batch_ = std::max(max_fmm_batch_, static_cast<int>(batch.size()));
      }
      for (Request* q : batch) {
        admitted_ -= q->est;
        in_flight_--;
        q->code = code;
        q->err = err;
        q->t_done = t;
        q->done = true;
        if (code == kOk) {
          completed_++;
          points_ += q->n;
          lat_ms_.push_back((q->t_done - q->t_submit) * 1e3);
          queue_ms_.push_back((q->t_admit - q->t_submit) * 1e3);
          last_done_ = std::max(last_done_, q->t_done);
        } else {
          failed_++;
        }
      }
    }
    cv_done_.notify_all();
    cv_work_.notify_all();
  }
}

void Server::run_fmm_batch(Slot& sl, const std::vector<Request*>& batch) {
  // Per request and rbf: the target tree, lists and far field up to L2L on
  // the slot's stream; then the near fields of the whole batch in shared
  // launches and its far-field L2Ps in shared launches (launch_deferred),
  // whatever the models; then each request's rbf terms, polynomial and
  // download. Every output gets the bits of the same request served alone.
  hipStream_t s = sl.s;
  while (sl.pos.size() < batch.size()) sl.pos.push_back(std::make_unique<Slot::Pos>());
  std::vector<const DeferredFmm*> jobs;
  for (size_t b = 0; b < batch.size(); b++) {
    Request& r = *batch[b];
    const int64_t n = r.n;
    if (n == 0) continue;
    ModelEntry& m = *r.model;
    Slot::Pos& p = *sl.pos[b];
    const size_t nr = m.owners.size();
    p.d_pts.upload(r.pts, static_cast<size_t>(n) * 3, s);
    p.d_out.resize(static_cast<size_t>(n));
    if (p.d_tmp.size() < nr) p.d_tmp.resize(nr);
    if (p.jobs.size() < nr) p.jobs.resize(nr);
    while (p.ev.size() < nr) {
      p.ev.push_back(std::make_unique<FmmEvaluator>(m.owners[p.ev.size()]->rbf(), nullptr, nullptr, false, s));
      // Requests overlap across slots; a second stream per slot evaluator would
      // only add hardware-queue sharing between the slots.
      p.ev.back()->set_overlap(false);
    }
    for (size_t k = 0; k < nr; k++) {
      FmmEvaluator& e = *p.ev[k];
      e.attach_source(m.owners[k].get());
      e.set_target_points(p.d_pts.get(), n, true);
      double* o = p.d_out.get();
      if (k > 0) {
        p.d_tmp[k].resize(static_cast<size_t>(n));
        o = p.d_tmp[k].get();
      }
      e.evaluate_deferred(o, n, &p.jobs[k]);
      if (p.jobs[k].valid) jobs.push_back(&p.jobs[k]);
    }
  }
  if (!jobs.empty()) launch_deferred(jobs, s);
  for (size_t b = 0; b < batch.size(); b++) {
    Request& r = *batch[b];
    const int64_t n = r.n;
    if (n == 0) continue;
    ModelEntry& m = *r.model;
    Slot::Pos& p = *sl.pos[b];
    for (size_t k = 1; k < m.owners.size(); k++) dev_axpy(1.0, p.d_tmp[k].get(), p.d_out.get(), n, s);
    if (m.model.l() > 0) launch_poly_eval(m.model.poly_degree, p.d_pts.get(), n, m.d_coeffs.get(), p.d_out.get(), s);
    p.d_out.download(r.out, static_cast<size_t>(n), s);
  }
  PAM_HIP(hipStreamSynchronize(s));
}

void Server::run_batch(Slot& sl, const std::vector<Request*>& batch) {
  std::vector<DirectBatchQuery> qs;
  std::vector<DirectBatchTerm> terms;
  for (Request* r : batch) {
    const ModelEntry& m = *r->model;
    DirectBatchQuery q{};
    q.pts = r->pts;
    q.out = r->out;
    q.n = r->n;
    q.first_term = static_cast<int>(terms.size());
    q.n_terms = static_cast<int>(m.owners.size());
    q.poly_degree = m.model.l() > 0 ? m.model.poly_degree : -1;
    q.coeffs = m.d_coeffs.get();
    qs.push_back(q);
    for (const auto& e : m.owners)
      terms.push_back(DirectBatchTerm{e->rbf(), e->d_source_points(), e->d_weights(), m.n});
  }
  sl.direct.run(qs, terms, sl.s);
}

void Server::wait(int64_t id) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) throw_invalid("unknown request id");
  Request* r = it->second.get();
  cv_done_.wait(lk, [&] { return r->done; });
  if (r->code != kOk) throw Error(r->code, r->err);
}

bool Server::done(int64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) throw_invalid("unknown request id");
  return it->second->done;
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return queue_.empty() && in_flight_ == 0; });
}

void Server::request_times(int64_t id, double* queue_ms, double* total_ms) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) throw_invalid("unknown request id");
  const Request& r = *it->second;
  if (!r.done) throw_runtime("the request has not completed");
  *queue_ms = (r.t_admit - r.t_submit) * 1e3;
  *total_ms = (r.t_done - r.t_submit) * 1e3;
}

ServeStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  ServeStats st;
  st.submitted = submitted_;
  st.completed = completed_;
  st.failed = failed_;
  st.points = points_;
  st.p50_ms = percentile(lat_ms_, 0.50);
  st.p99_ms = percentile(lat_ms_, 0.99);
  if (!lat_ms_.empty()) {
    double sum = 0, qs = 0;
    for (double v : lat_ms_) sum += v;
    for (double v : queue_ms_) qs += v;
    st.mean_ms = sum / static_cast<double>(lat_ms_.size());
    st.mean_queue_ms = qs / static_cast<double>(queue_ms_.size());
    st.max_ms = *std::max_element(lat_ms_.begin(), lat_ms_.end());
  }
  st.wall_s = first_submit_ >= 0 ? last_done_ - first_submit_ : 0.0;
  st.points_per_s = st.wall_s > 0 ? static_cast<double>(points_) / st.wall_s : 0.0;
  st.max_in_flight = max_in_flight_;
  st.resident_bytes = resident_;
  st.budget = opt_.hbm_budget;
  st.peak_admitted_bytes = peak_admitted_;
  st.batches = batches_;
  st.batched_requests = batched_requests_;
  st.max_batch = max_batch_;
  st.fmm_batches = fmm_batches_;
  st.fmm_batched_requests = fmm_batched_requests_;
  st.max_fmm_batch = max_fmm_batch_;
  return st;
}

void Server::reset_stats() {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto it = requests_.begin(); it != requests_.end();) {
    if (it->second->done) it = requests_.erase(it);
    else ++it;
  }
  lat_ms_.clear();
  queue_ms_.clear();
  completed_ = failed_ = submitted_ = points_ = 0;
  batches_ = batched_requests_ = 0;
  max_batch_ = 0;
  fmm_batches_ = fmm_batched_requests_ = 0;
  max_fmm_batch_ = 0;
  first_submit_ = -1;
  last_done_ = 0;
  max_in_flight_ = in_flight_;
  peak_admitted_ = admitted_;
}

}  // namespace pam
