// Request loop over fitted models kept resident in HBM. polatory has no
// server; the closest reference call is Interpolant::evaluate per request
// (interpolant.hpp:54-67), whose evaluator is rebuilt for every call.
//
// A model keeps one prepared source-side FMM evaluator per rbf (points,
// weights, source tree, multipole expansions), its polynomial coefficients,
// and the FMM configuration chosen for the requested relative accuracy.
// Requests are admitted in arrival order, when a slot is free and the HBM
// budget holds their working set. Every slot is a worker thread with its own
// stream whose evaluators attach to the model's source side, so the
// requests in flight overlap on the device while sharing the resident data.
// Requests are batched across models. Those that the FP64 direct sum answers
// (few query points against the model's centres): a slot takes the queued
// ones of any model into one shared direct-sum launch and one reduction.
// Those that need the FMM: a slot takes up to four queued ones of any model,
// builds each one's target tree, lists and far field up to L2L, then runs the
// near fields of the whole batch in shared launches and the far-field L2Ps in
// shared launches (one per precision and kernel kind, one per order).
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "interpolation.hpp"

namespace pam {

struct ServeOptions {
  int n_slots = 4;
  int64_t hbm_budget = 0;      // bytes for resident models + admitted requests; 0 = 90% of free HBM
  double rel_accuracy = 1e-6;  // relative to max |value| of the model at its centres; 0 = FP64 order 10
};

struct ServeStats {
  int64_t submitted = 0, completed = 0, failed = 0, points = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
  double mean_queue_ms = 0;
  double points_per_s = 0, wall_s = 0;
  int max_in_flight = 0;
  int64_t resident_bytes = 0, budget = 0, peak_admitted_bytes = 0;
  // shared direct-sum launches, the requests answered in them, and the most
  // requests in one
  int64_t batches = 0, batched_requests = 0;
  int max_batch = 0;
  // FMM batches (shared near-field and far-field launches), the requests in
  // them, and the most requests in one
  int64_t fmm_batches = 0, fmm_batched_requests = 0;
  int max_fmm_batch = 0;
};

class Server {
 public:
  explicit Server(const ServeOptions& opt);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Registers a fitted model: centres (n x 3), weights (n + l), and the box
  // that every query point must lie in. cfg != nullptr fixes the FMM
  // configuration instead of the accuracy search. Returns the model id.
  int add_model(const Model& model, const double* centers, int64_t n, const double* weights,
                const double* lo, const double* hi, const FmmConfig* cfg);
  const FmmConfig& model_config(int id) const;
  int64_t model_bytes(int id) const;

  // Queues a request; pts (n x 3) and out (n) are host buffers that must stay
  // valid until the request completes. Returns the request id.
  int64_t submit(int model, const double* pts, int64_t n, double* out);
  // Blocks until the request completes; rethrows its error.
  void wait(int64_t id);
  bool done(int64_t id);
  void drain();
  // Queue and run times of a completed request (ms).
  void request_times(int64_t id, double* queue_ms, double* total_ms);
  ServeStats stats();
  void reset_stats();

 private:
  struct ModelEntry;
  struct Request;
  struct Slot;
  void worker(int slot);
  int64_t estimate_bytes(const ModelEntry& m, int64_t n) const;
  void run_fmm_batch(Slot& slot, const std::vector<Request*>& batch);
  void run_batch(Slot& slot, const std::vector<Request*>& batch);
  static bool batchable(const Request& r);

  // Limits of one shared direct-sum launch.
  static constexpr size_t kMaxBatch = 64;
  static constexpr int64_t kMaxBatchRows = int64_t{1} << 19;
  // Limits of one FMM batch (shared near-field and L2P launches).
  static constexpr size_t kMaxFmmBatch = 4;
  static constexpr int64_t kMaxFmmBatchPoints = int64_t{1} << 21;

  ServeOptions opt_;
  std::vector<std::unique_ptr<ModelEntry>> models_;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::vector<std::thread> threads_;

  mutable std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Request*> queue_;
  std::unordered_map<int64_t, std::unique_ptr<Request>> requests_;
  int64_t next_id_ = 1;
  bool stop_ = false;
  int64_t resident_ = 0, admitted_ = 0, peak_admitted_ = 0;
  int in_flight_ = 0, max_in_flight_ = 0;
  std::vector<double> lat_ms_, queue_ms_;
  int64_t completed_ = 0, failed_ = 0, submitted_ = 0, points_ = 0;
  int64_t batches_ = 0, batched_requests_ = 0;
  int max_batch_ = 0;
  int64_t fmm_batches_ = 0, fmm_batched_requests_ = 0;
  int max_fmm_batch_ = 0;
  double first_submit_ = -1, last_done_ = 0;
};

}  // namespace pam
