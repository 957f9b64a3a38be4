// The batched march (see scan.hpp).
#include "scan.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "sync.hpp"

namespace sp {

// ------------------------------------------------------------------ worker pool

WorkerPool::WorkerPool(int threads) {
  for (int i = 0; i < threads; ++i) threads_.emplace_back([this] { work(); });
}

WorkerPool::~WorkerPool() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  wake_.notify_all();
  for (auto& t : threads_) t.join();
}

void WorkerPool::call(int i) {
  try {
    (*job_)(i);
  } catch (...) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!failure_) failure_ = std::current_exception();
  }
}

void WorkerPool::work() {
  long seen = 0;
  std::unique_lock<std::mutex> lock(mu_);
  for (;;) {
    wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
    if (stop_) return;
    seen = generation_;
    running_ += 1;
    while (next_ < n_) {
      const int i = next_++;
      lock.unlock();
      call(i);
      lock.lock();
    }
    running_ -= 1;
    if (running_ == 0) done_.notify_all();
  }
}

void WorkerPool::run(int n, const std::function<void(int)>& f) {
  if (threads_.empty() || n < 8) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::unique_lock<std::mutex> lock(mu_);
  job_ = &f;
  n_ = n;
  next_ = 0;
  failure_ = nullptr;
  generation_ += 1;
  wake_.notify_all();
  // The caller takes indices too.
  running_ += 1;
  while (next_ < n_) {
    const int i = next_++;
    lock.unlock();
    call(i);
    lock.lock();
  }
  running_ -= 1;
  done_.wait(lock, [&] { return running_ == 0 && next_ >= n_; });
  job_ = nullptr;
  n_ = 0;
  if (failure_) {
    std::exception_ptr e = failure_;
    failure_ = nullptr;
    std::rethrow_exception(e);
  }
}

// ------------------------------------------------------------------------ scan

namespace {

template <typename T>
T* device_upload(std::vector<void*>& owned, const T* host, std::size_t n) {
  T* p = nullptr;
  SP_HIP(hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(T)));
  owned.push_back(p);
  if (n) SP_HIP(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

double elapsed_ms(std::chrono::steady_clock::time_point since) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - since)
      .count();
}

}  // namespace

namespace {
// What a scan refuses, from the problem record alone (before anything is allocated).
std::string scan_refusal(const sp_problem_desc& d) {
  if (d.kind != 1 || d.method != 1)
    return "a scan instance runs the field grid under the exponential Adams march";
  if (d.n_modes < 1 || d.n_modes > kMaxBatchModes) return "a scan instance has 1 to 6 modes";
  if (d.n_raman) return "a scan instance has no Raman response";
  if (d.absorber_scheme != 0 || d.spatial_absorber) return "a scan instance has no absorber";
  return {};
}
}  // namespace

Scan::Scan(const sp_problem_desc* descs, int n) {
  SP_CHECK(descs && n >= 1, "a scan needs at least one instance");
  SP_CHECK(n <= 65536, "a scan holds at most 65536 instances");
  for (int i = 0; i < n; ++i) {
    const std::string why = scan_refusal(descs[i]);
    SP_CHECK(why.empty(), "scan instance " + std::to_string(i) + ": " + why);
  }
  try {
    SP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    inst_.reserve(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) {
      inst_.push_back(std::make_unique<FieldWorkspace>(descs[i], stream_));
      const std::string why = inst_.back()->batch_refusal();
      SP_CHECK(why.empty(), "scan instance " + std::to_string(i) + ": " + why);
    }
    SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&rec_host_), sizeof(AdamsRecord) * n));
    std::memset(rec_host_, 0, sizeof(AdamsRecord) * n);
    SP_HIP(hipMalloc(&rec_dev_, sizeof(AdamsRecord) * n));
    allocations_.push_back(rec_dev_);
    for (int i = 0; i < n; ++i) inst_[i]->batch_bind(rec_host_ + i, rec_dev_ + i);
    plan_launches();
    SP_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    const unsigned cores = std::max(1u, std::thread::hardware_concurrency());
    pool_ = std::make_unique<WorkerPool>(static_cast<int>(std::min(7u, cores - 1)));
    graph_ = capture_graph(stream_, [&] { enqueue_round(); });
  } catch (...) {
    release();
    throw;
  }
}

Scan::~Scan() { release(); }

void Scan::release() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_) (void)hipGraphExecDestroy(graph_);
  graph_ = nullptr;
  inst_.clear();
  for (void* p : allocations_) (void)hipFree(p);
  allocations_.clear();
  if (rec_host_) (void)hipHostFree(rec_host_);
  rec_host_ = nullptr;
  if (done_) (void)hipEventDestroy(done_);
  done_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  pool_.reset();
}

// The grids of every grouped launch: per instance, the workgroups its plain launch uses.
void Scan::plan_launches() {
  const int n = size();
  std::vector<BatchInstance> host(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) host[i] = inst_[i]->batch_instance();
  inst_dev_ = device_upload(allocations_, host.data(), host.size());

  std::vector<int> first;
  std::vector<char> fused(static_cast<std::size_t>(n), 0);
  auto table = [&](Group& g, auto blocks_of) {
    g.offset = static_cast<int>(first.size());
    int total = 0;
    for (int i = 0; i < n; ++i) {
      first.push_back(total);
      total += blocks_of(host[i].P);
    }
    first.push_back(total);
    g.blocks = total;
  };
  auto cn = [](const DeviceProblem& P) { return static_cast<std::size_t>(P.C) * P.N; };
  table(commit_, [&](const DeviceProblem& P) { return static_cast<int>(blocks_for(cn(P), 256)); });
  table(predict_, [&](const DeviceProblem& P) { return static_cast<int>(blocks_for(cn(P), kWave)); });
  const int per_group = kNodeWaves * kNodesPerWave;
  auto node_blocks = [&](const DeviceProblem& P) {
    return P.has_plasma ? scan_tiles(P.T) * ((P.Q + per_group - 1) / per_group) : 0;
  };
  table(rate_, node_blocks);
  table(current_, node_blocks);
  table(prefix_, [&](const DeviceProblem& P) { return P.has_plasma ? P.Q : 0; });
  for (int i = 0; i < n; ++i) fused[i] = fused_node_passes(host[i].P, host[i].B);
  int at = 0;
  table(prefix_current_, [&](const DeviceProblem& P) {
    return P.has_plasma && !fused[at++] ? P.Q : 0;
  });
  table(response_, [&](const DeviceProblem& P) {
    SP_CHECK(P.T % kWave == 0, "the node response takes 64-sample blocks");
    return P.T / kWave;
  });
  table(finish_, [&](const DeviceProblem& P) {
    return static_cast<int>(blocks_for(static_cast<std::size_t>(P.C) * (P.T / 4 + 1), kWave));
  });
  table(correct_, [&](const DeviceProblem& P) { return static_cast<int>(correct_blocks(P)); });
  table(norms_, [&](const DeviceProblem&) { return 1; });
  for (int i = 0; i < n; ++i)
    SP_CHECK(host[i].correct_blocks == first[correct_.offset + i + 1] - first[correct_.offset + i],
             "the corrector's workgroups");
  first_dev_ = device_upload(allocations_, first.data(), first.size());

  std::vector<std::vector<FftJob>> inv(static_cast<std::size_t>(n)), fwd(inv.size());
  std::vector<const FftPlan*> inv_plans, fwd_plans;
  for (int i = 0; i < n; ++i) {
    const int* gate = &rec_dev_[i].action;
    inv[i] = inst_[i]->batch_inverse_jobs(gate);
    fwd[i] = inst_[i]->batch_forward_jobs(gate);
    inv_plans.push_back(&inst_[i]->inverse_plan());
    fwd_plans.push_back(&inst_[i]->forward_plan());
  }
  inverse_ = group_passes(inv, inv_plans);
  forward_ = group_passes(fwd, fwd_plans);
  const char* env = std::getenv("SOLPROP_MI355X_AUTOTUNE");
  if (!(env && env[0] == '0')) {
    schedule_.clear();
    tune(inverse_, +1, "inverse");
    tune(forward_, -1, "forward");
    schedule_ = "[" + schedule_ + "]";
  }
}

// Re-plan a group's jobs at `log2G` sequences a workgroup.
void Scan::set_schedule(FftGroup& g, int log2G) {
  std::vector<FftJob> jobs = g.host;
  std::vector<int> first;
  int total = 0;
  for (std::size_t i = 0; i < jobs.size(); ++i) {
    jobs[i].groups = g.sequences[i] >> log2G;
    first.push_back(total);
    total += jobs[i].groups * jobs[i].rows;
  }
  first.push_back(total);
  SP_HIP(hipMemcpy(g.jobs, jobs.data(), jobs.size() * sizeof(FftJob), hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(g.first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
  g.log2G = log2G;
  g.blocks = total;
}

namespace {
std::mutex g_tuned_mu;
std::map<std::string, int> g_tuned;  // pass shape -> sequences a workgroup (log2)
}  // namespace

void Scan::tune(std::vector<FftGroup>& groups, int sign, const char* direction) {
  hipEvent_t t0 = nullptr, t1 = nullptr;
  SP_HIP(hipEventCreate(&t0));
  SP_HIP(hipEventCreate(&t1));
  try {
    for (std::size_t gi = 0; gi < groups.size(); ++gi) {
      FftGroup& g = groups[gi];
      // The key: direction, pass shape and every job's (sequences, rows).
      std::vector<std::pair<int, int>> shape;
      for (std::size_t i = 0; i < g.host.size(); ++i)
        shape.emplace_back(g.sequences[i], g.host[i].rows);
      std::sort(shape.begin(), shape.end());
      std::string key = std::string(direction) + ":" + std::to_string(g.log2L);
      for (const auto& sr : shape) key += ":" + std::to_string(sr.first) + "x" + std::to_string(sr.second);
      const int planned = g.log2G;
      std::string timings;
      int best = planned;
      bool cached = false;
      {
        std::lock_guard<std::mutex> lock(g_tuned_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
          best = it->second;
          cached = true;
        }
      }
      if (!cached) {
        // The timed jobs run ungated (the rounds' gates are closed before any load).
        std::vector<FftJob> open = g.host;
        for (FftJob& j : open) j.gate = nullptr;
        const std::vector<FftJob> saved = g.host;
        g.host = open;
        float best_ms = 0.0f;
        for (int cand = 0; cand <= 4; ++cand) {
          // Shapes with no kernel (more points a workgroup than LDS holds) are skipped.
          bool fits = true;
          for (int sq : g.sequences) fits = fits && sq % (1 << cand) == 0;
          if (!fits || !fft_grouped_available(g.log2L, cand)) continue;
          set_schedule(g, cand);
          for (int r = 0; r < 2; ++r)
            launch_fft_grouped(g.log2L, g.log2G, g.jobs, g.first, g.n, g.blocks, sign, stream_);
          SP_HIP(hipEventRecord(t0, stream_));
          constexpr int kReps = 5;
          for (int r = 0; r < kReps; ++r)
            launch_fft_grouped(g.log2L, g.log2G, g.jobs, g.first, g.n, g.blocks, sign, stream_);
          SP_HIP(hipEventRecord(t1, stream_));
          SP_HIP(hipEventSynchronize(t1));
          float ms = 0.0f;
          SP_HIP(hipEventElapsedTime(&ms, t0, t1));
          ms /= kReps;
          if (!timings.empty()) timings += ", ";
          timings += "\"" + std::to_string(1 << cand) + "\": " + std::to_string(1e3 * ms);
          if (best_ms == 0.0f || ms < best_ms) {
            best_ms = ms;
            best = cand;
          }
        }
        g.host = saved;
        std::lock_guard<std::mutex> lock(g_tuned_mu);
        g_tuned[key] = best;
      }
      set_schedule(g, best);
      if (!schedule_.empty()) schedule_ += ", ";
      schedule_ += "{\"direction\": \"" + std::string(direction) + "\", \"points\": " +
                   std::to_string(1 << g.log2L) + ", \"jobs\": " + std::to_string(g.n) +
                   ", \"planned\": " + std::to_string(1 << planned) + ", \"chosen\": " +
                   std::to_string(1 << best) + ", \"cached\": " + (cached ? "true" : "false") +
                   ", \"us\": {" + timings + "}}";
    }
  } catch (...) {
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    throw;
  }
  SP_HIP(hipStreamSynchronize(stream_));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
}

// The transforms' passes as grouped launches: every instance's first pass, one launch per
// pass shape, then every second pass (the four-step transforms' row passes).
std::vector<Scan::FftGroup> Scan::group_passes(
    const std::vector<std::vector<FftJob>>& per_instance,
    const std::vector<const FftPlan*>& plans) {
  std::vector<FftGroup> out;
  for (int pass = 0; pass < 2; ++pass) {
    std::map<std::pair<int, int>, std::vector<FftJob>> shapes;
    for (std::size_t i = 0; i < per_instance.size(); ++i) {
      if (pass >= plans[i]->passes()) continue;
      const auto key = std::make_pair(plans[i]->pass_log2L(pass), plans[i]->pass_log2G(pass));
      shapes[key].push_back(per_instance[i][static_cast<std::size_t>(pass)]);
    }
    for (const auto& kv : shapes) {
      FftGroup g;
      g.log2L = kv.first.first;
      g.log2G = kv.first.second;
      g.n = static_cast<int>(kv.second.size());
      g.host = kv.second;
      for (const FftJob& j : kv.second) g.sequences.push_back(j.groups << g.log2G);
      std::vector<int> first;
      int total = 0;
      for (const FftJob& j : kv.second) {
        first.push_back(total);
        total += j.groups * j.rows;
      }
      first.push_back(total);
      g.blocks = total;
      g.jobs = device_upload(allocations_, kv.second.data(), kv.second.size());
      g.first = device_upload(allocations_, first.data(), first.size());
      out.push_back(g);
    }
  }
  return out;
}

BatchGrid Scan::grid(const Group& g) const {
  return BatchGrid{inst_dev_, rec_dev_, first_dev_ + g.offset, size(), g.blocks};
}

// One round: the records up, then the trial's kernels over every instance.
void Scan::enqueue_round() {
  SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(AdamsRecord) * size(),
                        hipMemcpyHostToDevice, stream_));
  launch_batch_commit(grid(commit_), stream_);
  launch_batch_predict(grid(predict_), stream_);
  for (const FftGroup& g : inverse_)
    launch_fft_grouped(g.log2L, g.log2G, g.jobs, g.first, g.n, g.blocks, +1, stream_);
  launch_batch_rate(grid(rate_), stream_);
  launch_batch_prefix(grid(prefix_), false, stream_);
  launch_batch_current(grid(current_), stream_);
  launch_batch_prefix(grid(prefix_current_), true, stream_);
  launch_batch_response(grid(response_), stream_);
  for (const FftGroup& g : forward_)
    launch_fft_grouped(g.log2L, g.log2G, g.jobs, g.first, g.n, g.blocks, -1, stream_);
  launch_batch_finish(grid(finish_), stream_);
  launch_batch_correct(grid(correct_), stream_);
  launch_batch_norms(grid(norms_), stream_);
}

void Scan::load(int i, double z, const double* field) {
  SP_CHECK(i >= 0 && i < size(), "no such scan instance");
  inst_[static_cast<std::size_t>(i)]->load(z, field);
}

void Scan::state(int i, double* out) {
  SP_CHECK(i >= 0 && i < size(), "no such scan instance");
  inst_[static_cast<std::size_t>(i)]->read_state(out);
}

void Scan::advance(const double* targets, long work_limit,
                   std::vector<std::vector<double>>& rows, sp_march_result* out, int* status,
                   std::vector<std::string>& messages) {
  const int n = size();
  rows.assign(static_cast<std::size_t>(n), {});
  messages.assign(static_cast<std::size_t>(n), {});
  for (int i = 0; i < n; ++i)
    inst_[i]->batch_begin(targets[i], work_limit > 0 ? work_limit : 1, &rows[i]);
  std::vector<int> actions(static_cast<std::size_t>(n), kBatchIdle);
  for (;;) {
    auto t0 = std::chrono::steady_clock::now();
    pool_->run(n, [&](int i) { actions[i] = inst_[i]->batch_plan(); });
    const bool active =
        std::any_of(actions.begin(), actions.end(), [](int a) { return a != kBatchIdle; });
    timing_.host_ms += elapsed_ms(t0);
    if (!active) break;
    t0 = std::chrono::steady_clock::now();
    SP_HIP(hipGraphLaunch(graph_, stream_));
    SP_HIP(hipEventRecord(done_, stream_));
    wait_event(done_);
    timing_.device_ms += elapsed_ms(t0);
    timing_.rounds += 1;
    t0 = std::chrono::steady_clock::now();
    pool_->run(n, [&](int i) { inst_[i]->batch_conclude(actions[i]); });
    timing_.host_ms += elapsed_ms(t0);
  }
  for (int i = 0; i < n; ++i) status[i] = inst_[i]->batch_end(&out[i], &messages[i]);
  SP_HIP(hipStreamSynchronize(stream_));
}

long Scan::device_bytes() const {
  long total = 0;
  for (const auto& w : inst_) total += w->counters().device_bytes;
  return total;
}

}  // namespace sp
