// The extern "C" boundary (include/solprop_mi355x.h): status codes and messages in
// place of exceptions, host arrays in and out.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "cartesian.hpp"
#include "field_ws.hpp"
#include "scan.hpp"
#include "session.hpp"
#include "fno.hpp"
#include "gemm.hpp"
#include "trace.hpp"
#include "workspace.hpp"

namespace {
// The envelope route runs the embedded Runge-Kutta march; the modal field route the
// exponential Adams march (or the low-storage replay); the Cartesian route the replay.
std::unique_ptr<sp::Propagator> make_propagator(const sp_problem_desc& d) {
  if (d.kind == 1) return std::make_unique<sp::FieldWorkspace>(d);
  if (d.kind == 2) return std::make_unique<sp::CartesianWorkspace>(d);
  return std::make_unique<sp::Workspace>(d);
}
}  // namespace

struct sp_workspace {
  std::unique_ptr<sp::Propagator> owner;
  sp::Propagator& impl;
  int n_modes, n_omega;
  // An open session leases the workspace: its own march, load and trial refuse meanwhile.
  // Destroying a leased workspace defers its release until the session closes.
  std::mutex lease_mu;
  std::atomic<bool> leased{false};
  bool doomed = false;
  explicit sp_workspace(const sp_problem_desc& d)
      : owner(make_propagator(d)), impl(*owner), n_modes(d.n_modes), n_omega(d.n_omega) {}
};

struct sp_fno {
  std::unique_ptr<sp::fno::Network> impl;
};

struct sp_session {
  sp_workspace* ws;
  std::unique_ptr<sp::session::Session> impl;
};

namespace {

constexpr int kBatchThreads = 16;

void write_message(char* msg, std::size_t len, const char* text) {
  if (!msg || len == 0) return;
  std::snprintf(msg, len, "%s", text);
}

template <typename F>
int guarded(char* msg, std::size_t len, F&& body) {
  try {
    body();
    write_message(msg, len, "");
    return SP_OK;
  } catch (const sp::Error& e) {
    write_message(msg, len, e.what());
    return e.code;
  } catch (const sp::march::ConvergenceError& e) {
    write_message(msg, len, e.what());
    return SP_CONVERGENCE;
  } catch (const sp::session::LifecycleError& e) {
    write_message(msg, len, e.what());
    return SP_STATE;
  } catch (const std::invalid_argument& e) {
    write_message(msg, len, e.what());
    return SP_VALUE;
  } catch (const std::bad_alloc&) {
    write_message(msg, len, "host allocation failed");
    return SP_BACKEND;
  } catch (const std::exception& e) {
    write_message(msg, len, e.what());
    return SP_BACKEND;
  }
}

void refuse_leased(const sp_workspace* ws) {
  if (ws->leased.load())
    throw sp::Error(sp::kState, "the workspace is leased by an open session");
}

// A testing entry's stream and device buffers, released on every exit path.
struct DeviceScope {
  hipStream_t stream = nullptr;
  std::vector<void*> buffers;
  DeviceScope() { SP_HIP(hipStreamCreate(&stream)); }
  ~DeviceScope() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : buffers) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
  template <typename T>
  T* alloc(std::size_t bytes) {
    void* p = nullptr;
    SP_HIP(hipMalloc(&p, bytes));
    buffers.push_back(p);
    return static_cast<T*>(p);
  }
};

}  // namespace

extern "C" {

int sp_workspace_create(const sp_problem_desc* desc, sp_workspace** out, char* msg,
                        size_t msglen) {
  if (!desc || !out) {
    write_message(msg, msglen, "create needs a problem and an output handle");
    return SP_VALUE;
  }
  *out = nullptr;
  return guarded(msg, msglen, [&] { *out = new sp_workspace(*desc); });
}

void sp_workspace_destroy(sp_workspace* ws) {
  if (!ws) return;
  {
    std::lock_guard<std::mutex> lock(ws->lease_mu);
    if (ws->leased.load()) {
      ws->doomed = true;  // released when its session closes
      return;
    }
  }
  delete ws;
}

int sp_workspace_load(sp_workspace* ws, double z, const double* field, char* msg,
                      size_t msglen) {
  if (!ws) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    refuse_leased(ws);
    ws->impl.load(z, field);
  });
}

int sp_workspace_trial(sp_workspace* ws, double z, double dz, double* norm, char* msg,
                       size_t msglen) {
  if (!ws || !norm) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    refuse_leased(ws);
    *norm = ws->impl.probe_trial(z, dz);
  });
}

int sp_workspace_trial_field(sp_workspace* ws, double* out, char* msg, size_t msglen) {
  if (!ws || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] { ws->impl.read_trial(out); });
}

int sp_workspace_state(sp_workspace* ws, double* out, char* msg, size_t msglen) {
  if (!ws || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] { ws->impl.read_state(out); });
}

int sp_workspace_advance(sp_workspace* ws, double target, long work_limit, double* rows,
                         long max_rows, sp_march_result* out, char* msg, size_t msglen) {
  if (!ws || !out) return SP_VALUE;
  std::memset(out, 0, sizeof(*out));
  std::vector<double> held;
  const int status = guarded(msg, msglen, [&] {
    refuse_leased(ws);
    ws->impl.advance(target, work_limit > 0 ? work_limit : 1, held, out);
  });
  const long n = static_cast<long>(held.size() / 4);
  const long kept = rows ? (n < max_rows ? n : max_rows) : 0;
  if (kept > 0) std::memcpy(rows, held.data(), sizeof(double) * 4 * kept);
  out->n_rows = static_cast<int>(kept);
  return status;
}

int sp_batch_advance(sp_workspace* const* ws, int n, const double* targets, long work_limit,
                     double* rows, long max_rows, sp_march_result* out, int* status, char* msg,
                     size_t msglen) {
  if (!ws || n < 1 || !targets || !out || !status) return SP_VALUE;
  for (int i = 0; i < n; ++i)
    if (!ws[i]) return SP_VALUE;
  {
    // One workspace twice would share its stream, graphs and pinned record between two
    // threads: refused.
    std::vector<const sp_workspace*> seen(ws, ws + n);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) {
      write_message(msg, msglen, "a workspace appears more than once in the batch");
      return SP_VALUE;
    }
  }
  std::vector<std::vector<char>> messages(n, std::vector<char>(msglen > 0 ? msglen : 1, 0));
  // A fixed pool, not a thread per instance: at most kBatchThreads instances advance at
  // once, each waiting on its own stream with a backed-off poll.
  static sp::WorkerPool pool(kBatchThreads);
  static std::mutex one_batch;  // the pool runs one loop at a time
  std::lock_guard<std::mutex> lock(one_batch);
  pool.run(n, [&](int i) {
    status[i] = sp_workspace_advance(ws[i], targets[i], work_limit,
                                     rows ? rows + 4 * max_rows * i : nullptr,
                                     rows ? max_rows : 0, out + i, messages[i].data(),
                                     messages[i].size());
  });
  for (int i = 0; i < n; ++i)
    if (status[i] != SP_OK) {
      write_message(msg, msglen, messages[i].data());
      return status[i];
    }
  return SP_OK;
}

struct sp_scan {
  sp::Scan impl;
  sp_scan(const sp_problem_desc* d, int n) : impl(d, n) {}
};

int sp_scan_create(const sp_problem_desc* descs, int n, sp_scan** out, char* msg,
                   size_t msglen) {
  if (!descs || !out || n < 1) {
    write_message(msg, msglen, "a scan needs its problems and an output handle");
    return SP_VALUE;
  }
  *out = nullptr;
  return guarded(msg, msglen, [&] { *out = new sp_scan(descs, n); });
}

void sp_scan_destroy(sp_scan* scan) { delete scan; }

int sp_scan_load(sp_scan* scan, int i, double z, const double* field, char* msg,
                 size_t msglen) {
  if (!scan) return SP_VALUE;
  return guarded(msg, msglen, [&] { scan->impl.load(i, z, field); });
}

int sp_scan_advance(sp_scan* scan, const double* targets, long work_limit, double* rows,
                    long max_rows, sp_march_result* out, int* status, char* msg,
                    size_t msglen) {
  if (!scan || !targets || !out || !status) return SP_VALUE;
  const int n = scan->impl.size();
  std::vector<std::vector<double>> held;
  std::vector<std::string> messages;
  const int code = guarded(msg, msglen, [&] {
    scan->impl.advance(targets, work_limit, held, out, status, messages);
  });
  if (code != SP_OK) return code;
  for (int i = 0; i < n; ++i) {
    const long count = static_cast<long>(held[i].size() / 4);
    const long kept = rows ? std::min(count, max_rows) : 0;
    if (kept > 0) std::memcpy(rows + 4 * max_rows * i, held[i].data(), sizeof(double) * 4 * kept);
    out[i].n_rows = static_cast<int>(kept);
  }
  for (int i = 0; i < n; ++i)
    if (status[i] != SP_OK) {
      write_message(msg, msglen, ("scan instance " + std::to_string(i) + ": " + messages[i]).c_str());
      return status[i];
    }
  return SP_OK;
}

int sp_scan_state(sp_scan* scan, int i, double* out, char* msg, size_t msglen) {
  if (!scan || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] { scan->impl.state(i, out); });
}

int sp_scan_timing_of(sp_scan* scan, sp_scan_timing* out) {
  if (!scan || !out) return SP_VALUE;
  const sp::ScanTiming& t = scan->impl.timing();
  out->rounds = t.rounds;
  out->device_ms = t.device_ms;
  out->host_ms = t.host_ms;
  out->device_bytes = scan->impl.device_bytes();
  return SP_OK;
}

int sp_scan_schedule(sp_scan* scan, char* out, size_t len) {
  if (!scan || !out) return SP_VALUE;
  const std::string& s = scan->impl.schedule();
  if (s.size() + 1 > len) return SP_VALUE;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return SP_OK;
}

int sp_workspace_export(sp_workspace* ws, const double* power_weight, int optics,
                        double* out, char* msg, size_t msglen) {
  if (!ws || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] { ws->impl.export_state(power_weight, optics != 0, out); });
}

int sp_workspace_statistics(sp_workspace* ws, unsigned long long flags, const double* windows,
                            int n_windows, double* out, int capacity, int* count, char* msg,
                            size_t msglen) {
  if (!ws || !count || (capacity > 0 && !out)) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    const auto values = ws->impl.statistics(flags, windows, n_windows, sp::StatExtras{});
    SP_CHECK(static_cast<long>(values.size()) <= capacity, "the statistics output is too small");
    std::copy(values.begin(), values.end(), out);
    *count = static_cast<int>(values.size());
  });
}

int sp_workspace_statistics_request(sp_workspace* ws, const sp_statistics_request* request,
                                    double* out, int capacity, int* count, char* msg,
                                    size_t msglen) {
  if (!ws || !request || !count || (capacity > 0 && !out)) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    sp::StatExtras extras;
    extras.on_axis = request->on_axis;
    extras.rows = request->on_axis_rows;
    extras.intensity_factor = request->intensity_factor;
    const auto values =
        ws->impl.statistics(request->flags, request->windows, request->n_windows, extras);
    SP_CHECK(static_cast<long>(values.size()) <= capacity, "the statistics output is too small");
    std::copy(values.begin(), values.end(), out);
    *count = static_cast<int>(values.size());
  });
}

int sp_workspace_bind_monitor(sp_workspace* ws, const double* weights, const double* columns,
                              char* msg, size_t msglen) {
  if (!ws) return SP_VALUE;
  return guarded(msg, msglen, [&] { ws->impl.bind_monitor(weights, columns); });
}

int sp_workspace_monitor_reading(sp_workspace* ws, double* out, char* msg, size_t msglen) {
  if (!ws || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] { ws->impl.monitor_reading(out); });
}

int sp_workspace_counters(sp_workspace* ws, sp_counters* out) {
  if (!ws || !out) return SP_VALUE;
  *out = ws->impl.counters();
  return SP_OK;
}

int sp_workspace_profile(sp_workspace* ws, int enable) {
  if (!ws) return SP_VALUE;
  ws->impl.set_profile(enable != 0);
  return SP_OK;
}

int sp_workspace_timing(sp_workspace* ws, double* trial_us, long* trials) {
  if (!ws || !trial_us || !trials) return SP_VALUE;
  *trial_us = ws->impl.trial_us();
  *trials = ws->impl.timed();
  return SP_OK;
}

int sp_testing_source(sp_workspace* ws, double z, const double* lab, double* out,
                      double* streams, double* modal, char* msg, size_t msglen) {
  if (!ws) return SP_VALUE;
  return guarded(msg, msglen, [&] { ws->impl.source(z, lab, out, streams, modal); });
}

int sp_testing_buffer(sp_workspace* ws, const char* name, double* out, long count, long* size,
                      char* msg, size_t msglen) {
  if (!ws || !name || !size) return SP_VALUE;
  return guarded(msg, msglen, [&] { *size = ws->impl.buffer(name, out, count); });
}

int sp_workspace_phases(sp_workspace* ws, double* us, int capacity, int* count, char* names,
                        size_t names_len) {
  if (!ws || !count || (capacity > 0 && !us)) return SP_VALUE;
  const auto phases = ws->impl.phases();
  std::string joined;
  int n = 0;
  for (const auto& ph : phases) {
    if (n < capacity) us[n] = ph.second;
    if (!joined.empty()) joined += ",";
    joined += ph.first;
    n += 1;
  }
  *count = n;
  write_message(names, names_len, joined.c_str());
  return n > capacity ? SP_VALUE : SP_OK;
}

int sp_rk_trace(const sp_rk_scenario* scenario, char* out, size_t outlen) {
  if (!scenario || !out) return SP_VALUE;
  std::string trace;
  try {
    trace = sp::march::run_trace(*scenario);
  } catch (const std::exception& e) {
    std::snprintf(out, outlen, "%s", e.what());
    return SP_BACKEND;
  }
  if (trace.size() + 1 > outlen) return SP_VALUE;
  std::memcpy(out, trace.c_str(), trace.size() + 1);
  return SP_OK;
}

int sp_adams_trace(const sp_adams_scenario* scenario, char* out, size_t outlen) {
  if (!scenario || !out) return SP_VALUE;
  std::string trace;
  try {
    trace = sp::adams::run_adams_trace(*scenario);
  } catch (const std::exception& e) {
    std::snprintf(out, outlen, "%s", e.what());
    return SP_BACKEND;
  }
  if (trace.size() + 1 > outlen) return SP_VALUE;
  std::memcpy(out, trace.c_str(), trace.size() + 1);
  return SP_OK;
}

int sp_testing_fft(const double* in, double* out, long n, long batch, int sign, char* msg,
                   size_t msglen) {
  if (!in || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    SP_CHECK(sign == -1 || sign == 1, "sign must be -1 or +1");
    SP_CHECK(n > 0 && batch > 0, "n and batch must be positive");
    sp::FftPlan plan;
    plan.init(static_cast<std::size_t>(n), static_cast<std::size_t>(batch));
    const std::size_t bytes = sizeof(sp::cplx) * static_cast<std::size_t>(n * batch);
    DeviceScope scope;
    sp::cplx* data = scope.alloc<sp::cplx>(bytes);
    sp::cplx* scratch = scope.alloc<sp::cplx>(bytes);
    SP_HIP(hipMemcpy(data, in, bytes, hipMemcpyHostToDevice));
    plan.exec(data, scratch, sign, scope.stream);
    SP_HIP(hipStreamSynchronize(scope.stream));
    SP_HIP(hipMemcpy(out, data, bytes, hipMemcpyDeviceToHost));
  });
}

int sp_testing_gemm(long M, long N, long K, const double* A, const double* B, double* C,
                    char* msg, size_t msglen) {
  if (!A || !B || !C) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    SP_CHECK(M > 0 && N > 0 && K > 0 && M * K < (1L << 28) && K * N < (1L << 28) &&
                 M * N < (1L << 28),
             "GEMM dimensions must be positive and below 2^28 elements per operand");
    DeviceScope scope;
    double* a = scope.alloc<double>(sizeof(double) * M * K);
    double* b = scope.alloc<double>(sizeof(double) * K * N);
    double* c = scope.alloc<double>(sizeof(double) * M * N);
    SP_HIP(hipMemcpy(a, A, sizeof(double) * M * K, hipMemcpyHostToDevice));
    SP_HIP(hipMemcpy(b, B, sizeof(double) * K * N, hipMemcpyHostToDevice));
    sp::launch_dgemm(static_cast<int>(M), static_cast<int>(N), static_cast<int>(K), a, K, b, N,
                     c, N, scope.stream);
    SP_HIP(hipStreamSynchronize(scope.stream));
    SP_HIP(hipMemcpy(C, c, sizeof(double) * M * N, hipMemcpyDeviceToHost));
  });
}

// ------------------------------------------------------------------ sessions

namespace {
void fill_batch(const sp::session::Batch& b, sp_session_batch* out, double* attempts,
                long capacity, double* stats, int stats_capacity) {
  std::memset(out, 0, sizeof(*out));
  out->sequence = b.sequence;
  const sp::session::Event& e = b.event;
  out->event.kind = static_cast<int>(e.kind);
  out->event.observation = static_cast<int>(e.observation);
  out->event.z = e.z;
  out->event.plane_index = e.plane_index;
  out->event.has_stats = e.has_stats ? 1 : 0;
  out->event.ticket = e.ticket ? 1 : 0;
  out->event.generation = e.generation;
  out->event.outcome = static_cast<int>(e.outcome);
  out->event.reason = static_cast<int>(e.reason);
  out->event.category = static_cast<int>(e.category);
  out->event.stage = static_cast<int>(e.stage);
  out->event.floor_proposal = e.floor_proposal;
  std::snprintf(out->event.message, sizeof(out->event.message), "%s", e.message.c_str());
  const long rows = static_cast<long>(b.attempts.size() / sp::kWideColumns);
  SP_CHECK(rows <= capacity, "the attempts buffer is smaller than the batch's chunk");
  if (rows > 0) std::memcpy(attempts, b.attempts.data(), sizeof(double) * b.attempts.size());
  out->n_attempts = rows;
  const int n_stats = static_cast<int>(b.stats.size());
  SP_CHECK(n_stats <= stats_capacity, "the statistics buffer is smaller than the batch's");
  if (n_stats > 0) std::memcpy(stats, b.stats.data(), sizeof(double) * b.stats.size());
  out->n_stats = n_stats;
}
}  // namespace

int sp_session_start(sp_workspace* ws, double z, const double* field,
                     const sp_session_plan* plan, sp_session** out, char* msg, size_t msglen) {
  if (!ws || !plan || !out) {
    write_message(msg, msglen, "start needs a workspace, a plan and an output handle");
    return SP_VALUE;
  }
  *out = nullptr;
  return guarded(msg, msglen, [&] {
    SP_CHECK(plan->n_planes >= 1 && plan->planes, "a session needs its planes");
    SP_CHECK(plan->n_windows >= 0 && (plan->n_windows == 0 || plan->windows),
             "windows are (lo, hi) pairs");
    {
      std::lock_guard<std::mutex> lock(ws->lease_mu);
      if (ws->doomed) throw sp::Error(sp::kState, "the workspace is closed");
      bool expected = false;
      if (!ws->leased.compare_exchange_strong(expected, true))
        throw sp::Error(sp::kState, "the workspace is leased by an open session");
    }
    try {
      sp::session::Plan p;
      p.planes.assign(plan->planes, plan->planes + plan->n_planes);
      p.launch = plan->launch != 0;
      p.step_planes = plan->step_planes;
      p.statistics = plan->statistics;
      if (plan->n_windows)
        p.windows.assign(plan->windows, plan->windows + 2 * plan->n_windows);
      if (plan->attempt_capacity > 0) p.attempt_capacity = plan->attempt_capacity;
      if (plan->call_attempts > 0) p.call_attempts = plan->call_attempts;
      if (plan->call_seconds > 0.0) p.call_seconds = plan->call_seconds;
      auto session = std::make_unique<sp::session::Session>(ws->impl, ws->n_modes,
                                                            ws->n_omega, p, z, field);
      *out = new sp_session{ws, std::move(session)};
    } catch (...) {
      ws->leased.store(false);
      throw;
    }
  });
}

int sp_session_advance(sp_session* s, long work_limit, double deadline, sp_session_batch* out,
                       double* attempts, long capacity, double* stats, int stats_capacity,
                       char* msg, size_t msglen) {
  if (!s || !out || (capacity > 0 && !attempts) || (stats_capacity > 0 && !stats))
    return SP_VALUE;
  return guarded(msg, msglen, [&] {
    const sp::session::Batch b = s->impl->advance(work_limit, deadline);
    fill_batch(b, out, attempts, capacity, stats, stats_capacity);
  });
}

int sp_session_acknowledge(sp_session* s, unsigned long long sequence, char* msg,
                           size_t msglen) {
  if (!s) return SP_VALUE;
  return guarded(msg, msglen, [&] { s->impl->acknowledge(sequence); });
}

void sp_session_request_stop(sp_session* s) {
  if (s) s->impl->request_stop();
}

int sp_session_status_of(sp_session* s, sp_session_status* out) {
  if (!s || !out) return SP_VALUE;
  const sp::session::Status st = s->impl->status();
  out->z = st.z;
  out->plane_index = st.plane_index;
  out->attempts = st.attempts;
  out->rejected = st.rejected;
  out->state = static_cast<int>(st.state);
  out->outcome = static_cast<int>(st.outcome);
  out->generation = st.generation;
  out->march_calls = st.march_calls;
  out->batches = st.batches;
  out->exports = st.exports;
  out->export_bytes = st.export_bytes;
  return SP_OK;
}

int sp_session_export(sp_session* s, unsigned long long sequence, unsigned long long generation,
                      const double* power_weight, int optics, double* out, char* msg,
                      size_t msglen) {
  if (!s || !out) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    s->impl->export_field(sequence, generation, power_weight, optics != 0, out);
  });
}

int sp_session_fail_outside(sp_session* s, const char* reason, sp_session_batch* out,
                            double* attempts, long capacity, char* msg, size_t msglen) {
  if (!s || !out || (capacity > 0 && !attempts)) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    sp::session::Batch b;
    if (!s->impl->fail_outside(reason ? reason : "failed outside the session", &b))
      throw sp::session::LifecycleError("a batch is pending or the session has ended");
    fill_batch(b, out, attempts, capacity, nullptr, 0);
  });
}

void sp_session_close(sp_session* s) {
  if (!s) return;
  s->impl->close();
  sp_workspace* ws = s->ws;
  delete s;
  bool release = false;
  {
    std::lock_guard<std::mutex> lock(ws->lease_mu);
    ws->leased.store(false);
    release = ws->doomed;
  }
  if (release) delete ws;
}

// ---------------------------------------------------------- learned surrogate

int sp_fno_create(const sp_fno_shape* shape, int max_batch, sp_fno** out, char* msg,
                  size_t msglen) {
  if (!shape || !out) return SP_VALUE;
  *out = nullptr;
  return guarded(msg, msglen, [&] {
    sp::fno::Shape s;
    s.n_params = shape->n_params;
    s.width = shape->width;
    s.modes = shape->modes;
    s.layers = shape->layers;
    s.hidden = shape->hidden;
    s.n = shape->n;
    *out = new sp_fno{std::make_unique<sp::fno::Network>(s, max_batch)};
  });
}

void sp_fno_destroy(sp_fno* f) { delete f; }

long sp_fno_parameter_count(const sp_fno* f) { return f ? f->impl->parameter_count() : 0; }

int sp_fno_set_parameters(sp_fno* f, const float* flat, char* msg, size_t msglen) {
  if (!f || !flat) return SP_VALUE;
  return guarded(msg, msglen, [&] { f->impl->set_parameters(flat); });
}

int sp_fno_get_parameters(sp_fno* f, float* flat, char* msg, size_t msglen) {
  if (!f || !flat) return SP_VALUE;
  return guarded(msg, msglen, [&] { f->impl->get_parameters(flat); });
}

int sp_fno_forward(sp_fno* f, const float* spectra, const float* params, int batch, float* out,
                   char* msg, size_t msglen) {
  if (!f) return SP_VALUE;
  return guarded(msg, msglen, [&] { f->impl->forward(spectra, params, batch, out); });
}

int sp_fno_gradient(sp_fno* f, const float* spectra, const float* params, const float* target,
                    int batch, double* loss, float* grad, char* msg, size_t msglen) {
  if (!f || !loss) return SP_VALUE;
  return guarded(msg, msglen,
                 [&] { *loss = f->impl->gradient(spectra, params, target, batch, grad); });
}

int sp_fno_load_dataset(sp_fno* f, const float* spectra, const float* params,
                        const float* targets, long n, char* msg, size_t msglen) {
  if (!f) return SP_VALUE;
  return guarded(msg, msglen, [&] { f->impl->load_dataset(spectra, params, targets, n); });
}

int sp_fno_train(sp_fno* f, const int* indices, long steps, int batch, float lr, float* losses,
                 double* device_ms, char* msg, size_t msglen) {
  if (!f) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    f->impl->train(indices, steps, batch, lr, losses);
    if (device_ms) *device_ms = f->impl->train_ms();
  });
}

int sp_fno_reset_optimizer(sp_fno* f) {
  if (!f) return SP_VALUE;
  char msg[8];
  return guarded(msg, sizeof msg, [&] { f->impl->reset_optimizer(); });
}

int sp_testing_gemm_bf16(long M, long N, long K, const float* A, const float* B, float* C,
                         char* msg, size_t msglen) {
  if (!A || !B || !C) return SP_VALUE;
  return guarded(msg, msglen, [&] {
    SP_CHECK(M > 0 && N > 0 && K > 0 && M < (1L << 20) && N < (1L << 20) && K < (1L << 20),
             "GEMM dimensions must be positive");
    sp::fno::testing_gemm_bf16(static_cast<int>(M), static_cast<int>(N), static_cast<int>(K), A,
                               B, C);
  });
}

const char* sp_version(void) { return "solprop_mi355x 0.1 gfx950 fp64"; }

}  // extern "C"
