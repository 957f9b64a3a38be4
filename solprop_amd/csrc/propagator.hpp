// What the C ABI asks of a workspace, whichever route and march it runs: the
// envelope route under the embedded Runge-Kutta march (workspace.hpp) or the modal
// field route under the exponential Adams march (field_ws.hpp).
#pragma once

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

extern "C" {
#include "../../include/solprop_mi355x.h"
}

namespace sp {

// The bounds of one march call (the reference's per-call bounds, docs/dev/kernel-contract.md
// "What a native call takes"): the target, the most attempts, the accepted steps after
// which the caller has work, the call's wall time and a stop flag, each checked between
// attempts.
struct Bounds {
  double target = 0.0;
  long attempts = 1;
  long accepted = 0;      // 0: no such bound
  double seconds = -1.0;  // < 0: none
  const std::atomic<bool>* stop = nullptr;
};

// Attempt rows: 4 columns (end, step, norm, accepted) or kWideColumns with the
// attempt's order and its wall time since the call began, s (the norm then written
// as infinity when it is not finite).
constexpr int kWideColumns = 6;

// What a march's host records of a bounded call, and when it halts it.
struct AttemptLog {
  std::vector<double>* rows = nullptr;
  int columns = 4;
  long limit = -1, accept_limit = 0, attempts = 0, accepted = 0;
  double seconds = -1.0;
  const std::atomic<bool>* stop = nullptr;
  std::chrono::steady_clock::time_point began;
  int order = 0;

  void begin(std::vector<double>* r, int cols, const Bounds& b) {
    rows = r;
    columns = cols;
    limit = b.attempts;
    accept_limit = b.accepted;
    seconds = b.seconds;
    stop = b.stop;
    attempts = accepted = 0;
    began = std::chrono::steady_clock::now();
  }
  double elapsed() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - began).count();
  }
  void record(double end, double dz, double norm, bool acc) {
    if (rows) {
      rows->push_back(end);
      rows->push_back(dz);
      rows->push_back(columns == kWideColumns && !std::isfinite(norm) ? INFINITY : norm);
      rows->push_back(acc ? 1.0 : 0.0);
      if (columns == kWideColumns) {
        rows->push_back(static_cast<double>(order));
        rows->push_back(elapsed());
      }
    }
    attempts += 1;
    accepted += acc ? 1 : 0;
  }
  bool halt() const {
    if (limit >= 0 && attempts >= limit) return true;
    if (accept_limit > 0 && accepted >= accept_limit) return true;
    if (attempts > 0 && seconds >= 0.0 && elapsed() >= seconds) return true;
    return stop && stop->load(std::memory_order_acquire);
  }
};

// What the peak intensity reads beside the state: the on-axis synthesis, `rows` (1 or 2)
// rows of the C columns, and the intensity factor (the reference's `on_axis` and
// `intensity_factor`, native/src/general/statistics.cpp:517-603).
struct StatExtras {
  const double* on_axis = nullptr;
  int rows = 0;
  double intensity_factor = 1.0;
};

class Propagator {
 public:
  virtual ~Propagator() = default;
  virtual void load(double z, const double* field) = 0;
  // One trial from the committed state (a probe): the scaled norm.
  virtual double probe_trial(double z, double dz) = 0;
  virtual void read_trial(double* out) = 0;
  virtual void read_state(double* out) = 0;
  virtual void export_state(const double* weight, bool optics, double* out) = 0;
  // A bounded march call: rows of `columns` (4 or kWideColumns) doubles per attempt;
  // true when the target was reached.
  virtual bool advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                               sp_march_result* r) = 0;
  bool advance(double target, long work_limit, std::vector<double>& rows,
               sp_march_result* r) {
    Bounds b;
    b.target = target;
    b.attempts = work_limit;
    return advance_bounded(b, 4, rows, r);
  }
  // The order of the method's attempts when it is fixed (a Runge-Kutta pair), else 0.
  virtual int fixed_order() const { return 0; }
  virtual sp_counters counters() const = 0;
  virtual void set_profile(bool on) = 0;
  virtual double trial_us() const = 0;
  virtual long timed() const = 0;
  // The source at a lab-frame field at z (complex (C, N)); with `streams` (3, C, T)
  // and `modal` (C, T) given, the route's time-domain intermediates as well.
  virtual void source(double z, const double* lab, double* out, double* streams,
                      double* modal) {
    (void)z; (void)lab; (void)out; (void)streams; (void)modal;
    throw std::invalid_argument("source evaluation is built on the field route only");
  }
  // A named device buffer copied to the host (`count` doubles at most); the number of
  // doubles it holds.
  virtual long buffer(const char* name, double* out, long count) {
    (void)name; (void)out; (void)count;
    throw std::invalid_argument("buffer readback is built on the field route only");
  }
  // Plane statistics of the committed state (stats.hpp's flags, windows as (lo, hi)
  // pairs in rad/s, the peak intensity's inputs), packed in the reference's order.
  virtual std::vector<double> statistics(std::uint64_t flags, const double* windows,
                                         int n_windows, const StatExtras& extras) = 0;
  // The boundary monitor (weights (4, n_t) on the coarse axis, or null to unbind; the
  // columns' weights or null), and its reading of the committed state with the split
  // collar's sums since the last reading (10 doubles).
  virtual void bind_monitor(const double* weights, const double* columns) = 0;
  virtual void monitor_reading(double* out) = 0;
  // The average device time of each kernel (or transform) of a trial, microseconds.
  virtual std::vector<std::pair<const char*, double>> phases() const { return {}; }
};

}  // namespace sp
