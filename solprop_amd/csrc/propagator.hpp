// What the C ABI asks of a workspace, whichever route and march it runs: the
// envelope route under the embedded Runge-Kutta march (workspace.hpp) or the modal
// field route under the exponential Adams march (field_ws.hpp).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

extern "C" {
#include "../../include/solprop_mi355x.h"
}

namespace sp {

class Propagator {
 public:
  virtual ~Propagator() = default;
  virtual void load(double z, const double* field) = 0;
  // One trial from the committed state (a probe): the scaled norm.
  virtual double probe_trial(double z, double dz) = 0;
  virtual void read_trial(double* out) = 0;
  virtual void read_state(double* out) = 0;
  virtual void export_state(const double* weight, bool optics, double* out) = 0;
  virtual bool advance(double target, long work_limit, std::vector<double>& rows,
                       sp_march_result* r) = 0;
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
  // pairs in rad/s), packed in the reference's order.
  virtual std::vector<double> statistics(std::uint64_t flags, const double* windows,
                                         int n_windows) = 0;
  // The boundary monitor (weights (4, n_t) on the coarse axis, or null to unbind; the
  // columns' weights or null), and its reading of the committed state with the split
  // collar's sums since the last reading (10 doubles).
  virtual void bind_monitor(const double* weights, const double* columns) = 0;
  virtual void monitor_reading(double* out) = 0;
  // The average device time of each kernel (or transform) of a trial, microseconds.
  virtual std::vector<std::pair<const char*, double>> phases() const { return {}; }
};

}  // namespace sp
