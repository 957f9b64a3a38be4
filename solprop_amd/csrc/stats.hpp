// Plane statistics on the committed state (the reference's saved-plane statistics,
// native/src/general/statistics.cpp:106-360 and src/solprop/stats/definitions.py):
// per column the spectral centroid, the energy and the energy inside spectral windows,
// from one reduction over the stored bins; per mode the peak instantaneous power and
// the peak of the total power, from the columns' analytic signals on the coarse grid.
#pragma once

#include <cstdint>
#include <vector>

#include "device.hpp"
#include "fft.hpp"
#include "propagator.hpp"

namespace sp {

// The flags of native/include/solprop_native/problem/problem.hpp:44-60 this build forms
// (the plasma pair on the field routes, by the workspace).
enum : std::uint64_t {
  kStatOmega0 = 1ULL << 0U,
  kStatEnergy = 1ULL << 1U,
  kStatEnergyWindow = 1ULL << 2U,
  kStatPeakPower = 1ULL << 3U,
  kStatPeakPowerWindow = 1ULL << 4U,
  kStatPeakPowerTotal = 1ULL << 5U,
  // The total power trace sum_modes |a_m(t)|^2 (n_t values; the host measures the width,
  // native/src/general/statistics.cpp:80-93,291-294), packed after the plasma pair.
  kStatFwhm = 1ULL << 6U,
  // The field routes' plasma column at the committed state (read by the workspace).
  kStatElectronDensity = 1ULL << 7U,
  kStatPeakIonisationRate = 1ULL << 8U,
  // The largest on-axis intensity: the on-axis synthesis applied to the modes' analytic
  // signals, max over t of sum over its rows of the squared modulus, times the intensity
  // factor (native/src/general/statistics.cpp:517-603); packed after the plasma pair.
  kStatPeakIntensity = 1ULL << 9U,
  // The quadrature estimate of the last accepted step's source evaluation: its response
  // projected with the embedded coarse rule against the full one (trial_finish.inc:
  // 921-1028), read by the workspace; packed after the peak intensity.
  kStatQuadrature = 1ULL << 10U,
  // Each node's end-of-window ionised fraction and largest rate, (Q, 2), one channel
  // (statistics.cpp:636-650); packed last, after the power trace.
  kStatPopulations = 1ULL << 11U,
  // The plasma column's integration diagnostics: six maxima over the nodes (the exposure
  // increments' dip and rise, the loss current's work defect and the electrons freed,
  // the free-current integrand's level in the first 2K samples and its peak;
  // src/solprop/backends/numpy_ref/passes/plasma.py:488-552, statistics.cpp:184-186,
  // 684-697); packed last, after the populations.
  kStatPlasmaIntegration = 1ULL << 12U,
};
constexpr std::uint64_t kStatPlasma = kStatElectronDensity | kStatPeakIonisationRate;
// Every statistic the field routes' plasma column serves.
constexpr std::uint64_t kStatColumn = kStatPlasma | kStatPopulations | kStatPlasmaIntegration;
// The node of largest ionised fraction at the window's end (the first such node) and
// its largest rate: out[0] that fraction, out[1] the rate (numpy_ref/statistics.py:502-546).
void launch_plasma_densest(const double* frac, const double* rate, int Q, int T, double* out,
                           hipStream_t s);
// Every node's end-of-window fraction and largest rate into out (2 Q).
void launch_plasma_populations(const double* frac, const double* rate, int Q, int T,
                               double* out, hipStream_t s);
constexpr std::uint64_t kStatBuilt = kStatOmega0 | kStatEnergy | kStatEnergyWindow |
                                     kStatPeakPower | kStatPeakPowerWindow |
                                     kStatPeakPowerTotal | kStatFwhm | kStatPeakIntensity;
constexpr int kMaxStatWindows = 16;

// The boundary monitor's sums (src/solprop/monitor.py:493-606) over the time samples of
// C columns: P(t) = sum_c w_c |s_c(t)|^2 (w_c one without column weights), then with the
// monitor's four weight rows (low guard, high guard, collar, rate) either
//   kMonitorReading: [sum P, low.P, high.P, collar.P, rate.P], or
//   kMonitorSplit:   [sum P (1 - e^{-rate h}), sum H, low.H, high.H, collar.H] with
//                    H = P e^{-rate h / 2} (what the split takes, and the power it finds
//                    halfway through its factor),
// and kMonitorSplit adds its first sum to acc[0] and stores the other four in
// acc[1..4]; kMonitorReading writes the five to acc. Samples are real (stride one
// double), complex taken whole, or complex taken by their real part.
enum MonitorMode { kMonitorReading = 0, kMonitorSplit = 1 };
enum MonitorSamples { kSamplesReal = 0, kSamplesComplex = 1, kSamplesRealPart = 2 };
// The step h is read from the device at `h_at` when given (a record's accepted step,
// inside a graph), else taken from `h`. `partials` holds 5 doubles a block.
void launch_monitor_sums(const double* samples, MonitorSamples layout, int C, int n_t,
                         const double* columns, const double* weights, double h,
                         const double* h_at, MonitorMode mode, double* partials, double* acc,
                         hipStream_t s);
constexpr int kMonitorThreads = 256;
inline int monitor_blocks(int n_t) { return (n_t + kMonitorThreads - 1) / kMonitorThreads; }

// A bound monitor's device buffers: its weights (4, n_t), the columns' weights (C) or
// none, block partials, the last reading (5) and the split's running removal and last
// found energies (5). Binding uploads the weights and empties the split's sums.
class MonitorBuffers {
 public:
  ~MonitorBuffers();
  void bind(const double* weights, const double* columns, int C, int n_t, hipStream_t s);
  void unbind() { bound_ = false; }
  // Empty the reading and the split's sums (a state was loaded).
  void reset(hipStream_t s) {
    if (acc_) SP_HIP(hipMemsetAsync(acc_, 0, 10 * sizeof(double), s));
  }
  bool bound() const { return bound_; }
  const double* weights() const { return weights_; }
  const double* columns() const { return has_columns_ ? columns_ : nullptr; }
  double* partials() const { return partials_; }
  double* reading() const { return acc_; }
  double* split() const { return acc_ + 5; }
  // The last reading and the split's sums to the host (10 doubles), the split's
  // running removal then emptied (the host's accumulator drains it).
  void drain(double* out, hipStream_t s);

 private:
  bool bound_ = false, has_columns_ = false;
  int C_ = 0, n_t_ = 0;
  double *weights_ = nullptr, *columns_ = nullptr, *partials_ = nullptr, *acc_ = nullptr;
};

class PlaneStatistics {
 public:
  // C columns of N stored bins on a coarse grid of n_t samples spaced dt; field (real
  // time, N = n_t/2 + 1) or envelope (N = n_t) layout; omega (N) on the device.
  PlaneStatistics(int C, int N, int n_t, bool field, double dt, const double* omega);
  ~PlaneStatistics();
  PlaneStatistics(const PlaneStatistics&) = delete;
  PlaneStatistics& operator=(const PlaneStatistics&) = delete;
  // The requested statistics of the state y (C, N) in the reference's packing order:
  // omega0 (C), energy (C), energy_window (n_windows x C, window-major), peak_power (C),
  // peak_power_window (n_windows x C), peak_power_total (1). `windows` holds n_windows (lo, hi) pairs in rad/s. With
  // kStatFwhm the total power trace (n_t) goes to `trace`, which the caller packs after
  // the statistics that precede it.
  // With kStatPeakIntensity the largest on-axis intensity goes to `intensity`, from
  // `extras`' on-axis rows and factor.
  std::vector<double> run(const cplx* y, std::uint64_t flags, const double* windows,
                          int n_windows, hipStream_t s, std::vector<double>* trace = nullptr,
                          const StatExtras* extras = nullptr, double* intensity = nullptr);
  std::size_t bytes() const { return bytes_ + plan_.bytes(); }
  // The monitor's five sums of the state y (kMonitorReading) into acc (device, 5).
  void reading(const cplx* y, const double* columns, const double* weights, double* partials,
               double* acc, hipStream_t s);

 private:
  int C_, N_, n_t_;
  bool field_;
  double dt_;
  const double* omega_;
  FftPlan plan_;
  cplx *time_ = nullptr, *scratch_ = nullptr;
  double *windows_ = nullptr, *sums_ = nullptr, *peaks_ = nullptr, *trace_ = nullptr;
  double* on_axis_ = nullptr;  // (2, C) the peak intensity's synthesis rows
  double* intensity_ = nullptr;  // (peak blocks) its block maxima
  double* host_ = nullptr;
  int peak_blocks_ = 0;
  std::size_t bytes_ = 0;
};

}  // namespace sp
