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

namespace sp {

// The flags of native/include/solprop_native/problem/problem.hpp:44-60 this build forms.
enum : std::uint64_t {
  kStatOmega0 = 1ULL << 0U,
  kStatEnergy = 1ULL << 1U,
  kStatEnergyWindow = 1ULL << 2U,
  kStatPeakPower = 1ULL << 3U,
  kStatPeakPowerTotal = 1ULL << 5U,
};
constexpr std::uint64_t kStatBuilt =
    kStatOmega0 | kStatEnergy | kStatEnergyWindow | kStatPeakPower | kStatPeakPowerTotal;
constexpr int kMaxStatWindows = 16;

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
  // peak_power_total (1). `windows` holds n_windows (lo, hi) pairs in rad/s.
  std::vector<double> run(const cplx* y, std::uint64_t flags, const double* windows,
                          int n_windows, hipStream_t s);
  std::size_t bytes() const { return bytes_ + plan_.bytes(); }

 private:
  int C_, N_, n_t_;
  bool field_;
  double dt_;
  const double* omega_;
  FftPlan plan_;
  cplx *time_ = nullptr, *scratch_ = nullptr;
  double *windows_ = nullptr, *sums_ = nullptr, *peaks_ = nullptr;
  double* host_ = nullptr;
  int peak_blocks_ = 0;
  std::size_t bytes_ = 0;
};

}  // namespace sp
