// The free-space Cartesian route: an envelope on an (nx, ny, nt) grid in a homogeneous
// medium, marched by the low-storage RK4IP replay.
//
// The reference's free-space geometry (`prop_free`; its `kerr_beam_*` cases, four fixed
// RK4IP steps of a fused-silica Kerr beam, docs/dev/benchmarks.md:529-537; the
// low-storage engine, native/src/general/trial_low_storage.inc and
// src/solprop/backends/numpy_ref/engines.py:236-520). The state is the spectrum
// Y(ky, kx, w) = FFT3[E(y, x, t)] (unnormalised), row-major with w fastest: column
// iy nx + ix, x the faster transverse axis (src/solprop/translate/free.py:137). Its linear
// operator is one constant table D(ky, kx, w) (the flows exp(D dz) formed per element),
// its source S(Y) = p(ky, kx, w) FFT3[gamma |E|^2 E] with E = FFT3^-1[Y] / V, V = nx ny nt:
// the analytic (envelope) Kerr response at every point of the volume, its conjugate term
// for a response that keeps it. A source evaluation is five passes over the volume, each
// a Stockham transform of whole axis sequences staged in LDS: the t rows' inverse from the
// spectrum, the x axis's inverse, the y axis's inverse, response and forward in one pass,
// the x axis's forward, the t rows' forward with the prefactor applied as they are stored
// (the strided axes with a workgroup owning up to 16 neighbouring sequences, so that every
// load and store moves contiguous runs).
#pragma once

#include <utility>
#include <vector>

#include "device.hpp"
#include "fft.hpp"
#include "field.hpp"
#include "march.hpp"
#include "propagator.hpp"

namespace sp {

class CartesianWorkspace final : public Propagator {
 public:
  explicit CartesianWorkspace(const sp_problem_desc& desc);
  ~CartesianWorkspace() override;
  CartesianWorkspace(const CartesianWorkspace&) = delete;
  CartesianWorkspace& operator=(const CartesianWorkspace&) = delete;

  void load(double z, const double* field) override;
  double probe_trial(double z, double dz) override;
  void read_trial(double* out) override;
  void read_state(double* out) override;
  void export_state(const double* weight, bool optics, double* out) override;
  bool advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                       sp_march_result* r) override;
  sp_counters counters() const override;
  std::vector<double> statistics(std::uint64_t flags, const double* windows,
                                 int n_windows, const StatExtras& extras) override;
  void bind_monitor(const double* weights, const double* columns) override;
  void monitor_reading(double* out) override;
  // Profiling: trials launched directly with an event after every launch, each
  // launch's device time added to its phase.
  void set_profile(bool on) override { profile_ = on; }
  double trial_us() const override { return timed_ ? trial_ms_ * 1e3 / timed_ : 0.0; }
  long timed() const override { return timed_; }
  std::vector<std::pair<const char*, double>> phases() const override;
  // The source at a spectrum (complex (nx ny, nt)) into `out` (testing).
  void source(double z, const double* lab, double* out, double* streams,
              double* modal) override;
  long buffer(const char* name, double* out, long count) override;

 private:
  template <typename T>
  T* alloc(std::size_t n);
  void release();
  void enqueue_source(const cplx* spectrum);  // into B_.source
  void enqueue_step();
  double trial(double z, double dz);
  void wait();
  void mark(int phase);  // profiling: an event after the launch of `phase`
  void collect();        // profiling: the marked launches' times into their phases

  int nx_ = 0, ny_ = 0, nt_ = 0;
  int gt_ = 0, gy_ = 0;  // log2 of the sequences a workgroup owns: t rows, y axis
  std::size_t V_ = 0;
  DeviceProblem P_{};
  FieldBuffers B_{};
  std::vector<void*> allocations_;
  std::size_t device_bytes_ = 0;
  hipStream_t stream_ = nullptr;
  FftPlan plan_t_, plan_y_, plan_x_;
  cplx* field_ = nullptr;  // the (y, x, t) samples the transforms and the response act on
  cplx *acc_ = nullptr, *prev_ = nullptr, *save_ = nullptr;
  double kerr_self_ = 1.0, kerr_conj_ = 0.0, kerr_scale_ = 0.0;
  AdamsRecord* rec_dev_ = nullptr;   // (3), the flows' weights in records 0 and 1
  AdamsRecord* rec_host_ = nullptr;  // (3) pinned
  double* partials_ = nullptr;
  double* norms_host_ = nullptr;     // (3) mapped
  hipGraphExec_t graph_ = nullptr;
  hipEvent_t done_ = nullptr;
  bool use_graphs_ = true;
  std::vector<double> edges_, steps_;
  march::State state_;
  bool loaded_ = false;
  double stiffness_ = 0.0;
  long trials_ = 0, evaluations_ = 0, accepted_ = 0, graph_launches_ = 0;
  static constexpr int kPhases = 10;
  static constexpr int kMarks = 64;  // events a profiled trial records (it needs 27)
  bool profile_ = false, marking_ = false;
  std::vector<hipEvent_t> events_;
  std::vector<int> marked_;  // the phase each recorded event ends
  double phase_ms_[kPhases] = {};
  long phase_n_[kPhases] = {};
  double trial_ms_ = 0.0;
  long timed_ = 0;
};

}  // namespace sp
