// The device workspace: the problem's arrays resident in HBM, the stage buffers, the
// transform plans and the captured trial graphs (the reference's GeneralWorkspace,
// native/src/general/general_workspace.hpp, and native::Workspace,
// native/include/solprop_native/workspace/workspace.hpp:139).
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "device.hpp"
#include "fft.hpp"
#include "march.hpp"
#include "propagator.hpp"
#include "scalars.hpp"
#include "stats.hpp"

namespace sp {

class Workspace final : public Propagator {
 public:
  explicit Workspace(const sp_problem_desc& desc);
  ~Workspace() override;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;

  void load(double z, const double* field) override;
  // One trial from the committed state at z; every stage evaluated when `full`.
  double trial(double z, double dz, bool full);
  double probe_trial(double z, double dz) override { return trial(z, dz, true); }
  void accept(double dz);
  void read_state(double* out) override;
  void read_trial(double* out) override;
  void export_state(const double* weight, bool optics, double* out) override;
  bool advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                       sp_march_result* r) override;

  std::vector<double> statistics(std::uint64_t flags, const double* windows,
                                 int n_windows, const StatExtras& extras) override;
  void bind_monitor(const double* weights, const double* columns) override;
  void monitor_reading(double* out) override;
  void set_profile(bool on) override { profile_ = on; }
  double trial_us() const override { return timed_ ? time_total_ms_ * 1e3 / timed_ : 0.0; }
  long timed() const override { return timed_; }
  sp_counters counters() const override;

 private:
  enum Stage0 { kEvaluate = 0, kCopyRecycled = 1, kKeep = 2 };
  template <typename T>
  T* device_copy(const T* host, std::size_t n);
  template <typename T>
  T* device_alloc(std::size_t n);
  void fill_record(double z, double dz);
  void enqueue_trial(bool commit, Stage0 mode);
  // One stage; with `finish_previous` its entry also finishes the stage before it.
  void enqueue_stage(int stage, bool finish_previous);
  void enqueue_raman_kernel(int stage);
  void ensure_committed();
  // The commit of the held trial: the band mask and, under the split-step absorber, the
  // collar over the accepted step (R's, inside a graph; else `dz`) and the mask again.
  void enqueue_commit(const TrialRecord* R, double dz);
  double finish_trial();

  DeviceProblem P_{};
  std::vector<void*> allocations_;
  std::size_t device_bytes_ = 0;
  hipStream_t stream_ = nullptr;
  FftPlan fine_plan_, raman_plan_;
  // The forward transform over the polarisation's rows and the in-stage absorber's, and
  // the coarse grid's transform for the split-step absorber.
  FftPlan forward_plan_, coarse_plan_;
  bool in_stage_ = false, split_ = false;
  std::unique_ptr<PlaneStatistics> stats_;  // built at the first statistics call
  MonitorBuffers monitor_;
  double pending_dz_ = 0.0;
  // Host scalar paths.
  std::vector<ScalarPath> phi_paths_, pref_paths_, raman_paths_;
  ScalarPath density_, preion_;
  std::vector<double> tab_c_;
  bool fsal_ = false, raman_dynamic_ = false;
  // Device buffers.
  cplx *y_ = nullptr, *adv_ = nullptr, *recycled_ = nullptr, *einv_ = nullptr, *k_ = nullptr;
  cplx *fine_ = nullptr, *scratch_ = nullptr, *drive_ = nullptr, *H_ = nullptr;
  cplx* roots2T_ = nullptr;
  double* partials_ = nullptr;
  TrialRecord* rec_dev_ = nullptr;
  TrialRecord* rec_host_ = nullptr;
  double* partials_host_ = nullptr;
  int n_partial_blocks_ = 0;
  // Graphs keyed by (commit, stage-0 mode).
  std::map<int, hipGraphExec_t> graphs_;
  bool use_graphs_ = true;
  // March state.
  march::Policy policy_;
  march::State state_;
  double dz_init_ = 0.0;
  bool loaded_ = false, commit_pending_ = false, k0_valid_ = false, trial_held_ = false;
  double k0_z_ = 0.0;
  // Counters and timing.
  long trials_ = 0, evaluations_ = 0, accepted_ = 0, graph_launches_ = 0;
  bool profile_ = false;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  double time_total_ms_ = 0.0;
  long timed_ = 0;
};

}  // namespace sp
