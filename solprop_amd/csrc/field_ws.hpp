// The modal field-grid workspace under the exponential Adams march: the history's rows
// resident in HBM (the reference's resident engine, native/src/march/adams_march.cpp,
// and ExponentialHistory, src/solprop/backends/numpy_ref/multistep.py:806-1131), the
// trial captured as one HIP graph, the decisions on the host (adams_march.hpp).
#pragma once

#include <map>
#include <string>
#include <memory>
#include <vector>

#include "adams_march.hpp"
#include "device.hpp"
#include "fft.hpp"
#include "field.hpp"
#include "propagator.hpp"
#include "scalars.hpp"
#include "stats.hpp"

namespace sp {

class FieldWorkspace final : public Propagator, private adams::Engine {
 public:
  // `shared`, when given, is the stream the workspace runs on (a batched march's: the
  // workspace neither creates nor destroys a stream of its own).
  explicit FieldWorkspace(const sp_problem_desc& desc, hipStream_t shared = nullptr);
  ~FieldWorkspace() override;
  FieldWorkspace(const FieldWorkspace&) = delete;
  FieldWorkspace& operator=(const FieldWorkspace&) = delete;

  void load(double z, const double* field) override;
  double probe_trial(double z, double dz) override;
  void read_trial(double* out) override;
  void read_state(double* out) override;
  void export_state(const double* weight, bool optics, double* out) override;
  bool advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                       sp_march_result* r) override;
  sp_counters counters() const override;
  void source(double z, const double* lab, double* out, double* streams,
              double* modal) override;
  long buffer(const char* name, double* out, long count) override;
  std::vector<double> statistics(std::uint64_t flags, const double* windows,
                                 int n_windows, const StatExtras& extras) override;
  void bind_monitor(const double* weights, const double* columns) override;
  void monitor_reading(double* out) override;
  void set_profile(bool on) override { profile_ = on; }
  std::vector<std::pair<const char*, double>> phases() const override;
  bool phase_used(int i) const;

  // ---- batched marching (scan.hpp): the scan runs many workspaces' trials in one set of
  // grouped launches; each workspace plans its own march's next action and concludes it.
  // Why this problem cannot join a batched march (empty when it can).
  std::string batch_refusal() const;
  // Use a batch's record slot (pinned host and device) instead of the workspace's own.
  void batch_bind(AdamsRecord* host, AdamsRecord* device);
  // The grouped kernels' view of this instance, and its transforms as grouped-launch jobs
  // gated by its record's action.
  BatchInstance batch_instance() const;
  std::vector<FftJob> batch_inverse_jobs(const int* gate) const;
  std::vector<FftJob> batch_forward_jobs(const int* gate) const;
  const FftPlan& inverse_plan() const { return inv_plan_; }
  const FftPlan& forward_plan() const { return fwd_plan_; }
  const DeviceProblem& device_problem() const { return P_; }
  // One advance call's course: begin, then per round plan (fills the record; returns the
  // action) and, after the round's launches, conclude; finished() once the march stops.
  void batch_begin(double target, long work_limit, std::vector<double>* rows);
  int batch_plan();
  void batch_conclude(int action);
  bool batch_finished() const { return batch_finished_; }
  bool batch_pending() const { return pending_slot_ >= 0; }
  // The call's outcome: the status (SP_OK or the march's failure) and its message, and
  // the result record; `batch_end` flushes a commit no round carried.
  int batch_end(sp_march_result* r, std::string* message);
  static constexpr int kPhases = 16;
  double trial_us() const override { return timed_ ? time_total_ms_ * 1e3 / timed_ : 0.0; }
  long timed() const override { return timed_; }

 private:
  // adams::Engine
  void start(double z) override;
  std::vector<double> trial(double z, double dz, const adams::Stencil& s) override;
  void accept(double dz, int free, const adams::Stencil& s,
              const std::vector<adams::Entry>& kept) override;

  void allocate(const sp_problem_desc& d);
  void release();
  template <typename T>
  T* alloc(std::size_t n);
  template <typename T>
  T* upload(const T* host, std::size_t n);
  void fill_position(double z, AdamsRecord* into = nullptr);
  // The source at a lab-frame spectrum into `out`, its position's scalars read from `R`
  // (the workspace's record when null).
  // The source evaluation into `out`; without `finish` it stops at the forward transform's
  // streams (the corrector then finishes each bin itself).
  // The source evaluation's response spectrum at B.lab (the finish without the absorber
  // term) into `out`, under `P` (whose projection may be the coarse rule's).
  void enqueue_response_spectrum(const DeviceProblem& P, cplx* out);
  // The quadrature estimate of the last accepted step's source evaluation, or NaN.
  double quadrature_estimate();
  void enqueue_source(const cplx* spectrum, cplx* out, const AdamsRecord* R = nullptr,
                      bool finish = true);
  void enqueue_synthesis();
  void enqueue_projection();
  void enqueue_raman_kernel(const AdamsRecord* R);
  void enqueue_split(const AdamsRecord* R, double dz);
  void enqueue_raman(const cplx* spectrum, const AdamsRecord* R);
  // ---- the low-storage RK4IP replay (method 2): a fixed schedule, every step accepted.
  // One step is one captured graph: the three positions' records up, then the source at
  // z, m, m and z + h, each followed by its register pass, and the passes' masked sums.
  void enqueue_ls_step();
  double ls_trial(double z, double dz);  // 0.0, or the state's non-finite squared norm
  void ls_accept(double dz);
  bool ls_advance(const Bounds& b, int columns, std::vector<double>& rows, sp_march_result* r);
  bool low_storage_ = false;
  std::vector<double> edges_, steps_;
  cplx *ls_acc_ = nullptr, *ls_prev_ = nullptr, *ls_save_ = nullptr;
  AdamsRecord* ls_rec_dev_ = nullptr;   // (3): z, the midpoint, z + h
  AdamsRecord* ls_rec_host_ = nullptr;  // (3) pinned
  double* ls_partials_ = nullptr;       // (3, blocks)
  hipGraphExec_t ls_graph_ = nullptr;
  march::State ls_state_;
  double stiffness_ = 0.0;              // h rho of the last step (NaN when undefined)
  enum Graph { kTrial = 1, kStart = 2, kCommitTrial = 3 };
  void enqueue_trial(bool commit);
  void enqueue_start();
  void enqueue(int key);
  void run(int key);
  void wait();
  void flush();
  void mark(int i);
  void fill_trial(double z, double dz, const adams::Stencil& s);
  std::vector<double> trial_norms();

  DeviceProblem P_{};
  PlasmaTable pt_{};
  FieldBuffers B_{};
  std::vector<void*> allocations_;
  std::size_t device_bytes_ = 0;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;
  FftPlan inv_plan_, fwd_plan_;
  cplx* scratch_ = nullptr;
  AdamsRecord* rec_dev_ = nullptr;
  AdamsRecord* rec_host_ = nullptr;
  AdamsRecord* rec_host_dev_ = nullptr;  // rec_host_ as the device addresses it
  AdamsRecord* own_rec_host_ = nullptr;  // the workspace's pinned record (freed at release)
  double* norms_host_ = nullptr;
  int n_blocks_ = 0;
  std::map<int, hipGraphExec_t> graphs_;
  bool use_graphs_ = true;
  // Host scalar paths.
  std::vector<ScalarPath> phi_paths_, rate_paths_, pref_paths_, raman_paths_;
  bool raman_dynamic_ = false;  // Raman rates vary along z: the kernel is rebuilt per source
  FftPlan raman_plan_;          // T points, Q rows: the doubled-grid drive
  FftPlan raman_kernel_plan_;   // T points, one row: the kernel
  cplx* raman_scratch_ = nullptr;
  cplx* raman_kernel_z_ = nullptr;
  ScalarPath density_, preion_;
  double rtol_ = 1e-6, atol_ = 0.0;
  bool compensated_ = false;
  bool dense_ = false;          // synthesis and projection as GEMMs (more than kMaxModes modes)
  bool in_stage_ = false;       // the in-stage absorber: a stream of its own
  bool split_ = false;          // the split-step absorber: applied at every commit
  bool spatial_ = false;        // the transverse collar's matrix (dense route)
  int n_t_ = 0;                 // coarse samples (the statistics' grid)
  std::unique_ptr<PlaneStatistics> stats_;  // built at the first statistics call
  double* populations_ = nullptr;            // (2 Q), at the first populations call
  double* diagnostics_ = nullptr;            // (6 Q + 6), at the first diagnostics call
  // The embedded coarse rule's projection (C, Q), or null; the quadrature estimate's
  // spectra (C N each) and value, at its first call. The estimate repeats the source
  // evaluation of the last accepted trial (its field stays in B.lab until the next
  // trial) at that trial's end, quad_z_.
  double* projection_coarse_ = nullptr;
  cplx* quad_fine_ = nullptr;
  cplx* quad_coarse_ = nullptr;
  double* quad_out_ = nullptr;
  bool quad_ready_ = false;
  double trial_end_ = 0.0, quad_z_ = 0.0;
  MonitorBuffers monitor_;
  double* spatial_matrix_ = nullptr;  // (C, C) row-major
  bool absorbing() const { return in_stage_ || spatial_; }
  void enqueue_absorbers();
  double pending_dz_ = 0.0;     // the accepted step a pending commit finishes
  // The march, and the host it reports to: rows and the attempt limit of one call.
  struct CallHost final : adams::Host {
    AttemptLog log;
    void record(double end, double dz, double norm, bool acc) override {
      log.record(end, dz, norm, acc);
    }
    void attempt_order(int order) override { log.order = order; }
    bool halt() override { return log.halt(); }
  };
  adams::Policy policy_;
  CallHost host_;
  std::unique_ptr<adams::March> march_;
  std::vector<std::vector<double>> slot_rates_;  // D(z_j) weights by store slot
  std::vector<double> held_rate_end_;
  double z_ = 0.0, dz_init_ = 0.0;
  bool loaded_ = false;
  // A batched advance call's state.
  adams::Request request_;
  double batch_target_ = 0.0;
  bool batch_finished_ = true, batch_done_ = false;
  int batch_status_ = 0;
  std::string batch_message_;
  // Counters and timing.
  long trials_ = 0, evaluations_ = 0, accepted_ = 0, graph_launches_ = 0;
  int pending_slot_ = -1;  // an accepted step's store slot, committed with the next trial
  int width_ = 2;          // the current trial's stencil width
  hipEvent_t done_ = nullptr;
  bool profile_ = false, marking_ = false;
  hipEvent_t marks_[kPhases + 1] = {};
  double phase_ms_[kPhases] = {};
  double time_total_ms_ = 0.0;
  long timed_ = 0;
};

}  // namespace sp
