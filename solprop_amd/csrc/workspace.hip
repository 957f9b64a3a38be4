// The device workspace (see workspace.hpp). Host code: uploads, the per-trial scalar
// record, the stage sequence, graph capture and the march's engine.
#include "workspace.hpp"
#include "sync.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "kernels.hpp"

namespace sp {

// -------------------------------------------------------------------- set-up

template <typename T>
T* Workspace::device_alloc(std::size_t n) {
  T* p = nullptr;
  if (n == 0) n = 1;
  SP_HIP(hipMalloc(&p, n * sizeof(T)));
  allocations_.push_back(p);
  device_bytes_ += n * sizeof(T);
  return p;
}

template <typename T>
T* Workspace::device_copy(const T* host, std::size_t n) {
  T* p = device_alloc<T>(n);
  if (host && n) SP_HIP(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

Workspace::Workspace(const sp_problem_desc& d) {
  // Refusals first, before anything is allocated.
  SP_CHECK(d.kind == 0, "the Runge-Kutta workspace runs the envelope route");
  SP_CHECK(d.method == 0, "the envelope route marches with the embedded Runge-Kutta pair");
  SP_CHECK(d.n_to >= 16 && (d.n_to & (d.n_to - 1)) == 0, "n_to must be a power of two >= 16");
  SP_CHECK(d.n_t >= 4 && d.n_to % d.n_t == 0, "n_to must be a multiple of n_t");
  SP_CHECK(d.n_omega == d.n_t, "an envelope grid has n_omega = n_t bins");
  SP_CHECK(d.n_modes >= 1 && d.n_nodes == 1, "the envelope route runs one node");
  SP_CHECK(d.n_modes == 1, "the envelope route runs one mode");
  SP_CHECK(d.n_phi >= 1 && d.n_phi <= kMaxTerms, "1..4 phi terms");
  SP_CHECK(d.n_prefactor >= 1 && d.n_prefactor <= kMaxTerms, "1..4 prefactor terms");
  SP_CHECK(d.n_raman >= 0 && d.n_raman <= kMaxTerms, "0..4 Raman rows");
  SP_CHECK(d.absorber_scheme >= 0 && d.absorber_scheme <= 2,
           "absorber scheme 0 (none), 1 (in-stage) or 2 (split-step)");
  SP_CHECK(d.absorber_scheme == 0 || d.absorber_rate, "the absorber's rate is missing");
  SP_CHECK(d.absorber_scheme != 2 || (d.n_t >= 16 && (d.n_t & (d.n_t - 1)) == 0),
           "the split-step absorber transforms the coarse grid: n_t a power of two >= 16");
  SP_CHECK(d.n_stages >= 1 && d.n_stages <= kMaxStages, "1..8 Runge-Kutta stages");
  SP_CHECK(!d.has_plasma, "the envelope route carries no field-grid plasma");
  SP_CHECK(d.crop_index && d.active_mask && d.norm_mask && d.omega_window && d.omega &&
               d.time_window && d.phi_terms && d.prefactor_terms && d.synthesis &&
               d.projection && d.tableau_a && d.tableau_b && d.tableau_c && d.tableau_e,
           "a required problem array is missing");
  const int C = d.n_modes, N = d.n_omega, T = d.n_to;
  for (int b = 0; b < N; ++b)
    SP_CHECK(d.crop_index[b] >= 0 && d.crop_index[b] < T, "crop_index outside the fine grid");

  for (int t = 0; t < d.n_phi; ++t) {
    phi_paths_.emplace_back();
    phi_paths_.back().assign(d.phi_paths[t]);
  }
  for (int t = 0; t < d.n_prefactor; ++t) {
    pref_paths_.emplace_back();
    pref_paths_.back().assign(d.prefactor_paths[t]);
  }
  for (int r = 0; r < d.n_raman; ++r) {
    raman_paths_.emplace_back();
    raman_paths_.back().assign(d.raman_rates[r]);
    raman_dynamic_ = raman_dynamic_ || !raman_paths_.back().constant();
  }
  density_.assign(d.density);
  preion_.assign(d.preionisation);
  tab_c_.assign(d.tableau_c, d.tableau_c + d.n_stages);
  in_stage_ = d.absorber_scheme == 1;
  split_ = d.absorber_scheme == 2;
  // The split collar changes the committed state after the trial: its last stage is no
  // longer the new state's first, and every stage is evaluated
  // (src/solprop/backends/numpy_ref/engines.py:99-178).
  fsal_ = d.first_same_as_last != 0 && !split_;

  // Controller and landings (the run's policy, src/solprop/backends/records.py:227).
  auto& c = policy_.controller;
  c.rtol = d.rtol;
  c.atol = d.atol;
  c.safety = d.safety;
  c.max_factor = d.max_factor;
  c.min_factor = d.min_factor;
  c.beta1 = d.beta1;
  c.beta2 = d.beta2;
  c.lag = d.lag;
  c.reject_free = d.reject_free != 0;
  c.estimator_order = d.estimator_order;
  policy_.dz_min = d.dz_min;
  policy_.dz_max = d.dz_max;
  policy_.auto_init = d.auto_init != 0;
  policy_.stretch = d.stretch;
  policy_.z_end = d.z_end;
  policy_.breakpoints = {0.0, d.z_end};
  dz_init_ = d.dz_init > 0.0 ? d.dz_init : 1e-4 * d.z_end;
  state_.max_norms = static_cast<std::size_t>(c.lag + 2);

  SP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  const char* env = std::getenv("SOLPROP_MI355X_GRAPH");
  use_graphs_ = !(env && env[0] == '0');

  P_.kind = d.kind;
  P_.C = C;
  P_.N = N;
  P_.T = T;
  P_.NF = T;
  P_.n_phi = d.n_phi;
  P_.n_pref = d.n_prefactor;
  P_.Q = d.n_nodes;
  P_.n_raman = d.n_raman;
  P_.has_kerr = d.has_kerr;
  P_.kerr_real = d.kerr_real;
  P_.has_plasma = 0;
  P_.noise_modified = d.noise_model == 2 && d.noise != nullptr;
  P_.kerr_coefficient = d.kerr_coefficient;
  P_.kerr_self = d.kerr_self;
  P_.kerr_conjugate = d.kerr_conjugate;
  P_.raman_fraction = d.raman_fraction;
  P_.raman_epsilon0 = d.raman_epsilon0;
  P_.dt_fine = d.dt_fine;
  P_.inv_T = 1.0 / T;
  P_.inv_2T = 1.0 / (2.0 * T);
  P_.n_stages = d.n_stages;

  std::vector<int> f2c(T, -1), crop(N);
  for (int b = 0; b < N; ++b) {
    crop[b] = static_cast<int>(d.crop_index[b]);
    f2c[crop[b]] = b;
  }
  P_.fine_to_coarse = device_copy(f2c.data(), f2c.size());
  P_.crop = device_copy(crop.data(), crop.size());
  P_.active = device_copy(d.active_mask, N);
  P_.normmask = device_copy(d.norm_mask, N);
  P_.omega_window = device_copy(d.omega_window, N);
  P_.omega = device_copy(d.omega, N);
  P_.t_window = device_copy(d.time_window, T);
  P_.t_deriv = d.time_window_derivative ? device_copy(d.time_window_derivative, T)
                                        : device_alloc<double>(T);
  const std::size_t CN = static_cast<std::size_t>(C) * N;
  P_.phi = device_copy(reinterpret_cast<const cplx*>(d.phi_terms), d.n_phi * CN);
  P_.pref = device_copy(reinterpret_cast<const cplx*>(d.prefactor_terms), d.n_prefactor * CN);
  P_.noise = P_.noise_modified ? device_copy(reinterpret_cast<const cplx*>(d.noise), CN)
                               : nullptr;
  P_.synthesis = device_copy(d.synthesis, static_cast<std::size_t>(d.n_nodes) * C);
  P_.projection = device_copy(d.projection, static_cast<std::size_t>(d.n_nodes) * C);
  P_.raman_rows = d.n_raman ? device_copy(d.raman_undamped, d.n_raman * 2 * std::size_t(T))
                            : nullptr;
  P_.absorber_rate =
      d.absorber_scheme ? device_copy(d.absorber_rate, static_cast<std::size_t>(in_stage_ ? T : N))
                        : nullptr;
  P_.absorber_stream = in_stage_ ? C : -1;
  P_.tab_a = device_copy(d.tableau_a, static_cast<std::size_t>(d.n_stages) * d.n_stages);
  P_.tab_b = device_copy(d.tableau_b, d.n_stages);
  P_.tab_e = device_copy(d.tableau_e, d.n_stages);

  y_ = device_alloc<cplx>(CN);
  adv_ = device_alloc<cplx>(CN);
  recycled_ = device_alloc<cplx>(CN);
  einv_ = device_alloc<cplx>(CN);
  k_ = device_alloc<cplx>(CN * d.n_stages);
  const std::size_t rows = in_stage_ ? 2 * C : C;
  fine_ = device_alloc<cplx>(rows * T);
  scratch_ = device_alloc<cplx>(rows * T);
  fine_plan_.init(T, C);
  if (in_stage_) forward_plan_.init(T, rows);
  if (split_) coarse_plan_.init(N, C);
  if (d.n_raman) {
    drive_ = device_alloc<cplx>(T);
    H_ = device_alloc<cplx>(T + 1);
    const auto roots = unit_roots(2 * static_cast<std::size_t>(T));
    roots2T_ = device_copy(roots.data(), roots.size());
    raman_plan_.init(T, 1);
  }
  n_partial_blocks_ = static_cast<int>(accept_blocks(P_));
  partials_ = device_alloc<double>(3 * static_cast<std::size_t>(n_partial_blocks_));
  rec_dev_ = device_alloc<TrialRecord>(1);
  SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&rec_host_), sizeof(TrialRecord)));
  std::memset(rec_host_, 0, sizeof(TrialRecord));
  SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&partials_host_),
                       3 * sizeof(double) * n_partial_blocks_));
  SP_HIP(hipEventCreate(&ev0_));
  SP_HIP(hipEventCreate(&ev1_));
}

Workspace::~Workspace() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  for (void* p : allocations_) (void)hipFree(p);
  if (rec_host_) (void)hipHostFree(rec_host_);
  if (partials_host_) (void)hipHostFree(partials_host_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// --------------------------------------------------------------- trial sequence

void Workspace::fill_record(double z, double dz) {
  TrialRecord& r = *rec_host_;
  r.dz = dz;
  auto fill = [&](StageRow& row, double zi) {
    for (int t = 0; t < P_.n_phi; ++t) {
      row.phi_w[t] = path_increment(phi_paths_[t], z, zi);
      row.noise_w[t] = phi_paths_[t](zi);
    }
    for (int t = 0; t < P_.n_pref; ++t) row.pref_w[t] = pref_paths_[t](zi);
    for (int q = 0; q < P_.n_raman; ++q) row.raman_rate[q] = raman_paths_[q](zi);
    row.density = density_(zi);
    row.preionisation = preion_(zi);
    row.taper = 1.0;
  };
  for (int i = 0; i < P_.n_stages; ++i) fill(r.stage[i], stage_position(z, tab_c_[i], dz));
  fill(r.whole, stage_position(z, 1.0, dz));
}

void Workspace::enqueue_raman_kernel(int stage) {
  launch_raman_kernel_fill(P_, rec_dev_, stage, drive_, stream_);
  raman_plan_.exec(drive_, scratch_, -1, stream_);
  launch_raman_kernel_post(drive_, H_, roots2T_, P_.T, P_.dt_fine, stream_);
}

void Workspace::enqueue_stage(int stage, bool finish_previous) {
  // The previous stage's finish rides in this stage's entry, and the last stage's in the
  // step's acceptance.
  if (finish_previous)
    launch_rk_finish_entry(P_, rec_dev_, y_, k_, fine_, einv_, stage, stream_);
  else
    launch_rk_entry_envelope(P_, rec_dev_, y_, k_, fine_, einv_, stage, stream_);
  fine_plan_.exec(fine_, scratch_, +1, stream_);
  if (P_.n_raman) {
    if (raman_dynamic_) enqueue_raman_kernel(stage);
    launch_env_raman_drive(P_, rec_dev_, fine_, drive_, stage, stream_);
    raman_plan_.exec(drive_, scratch_, -1, stream_);
    launch_raman_product(drive_, H_, roots2T_, P_.T, 1, stream_);
    raman_plan_.exec(drive_, scratch_, +1, stream_);
  }
  launch_env_response(P_, rec_dev_, fine_, drive_, stage, stream_);
  (in_stage_ ? forward_plan_ : fine_plan_).exec(fine_, scratch_, -1, stream_);
}

void Workspace::enqueue_commit(const TrialRecord* R, double dz) {
  launch_commit_masked(P_, adv_, y_, stream_);
  if (!split_) return;
  // fine_ is free between trials: it holds the coarse grid's time samples here.
  launch_env_split_pack(P_, y_, fine_, stream_);
  coarse_plan_.exec(fine_, scratch_, +1, stream_);
  // What the split takes and finds, from the masked state before it (a bound monitor).
  if (monitor_.bound())
    launch_monitor_sums(reinterpret_cast<const double*>(fine_), kSamplesComplex, P_.C, P_.N,
                        monitor_.columns(), monitor_.weights(), dz,
                        R ? &R->accepted_dz : nullptr, kMonitorSplit, monitor_.partials(),
                        monitor_.split(), stream_);
  launch_env_split_collar(P_, R, dz, fine_, stream_);
  coarse_plan_.exec(fine_, scratch_, -1, stream_);
  launch_commit_masked(P_, fine_, y_, stream_);
}

void Workspace::enqueue_trial(bool commit, Stage0 mode) {
  const long CN = static_cast<long>(P_.C) * P_.N;
  if (commit) enqueue_commit(rec_dev_, 0.0);
  if (mode == kCopyRecycled) launch_copy_points(recycled_, k_, CN, stream_);
  const int first = mode == kEvaluate ? 0 : 1;
  for (int i = first; i < P_.n_stages; ++i) enqueue_stage(i, i > first);
  launch_rk_accept(P_, rec_dev_, y_, k_, adv_, recycled_, partials_, fine_, einv_, stream_);
}

double Workspace::finish_trial() {
  SP_HIP(hipMemcpyAsync(partials_host_, partials_, 3 * sizeof(double) * n_partial_blocks_,
                        hipMemcpyDeviceToHost, stream_));
  if (profile_) SP_HIP(hipEventRecord(ev1_, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
  if (profile_) {
    float ms = 0.0f;
    SP_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    time_total_ms_ += ms;
    timed_ += 1;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int nb = n_partial_blocks_;
  for (int i = 0; i < nb; ++i) s0 = s0 + partials_host_[i];
  for (int i = 0; i < nb; ++i) s1 = s1 + partials_host_[nb + i];
  for (int i = 0; i < nb; ++i) s2 = s2 + partials_host_[2 * nb + i];
  const double norm_y = std::sqrt(s0), norm_new = std::sqrt(s1);
  const double scale =
      policy_.controller.rtol * (norm_new > norm_y ? norm_new : norm_y) + policy_.controller.atol;
  return std::sqrt(s2) / scale;
}

double Workspace::trial(double z, double dz, bool full) {
  if (!loaded_) throw Error(kState, "load a state before taking a trial");
  SP_CHECK(std::isfinite(dz) && dz > 0.0, "dz must be finite and positive");
  fill_record(z, dz);
  const bool commit = commit_pending_;
  rec_host_->accepted_dz = commit ? pending_dz_ : 0.0;
  Stage0 mode = kEvaluate;
  if (!full && fsal_) {
    if (commit && k0_valid_)
      mode = kCopyRecycled;
    else if (!commit && k0_valid_ && k0_z_ == z)
      mode = kKeep;
  }
  SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(TrialRecord), hipMemcpyHostToDevice,
                        stream_));
  if (profile_) SP_HIP(hipEventRecord(ev0_, stream_));
  const int key = (commit ? 4 : 0) + static_cast<int>(mode);
  if (use_graphs_) {
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
      const hipGraphExec_t exec =
          capture_graph(stream_, [&] { enqueue_trial(commit, mode); });
      it = graphs_.emplace(key, exec).first;
    }
    SP_HIP(hipGraphLaunch(it->second, stream_));
    graph_launches_ += 1;
  } else {
    enqueue_trial(commit, mode);
  }
  commit_pending_ = false;
  const double norm = finish_trial();
  trials_ += 1;
  evaluations_ += P_.n_stages - (mode == kEvaluate ? 0 : 1);
  k0_valid_ = fsal_;
  k0_z_ = z;
  trial_held_ = true;
  return norm;
}

void Workspace::accept(double dz) {
  if (!trial_held_) throw Error(kState, "accept without a held trial");
  commit_pending_ = true;
  pending_dz_ = dz;
  trial_held_ = false;
  accepted_ += 1;
}

void Workspace::ensure_committed() {
  if (!commit_pending_) return;
  enqueue_commit(nullptr, pending_dz_);
  SP_HIP(hipStreamSynchronize(stream_));
  commit_pending_ = false;
  // k[0] is still the committed state's source only if it was recycled into place.
  if (fsal_) {
    launch_copy_points(recycled_, k_, static_cast<long>(P_.C) * P_.N, stream_);
    SP_HIP(hipStreamSynchronize(stream_));
    k0_valid_ = true;
    k0_z_ = state_.z;
  }
}

void Workspace::load(double z, const double* field) {
  SP_CHECK(field != nullptr, "load needs a field");
  const std::size_t CN = static_cast<std::size_t>(P_.C) * P_.N;
  monitor_.reset(stream_);
  SP_HIP(hipMemcpyAsync(y_, field, CN * sizeof(cplx), hipMemcpyHostToDevice, stream_));
  if (P_.n_raman && !raman_dynamic_) {
    fill_record(z, 1.0);
    SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(TrialRecord), hipMemcpyHostToDevice,
                          stream_));
    enqueue_raman_kernel(0);
  }
  SP_HIP(hipStreamSynchronize(stream_));
  loaded_ = true;
  commit_pending_ = false;
  k0_valid_ = false;
  trial_held_ = false;
  state_ = march::State{};
  state_.z = z;
  state_.proposal = dz_init_;
  state_.max_norms = static_cast<std::size_t>(policy_.controller.lag + 2);
}

void Workspace::read_state(double* out) {
  if (!loaded_) throw Error(kState, "no state is loaded");
  ensure_committed();
  SP_HIP(hipMemcpy(out, y_, sizeof(cplx) * P_.C * P_.N, hipMemcpyDeviceToHost));
}

void Workspace::read_trial(double* out) {
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, adv_, sizeof(cplx) * P_.C * P_.N, hipMemcpyDeviceToHost));
}

void Workspace::export_state(const double* weight, bool optics, double* out) {
  read_state(out);
  const long n = static_cast<long>(P_.C) * P_.N;
  for (long i = 0; i < n; ++i) {
    double re = out[2 * i], im = out[2 * i + 1];
    if (weight) {
      const double s = std::sqrt(weight[i]);
      re = re * s;
      im = im * s;
    }
    if (optics) {
      im = -im;
      if ((i % P_.N) & 1) {
        re = -re;
        im = -im;
      }
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
  }
}

// ------------------------------------------------------------------- the march

namespace {
struct WorkspaceEngine final : march::Engine {
  Workspace& ws;
  explicit WorkspaceEngine(Workspace& w) : ws(w) {}
  double trial(double z, double dz) override { return ws.trial(z, dz, false); }
  void accept(double dz) override { ws.accept(dz); }
};

struct SessionHost final : march::Host {
  AttemptLog& log;
  explicit SessionHost(AttemptLog& l) : log(l) {}
  void record(double end, double dz, double norm, bool acc) override {
    log.record(end, dz, norm, acc);
  }
  bool halt() override { return log.halt(); }
};
}  // namespace

bool Workspace::advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                                sp_march_result* r) {
  if (!loaded_) throw Error(kState, "load a state before advancing");
  WorkspaceEngine engine(*this);
  AttemptLog log;
  log.begin(&rows, columns, b);
  // A Runge-Kutta pair's attempts are all of its order: the estimator's plus one.
  log.order = policy_.controller.estimator_order + 1;
  SessionHost host(log);
  bool done = false;
  try {
    done = march::march_to(policy_, engine, host, state_, b.target) &&
           march::reached(state_.z, b.target);
  } catch (...) {
    r->z = state_.z;
    r->proposal = state_.proposal;
    r->attempts = log.attempts;
    r->accepted = log.accepted;
    r->evaluations = evaluations_;
    throw;
  }
  r->z = state_.z;
  r->proposal = state_.proposal;
  r->done = done ? 1 : 0;
  r->attempts = log.attempts;
  r->accepted = log.accepted;
  r->evaluations = evaluations_;
  return done;
}

void Workspace::bind_monitor(const double* weights, const double* columns) {
  SP_HIP(hipStreamSynchronize(stream_));
  // The captured trials hold the monitor's kernels or not: capture them again.
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
  if (!weights) {
    monitor_.unbind();
    return;
  }
  monitor_.bind(weights, columns, P_.C, P_.N, stream_);
}

void Workspace::monitor_reading(double* out) {
  SP_CHECK(monitor_.bound(), "no boundary monitor is bound");
  if (!loaded_) throw Error(kState, "no state is loaded");
  ensure_committed();
  if (!stats_)
    stats_ = std::make_unique<PlaneStatistics>(P_.C, P_.N, P_.N, false,
                                               P_.dt_fine * P_.T / P_.N, P_.omega);
  stats_->reading(y_, monitor_.columns(), monitor_.weights(), monitor_.partials(),
                  monitor_.reading(), stream_);
  monitor_.drain(out, stream_);
}

std::vector<double> Workspace::statistics(std::uint64_t flags, const double* windows,
                                          int n_windows, const StatExtras& extras) {
  if (!loaded_) throw Error(kState, "no state is loaded");
  ensure_committed();
  // The coarse grid: N samples spaced by the fine spacing times T / N.
  if (!stats_)
    stats_ = std::make_unique<PlaneStatistics>(P_.C, P_.N, P_.N, false,
                                               P_.dt_fine * P_.T / P_.N, P_.omega);
  SP_CHECK(!(flags & kStatColumn),
           "the plasma statistics read the field-grid plasma column, and the envelope route "
           "carries none");
  std::vector<double> trace;
  double intensity = 0.0;
  std::vector<double> out = stats_->run(y_, flags & ~kStatQuadrature, windows, n_windows,
                                        stream_, &trace, &extras, &intensity);
  if (flags & kStatPeakIntensity) out.push_back(intensity);
  // One mode on one node: no radial rule, so no embedded pair to compare against.
  if (flags & kStatQuadrature) out.push_back(std::nan(""));
  out.insert(out.end(), trace.begin(), trace.end());
  return out;
}

sp_counters Workspace::counters() const {
  sp_counters c{};
  c.trials = trials_;
  c.evaluations = evaluations_;
  c.accepted = accepted_;
  c.graph_launches = graph_launches_;
  c.device_bytes = static_cast<long>(device_bytes_ + fine_plan_.bytes() + raman_plan_.bytes() +
                                     forward_plan_.bytes() + coarse_plan_.bytes() +
                                     (stats_ ? stats_->bytes() : 0));
  return c;
}

}  // namespace sp
