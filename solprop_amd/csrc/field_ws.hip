// The modal field-grid workspace under the exponential Adams march (see field_ws.hpp).
#include "field_ws.hpp"
#include "gemm.hpp"
#include "kernels.hpp"
#include "sync.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>

namespace sp {

template <typename T>
T* FieldWorkspace::alloc(std::size_t n) {
  T* p = nullptr;
  if (n == 0) n = 1;
  SP_HIP(hipMalloc(&p, n * sizeof(T)));
  allocations_.push_back(p);
  device_bytes_ += n * sizeof(T);
  return p;
}

template <typename T>
T* FieldWorkspace::upload(const T* host, std::size_t n) {
  T* p = alloc<T>(n);
  if (host && n) SP_HIP(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

FieldWorkspace::FieldWorkspace(const sp_problem_desc& d, hipStream_t shared) {
  if (shared) {
    stream_ = shared;
    own_stream_ = false;
  }
  // Refusals first, before anything is allocated.
  SP_CHECK(d.kind == 1 && (d.method == 1 || d.method == 2),
           "the field workspace runs the real field grid under the exponential Adams march "
           "or the low-storage RK4IP replay");
  if (d.method == 2) {
    SP_CHECK(d.n_schedule >= 1 && d.schedule_edges, "the low-storage replay needs a schedule");
    for (int i = 0; i < d.n_schedule; ++i)
      SP_CHECK(d.schedule_edges[i + 1] > d.schedule_edges[i],
               "the schedule's edges must increase");
    SP_CHECK(d.absorber_scheme == 0 && !d.spatial_absorber,
             "the low-storage replay is built without absorbers");
  }
  const int C = d.n_modes, N = d.n_omega, T = d.n_to, Q = d.n_nodes;
  SP_CHECK(T >= 256 && (T & (T - 1)) == 0, "n_to must be a power of two >= 256");
  SP_CHECK(d.n_t >= 4 && T % d.n_t == 0, "n_to must be a multiple of n_t");
  SP_CHECK(N == d.n_t / 2 + 1, "a field grid has n_omega = n_t/2 + 1 bins");
  SP_CHECK(C >= 1, "at least one mode");
  SP_CHECK(Q >= 1, "at least one quadrature node");
  SP_CHECK(d.n_phi >= 1 && d.n_phi <= kMaxTerms, "1..4 phi terms");
  SP_CHECK(d.n_prefactor >= 1 && d.n_prefactor <= kMaxTerms, "1..4 prefactor terms");
  SP_CHECK(d.n_raman >= 0 && d.n_raman <= kMaxTerms, "0..4 Raman rows");
  SP_CHECK(d.n_raman == 0 || d.raman_undamped, "the Raman rows are missing");
  SP_CHECK(d.absorber_scheme >= 0 && d.absorber_scheme <= 2,
           "absorber scheme 0 (none), 1 (in-stage) or 2 (split-step)");
  SP_CHECK(d.absorber_scheme == 0 || d.absorber_rate, "the absorber's rate is missing");
  SP_CHECK(d.absorber_scheme != 2 || d.n_t == d.n_to,
           "the split-step absorber needs n_to == n_t (its collar acts on the coarse grid)");
  SP_CHECK(!d.spatial_absorber || d.n_modes > kMaxModes,
           "the transverse collar's matrix runs on the dense route (more than 8 modes)");
  SP_CHECK(!d.has_plasma || (d.n_taps >= 1 && d.n_taps <= kMaxBandHalf + 1 && d.taps),
           "the running-integral rule needs 1 (the trapezoid) to 9 (the band rule) taps");
  SP_CHECK(d.adams_order >= 1 && d.adams_order + 1 < kMaxAdamsWidth, "Adams order 1..6");
  SP_CHECK(d.noise_model != 2, "the modified noise model is not built on the field route");
  SP_CHECK(d.crop_index && d.active_mask && d.norm_mask && d.omega_window && d.omega &&
               d.time_window && d.time_window_derivative && d.phi_terms &&
               d.prefactor_terms && d.synthesis && d.projection,
           "a required problem array is missing");
  for (int b = 0; b < N; ++b)
    SP_CHECK(d.crop_index[b] == b, "a field grid's coarse bins lead the fine spectrum");
  SP_CHECK(!d.has_plasma || (d.rate_intervals >= 1 && d.rate_coefficients),
           "the rate table is missing");

  for (int t = 0; t < d.n_phi; ++t) {
    phi_paths_.emplace_back();
    phi_paths_.back().assign(d.phi_paths[t]);
    rate_paths_.emplace_back();
    rate_paths_.back().assign(d.phi_rate_paths[t]);
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
  rtol_ = d.rtol;
  atol_ = d.atol;
  compensated_ = d.n_phi > 1;
  dense_ = C > kMaxModes;

  policy_.order = d.adams_order;
  policy_.variable_order = d.adams_variable_order != 0;
  policy_.safety = d.adams_safety;
  policy_.max_growth = d.adams_max_growth;
  policy_.min_factor = d.adams_min_factor;
  policy_.equal_approach = d.adams_equal_approach != 0;
  policy_.dz_min = d.dz_min;
  policy_.dz_max = d.dz_max;
  policy_.stretch = d.stretch;
  policy_.z_end = d.z_end;
  policy_.breakpoints = {0.0, d.z_end};
  dz_init_ = d.dz_init > 0.0 ? d.dz_init : 1e-4 * d.z_end;

  low_storage_ = d.method == 2;
  if (low_storage_) {
    edges_.assign(d.schedule_edges, d.schedule_edges + d.n_schedule + 1);
    for (int i = 0; i < d.n_schedule; ++i) steps_.push_back(edges_[i + 1] - edges_[i]);
  }
  const char* env = std::getenv("SOLPROP_MI355X_GRAPH");
  use_graphs_ = !(env && env[0] == '0');
  try {
    allocate(d);
  } catch (...) {
    release();
    throw;
  }
}

void FieldWorkspace::allocate(const sp_problem_desc& d) {
  const int C = d.n_modes, N = d.n_omega, T = d.n_to, Q = d.n_nodes;
  if (own_stream_) SP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));

  P_.kind = 1;
  P_.C = C;
  P_.N = N;
  P_.T = T;
  P_.NF = T / 2 + 1;
  P_.n_phi = d.n_phi;
  P_.n_pref = d.n_prefactor;
  P_.Q = Q;
  P_.has_kerr = d.has_kerr;
  P_.kerr_real = 1;
  P_.has_plasma = d.has_plasma;
  P_.kerr_coefficient = d.kerr_coefficient;
  P_.plasma_cmr = d.charge_mass_ratio;
  P_.plasma_ip = d.ionisation_potential;
  P_.dt_fine = d.dt_fine;
  P_.inv_T = 1.0 / T;
  P_.inv_2T = 1.0 / (2.0 * T);
  P_.n_raman = d.n_raman;
  P_.raman_fraction = d.raman_fraction;
  P_.raman_epsilon0 = d.raman_epsilon0;
  P_.raman_rows = d.n_raman ? upload(d.raman_undamped, d.n_raman * 2UL * T) : nullptr;
  P_.raman_analytic = d.n_raman && !d.raman_thg ? 1 : 0;
  in_stage_ = d.absorber_scheme == 1;
  split_ = d.absorber_scheme == 2;
  spatial_ = d.spatial_absorber != nullptr;
  P_.absorber_rate = d.absorber_scheme ? upload(d.absorber_rate, static_cast<std::size_t>(T))
                                       : nullptr;
  spatial_matrix_ =
      spatial_ ? upload(d.spatial_absorber, static_cast<std::size_t>(C) * C) : nullptr;
  // The in-stage absorbers' stream (the temporal term, the transverse collar's product,
  // or their sum) follows the polarisation and, with plasma, the currents.
  P_.absorber_stream = absorbing() ? (d.has_plasma ? 3 : 1) : -1;
  std::vector<int> crop(N);
  for (int b = 0; b < N; ++b) crop[b] = b;
  P_.crop = upload(crop.data(), crop.size());
  P_.active = upload(d.active_mask, N);
  P_.normmask = upload(d.norm_mask, N);
  P_.omega_window = upload(d.omega_window, N);
  P_.omega = upload(d.omega, N);
  n_t_ = d.n_t;
  P_.t_window = upload(d.time_window, T);
  P_.t_deriv = upload(d.time_window_derivative, T);
  const std::size_t CN = static_cast<std::size_t>(C) * N;
  P_.phi = upload(reinterpret_cast<const cplx*>(d.phi_terms), d.n_phi * CN);
  P_.pref = upload(reinterpret_cast<const cplx*>(d.prefactor_terms), d.n_prefactor * CN);
  P_.synthesis = upload(d.synthesis, static_cast<std::size_t>(Q) * C);
  P_.projection = upload(d.projection, static_cast<std::size_t>(Q) * C);
  if (d.projection_coarse)
    projection_coarse_ = upload(d.projection_coarse, static_cast<std::size_t>(Q) * C);
  if (d.has_plasma) {
    pt_.log_field_min = d.log_field_min;
    pt_.d_log_field = d.d_log_field;
    pt_.field_min = d.field_min;
    pt_.field_max = d.field_max;
    pt_.rate_max = d.rate_max;
    pt_.charge_mass_ratio = d.charge_mass_ratio;
    pt_.ionisation_potential = d.ionisation_potential;
    pt_.n_intervals = d.rate_intervals;
    pt_.n_taps = d.n_taps;
    for (int k = 0; k < d.n_taps; ++k) pt_.taps[k] = d.taps[k];
    B_.rate_c = upload(d.rate_coefficients, 4 * static_cast<std::size_t>(d.rate_intervals));
  }

  // Store slots, and the widest stencil's carried rows.
  const int S = policy_.order + 2;
  const int W = S;
  B_.y = alloc<cplx>(CN);
  B_.store = alloc<cplx>(S * CN);
  B_.states = compensated_ ? alloc<cplx>(S * CN) : nullptr;
  B_.carried = compensated_ ? alloc<cplx>(W * CN) : nullptr;
  B_.lab = alloc<cplx>(CN);
  B_.source = alloc<cplx>(CN);
  B_.advanced = alloc<cplx>(CN);
  B_.modal_t = alloc<cplx>(static_cast<std::size_t>(C) * (T / 2));
  const std::size_t n_rows = (absorbing() && d.has_plasma ? 4UL : 3UL) * C;
  B_.streams = alloc<cplx>(n_rows * (T / 2));
  // Streams a run never writes (the currents without plasma on the dense route) read as
  // zeros.
  SP_HIP(hipMemset(B_.streams, 0, n_rows * (T / 2) * sizeof(cplx)));
  const std::size_t tiles = static_cast<std::size_t>(scan_tiles(T));
  B_.rate = d.has_plasma ? alloc<double>(static_cast<std::size_t>(Q) * T) : nullptr;
  B_.frac = d.has_plasma ? alloc<double>(static_cast<std::size_t>(Q) * T) : nullptr;
  B_.free_local = d.has_plasma ? alloc<double>(static_cast<std::size_t>(Q) * T) : nullptr;
  B_.tile_x = alloc<double>(static_cast<std::size_t>(Q) * tiles);
  B_.tile_f = alloc<double>(static_cast<std::size_t>(Q) * tiles);
  // The band passes' within-tile exposure: the fused passes read it under either rule,
  // and so do the integration diagnostics; the fused passes write each node group's
  // projections and projected totals.
  if (d.has_plasma) B_.x_local = alloc<double>(static_cast<std::size_t>(Q) * T);
  if (d.has_plasma && !dense_) {
    const std::size_t groups = static_cast<std::size_t>(node_groups(Q));
    B_.group_streams = alloc<double>(groups * 2 * C * T);
    B_.tile_p = alloc<double>(groups * C * tiles);
    B_.zero_tiles = alloc<unsigned char>(static_cast<std::size_t>(Q) * tiles);
  }
  if (dense_) {
    B_.node_field = alloc<double>(static_cast<std::size_t>(Q) * T);
    B_.node_streams = alloc<double>(static_cast<std::size_t>(Q) * (d.has_plasma ? 2 : 1) * T);
  }
  n_blocks_ = static_cast<int>(correct_blocks(P_));
  // The corrector's block sums stay on the device; their totals land in mapped pinned
  // host memory, which the host reads after the trial's one synchronisation.
  B_.partials = alloc<double>(static_cast<std::size_t>(1 + kMaxScored) * n_blocks_);
  SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&norms_host_), sizeof(double) * (1 + kMaxScored),
                       hipHostMallocMapped));
  SP_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&B_.norms), norms_host_, 0));
  const auto roots = unit_roots(static_cast<std::size_t>(T));
  B_.roots_T = upload(roots.data(), roots.size());
  // Without plasma only the polarisation stream is transformed (the finish reads the
  // current streams on plasma runs alone). A two-pass transform needs scratch for every
  // row it transforms.
  const std::size_t forward_rows = ((d.has_plasma ? 3 : 1) + (absorbing() ? 1 : 0)) * C;
  scratch_ = alloc<cplx>((forward_rows > 3UL * C ? forward_rows : 3UL * C) * (T / 2));
  inv_plan_.init(T / 2, C);
  fwd_plan_.init(T / 2, forward_rows);
  if (d.n_raman) {
    // The doubled-grid drive of every node, its transform scratch, the kernel's packed
    // samples and spectrum, and the roots of 2T for the real-transform packing.
    B_.raman_drive = alloc<cplx>(static_cast<std::size_t>(Q) * T);
    if (P_.raman_analytic) {
      B_.modal_h = alloc<cplx>(static_cast<std::size_t>(C) * (T / 2));
      if (dense_) B_.node_hilbert = alloc<double>(static_cast<std::size_t>(Q) * T);
    }
    raman_scratch_ = alloc<cplx>(static_cast<std::size_t>(Q) * T);
    raman_kernel_z_ = alloc<cplx>(T);
    B_.raman_H = alloc<cplx>(T + 1UL);
    const auto roots2 = unit_roots(2 * static_cast<std::size_t>(T));
    B_.roots_2T = upload(roots2.data(), roots2.size());
    raman_plan_.init(T, Q);
    raman_kernel_plan_.init(T, 1);
  }
  rec_dev_ = alloc<AdamsRecord>(1);
  SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&own_rec_host_), sizeof(AdamsRecord)));
  rec_host_ = own_rec_host_;
  SP_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&rec_host_dev_), rec_host_, 0));
  std::memset(rec_host_, 0, sizeof(AdamsRecord));
  if (low_storage_) {
    const std::size_t CN = static_cast<std::size_t>(C) * N;
    ls_acc_ = alloc<cplx>(CN);
    ls_prev_ = alloc<cplx>(CN);
    ls_save_ = alloc<cplx>(CN);
    ls_rec_dev_ = alloc<AdamsRecord>(3);
    SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&ls_rec_host_), 3 * sizeof(AdamsRecord)));
    std::memset(ls_rec_host_, 0, 3 * sizeof(AdamsRecord));
    ls_partials_ = alloc<double>(ls_partials(P_));
  }
  for (hipEvent_t& e : marks_) SP_HIP(hipEventCreate(&e));
  SP_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
  slot_rates_.assign(S, std::vector<double>(kMaxTerms, 0.0));
  if (d.n_raman && !raman_dynamic_) {
    // Constant rates: the kernel's spectrum once, here.
    fill_position(0.0);
    SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(AdamsRecord), hipMemcpyHostToDevice,
                          stream_));
    enqueue_raman_kernel(rec_dev_);
    SP_HIP(hipStreamSynchronize(stream_));
  }
}

FieldWorkspace::~FieldWorkspace() { release(); }

void FieldWorkspace::release() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  if (ls_graph_) (void)hipGraphExecDestroy(ls_graph_);
  ls_graph_ = nullptr;
  if (ls_rec_host_) (void)hipHostFree(ls_rec_host_);
  ls_rec_host_ = nullptr;
  for (void* p : allocations_) (void)hipFree(p);
  if (own_rec_host_) (void)hipHostFree(own_rec_host_);
  if (norms_host_) (void)hipHostFree(norms_host_);
  for (hipEvent_t& e : marks_) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (done_) (void)hipEventDestroy(done_);
  done_ = nullptr;
  if (stream_ && own_stream_) (void)hipStreamDestroy(stream_);
  graphs_.clear();
  allocations_.clear();
  rec_host_ = nullptr;
  rec_host_dev_ = nullptr;
  own_rec_host_ = nullptr;
  norms_host_ = nullptr;
  stream_ = nullptr;
}

// ------------------------------------------------------------- device sequences

namespace {
// The kernels (or transform plans) of one trial, in launch order.
constexpr const char* kPhaseNames[FieldWorkspace::kPhases] = {
    "adams_predict",      "field_scatter",        "inverse_fft",
    "synthesis_gemm",     "raman",                "plasma_rate",          "plasma_tile_prefix",
    "plasma_current",       "plasma_tile_prefix_current", "node_response",
    "projection_gemm",    "absorber",             "forward_fft",          "field_finish",
    "adams_correct",      "adams_norms"};
}  // namespace

void FieldWorkspace::mark(int i) {
  if (marking_) SP_HIP(hipEventRecord(marks_[i], stream_));
}

// The dense route's GEMMs: node field = S (Q x C) . modal fields (C x T), and the
// streams = W (C x Q) . node streams with the time collars applied.
void FieldWorkspace::enqueue_synthesis() {
  if (!dense_) return;
  launch_dgemm(P_.Q, P_.T, P_.C, P_.synthesis, P_.C,
               reinterpret_cast<const double*>(B_.modal_t), P_.T, B_.node_field, P_.T,
               stream_);
}

void FieldWorkspace::enqueue_projection() {
  if (!dense_) return;
  launch_collar_projection(P_.C, P_.T, P_.Q, P_.projection, B_.node_streams,
                           P_.has_plasma != 0, P_.t_window, P_.t_deriv,
                           reinterpret_cast<double*>(B_.streams), stream_);
}

// The in-stage absorbers' stream: the transverse collar's matrix times the modal time
// fields (a C x C by C x T GEMM on the matrix cores), then the temporal term
// -alpha(t) a_c(t) / 2 added to it (or written, without a collar).
void FieldWorkspace::enqueue_absorbers() {
  double* stream = reinterpret_cast<double*>(B_.streams) +
                   static_cast<std::size_t>(P_.absorber_stream) * P_.C * P_.T;
  if (spatial_)
    launch_dgemm(P_.C, P_.T, P_.C, spatial_matrix_, P_.C,
                 reinterpret_cast<const double*>(B_.modal_t), P_.T, stream, P_.T, stream_);
  if (in_stage_) launch_field_absorber_stream(P_, B_, spatial_, stream_);
}

// The Raman kernel's spectrum at the record's rates: H = rfft(h) dt on the doubled grid.
void FieldWorkspace::enqueue_raman_kernel(const AdamsRecord* R) {
  launch_field_raman_kernel_fill(P_, R, raman_kernel_z_, stream_);
  raman_kernel_plan_.exec(raman_kernel_z_, raman_scratch_, -1, stream_);
  launch_raman_kernel_post(raman_kernel_z_, B_.raman_H, B_.roots_2T, P_.T, P_.dt_fine, stream_);
}

// The Raman convolution of every node's drive: E² on the doubled grid, transformed,
// times the kernel's spectrum, transformed back (the node passes read it).
void FieldWorkspace::enqueue_raman(const cplx* spectrum, const AdamsRecord* R) {
  if (!P_.n_raman) return;
  if (raman_dynamic_) enqueue_raman_kernel(R);
  if (P_.raman_analytic) {
    // The modes' Hilbert transforms (a real inverse transform of -i Y), synthesised at
    // the nodes by the drive pass, or by a GEMM on the dense route.
    launch_field_scatter_hilbert(P_, spectrum, B_, stream_);
    inv_plan_.exec(B_.modal_h, scratch_, +1, stream_);
    if (dense_)
      launch_dgemm(P_.Q, P_.T, P_.C, P_.synthesis, P_.C,
                   reinterpret_cast<const double*>(B_.modal_h), P_.T, B_.node_hilbert, P_.T,
                   stream_);
  }
  launch_field_raman_drive(P_, B_, stream_);
  raman_plan_.exec(B_.raman_drive, raman_scratch_, -1, stream_);
  launch_raman_product(B_.raman_drive, B_.raman_H, B_.roots_2T, P_.T, P_.Q, stream_);
  raman_plan_.exec(B_.raman_drive, raman_scratch_, +1, stream_);
}

void FieldWorkspace::enqueue_source(const cplx* spectrum, cplx* out, const AdamsRecord* R,
                                    bool finish) {
  const AdamsRecord* rec = R ? R : rec_dev_;
  mark(2);
  // The half spectrum is packed for the real inverse transform as its first pass loads it.
  inv_plan_.exec_packed(PackedSpectrum{spectrum, B_.roots_T, P_.N, P_.inv_T}, B_.modal_t,
                        scratch_, +1, stream_);
  mark(3);
  enqueue_synthesis();
  mark(4);
  enqueue_raman(spectrum, rec);
  mark(5);
  if (fused_node_passes(P_, B_))
    launch_node_passes_fused(P_, rec, pt_, B_, stream_, marking_ ? marks_ + 6 : nullptr);
  else
    launch_node_passes(P_, rec, pt_, B_, stream_, marking_ ? marks_ + 6 : nullptr);
  enqueue_projection();
  mark(11);
  enqueue_absorbers();
  mark(12);
  fwd_plan_.exec(B_.streams, scratch_, -1, stream_);
  mark(13);
  if (finish) launch_field_finish(P_, rec, B_, out, stream_);
  mark(14);
}

void FieldWorkspace::enqueue_trial(bool commit) {
  // An accepted step's commit rides at the head of the next trial (its slot is in the
  // record), so that the host launches one graph per attempt; without the split
  // absorber, which acts on the committed field, the predictor applies it bin by bin.
  const bool separate = commit && split_;
  if (separate) {
    launch_adams_commit(P_, rec_dev_, B_, stream_);
    enqueue_split(rec_dev_, 0.0);
  }
  mark(0);
  launch_adams_predict(P_, rec_dev_, B_, width_, stream_, commit && !separate);
  mark(1);
  // The corrector finishes each bin's source itself (into B.source, for the commit).
  enqueue_source(B_.lab, B_.source, nullptr, false);
  launch_adams_correct(P_, rec_dev_, B_, width_, stream_, true);
  mark(15);
  launch_adams_norms(P_, B_, stream_);
  mark(16);
}

void FieldWorkspace::enqueue_start() {
  enqueue_source(B_.y, B_.source);
  launch_adams_start(P_, rec_dev_, B_, stream_);
}

void FieldWorkspace::enqueue(int key) {
  // The record's upload is the graph's first node: the host fills the pinned record,
  // then launches one graph.
  launch_record_fetch(rec_host_dev_, rec_dev_, 1, stream_);
  const int kind = key & 15;
  if (kind == kStart)
    enqueue_start();
  else
    enqueue_trial(kind == kCommitTrial);
}

void FieldWorkspace::run(int key) {
  // While profiling, the trial runs as direct launches with an event between kernels
  // (events recorded by graph nodes cannot be timed); the kernels are the same.
  marking_ = profile_ && (key & 15) != kStart;
  if (!use_graphs_ || marking_) {
    enqueue(key);
    return;
  }
  auto it = graphs_.find(key);
  if (it == graphs_.end()) {
    const hipGraphExec_t exec = capture_graph(stream_, [&] { enqueue(key); });
    it = graphs_.emplace(key, exec).first;
  }
  SP_HIP(hipGraphLaunch(it->second, stream_));
  graph_launches_ += 1;
}

void FieldWorkspace::wait() {
  // Poll the stream's last event: the host's next decision needs this trial's sums as
  // soon as they land (a blocking wait adds a wake-up); past a short spin the poll
  // yields and backs off, and a device that never finishes is reported.
  SP_HIP(hipEventRecord(done_, stream_));
  wait_event(done_);
}

void FieldWorkspace::flush() {
  if (pending_slot_ < 0) return;
  launch_adams_accept(P_, pending_slot_, compensated_, B_, stream_);
  if (split_) enqueue_split(nullptr, pending_dz_);
  pending_slot_ = -1;
}

// The split-step absorber on the committed state (the workspace's finish: the mask, the
// collar over the accepted step, the mask again): y to the time samples through B.modal_t,
// times exp(-alpha h / 2), and back. h is the record's accepted step when R is given
// (inside a graph), else `dz`.
void FieldWorkspace::enqueue_split(const AdamsRecord* R, double dz) {
  launch_field_split_pack(P_, B_, stream_);
  inv_plan_.exec(B_.modal_t, scratch_, +1, stream_);
  // What the split takes and finds, from the masked state before it (a bound monitor;
  // the split needs n_to == n_t, so these are the coarse samples).
  if (monitor_.bound())
    launch_monitor_sums(reinterpret_cast<const double*>(B_.modal_t), kSamplesReal, P_.C, P_.T,
                        monitor_.columns(), monitor_.weights(), dz,
                        R ? &R->accepted_dz : nullptr, kMonitorSplit, monitor_.partials(),
                        monitor_.split(), stream_);
  launch_field_split_collar(P_, R, dz, B_, stream_);
  inv_plan_.exec(B_.modal_t, scratch_, -1, stream_);
  launch_field_split_unpack(P_, B_, stream_);
}

// The phases this problem's trial launches work in; the others are two events with
// nothing between them and are not reported.
bool FieldWorkspace::phase_used(int i) const {
  switch (i) {
    case 3:   // synthesis_gemm
    case 10:  // projection_gemm
      return dense_;
    case 1: return false;  // the packing rides the inverse transform's first pass
    case 4: return P_.n_raman > 0;
    case 5: case 6: case 7: case 8: return P_.has_plasma != 0;
    case 9: return !fused_node_passes(P_, B_);  // the fused passes have no response of its own
    case 11: return absorbing();
    case 13: return false;  // a trial's corrector finishes the source
    default: return true;
  }
}

std::vector<std::pair<const char*, double>> FieldWorkspace::phases() const {
  std::vector<std::pair<const char*, double>> out;
  const bool fused = fused_node_passes(P_, B_);
  for (int i = 0; i < kPhases; ++i) {
    if (!phase_used(i)) continue;
    // The fused passes' third and fourth launches: the free current with the response,
    // and the combination.
    const char* name = kPhaseNames[i];
    if (fused && i == 7) name = "plasma_current_response";
    if (fused && i == 8) name = "plasma_combine";
    out.emplace_back(name, timed_ ? phase_ms_[i] * 1e3 / timed_ : 0.0);
  }
  return out;
}

// -------------------------------------------------------------- host scalars

void FieldWorkspace::fill_position(double zp, AdamsRecord* into) {
  AdamsRecord& r = into ? *into : *rec_host_;
  for (int t = 0; t < P_.n_pref; ++t) r.pref_w[t] = pref_paths_[t](zp);
  r.density = density_(zp);
  r.preionisation = preion_(zp);
  for (int i = 0; i < P_.n_raman; ++i) r.raman_rate[i] = raman_paths_[i](zp);
  r.compensated = compensated_ ? 1 : 0;
}

// ------------------------------------------------------------------ the engine

void FieldWorkspace::start(double z) {
  flush();
  fill_position(z);
  for (int t = 0; t < P_.n_phi; ++t) slot_rates_[0][t] = rate_paths_[t](z);
  run(kStart);
  SP_HIP(hipStreamSynchronize(stream_));
  evaluations_ += 1;
}

void FieldWorkspace::fill_trial(double z, double dz, const adams::Stencil& s) {
  AdamsRecord& r = *rec_host_;
  const double end = z + dz;
  r.dz = dz;
  for (int t = 0; t < P_.n_phi; ++t) {
    r.x_w[t] = path_increment(phi_paths_[t], z, end);
    r.rate_end_w[t] = rate_paths_[t](end);
  }
  fill_position(end);
  r.order = s.order;
  r.n_orders = static_cast<int>(s.orders.size());
  int top = 0;
  for (int i = 0; i < r.n_orders; ++i) {
    r.orders[i] = s.orders[i];
    top = std::max(top, s.orders[i]);
  }
  r.width = 1 + top;
  r.n_read = r.width - 1;
  SP_CHECK(r.width >= 2 && r.width <= kMaxAdamsWidth - 1, "stencil width outside 2..7");
  width_ = r.width;
  // Unused slots point at row 0 and unused nodes are zero: the kernels read the whole
  // width with zero coefficients past each order.
  std::fill(r.slots, r.slots + kMaxAdamsWidth, 0);
  std::fill(r.nodes, r.nodes + kMaxAdamsWidth, 0.0);
  SP_CHECK(r.width <= kMaxAdamsWidth, "the stencil is wider than the device record");
  for (int j = 0; j < r.n_read; ++j) {
    r.slots[j] = s.rows[j];
    r.nodes[j] = s.nodes[j];
    for (int t = 0; t < kMaxTerms; ++t) r.slot_rate_w[j][t] = slot_rates_[s.rows[j]][t];
  }
  double rows[kMaxAdamsWidth * kMaxAdamsWidth];
  double points[kMaxAdamsWidth + 1];
  double weights[kMaxAdamsWidth + 1];
  std::memset(r.pred_c, 0, sizeof r.pred_c);
  std::memset(r.corr_c, 0, sizeof r.corr_c);
  std::memset(r.rows_c, 0, sizeof r.rows_c);
  std::memset(r.comp_pred_w, 0, sizeof r.comp_pred_w);
  std::memset(r.comp_corr_w, 0, sizeof r.comp_corr_w);
  // The predictor's basis on the order newest nodes.
  lagrange_basis(s.nodes.data(), s.order, 1.0, false, rows, nullptr);
  for (int j = 0; j < s.order; ++j)
    for (int k = 0; k < s.order; ++k) r.pred_c[j][k] = rows[j * s.order + k];
  // The corrector through the new sample, [1, nodes[:order]].
  points[0] = 1.0;
  for (int j = 0; j < s.order; ++j) points[1 + j] = s.nodes[j];
  lagrange_basis(points, s.order + 1, 1.0, false, rows, nullptr);
  for (int j = 0; j <= s.order; ++j)
    for (int k = 0; k <= s.order; ++k) r.corr_c[j][k] = rows[j * (s.order + 1) + k];
  // Each scored order's defect: its corrector less its predictor; the compensation's
  // ordinary Adams weights of both.
  for (int i = 0; i < r.n_orders; ++i) {
    const int k = r.orders[i];
    points[0] = 1.0;
    for (int j = 0; j < k; ++j) points[1 + j] = s.nodes[j];
    lagrange_basis(points, k + 1, 1.0, false, rows, nullptr);
    for (int j = 0; j <= k; ++j)
      for (int c = 0; c <= k; ++c) r.rows_c[i][j][c] = rows[j * (k + 1) + c];
    lagrange_basis(s.nodes.data(), k, 1.0, false, rows, nullptr);
    for (int j = 0; j < k; ++j)
      for (int c = 0; c < k; ++c) r.rows_c[i][1 + j][c] -= rows[j * k + c];
    lagrange_basis(s.nodes.data(), k, 1.0, true, nullptr, weights);
    for (int j = 0; j < k; ++j) r.comp_pred_w[i][j] = weights[j];
    lagrange_basis(points, k + 1, 1.0, true, nullptr, weights);
    for (int j = 0; j <= k; ++j) r.comp_corr_w[i][j] = weights[j];
  }
  r.free_slot = pending_slot_;
  held_rate_end_.assign(r.rate_end_w, r.rate_end_w + kMaxTerms);
}

std::vector<double> FieldWorkspace::trial_norms() {
  const AdamsRecord& r = *rec_host_;
  const double magnitude = std::sqrt(norms_host_[0]);
  const double divisor = rtol_ * magnitude + atol_;
  std::vector<double> norms(r.n_orders);
  for (int i = 0; i < r.n_orders; ++i) norms[i] = std::sqrt(norms_host_[1 + i]) / divisor;
  trials_ += 1;
  evaluations_ += 1;
  return norms;
}

std::vector<double> FieldWorkspace::trial(double z, double dz, const adams::Stencil& s) {
  fill_trial(z, dz, s);
  // The trial's source evaluation replaces B.lab: the last accepted one's is gone.
  quad_ready_ = false;
  trial_end_ = z + dz;
  run((pending_slot_ >= 0 ? kCommitTrial : kTrial) + 16 * width_);
  pending_slot_ = -1;
  if (marking_)
    SP_HIP(hipStreamSynchronize(stream_));
  else
    wait();
  if (marking_) {
    float ms = 0.0f;
    for (int i = 0; i < kPhases; ++i) {
      SP_HIP(hipEventElapsedTime(&ms, marks_[i], marks_[i + 1]));
      phase_ms_[i] += ms;
    }
    SP_HIP(hipEventElapsedTime(&ms, marks_[0], marks_[kPhases]));
    time_total_ms_ += ms;
    timed_ += 1;
  }
  return trial_norms();
}

void FieldWorkspace::accept(double dz, int free, const adams::Stencil&,
                            const std::vector<adams::Entry>&) {
  // Deferred into the next trial's graph, or flushed before anything reads the state.
  pending_slot_ = free;
  pending_dz_ = dz;
  rec_host_->accepted_dz = dz;
  slot_rates_[free] = held_rate_end_;
  accepted_ += 1;
  quad_ready_ = true;
  quad_z_ = trial_end_;
}

// ------------------------------------------------ the low-storage RK4IP replay

void FieldWorkspace::enqueue_ls_step() {
  SP_HIP(hipMemcpyAsync(ls_rec_dev_, ls_rec_host_, 3 * sizeof(AdamsRecord),
                        hipMemcpyHostToDevice, stream_));
  // Pass 1 reads the source at the committed state, z; passes 2 and 3 the stage field at
  // the midpoint; pass 4 at the end.
  // Each pass finishes its own bins' source (no finish launch between them).
  enqueue_source(B_.y, B_.source, ls_rec_dev_, false);
  launch_ls_pass(1, P_, ls_rec_dev_, B_, ls_acc_, ls_prev_, ls_partials_, stream_, true);
  enqueue_source(B_.lab, B_.source, ls_rec_dev_ + 1, false);
  launch_ls_pass(2, P_, ls_rec_dev_, B_, ls_acc_, ls_prev_, ls_partials_, stream_, true);
  enqueue_source(B_.lab, B_.source, ls_rec_dev_ + 1, false);
  launch_ls_pass(3, P_, ls_rec_dev_, B_, ls_acc_, ls_prev_, ls_partials_, stream_, true);
  enqueue_source(B_.lab, B_.source, ls_rec_dev_ + 2, false);
  launch_ls_pass(4, P_, ls_rec_dev_, B_, ls_acc_, ls_prev_, ls_partials_, stream_, true);
  launch_ls_norms(P_, ls_partials_, B_.norms, stream_);
}

double FieldWorkspace::ls_trial(double z, double dz) {
  const double mid = z + 0.5 * dz, end = z + dz;
  // The step's last source evaluation (its field in B.lab, at `end`) is the one a
  // quadrature estimate repeats once the step is accepted.
  quad_ready_ = false;
  trial_end_ = end;
  // The three positions' scalars (each evaluation reads its own position), and the
  // flows' weights: Phi(mid) - Phi(z) in record 0, Phi(end) - Phi(mid) in record 1.
  const double at[3] = {z, mid, end};
  for (int i = 0; i < 3; ++i) fill_position(at[i], ls_rec_host_ + i);
  ls_rec_host_[0].dz = dz;
  for (int t = 0; t < P_.n_phi; ++t) {
    ls_rec_host_[0].x_w[t] = path_increment(phi_paths_[t], z, mid);
    ls_rec_host_[1].x_w[t] = path_increment(phi_paths_[t], mid, end);
  }
  marking_ = false;
  if (use_graphs_) {
    if (!ls_graph_) ls_graph_ = capture_graph(stream_, [&] { enqueue_ls_step(); });
    SP_HIP(hipGraphLaunch(ls_graph_, stream_));
    graph_launches_ += 1;
  } else {
    enqueue_ls_step();
  }
  wait();
  trials_ += 1;
  evaluations_ += 4;
  const double lower = std::sqrt(norms_host_[0]), upper = std::sqrt(norms_host_[1]);
  stiffness_ = lower > 0.0 ? 2.0 * upper / lower : std::nan("");
  const double squared = norms_host_[2];
  // The history's norm is "not estimated"; a non-finite state is the replay's divergence.
  return std::isfinite(squared) ? 0.0 : squared;
}

void FieldWorkspace::ls_accept(double dz) {
  // The hard mask was applied by the step's last pass; nothing else finishes a step here.
  (void)dz;
  accepted_ += 1;
  quad_ready_ = true;
  quad_z_ = trial_end_;
}

namespace {
struct LsEngine final : march::Engine {
  std::function<double(double, double)> step;
  std::function<void(double)> commit;
  double trial(double z, double dz) override { return step(z, dz); }
  void accept(double dz) override { commit(dz); }
};

struct LsHost final : march::Host {
  AttemptLog& log;
  explicit LsHost(AttemptLog& l) : log(l) {}
  void record(double end, double dz, double norm, bool acc) override {
    log.record(end, dz, norm, acc);
  }
  bool halt() override { return log.halt(); }
};
}  // namespace

bool FieldWorkspace::ls_advance(const Bounds& b, int columns, std::vector<double>& rows,
                                sp_march_result* r) {
  LsEngine engine;
  engine.step = [this](double z, double dz) { return ls_trial(z, dz); };
  engine.commit = [this](double dz) { ls_accept(dz); };
  AttemptLog log;
  log.begin(&rows, columns, b);
  log.order = 4;  // RK4IP
  LsHost host(log);
  // The replay stops at the target or at the schedule's end, whichever comes first.
  const double target = std::min(b.target, edges_.back());
  bool done = false;
  try {
    done = march::replay_to(engine, host, ls_state_, steps_, edges_, target) &&
           march::reached(ls_state_.z, target);
  } catch (...) {
    z_ = ls_state_.z;
    r->z = ls_state_.z;
    r->attempts = log.attempts;
    r->accepted = log.accepted;
    r->evaluations = evaluations_;
    throw;
  }
  z_ = ls_state_.z;
  r->z = ls_state_.z;
  r->proposal = ls_state_.cursor < static_cast<long>(steps_.size())
                    ? steps_[static_cast<std::size_t>(ls_state_.cursor)]
                    : 0.0;
  r->done = done ? 1 : 0;
  r->attempts = log.attempts;
  r->accepted = log.accepted;
  r->evaluations = evaluations_;
  return done;
}

// -------------------------------------------------------------- the interface

void FieldWorkspace::load(double z, const double* field) {
  SP_CHECK(field != nullptr, "load needs a field");
  const std::size_t CN = static_cast<std::size_t>(P_.C) * P_.N;
  pending_slot_ = -1;
  monitor_.reset(stream_);
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(B_.y, field, CN * sizeof(cplx), hipMemcpyHostToDevice));
  adams::Engine& engine = *this;
  march_ = std::make_unique<adams::March>(policy_, engine, host_, z, dz_init_);
  if (low_storage_) {
    // A replay resumes at one of its schedule's edges.
    long cursor = -1;
    for (std::size_t i = 0; i < edges_.size(); ++i)
      if (march::reached(z, edges_[i]) && march::reached(edges_[i], z))
        cursor = static_cast<long>(i);
    SP_CHECK(cursor >= 0, "a replay loads its state at one of its schedule's edges");
    ls_state_ = march::State{};
    ls_state_.z = edges_[static_cast<std::size_t>(cursor)];
    ls_state_.cursor = cursor;
  }
  z_ = z;
  loaded_ = true;
  quad_ready_ = false;
}

double FieldWorkspace::probe_trial(double z, double dz) {
  if (!loaded_) throw Error(kState, "load a state before taking a trial");
  if (low_storage_) {
    // One step from the committed state, which is kept: the step's end into `advanced`.
    const std::size_t bytes = sizeof(cplx) * P_.C * P_.N;
    SP_HIP(hipMemcpyAsync(ls_save_, B_.y, bytes, hipMemcpyDeviceToDevice, stream_));
    const double norm = ls_trial(z, dz);
    SP_HIP(hipMemcpyAsync(B_.advanced, B_.y, bytes, hipMemcpyDeviceToDevice, stream_));
    SP_HIP(hipMemcpyAsync(B_.y, ls_save_, bytes, hipMemcpyDeviceToDevice, stream_));
    SP_HIP(hipStreamSynchronize(stream_));
    return norm;
  }
  if (march_->history.empty()) {
    march_->z = z;
    start(z);
    march_->history.assign(1, adams::Entry{z, 0, 0});
    march_->ledger.startup += 1;
  }
  const adams::Stencil s = adams::stencil(policy_, z, dz, march_->history, march_->order,
                                          march_->last_rejected);
  return trial(z, dz, s)[0];
}

void FieldWorkspace::read_trial(double* out) {
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, B_.advanced, sizeof(cplx) * P_.C * P_.N, hipMemcpyDeviceToHost));
}

void FieldWorkspace::read_state(double* out) {
  if (!loaded_) throw Error(kState, "no state is loaded");
  flush();
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, B_.y, sizeof(cplx) * P_.C * P_.N, hipMemcpyDeviceToHost));
}

void FieldWorkspace::export_state(const double* weight, bool optics, double* out) {
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

bool FieldWorkspace::advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                                     sp_march_result* r) {
  if (!loaded_) throw Error(kState, "load a state before advancing");
  if (low_storage_) return ls_advance(b, columns, rows, r);
  host_.log.begin(&rows, columns, b);
  bool done = false;
  try {
    done = march_->march_to(b.target) && march::reached(march_->z, b.target);
  } catch (...) {
    host_.log.rows = nullptr;
    z_ = march_->z;
    r->z = march_->z;
    r->proposal = march_->proposal;
    r->attempts = host_.log.attempts;
    r->accepted = host_.log.accepted;
    r->evaluations = evaluations_;
    flush();
    throw;
  }
  flush();
  host_.log.rows = nullptr;
  z_ = march_->z;
  r->z = march_->z;
  r->proposal = march_->proposal;
  r->done = done ? 1 : 0;
  r->attempts = host_.log.attempts;
  r->accepted = host_.log.accepted;
  r->evaluations = evaluations_;
  return done;
}

void FieldWorkspace::source(double z, const double* lab, double* out, double* streams,
                            double* modal) {
  SP_CHECK(lab && out, "the source needs a field and an output");
  const std::size_t CN = static_cast<std::size_t>(P_.C) * P_.N;
  const std::size_t CT = static_cast<std::size_t>(P_.C) * P_.T;
  flush();
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(B_.lab, lab, CN * sizeof(cplx), hipMemcpyHostToDevice));
  fill_position(z);
  SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(AdamsRecord), hipMemcpyHostToDevice,
                        stream_));
  marking_ = false;
  inv_plan_.exec_packed(PackedSpectrum{B_.lab, B_.roots_T, P_.N, P_.inv_T}, B_.modal_t,
                        scratch_, +1, stream_);
  enqueue_synthesis();
  enqueue_raman(B_.lab, rec_dev_);
  // The unfused passes leave every node's rate, fraction and free current for inspection
  // (buffer()); the streams are then those of the fused passes where the trial runs them.
  launch_node_passes(P_, rec_dev_, pt_, B_, stream_);
  if (fused_node_passes(P_, B_)) launch_node_passes_fused(P_, rec_dev_, pt_, B_, stream_);
  enqueue_projection();
  enqueue_absorbers();
  SP_HIP(hipStreamSynchronize(stream_));
  if (modal) SP_HIP(hipMemcpy(modal, B_.modal_t, CT * sizeof(double), hipMemcpyDeviceToHost));
  if (streams)
    SP_HIP(hipMemcpy(streams, B_.streams, 3 * CT * sizeof(double), hipMemcpyDeviceToHost));
  fwd_plan_.exec(B_.streams, scratch_, -1, stream_);
  launch_field_finish(P_, rec_dev_, B_, B_.source, stream_);
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, B_.source, CN * sizeof(cplx), hipMemcpyDeviceToHost));
}

long FieldWorkspace::buffer(const char* name, double* out, long count) {
  SP_CHECK(name != nullptr, "a buffer name is required");
  if (std::strcmp(name, "stiffness") == 0) {
    // The low-storage replay's last step's h rho (a host value).
    if (out && count > 0) out[0] = stiffness_;
    return 1;
  }
  const long CN2 = 2L * P_.C * P_.N;
  const long S = policy_.order + 2;
  const struct {
    const char* name;
    const void* ptr;
    long doubles;
  } table[] = {
      {"y", B_.y, CN2},           {"lab", B_.lab, CN2},
      {"source", B_.source, CN2}, {"advanced", B_.advanced, CN2},
      {"store", B_.store, S * CN2},
      {"states", B_.states, B_.states ? S * CN2 : 0},
      {"carried", B_.carried, B_.carried ? S * CN2 : 0},
      {"node_field", B_.node_field, B_.node_field ? static_cast<long>(P_.Q) * P_.T : 0},
      {"rate", B_.rate, B_.rate ? static_cast<long>(P_.Q) * P_.T : 0},
      {"frac", B_.frac, B_.frac ? static_cast<long>(P_.Q) * P_.T : 0},
      {"free_local", B_.free_local, B_.free_local ? static_cast<long>(P_.Q) * P_.T : 0},
      {"x_local", B_.x_local, B_.x_local ? static_cast<long>(P_.Q) * P_.T : 0},
      {"tile_x", B_.tile_x, static_cast<long>(P_.Q) * scan_tiles(P_.T)},
      {"tile_f", B_.tile_f, static_cast<long>(P_.Q) * scan_tiles(P_.T)},
  };
  flush();
  for (const auto& row : table) {
    if (std::strcmp(row.name, name) != 0) continue;
    if (out && count > 0 && row.ptr) {
      SP_HIP(hipStreamSynchronize(stream_));
      SP_HIP(hipMemcpy(out, row.ptr, sizeof(double) * std::min(count, row.doubles),
                       hipMemcpyDeviceToHost));
    }
    return row.doubles;
  }
  throw std::invalid_argument(std::string("no device buffer named ") + name);
}

void FieldWorkspace::bind_monitor(const double* weights, const double* columns) {
  SP_HIP(hipStreamSynchronize(stream_));
  // The captured trials hold the monitor's kernels or not: capture them again.
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
  if (!weights) {
    monitor_.unbind();
    return;
  }
  monitor_.bind(weights, columns, P_.C, n_t_, stream_);
}

void FieldWorkspace::monitor_reading(double* out) {
  SP_CHECK(monitor_.bound(), "no boundary monitor is bound");
  if (!loaded_) throw Error(kState, "no state is loaded");
  flush();
  if (!stats_)
    stats_ = std::make_unique<PlaneStatistics>(P_.C, P_.N, n_t_, true,
                                               P_.dt_fine * P_.T / n_t_, P_.omega);
  stats_->reading(B_.y, monitor_.columns(), monitor_.weights(), monitor_.partials(),
                  monitor_.reading(), stream_);
  monitor_.drain(out, stream_);
}

std::vector<double> FieldWorkspace::statistics(std::uint64_t flags, const double* windows,
                                               int n_windows, const StatExtras& extras) {
  if (!loaded_) throw Error(kState, "no state is loaded");
  flush();
  // The coarse grid: n_t samples spaced by the fine spacing times n_to / n_t.
  if (!stats_)
    stats_ = std::make_unique<PlaneStatistics>(P_.C, P_.N, n_t_, true,
                                               P_.dt_fine * P_.T / n_t_, P_.omega);
  std::vector<double> trace;
  double intensity = 0.0;
  std::vector<double> out =
      stats_->run(B_.y, flags & ~(kStatColumn | kStatQuadrature), windows, n_windows, stream_,
                  &trace, &extras,
                  &intensity);
  std::vector<double> populations, diagnostics;
  if (flags & kStatColumn) {
    SP_CHECK(P_.has_plasma, "the plasma statistics need a compiled plasma response, and this "
                            "problem declares none");
    // The plasma column at the committed state and its position: the node passes on its
    // modal fields, then the densest node's end fraction and largest rate.
    fill_position(z_);
    SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(AdamsRecord), hipMemcpyHostToDevice,
                          stream_));
    launch_field_scatter(P_, B_.y, B_, stream_);
    inv_plan_.exec(B_.modal_t, scratch_, +1, stream_);
    enqueue_synthesis();
    launch_node_passes(P_, rec_dev_, pt_, B_, stream_, nullptr,
                       (flags & kStatPlasmaIntegration) != 0);
    launch_plasma_densest(B_.frac, B_.rate, P_.Q, P_.T, B_.partials, stream_);
    double densest[2];
    SP_HIP(hipMemcpyAsync(densest, B_.partials, sizeof(densest), hipMemcpyDeviceToHost,
                          stream_));
    if (flags & kStatPopulations) {
      // Each node's end-of-window fraction and largest rate.
      if (!populations_) populations_ = alloc<double>(2 * static_cast<std::size_t>(P_.Q));
      launch_plasma_populations(B_.frac, B_.rate, P_.Q, P_.T, populations_, stream_);
      populations.assign(2 * static_cast<std::size_t>(P_.Q), 0.0);
      SP_HIP(hipMemcpyAsync(populations.data(), populations_,
                            populations.size() * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    }
    if (flags & kStatPlasmaIntegration) {
      // Each node's six parts, then their maxima over the nodes.
      if (!diagnostics_) diagnostics_ = alloc<double>(6 * static_cast<std::size_t>(P_.Q) + 6);
      launch_plasma_diagnostics(P_, pt_, B_, diagnostics_, diagnostics_ + 6L * P_.Q, stream_);
      diagnostics.assign(6, 0.0);
      SP_HIP(hipMemcpyAsync(diagnostics.data(), diagnostics_ + 6L * P_.Q, 6 * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
    }
    SP_HIP(hipStreamSynchronize(stream_));
    if (flags & kStatElectronDensity) out.push_back(density_(z_) * densest[0]);
    if (flags & kStatPeakIonisationRate) out.push_back(densest[1]);
  }
  if (flags & kStatPeakIntensity) out.push_back(intensity);
  if (flags & kStatQuadrature) out.push_back(quadrature_estimate());
  out.insert(out.end(), trace.begin(), trace.end());
  out.insert(out.end(), populations.begin(), populations.end());
  out.insert(out.end(), diagnostics.begin(), diagnostics.end());
  return out;
}

void FieldWorkspace::enqueue_response_spectrum(const DeviceProblem& P, cplx* out) {
  inv_plan_.exec_packed(PackedSpectrum{B_.lab, B_.roots_T, P_.N, P_.inv_T}, B_.modal_t,
                        scratch_, +1, stream_);
  enqueue_synthesis();
  enqueue_raman(B_.lab, rec_dev_);
  if (fused_node_passes(P, B_))
    launch_node_passes_fused(P, rec_dev_, pt_, B_, stream_);
  else
    launch_node_passes(P, rec_dev_, pt_, B_, stream_);
  if (dense_)
    launch_collar_projection(P.C, P.T, P.Q, P.projection, B_.node_streams, P.has_plasma != 0,
                             P.t_window, P.t_deriv, reinterpret_cast<double*>(B_.streams),
                             stream_);
  fwd_plan_.exec(B_.streams, scratch_, -1, stream_);
  launch_field_finish(P, rec_dev_, B_, out, stream_);
}

double FieldWorkspace::quadrature_estimate() {
  // NaN without an embedded rule or before an accepted step since the state was loaded.
  if (!projection_coarse_ || !quad_ready_) return std::nan("");
  const std::size_t CN = static_cast<std::size_t>(P_.C) * P_.N;
  if (!quad_fine_) {
    quad_fine_ = alloc<cplx>(CN);
    quad_coarse_ = alloc<cplx>(CN);
    quad_out_ = alloc<double>(1);
  }
  // The accepted trial's position terms (prefactor, density, preionisation, Raman rates).
  fill_position(quad_z_);
  SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, sizeof(AdamsRecord), hipMemcpyHostToDevice,
                        stream_));
  DeviceProblem P = P_;
  P.absorber_stream = -1;  // the responses alone: no absorber term in either finish
  enqueue_response_spectrum(P, quad_fine_);
  P.projection = projection_coarse_;
  enqueue_response_spectrum(P, quad_coarse_);
  launch_quadrature_score(quad_fine_, quad_coarse_, static_cast<long>(CN), quad_out_, stream_);
  double value = 0.0;
  SP_HIP(hipMemcpyAsync(&value, quad_out_, sizeof(double), hipMemcpyDeviceToHost, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
  return value;
}

// ------------------------------------------------------------ batched marching

std::string FieldWorkspace::batch_refusal() const {
  if (dense_) return "a batched march runs the modal route (at most 6 modes), not the dense one";
  if (low_storage_) return "a batched march instance runs the exponential Adams march";
  if (P_.C > kMaxBatchModes) return "a batched march instance has at most 6 modes";
  if (P_.n_raman) return "a batched march instance has no Raman response";
  if (absorbing() || split_) return "a batched march instance has no absorber";
  if (correct_threads(P_) != kWave) return "the instance's corrector needs one-wave workgroups";
  return {};
}

void FieldWorkspace::batch_bind(AdamsRecord* host, AdamsRecord* device) {
  *host = *rec_host_;
  rec_host_ = host;
  SP_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&rec_host_dev_), rec_host_, 0));
  rec_dev_ = device;
  // The workspace's own graphs read its own record: they are not used again.
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
  use_graphs_ = false;
}

BatchInstance FieldWorkspace::batch_instance() const {
  BatchInstance in{};
  in.P = P_;
  in.B = B_;
  in.pt = pt_;
  in.correct_blocks = n_blocks_;
  return in;
}

std::vector<FftJob> FieldWorkspace::batch_inverse_jobs(const int* gate) const {
  const PackedSpectrum packed{B_.lab, B_.roots_T, P_.N, P_.inv_T};
  // The plan copies the spectrum's fields into the jobs; `packed` may go.
  return inv_plan_.jobs(B_.modal_t, scratch_, &packed, gate);
}

std::vector<FftJob> FieldWorkspace::batch_forward_jobs(const int* gate) const {
  return fwd_plan_.jobs(B_.streams, scratch_, nullptr, gate);
}

void FieldWorkspace::batch_begin(double target, long work_limit, std::vector<double>* rows) {
  if (!loaded_) throw Error(kState, "load a state before advancing");
  Bounds b;
  b.target = target;
  b.attempts = work_limit;
  host_.log.begin(rows, 4, b);
  batch_target_ = target;
  batch_finished_ = false;
  batch_done_ = false;
  batch_status_ = kOk;
  batch_message_.clear();
}

int FieldWorkspace::batch_plan() {
  AdamsRecord& r = *rec_host_;
  // A pending commit rides at the head of the round whatever the action.
  r.free_slot = pending_slot_;
  r.action = kBatchIdle;
  if (batch_finished_) return kBatchIdle;
  try {
    switch (march_->plan(batch_target_, &request_)) {
      case adams::Request::kDone:
        batch_finished_ = true;
        batch_done_ = march::reached(march_->z, batch_target_);
        return kBatchIdle;
      case adams::Request::kHalted:
        batch_finished_ = true;
        return kBatchIdle;
      case adams::Request::kStart:
        fill_position(request_.z);
        for (int t = 0; t < P_.n_phi; ++t) slot_rates_[0][t] = rate_paths_[t](request_.z);
        r.action = kBatchStart;
        return kBatchStart;
      case adams::Request::kTrial:
        fill_trial(request_.z, request_.dz, request_.stencil);
        r.action = kBatchTrial;
        return kBatchTrial;
    }
  } catch (const Error& e) {
    batch_status_ = e.code;
    batch_message_ = e.what();
  } catch (const march::ConvergenceError& e) {
    batch_status_ = kConvergence;
    batch_message_ = e.what();
  } catch (const std::exception& e) {
    batch_status_ = kBackend;
    batch_message_ = e.what();
  }
  batch_finished_ = true;
  r.action = kBatchIdle;
  return kBatchIdle;
}

void FieldWorkspace::batch_conclude(int action) {
  // The round carried the pending commit (if any) at its head.
  pending_slot_ = -1;
  if (action == kBatchStart) {
    evaluations_ += 1;
    march_->started();
    return;
  }
  if (action != kBatchTrial) return;
  try {
    march_->conclude(request_, trial_norms());
  } catch (const Error& e) {
    batch_status_ = e.code;
    batch_message_ = e.what();
    batch_finished_ = true;
  } catch (const march::ConvergenceError& e) {
    batch_status_ = kConvergence;
    batch_message_ = e.what();
    batch_finished_ = true;
  } catch (const std::exception& e) {
    batch_status_ = kBackend;
    batch_message_ = e.what();
    batch_finished_ = true;
  }
}

int FieldWorkspace::batch_end(sp_march_result* out, std::string* message) {
  // An accepted step the last round concluded is committed here, on the batch's stream.
  flush();
  host_.log.rows = nullptr;
  z_ = march_->z;
  out->z = march_->z;
  out->proposal = march_->proposal;
  out->done = batch_done_ ? 1 : 0;
  out->attempts = host_.log.attempts;
  out->accepted = host_.log.accepted;
  out->evaluations = evaluations_;
  if (message) *message = batch_message_;
  return batch_status_;
}

sp_counters FieldWorkspace::counters() const {
  sp_counters c{};
  c.trials = trials_;
  c.evaluations = evaluations_;
  c.accepted = accepted_;
  c.graph_launches = graph_launches_;
  c.device_bytes =
      static_cast<long>(device_bytes_ + inv_plan_.bytes() + fwd_plan_.bytes() +
                        raman_plan_.bytes() + raman_kernel_plan_.bytes() +
                        (stats_ ? stats_->bytes() : 0));
  return c;
}

}  // namespace sp
