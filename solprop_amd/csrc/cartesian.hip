// The free-space Cartesian route (see cartesian.hpp).
#include "cartesian.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>

#include "fft_lds.hpp"
#include "sync.hpp"

namespace sp {
namespace {

struct KerrArgs {
  double coefficient, self, conj_term;
};

__device__ inline cplx kerr_point(cplx e, const KerrArgs& k) {
  cplx p = scale(k.self * norm2(e), e);
  if (k.conj_term != 0.0) p = add(p, scale(k.conj_term, mul(mul(e, e), conj(e))));
  return scale(k.coefficient, p);
}

// The y axis's inverse transform, the response at every point and the forward transform
// in one pass: G neighbouring y-sequences (points `stride` apart) staged in LDS, in place.
// Each workgroup reads and writes only its own sequences.
template <int LOG2L, int G>
__global__ void __launch_bounds__((Shape<LOG2L, G>::NT))
    cart_y_response(cplx* f, long stride, const cplx* __restrict__ tw, KerrArgs k) {
  using S = Shape<LOG2L, G>;
  constexpr int E = S::E, NT = S::NT, STRIDE = S::STRIDE;
  __shared__ cplx buf[G * STRIDE];
  const long s0 = static_cast<long>(blockIdx.x) * G;
  for (int e = threadIdx.x; e < E; e += NT) {
    const int s = e % G, j = e / G;
    buf[s * STRIDE + j] = f[j * stride + s0 + s];
  }
  __syncthreads();
  lds_transform<LOG2L, G>(buf, tw, +1);
  for (int e = threadIdx.x; e < E; e += NT) {
    const int s = e % G, j = e / G;
    buf[s * STRIDE + j] = kerr_point(buf[s * STRIDE + j], k);
  }
  __syncthreads();
  lds_transform<LOG2L, G>(buf, tw, -1);
  for (int e = threadIdx.x; e < E; e += NT) {
    const int s = e % G, j = e / G;
    f[j * stride + s0 + s] = buf[s * STRIDE + j];
  }
}

// The t axis (contiguous rows of L): the inverse transform of G rows of the spectrum into
// the field, and the forward transform of G rows of the field times the prefactor into
// the source.
template <int LOG2L, int G>
__global__ void __launch_bounds__((Shape<LOG2L, G>::NT))
    cart_t_inverse(const cplx* __restrict__ spectrum, cplx* __restrict__ f,
                   const cplx* __restrict__ tw) {
  using S = Shape<LOG2L, G>;
  constexpr int L = S::L, E = S::E, NT = S::NT, STRIDE = S::STRIDE;
  __shared__ cplx buf[G * STRIDE];
  const long base = static_cast<long>(blockIdx.x) * E;
  for (int e = threadIdx.x; e < E; e += NT) buf[(e / L) * STRIDE + e % L] = spectrum[base + e];
  __syncthreads();
  lds_transform<LOG2L, G>(buf, tw, +1);
  for (int e = threadIdx.x; e < E; e += NT) f[base + e] = buf[(e / L) * STRIDE + e % L];
}

template <int LOG2L, int G>
__global__ void __launch_bounds__((Shape<LOG2L, G>::NT))
    cart_t_forward(const cplx* __restrict__ f, const cplx* __restrict__ pref,
                   cplx* __restrict__ out, const cplx* __restrict__ tw) {
  using S = Shape<LOG2L, G>;
  constexpr int L = S::L, E = S::E, NT = S::NT, STRIDE = S::STRIDE;
  __shared__ cplx buf[G * STRIDE];
  const long base = static_cast<long>(blockIdx.x) * E;
  for (int e = threadIdx.x; e < E; e += NT) buf[(e / L) * STRIDE + e % L] = f[base + e];
  __syncthreads();
  lds_transform<LOG2L, G>(buf, tw, -1);
  for (int e = threadIdx.x; e < E; e += NT)
    out[base + e] = mul(pref[base + e], buf[(e / L) * STRIDE + e % L]);
}

template <int LOG2L, int G>
struct YResponse {
  static void launch(unsigned blocks, hipStream_t s, cplx* f, long stride, const cplx* tw,
                     KerrArgs k) {
    hipLaunchKernelGGL((cart_y_response<LOG2L, G>), dim3(blocks), dim3(Shape<LOG2L, G>::NT), 0,
                       s, f, stride, tw, k);
  }
};

template <int LOG2L, int G>
struct TInverse {
  static void launch(unsigned blocks, hipStream_t s, const cplx* spectrum, cplx* f,
                     const cplx* tw) {
    hipLaunchKernelGGL((cart_t_inverse<LOG2L, G>), dim3(blocks), dim3(Shape<LOG2L, G>::NT), 0,
                       s, spectrum, f, tw);
  }
};

template <int LOG2L, int G>
struct TForward {
  static void launch(unsigned blocks, hipStream_t s, const cplx* f, const cplx* pref, cplx* out,
                     const cplx* tw) {
    hipLaunchKernelGGL((cart_t_forward<LOG2L, G>), dim3(blocks), dim3(Shape<LOG2L, G>::NT), 0,
                       s, f, pref, out, tw);
  }
};

template <template <int, int> class Op, int LOG2L, typename... A>
void dispatch_g(int log2g, A... a) {
  constexpr int L = 1 << LOG2L;
  switch (log2g) {
    case 0: Op<LOG2L, 1>::launch(a...); return;
    case 1: if constexpr (L * 2 <= kMaxElems) { Op<LOG2L, 2>::launch(a...); return; } break;
    case 2: if constexpr (L * 4 <= kMaxElems) { Op<LOG2L, 4>::launch(a...); return; } break;
    case 3: if constexpr (L * 8 <= kMaxElems) { Op<LOG2L, 8>::launch(a...); return; } break;
    case 4: if constexpr (L * 16 <= kMaxElems) { Op<LOG2L, 16>::launch(a...); return; } break;
    default: break;
  }
  throw Error(kValue, "no Cartesian pass kernel for this shape");
}

template <template <int, int> class Op, typename... A>
void dispatch(int log2l, int log2g, A... a) {
  switch (log2l) {
    case 2: dispatch_g<Op, 2>(log2g, a...); break;
    case 3: dispatch_g<Op, 3>(log2g, a...); break;
    case 4: dispatch_g<Op, 4>(log2g, a...); break;
    case 5: dispatch_g<Op, 5>(log2g, a...); break;
    case 6: dispatch_g<Op, 6>(log2g, a...); break;
    case 7: dispatch_g<Op, 7>(log2g, a...); break;
    case 8: dispatch_g<Op, 8>(log2g, a...); break;
    case 9: dispatch_g<Op, 9>(log2g, a...); break;
    case 10: dispatch_g<Op, 10>(log2g, a...); break;
    case 11: dispatch_g<Op, 11>(log2g, a...); break;
    case 12: dispatch_g<Op, 12>(log2g, a...); break;
    default: throw Error(kValue, "a Cartesian axis has 4 to 4096 points");
  }
  SP_HIP(hipGetLastError());
}

int log2_of(int n) {
  int m = 0;
  while ((1 << m) < n) ++m;
  return m;
}

// Sequences a workgroup owns on an axis of n points with `count` sequences: contiguous
// runs of up to 16 points across the sequences (a strided axis), or for the contiguous t
// rows enough rows for 256 points at least; never more than `max_points` points.
int axis_log2g(int n, long count, bool strided, int max_points) {
  int g = 0;
  const long want = strided ? 16 : std::max(1, 256 / n);
  const int cap = std::min(max_points, kMaxElems);
  while ((1L << (g + 1)) <= want && n * (1 << (g + 1)) <= cap && (1L << (g + 1)) <= count &&
         count % (1L << (g + 1)) == 0)
    ++g;
  return g;
}

enum Phase : int {
  kUpload, kPassT, kPassX, kPassY, kFinish, kLs1, kLs2, kLs3, kLs4, kNorms
};
const char* const kPhaseNames[] = {"record upload", "t inverse", "x pass", "y response",
                                   "t forward", "ls pass 1", "ls pass 2", "ls pass 3",
                                   "ls pass 4", "norms"};

struct ReplayEngine final : march::Engine {
  std::function<double(double, double)> step;
  std::function<void(double)> commit;
  double trial(double z, double dz) override { return step(z, dz); }
  void accept(double dz) override { commit(dz); }
};

struct ReplayHost final : march::Host {
  AttemptLog& log;
  explicit ReplayHost(AttemptLog& l) : log(l) {}
  void record(double end, double dz, double norm, bool acc) override {
    log.record(end, dz, norm, acc);
  }
  bool halt() override { return log.halt(); }
};

}  // namespace

template <typename T>
T* CartesianWorkspace::alloc(std::size_t n) {
  T* p = nullptr;
  if (n == 0) n = 1;
  SP_HIP(hipMalloc(&p, n * sizeof(T)));
  allocations_.push_back(p);
  device_bytes_ += n * sizeof(T);
  return p;
}

CartesianWorkspace::CartesianWorkspace(const sp_problem_desc& d) {
  // Refusals first, before anything is allocated.
  SP_CHECK(d.kind == 2, "the Cartesian workspace runs the free-space (x, y, t) grid");
  SP_CHECK(d.method == 2 && d.n_schedule >= 1 && d.schedule_edges,
           "the Cartesian route runs the low-storage RK4IP replay of a schedule");
  nx_ = d.grid_nx;
  ny_ = d.grid_ny;
  nt_ = d.n_omega;
  auto pow2 = [](int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; };
  SP_CHECK(pow2(nx_, 4, 4096) && pow2(ny_, 4, 4096) && pow2(nt_, 4, 4096),
           "the Cartesian grid's nx, ny and nt are powers of two from 4 to 4096");
  SP_CHECK(d.n_modes == nx_ * ny_, "a Cartesian grid's columns are its nx ny points");
  SP_CHECK(d.n_t == nt_ && d.n_to == nt_, "a Cartesian grid's time axis is n_t = n_to");
  SP_CHECK(d.n_phi == 1 && d.phi_terms, "the Cartesian route's operator is one constant table");
  SP_CHECK(d.n_prefactor == 1 && d.prefactor_terms, "one constant prefactor table");
  SP_CHECK(!d.has_plasma && d.n_raman == 0 && d.absorber_scheme == 0 && !d.spatial_absorber,
           "the Cartesian route is built with the Kerr response alone");
  SP_CHECK(d.has_kerr == 0 || d.kerr_real == 0, "the Cartesian envelope takes the analytic Kerr form");
  SP_CHECK(d.phi_paths[0].kind == 1, "the Cartesian operator's weight is W(z) = z");
  SP_CHECK(d.prefactor_paths[0].kind == 0 && d.density.kind == 0,
           "the Cartesian prefactor and density are constant");
  for (int i = 0; i < d.n_schedule; ++i)
    SP_CHECK(d.schedule_edges[i + 1] > d.schedule_edges[i], "the schedule's edges must increase");
  V_ = static_cast<std::size_t>(nx_) * ny_ * nt_;
  SP_CHECK(V_ <= (std::size_t{1} << 31), "the volume holds at most 2^31 samples");
  edges_.assign(d.schedule_edges, d.schedule_edges + d.n_schedule + 1);
  for (int i = 0; i < d.n_schedule; ++i) steps_.push_back(edges_[i + 1] - edges_[i]);
  const double v = static_cast<double>(V_);
  kerr_self_ = d.has_kerr ? d.kerr_self : 0.0;
  kerr_conj_ = d.has_kerr ? d.kerr_conjugate : 0.0;
  // gamma = the Kerr coefficient times the density times the prefactor's constant weight.
  kerr_scale_ = d.has_kerr ? d.kerr_coefficient * d.density.value * d.prefactor_paths[0].value /
                                 (v * v * v)
                           : 0.0;
  const char* env = std::getenv("SOLPROP_MI355X_GRAPH");
  use_graphs_ = !(env && env[0] == '0');
  // The strided passes' workgroup tile, points (a launch shape only: every sequence's
  // arithmetic is the same whichever is chosen).
  const char* tile = std::getenv("SOLPROP_MI355X_CARTESIAN_TILE");
  const int max_points = tile ? std::atoi(tile) : 4096;
  SP_CHECK(max_points >= 4 && max_points <= 4096,
           "SOLPROP_MI355X_CARTESIAN_TILE is 4 to 4096 points");
  const long rows = static_cast<long>(nx_) * ny_;
  gt_ = axis_log2g(nt_, rows, false, 4096);
  gy_ = axis_log2g(ny_, static_cast<long>(nx_) * nt_, true, max_points);
  try {
    SP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const int C = nx_ * ny_, N = nt_;
    P_.kind = 2;
    P_.C = C;
    P_.N = N;
    P_.T = N;
    P_.NF = N;
    P_.n_phi = 1;
    P_.n_pref = 1;
    P_.Q = 1;
    std::vector<unsigned char> ones(static_cast<std::size_t>(N), 1);
    auto* active = alloc<unsigned char>(N);
    SP_HIP(hipMemcpy(active, ones.data(), N, hipMemcpyHostToDevice));
    P_.active = active;
    auto* nm = alloc<unsigned char>(N);
    SP_HIP(hipMemcpy(nm, d.norm_mask, N, hipMemcpyHostToDevice));
    P_.normmask = nm;
    auto* phi = alloc<cplx>(V_);
    SP_HIP(hipMemcpy(phi, d.phi_terms, V_ * sizeof(cplx), hipMemcpyHostToDevice));
    P_.phi = phi;
    auto* pref = alloc<cplx>(V_);
    SP_HIP(hipMemcpy(pref, d.prefactor_terms, V_ * sizeof(cplx), hipMemcpyHostToDevice));
    P_.pref = pref;
    B_.y = alloc<cplx>(V_);
    B_.lab = alloc<cplx>(V_);
    B_.source = alloc<cplx>(V_);
    B_.advanced = alloc<cplx>(V_);
    field_ = alloc<cplx>(V_);
    acc_ = alloc<cplx>(V_);
    prev_ = alloc<cplx>(V_);
    save_ = alloc<cplx>(V_);
    plan_t_.init(static_cast<std::size_t>(nt_), static_cast<std::size_t>(nx_) * ny_);
    // Columns iy nx + ix: the x axis has stride nt within each of ny blocks, the y axis
    // stride nx nt.
    plan_x_.init_axis(static_cast<std::size_t>(nx_), static_cast<std::size_t>(nt_),
                      static_cast<std::size_t>(ny_), static_cast<std::size_t>(max_points));
    plan_y_.init_axis(static_cast<std::size_t>(ny_), static_cast<std::size_t>(nx_) * nt_, 1);
    rec_dev_ = alloc<AdamsRecord>(3);
    SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&rec_host_), 3 * sizeof(AdamsRecord)));
    std::memset(rec_host_, 0, 3 * sizeof(AdamsRecord));
    partials_ = alloc<double>(ls_partials(P_));
    SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&norms_host_), 4 * sizeof(double),
                         hipHostMallocMapped));
    SP_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&B_.norms), norms_host_, 0));
    SP_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    events_.assign(kMarks, nullptr);
    for (auto& e : events_) SP_HIP(hipEventCreate(&e));
  } catch (...) {
    release();
    throw;
  }
}

CartesianWorkspace::~CartesianWorkspace() { release(); }

void CartesianWorkspace::release() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_) (void)hipGraphExecDestroy(graph_);
  graph_ = nullptr;
  for (void* p : allocations_) (void)hipFree(p);
  allocations_.clear();
  if (rec_host_) (void)hipHostFree(rec_host_);
  rec_host_ = nullptr;
  if (norms_host_) (void)hipHostFree(norms_host_);
  norms_host_ = nullptr;
  if (done_) (void)hipEventDestroy(done_);
  done_ = nullptr;
  for (auto& e : events_)
    if (e) (void)hipEventDestroy(e);
  events_.clear();
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
}

void CartesianWorkspace::mark(int phase) {
  if (!marking_) return;
  const std::size_t k = marked_.size() + 1;
  SP_CHECK(k < events_.size(), "a profiled trial records at most 63 launches");
  SP_HIP(hipEventRecord(events_[k], stream_));
  marked_.push_back(phase);
}

void CartesianWorkspace::collect() {
  for (std::size_t k = 0; k < marked_.size(); ++k) {
    float ms = 0.0f;
    SP_HIP(hipEventElapsedTime(&ms, events_[k], events_[k + 1]));
    phase_ms_[marked_[k]] += ms;
    phase_n_[marked_[k]] += 1;
  }
  float total = 0.0f;
  if (!marked_.empty()) SP_HIP(hipEventElapsedTime(&total, events_[0], events_[marked_.size()]));
  trial_ms_ += total;
  timed_ += 1;
  marked_.clear();
}

std::vector<std::pair<const char*, double>> CartesianWorkspace::phases() const {
  std::vector<std::pair<const char*, double>> out;
  for (int i = 0; i < kPhases; ++i)
    if (phase_n_[i] > 0) out.emplace_back(kPhaseNames[i], phase_ms_[i] * 1e3 / phase_n_[i]);
  return out;
}

void CartesianWorkspace::enqueue_source(const cplx* spectrum) {
  // E = FFT3^-1[Y] over t, x, then y; the response between the y axis's inverse and
  // forward transforms in one pass; then forward over x and t, the prefactor applied as
  // the t rows are stored.
  const long rows = static_cast<long>(nx_) * ny_;
  const int lt = log2_of(nt_), ly = log2_of(ny_);
  const int gt = gt_, gy = gy_;
  const long y_seqs = static_cast<long>(nx_) * nt_;
  const KerrArgs k{kerr_scale_, kerr_self_, kerr_conj_};
  dispatch<TInverse>(lt, gt, static_cast<unsigned>(rows >> gt), stream_, spectrum, field_,
                     plan_t_.roots());
  mark(kPassT);
  plan_x_.exec(field_, nullptr, +1, stream_);
  mark(kPassX);
  dispatch<YResponse>(ly, gy, static_cast<unsigned>(y_seqs >> gy), stream_, field_, y_seqs,
                      plan_y_.roots(), k);
  mark(kPassY);
  plan_x_.exec(field_, nullptr, -1, stream_);
  mark(kPassX);
  dispatch<TForward>(lt, gt, static_cast<unsigned>(rows >> gt), stream_,
                     static_cast<const cplx*>(field_), P_.pref, B_.source, plan_t_.roots());
  mark(kFinish);
}

void CartesianWorkspace::enqueue_step() {
  if (marking_) SP_HIP(hipEventRecord(events_[0], stream_));
  SP_HIP(hipMemcpyAsync(rec_dev_, rec_host_, 3 * sizeof(AdamsRecord), hipMemcpyHostToDevice,
                        stream_));
  mark(kUpload);
  enqueue_source(B_.y);
  launch_ls_pass(1, P_, rec_dev_, B_, acc_, prev_, partials_, stream_);
  mark(kLs1);
  enqueue_source(B_.lab);
  launch_ls_pass(2, P_, rec_dev_, B_, acc_, prev_, partials_, stream_);
  mark(kLs2);
  enqueue_source(B_.lab);
  launch_ls_pass(3, P_, rec_dev_, B_, acc_, prev_, partials_, stream_);
  mark(kLs3);
  enqueue_source(B_.lab);
  launch_ls_pass(4, P_, rec_dev_, B_, acc_, prev_, partials_, stream_);
  mark(kLs4);
  launch_ls_norms(P_, partials_, B_.norms, stream_);
  mark(kNorms);
}

void CartesianWorkspace::wait() {
  SP_HIP(hipEventRecord(done_, stream_));
  wait_event(done_);
}

double CartesianWorkspace::trial(double z, double dz) {
  // The constant operator's increments are the distances: Phi(z_m) - Phi(z) = z_m - z and
  // Phi(z + h) - Phi(z_m) = z + h - z_m, times the one table.
  const double mid = z + 0.5 * dz, end = z + dz;
  rec_host_[0].dz = dz;
  rec_host_[0].x_w[0] = mid - z;
  rec_host_[1].x_w[0] = end - mid;
  if (profile_) {
    marking_ = true;
    marked_.clear();
    try {
      enqueue_step();
    } catch (...) {
      marking_ = false;
      throw;
    }
    marking_ = false;
    wait();
    collect();
  } else if (use_graphs_) {
    if (!graph_) graph_ = capture_graph(stream_, [&] { enqueue_step(); });
    SP_HIP(hipGraphLaunch(graph_, stream_));
    graph_launches_ += 1;
    wait();
  } else {
    enqueue_step();
    wait();
  }
  trials_ += 1;
  evaluations_ += 4;
  const double lower = std::sqrt(norms_host_[0]), upper = std::sqrt(norms_host_[1]);
  stiffness_ = lower > 0.0 ? 2.0 * upper / lower : std::nan("");
  const double squared = norms_host_[2];
  return std::isfinite(squared) ? 0.0 : squared;
}

void CartesianWorkspace::load(double z, const double* field) {
  SP_CHECK(field != nullptr, "load needs a field");
  long cursor = -1;
  for (std::size_t i = 0; i < edges_.size(); ++i)
    if (march::reached(z, edges_[i]) && march::reached(edges_[i], z)) cursor = static_cast<long>(i);
  SP_CHECK(cursor >= 0, "a replay loads its state at one of its schedule's edges");
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(B_.y, field, V_ * sizeof(cplx), hipMemcpyHostToDevice));
  state_ = march::State{};
  state_.z = edges_[static_cast<std::size_t>(cursor)];
  state_.cursor = cursor;
  loaded_ = true;
}

double CartesianWorkspace::probe_trial(double z, double dz) {
  if (!loaded_) throw Error(kState, "load a state before taking a trial");
  SP_HIP(hipMemcpyAsync(save_, B_.y, V_ * sizeof(cplx), hipMemcpyDeviceToDevice, stream_));
  // The transforms' scratch is `save_` only for a four-step time axis (nt > 4096, refused).
  const double norm = trial(z, dz);
  SP_HIP(hipMemcpyAsync(B_.advanced, B_.y, V_ * sizeof(cplx), hipMemcpyDeviceToDevice, stream_));
  SP_HIP(hipMemcpyAsync(B_.y, save_, V_ * sizeof(cplx), hipMemcpyDeviceToDevice, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
  return norm;
}

void CartesianWorkspace::read_trial(double* out) {
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, B_.advanced, V_ * sizeof(cplx), hipMemcpyDeviceToHost));
}

void CartesianWorkspace::read_state(double* out) {
  if (!loaded_) throw Error(kState, "no state is loaded");
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, B_.y, V_ * sizeof(cplx), hipMemcpyDeviceToHost));
}

void CartesianWorkspace::export_state(const double* weight, bool optics, double* out) {
  read_state(out);
  const long n = static_cast<long>(V_);
  for (long i = 0; i < n; ++i) {
    double re = out[2 * i], im = out[2 * i + 1];
    if (weight) {
      const double s = std::sqrt(weight[i]);
      re = re * s;
      im = im * s;
    }
    if (optics) {
      im = -im;
      if ((i % nt_) & 1) {
        re = -re;
        im = -im;
      }
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
  }
}

bool CartesianWorkspace::advance_bounded(const Bounds& b, int columns, std::vector<double>& rows,
                                         sp_march_result* r) {
  if (!loaded_) throw Error(kState, "load a state before advancing");
  ReplayEngine engine;
  engine.step = [this](double z, double dz) { return trial(z, dz); };
  engine.commit = [this](double) { accepted_ += 1; };
  AttemptLog log;
  log.begin(&rows, columns, b);
  log.order = 4;
  ReplayHost host(log);
  const double target = std::min(b.target, edges_.back());
  bool done = false;
  try {
    done = march::replay_to(engine, host, state_, steps_, edges_, target) &&
           march::reached(state_.z, target);
  } catch (...) {
    r->z = state_.z;
    r->attempts = log.attempts;
    r->accepted = log.accepted;
    r->evaluations = evaluations_;
    throw;
  }
  r->z = state_.z;
  r->proposal = state_.cursor < static_cast<long>(steps_.size())
                    ? steps_[static_cast<std::size_t>(state_.cursor)]
                    : 0.0;
  r->done = done ? 1 : 0;
  r->attempts = log.attempts;
  r->accepted = log.accepted;
  r->evaluations = evaluations_;
  return done;
}

sp_counters CartesianWorkspace::counters() const {
  sp_counters c{};
  c.trials = trials_;
  c.evaluations = evaluations_;
  c.accepted = accepted_;
  c.graph_launches = graph_launches_;
  c.device_bytes = static_cast<long>(device_bytes_ + plan_t_.bytes() + plan_y_.bytes() +
                                     plan_x_.bytes());
  return c;
}

std::vector<double> CartesianWorkspace::statistics(std::uint64_t, const double*, int,
                                                   const StatExtras&) {
  throw std::invalid_argument("plane statistics are not built on the Cartesian route");
}

void CartesianWorkspace::bind_monitor(const double*, const double*) {
  throw std::invalid_argument("the boundary monitor is not built on the Cartesian route");
}

void CartesianWorkspace::monitor_reading(double*) {
  throw std::invalid_argument("the boundary monitor is not built on the Cartesian route");
}

void CartesianWorkspace::source(double z, const double* lab, double* out, double* streams,
                                double* modal) {
  (void)z;
  SP_CHECK(lab && out && !streams && !modal, "the Cartesian source takes a spectrum and an output");
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(B_.lab, lab, V_ * sizeof(cplx), hipMemcpyHostToDevice));
  enqueue_source(B_.lab);
  SP_HIP(hipStreamSynchronize(stream_));
  SP_HIP(hipMemcpy(out, B_.source, V_ * sizeof(cplx), hipMemcpyDeviceToHost));
}

long CartesianWorkspace::buffer(const char* name, double* out, long count) {
  SP_CHECK(name != nullptr, "a buffer name is required");
  if (std::strcmp(name, "stiffness") == 0) {
    if (out && count > 0) out[0] = stiffness_;
    return 1;
  }
  throw std::invalid_argument(std::string("no buffer named ") + name + " on the Cartesian route");
}

}  // namespace sp
