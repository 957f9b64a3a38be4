// Batched FP64 FFTs for gfx950: Stockham passes staged in LDS (radix-16 or radix-4/2 stages).
//
// A transform of n = 2^m points is one LDS-resident pass when n <= 4096 and the
// four-step n = n1 * n2 split (column pass with the n-point twiddle, then a row pass
// with a transposed store) otherwise. Every
// pass moves whole sequences through LDS with coalesced HBM reads and writes: a
// workgroup owns G neighbouring sequences, so a strided (column) pass still reads G
// contiguous complex doubles per row. Sequences of 1024 points or more take radix-16
// stages whose first reads HBM and whose last writes it (fft_lds.hpp).
//
// Transforms are unnormalised in both directions (sign -1 forward, +1 inverse), as the
// reference's native plans are; every scale (1/n, pad and crop) is folded into the
// kernels around them.
#pragma once

#include <vector>

#include "common.hpp"

namespace sp {

// One LDS pass: per batch row b, per sequence s, transform the L points
// in[b*batch + j*in_sj + s*in_ss] (j < L) and store point k at
// out[b*batch_out + k*out_sk + s*out_ss], optionally times W_N^{s*k}.
struct FftPass {
  int log2L = 0;
  int log2G = 0;            // log2 of the sequences a workgroup owns
  long in_sj = 1, in_ss = 0, out_sk = 1, out_ss = 0;
  long batch_in = 0, batch_out = 0;
  int n_seq = 1;            // sequences per batch row
  bool global_twiddle = false;
  int log2N = 0;            // length of the global twiddle table
};

// The strides and scalings of one pass's launch (what the pass kernel reads). When
// `spec` is set, the first pass loads its input from a half spectrum instead: row r's
// point k is c2r_pre(X[k], X[M - k], W_2M^k) of X = spec row r (spec_n bins, zero beyond)
// scaled by spec_scale, DC and Nyquist taken real (the packing of a real inverse
// transform of 2M points as a complex one of M).
struct FftPassArgs {
  long in_sj, in_ss, out_sk, out_ss, batch_in, batch_out;
  int global_twiddle;
  int mask_n;  // N - 1 for the global twiddle index
  const cplx* spec;
  const cplx* spec_roots;  // W_2M^k, k < M
  int spec_n;
  double spec_scale;
  long spec_m;             // M, the points of a row
};

// One pass of one transform in a grouped launch, which runs the same pass shape for
// many independent transforms (one per instance of a batched march): the pass's
// operands and strides, its workgroups per batch row and its rows. The job runs when
// `*gate` is non-zero (a null gate always runs).
struct FftJob {
  const cplx* in;
  cplx* out;
  const cplx* tw_l;
  const cplx* tw_n;
  FftPassArgs a;
  int groups;
  int rows;
  const int* gate;
};

// A half spectrum for FftPlan::exec_packed: `n_bins` bins a row (rows contiguous),
// scaled by `scale`, with the unit roots W_2M^k of the real transform of 2M points.
struct PackedSpectrum {
  const cplx* spectrum;
  const cplx* roots;
  long n_bins;
  double scale;
};

class FftPlan {
 public:
  FftPlan() = default;
  ~FftPlan();
  FftPlan(const FftPlan&) = delete;
  FftPlan& operator=(const FftPlan&) = delete;

  // A complex transform of n points over `batch` contiguous rows of n.
  void init(std::size_t n, std::size_t batch);
  // The same plan with its twiddle tables in device memory the caller owns, at least
  // table_points(n) points at `tables`: nothing is allocated or freed. The tables hold
  // nothing until upload_tables has run (on `stream`, which it then synchronises).
  void init_external(std::size_t n, std::size_t batch, cplx* tables);
  // A transform along one axis of a row-major array (outer, n, inner): `outer` blocks of
  // n points each `inner` apart, every one of the `inner` sequences of a block
  // transformed (n <= 4096, one LDS pass; in place through exec). A workgroup owns up to
  // 16 neighbouring sequences, at most `max_points` points in all.
  void init_axis(std::size_t n, std::size_t inner, std::size_t outer,
                 std::size_t max_points = 4096);
  void upload_tables(hipStream_t stream) const;
  static std::size_t table_points(std::size_t n);
  // In place on data (rows of n); `scratch` holds n*batch points for the four-step.
  void exec(cplx* data, cplx* scratch, int sign, hipStream_t stream) const;
  // The same transform with its input packed from a half spectrum on the way in (the
  // first pass reads the spectrum; `data` need not hold anything), the result in `data`.
  void exec_packed(const PackedSpectrum& spectrum, cplx* data, cplx* scratch, int sign,
                   hipStream_t stream) const;
  // The passes exec (packed == null) or exec_packed would launch on these buffers, in
  // order, as grouped-launch jobs, and each pass's shape (log2 of its length and of the
  // sequences a workgroup owns).
  std::vector<FftJob> jobs(cplx* data, cplx* scratch, const PackedSpectrum* packed,
                           const int* gate) const;
  int passes() const { return static_cast<int>(passes_.size()); }
  int pass_log2L(int i) const { return passes_[static_cast<std::size_t>(i)].log2L; }
  int pass_log2G(int i) const { return passes_[static_cast<std::size_t>(i)].log2G; }
  std::size_t n() const { return n_; }
  std::size_t batch() const { return batch_; }
  bool two_pass() const { return passes_.size() == 2; }
  // Unit roots W_n^t = exp(-2 pi i t / n), t < n (device table), for pre/post passes.
  const cplx* roots() const { return roots_n_; }
  std::size_t bytes() const { return bytes_; }

 private:
  void layout(std::size_t n, std::size_t batch);
  void launch(const FftPass& p, const cplx* in, cplx* out, int sign, hipStream_t s) const;
  std::size_t n_ = 0, batch_ = 0, bytes_ = 0;
  bool owns_tables_ = true;
  std::vector<FftPass> passes_;
  mutable const PackedSpectrum* spec_ = nullptr;  // set for the first pass of exec_packed
  cplx* roots_n_ = nullptr;   // length n
  cplx* roots_l_[2] = {nullptr, nullptr};  // per-pass local tables, length L
};

// One grouped launch of a pass shape over `n_jobs` jobs (device array `jobs`): job i
// owns workgroups first[i] .. first[i+1] - 1 (device array, first[n_jobs] = `blocks`).
void launch_fft_grouped(int log2L, int log2G, const FftJob* jobs, const int* first, int n_jobs,
                        int blocks, int sign, hipStream_t stream);
// Whether a grouped pass kernel exists for this shape.
bool fft_grouped_available(int log2L, int log2G);

// Host-computed unit roots exp(-2 pi i t / n), t < n, correctly rounded through
// extended precision and the octant symmetries.
std::vector<cplx> unit_roots(std::size_t n);

}  // namespace sp
