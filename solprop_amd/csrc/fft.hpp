// Batched FP64 FFTs for gfx950: Stockham radix-4/2 passes staged in LDS.
//
// A transform of n = 2^m points is one LDS-resident pass when n <= 4096 and the
// four-step n = n1 * n2 split (column pass with the n-point twiddle, then a row pass
// with a transposed store) otherwise. Every pass moves whole sequences through LDS
// with coalesced HBM reads and writes: a workgroup owns G neighbouring sequences, so a
// strided (column) pass still reads G contiguous complex doubles per row.
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
  // In place on data (rows of n); `scratch` holds n*batch points for the four-step.
  void exec(cplx* data, cplx* scratch, int sign, hipStream_t stream) const;
  // The same transform with its input packed from a half spectrum on the way in (the
  // first pass reads the spectrum; `data` need not hold anything), the result in `data`.
  void exec_packed(const PackedSpectrum& spectrum, cplx* data, cplx* scratch, int sign,
                   hipStream_t stream) const;
  std::size_t n() const { return n_; }
  std::size_t batch() const { return batch_; }
  bool two_pass() const { return passes_.size() == 2; }
  // Unit roots W_n^t = exp(-2 pi i t / n), t < n (device table), for pre/post passes.
  const cplx* roots() const { return roots_n_; }
  std::size_t bytes() const { return bytes_; }

 private:
  void launch(const FftPass& p, const cplx* in, cplx* out, int sign, hipStream_t s) const;
  std::size_t n_ = 0, batch_ = 0, bytes_ = 0;
  std::vector<FftPass> passes_;
  mutable const PackedSpectrum* spec_ = nullptr;  // set for the first pass of exec_packed
  cplx* roots_n_ = nullptr;   // length n
  cplx* roots_l_[2] = {nullptr, nullptr};  // per-pass local tables, length L
};

// Host-computed unit roots exp(-2 pi i t / n), t < n, correctly rounded through
// extended precision and the octant symmetries.
std::vector<cplx> unit_roots(std::size_t n);

}  // namespace sp
