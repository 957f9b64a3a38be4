// The modal field-grid route's kernels for gfx950, FP64 (see field.hpp).
#include "field.hpp"

#include <hip/hip_ext.h>

#include <cmath>

namespace sp {

namespace {

constexpr int kThreads = 256;
constexpr double kRecipFactorial[10] = {1.0,          1.0,          0.5,
                                        1.0 / 6.0,    1.0 / 24.0,   1.0 / 120.0,
                                        1.0 / 720.0,  1.0 / 5040.0, 1.0 / 40320.0,
                                        1.0 / 362880.0};
constexpr double kFactorial[9] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320};
// 1/n, rounded to nearest, for the series' divisors n = m + order <= 30.
constexpr double kRecipInteger[32] = {
    0.0,        1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,  1.0 / 6,  1.0 / 7,
    1.0 / 8,    1.0 / 9,    1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15,
    1.0 / 16,   1.0 / 17,   1.0 / 18, 1.0 / 19, 1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23,
    1.0 / 24,   1.0 / 25,   1.0 / 26, 1.0 / 27, 1.0 / 28, 1.0 / 29, 1.0 / 30, 1.0 / 31};

// Division by a fixed complex b by Smith's algorithm, as NumPy divides complex numbers:
// the ratio and scale depend on b alone, so a divisor used several times is prepared once.
struct Divisor {
  double rat, scl;
  bool real_major;  // |b.x| >= |b.y|
};

__device__ inline Divisor divisor(cplx b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double rat = b.y / b.x;
    return {rat, 1.0 / (b.x + b.y * rat), true};
  }
  const double rat = b.x / b.y;
  return {rat, 1.0 / (b.y + b.x * rat), false};
}

__device__ inline cplx divide(cplx a, const Divisor& d) {
  if (d.real_major) return make((a.x + a.y * d.rat) * d.scl, (a.y - a.x * d.rat) * d.scl);
  return make((a.x * d.rat + a.y) * d.scl, (a.y * d.rat - a.x) * d.scl);
}

__device__ inline cplx cdiv(cplx a, cplx b) { return divide(a, divisor(b)); }

// phi_0..phi_kOrder at one argument (src/solprop/backends/exponential.py:131-226):
// scaled Taylor series, downward recurrence and doubling near zero; upward recurrence
// from exp(x) beyond |x| = 3. The order is a template argument, so that every loop has
// a static bound and `out` stays in registers.
template <int kOrder>
__device__ void phi_functions(cplx x, cplx (&out)[kOrder + 1]) {
  if (x.x == 0.0 && x.y == 0.0) {
    // The series at zero, exactly: phi_k(0) = 1/k! (what the Taylor path rounds to).
#pragma unroll
    for (int k = 0; k <= kOrder; ++k) out[k] = make(kRecipFactorial[k], 0.0);
    return;
  }
  const double magnitude = hypot(x.x, x.y);
  if (magnitude > 3.0) {
    out[0] = cexp(x);
    const Divisor by_x = divisor(x);
#pragma unroll
    for (int k = 1; k <= kOrder; ++k)
      out[k] = divide(sub(out[k - 1], make(kRecipFactorial[k - 1], 0.0)), by_x);
    return;
  }
  int squarings = 0;
  if (magnitude > 0.5) squarings = static_cast<int>(ceil(log2(fmax(magnitude, 0.5) / 0.5)));
  // Scaling by powers of two is exact: ldexp gives the quotient bit for bit.
  const cplx y = make(ldexp(x.x, -squarings), ldexp(x.y, -squarings));
  cplx term = make(kRecipFactorial[kOrder], 0.0);
  cplx total = term;
  // term_m = term_{m-1} y / (m + kOrder), the division as a product with the rounded
  // reciprocal (within one ulp of the quotient per term; no division in the loop).
#pragma unroll
  for (int m = 1; m < 24; ++m) {
    const cplx ty = mul(term, y);
    const double r = kRecipInteger[m + kOrder];
    term = make(ty.x * r, ty.y * r);
    total = add(total, term);
  }
  out[kOrder] = total;
#pragma unroll
  for (int k = kOrder - 1; k >= 1; --k)
    out[k] = add(make(kRecipFactorial[k], 0.0), mul(y, out[k + 1]));
  out[0] = cexp(y);
  for (int level = 0; level < squarings; ++level) {
    cplx doubled[kOrder + 1];
    doubled[0] = mul(out[0], out[0]);
#pragma unroll
    for (int k = 1; k <= kOrder; ++k) {
      cplx t = mul(out[0], out[k]);
#pragma unroll
      for (int j = 1; j <= k; ++j) t = add(t, scale(kRecipFactorial[k - j], out[j]));
      doubled[k] = make(ldexp(t.x, -k), ldexp(t.y, -k));
    }
#pragma unroll
    for (int k = 0; k <= kOrder; ++k) out[k] = doubled[k];
  }
}

__device__ inline cplx weighted(const double* w, const cplx* terms, int n, long stride,
                                long at) {
  cplx total = make(0.0, 0.0);
  for (int t = 0; t < n; ++t) total = add(total, scale(w[t], terms[t * stride + at]));
  return total;
}

// dz Σ_k M_k (Σ_j c_jk X_j) over a W x W table (coefficients [j][k], zero where the
// stencil is narrower), in increasing k and j.
template <int W>
__device__ inline cplx apply_moments(const cplx (&M)[W], const double (*c)[kMaxAdamsWidth],
                                     const cplx (&X)[W], double dz) {
  cplx total = make(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    cplx combined = make(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < W; ++j) combined = add(combined, scale(c[j][k], X[j]));
    total = add(total, mul(M[k], combined));
  }
  return scale(dz, total);
}

__device__ inline cplx r2c_post(cplx za, cplx zb, cplx w) {
  const cplx e = scale(0.5, add(za, conj(zb)));
  const cplx o = scale(0.5, rot(sub(za, conj(zb)), -1));
  return add(e, mul(w, o));
}


// The rate table (src/solprop/translate/ionisation.py:242-313): exactly zero below
// field_min, held above field_max, exp(cubic in log|E|) between. Evaluated without
// contraction, in the table's own order of roundings: the rate's log-slope (tens to
// hundreds near threshold) multiplies any rounding of the cubic.
__device__ inline double rate_of(double magnitude, const PlasmaTable& pt, const double* c) {
#pragma clang fp contract(off)
  if (magnitude > pt.field_max) return pt.rate_max;
  if (!(magnitude >= pt.field_min)) return 0.0;
  const double lf = log(magnitude);
  int index = static_cast<int>((lf - pt.log_field_min) / pt.d_log_field);
  index = index < 0 ? 0 : (index > pt.n_intervals - 1 ? pt.n_intervals - 1 : index);
  const double local = lf - (pt.log_field_min + index * pt.d_log_field);
  const int n = pt.n_intervals;
  const double v = ((c[index] * local + c[n + index]) * local + c[2 * n + index]) * local +
                   c[3 * n + index];
  return exp(v);
}

// The ionised fraction after an exposure x from f0, 1 - (1 - f0) e^{-x}, without
// contraction (src/solprop/backends/numpy_ref/passes/plasma.py:104-191).
__device__ inline double fraction_of(double f0, double x) {
#pragma clang fp contract(off)
  return 1.0 - (1.0 - f0) * exp(-x);
}

// Cross-lane moves of a double by DPP (two 32-bit halves, the same control), lanes with
// no source or outside `kRowMask` reading zero. Controls (GFX9 DPP): row_shr:n is
// 0x110 + n, row_bcast:15 is 0x142, row_bcast:31 is 0x143, wave_shl:1 is 0x130 (lane i
// reads lane i + 1), wave_shr:1 is 0x138 (lane i reads lane i - 1).
template <int kCtrl, int kRowMask, bool kBound>
__device__ inline double dpp_move(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_update_dpp(0, lo, kCtrl, kRowMask, 0xf, kBound);
  const int rhi = __builtin_amdgcn_update_dpp(0, hi, kCtrl, kRowMask, 0xf, kBound);
  return __hiloint2double(rhi, rlo);
}

// Inclusive scan across one wavefront in a fixed order: Hillis-Steele within each row of
// 16 lanes, then row 0's total into row 1 and row 2's into row 3, then rows 0-1's total
// into rows 2 and 3. The same on every run.
__device__ inline double wave_scan(double v) {
  v = v + dpp_move<0x111, 0xf, true>(v);
  v = v + dpp_move<0x112, 0xf, true>(v);
  v = v + dpp_move<0x114, 0xf, true>(v);
  v = v + dpp_move<0x118, 0xf, true>(v);
  v = v + dpp_move<0x142, 0xa, false>(v);
  v = v + dpp_move<0x143, 0xc, false>(v);
  return v;
}

// The value of lane + 1 (zero in the last lane) and of lane - 1 (zero in the first).
__device__ inline double from_next_lane(double v) { return dpp_move<0x130, 0xf, true>(v); }
__device__ inline double from_previous_lane(double v) { return dpp_move<0x138, 0xf, true>(v); }

// Registers a lane holds for its sample's modal fields (one unused for the dense route).
constexpr int held(int C) { return C > 0 ? C : 1; }

// The field at a node from the modal fields of this lane's sample, Σ_m S[m] a_m, summed
// in mode order.
template <int C>
__device__ inline double synthesise_at(const double* S, const double (&am)[held(C)]) {
  double e = 0.0;
#pragma unroll
  for (int m = 0; m < C; ++m) e = e + S[m] * am[m];
  return e;
}

// The modal fields (C, T) real at sample t, zero past the grid's end.
template <int C>
__device__ inline void load_modal(const FieldBuffers& B, int T, int t, double (&am)[held(C)]) {
  const double* a = reinterpret_cast<const double*>(B.modal_t);
#pragma unroll
  for (int m = 0; m < C; ++m) am[m] = t < T ? a[static_cast<long>(m) * T + t] : 0.0;
}

// The field at node q and this lane's sample t: synthesised from the lane's modal
// fields, or on the dense route (C = 0) read from the node field. Zero past the end.
template <int C>
__device__ inline double field_at(const DeviceProblem& P, const FieldBuffers& B,
                                  const double (&am)[held(C)], int q, int t) {
  if constexpr (C > 0) {
    return synthesise_at<C>(P.synthesis + static_cast<long>(q) * C, am);
  } else {
    return t < P.T ? B.node_field[static_cast<long>(q) * P.T + t] : 0.0;
  }
}

// One node's exposure and free current over one scan tile (see field.hpp): the rate
// read back, the exposure increments' scan plus the tile's offset, the fraction, the
// free-current increments' scan. Lanes past the grid's end hold zeros.
struct TileColumn {
  double w, f;         // rate and fraction at this lane's sample
  double j_incl;       // inclusive free-current scan within the tile (lane 62: its total)
  double j_excl;       // exclusive free-current scan within the tile
};

__device__ inline TileColumn tile_column(const DeviceProblem& P, double e, const double* w_row,
                                         double x_off, double f0, int t, int lane) {
  const int T = P.T;
  const double dt = P.dt_fine;
  TileColumn c;
  c.w = t < T ? w_row[t] : 0.0;
  const bool step = lane < kTileSpan && t + 1 < T;  // this lane's increment reaches t + 1
  const double w_next = from_next_lane(c.w);
  const double inc = step ? (dt * (w_next + c.w)) / 2.0 : 0.0;
  const double excl = from_previous_lane(wave_scan(inc));
  c.f = fraction_of(f0, x_off + excl);
  const double g = (c.f * P.plasma_cmr) * e;
  const double g_next = from_next_lane(g);
  const double jinc = step ? (dt * (g_next + g)) / 2.0 : 0.0;
  c.j_incl = wave_scan(jinc);
  c.j_excl = from_previous_lane(c.j_incl);
  return c;
}
}  // namespace

// The step's operator at one bin: x = Σ_t w_t Φ_t in term order, the moments
// M_k = k! φ_{k+1}(x) and the flow exp(x) (exactly 1 where x = 0: inactive bins skip
// the transcendental). The predictor and the corrector both form them from the Φ terms,
// so neither array is stored.
template <int W>
__device__ inline cplx step_operator(const DeviceProblem& P, const AdamsRecord& r,
                                     const cplx (&phi_t)[kMaxTerms], cplx (&M)[W],
                                     cplx& flow) {
  cplx x = make(0.0, 0.0);
#pragma unroll
  for (int t = 0; t < kMaxTerms; ++t)
    if (t < P.n_phi) x = add(x, scale(r.x_w[t], phi_t[t]));
  cplx phi[W + 1];
  phi_functions<W>(x, phi);
#pragma unroll
  for (int k = 0; k < W; ++k) M[k] = scale(kFactorial[k], phi[k + 1]);
  const bool still = x.x == 0.0 && x.y == 0.0;
  flow = still ? make(1.0, 0.0) : cexp(x);
  return x;
}

// ------------------------------------------------------------- Adams predictor

// The predictor for a stencil of width W (moments M_0..M_{W-1}, W - 1 stored rows). One
// wavefront per workgroup; every global operand is loaded before any arithmetic.
// Coefficients past the stencil's order are zero in the record.
template <int W, bool Commit = false>
__device__ __forceinline__ void predict_body(const DeviceProblem& P, const AdamsRecord* __restrict__ R,
                                    const FieldBuffers& B, long idx) {
  const long CN = static_cast<long>(P.C) * P.N;
  if (idx >= CN) return;
  const AdamsRecord& r = *R;
  const double dz = r.dz;
  const bool compensated = r.compensated != 0;
  // Loads.
  cplx phi_t[kMaxTerms], F[W], U[W];
#pragma unroll
  for (int t = 0; t < kMaxTerms; ++t) phi_t[t] = t < P.n_phi ? P.phi[t * CN + idx] : make(0.0, 0.0);
  cplx y;
  if constexpr (Commit) {
    // The accepted step's commit for this bin (the commit below, element for element):
    // its source row and state enter the store at the free slot and its advanced field
    // becomes y. The stencil's loads below follow the stores in this thread, so they
    // read the committed rows.
    const int slot = r.free_slot;
    B.store[slot * CN + idx] = B.source[idx];
    if (compensated) B.states[slot * CN + idx] = B.lab[idx];
    y = P.active[idx % P.N] ? B.advanced[idx] : make(0.0, 0.0);
    B.y[idx] = y;
  } else {
    y = B.y[idx];
  }
#pragma unroll
  for (int j = 0; j < W; ++j) {
    F[j] = j < W - 1 ? B.store[r.slots[j] * CN + idx] : make(0.0, 0.0);
    U[j] = compensated && j < W - 1 ? B.states[r.slots[j] * CN + idx] : make(0.0, 0.0);
  }
  cplx M[W], flow;
  const cplx x = step_operator<W>(P, r, phi_t, M, flow);
  cplx lab = add(mul(flow, y), apply_moments<W>(M, r.pred_c, F, dz));
  if (compensated) {
    const cplx reference = make(x.x / dz, x.y / dz);
    cplx total = make(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < W - 1; ++j) {
      cplx rate = make(0.0, 0.0);
#pragma unroll
      for (int t = 0; t < kMaxTerms; ++t)
        if (t < P.n_phi) rate = add(rate, scale(r.slot_rate_w[j][t], phi_t[t]));
      const cplx term = mul(sub(rate, reference), U[j]);
      // A zero term (a bin with no field) carries zero whatever it is transported by.
      const bool empty = term.x == 0.0 && term.y == 0.0;
      const cplx span = scale(1.0 - r.nodes[j], x);
      const cplx carried = empty ? make(0.0, 0.0) : mul(cexp(span), term);
      B.carried[j * CN + idx] = carried;
      total = add(total, scale(r.comp_pred_w[0][j], carried));
    }
    lab = add(lab, scale(dz, total));
  }
  B.lab[idx] = lab;
}

template <int W, bool Commit>
__global__ void __launch_bounds__(kWave)
    adams_predict(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  predict_body<W, Commit>(P, R, B, static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------- the source's transforms

// The lab-frame modal spectrum (1/T folded in) packed for the real inverse transform
// of T points as a complex one of T/2.
// With Hilbert set, the spectrum of the Hilbert transform instead: -i Y strictly
// between DC and Nyquist, zero at both (fourier.py:243-300 keeps those two bins in the
// analytic signal's real part only).
template <bool Hilbert>
__global__ void field_scatter(DeviceProblem P, const cplx* spectrum, cplx* modal_t,
                              const cplx* roots_T) {
  const int half = P.T / 2;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.C) * half) return;
  const int c = static_cast<int>(idx / half), k = static_cast<int>(idx % half);
  const long row = static_cast<long>(c) * P.N;
  const int kb = half - k;
  cplx xa = k < P.N ? scale(P.inv_T, spectrum[row + k]) : make(0.0, 0.0);
  cplx xb = kb < P.N ? scale(P.inv_T, spectrum[row + kb]) : make(0.0, 0.0);
  // A real transform's zero-frequency and Nyquist bins are real: their imaginary parts
  // are discarded, as a c2r transform does (numpy.fft.irfft, FFTW, rocFFT).
  if (k == 0) {
    xa.y = 0.0;
    xb.y = 0.0;
  }
  if constexpr (Hilbert) {
    xa = k == 0 ? make(0.0, 0.0) : rot(xa, -1);
    xb = k == 0 ? make(0.0, 0.0) : rot(xb, -1);
  }
  modal_t[idx] = c2r_pre(xa, xb, roots_T[k]);
}

// Pass A: the rate at every node and sample, stored, and the exposure increments' sum
// over each tile. A workgroup takes one tile and kNodeWaves * kNodesPerWave nodes; each
// wave runs its nodes with the tile's modal fields in registers and the rate table in
// LDS (read from global memory above kLdsTableIntervals intervals).
template <int C>
__device__ __forceinline__ void rate_tiles_body(const DeviceProblem& P, const PlasmaTable& pt,
                                       const FieldBuffers& B, int tile, int group,
                                       double* table) {
  const bool in_lds = pt.n_intervals <= kLdsTableIntervals;
  if (in_lds)
    for (int i = threadIdx.x; i < 4 * pt.n_intervals; i += blockDim.x) table[i] = B.rate_c[i];
  __syncthreads();
  const double* coefficients = in_lds ? table : B.rate_c;
  const int T = P.T, tiles = scan_tiles(T);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int t = tile * kTileSpan + lane;
  const bool inside = t < T;
  // The shared last sample is stored by the next tile's first lane.
  const bool stores = inside && (lane < kTileSpan || t == T - 1);
  const bool step = lane < kTileSpan && t + 1 < T;
  double am[held(C)];
  load_modal<C>(B, T, t, am);
  const int q0 = group * kNodeWaves * kNodesPerWave;
#pragma unroll
  for (int i = 0; i < kNodesPerWave; ++i) {
    const int q = q0 + i * kNodeWaves + wave;
    if (q < P.Q) {  // uniform across the wave
      const double e = field_at<C>(P, B, am, q, t);
      const double w = inside ? rate_of(fabs(e), pt, coefficients) : 0.0;
      if (stores) B.rate[static_cast<long>(q) * T + t] = w;
      const double w_next = from_next_lane(w);
      const double inc = step ? (P.dt_fine * (w_next + w)) / 2.0 : 0.0;
      const double incl = wave_scan(inc);
      // The tile's total is lane 62's inclusive sum (lane 63 adds no increment): the next
      // tile's first sample then sits at exactly this tile's offset plus that sum.
      if (lane == kTileSpan - 1) B.tile_x[static_cast<long>(q) * tiles + tile] = incl;
    }
  }
}

template <int C>
__global__ void __launch_bounds__(kWave * kNodeWaves)
    plasma_rate_tiles(DeviceProblem P, PlasmaTable pt, FieldBuffers B) {
  __shared__ double table[4 * kLdsTableIntervals];
  rate_tiles_body<C>(P, pt, B, blockIdx.x, blockIdx.y, table);
}

// ---------------------------------------------- band running-integral rule

// Samples a wave's window holds: its tile and a halo of K each side.
constexpr int kBandSpan = kWave + 2 * kMaxBandHalf;

// The band rule's filter at a lane of the wave's window (window[j] is sample
// 63 tile - K + j): c_0 y_n + Σ_k c_k (y_{n-k} + y_{n+k}), the centre first and each
// pair added before it is weighted (src/solprop/nonlinear/ionisation.py:611-653).
__device__ inline double band_filter(const PlasmaTable& pt, const double* window, int lane) {
  const int K = pt.n_taps - 1;
  const double* y = window + K + lane;
  double f = pt.taps[0] * y[0];
#pragma unroll 2
  for (int k = 1; k <= K; ++k) f = f + pt.taps[k] * (y[-k] + y[k]);
  return f;
}

// The rate at node q and sample t of the grid.
template <int C>
__device__ inline double rate_sample(const DeviceProblem& P, const PlasmaTable& pt,
                                     const double* coefficients, const FieldBuffers& B, int q,
                                     int t) {
  double am[held(C)];
  load_modal<C>(B, P.T, t, am);
  return rate_of(fabs(field_at<C>(P, B, am, q, t)), pt, coefficients);
}

// The free-current integrand (f e²/m) E at node q and sample t, the fraction from the
// floored exposure (the owning tile's offset plus the within-tile exposure).
template <int C>
__device__ inline double current_sample(const DeviceProblem& P, const FieldBuffers& B, int q,
                                        int t, double f0, int tiles) {
  const int owner = min(t / kTileSpan, tiles - 1);
  const double x = B.tile_x[static_cast<long>(q) * tiles + owner] +
                   B.x_local[static_cast<long>(q) * P.T + t];
  const double f = fraction_of(f0, x > 0.0 ? x : 0.0);
  double am[held(C)];
  load_modal<C>(B, P.T, t, am);
  return (f * P.plasma_cmr) * field_at<C>(P, B, am, q, t);
}

// A zero in device memory that nothing writes: loads a flag redirects read it.
__device__ double kZeroSample = 0.0;

// The same where the rate pass ran sparse: a flagged tile's within-tile exposure is zero.
template <int C>
__device__ inline double current_sample_sparse(const DeviceProblem& P, const FieldBuffers& B,
                                               int q, int t, double f0, int tiles) {
  const int owner = min(t / kTileSpan, tiles - 1);
  const long at = static_cast<long>(q) * tiles + owner;
  const double* local = B.zero_tiles[at] ? &kZeroSample : B.x_local + static_cast<long>(q) * P.T + t;
  const double x = B.tile_x[at] + *local;
  const double f = fraction_of(f0, x > 0.0 ? x : 0.0);
  double am[held(C)];
  load_modal<C>(B, P.T, t, am);
  return (f * P.plasma_cmr) * field_at<C>(P, B, am, q, t);
}

// Point reflection past the grid's ends: y(-m) = 2 y(0) - y(m), y(T-1+m) = 2 y(T-1) -
// y(T-1-m) (ionisation.py:642-644). Samples past T-1+K lie outside every owned
// sample's stencil and read as zero.
template <class Sample>
__device__ inline double reflected(int T, int K, int t, const Sample& sample) {
  if (t > T - 1 + K) return 0.0;
  if (t < 0) return 2.0 * sample(0) - sample(-t);
  if (t > T - 1) return 2.0 * sample(T - 1) - sample(2 * (T - 1) - t);
  return sample(t);
}

// A tile's halo of 2K samples: slot h < K is sample 63 tile - K + h, left of the tile's
// 64 lanes; slot h >= K is sample 63 tile + 64 + (h - K), right of them.
__device__ inline int halo_sample(int tile, int K, int h) {
  return h < K ? tile * kTileSpan - K + h : tile * kTileSpan + kWave + (h - K);
}

// Where halo slot h sits in the window (window[j] is sample 63 tile - K + j).
__device__ inline int halo_slot(int K, int h) { return h < K ? h : kWave + h; }

// Orders a wave's LDS writes before its lanes read each other's values: the LDS serves
// one wave's instructions in issue order, so only the compiler has to be held back.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether a node's whole filter window is zero: every lane's own value and the node's 2K
// halo values (wave-uniform). With a positive centre tap the filter of such a window is
// +0.0 at every lane, and so are the trapezoid increments and their scans, so a pass
// may store those zeros without forming them (the same bits): away from the pulse the
// rate vanishes, and before it the free-current integrand.
__device__ inline bool window_is_zero(const PlasmaTable& pt, double own,
                                      const double* halo_row, int lane) {
  const double h = lane < 2 * (pt.n_taps - 1) ? halo_row[lane] : 0.0;
  return pt.taps[0] > 0.0 && __ballot(own != 0.0 || h != 0.0) == 0;
}

// The band passes run each wave's nodes as the trapezoid passes do (modal fields of the
// lane's sample held in registers across them), with the samples a node needs beyond
// its own lanes evaluated first for every node of the wave: the filter's halo of 2K
// samples and, in the last tile, the K reflections past the grid's end that the tile's
// own lanes hold (slot 2K + r is sample T + r). Keeping those evaluations out of the
// node loop keeps its registers few.
constexpr int kBandSlots = 3 * kMaxBandHalf;

// Slot h of a node's extra samples in a tile with 2K halo slots.
__device__ inline int extra_sample(int tile, int K, int T, int h) {
  return h < 2 * K ? halo_sample(tile, K, h) : T + (h - 2 * K);
}

// Up to four modes the band passes ask for six (rate) and eight (current) waves a SIMD:
// at most 80 and 64 VGPRs, with no scratch; more modes hold more registers and ask for
// fewer waves, again with no scratch.

// Pass A under the band rule: the raw rate stored for the owned samples, the filtered
// rate's trapezoid increments scanned, the within-tile exposure stored and the tile's
// sum.
// LDS a band pass's workgroup holds: each wave's window and its nodes' extra samples.
template <int NPW>
struct BandLdsN {
  double windows[kNodeWaves][kBandSpan];
  double halos[kNodeWaves][NPW * kBandSlots];
};
using BandLds = BandLdsN<kNodesPerWave>;

// The nodes a wave of the rate pass runs: on the modal route half as many as the other
// band passes, for more waves to hide the loads' latency (at 2^14 samples and 128 nodes
// 15.7 us against 17.5 with four and 19.3 with one); on the dense route (C = 0) as many
// (26.8 us against 24.5 with two at the radial grid).
template <int C>
constexpr int rate_nodes_per_wave() {
  return C == 0 ? kNodesPerWave : kNodesPerWave / 2;
}
template <int C>
inline int rate_groups(int Q) {
  constexpr int per_group = kNodeWaves * rate_nodes_per_wave<C>();
  return (Q + per_group - 1) / per_group;
}

template <int C, int NPW = kNodesPerWave>
__device__ __forceinline__ void rate_band_body(const DeviceProblem& P, const PlasmaTable& pt,
                                      const FieldBuffers& B, int tile, int group, double* table,
                                      BandLdsN<NPW>& lds, bool sparse) {
  auto& windows = lds.windows;
  auto& halos = lds.halos;
  (void)table;
  // The rate table is read through the vector caches: most workgroups lie where the field
  // is below the table's threshold and never read it, so no workgroup waits to stage it.
  const double* coefficients = B.rate_c;
  const int T = P.T, tiles = scan_tiles(T), K = pt.n_taps - 1, H = 2 * K;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int t = tile * kTileSpan + lane;
  const bool stores = t < T && (lane < kTileSpan || t == T - 1);
  const bool step = lane < kTileSpan && t + 1 < T;
  double* window = windows[wave];
  double* halo = halos[wave];
  const int q0 = group * kNodeWaves * NPW + wave;
  const int S = H + (tile == tiles - 1 ? K : 0);  // extra samples a node needs
  // The lane's own modal fields first, so that their loads overlap the halo's (up to four
  // modes: more held across the halo would spill).
  double am[held(C)];
  if constexpr (C <= 4) load_modal<C>(B, T, t, am);
  for (int idx = lane; idx < NPW * S; idx += kWave) {
    const int i = idx / S, h = idx - i * S;
    const int q = q0 + i * kNodeWaves;
    double v = 0.0;
    if (q < P.Q) {
      auto sample = [&](int u) { return rate_sample<C>(P, pt, coefficients, B, q, u); };
      v = reflected(T, K, extra_sample(tile, K, T, h), sample);
    }
    halo[i * kBandSlots + h] = v;
  }
  wave_sync();
  if constexpr (C > 4) load_modal<C>(B, T, t, am);
  for (int i = 0; i < NPW; ++i) {
    const int q = q0 + i * kNodeWaves;
    if (q >= P.Q) break;  // uniform across the wave, and so for every later node
    double w = 0.0;
    if (t < T)
      w = rate_of(fabs(field_at<C>(P, B, am, q, t)), pt, coefficients);
    else if (t - T < K)  // the last tile's lanes past the end: their reflections
      w = halo[i * kBandSlots + H + (t - T)];
    const long row = static_cast<long>(q) * T;
    const bool zero = window_is_zero(pt, w, halo + i * kBandSlots, lane);
    if (B.zero_tiles && lane == 0) B.zero_tiles[static_cast<long>(q) * tiles + tile] = zero;
    if (stores && !(zero && sparse)) B.rate[row + t] = w;
    if (zero) {
      if (lane == kTileSpan - 1) B.tile_x[static_cast<long>(q) * tiles + tile] = 0.0;
      if (stores && !sparse) B.x_local[row + t] = 0.0;
      continue;
    }
    wave_sync();  // the previous node's window has been read
    window[K + lane] = w;
    if (lane < H) window[halo_slot(K, lane)] = halo[i * kBandSlots + lane];
    wave_sync();
    // Cross-lane moves with every lane of the wave active: a lane masked off reads as
    // zero to its neighbours.
    const double y = band_filter(pt, window, lane);
    const double y_next = from_next_lane(y);
    const double inc = step ? (P.dt_fine * (y_next + y)) / 2.0 : 0.0;
    const double incl = wave_scan(inc);
    const double excl = from_previous_lane(incl);
    if (lane == kTileSpan - 1) B.tile_x[static_cast<long>(q) * tiles + tile] = incl;
    if (stores) B.x_local[row + t] = excl;
  }
}

template <int C>
__global__ void __launch_bounds__(kWave * kNodeWaves)
    __attribute__((amdgpu_waves_per_eu(C <= 4 ? 6 : 4)))
    plasma_rate_band(DeviceProblem P, PlasmaTable pt, FieldBuffers B, int sparse) {
  __shared__ BandLdsN<rate_nodes_per_wave<C>()> lds;
  rate_band_body<C, rate_nodes_per_wave<C>()>(P, pt, B, blockIdx.x, blockIdx.y, nullptr, lds,
                                              sparse != 0);
}

// Pass C under the band rule: the free-current integrand (the fraction from the stored
// exposure) filtered and its increments scanned; the fraction and the within-tile free
// current stored for the owned samples, and the tile's sum.
template <int C>
__device__ __forceinline__ void current_band_body(const DeviceProblem& P, const PlasmaTable& pt,
                                         const AdamsRecord* __restrict__ R,
                                         const FieldBuffers& B, int tile, int group,
                                         BandLds& lds) {
  auto& windows = lds.windows;
  auto& halos = lds.halos;
  const int T = P.T, tiles = scan_tiles(T), K = pt.n_taps - 1, H = 2 * K;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int t = tile * kTileSpan + lane;
  const int owner = min(t / kTileSpan, tiles - 1);
  const bool stores = t < T && (lane < kTileSpan || t == T - 1);
  const bool step = lane < kTileSpan && t + 1 < T;
  const double f0 = R->preionisation;
  double* window = windows[wave];
  double* halo = halos[wave];
  const int q0 = group * kNodeWaves * kNodesPerWave + wave;
  const int S = H + (tile == tiles - 1 ? K : 0);  // extra samples a node needs
  for (int idx = lane; idx < kNodesPerWave * S; idx += kWave) {
    const int i = idx / S, h = idx - i * S;
    const int q = q0 + i * kNodeWaves;
    double v = 0.0;
    if (q < P.Q) {
      auto sample = [&](int u) { return current_sample<C>(P, B, q, u, f0, tiles); };
      v = reflected(T, K, extra_sample(tile, K, T, h), sample);
    }
    halo[i * kBandSlots + h] = v;
  }
  wave_sync();
  double am[held(C)];
  load_modal<C>(B, T, t, am);
  for (int i = 0; i < kNodesPerWave; ++i) {
    const int q = q0 + i * kNodeWaves;
    if (q >= P.Q) break;  // uniform across the wave, and so for every later node
    const long row = static_cast<long>(q) * T;
    double f = 0.0, y = 0.0;
    if (t < T) {
      const double x = B.tile_x[static_cast<long>(q) * tiles + owner] + B.x_local[row + t];
      f = fraction_of(f0, x > 0.0 ? x : 0.0);
      y = (f * P.plasma_cmr) * field_at<C>(P, B, am, q, t);
    } else if (t - T < K) {  // the last tile's lanes past the end: their reflections
      y = halo[i * kBandSlots + H + (t - T)];
    }
    wave_sync();  // the previous node's window has been read
    window[K + lane] = y;
    if (lane < H) window[halo_slot(K, lane)] = halo[i * kBandSlots + lane];
    wave_sync();
    // Cross-lane moves with every lane of the wave active (see the rate pass).
    const double yf = band_filter(pt, window, lane);
    const double y_next = from_next_lane(yf);
    const double inc = step ? (P.dt_fine * (y_next + yf)) / 2.0 : 0.0;
    const double incl = wave_scan(inc);
    const double excl = from_previous_lane(incl);
    if (lane == kTileSpan - 1) B.tile_f[static_cast<long>(q) * tiles + tile] = incl;
    if (stores) {
      B.frac[row + t] = f;
      B.free_local[row + t] = excl;
    }
  }
}

template <int C>
__global__ void __launch_bounds__(kWave * kNodeWaves)
    __attribute__((amdgpu_waves_per_eu(C <= 4 ? 8 : (C < 8 ? 5 : 4))))
    plasma_current_band(DeviceProblem P, PlasmaTable pt, const AdamsRecord* __restrict__ R,
                        FieldBuffers B) {
  __shared__ BandLds lds;
  current_band_body<C>(P, pt, R, B, blockIdx.x, blockIdx.y, lds);
}

// ------------------------------ the free-current pass fused with the node response
//
// On the fused-node route the free current F(q, t) = F_off(q, tile) + F_local(q, t)
// enters the streams only through the projection Σ_q W[m, q] F(q, t), which is linear:
// the fused pass projects each node's tile-local current (with the loss current and the
// polarisation, which need no offset) over the group's nodes as it scans, and projects
// the nodes' tile totals; the combination adds Σ_{j < tile} of the projected totals
// once every tile's total is known. No per-node array is written after the exposure.

// Pass C fused with pass E: the free-current integrand filtered and scanned as in the
// band pass (either rule: the trapezoid is one tap), then at the lane's sample the
// polarisation (Kerr and Raman) and the current F_local + I_p W (1 - f) E / |E|², each
// projected onto the modes and summed over the wave's nodes in order, then over the
// waves in order (`red`, 2C rows of kWave for each wave). The group's rows of B.group_streams receive
// the sums at the tile's owned samples, and B.tile_p its nodes' free-current tile
// totals projected onto the modes (nodes in order; `totals` holds kNodesPerGroup).
template <int C>
__device__ __forceinline__ void current_response_body(const DeviceProblem& P,
                                                      const PlasmaTable& pt,
                                                      const AdamsRecord* __restrict__ R,
                                                      const FieldBuffers& B, int tile,
                                                      int group, BandLds& lds, double* red,
                                                      double* totals) {
  auto& windows = lds.windows;
  auto& halos = lds.halos;
  const int T = P.T, tiles = scan_tiles(T), K = pt.n_taps - 1, H = 2 * K;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int t = tile * kTileSpan + lane;
  const int owner = min(t / kTileSpan, tiles - 1);
  const bool stores = t < T && (lane < kTileSpan || t == T - 1);
  const bool step = lane < kTileSpan && t + 1 < T;
  const double f0 = R->preionisation;
  const double density = R->density;
  const double kerr = P.has_kerr ? P.kerr_coefficient * density : 0.0;
  double* window = windows[wave];
  double* halo = halos[wave];
  const int base = group * kNodesPerGroup;
  const int q0 = base + wave;
  const int S = H + (tile == tiles - 1 ? K : 0);  // extra samples a node needs
  // Every node's exposure and rate at the lane's sample are loaded before the halo is
  // evaluated, so that their latency overlaps it (the node loop's LDS fences would
  // otherwise hold each node's loads until the previous node is done).
  // A tile the rate pass flagged holds no stores: its rate and within-tile exposure are
  // zero, and the exposure there is the tile's offset (x + 0.0 is x).
  // No load sits under a branch (a branch would make each wait for the one before): the
  // flags and offsets are loaded from clamped addresses, all together; then the stored
  // tiles' entries, where a flagged tile reads a constant zero instead.
  const int tc = min(t, T - 1);
  unsigned char zero[kNodesPerWave];
  double xs[kNodesPerWave], ws[kNodesPerWave];
#pragma unroll
  for (int i = 0; i < kNodesPerWave; ++i) {
    const long tq = static_cast<long>(min(q0 + i * kNodeWaves, P.Q - 1)) * tiles + owner;
    zero[i] = B.zero_tiles[tq];
    xs[i] = B.tile_x[tq];
  }
#pragma unroll
  for (int i = 0; i < kNodesPerWave; ++i) {
    const int q = q0 + i * kNodeWaves;
    const bool here = q < P.Q && t < T;
    const long at = static_cast<long>(min(q, P.Q - 1)) * T + tc;
    const double* x_at = zero[i] ? &kZeroSample : B.x_local + at;
    const double* w_at = zero[i] ? &kZeroSample : B.rate + at;
    const double local = *x_at, w = *w_at;
    xs[i] = here ? xs[i] + local : 0.0;  // a flagged tile's offset plus 0.0
    ws[i] = here ? w : 0.0;
  }
  for (int idx = lane; idx < kNodesPerWave * S; idx += kWave) {
    const int i = idx / S, h = idx - i * S;
    const int q = q0 + i * kNodeWaves;
    double v = 0.0;
    if (q < P.Q) {
      auto sample = [&](int u) { return current_sample_sparse<C>(P, B, q, u, f0, tiles); };
      v = reflected(T, K, extra_sample(tile, K, T, h), sample);
    }
    halo[i * kBandSlots + h] = v;
  }
  wave_sync();
  double am[held(C)];
  load_modal<C>(B, T, t, am);
  double acc_p[C], acc_c[C];
#pragma unroll
  for (int m = 0; m < C; ++m) acc_p[m] = acc_c[m] = 0.0;
#pragma unroll
  for (int i = 0; i < kNodesPerWave; ++i) {
    const int q = q0 + i * kNodeWaves;
    if (q >= P.Q) break;  // uniform across the wave, and so for every later node
    double f = 0.0, e = 0.0, y = 0.0;
    if (t < T) {
      const double x = xs[i];
      f = fraction_of(f0, x > 0.0 ? x : 0.0);
      e = field_at<C>(P, B, am, q, t);
      y = (f * P.plasma_cmr) * e;
    } else if (t - T < K) {  // the last tile's lanes past the end: their reflections
      y = halo[i * kBandSlots + H + (t - T)];
    }
    // Before the pulse the integrand vanishes: the scan is zero, and a zero of either sign
    // leaves the projections as they are.
    double local = 0.0, incl = 0.0;
    if (!window_is_zero(pt, y, halo + i * kBandSlots, lane)) {
      wave_sync();  // the previous node's window has been read
      window[K + lane] = y;
      if (lane < H) window[halo_slot(K, lane)] = halo[i * kBandSlots + lane];
      wave_sync();
      // Cross-lane moves with every lane of the wave active (see the rate pass).
      const double yf = band_filter(pt, window, lane);
      const double y_next = from_next_lane(yf);
      const double inc = step ? (P.dt_fine * (y_next + yf)) / 2.0 : 0.0;
      incl = wave_scan(inc);
      local = from_previous_lane(incl);
    }
    if (lane == kTileSpan - 1) totals[i * kNodeWaves + wave] = incl;
    if (t < T) {
      const double w = ws[i];
      const double squared = fabs(e) * fabs(e);
      double loss = 0.0;
      if (w > 0.0 && squared > 0.0) loss = (P.plasma_ip * w * (1.0 - f)) / squared;
      const double current = local + loss * e;
      double pol = (kerr * (e * e)) * e;
      if (P.n_raman > 0) {
        const double* conv = reinterpret_cast<const double*>(B.raman_drive);
        pol = pol + ((P.raman_epsilon0 * density) * e) * (conv[2L * q * T + t] * P.inv_2T);
      }
#pragma unroll
      for (int m = 0; m < C; ++m) {
        const double wmq = P.projection[static_cast<long>(m) * P.Q + q];
        acc_p[m] = acc_p[m] + wmq * pol;
        acc_c[m] = acc_c[m] + wmq * current;
      }
    }
  }
  // Each wave's projections to its own rows of `red`, then the waves' sums in wave order.
  const int waves = static_cast<int>(blockDim.x) / kWave;
#pragma unroll
  for (int m = 0; m < C; ++m) {
    red[(wave * 2 * C + m) * kWave + lane] = acc_p[m];
    red[(wave * 2 * C + C + m) * kWave + lane] = acc_c[m];
  }
  __syncthreads();
  if (stores)
    for (int j = wave; j < 2 * C; j += waves) {
      double sum = red[j * kWave + lane];
      for (int w = 1; w < waves; ++w) sum = sum + red[(w * 2 * C + j) * kWave + lane];
      B.group_streams[(static_cast<long>(group) * 2 * C + j) * T + t] = sum;
    }
  if (static_cast<int>(threadIdx.x) < C) {
    const int m = threadIdx.x;
    double s = 0.0;
    for (int n = 0; n < kNodesPerGroup && base + n < P.Q; ++n)
      s = s + P.projection[static_cast<long>(m) * P.Q + base + n] * totals[n];
    B.tile_p[(static_cast<long>(group) * C + m) * tiles + tile] = s;
  }
}

// Five waves a SIMD up to four modes (at most 96 VGPRs): the 32 waves of each of T/63
// tiles then take two rounds of the machine at 2^14 samples rather than three.
template <int C>
__global__ void __launch_bounds__(kWave * kNodeWaves)
    __attribute__((amdgpu_waves_per_eu(C <= 4 ? 5 : 4)))
    plasma_current_response(DeviceProblem P, PlasmaTable pt, const AdamsRecord* __restrict__ R,
                            FieldBuffers B) {
  __shared__ BandLds lds;
  __shared__ double red[kNodeWaves * 2 * C * kWave];
  __shared__ double totals[kNodesPerGroup];
  current_response_body<C>(P, pt, R, B, blockIdx.x, blockIdx.y, lds, red, totals);
}

// The sum of a double over the wave's 64 lanes in a fixed order, in every lane.
__device__ inline double wave_total(double v) {
  const double s = wave_scan(v);
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(s), kWave - 1),
                          __builtin_amdgcn_readlane(__double2loint(s), kWave - 1));
}

// The combination over one block of 64 samples: the projected free-current offset of
// each mode at the block's (at most two) tiles, Σ_{j < k} Σ_g tile_p[g, m, j] (lane l
// sums the tiles j ≡ l mod 64 in order, each tile's groups in order, then the wave's
// total in a fixed order, so every block that needs tile k forms the same value); then at
// each sample the groups' projections summed in order, the offset added to the current,
// and the streams written: g P, g ρ J, g' ρ J. Any number of waves; `offsets` holds 2C.
constexpr int kCombineRun = 8;  // tiles a lane of the combination holds (up to 512 tiles)

template <int C>
__device__ __forceinline__ void combine_body(const DeviceProblem& P,
                                             const AdamsRecord* __restrict__ R,
                                             const FieldBuffers& B, int block,
                                             double* offsets) {
  const int T = P.T, tiles = scan_tiles(T), G = node_groups(P.Q);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int waves = static_cast<int>(blockDim.x) / kWave;
  const int t0 = block * kWave;
  const int k0 = min(t0 / kTileSpan, tiles - 1);
  // The second tile's sums are the first's with tile k0 added last by its lane, which is
  // where the direct sum over j < k0 + 1 adds it: the same value either way.
  const int k1 = min(k0 + 1, tiles - 1);
  for (int m = wave; m < C; m += waves) {
    double s = 0.0, s1;
    if (tiles <= kCombineRun * kWave) {  // uniform: each lane's tiles held in registers
      // Every (group, tile) entry is loaded before any is added (clamped addresses, no
      // branch), a group at a time; then the same additions in the same order.
      double v[kCombineRun];
#pragma unroll
      for (int r = 0; r < kCombineRun; ++r) v[r] = 0.0;
      for (int g = 0; g < G; ++g) {
        const double* row = B.tile_p + (static_cast<long>(g) * C + m) * tiles;
        double u[kCombineRun];
#pragma unroll
        for (int r = 0; r < kCombineRun; ++r) u[r] = row[min(lane + r * kWave, tiles - 1)];
#pragma unroll
        for (int r = 0; r < kCombineRun; ++r) v[r] = v[r] + u[r];
      }
      double extra = 0.0;
#pragma unroll
      for (int r = 0; r < kCombineRun; ++r) {
        const int j = lane + r * kWave;
        if (j < k0) s = s + v[r];
        if (j == k0) extra = v[r];
      }
      s1 = k1 > k0 && lane == k0 % kWave ? s + extra : s;
    } else {
      for (int j = lane; j < k0; j += kWave) {
        double tj = 0.0;
        for (int g = 0; g < G; ++g) tj = tj + B.tile_p[(static_cast<long>(g) * C + m) * tiles + j];
        s = s + tj;
      }
      s1 = s;
      if (k1 > k0 && lane == k0 % kWave) {
        double tj = 0.0;
        for (int g = 0; g < G; ++g) tj = tj + B.tile_p[(static_cast<long>(g) * C + m) * tiles + k0];
        s1 = s + tj;
      }
    }
    const double o0 = wave_total(s), o1 = wave_total(s1);
    if (lane == 0) {
      offsets[m] = o0;
      offsets[C + m] = o1;
    }
  }
  __syncthreads();
  const int t = t0 + lane;  // < T: the grid is T/64 blocks
  const int kk = min(t / kTileSpan, tiles - 1) - k0;
  const double g = P.t_window[t], gd = P.t_deriv[t], density = R->density;
  double* streams = reinterpret_cast<double*>(B.streams);
  for (int m = wave; m < C; m += waves) {
    double sp = 0.0, sc = 0.0;
    for (int gi = 0; gi < G; ++gi) {
      sp = sp + B.group_streams[(static_cast<long>(gi) * 2 * C + m) * T + t];
      sc = sc + B.group_streams[(static_cast<long>(gi) * 2 * C + C + m) * T + t];
    }
    const double current = density * (sc + offsets[kk * C + m]);
    streams[static_cast<long>(m) * T + t] = g * sp;
    streams[static_cast<long>(C + m) * T + t] = g * current;
    streams[static_cast<long>(2 * C + m) * T + t] = gd * current;
  }
}

template <int C>
__global__ void __launch_bounds__(kWave * 4)
    plasma_response_combine(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  __shared__ double offsets[2 * C];
  combine_body<C>(P, R, B, blockIdx.x, offsets);
}

// Passes B and D: exclusive prefix of the tile sums of each node, one wavefront per
// node. Each lane sums a run of consecutive tiles in order, the wave scans the run
// totals, and each lane then writes its run's offsets in order: a fixed order, so the
// offsets are the same on every run.
// Runs of up to kPrefixRun tiles a lane (grids of up to 512 tiles, 32257 samples) are
// loaded into registers at once, so that their loads overlap; longer runs take the loop.
// The additions are the same, in the same order, either way.
constexpr int kPrefixRun = 8;

__device__ __forceinline__ void tile_prefix_body(int tiles, double* sums, int node) {
  const int lane = threadIdx.x;
  double* row = sums + static_cast<long>(node) * tiles;
  const int run = (tiles + kWave - 1) / kWave;
  const int k0 = lane * run;
  const int k1 = k0 + run < tiles ? k0 + run : tiles;
  if (run <= kPrefixRun) {  // uniform across the wave
    double v[kPrefixRun];
#pragma unroll
    for (int r = 0; r < kPrefixRun; ++r) v[r] = k0 + r < k1 ? row[k0 + r] : 0.0;
    double total = 0.0;
#pragma unroll
    for (int r = 0; r < kPrefixRun; ++r)
      if (k0 + r < k1) total = total + v[r];
    double running = from_previous_lane(wave_scan(total));
#pragma unroll
    for (int r = 0; r < kPrefixRun; ++r)
      if (k0 + r < k1) {
        row[k0 + r] = running;
        running = running + v[r];
      }
    return;
  }
  double total = 0.0;
  for (int k = k0; k < k1; ++k) total = total + row[k];
  double running = from_previous_lane(wave_scan(total));
  for (int k = k0; k < k1; ++k) {
    const double here = row[k];
    row[k] = running;
    running = running + here;
  }
}

__global__ void __launch_bounds__(kWave) plasma_tile_prefix(int tiles, double* sums) {
  tile_prefix_body(tiles, sums, blockIdx.x);
}

// Pass C: the free-current increments' sum over each tile, and each sample's fraction
// and within-tile free current for the response pass. Same layout as pass A.
template <int C>
__device__ __forceinline__ void current_tiles_body(const DeviceProblem& P,
                                          const AdamsRecord* __restrict__ R,
                                          const FieldBuffers& B, int tile, int group) {
  const int T = P.T, tiles = scan_tiles(T);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int t = tile * kTileSpan + lane;
  const bool stores = t < T && (lane < kTileSpan || t == T - 1);
  const double f0 = R->preionisation;
  double am[held(C)];
  load_modal<C>(B, T, t, am);
  const int q0 = group * kNodeWaves * kNodesPerWave;
#pragma unroll
  for (int i = 0; i < kNodesPerWave; ++i) {
    const int q = q0 + i * kNodeWaves + wave;
    if (q < P.Q) {  // uniform across the wave
      const double e = field_at<C>(P, B, am, q, t);
      const long row = static_cast<long>(q) * tiles + tile;
      const TileColumn c = tile_column(P, e, B.rate + static_cast<long>(q) * T, B.tile_x[row],
                                       f0, t, lane);
      if (lane == kTileSpan - 1) B.tile_f[row] = c.j_incl;
      if (stores) {
        B.frac[static_cast<long>(q) * T + t] = c.f;
        B.free_local[static_cast<long>(q) * T + t] = c.j_excl;
      }
    }
  }
}

template <int C>
__global__ void __launch_bounds__(kWave * kNodeWaves)
    plasma_current_tiles(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  current_tiles_body<C>(P, R, B, blockIdx.x, blockIdx.y);
}

// Pass E: Kerr, the fraction, the free and loss currents and the collars at every node,
// projected onto the modes. A workgroup takes 64 consecutive samples, one a lane (its
// own blocks, not the scan tiles: T/64 workgroups, T a power of two); each wave takes
// every kResponseWaves-th node, and the waves' projections are summed in wave order in
// LDS and written once. A sample's free current is its owning scan tile's offset plus
// the within-tile current that tile stored: tile t / 63, the last tile for the grid's
// last sample.
template <int C>
__device__ __forceinline__ void node_response_body(const DeviceProblem& P,
                                          const AdamsRecord* __restrict__ R,
                                          const FieldBuffers& B, int block, double* red) {
  const int T = P.T, tiles = scan_tiles(T);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int t = block * kWave + lane;  // < T: the grid is T/64 workgroups
  const int tile = min(t / kTileSpan, tiles - 1);
  const double density = R->density;
  const double g = P.t_window[t], gd = P.t_deriv[t];
  double am[C];
  load_modal<C>(B, T, t, am);
  double acc[3][C];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int m = 0; m < C; ++m) acc[s][m] = 0.0;
#pragma unroll(C <= 4 ? 2 : 1)
  for (int q = wave; q < P.Q; q += kResponseWaves) {
    const double e = synthesise_at<C>(P.synthesis + static_cast<long>(q) * C, am);
    double current = 0.0;
    if (P.has_plasma) {
      const long at = static_cast<long>(q) * T + t;
      const double w = B.rate[at], f = B.frac[at];
      const double F = B.tile_f[static_cast<long>(q) * tiles + tile] + B.free_local[at];
      const double squared = fabs(e) * fabs(e);
      double loss = 0.0;
      if (w > 0.0 && squared > 0.0) loss = (P.plasma_ip * w * (1.0 - f)) / squared;
      current = density * (F + loss * e);
    }
    double pol = 0.0;
    if (P.has_kerr) {
      const double factor = P.kerr_coefficient * density;
      pol = (factor * (e * e)) * e;
    }
    if (P.n_raman > 0) {
      const double* conv = reinterpret_cast<const double*>(B.raman_drive);
      const double cv = conv[2L * q * T + t] * P.inv_2T;
      pol = pol + ((P.raman_epsilon0 * density) * e) * cv;
    }
    pol = pol * g;
    const double gated = g * current, edge = gd * current;
#pragma unroll
    for (int m = 0; m < C; ++m) {
      const double wmq = P.projection[static_cast<long>(m) * P.Q + q];
      acc[0][m] = acc[0][m] + wmq * pol;
      acc[1][m] = acc[1][m] + wmq * gated;
      acc[2][m] = acc[2][m] + wmq * edge;
    }
  }
  // Sum the waves' projections in wave order.
  for (int w = 0; w < kResponseWaves; ++w) {
    if (wave == w)
#pragma unroll
      for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int m = 0; m < C; ++m) {
          double& cell = red[(s * C + m) * kWave + lane];
          cell = w == 0 ? acc[s][m] : cell + acc[s][m];
        }
    __syncthreads();
  }
  if (wave == 0) {
    double* streams = reinterpret_cast<double*>(B.streams);
#pragma unroll
    for (int j = 0; j < 3 * C; ++j) streams[static_cast<long>(j) * T + t] = red[j * kWave + lane];
  }
}

template <int C>
__global__ void __launch_bounds__(kWave * kResponseWaves)
    node_response(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  __shared__ double red[3 * C * kWave];
  node_response_body<C>(P, R, B, blockIdx.x, red);
}

// Pass E of the dense route, one thread per node and sample: the un-gated Kerr
// polarisation and the total current J = ρ (F + I_p W (1 - f) E / |E|²), laid out
// (Q, 2, T) for the projection GEMM, which applies g and g' (as (Q, T) polarisation
// alone without plasma). A sample's fraction and within-tile current are those its
// owning tile stored: tile t / 63, the last tile for the grid's last sample.
__global__ void node_response_dense(DeviceProblem P, const AdamsRecord* __restrict__ R,
                                    FieldBuffers B) {
  const int T = P.T, tiles = scan_tiles(T);
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.Q) * T) return;
  const int q = static_cast<int>(idx / T), t = static_cast<int>(idx % T);
  const double density = R->density;
  const double e = B.node_field[idx];
  double pol = 0.0;
  if (P.has_kerr) {
    const double factor = P.kerr_coefficient * density;
    pol = (factor * (e * e)) * e;
  }
  if (P.n_raman > 0) {
    const double* conv = reinterpret_cast<const double*>(B.raman_drive);
    const double cv = conv[2L * q * T + t] * P.inv_2T;
    pol = pol + ((P.raman_epsilon0 * density) * e) * cv;
  }
  if (!P.has_plasma) {
    B.node_streams[idx] = pol;
    return;
  }
  const int tile = min(t / kTileSpan, tiles - 1);
  const double w = B.rate[idx], f = B.frac[idx];
  const double F = B.tile_f[static_cast<long>(q) * tiles + tile] + B.free_local[idx];
  const double squared = fabs(e) * fabs(e);
  double loss = 0.0;
  if (w > 0.0 && squared > 0.0) loss = (P.plasma_ip * w * (1.0 - f)) / squared;
  const double current = density * (F + loss * e);
  const long row = 2L * q * T;
  B.node_streams[row + t] = pol;
  B.node_streams[row + T + t] = current;
}

// ------------------------------------------------------------- Raman response

// The damped kernel h(t) = fraction Σ_r undamped_r(t) e^{-Γ_r t} on the doubled grid at
// the record's rates, two reals a point: z[m] = (h(2m), h(2m+1)), m < T.
__global__ void field_raman_kernel_fill(DeviceProblem P, const AdamsRecord* __restrict__ R,
                                        cplx* z) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= P.T) return;
  const long n2 = 2L * P.T;
  double h[2];
  for (int j = 0; j < 2; ++j) {
    const long t = 2L * m + j;
    const double delay = static_cast<double>(t) * P.dt_fine;
    double total = 0.0;
    for (int r = 0; r < P.n_raman; ++r)
      total = total + P.raman_rows[r * n2 + t] * exp(-R->raman_rate[r] * delay);
    h[j] = P.raman_fraction * total;
  }
  z[m] = make(h[0], h[1]);
}

// The drive E² of node q on the doubled grid, two reals a point: (E²(2m), E²(2m+1)) for
// 2m < T, zeros on the padding half. The field is synthesised from the modal fields
// (C modes), or on the dense route (C = 0) read from the node field.
template <int C>
__global__ void field_raman_drive(DeviceProblem P, FieldBuffers B) {
  const int T = P.T;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.Q) * T) return;
  const int q = static_cast<int>(idx / T), m = static_cast<int>(idx % T);
  if (2 * m >= T) {
    B.raman_drive[idx] = make(0.0, 0.0);
    return;
  }
  double e[2];
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * m + j;
    double am[held(C)];
    load_modal<C>(B, T, t, am);
    const double field = field_at<C>(P, B, am, q, t);
    if (!P.raman_analytic) {
      e[j] = field * field;
      continue;
    }
    // (E^2 + H[E]^2) / 2, H[E] synthesised from the modes' Hilbert transforms as E is
    // from their fields (the transform is linear), or read on the dense route.
    double hilbert;
    if constexpr (C > 0) {
      const double* h = reinterpret_cast<const double*>(B.modal_h);
      double hm[C];
#pragma unroll
      for (int c = 0; c < C; ++c) hm[c] = h[static_cast<long>(c) * T + t];
      hilbert = synthesise_at<C>(P.synthesis + static_cast<long>(q) * C, hm);
    } else {
      hilbert = B.node_hilbert[static_cast<long>(q) * T + t];
    }
    e[j] = 0.5 * (field * field + hilbert * hilbert);
  }
  B.raman_drive[idx] = make(e[0], e[1]);
}

// One coarse bin b of mode c from its two packed entries per stream (za at b, zb at its
// mirror): unpack the stream spectra, assemble the held current on the active bins and
// normalise, k = (p W) (P~ + G~/(iω) - E~/(iω)^2).
__device__ inline cplx finish_value(const DeviceProblem& P, const AdamsRecord& r,
                                    const FieldBuffers& B, int c, int b, const cplx (&za)[4],
                                    const cplx (&zb)[4], int n_streams) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(c) * P.N + b;
  cplx X[3] = {};
#pragma unroll
  for (int s = 0; s < 3; ++s)
    if (s < n_streams) X[s] = r2c_post(za[s], zb[s], B.roots_T[b]);
  cplx spectrum = X[0];
  if (P.has_plasma && P.active[b]) {
    const cplx q = make(0.0, P.omega[b]);
    const cplx q2 = mul(q, q);
    spectrum = add(spectrum, sub(cdiv(X[1], q), cdiv(X[2], q2)));
  }
  const cplx pref = weighted(r.pref_w, P.pref, P.n_pref, CN, idx);
  cplx value = mul(scale(P.omega_window[b], pref), spectrum);
  // The in-stage absorber's spectrum, on the active band, without the prefactor. Its
  // stream follows the polarisation (and the two currents with plasma).
  if (P.absorber_stream >= 0 && P.active[b]) {
    const cplx a = P.has_plasma ? za[3] : za[1], ab = P.has_plasma ? zb[3] : zb[1];
    value = add(value, r2c_post(a, ab, B.roots_T[b]));
  }
  return value;
}

__device__ inline void finish_bin(const DeviceProblem& P, const AdamsRecord& r,
                                  const FieldBuffers& B, int c, int b, const cplx (&za)[4],
                                  const cplx (&zb)[4], int n_streams, cplx* out) {
  out[static_cast<long>(c) * P.N + b] = finish_value(P, r, B, c, b, za, zb, n_streams);
}

// The finish of one kept bin b < N of mode c alone: its two packed entries per stream
// are row[b mod T/2] and row[(T/2 - b) mod T/2] (the Nyquist bin T/2 reads entry 0
// twice, as its pair's thread does above).
__device__ inline cplx finish_at(const DeviceProblem& P, const AdamsRecord& r,
                                 const FieldBuffers& B, int c, int b) {
  const int half = P.T / 2;
  const int ka = b % half, kb = (half - b) % half;
  const int n_streams = P.has_plasma ? 3 : 1;
  cplx za[4] = {}, zb[4] = {};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s < n_streams || s == P.absorber_stream) {
      const cplx* row = B.streams + (static_cast<long>(s) * P.C + c) * half;
      za[s] = row[ka];
      zb[s] = row[kb];
    }
  }
  return finish_value(P, r, B, c, b, za, zb, n_streams);
}

// The finish over bin pairs: the packed real transform holds bins k and T/2 - k in the
// same two entries (k = 0 pairs with the Nyquist bin T/2, both from entry 0), so one
// thread reads them once and writes both bins that the coarse grid keeps.
__device__ __forceinline__ void finish_body(const DeviceProblem& P, const AdamsRecord* __restrict__ R,
                                   const FieldBuffers& B, cplx* out, long idx) {
  const int half = P.T / 2;
  const int pairs = half / 2 + 1;
  if (idx >= static_cast<long>(P.C) * pairs) return;
  const int c = static_cast<int>(idx / pairs), k = static_cast<int>(idx % pairs);
  const int mirror = half - k;  // the partner bin, T/2 for k = 0
  if (k >= P.N) return;         // neither bin is kept (mirror >= k)
  const int ka = k, kb = mirror % half;
  const int n_streams = P.has_plasma ? 3 : 1;  // the currents are transformed with plasma
  cplx za[4] = {}, zb[4] = {};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s < n_streams || s == P.absorber_stream) {
      const cplx* row = B.streams + (static_cast<long>(s) * P.C + c) * half;
      za[s] = row[ka];
      zb[s] = row[kb];
    }
  }
  finish_bin(P, *R, B, c, k, za, zb, n_streams, out);
  if (mirror != k && mirror < P.N) finish_bin(P, *R, B, c, mirror, zb, za, n_streams, out);
}

__global__ void field_finish(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B,
                             cplx* out) {
  finish_body(P, R, B, out, static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------ temporal absorber

// The in-stage term -alpha(t) a_c(t) / 2 of every mode, as a stream of its own, or added
// to what the stream holds (the transverse collar's product).
__global__ void field_absorber_stream(DeviceProblem P, FieldBuffers B, bool accumulate) {
  const int T = P.T;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.C) * T) return;
  const int t = static_cast<int>(idx % T);
  const double* modal = reinterpret_cast<const double*>(B.modal_t);
  double* stream = reinterpret_cast<double*>(B.streams) +
                   static_cast<long>(P.absorber_stream) * P.C * T;
  const double term = (-0.5 * P.absorber_rate[t]) * modal[idx];
  stream[idx] = accumulate ? term + stream[idx] : term;
}

// The split-step factor on the committed state's time samples: exp(-alpha(t) h / 2).
__global__ void field_split_collar(DeviceProblem P, const AdamsRecord* __restrict__ R,
                                   double dz, FieldBuffers B) {
  const int T = P.T;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.C) * T) return;
  const int t = static_cast<int>(idx % T);
  const double h = R ? R->accepted_dz : dz;
  double* modal = reinterpret_cast<double*>(B.modal_t);
  modal[idx] = modal[idx] * exp((-0.5 * P.absorber_rate[t]) * h);
}

// The collared state back on the coarse bins, the band imposed again.
__global__ void field_split_unpack(DeviceProblem P, FieldBuffers B) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= CN) return;
  const int c = static_cast<int>(idx / P.N), b = static_cast<int>(idx % P.N);
  const int half = P.T / 2;
  const cplx* row = B.modal_t + static_cast<long>(c) * half;
  const cplx x = r2c_post(row[b % half], row[(half - b) % half], B.roots_T[b]);
  B.y[idx] = P.active[b] ? x : make(0.0, 0.0);
}

// ------------------------------------------------------------- Adams corrector

// The corrector, the defects of the scored orders and their masked squared sums, for a
// stencil of width W. One wavefront per workgroup, operands loaded first; each wave's
// sums are reduced across its lanes in a fixed order and written as that workgroup's
// partials. Coefficients past each order are zero in the record.
// The corrector's partial sums of one workgroup, in its waves' LDS cells.
using WaveSums = double[1 + kMaxScored][kCorrectThreads / kWave];

template <int W, bool Finish = false>
__device__ __forceinline__ void correct_body(const DeviceProblem& P, const AdamsRecord* __restrict__ R,
                                    const FieldBuffers& B, int block, int blocks,
                                    WaveSums& wave_sums) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(block) * blockDim.x + threadIdx.x;
  const AdamsRecord& r = *R;
  double v[1 + kMaxScored] = {0.0, 0.0, 0.0, 0.0};
  if (idx < CN) {
    const int b = static_cast<int>(idx % P.N);
    const double dz = r.dz;
    const int n_orders = r.n_orders;
    const bool compensated = r.compensated != 0;
    // Loads.
    cplx X[W], carried[W], phi_t[kMaxTerms];
    if constexpr (Finish) {
      // The source evaluation's finish for this bin, stored for the commit as the finish
      // kernel would store it.
      X[0] = finish_at(P, r, B, static_cast<int>(idx / P.N), b);
      B.source[idx] = X[0];
    } else {
      X[0] = B.source[idx];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
      if (k >= 1) X[k] = B.store[r.slots[k - 1] * CN + idx];
      carried[k] = compensated && k < W - 1 ? B.carried[k * CN + idx] : make(0.0, 0.0);
    }
#pragma unroll
    for (int t = 0; t < kMaxTerms; ++t)
      phi_t[t] = t < P.n_phi ? P.phi[t * CN + idx] : make(0.0, 0.0);
    const cplx y = B.y[idx];
    const cplx lab = compensated ? B.lab[idx] : make(0.0, 0.0);
    const bool scored = P.normmask[b] != 0;
    cplx M[W], flow;
    const cplx x = step_operator<W>(P, r, phi_t, M, flow);

    cplx adv = add(mul(flow, y), apply_moments<W>(M, r.corr_c, X, dz));
    cplx comp_def[kMaxScored];
#pragma unroll
    for (int i = 0; i < kMaxScored; ++i) comp_def[i] = make(0.0, 0.0);
    if (compensated) {
      cplx rate_end = make(0.0, 0.0);
#pragma unroll
      for (int t = 0; t < kMaxTerms; ++t)
        if (t < P.n_phi) rate_end = add(rate_end, scale(r.rate_end_w[t], phi_t[t]));
      const cplx reference = make(x.x / dz, x.y / dz);
      const cplx endpoint = mul(sub(rate_end, reference), lab);
#pragma unroll
      for (int i = 0; i < kMaxScored; ++i) {
        if (i < n_orders) {
          cplx corr = make(0.0, 0.0), pred = make(0.0, 0.0);
#pragma unroll
          for (int j = 0; j < W - 1; ++j)
            corr = add(corr, scale(r.comp_corr_w[i][1 + j], carried[j]));
          corr = scale(dz, add(corr, scale(r.comp_corr_w[i][0], endpoint)));
#pragma unroll
          for (int j = 0; j < W - 1; ++j) pred = add(pred, scale(r.comp_pred_w[i][j], carried[j]));
          pred = scale(dz, pred);
          comp_def[i] = sub(corr, pred);
          if (i == 0) adv = add(adv, corr);
        }
      }
    }
    B.advanced[idx] = adv;
    if (scored) {
      v[0] = norm2(y);
#pragma unroll
      for (int i = 0; i < kMaxScored; ++i) {
        if (i < n_orders) {
          cplx d = apply_moments<W>(M, r.rows_c[i], X, dz);
          if (compensated) d = add(d, comp_def[i]);
          v[1 + i] = norm2(d);
        }
      }
    }
  }
  // Each wave's sums across its lanes, then the workgroup's waves in order.
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int s = 0; s < 1 + kMaxScored; ++s) {
    const double total = wave_scan(v[s]);
    if (lane == kWave - 1) wave_sums[s][wave] = total;
  }
  __syncthreads();
  if (threadIdx.x < 1 + kMaxScored) {
    const int s = threadIdx.x;
    double total = wave_sums[s][0];
    for (int w = 1; w < static_cast<int>(blockDim.x) / kWave; ++w) total = total + wave_sums[s][w];
    B.partials[s * blocks + block] = total;
  }
}

template <int W, bool Finish>
__global__ void __launch_bounds__(kCorrectThreads)
    adams_correct(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  __shared__ WaveSums wave_sums;
  correct_body<W, Finish>(P, R, B, blockIdx.x, gridDim.x, wave_sums);
}

// The corrector's block sums reduced to the 1 + kMaxScored totals in a fixed order, written
// to the host's mapped memory. One workgroup: a group of kNormGroup threads per total,
// thread j summing blocks j, j + kNormGroup, ... in order (coalesced across the group),
// then a fixed pairwise tree in LDS.
constexpr int kNormGroup = 256;
constexpr int kNormThreads = kNormGroup * (1 + kMaxScored);
__device__ __forceinline__ void norms_body(const double* __restrict__ partials, int blocks,
                                  double* norms, double* tree) {
  const int s = threadIdx.x / kNormGroup, j = threadIdx.x % kNormGroup;
  const double* row = partials + static_cast<long>(s) * blocks;
  double total = 0.0;
#pragma unroll 8
  for (int b = j; b < blocks; b += kNormGroup) total = total + row[b];
  tree[threadIdx.x] = total;
  __syncthreads();
  for (int half = kNormGroup / 2; half > 0; half /= 2) {
    if (j < half) tree[threadIdx.x] = tree[threadIdx.x] + tree[threadIdx.x + half];
    __syncthreads();
  }
  if (j == 0) norms[s] = tree[threadIdx.x];
}

__global__ void __launch_bounds__(kNormThreads)
    adams_norms(const double* __restrict__ partials, int blocks, double* norms) {
  __shared__ double tree[kNormThreads];
  norms_body(partials, blocks, norms, tree);
}

__global__ void adams_accept(DeviceProblem P, int slot, int compensated, FieldBuffers B) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= CN) return;
  B.store[slot * CN + idx] = B.source[idx];
  if (compensated) B.states[slot * CN + idx] = B.lab[idx];
  B.y[idx] = P.active[idx % P.N] ? B.advanced[idx] : make(0.0, 0.0);
}

// The same commit, its slot read from the record: the head of the next trial's graph.
__device__ __forceinline__ void commit_body(const DeviceProblem& P, const AdamsRecord* __restrict__ R,
                                   const FieldBuffers& B, long idx) {
  const long CN = static_cast<long>(P.C) * P.N;
  if (idx >= CN) return;
  const int slot = R->free_slot;
  B.store[slot * CN + idx] = B.source[idx];
  if (R->compensated) B.states[slot * CN + idx] = B.lab[idx];
  B.y[idx] = P.active[idx % P.N] ? B.advanced[idx] : make(0.0, 0.0);
}

__global__ void adams_commit(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  commit_body(P, R, B, static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

__device__ __forceinline__ void start_body(const DeviceProblem& P, const AdamsRecord* __restrict__ R,
                                  const FieldBuffers& B, long idx) {
  const long CN = static_cast<long>(P.C) * P.N;
  if (idx >= CN) return;
  B.store[idx] = B.source[idx];
  if (R->compensated) B.states[idx] = B.y[idx];
}

__global__ void adams_start(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B) {
  start_body(P, R, B, static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------ batched marches

// The grouped kernels (field.hpp): each workgroup finds its instance by binary search
// over the launch's first workgroups, reads the instance's record and runs the plain
// kernel's body for its local workgroup, with the instance's mode count and stencil
// width as a switch (uniform across the workgroup).

namespace {

struct Owned {
  const BatchInstance* inst;
  const AdamsRecord* R;
  int local;
};

// The grouped kernels take the instance table and the first workgroups as read-only,
// unaliased arguments, so that the uniform reads of an instance's problem and record
// are scalar loads, as they are from the plain kernels' arguments.
#define SP_BATCH_ARGS                                                                \
  const BatchInstance* __restrict__ insts, const AdamsRecord* __restrict__ records, \
      const int* __restrict__ first, int n_inst

__device__ inline Owned owned(const BatchInstance* __restrict__ insts,
                              const AdamsRecord* __restrict__ records,
                              const int* __restrict__ first, int n_inst) {
  const int i = owner_of(first, n_inst, blockIdx.x);
  return {insts + i, records + i, static_cast<int>(blockIdx.x) - first[i]};
}

template <template <int> class Body, typename... Args>
__device__ __forceinline__ void with_width(int width, Args&&... args) {
  switch (width) {
    case 2: Body<2>::run(args...); break;
    case 3: Body<3>::run(args...); break;
    case 4: Body<4>::run(args...); break;
    case 5: Body<5>::run(args...); break;
    case 6: Body<6>::run(args...); break;
    default: Body<7>::run(args...); break;
  }
}

template <template <int> class Body, typename... Args>
__device__ __forceinline__ void with_modes(int modes, Args&&... args) {
  switch (modes) {
    case 1: Body<1>::run(args...); break;
    case 2: Body<2>::run(args...); break;
    case 3: Body<3>::run(args...); break;
    case 4: Body<4>::run(args...); break;
    case 5: Body<5>::run(args...); break;
    default: Body<6>::run(args...); break;
  }
}

template <int W>
struct PredictAt {
  __device__ __forceinline__ static void run(const BatchInstance& in, const AdamsRecord* __restrict__ R,
                             long idx) {
    predict_body<W>(in.P, R, in.B, idx);
  }
};

template <int W>
struct CorrectAt {
  __device__ __forceinline__ static void run(const BatchInstance& in, const AdamsRecord* __restrict__ R,
                             int block, WaveSums& sums) {
    correct_body<W>(in.P, R, in.B, block, in.correct_blocks, sums);
  }
};

template <int C>
struct RateAt {
  __device__ __forceinline__ static void run(const BatchInstance& in, int tile, int group, double* table,
                             BandLds& lds) {
    // The fused passes read the within-tile exposure, which the band pass stores (under
    // either rule: one tap is the trapezoid).
    const bool fused = fused_node_passes(in.P, in.B);
    if (in.pt.n_taps > 1 || fused)
      rate_band_body<C>(in.P, in.pt, in.B, tile, group, table, lds, fused);
    else
      rate_tiles_body<C>(in.P, in.pt, in.B, tile, group, table);
  }
};

template <int C>
struct CurrentAt {
  __device__ __forceinline__ static void run(const BatchInstance& in, const AdamsRecord* __restrict__ R,
                             int tile, int group, BandLds& lds, double* red, double* totals) {
    if (fused_node_passes(in.P, in.B))
      current_response_body<C>(in.P, in.pt, R, in.B, tile, group, lds, red, totals);
    else if (in.pt.n_taps > 1)
      current_band_body<C>(in.P, in.pt, R, in.B, tile, group, lds);
    else
      current_tiles_body<C>(in.P, R, in.B, tile, group);
  }
};

template <int C>
struct ResponseAt {
  __device__ __forceinline__ static void run(const BatchInstance& in, const AdamsRecord* __restrict__ R,
                             int block, double* red) {
    if (fused_node_passes(in.P, in.B))
      combine_body<C>(in.P, R, in.B, block, red);
    else
      node_response_body<C>(in.P, R, in.B, block, red);
  }
};

}  // namespace

__global__ void __launch_bounds__(kThreads) batch_commit(SP_BATCH_ARGS) {
  const Owned o = owned(insts, records, first, n_inst);
  if (o.R->free_slot < 0) return;
  commit_body(o.inst->P, o.R, o.inst->B, static_cast<long>(o.local) * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(kWave) batch_predict(SP_BATCH_ARGS) {
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  const int action = o.R->action;
  const long idx = static_cast<long>(o.local) * blockDim.x + threadIdx.x;
  if (action == kBatchStart) {
    // The start's source is evaluated at the state itself.
    if (idx < static_cast<long>(in.P.C) * in.P.N) in.B.lab[idx] = in.B.y[idx];
    return;
  }
  if (action != kBatchTrial) return;
  with_width<PredictAt>(o.R->width, in, o.R, idx);
}

__global__ void __launch_bounds__(kWave * kNodeWaves) batch_rate(SP_BATCH_ARGS) {
  __shared__ double table[4 * kLdsTableIntervals];
  __shared__ BandLds lds;
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  if (o.R->action == kBatchIdle) return;
  const int tiles = scan_tiles(in.P.T);
  with_modes<RateAt>(in.P.C, in, o.local % tiles, o.local / tiles, table, lds);
}

__global__ void __launch_bounds__(kWave) batch_prefix(SP_BATCH_ARGS, int current) {
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  if (o.R->action == kBatchIdle) return;
  tile_prefix_body(scan_tiles(in.P.T), current ? in.B.tile_f : in.B.tile_x, o.local);
}

__global__ void __launch_bounds__(kWave * kNodeWaves) batch_current(SP_BATCH_ARGS) {
  __shared__ BandLds lds;
  __shared__ double red[kNodeWaves * 2 * kMaxBatchModes * kWave];
  __shared__ double totals[kNodesPerGroup];
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  if (o.R->action == kBatchIdle) return;
  const int tiles = scan_tiles(in.P.T);
  with_modes<CurrentAt>(in.P.C, in, o.R, o.local % tiles, o.local / tiles, lds, red, totals);
}

__global__ void __launch_bounds__(kWave * kResponseWaves) batch_response(SP_BATCH_ARGS) {
  __shared__ double red[3 * kMaxBatchModes * kWave];
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  if (o.R->action == kBatchIdle) return;
  with_modes<ResponseAt>(in.P.C, in, o.R, o.local, red);
}

__global__ void __launch_bounds__(kWave) batch_finish(SP_BATCH_ARGS) {
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  if (o.R->action == kBatchIdle) return;
  finish_body(in.P, o.R, in.B, in.B.source, static_cast<long>(o.local) * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(kWave) batch_correct(SP_BATCH_ARGS) {
  __shared__ WaveSums sums;
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  const int action = o.R->action;
  if (action == kBatchStart) {
    start_body(in.P, o.R, in.B, static_cast<long>(o.local) * blockDim.x + threadIdx.x);
    return;
  }
  if (action != kBatchTrial) return;
  with_width<CorrectAt>(o.R->width, in, o.R, o.local, sums);
}

__global__ void __launch_bounds__(kNormThreads) batch_norms(SP_BATCH_ARGS) {
  __shared__ double tree[kNormThreads];
  const Owned o = owned(insts, records, first, n_inst);
  const BatchInstance& in = *o.inst;
  if (o.R->action != kBatchTrial) return;
  norms_body(in.B.partials, in.correct_blocks, in.B.norms, tree);
}

// ------------------------------------------------- the low-storage RK4IP passes
//
// The four register passes of one RK4IP step (J. Hult, J. Lightwave Technol. 25, 3770
// (2007), in the three-register form of the reference's low-storage engines:
// src/solprop/backends/numpy_ref/engines.py:432-480), each after its source evaluation
// S = B.source: A = B.y (the state), Bacc the accumulator, G the previous source (for the
// stiffness indicator), C = B.lab the stage field the next source reads. R[0].x_w and
// R[1].x_w are the weights of Phi(z_m) - Phi(z) and Phi(z + h) - Phi(z_m); the flows
// P = exp(.) and Q = exp(.) are formed per element from the operator's terms. The passes
// with a norm leave one masked sum per workgroup: pass 2 |S2 - g|^2, pass 3 |S3 - S2|^2,
// pass 4 |A|^2 (the state's, after the hard mask).
__device__ inline cplx ls_flow(const DeviceProblem& P, const double* w, long idx, long CN) {
  cplx x = make(0.0, 0.0);
#pragma unroll
  for (int t = 0; t < kMaxTerms; ++t)
    if (t < P.n_phi) x = add(x, scale(w[t], P.phi[t * CN + idx]));
  return cexp(x);
}

// With Finish (the field route) a pass finishes its bin's source from the forward
// transform's streams itself, at its evaluation's record (R[0] for pass 1, R[1] for passes
// 2 and 3, R[2] for pass 4), instead of reading B.source.
template <int Pass, bool Finish>
__global__ void __launch_bounds__(kLsThreads)
    ls_pass(DeviceProblem P, const AdamsRecord* __restrict__ R, FieldBuffers B, cplx* acc,
            cplx* prev, double* partials) {
  __shared__ double tree[kLsThreads];
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const double h = R[0].dz;
  double sq = 0.0;
  if (idx < CN) {
    constexpr int at = Pass == 1 ? 0 : (Pass == 4 ? 2 : 1);
    const cplx s = Finish ? finish_at(P, R[at], B, static_cast<int>(idx / P.N),
                                      static_cast<int>(idx % P.N))
                          : B.source[idx];
    const bool scored = P.normmask[idx % P.N] != 0;
    if constexpr (Pass == 1) {
      const cplx flow = ls_flow(P, R[0].x_w, idx, CN);
      const cplx a = mul(flow, B.y[idx]);
      const cplx g = mul(flow, s);
      B.y[idx] = a;
      prev[idx] = g;
      acc[idx] = add(a, scale(h / 6.0, g));
      B.lab[idx] = add(a, scale(0.5 * h, g));
    } else if constexpr (Pass == 2) {
      const cplx d = sub(s, prev[idx]);
      if (scored) sq = norm2(d);
      acc[idx] = add(acc[idx], scale(h / 3.0, s));
      B.lab[idx] = add(B.y[idx], scale(0.5 * h, s));
      prev[idx] = s;
    } else if constexpr (Pass == 3) {
      const cplx d = sub(s, prev[idx]);
      if (scored) sq = norm2(d);
      acc[idx] = add(acc[idx], scale(h / 3.0, s));
      const cplx flow = ls_flow(P, R[1].x_w, idx, CN);
      B.lab[idx] = mul(flow, add(B.y[idx], scale(h, s)));
    } else {
      const cplx flow = ls_flow(P, R[1].x_w, idx, CN);
      cplx a = add(mul(flow, acc[idx]), scale(h / 6.0, s));
      if (!P.active[idx % P.N]) a = make(0.0, 0.0);
      B.y[idx] = a;
      if (scored) sq = norm2(a);
    }
  }
  if constexpr (Pass > 1) {
    tree[threadIdx.x] = sq;
    __syncthreads();
    for (int half = kLsThreads / 2; half > 0; half /= 2) {
      if (threadIdx.x < half) tree[threadIdx.x] = tree[threadIdx.x] + tree[threadIdx.x + half];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[static_cast<long>(Pass - 2) * gridDim.x + blockIdx.x] = tree[0];
  }
}

// Above kLsRun workgroup sums a first level sums each run of kLsRun of them (workgroup
// (run, k)), in the same fixed order, into `out` (3 rows of `runs`).
__global__ void __launch_bounds__(kLsThreads)
    ls_runs(const double* __restrict__ partials, int blocks, double* out, int runs) {
  __shared__ double tree[kLsThreads];
  const int run = blockIdx.x, k = blockIdx.y;
  const int b0 = run * kLsRun, b1 = min(b0 + kLsRun, blocks);
  double total = 0.0;
  for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += kLsThreads)
    total = total + partials[static_cast<long>(k) * blocks + b];
  tree[threadIdx.x] = total;
  __syncthreads();
  for (int half = kLsThreads / 2; half > 0; half /= 2) {
    if (threadIdx.x < half) tree[threadIdx.x] = tree[threadIdx.x] + tree[threadIdx.x + half];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[static_cast<long>(k) * runs + run] = tree[0];
}

// The three sums from the passes' workgroup sums, in a fixed order, into `norms`.
__global__ void __launch_bounds__(kLsThreads)
    ls_norms(const double* __restrict__ partials, int blocks, double* norms) {
  __shared__ double tree[kLsThreads];
  for (int k = 0; k < 3; ++k) {
    double total = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kLsThreads) total = total + partials[k * blocks + b];
    tree[threadIdx.x] = total;
    __syncthreads();
    for (int half = kLsThreads / 2; half > 0; half /= 2) {
      if (threadIdx.x < half) tree[threadIdx.x] = tree[threadIdx.x] + tree[threadIdx.x + half];
      __syncthreads();
    }
    if (threadIdx.x == 0) norms[k] = tree[0];
    __syncthreads();
  }
}

// -------------------------------------------------------------------- launchers

// One wave per workgroup while that leaves at most 4096 workgroups (small grids are
// latency-bound: the most workgroups), else four, which keeps the block sums few.
unsigned correct_threads(const DeviceProblem& P) {
  const std::size_t n = static_cast<std::size_t>(P.C) * P.N;
  return n <= 4096UL * kWave ? kWave : kCorrectThreads;
}

unsigned correct_blocks(const DeviceProblem& P) {
  return blocks_for(static_cast<std::size_t>(P.C) * P.N, correct_threads(P));
}

namespace {
void check() { SP_HIP(hipGetLastError()); }
unsigned cn_blocks(const DeviceProblem& P) {
  return blocks_for(static_cast<std::size_t>(P.C) * P.N, kThreads);
}
}  // namespace

namespace {
template <bool Commit>
void launch_predict(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B, int width,
                    hipStream_t s) {
  const dim3 grid(blocks_for(static_cast<std::size_t>(P.C) * P.N, kWave)), block(kWave);
  switch (width) {
    case 2: hipLaunchKernelGGL((adams_predict<2, Commit>), grid, block, 0, s, P, R, B); break;
    case 3: hipLaunchKernelGGL((adams_predict<3, Commit>), grid, block, 0, s, P, R, B); break;
    case 4: hipLaunchKernelGGL((adams_predict<4, Commit>), grid, block, 0, s, P, R, B); break;
    case 5: hipLaunchKernelGGL((adams_predict<5, Commit>), grid, block, 0, s, P, R, B); break;
    case 6: hipLaunchKernelGGL((adams_predict<6, Commit>), grid, block, 0, s, P, R, B); break;
    default: hipLaunchKernelGGL((adams_predict<7, Commit>), grid, block, 0, s, P, R, B); break;
  }
  check();
}
}  // namespace

void launch_adams_predict(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                          int width, hipStream_t s, bool commit) {
  SP_CHECK(width >= 2 && width <= kMaxAdamsWidth - 1, "stencil width 2..7");
  if (commit)
    launch_predict<true>(P, R, B, width, s);
  else
    launch_predict<false>(P, R, B, width, s);
}

void launch_field_scatter(const DeviceProblem& P, const cplx* spectrum, FieldBuffers B,
                          hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * (P.T / 2);
  hipLaunchKernelGGL(field_scatter<false>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                     P, spectrum, B.modal_t, B.roots_T);
  check();
}

void launch_field_scatter_hilbert(const DeviceProblem& P, const cplx* spectrum, FieldBuffers B,
                                  hipStream_t s) {
  SP_CHECK(B.modal_h, "the Hilbert transforms' buffer");
  const std::size_t n = static_cast<std::size_t>(P.C) * (P.T / 2);
  hipLaunchKernelGGL(field_scatter<true>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                     P, spectrum, B.modal_h, B.roots_T);
  check();
}

namespace {
template <int C>
void launch_node_passes_for(const DeviceProblem& P, const AdamsRecord* R,
                            const PlasmaTable& pt, FieldBuffers B, hipStream_t s,
                            hipEvent_t* marks, bool exposure) {
  const int tiles = scan_tiles(P.T);
  const int per_block = kNodeWaves * kNodesPerWave;
  const dim3 node_grid(tiles, (P.Q + per_block - 1) / per_block);
  const dim3 node_block(kWave * kNodeWaves);
  auto mark = [&](int i) {
    if (marks) SP_HIP(hipEventRecord(marks[i], s));
  };
  const bool band = pt.n_taps > 1 || exposure;
  if (P.has_plasma) {
    SP_CHECK(!band || (B.x_local && pt.n_taps <= kMaxBandHalf + 1 && 64 + 2 * (pt.n_taps - 1) <= kBandSpan),
             "the band rule needs its exposure buffer and at most 8 taps each side");
    if (band)
      hipLaunchKernelGGL(plasma_rate_band<C>, dim3(tiles, rate_groups<C>(P.Q)), node_block, 0, s,
                         P, pt, B, 0);
    else
      hipLaunchKernelGGL(plasma_rate_tiles<C>, node_grid, node_block, 0, s, P, pt, B);
    check();
    mark(0);
    hipLaunchKernelGGL(plasma_tile_prefix, dim3(P.Q), dim3(kWave), 0, s, tiles, B.tile_x);
    check();
    mark(1);
    if (band)
      hipLaunchKernelGGL(plasma_current_band<C>, node_grid, node_block, 0, s, P, pt, R, B);
    else
      hipLaunchKernelGGL(plasma_current_tiles<C>, node_grid, node_block, 0, s, P, R, B);
    check();
    mark(2);
    hipLaunchKernelGGL(plasma_tile_prefix, dim3(P.Q), dim3(kWave), 0, s, tiles, B.tile_f);
    check();
    mark(3);
  } else {
    for (int i = 0; i < 4; ++i) mark(i);
  }
  if constexpr (C > 0) {
    SP_CHECK(P.T % kWave == 0, "the node response takes 64-sample blocks");
    hipLaunchKernelGGL(node_response<C>, dim3(P.T / kWave), dim3(kWave * kResponseWaves), 0,
                       s, P, R, B);
  } else {
    const std::size_t n = static_cast<std::size_t>(P.Q) * P.T;
    hipLaunchKernelGGL(node_response_dense, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0,
                       s, P, R, B);
  }
  check();
  mark(4);
}
}  // namespace

void launch_node_passes(const DeviceProblem& P, const AdamsRecord* R, const PlasmaTable& pt,
                        FieldBuffers B, hipStream_t s, hipEvent_t* marks, bool exposure) {
  SP_CHECK(P.T >= 2, "the fine grid needs two samples");
  if (B.node_field) {
    SP_CHECK(B.node_streams, "the dense route needs its node streams");
    launch_node_passes_for<0>(P, R, pt, B, s, marks, exposure);
    return;
  }
  SP_CHECK(P.C >= 1 && P.C <= kMaxModes, "the node passes project 1..8 modes");
  switch (P.C) {
    case 1: launch_node_passes_for<1>(P, R, pt, B, s, marks, exposure); break;
    case 2: launch_node_passes_for<2>(P, R, pt, B, s, marks, exposure); break;
    case 3: launch_node_passes_for<3>(P, R, pt, B, s, marks, exposure); break;
    case 4: launch_node_passes_for<4>(P, R, pt, B, s, marks, exposure); break;
    case 5: launch_node_passes_for<5>(P, R, pt, B, s, marks, exposure); break;
    case 6: launch_node_passes_for<6>(P, R, pt, B, s, marks, exposure); break;
    case 7: launch_node_passes_for<7>(P, R, pt, B, s, marks, exposure); break;
    default: launch_node_passes_for<8>(P, R, pt, B, s, marks, exposure); break;
  }
}

// ------------------------------------------ the plasma integration diagnostics

namespace {
constexpr int kDiagThreads = 256;
constexpr int kDiagParts = 6;

// One node's parts (see launch_plasma_diagnostics). The sum of the work integrand runs
// over each thread's samples in order, then the wave's fixed-order total, then the waves
// in order; minima and maxima need no order.
template <int C>
__global__ void __launch_bounds__(kDiagThreads)
    plasma_diagnose(DeviceProblem P, PlasmaTable pt, FieldBuffers B, double* parts) {
  __shared__ double lds[kDiagParts][kDiagThreads / kWave];
  const int q = blockIdx.x;
  const int T = P.T, tiles = scan_tiles(T), reach = 2 * (pt.n_taps - 1);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const long row = static_cast<long>(q) * T;
  const double* offsets = B.tile_x + static_cast<long>(q) * tiles;
  const double dt = P.dt_fine;
  auto exposure = [&](int t) { return offsets[min(t / kTileSpan, tiles - 1)] + B.x_local[row + t]; };
  auto work = [&](int t) { return B.rate[row + t] * (1.0 - B.frac[row + t]); };
  double lo = INFINITY, hi = -INFINITY, paid = 0.0, early = 0.0, peak = 0.0;
  for (int t = threadIdx.x; t < T; t += kDiagThreads) {
    if (t + 1 < T) {
      const double step = exposure(t + 1) - exposure(t);
      lo = fmin(lo, step);
      hi = fmax(hi, step);
      paid = paid + (dt * (work(t + 1) + work(t))) / 2.0;
    }
    double am[held(C)];
    load_modal<C>(B, T, t, am);
    const double level = fabs(B.frac[row + t] * field_at<C>(P, B, am, q, t));
    peak = fmax(peak, level);
    if (t < reach) early = fmax(early, level);
  }
  // Wave totals: the sum in a fixed order, the extremes by their own reductions.
  paid = wave_total(paid);
  for (int offset = kWave / 2; offset > 0; offset /= 2) {
    lo = fmin(lo, __shfl_xor(lo, offset));
    hi = fmax(hi, __shfl_xor(hi, offset));
    early = fmax(early, __shfl_xor(early, offset));
    peak = fmax(peak, __shfl_xor(peak, offset));
  }
  if (lane == 0) {
    lds[0][wave] = lo;
    lds[1][wave] = hi;
    lds[2][wave] = paid;
    lds[3][wave] = early;
    lds[4][wave] = peak;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double l = lds[0][0], h = lds[1][0], w = lds[2][0], e = lds[3][0], k = lds[4][0];
    for (int v = 1; v < kDiagThreads / kWave; ++v) {
      l = fmin(l, lds[0][v]);
      h = fmax(h, lds[1][v]);
      w = w + lds[2][v];
      e = fmax(e, lds[3][v]);
      k = fmax(k, lds[4][v]);
    }
    const double freed = B.frac[row + T - 1] - B.frac[row];
    double* out = parts + static_cast<long>(q) * kDiagParts;
    out[0] = fmax(-l, 0.0);
    out[1] = h;
    out[2] = fabs(w - freed);
    out[3] = freed;
    out[4] = e;
    out[5] = k;
  }
}

// The largest of each part over the nodes.
__global__ void __launch_bounds__(kWave)
    plasma_diagnose_max(const double* __restrict__ parts, int Q, double* out) {
  const int part = threadIdx.x;
  if (part >= kDiagParts) return;
  double largest = parts[part];
  for (int q = 1; q < Q; ++q) largest = fmax(largest, parts[static_cast<long>(q) * kDiagParts + part]);
  out[part] = largest;
}

template <int C>
void launch_diagnose_for(const DeviceProblem& P, const PlasmaTable& pt, FieldBuffers B,
                         double* parts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(plasma_diagnose<C>, dim3(P.Q), dim3(kDiagThreads), 0, s, P, pt, B, parts);
  SP_HIP(hipGetLastError());
  hipLaunchKernelGGL(plasma_diagnose_max, dim3(1), dim3(kWave), 0, s, parts, P.Q, out);
  SP_HIP(hipGetLastError());
}
}  // namespace

void launch_plasma_diagnostics(const DeviceProblem& P, const PlasmaTable& pt, FieldBuffers B,
                               double* parts, double* out, hipStream_t s) {
  SP_CHECK(P.has_plasma && B.x_local && B.rate && B.frac && B.tile_x && P.T >= 2 && P.Q >= 1,
           "the integration diagnostics read a stored plasma column with its exposure");
  if (B.node_field) {
    launch_diagnose_for<0>(P, pt, B, parts, out, s);
    return;
  }
  switch (P.C) {
    case 1: launch_diagnose_for<1>(P, pt, B, parts, out, s); break;
    case 2: launch_diagnose_for<2>(P, pt, B, parts, out, s); break;
    case 3: launch_diagnose_for<3>(P, pt, B, parts, out, s); break;
    case 4: launch_diagnose_for<4>(P, pt, B, parts, out, s); break;
    case 5: launch_diagnose_for<5>(P, pt, B, parts, out, s); break;
    case 6: launch_diagnose_for<6>(P, pt, B, parts, out, s); break;
    case 7: launch_diagnose_for<7>(P, pt, B, parts, out, s); break;
    default: launch_diagnose_for<8>(P, pt, B, parts, out, s); break;
  }
}

namespace {
template <int C>
void launch_node_passes_fused_for(const DeviceProblem& P, const AdamsRecord* R,
                                  const PlasmaTable& pt, FieldBuffers B, hipStream_t s,
                                  hipEvent_t* marks) {
  const int tiles = scan_tiles(P.T);
  const dim3 node_grid(tiles, node_groups(P.Q));
  const dim3 node_block(kWave * kNodeWaves);
  auto mark = [&](int i) {
    if (marks) SP_HIP(hipEventRecord(marks[i], s));
  };
  hipLaunchKernelGGL(plasma_rate_band<C>, dim3(tiles, rate_groups<C>(P.Q)), node_block, 0, s, P,
                     pt, B, 1);
  check();
  mark(0);
  hipLaunchKernelGGL(plasma_tile_prefix, dim3(P.Q), dim3(kWave), 0, s, tiles, B.tile_x);
  check();
  if (marks) {
    // The dominant pass is timed by its own start and end rather than by events recorded
    // between launches (which add their own latency to the interval).
    hipExtLaunchKernelGGL(plasma_current_response<C>, node_grid, node_block, 0, s, marks[1],
                          marks[2], 0, P, pt, R, B);
  } else {
    hipLaunchKernelGGL(plasma_current_response<C>, node_grid, node_block, 0, s, P, pt, R, B);
  }
  check();
  hipLaunchKernelGGL(plasma_response_combine<C>, dim3(P.T / kWave), dim3(kWave * 4), 0, s, P,
                     R, B);
  check();
  mark(3);
  mark(4);
}
}  // namespace

void launch_node_passes_fused(const DeviceProblem& P, const AdamsRecord* R,
                              const PlasmaTable& pt, FieldBuffers B, hipStream_t s,
                              hipEvent_t* marks) {
  SP_CHECK(fused_node_passes(P, B), "the fused node passes need plasma, 1..8 modes, the "
                                    "exposure buffer and the group streams");
  SP_CHECK(P.T >= 2 && P.T % kWave == 0, "the fused node passes take 64-sample blocks");
  SP_CHECK(pt.n_taps >= 1 && pt.n_taps <= kMaxBandHalf + 1,
           "the running-integral rule has 1 to 9 taps");
  switch (P.C) {
    case 1: launch_node_passes_fused_for<1>(P, R, pt, B, s, marks); break;
    case 2: launch_node_passes_fused_for<2>(P, R, pt, B, s, marks); break;
    case 3: launch_node_passes_fused_for<3>(P, R, pt, B, s, marks); break;
    case 4: launch_node_passes_fused_for<4>(P, R, pt, B, s, marks); break;
    case 5: launch_node_passes_fused_for<5>(P, R, pt, B, s, marks); break;
    case 6: launch_node_passes_fused_for<6>(P, R, pt, B, s, marks); break;
    case 7: launch_node_passes_fused_for<7>(P, R, pt, B, s, marks); break;
    default: launch_node_passes_fused_for<8>(P, R, pt, B, s, marks); break;
  }
}

void launch_field_absorber_stream(const DeviceProblem& P, FieldBuffers B, bool accumulate,
                                  hipStream_t s) {
  SP_CHECK(P.absorber_rate && P.absorber_stream >= 1 && P.absorber_stream <= 3,
           "the in-stage absorber's rate and stream");
  const std::size_t n = static_cast<std::size_t>(P.C) * P.T;
  hipLaunchKernelGGL(field_absorber_stream, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0,
                     s, P, B, accumulate);
  check();
}

void launch_field_split_pack(const DeviceProblem& P, FieldBuffers B, hipStream_t s) {
  launch_field_scatter(P, B.y, B, s);
}

void launch_field_split_collar(const DeviceProblem& P, const AdamsRecord* R, double dz,
                               FieldBuffers B, hipStream_t s) {
  SP_CHECK(P.absorber_rate, "the split-step absorber's rate");
  const std::size_t n = static_cast<std::size_t>(P.C) * P.T;
  hipLaunchKernelGGL(field_split_collar, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                     P, R, dz, B);
  check();
}

void launch_field_split_unpack(const DeviceProblem& P, FieldBuffers B, hipStream_t s) {
  SP_CHECK(P.N == P.T / 2 + 1, "the split-step absorber needs the coarse grid to be the fine one");
  hipLaunchKernelGGL(field_split_unpack, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, B);
  check();
}

void launch_field_raman_kernel_fill(const DeviceProblem& P, const AdamsRecord* R, cplx* z,
                                    hipStream_t s) {
  hipLaunchKernelGGL(field_raman_kernel_fill, dim3(blocks_for(P.T, kThreads)), dim3(kThreads),
                     0, s, P, R, z);
  check();
}

void launch_field_raman_drive(const DeviceProblem& P, FieldBuffers B, hipStream_t s) {
  SP_CHECK(B.raman_drive, "the Raman drive buffer");
  const dim3 grid(blocks_for(static_cast<std::size_t>(P.Q) * P.T, kThreads)), block(kThreads);
  if (B.node_field) {
    hipLaunchKernelGGL(field_raman_drive<0>, grid, block, 0, s, P, B);
  } else {
    switch (P.C) {
      case 1: hipLaunchKernelGGL(field_raman_drive<1>, grid, block, 0, s, P, B); break;
      case 2: hipLaunchKernelGGL(field_raman_drive<2>, grid, block, 0, s, P, B); break;
      case 3: hipLaunchKernelGGL(field_raman_drive<3>, grid, block, 0, s, P, B); break;
      case 4: hipLaunchKernelGGL(field_raman_drive<4>, grid, block, 0, s, P, B); break;
      case 5: hipLaunchKernelGGL(field_raman_drive<5>, grid, block, 0, s, P, B); break;
      case 6: hipLaunchKernelGGL(field_raman_drive<6>, grid, block, 0, s, P, B); break;
      case 7: hipLaunchKernelGGL(field_raman_drive<7>, grid, block, 0, s, P, B); break;
      default: hipLaunchKernelGGL(field_raman_drive<8>, grid, block, 0, s, P, B); break;
    }
  }
  check();
}

void launch_field_finish(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                         cplx* out, hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * (P.T / 4 + 1);
  SP_CHECK(P.N <= P.T / 2 + 1, "the coarse grid lies within the fine one");
  hipLaunchKernelGGL(field_finish, dim3(blocks_for(n, kWave)), dim3(kWave), 0, s, P, R, B,
                     out);
  check();
}

namespace {
template <bool Finish>
void launch_correct(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B, int width,
                    hipStream_t s) {
  const dim3 grid(correct_blocks(P)), block(correct_threads(P));
  switch (width) {
    case 2: hipLaunchKernelGGL((adams_correct<2, Finish>), grid, block, 0, s, P, R, B); break;
    case 3: hipLaunchKernelGGL((adams_correct<3, Finish>), grid, block, 0, s, P, R, B); break;
    case 4: hipLaunchKernelGGL((adams_correct<4, Finish>), grid, block, 0, s, P, R, B); break;
    case 5: hipLaunchKernelGGL((adams_correct<5, Finish>), grid, block, 0, s, P, R, B); break;
    case 6: hipLaunchKernelGGL((adams_correct<6, Finish>), grid, block, 0, s, P, R, B); break;
    default: hipLaunchKernelGGL((adams_correct<7, Finish>), grid, block, 0, s, P, R, B); break;
  }
  check();
}
}  // namespace

void launch_adams_correct(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                          int width, hipStream_t s, bool finish) {
  SP_CHECK(width >= 2 && width <= kMaxAdamsWidth - 1, "stencil width 2..7");
  if (finish)
    launch_correct<true>(P, R, B, width, s);
  else
    launch_correct<false>(P, R, B, width, s);
}

void launch_adams_norms(const DeviceProblem& P, FieldBuffers B, hipStream_t s) {
  hipLaunchKernelGGL(adams_norms, dim3(1), dim3(kNormThreads), 0, s, B.partials,
                     static_cast<int>(correct_blocks(P)), B.norms);
  check();
}

void launch_adams_accept(const DeviceProblem& P, int slot, bool compensated, FieldBuffers B,
                         hipStream_t s) {
  hipLaunchKernelGGL(adams_accept, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, slot,
                     compensated ? 1 : 0, B);
  check();
}

void launch_adams_commit(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                         hipStream_t s) {
  hipLaunchKernelGGL(adams_commit, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, R, B);
  check();
}

namespace {
// The record's upload as a kernel node: each thread copies 8-byte words of the pinned
// host record, read at system scope (no cache can hold an earlier launch's record), into
// device memory. A copy node would be a blit launch with a wait before the next kernel.
constexpr int kFetchThreads = 512;
static_assert(sizeof(AdamsRecord) % sizeof(unsigned long long) == 0,
              "the record is copied in 8-byte words");

__global__ void __launch_bounds__(kFetchThreads)
    record_fetch(const unsigned long long* host, unsigned long long* device, int words) {
  for (int i = threadIdx.x; i < words; i += kFetchThreads)
    device[i] = __hip_atomic_load(host + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

namespace {
constexpr int kScoreThreads = 256;

// Each thread sums its strided bins in order, then a fixed pairwise tree in LDS.
__global__ void __launch_bounds__(kScoreThreads)
    quadrature_score(const cplx* __restrict__ fine, const cplx* __restrict__ coarse, long n,
                     double* out) {
  __shared__ double diff[kScoreThreads], base[kScoreThreads];
  const int tid = threadIdx.x;
  double d = 0.0, f = 0.0;
  for (long i = tid; i < n; i += kScoreThreads) {
    const cplx a = fine[i], b = coarse[i];
    const double dx = b.x - a.x, dy = b.y - a.y;
    d = d + (dx * dx + dy * dy);
    f = f + (a.x * a.x + a.y * a.y);
  }
  diff[tid] = d;
  base[tid] = f;
  __syncthreads();
  for (int half = kScoreThreads / 2; half > 0; half /= 2) {
    if (tid < half) {
      diff[tid] = diff[tid] + diff[tid + half];
      base[tid] = base[tid] + base[tid + half];
    }
    __syncthreads();
  }
  if (tid == 0) out[0] = base[0] != 0.0 ? sqrt(diff[0] / base[0]) : __builtin_nan("");
}
}  // namespace

void launch_quadrature_score(const cplx* fine, const cplx* coarse, long n, double* out,
                             hipStream_t s) {
  hipLaunchKernelGGL(quadrature_score, dim3(1), dim3(kScoreThreads), 0, s, fine, coarse, n,
                     out);
  check();
}

void launch_record_fetch(const AdamsRecord* host, AdamsRecord* device, int count,
                         hipStream_t s) {
  const int words = static_cast<int>(count * sizeof(AdamsRecord) / sizeof(unsigned long long));
  hipLaunchKernelGGL(record_fetch, dim3(1), dim3(kFetchThreads), 0, s,
                     reinterpret_cast<const unsigned long long*>(host),
                     reinterpret_cast<unsigned long long*>(device), words);
  check();
}

void launch_adams_start(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                        hipStream_t s) {
  hipLaunchKernelGGL(adams_start, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, R, B);
  check();
}

}  // namespace sp

namespace sp {

namespace {
template <typename K>
void grouped(K kernel, const BatchGrid& g, unsigned threads, hipStream_t s) {
  if (g.blocks <= 0) return;
  hipLaunchKernelGGL(kernel, dim3(g.blocks), dim3(threads), 0, s, g.inst, g.records, g.first,
                     g.n);
  SP_HIP(hipGetLastError());
}
}  // namespace

void launch_batch_commit(const BatchGrid& g, hipStream_t s) { grouped(batch_commit, g, kThreads, s); }
void launch_batch_predict(const BatchGrid& g, hipStream_t s) { grouped(batch_predict, g, kWave, s); }
void launch_batch_rate(const BatchGrid& g, hipStream_t s) {
  grouped(batch_rate, g, kWave * kNodeWaves, s);
}
void launch_batch_prefix(const BatchGrid& g, bool current, hipStream_t s) {
  if (g.blocks <= 0) return;
  hipLaunchKernelGGL(batch_prefix, dim3(g.blocks), dim3(kWave), 0, s, g.inst, g.records,
                     g.first, g.n, current ? 1 : 0);
  SP_HIP(hipGetLastError());
}
void launch_batch_current(const BatchGrid& g, hipStream_t s) {
  grouped(batch_current, g, kWave * kNodeWaves, s);
}
void launch_batch_response(const BatchGrid& g, hipStream_t s) {
  grouped(batch_response, g, kWave * kResponseWaves, s);
}
void launch_batch_finish(const BatchGrid& g, hipStream_t s) { grouped(batch_finish, g, kWave, s); }
void launch_batch_correct(const BatchGrid& g, hipStream_t s) { grouped(batch_correct, g, kWave, s); }
void launch_batch_norms(const BatchGrid& g, hipStream_t s) { grouped(batch_norms, g, kNormThreads, s); }

unsigned ls_blocks(const DeviceProblem& P) {
  return blocks_for(static_cast<std::size_t>(P.C) * P.N, kLsThreads);
}

std::size_t ls_partials(const DeviceProblem& P) {
  const std::size_t blocks = ls_blocks(P);
  const std::size_t runs = (blocks + kLsRun - 1) / kLsRun;
  return 3 * blocks + (blocks > kLsRun ? 3 * runs : 0);
}

namespace {
template <bool Finish>
void launch_ls_pass_for(int pass, const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                        cplx* acc, cplx* prev, double* partials, hipStream_t s) {
  const dim3 grid(ls_blocks(P)), block(kLsThreads);
  switch (pass) {
    case 1: hipLaunchKernelGGL((ls_pass<1, Finish>), grid, block, 0, s, P, R, B, acc, prev, partials); break;
    case 2: hipLaunchKernelGGL((ls_pass<2, Finish>), grid, block, 0, s, P, R, B, acc, prev, partials); break;
    case 3: hipLaunchKernelGGL((ls_pass<3, Finish>), grid, block, 0, s, P, R, B, acc, prev, partials); break;
    default: hipLaunchKernelGGL((ls_pass<4, Finish>), grid, block, 0, s, P, R, B, acc, prev, partials); break;
  }
  SP_HIP(hipGetLastError());
}
}  // namespace

void launch_ls_pass(int pass, const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                    cplx* acc, cplx* prev, double* partials, hipStream_t s, bool finish) {
  if (finish)
    launch_ls_pass_for<true>(pass, P, R, B, acc, prev, partials, s);
  else
    launch_ls_pass_for<false>(pass, P, R, B, acc, prev, partials, s);
}

void launch_ls_norms(const DeviceProblem& P, const double* partials, double* norms,
                     hipStream_t s) {
  const int blocks = static_cast<int>(ls_blocks(P));
  if (blocks <= kLsRun) {
    hipLaunchKernelGGL(ls_norms, dim3(1), dim3(kLsThreads), 0, s, partials, blocks, norms);
  } else {
    // Two levels: a million workgroup sums would take one workgroup milliseconds.
    const int runs = (blocks + kLsRun - 1) / kLsRun;
    double* level = const_cast<double*>(partials) + 3L * blocks;
    hipLaunchKernelGGL(ls_runs, dim3(runs, 3), dim3(kLsThreads), 0, s, partials, blocks, level,
                       runs);
    check();
    hipLaunchKernelGGL(ls_norms, dim3(1), dim3(kLsThreads), 0, s, level, runs, norms);
  }
  SP_HIP(hipGetLastError());
}

}  // namespace sp
