// The modal field-grid route's kernels for gfx950, FP64 (see field.hpp).
#include "field.hpp"

#include <cmath>

namespace sp {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxModes = 8;
constexpr double kRecipFactorial[10] = {1.0,          1.0,          0.5,
                                        1.0 / 6.0,    1.0 / 24.0,   1.0 / 120.0,
                                        1.0 / 720.0,  1.0 / 5040.0, 1.0 / 40320.0,
                                        1.0 / 362880.0};
constexpr double kFactorial[9] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320};

__device__ inline cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm, as NumPy divides complex numbers.
  if (fabs(b.x) >= fabs(b.y)) {
    const double rat = b.y / b.x;
    const double scl = 1.0 / (b.x + b.y * rat);
    return make((a.x + a.y * rat) * scl, (a.y - a.x * rat) * scl);
  }
  const double rat = b.x / b.y;
  const double scl = 1.0 / (b.y + b.x * rat);
  return make((a.x * rat + a.y) * scl, (a.y * rat - a.x) * scl);
}

// phi_0..phi_order at one argument (src/solprop/backends/exponential.py:131-226):
// scaled Taylor series, downward recurrence and doubling near zero; upward recurrence
// from exp(x) beyond |x| = 3.
__device__ void phi_functions(cplx x, int order, cplx* out) {
  const double magnitude = hypot(x.x, x.y);
  if (magnitude > 3.0) {
    out[0] = cexp(x);
    for (int k = 1; k <= order; ++k)
      out[k] = cdiv(sub(out[k - 1], make(kRecipFactorial[k - 1], 0.0)), x);
    return;
  }
  int squarings = 0;
  if (magnitude > 0.5) squarings = static_cast<int>(ceil(log2(fmax(magnitude, 0.5) / 0.5)));
  const double factor = exp2(static_cast<double>(squarings));
  const cplx y = make(x.x / factor, x.y / factor);
  cplx term = make(kRecipFactorial[order], 0.0);
  cplx total = term;
  for (int m = 1; m < 24; ++m) {
    const cplx ty = mul(term, y);
    const double d = static_cast<double>(m + order);
    term = make(ty.x / d, ty.y / d);
    total = add(total, term);
  }
  out[order] = total;
  for (int k = order - 1; k >= 1; --k)
    out[k] = add(make(kRecipFactorial[k], 0.0), mul(y, out[k + 1]));
  out[0] = cexp(y);
  for (int level = 0; level < squarings; ++level) {
    cplx doubled[kMaxAdamsWidth + 1];
    doubled[0] = mul(out[0], out[0]);
    for (int k = 1; k <= order; ++k) {
      cplx t = mul(out[0], out[k]);
      for (int j = 1; j <= k; ++j) t = add(t, scale(kRecipFactorial[k - j], out[j]));
      const double p2 = exp2(static_cast<double>(k));
      doubled[k] = make(t.x / p2, t.y / p2);
    }
    for (int k = 0; k <= order; ++k) out[k] = doubled[k];
  }
}

__device__ inline cplx weighted(const double* w, const cplx* terms, int n, long stride,
                                long at) {
  cplx total = make(0.0, 0.0);
  for (int t = 0; t < n; ++t) total = add(total, scale(w[t], terms[t * stride + at]));
  return total;
}

// dz Σ_k M_k (Σ_j c_jk X_j) over n_k columns and n_j rows (coefficients [j][k]).
__device__ inline cplx apply_moments(const cplx* M, const double (*c)[kMaxAdamsWidth],
                                     const cplx* X, int n_j, int n_k, double dz) {
  cplx total = make(0.0, 0.0);
  for (int k = 0; k < n_k; ++k) {
    cplx combined = make(0.0, 0.0);
    for (int j = 0; j < n_j; ++j) {
      const double w = c[j][k];
      if (w != 0.0) combined = add(combined, scale(w, X[j]));
    }
    total = add(total, mul(M[k], combined));
  }
  return scale(dz, total);
}

__device__ inline cplx r2c_post(cplx za, cplx zb, cplx w) {
  const cplx e = scale(0.5, add(za, conj(zb)));
  const cplx o = scale(0.5, rot(sub(za, conj(zb)), -1));
  return add(e, mul(w, o));
}

__device__ inline cplx c2r_pre(cplx ya, cplx yb, cplx w) {
  const cplx s = add(ya, conj(yb));
  const cplx d = sub(ya, conj(yb));
  return add(s, rot(mul(conj(w), d), +1));
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

// Scan state of one workgroup of kScanTile threads (two waves).
struct ScanLds {
  double waves[2];
  double incl[kScanTile];
  double edge[2][2];  // the first lane's values of each wave, for the neighbour exchange
};

// Inclusive scan in a fixed order (a shuffle ladder in each wave, then the wave
// totals); the exclusive value is the previous thread's inclusive one, bit for bit.
__device__ inline double tile_scan(double v, ScanLds& s, double* exclusive) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = 1; d < 64; d <<= 1) {
    const double n = __shfl_up(v, d, 64);
    if (lane >= d) v = v + n;
  }
  if (lane == 63) s.waves[wave] = v;
  __syncthreads();
  double offset = 0.0;
  for (int w = 0; w < wave; ++w) offset = offset + s.waves[w];
  const double incl = wave ? offset + v : v;
  s.incl[threadIdx.x] = incl;
  __syncthreads();
  *exclusive = threadIdx.x ? s.incl[threadIdx.x - 1] : 0.0;
  __syncthreads();
  return incl;
}

// Two values of thread + 1 within the tile; the tile's last thread takes `tail_*`.
__device__ inline void next_of(double a, double b, ScanLds& s, double tail_a, double tail_b,
                               double* na, double* nb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double xa = __shfl_down(a, 1, 64), xb = __shfl_down(b, 1, 64);
  if (lane == 0) {
    s.edge[wave][0] = a;
    s.edge[wave][1] = b;
  }
  __syncthreads();
  if (lane == 63) {
    if (threadIdx.x == kScanTile - 1) {
      xa = tail_a;
      xb = tail_b;
    } else {
      xa = s.edge[wave + 1][0];
      xb = s.edge[wave + 1][1];
    }
  }
  __syncthreads();
  *na = xa;
  *nb = xb;
}

__device__ inline double node_field(const double* S, const double* a, int C, int T, int t) {
  double e = 0.0;
  for (int m = 0; m < C; ++m) e = e + S[m] * a[static_cast<long>(m) * T + t];
  return e;
}

}  // namespace

// ------------------------------------------------------------- Adams predictor

__global__ void adams_predict(DeviceProblem P, const AdamsRecord* R, FieldBuffers B) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long CN = static_cast<long>(P.C) * P.N;
  if (idx >= CN) return;
  const AdamsRecord& r = *R;
  const double dz = r.dz;
  const cplx x = weighted(r.x_w, P.phi, P.n_phi, CN, idx);
  cplx phi[kMaxAdamsWidth + 1];
  phi_functions(x, r.width, phi);
  cplx M[kMaxAdamsWidth];
  for (int k = 0; k < r.width; ++k) {
    M[k] = scale(kFactorial[k], phi[k + 1]);
    B.moments[k * CN + idx] = M[k];
  }
  const cplx flow = cexp(x);
  B.flow[idx] = flow;
  cplx F[kMaxAdamsWidth];
  for (int j = 0; j < r.order; ++j) F[j] = B.store[r.slots[j] * CN + idx];
  cplx lab = add(mul(flow, B.y[idx]), apply_moments(M, r.pred_c, F, r.order, r.order, dz));
  if (r.compensated) {
    const cplx reference = make(x.x / dz, x.y / dz);
    cplx carried[kMaxAdamsWidth];
    for (int j = 0; j < r.n_read; ++j) {
      const int slot = r.slots[j];
      const cplx rate = weighted(r.slot_rate_w[j], P.phi, P.n_phi, CN, idx);
      const cplx term = mul(sub(rate, reference), B.states[slot * CN + idx]);
      const cplx span = scale(1.0 - r.nodes[j], x);
      carried[j] = mul(cexp(span), term);
      B.carried[j * CN + idx] = carried[j];
    }
    cplx total = make(0.0, 0.0);
    for (int j = 0; j < r.order; ++j) total = add(total, scale(r.comp_pred_w[0][j], carried[j]));
    lab = add(lab, scale(dz, total));
  }
  B.lab[idx] = lab;
}

// ---------------------------------------------------- the source's transforms

// The lab-frame modal spectrum (1/T folded in) packed for the real inverse transform
// of T points as a complex one of T/2.
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
  modal_t[idx] = c2r_pre(xa, xb, roots_T[k]);
}

// Pass A: the exposure's trapezoid increments over each tile of each node; the tile's
// total, its last inclusive value.
__global__ void __launch_bounds__(kScanTile)
    plasma_exposure_tiles(DeviceProblem P, PlasmaTable pt, FieldBuffers B) {
  __shared__ ScanLds lds;
  const int T = P.T, tiles = T / kScanTile;
  const int tile = blockIdx.x, q0 = blockIdx.y * kNodeBlock;
  const int t = tile * kScanTile + threadIdx.x;
  const bool last = threadIdx.x == kScanTile - 1;
  const double* a = reinterpret_cast<const double*>(B.modal_t);
  const double dt = P.dt_fine;
  for (int q = q0; q < q0 + kNodeBlock && q < P.Q; ++q) {
    const double* S = P.synthesis + static_cast<long>(q) * P.C;
    const double e = node_field(S, a, P.C, T, t);
    const double w = rate_of(fabs(e), pt, B.rate_c);
    double tail = 0.0;
    if (last && t + 1 < T) tail = rate_of(fabs(node_field(S, a, P.C, T, t + 1)), pt, B.rate_c);
    double w_next, unused;
    next_of(w, 0.0, lds, tail, 0.0, &w_next, &unused);
    const double inc = t + 1 < T ? (dt * (w_next + w)) / 2.0 : 0.0;
    double excl;
    const double incl = tile_scan(inc, lds, &excl);
    if (last) B.tile_x[static_cast<long>(q) * tiles + tile] = incl;
  }
}

// Passes B and D: exclusive prefix of the tile sums of each node, in tile order.
__global__ void plasma_tile_prefix(int Q, int tiles, double* sums) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  double running = 0.0;
  double* row = sums + static_cast<long>(q) * tiles;
  for (int k = 0; k < tiles; ++k) {
    const double s = row[k];
    row[k] = running;
    running = running + s;
  }
}

// Pass C (kFinal false): the free-current increments' tile sums. Pass E (kFinal true):
// the whole column -- Kerr, fraction, free and loss currents, the collars -- projected
// onto the modes over this workgroup's block of nodes.
template <bool kFinal>
__global__ void __launch_bounds__(kScanTile)
    node_response(DeviceProblem P, const AdamsRecord* R, PlasmaTable pt, FieldBuffers B) {
  __shared__ ScanLds lds;
  const int T = P.T, tiles = T / kScanTile;
  const int tile = blockIdx.x, q0 = blockIdx.y * kNodeBlock;
  const int t = tile * kScanTile + threadIdx.x;
  const bool last = threadIdx.x == kScanTile - 1;
  const double* a = reinterpret_cast<const double*>(B.modal_t);
  const double dt = P.dt_fine;
  const double density = R->density, f0 = R->preionisation;
  const double g = P.t_window[t], gd = P.t_deriv[t];
  double acc[3][kMaxModes];
  for (int s = 0; s < 3; ++s)
    for (int m = 0; m < kMaxModes; ++m) acc[s][m] = 0.0;
  for (int q = q0; q < q0 + kNodeBlock && q < P.Q; ++q) {
    const double* S = P.synthesis + static_cast<long>(q) * P.C;
    const double e = node_field(S, a, P.C, T, t);
    double w = 0.0, f = f0, F = 0.0;
    if (P.has_plasma) {
      w = rate_of(fabs(e), pt, B.rate_c);
      double tail_e = 0.0, tail_w = 0.0;
      if (last && t + 1 < T) {
        tail_e = node_field(S, a, P.C, T, t + 1);
        tail_w = rate_of(fabs(tail_e), pt, B.rate_c);
      }
      double e_next, w_next;
      next_of(e, w, lds, tail_e, tail_w, &e_next, &w_next);
      const double inc = t + 1 < T ? (dt * (w_next + w)) / 2.0 : 0.0;
      double excl;
      const double incl = tile_scan(inc, lds, &excl);
      const double off = B.tile_x[static_cast<long>(q) * tiles + tile];
      const double x = off + excl;
      const double x_next = off + incl;
      f = fraction_of(f0, x);
      const double f_next = fraction_of(f0, x_next);
      const double gval = (f * P.plasma_cmr) * e;
      const double g_next = (f_next * P.plasma_cmr) * e_next;
      const double jinc = t + 1 < T ? (dt * (g_next + gval)) / 2.0 : 0.0;
      double jexcl;
      const double jincl = tile_scan(jinc, lds, &jexcl);
      if (!kFinal) {
        if (last) B.tile_f[static_cast<long>(q) * tiles + tile] = jincl;
        continue;
      }
      F = B.tile_f[static_cast<long>(q) * tiles + tile] + jexcl;
    }
    if (!kFinal) continue;
    double pol = 0.0;
    if (P.has_kerr) {
      const double factor = P.kerr_coefficient * density;
      pol = (factor * (e * e)) * e;
    }
    pol = pol * g;
    double current = 0.0;
    if (P.has_plasma) {
      const double squared = fabs(e) * fabs(e);
      double loss = 0.0;
      if (w > 0.0 && squared > 0.0) loss = (P.plasma_ip * w * (1.0 - f)) / squared;
      current = density * (F + loss * e);
    }
    const double gated = g * current, edge = gd * current;
    for (int m = 0; m < P.C; ++m) {
      const double wmq = P.projection[static_cast<long>(m) * P.Q + q];
      acc[0][m] = acc[0][m] + wmq * pol;
      acc[1][m] = acc[1][m] + wmq * gated;
      acc[2][m] = acc[2][m] + wmq * edge;
    }
  }
  if (kFinal) {
    const long block = blockIdx.y;
    for (int s = 0; s < 3; ++s)
      for (int m = 0; m < P.C; ++m)
        B.partial[((block * 3 + s) * P.C + m) * static_cast<long>(T) + t] = acc[s][m];
  }
}

// Sum the node blocks' partial streams in block order into the packed stream rows.
__global__ void stream_reduce(DeviceProblem P, FieldBuffers B, int n_blocks) {
  const long n = 3L * P.C * P.T;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  double v = 0.0;
  for (int b = 0; b < n_blocks; ++b) v = v + B.partial[b * n + idx];
  reinterpret_cast<double*>(B.streams)[idx] = v;
}

// Unpack the three stream spectra, assemble the held current on the active bins and
// normalise: k = (p W) (P~ + G~/(iω) - E~/(iω)^2).
__global__ void field_finish(DeviceProblem P, const AdamsRecord* R, FieldBuffers B, cplx* out) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= CN) return;
  const int c = static_cast<int>(idx / P.N), b = static_cast<int>(idx % P.N);
  const int half = P.T / 2;
  const int k = P.crop[b];
  const int ka = k % half, kb = (half - k) % half;
  cplx X[3];
  for (int s = 0; s < 3; ++s) {
    const cplx* row = B.streams + (static_cast<long>(s) * P.C + c) * half;
    X[s] = r2c_post(row[ka], row[kb], B.roots_T[k]);
  }
  cplx spectrum = X[0];
  if (P.has_plasma && P.active[b]) {
    const cplx q = make(0.0, P.omega[b]);
    const cplx q2 = mul(q, q);
    spectrum = add(spectrum, sub(cdiv(X[1], q), cdiv(X[2], q2)));
  }
  const cplx pref = weighted(R->pref_w, P.pref, P.n_pref, CN, idx);
  out[idx] = mul(scale(P.omega_window[b], pref), spectrum);
}

// ------------------------------------------------------------- Adams corrector

__global__ void __launch_bounds__(kThreads)
    adams_correct(DeviceProblem P, const AdamsRecord* R, FieldBuffers B) {
  __shared__ double red[1 + kMaxScored][kThreads];
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const AdamsRecord& r = *R;
  double v[1 + kMaxScored] = {0.0, 0.0, 0.0, 0.0};
  if (idx < CN) {
    const int b = static_cast<int>(idx % P.N);
    const double dz = r.dz;
    cplx M[kMaxAdamsWidth], X[kMaxAdamsWidth];
    for (int k = 0; k < r.width; ++k) M[k] = B.moments[k * CN + idx];
    X[0] = B.source[idx];
    for (int j = 0; j < r.width - 1; ++j) X[1 + j] = B.store[r.slots[j] * CN + idx];
    const cplx y = B.y[idx];
    cplx adv = add(mul(B.flow[idx], y), apply_moments(M, r.corr_c, X, r.width, r.width, dz));
    cplx comp_def[kMaxScored] = {};
    if (r.compensated) {
      const cplx x = weighted(r.x_w, P.phi, P.n_phi, CN, idx);
      const cplx reference = make(x.x / dz, x.y / dz);
      const cplx rate_end = weighted(r.rate_end_w, P.phi, P.n_phi, CN, idx);
      const cplx endpoint = mul(sub(rate_end, reference), B.lab[idx]);
      cplx carried[kMaxAdamsWidth];
      for (int j = 0; j < r.n_read; ++j) carried[j] = B.carried[j * CN + idx];
      for (int i = 0; i < r.n_orders; ++i) {
        const int k = r.orders[i];
        cplx corr = make(0.0, 0.0), pred = make(0.0, 0.0);
        for (int j = 0; j < k; ++j) corr = add(corr, scale(r.comp_corr_w[i][1 + j], carried[j]));
        corr = scale(dz, add(corr, scale(r.comp_corr_w[i][0], endpoint)));
        for (int j = 0; j < k; ++j) pred = add(pred, scale(r.comp_pred_w[i][j], carried[j]));
        pred = scale(dz, pred);
        comp_def[i] = sub(corr, pred);
        if (i == 0) adv = add(adv, corr);
      }
    }
    B.advanced[idx] = adv;
    if (P.normmask[b]) {
      v[0] = norm2(y);
      for (int i = 0; i < r.n_orders; ++i) {
        cplx d = apply_moments(M, r.rows_c[i], X, r.width, r.width, dz);
        if (r.compensated) d = add(d, comp_def[i]);
        v[1 + i] = norm2(d);
      }
    }
  }
  const int tid = threadIdx.x;
  for (int s = 0; s < 1 + kMaxScored; ++s) red[s][tid] = v[s];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (tid < w)
      for (int s = 0; s < 1 + kMaxScored; ++s) red[s][tid] = red[s][tid] + red[s][tid + w];
    __syncthreads();
  }
  if (tid == 0)
    for (int s = 0; s < 1 + kMaxScored; ++s) B.partials[s * gridDim.x + blockIdx.x] = red[s][0];
}

__global__ void adams_accept(DeviceProblem P, const AdamsRecord* R, FieldBuffers B) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= CN) return;
  const int slot = R->free_slot;
  B.store[slot * CN + idx] = B.source[idx];
  if (R->compensated) B.states[slot * CN + idx] = B.lab[idx];
  B.y[idx] = P.active[idx % P.N] ? B.advanced[idx] : make(0.0, 0.0);
}

__global__ void adams_start(DeviceProblem P, const AdamsRecord* R, FieldBuffers B) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= CN) return;
  B.store[idx] = B.source[idx];
  if (R->compensated) B.states[idx] = B.y[idx];
}

// -------------------------------------------------------------------- launchers

unsigned correct_blocks(const DeviceProblem& P) {
  return blocks_for(static_cast<std::size_t>(P.C) * P.N, kThreads);
}

namespace {
void check() { SP_HIP(hipGetLastError()); }
unsigned cn_blocks(const DeviceProblem& P) {
  return blocks_for(static_cast<std::size_t>(P.C) * P.N, kThreads);
}
}  // namespace

void launch_adams_predict(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                          hipStream_t s) {
  hipLaunchKernelGGL(adams_predict, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, R, B);
  check();
}

void launch_field_scatter(const DeviceProblem& P, const cplx* spectrum, FieldBuffers B,
                          hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * (P.T / 2);
  hipLaunchKernelGGL(field_scatter, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P,
                     spectrum, B.modal_t, B.roots_T);
  check();
}

void launch_node_passes(const DeviceProblem& P, const AdamsRecord* R, const PlasmaTable& pt,
                        FieldBuffers B, hipStream_t s, hipEvent_t* marks) {
  SP_CHECK(P.C <= kMaxModes, "the node passes project at most 8 modes");
  SP_CHECK(P.T % kScanTile == 0, "the fine grid must hold whole scan tiles");
  const int tiles = P.T / kScanTile;
  const int node_blocks = (P.Q + kNodeBlock - 1) / kNodeBlock;
  const dim3 grid(tiles, node_blocks);
  auto mark = [&](int i) {
    if (marks) SP_HIP(hipEventRecord(marks[i], s));
  };
  if (P.has_plasma) {
    hipLaunchKernelGGL(plasma_exposure_tiles, grid, dim3(kScanTile), 0, s, P, pt, B);
    check();
    mark(0);
    hipLaunchKernelGGL(plasma_tile_prefix, dim3(blocks_for(P.Q, 64)), dim3(64), 0, s, P.Q,
                       tiles, B.tile_x);
    check();
    mark(1);
    hipLaunchKernelGGL(node_response<false>, grid, dim3(kScanTile), 0, s, P, R, pt, B);
    check();
    mark(2);
    hipLaunchKernelGGL(plasma_tile_prefix, dim3(blocks_for(P.Q, 64)), dim3(64), 0, s, P.Q,
                       tiles, B.tile_f);
    check();
    mark(3);
  } else {
    for (int i = 0; i < 4; ++i) mark(i);
  }
  hipLaunchKernelGGL(node_response<true>, grid, dim3(kScanTile), 0, s, P, R, pt, B);
  check();
  mark(4);
  const std::size_t n = 3UL * P.C * P.T;
  hipLaunchKernelGGL(stream_reduce, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P, B,
                     node_blocks);
  check();
}

void launch_field_finish(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                         cplx* out, hipStream_t s) {
  hipLaunchKernelGGL(field_finish, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, R, B, out);
  check();
}

void launch_adams_correct(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                          hipStream_t s) {
  hipLaunchKernelGGL(adams_correct, dim3(correct_blocks(P)), dim3(kThreads), 0, s, P, R, B);
  check();
}

void launch_adams_accept(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                         hipStream_t s) {
  hipLaunchKernelGGL(adams_accept, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, R, B);
  check();
}

void launch_adams_start(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                        hipStream_t s) {
  hipLaunchKernelGGL(adams_start, dim3(cn_blocks(P)), dim3(kThreads), 0, s, P, R, B);
  check();
}

}  // namespace sp
