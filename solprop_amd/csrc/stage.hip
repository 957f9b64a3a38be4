// Stage kernels of the interaction-picture Runge-Kutta trial, FP64, for gfx950.
//
// One stage on the envelope route is: entry (combine, propagate, add noise, scatter
// onto the fine grid with 1/T) -> inverse FFT -> Raman drive -> FFT -> kernel product
// -> inverse FFT -> response (Kerr + Raman, collar, projection) -> FFT -> finish
// (prefactor, spectral collar, back-propagate). The accept pass forms the solution and
// the embedded error from the stored stages, the step's end state and three masked
// partial sums in a fixed-order tree, so that repeats agree to the bit.
#include "kernels.hpp"

namespace sp {

namespace {

__device__ inline cplx phase_sum(const double* w, const cplx* terms, int n_terms, long stride,
                                 long at) {
  cplx total = make(0.0, 0.0);
  for (int t = 0; t < n_terms; ++t) total = add(total, scale(w[t], terms[t * stride + at]));
  return total;
}

// The stage's lab-frame field at one coarse bin, and exp(-ΔΦ_i) there.
__device__ inline cplx stage_lab(const DeviceProblem& P, const TrialRecord* R, const cplx* y,
                                 const cplx* k, int stage, int c, int b, cplx* back) {
  const StageRow& row = R->stage[stage];
  const long CN = static_cast<long>(P.C) * P.N;
  const long cb = static_cast<long>(c) * P.N + b;
  const cplx dphi = phase_sum(row.phi_w, P.phi, P.n_phi, CN, cb);
  cplx fwd;
  cexp_pair(dphi, &fwd, back);
  cplx comb = make(0.0, 0.0);
  for (int j = 0; j < stage; ++j)
    comb = add(comb, scale(P.tab_a[stage * P.n_stages + j], k[j * CN + cb]));
  cplx u = mul(fwd, add(y[cb], scale(R->dz, comb)));
  if (P.noise_modified) {
    double phase = 0.0;
    for (int t = 0; t < P.n_phi; ++t) phase = phase + row.noise_w[t] * P.phi[t * CN + cb].y;
    double s, co;
    sincos(phase, &s, &co);
    u = add(u, mul(P.noise[cb], make(co, s)));
  }
  return u;
}

__device__ inline cplx r2c_post(cplx za, cplx zb, cplx w) {
  // X = E + W O with E = (Za + conj Zb)/2, O = -i (Za - conj Zb)/2.
  const cplx e = scale(0.5, add(za, conj(zb)));
  const cplx o = scale(0.5, rot(sub(za, conj(zb)), -1));
  return add(e, mul(w, o));
}

}  // namespace

// ---------------------------------------------------------------- stage entry

__global__ void rk_entry_envelope(DeviceProblem P, const TrialRecord* R, const cplx* y,
                                  const cplx* k, cplx* fine, cplx* einv, int stage) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long total = static_cast<long>(P.C) * P.T;
  if (idx >= total) return;
  const int c = static_cast<int>(idx / P.T);
  const int f = static_cast<int>(idx % P.T);
  const int b = P.fine_to_coarse[f];
  if (b < 0) {
    fine[idx] = make(0.0, 0.0);
    return;
  }
  cplx back;
  const cplx u = stage_lab(P, R, y, k, stage, c, b, &back);
  fine[idx] = scale(P.inv_T, u);
  einv[static_cast<long>(c) * P.N + b] = back;
}

// ------------------------------------------------------- envelope response, Raman

__global__ void env_raman_drive(DeviceProblem P, const TrialRecord* R, const cplx* fine,
                                cplx* drive, int stage) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= P.T) return;
  if (m >= P.T / 2) {
    drive[m] = make(0.0, 0.0);
    return;
  }
  const double s = P.synthesis[0] * R->stage[stage].taper;
  const cplx e0 = scale(s, fine[2 * m]), e1 = scale(s, fine[2 * m + 1]);
  drive[m] = make(0.5 * (e0.x * e0.x + e0.y * e0.y), 0.5 * (e1.x * e1.x + e1.y * e1.y));
}

// One pair (k, T-k) of the packed doubled-grid transform of row blockIdx.y: unpack the
// real transform, multiply by the kernel spectrum, and pack again for the inverse, in
// place.
__global__ void raman_product(cplx* z, const cplx* H, const cplx* roots2T, int T) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > T / 2) return;
  z += static_cast<long>(blockIdx.y) * T;
  const int kk = (T - k) % T;
  const cplx za = z[k], zb = z[kk];
  const cplx xa = r2c_post(za, zb, roots2T[k]);
  const cplx xb = r2c_post(zb, za, roots2T[T - k]);
  const cplx ya = mul(H[k], xa), yb = mul(H[T - k], xb);
  z[k] = c2r_pre(ya, yb, roots2T[k]);
  if (k > 0 && k != T - k) z[T - k] = c2r_pre(yb, ya, roots2T[T - k]);
}

// Damped Raman rows on the doubled grid, packed two reals per complex point.
__global__ void raman_kernel_fill(DeviceProblem P, const TrialRecord* R, int stage,
                                  cplx* z) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= P.T) return;
  const long n2 = 2L * P.T;
  double h[2];
  for (int q = 0; q < 2; ++q) {
    const long t = 2L * m + q;
    const double delay = static_cast<double>(t) * P.dt_fine;
    double total = 0.0;
    for (int r = 0; r < P.n_raman; ++r)
      total = total + P.raman_rows[r * n2 + t] * exp(-R->stage[stage].raman_rate[r] * delay);
    h[q] = P.raman_fraction * total;
  }
  z[m] = make(h[0], h[1]);
}

// H[k] = rfft(h)[k] dt for k in [0, T].
__global__ void raman_kernel_post(const cplx* z, cplx* H, const cplx* roots2T, int T,
                                  double dt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > T) return;
  const cplx x = r2c_post(z[k % T], z[(T - k) % T], roots2T[k]);
  H[k] = scale(dt, x);
}

__global__ void env_response(DeviceProblem P, const TrialRecord* R, cplx* fine,
                             const cplx* conv, int stage) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.T) return;
  const StageRow& row = R->stage[stage];
  // The in-stage absorber's term from the modal field, before the polarisation takes
  // its place (src/solprop/backends/numpy_ref/passes/stepping.py:286-345).
  if (P.absorber_stream >= 0)
    fine[static_cast<long>(P.absorber_stream) * P.T + t] =
        scale(-0.5 * P.absorber_rate[t], fine[t]);
  const cplx e = scale(P.synthesis[0] * row.taper, fine[t]);
  cplx pol = make(0.0, 0.0);
  if (P.has_kerr) {
    const double factor = P.kerr_coefficient * row.density;
    const cplx squares = mul(e, e);
    const double modulus = hypot(e.x, e.y);
    const double moduli = modulus * modulus;
    const cplx self = scale(P.kerr_self * moduli, e);
    const cplx conjugate = mul(scale(P.kerr_conjugate, squares), conj(e));
    pol = add(pol, scale(factor, add(self, conjugate)));
  }
  if (P.n_raman > 0) {
    const cplx packed = conv[t / 2];
    const double cv = ((t & 1) ? packed.y : packed.x) * P.inv_2T;
    pol = add(pol, scale(cv, scale(P.raman_epsilon0 * row.density, e)));
  }
  pol = scale(P.t_window[t], pol);
  fine[t] = scale(P.projection[0] / row.taper, pol);
}

// ------------------------------------------------------------------ stage finish

// Stage `stage`'s derivative at coarse bin b of mode c from its transformed response:
// k = einv (p W P~ + A).
__device__ inline cplx stage_finish(const DeviceProblem& P, const TrialRecord* R,
                                    const cplx* fine, const cplx* einv, int stage, int c,
                                    int b) {
  const long CN = static_cast<long>(P.C) * P.N;
  const long cb = static_cast<long>(c) * P.N + b;
  const cplx pref = phase_sum(R->stage[stage].pref_w, P.pref, P.n_pref, CN, cb);
  const cplx pt = fine[static_cast<long>(c) * P.NF + P.crop[b]];
  cplx normalised = mul(scale(P.omega_window[b], pref), pt);
  // The absorber's term joins unscaled, on the active band (stepping.py:88-145).
  if (P.absorber_stream >= 0 && P.active[b])
    normalised = add(normalised,
                     fine[static_cast<long>(P.absorber_stream + c) * P.NF + P.crop[b]]);
  return mul(einv[cb], normalised);
}

// The previous stage's finish fused into this stage's entry. Each coarse bin b is the
// crop of exactly one fine bin f, whose thread reads the previous stage's transformed
// response at f and that stage's einv at b before it writes this stage's input at f and
// einv at b; the stage input then reads the derivative it just stored.
__global__ void rk_finish_entry(DeviceProblem P, const TrialRecord* R, const cplx* y, cplx* k,
                                cplx* fine, cplx* einv, int stage) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long total = static_cast<long>(P.C) * P.T;
  if (idx >= total) return;
  const int c = static_cast<int>(idx / P.T);
  const int f = static_cast<int>(idx % P.T);
  const int b = P.fine_to_coarse[f];
  if (b < 0) {
    fine[idx] = make(0.0, 0.0);
    return;
  }
  const long CN = static_cast<long>(P.C) * P.N;
  const long cb = static_cast<long>(c) * P.N + b;
  k[(stage - 1) * CN + cb] = stage_finish(P, R, fine, einv, stage - 1, c, b);
  cplx back;
  const cplx u = stage_lab(P, R, y, k, stage, c, b, &back);
  fine[idx] = scale(P.inv_T, u);
  einv[cb] = back;
}

// ----------------------------------------------------------------- step accept

template <int NT>
__device__ inline void block_sum3(double v0, double v1, double v2, double* out, int nblocks) {
  __shared__ double s0[NT], s1[NT], s2[NT];
  const int tid = threadIdx.x;
  s0[tid] = v0;
  s1[tid] = v1;
  s2[tid] = v2;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (tid < w) {
      s0[tid] = s0[tid] + s0[tid + w];
      s1[tid] = s1[tid] + s1[tid + w];
      s2[tid] = s2[tid] + s2[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[blockIdx.x] = s0[0];
    out[nblocks + blockIdx.x] = s1[0];
    out[2 * nblocks + blockIdx.x] = s2[0];
  }
}

__global__ void __launch_bounds__(kAcceptThreads)
    rk_accept(DeviceProblem P, const TrialRecord* R, const cplx* y, cplx* k, cplx* adv,
              cplx* recycled, double* partials, const cplx* fine, const cplx* einv) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long CN = static_cast<long>(P.C) * P.N;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (idx < CN) {
    const int b = static_cast<int>(idx % P.N);
    // The last stage's finish first (stage_finish, bin by bin), its derivative stored.
    const int last = P.n_stages - 1;
    k[last * CN + idx] = stage_finish(P, R, fine, einv, last, static_cast<int>(idx / P.N), b);
    cplx sol = make(0.0, 0.0), err = make(0.0, 0.0);
    for (int i = 0; i < P.n_stages; ++i) {
      const cplx ki = k[i * CN + idx];
      sol = add(sol, scale(P.tab_b[i], ki));
      err = add(err, scale(P.tab_e[i], ki));
    }
    const cplx whole = cexp(phase_sum(R->whole.phi_w, P.phi, P.n_phi, CN, idx));
    const double dz = R->dz;
    const cplx yn = mul(whole, add(y[idx], scale(dz, sol)));
    const cplx r = mul(whole, scale(dz, err));
    adv[idx] = yn;
    recycled[idx] = mul(whole, k[(P.n_stages - 1) * CN + idx]);
    if (P.normmask[b]) {
      v0 = norm2(y[idx]);
      v1 = norm2(yn);
      v2 = norm2(r);
    }
  }
  block_sum3<kAcceptThreads>(v0, v1, v2, partials, gridDim.x);
}

__global__ void commit_masked(DeviceProblem P, const cplx* adv, cplx* y) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long CN = static_cast<long>(P.C) * P.N;
  if (idx >= CN) return;
  y[idx] = P.active[idx % P.N] ? adv[idx] : make(0.0, 0.0);
}

// The split-step absorber on the committed envelope (the workspace's finish,
// src/solprop/backends/numpy_ref/workspace.py:1021-1106): the masked state to time
// samples (1/N folded in before the unscaled inverse transform), times exp(-alpha h / 2)
// with h the accepted step (the record's, inside a graph; else `dz`), and, after the
// forward transform, back onto the state with the band imposed again.
__global__ void env_split_pack(DeviceProblem P, const cplx* y, cplx* buf) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.C) * P.N) return;
  buf[idx] = scale(1.0 / P.N, y[idx]);
}

__global__ void env_split_collar(DeviceProblem P, const TrialRecord* R, double dz, cplx* buf) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(P.C) * P.N) return;
  const double h = R ? R->accepted_dz : dz;
  buf[idx] = scale(exp((-0.5 * P.absorber_rate[idx % P.N]) * h), buf[idx]);
}

__global__ void copy_points(const cplx* src, cplx* dst, long n) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx < n) dst[idx] = src[idx];
}

// -------------------------------------------------------------------- launchers

namespace {
constexpr int kThreads = 256;
void check_launch() { SP_HIP(hipGetLastError()); }
}  // namespace

void launch_rk_entry_envelope(const DeviceProblem& P, const TrialRecord* R, const cplx* y,
                              const cplx* k, cplx* fine, cplx* einv, int stage, hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * P.T;
  hipLaunchKernelGGL(rk_entry_envelope, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P,
                     R, y, k, fine, einv, stage);
  check_launch();
}

void launch_env_raman_drive(const DeviceProblem& P, const TrialRecord* R, const cplx* fine,
                            cplx* drive, int stage, hipStream_t s) {
  hipLaunchKernelGGL(env_raman_drive, dim3(blocks_for(P.T, kThreads)), dim3(kThreads), 0, s, P,
                     R, fine, drive, stage);
  check_launch();
}

void launch_raman_product(cplx* z, const cplx* H, const cplx* roots2T, int T, int rows,
                          hipStream_t s) {
  hipLaunchKernelGGL(raman_product, dim3(blocks_for(T / 2 + 1, kThreads), rows),
                     dim3(kThreads), 0, s, z, H, roots2T, T);
  check_launch();
}

void launch_raman_kernel_fill(const DeviceProblem& P, const TrialRecord* R, int stage, cplx* z,
                              hipStream_t s) {
  hipLaunchKernelGGL(raman_kernel_fill, dim3(blocks_for(P.T, kThreads)), dim3(kThreads), 0, s,
                     P, R, stage, z);
  check_launch();
}

void launch_raman_kernel_post(const cplx* z, cplx* H, const cplx* roots2T, int T, double dt,
                              hipStream_t s) {
  hipLaunchKernelGGL(raman_kernel_post, dim3(blocks_for(T + 1, kThreads)), dim3(kThreads), 0,
                     s, z, H, roots2T, T, dt);
  check_launch();
}

void launch_env_response(const DeviceProblem& P, const TrialRecord* R, cplx* fine,
                         const cplx* conv, int stage, hipStream_t s) {
  hipLaunchKernelGGL(env_response, dim3(blocks_for(P.T, kThreads)), dim3(kThreads), 0, s, P, R,
                     fine, conv, stage);
  check_launch();
}

void launch_rk_finish_entry(const DeviceProblem& P, const TrialRecord* R, const cplx* y,
                            cplx* k, cplx* fine, cplx* einv, int stage, hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * P.T;
  hipLaunchKernelGGL(rk_finish_entry, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P,
                     R, y, k, fine, einv, stage);
  check_launch();
}

void launch_rk_accept(const DeviceProblem& P, const TrialRecord* R, const cplx* y, cplx* k,
                      cplx* adv, cplx* recycled, double* partials, const cplx* fine,
                      const cplx* einv, hipStream_t s) {
  hipLaunchKernelGGL(rk_accept, dim3(accept_blocks(P)), dim3(kAcceptThreads), 0, s, P, R, y, k,
                     adv, recycled, partials, fine, einv);
  check_launch();
}

void launch_commit_masked(const DeviceProblem& P, const cplx* adv, cplx* y, hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * P.N;
  hipLaunchKernelGGL(commit_masked, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P, adv,
                     y);
  check_launch();
}

void launch_env_split_pack(const DeviceProblem& P, const cplx* y, cplx* buf, hipStream_t s) {
  const std::size_t n = static_cast<std::size_t>(P.C) * P.N;
  hipLaunchKernelGGL(env_split_pack, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                     P, y, buf);
  check_launch();
}

void launch_env_split_collar(const DeviceProblem& P, const TrialRecord* R, double dz, cplx* buf,
                             hipStream_t s) {
  SP_CHECK(P.absorber_rate, "the split-step absorber's rate");
  const std::size_t n = static_cast<std::size_t>(P.C) * P.N;
  hipLaunchKernelGGL(env_split_collar, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     s, P, R, dz, buf);
  check_launch();
}

void launch_copy_points(const cplx* src, cplx* dst, long n, hipStream_t s) {
  hipLaunchKernelGGL(copy_points, dim3(blocks_for(static_cast<std::size_t>(n), kThreads)),
                     dim3(kThreads), 0, s, src, dst, n);
  check_launch();
}

}  // namespace sp
