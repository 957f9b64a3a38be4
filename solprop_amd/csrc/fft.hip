// Stockham FFT passes staged in LDS (radix-16 stages in registers for large workgroups, radix-4/2 otherwise), FP64, for gfx950.
#include "fft.hpp"
#include "fft_lds.hpp"

#include <algorithm>
#include <cmath>

namespace sp {
namespace {

using PassArgs = FftPassArgs;

// The packed point at global index g (row g / M, point g % M) from the half spectrum.
// M is the transform's length, a power of two: a shift and a mask find the row and point.
__device__ inline cplx packed_point(const PassArgs& a, long g) {
  const long M = a.spec_m;
  const long r = g >> __builtin_ctzl(static_cast<unsigned long>(M)), k = g & (M - 1);
  const cplx* row = a.spec + r * a.spec_n;
  const long kb = M - k;
  cplx xa = k < a.spec_n ? scale(a.spec_scale, row[k]) : make(0.0, 0.0);
  cplx xb = kb < a.spec_n ? scale(a.spec_scale, row[kb]) : make(0.0, 0.0);
  if (k == 0) {  // DC and Nyquist are real in a real transform (as irfft discards)
    xa.y = 0.0;
    xb.y = 0.0;
  }
  return c2r_pre(xa, xb, a.spec_roots[k]);
}

// The pass's workgroup shape: radix-16 stages on padded rows for sequences of at least
// 1024 points, radix-4 on plain rows otherwise. LDS_POINTS is the complex points of the
// workgroup's LDS buffer (a whole-point hand-over through half of it, real parts then
// imaginary ones, let more workgroups share a CU and ran slower).
template <int LOG2L, int G>
struct PassShape {
  static constexpr bool r16 = Shape16<LOG2L, G>::use;
  static constexpr int L = 1 << LOG2L;
  static constexpr int E = L * G;
  static constexpr int NT = r16 ? Shape16<LOG2L, G>::NT : Shape<LOG2L, G>::NT;
  static constexpr int STRIDE = r16 ? Shape16<LOG2L, G>::STRIDE : Shape<LOG2L, G>::STRIDE;
  // The radix-4 shapes of at least 256 points also stage the pass's L twiddles after the
  // rows (shorter sequences take theirs from the table: staging them measured slower).
  static constexpr bool stage_tw = !r16 && L >= 256;
  static constexpr int LDS_POINTS = G * STRIDE + (stage_tw ? L : 0);
};

// A radix-16 pass (see the radix-16 stages in fft_lds.hpp) keeps each stage's values in
// registers: a thread runs PER butterflies of R points. The first stage reads HBM (point
// j of sequence s at the pass's input strides, or packed from the half spectrum), the
// last writes HBM (point k = i + j p, times the four-step twiddle where asked), and each
// stage hands its outputs to the next through the padded LDS rows.
template <int LOG2L, int G, int R>
struct Stage16 {
  static constexpr int NB = Shape16<LOG2L, G>::E / R;
  static constexpr int PER = (NB + Shape16<LOG2L, G>::NT - 1) / Shape16<LOG2L, G>::NT;
};

// A thread's values of one stage.
template <int LOG2L, int G, int R>
using Regs16 = cplx[Stage16<LOG2L, G, R>::PER][R];

// The radix of stage `Stage` of a pass over 2^LOG2L points: 16 while whole radix-16
// stages remain, then 2, 4 or 8 for what remains of L.
template <int LOG2L, int Stage>
struct Radix16 {
  static constexpr int full = LOG2L / 4;
  static constexpr int rest = (1 << LOG2L) >> (4 * full);
  static constexpr int stages = full + (rest > 1 ? 1 : 0);
  static constexpr int R = Stage < full ? 16 : rest;
  static constexpr bool last = Stage == stages - 1;
};

template <int LOG2L, int G, int R>
__device__ __forceinline__ void r16_load(const cplx* src, const PassArgs& a, long s0, int row,
                                         Regs16<LOG2L, G, R>& v) {
  constexpr int L = 1 << LOG2L, NT = Shape16<LOG2L, G>::NT;
  constexpr int NB = Stage16<LOG2L, G, R>::NB, PER = Stage16<LOG2L, G, R>::PER;
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int b = tid + r * NT;
    if (b < NB) {
      const int s = b / (L / R), i = b % (L / R);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const long at = (i + j * (L / R)) * a.in_sj + (s0 + s) * a.in_ss;
        v[r][j] = a.spec ? packed_point(a, static_cast<long>(row) * a.batch_in + at) : src[at];
      }
    }
  }
}

// The stage's twiddles W_{pR}^{j k} (k = i mod p) and its butterflies.
template <int LOG2L, int G, int R>
__device__ __forceinline__ void r16_butterflies(Regs16<LOG2L, G, R>& v,
                                                const cplx* __restrict__ tw_l, int sign, int p) {
  constexpr int L = 1 << LOG2L, NT = Shape16<LOG2L, G>::NT;
  constexpr int NB = Stage16<LOG2L, G, R>::NB, PER = Stage16<LOG2L, G, R>::PER;
  const int tid = threadIdx.x;
  const int step = L / (p * R);
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int b = tid + r * NT;
    if (b < NB) {
      const int k = (b % (L / R)) & (p - 1);
      if (p > 1)
#pragma unroll
        for (int j = 1; j < R; ++j) v[r][j] = mul(v[r][j], twiddle(tw_l, j * k * step, sign));
      dft<R>(v[r], sign);
    }
  }
}

// Where output j of butterfly b of a stage with span p lands in its sequence.
template <int LOG2L, int R>
__device__ __forceinline__ int r16_out_point(int b, int j, int p) {
  constexpr int L = 1 << LOG2L;
  const int i = b % (L / R), k = i & (p - 1);
  return (i - k) * R + k + j * p;
}

template <int LOG2L, int G, int R>
__device__ __forceinline__ void r16_store(cplx* dst, const cplx* __restrict__ tw_n,
                                          const PassArgs& a, int sign, long s0,
                                          const Regs16<LOG2L, G, R>& v, int p) {
  constexpr int L = 1 << LOG2L, NT = Shape16<LOG2L, G>::NT;
  constexpr int NB = Stage16<LOG2L, G, R>::NB, PER = Stage16<LOG2L, G, R>::PER;
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int b = tid + r * NT;
    if (b < NB) {
      const int s = b / (L / R);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int point = r16_out_point<LOG2L, R>(b, j, p);
        cplx value = v[r][out_at<R>(j)];
        if (a.global_twiddle) {
          const long idx = ((s0 + s) * static_cast<long>(point)) & a.mask_n;
          value = mul(value, twiddle(tw_n, static_cast<int>(idx), sign));
        }
        dst[point * a.out_sk + (s0 + s) * a.out_ss] = value;
      }
    }
  }
}

// An LDS hand-over: a stage's outputs `w` (radix RW, span pw) written to the rows, a
// barrier, and the next stage's inputs `v` (radix RR) read back.
template <int LOG2L, int G, int RW, int RR>
__device__ __forceinline__ void r16_move(const Regs16<LOG2L, G, RW>& w, int pw,
                                         Regs16<LOG2L, G, RR>& v, cplx* lds) {
  constexpr int L = 1 << LOG2L, NT = Shape16<LOG2L, G>::NT, STRIDE = Shape16<LOG2L, G>::STRIDE;
  constexpr int NBW = Stage16<LOG2L, G, RW>::NB, PERW = Stage16<LOG2L, G, RW>::PER;
  constexpr int NBR = Stage16<LOG2L, G, RR>::NB, PERR = Stage16<LOG2L, G, RR>::PER;
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < PERW; ++r) {
    const int b = tid + r * NT;
    if (b < NBW) {
      const int s = b / (L / RW);
#pragma unroll
      for (int j = 0; j < RW; ++j)
        lds[s * STRIDE + padded(r16_out_point<LOG2L, RW>(b, j, pw))] = w[r][out_at<RW>(j)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PERR; ++r) {
    const int b = tid + r * NT;
    if (b < NBR) {
      const int s = b / (L / RR), i = b % (L / RR);
#pragma unroll
      for (int j = 0; j < RR; ++j) v[r][j] = lds[s * STRIDE + padded(i + j * (L / RR))];
    }
  }
}

// The stages from `Stage` on, its inputs in `v` and its span p; `lds` holds
// PassShape::LDS_POINTS points. A hand-over after the first starts with a barrier: the
// previous hand-over's reads are done before the rows are written again.
template <int LOG2L, int G, int Stage>
__device__ __forceinline__ void r16_stages(Regs16<LOG2L, G, Radix16<LOG2L, Stage>::R>& v,
                                           cplx* dst, const cplx* __restrict__ tw_l,
                                           const cplx* __restrict__ tw_n, const PassArgs& a,
                                           int sign, long s0, int p, cplx* lds) {
  constexpr int R = Radix16<LOG2L, Stage>::R;
  r16_butterflies<LOG2L, G, R>(v, tw_l, sign, p);
  if constexpr (Radix16<LOG2L, Stage>::last) {
    r16_store<LOG2L, G, R>(dst, tw_n, a, sign, s0, v, p);
  } else {
    constexpr int RN = Radix16<LOG2L, Stage + 1>::R;
    Regs16<LOG2L, G, RN> next;
    if constexpr (Stage > 0) __syncthreads();
    r16_move<LOG2L, G, R, RN>(v, p, next, lds);
    r16_stages<LOG2L, G, Stage + 1>(next, dst, tw_l, tw_n, a, sign, s0, p * R, lds);
  }
}

template <int LOG2L, int G>
__device__ __forceinline__ void fft_pass_body(const cplx* in, cplx* out, const cplx* __restrict__ tw_l,
                                     const cplx* __restrict__ tw_n, const PassArgs& a, int sign,
                                     int group, int row, cplx* buf) {
  using S = PassShape<LOG2L, G>;
  if constexpr (S::r16) {
    const long s0 = static_cast<long>(group) * G;
    constexpr int R0 = Radix16<LOG2L, 0>::R;
    Regs16<LOG2L, G, R0> v;
    r16_load<LOG2L, G, R0>(in + static_cast<long>(row) * a.batch_in, a, s0, row, v);
    r16_stages<LOG2L, G, 0>(v, out + static_cast<long>(row) * a.batch_out, tw_l, tw_n, a, sign,
                            s0, 1, buf);
    return;
  }
  constexpr int L = S::L, E = S::E, NT = S::NT, STRIDE = S::STRIDE;
  const int tid = threadIdx.x;
  const long s0 = static_cast<long>(group) * G;
  const cplx* src = in + static_cast<long>(row) * a.batch_in;
  cplx* dst = out + static_cast<long>(row) * a.batch_out;

  // Where store e of the workgroup goes: sequence s, point k.
  auto store_at = [&](int e, int& sq, int& k) {
    if (a.out_ss == 1) {
      sq = e % G;
      k = e / G;
    } else {
      k = e % L;
      sq = e / L;
    }
  };
  // Loads that do not wait on the data are issued with it: the stages' twiddles into LDS
  // after the rows (a stage would otherwise wait for its own after each barrier) and,
  // for a few stores a thread, the four-step twiddles into registers.
  constexpr int PE = (E + NT - 1) / NT;
  constexpr bool held = PE <= 4;
  cplx* tw = S::stage_tw ? buf + G * STRIDE : nullptr;
  cplx gw[held ? PE : 1];
  if (held && a.global_twiddle) {
#pragma unroll
    for (int r = 0; r < PE; ++r) {
      const int e = tid + r * NT;
      if (e < E) {
        int sq, k;
        store_at(e, sq, k);
        const long idx = ((s0 + sq) * static_cast<long>(k)) & a.mask_n;
        gw[r] = twiddle(tw_n, static_cast<int>(idx), sign);
      }
    }
  }
  // Load G sequences, the contiguous axis fastest across the threads.
  for (int e = tid; e < E; e += NT) {
    int s, j;
    if (a.in_ss == 1) {
      s = e % G;
      j = e / G;
    } else {
      j = e % L;
      s = e / L;
    }
    const long at = j * a.in_sj + (s0 + s) * a.in_ss;
    buf[s * STRIDE + j] =
        a.spec ? packed_point(a, static_cast<long>(row) * a.batch_in + at) : src[at];
  }
  if constexpr (S::stage_tw)
    for (int e = tid; e < L; e += NT) tw[e] = tw_l[e];
  __syncthreads();

  lds_transform<LOG2L, G>(buf, S::stage_tw ? tw : tw_l, sign);

  // Store, with the four-step twiddle W_N^{s k} where asked.
#pragma unroll
  for (int r = 0; r < PE; ++r) {
    const int e = tid + r * NT;
    if (e < E) {
      int sq, k;
      store_at(e, sq, k);
      cplx value = buf[sq * STRIDE + k];
      if (a.global_twiddle) {
        if constexpr (held) {
          value = mul(value, gw[r]);
        } else {
          const long idx = ((s0 + sq) * static_cast<long>(k)) & a.mask_n;
          value = mul(value, twiddle(tw_n, static_cast<int>(idx), sign));
        }
      }
      dst[k * a.out_sk + (s0 + sq) * a.out_ss] = value;
    }
  }
}

template <int LOG2L, int G>
__global__ void __launch_bounds__((PassShape<LOG2L, G>::NT))
    fft_pass_kernel(const cplx* in, cplx* out,
                    const cplx* __restrict__ tw_l, const cplx* __restrict__ tw_n,
                    PassArgs a, int sign) {
  using S = PassShape<LOG2L, G>;
  __shared__ cplx buf[S::LDS_POINTS];
  fft_pass_body<LOG2L, G>(in, out, tw_l, tw_n, a, sign, blockIdx.x, blockIdx.y, buf);
}

// The same pass over the jobs of a grouped launch: each workgroup finds its job by the
// jobs' first workgroups (a binary search, uniform across the workgroup) and runs that
// job's sequence group and row.
template <int LOG2L, int G>
__global__ void __launch_bounds__((PassShape<LOG2L, G>::NT))
    fft_pass_grouped(const FftJob* __restrict__ jobs, const int* __restrict__ first, int n_jobs,
                     int sign) {
  using S = PassShape<LOG2L, G>;
  __shared__ cplx buf[S::LDS_POINTS];
  const int block = blockIdx.x;
  const int i = owner_of(first, n_jobs, block);
  const FftJob& job = jobs[i];
  if (job.gate && *job.gate == 0) return;
  const int local = block - first[i];
  fft_pass_body<LOG2L, G>(job.in, job.out, job.tw_l, job.tw_n, job.a, sign, local % job.groups,
                          local / job.groups, buf);
}

using GroupedFn = void (*)(const FftJob*, const int*, int, int);

// The pass shapes that exist: workgroups of at most kMaxElems points.
template <int LOG2L, int G>
constexpr bool pass_exists() {
  return (1 << LOG2L) * G <= kMaxElems;
}

template <int LOG2L, int G>
constexpr GroupedFn pick_grouped() {
  if constexpr (pass_exists<LOG2L, G>())
    return fft_pass_grouped<LOG2L, G>;
  else
    return nullptr;
}

template <int LOG2L>
GroupedFn grouped_for_g(int log2g) {
  switch (log2g) {
    case 0: return pick_grouped<LOG2L, 1>();
    case 1: return pick_grouped<LOG2L, 2>();
    case 2: return pick_grouped<LOG2L, 4>();
    case 3: return pick_grouped<LOG2L, 8>();
    case 4: return pick_grouped<LOG2L, 16>();
    default: return nullptr;
  }
}

GroupedFn grouped_for(int log2l, int log2g) {
  switch (log2l) {
    case 5: return grouped_for_g<5>(log2g);
    case 6: return grouped_for_g<6>(log2g);
    case 7: return grouped_for_g<7>(log2g);
    case 8: return grouped_for_g<8>(log2g);
    case 9: return grouped_for_g<9>(log2g);
    case 10: return grouped_for_g<10>(log2g);
    case 11: return grouped_for_g<11>(log2g);
    case 12: return grouped_for_g<12>(log2g);
    default: return nullptr;
  }
}

using KernelFn = void (*)(const cplx*, cplx*, const cplx*, const cplx*, PassArgs, int);

template <int LOG2L, int G>
constexpr KernelFn pick() {
  if constexpr (pass_exists<LOG2L, G>())
    return fft_pass_kernel<LOG2L, G>;
  else
    return nullptr;
}

template <int LOG2L>
KernelFn kernel_for_g(int log2g) {
  switch (log2g) {
    case 0: return pick<LOG2L, 1>();
    case 1: return pick<LOG2L, 2>();
    case 2: return pick<LOG2L, 4>();
    case 3: return pick<LOG2L, 8>();
    case 4: return pick<LOG2L, 16>();
    case 5: return pick<LOG2L, 32>();
    case 6: return pick<LOG2L, 64>();
    default: return nullptr;
  }
}

KernelFn kernel_for(int log2l, int log2g) {
  switch (log2l) {
    case 2: return kernel_for_g<2>(log2g);
    case 3: return kernel_for_g<3>(log2g);
    case 4: return kernel_for_g<4>(log2g);
    case 5: return kernel_for_g<5>(log2g);
    case 6: return kernel_for_g<6>(log2g);
    case 7: return kernel_for_g<7>(log2g);
    case 8: return kernel_for_g<8>(log2g);
    case 9: return kernel_for_g<9>(log2g);
    case 10: return kernel_for_g<10>(log2g);
    case 11: return kernel_for_g<11>(log2g);
    case 12: return kernel_for_g<12>(log2g);
    default: return nullptr;
  }
}

int threads_for(int log2l, int log2g) {
  const int e = 1 << (log2l + log2g);
  // The radix-16 shapes (Shape16::use): a thread a 16-point butterfly, at most 256.
  if (log2l >= 10) return e / 16 >= 256 ? 256 : e / 16;
  const int nt = e / 4;
  return nt >= 256 ? 256 : (nt >= 1 ? nt : 1);
}

int ilog2(std::size_t n) {
  int m = 0;
  while ((std::size_t{1} << m) < n) ++m;
  return m;
}

// Sequences per workgroup: aim for 256 points (one wavefront of radix-4 butterflies),
// never more than the sequences there are, and a divisor of their count. Small
// workgroups keep the short transforms of a few modes spread over many CUs.
int pick_log2g(int log2l, std::size_t n_seq) {
  int g = 0;
  while (g < 6 && (1 << (log2l + g + 1)) <= 256 && (std::size_t{1} << (g + 1)) <= n_seq &&
         n_seq % (std::size_t{1} << (g + 1)) == 0)
    ++g;
  return g;
}

}  // namespace

std::vector<cplx> unit_roots(std::size_t n) {
  // exp(-2 pi i t/n): the angle is reduced to an octant offset in [0, pi/4] in long
  // double (exact, n a power of two) and mapped by the octant symmetries, reflecting
  // the odd octants so that no cancellation occurs.
  std::vector<cplx> w(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (std::size_t t = 0; t < n; ++t) {
    const std::size_t q = (8 * t) / n;  // octant 0..7
    const long double frac = static_cast<long double>(t) / static_cast<long double>(n);
    long double c, s;
    if (q % 2 == 0) {
      const long double r = two_pi * (frac - static_cast<long double>(q) / 8.0L);
      const long double cr = std::cos(r), sr = std::sin(r);
      switch (q) {
        case 0: c = cr; s = sr; break;
        case 2: c = -sr; s = cr; break;
        case 4: c = -cr; s = -sr; break;
        default: c = sr; s = -cr; break;
      }
    } else {
      const long double r = two_pi * (static_cast<long double>(q + 1) / 8.0L - frac);
      const long double cr = std::cos(r), sr = std::sin(r);
      switch (q) {
        case 1: c = sr; s = cr; break;
        case 3: c = -cr; s = sr; break;
        case 5: c = -sr; s = -cr; break;
        default: c = cr; s = -sr; break;
      }
    }
    w[t] = make(static_cast<double>(c), -static_cast<double>(s));
  }
  return w;
}

FftPlan::~FftPlan() {
  if (!owns_tables_) return;
  if (roots_n_) (void)hipFree(roots_n_);
  for (auto* p : roots_l_)
    if (p) (void)hipFree(p);
}

void FftPlan::layout(std::size_t n, std::size_t batch) {
  SP_CHECK(n >= 4 && (n & (n - 1)) == 0, "FFT length must be a power of two >= 4");
  SP_CHECK(n <= (std::size_t{1} << 24), "FFT length above 2^24 is not supported");
  SP_CHECK(batch >= 1, "FFT batch must be positive");
  n_ = n;
  batch_ = batch;
  const int m = ilog2(n);
  passes_.clear();
  if (n <= static_cast<std::size_t>(kMaxElems)) {
    FftPass p;
    p.log2L = m;
    p.log2G = pick_log2g(m, batch);
    p.in_sj = 1;
    p.in_ss = static_cast<long>(n);
    p.out_sk = 1;
    p.out_ss = static_cast<long>(n);
    p.batch_in = p.batch_out = 0;
    p.n_seq = static_cast<int>(batch);
    passes_.push_back(p);
  } else {
    const int m1 = m / 2, m2 = m - m1;
    const long n1 = 1L << m1, n2 = 1L << m2;
    FftPass p1;  // columns: length n1 over j1 (stride n2), sequences j2
    p1.log2L = m1;
    p1.log2G = pick_log2g(m1, static_cast<std::size_t>(n2));
    p1.in_sj = n2;
    p1.in_ss = 1;
    p1.out_sk = n2;
    p1.out_ss = 1;
    p1.batch_in = p1.batch_out = static_cast<long>(n);
    p1.n_seq = static_cast<int>(n2);
    p1.global_twiddle = true;
    p1.log2N = m;
    FftPass p2;  // rows: length n2 over j2, sequences k1 (stride n2), transposed store
    p2.log2L = m2;
    p2.log2G = pick_log2g(m2, static_cast<std::size_t>(n1));
    p2.in_sj = 1;
    p2.in_ss = n2;
    p2.out_sk = n1;
    p2.out_ss = 1;
    p2.batch_in = p2.batch_out = static_cast<long>(n);
    p2.n_seq = static_cast<int>(n1);
    passes_.push_back(p1);
    passes_.push_back(p2);
  }
  for (const auto& p : passes_)
    SP_CHECK(kernel_for(p.log2L, p.log2G) != nullptr, "no FFT pass kernel for this shape");
}

void FftPlan::init_axis(std::size_t n, std::size_t inner, std::size_t outer,
                        std::size_t max_points) {
  SP_CHECK(n >= 4 && (n & (n - 1)) == 0 && n <= static_cast<std::size_t>(kMaxElems),
           "an axis transform has a power-of-two length from 4 to 4096");
  SP_CHECK(inner >= 1 && outer >= 1 && inner <= (std::size_t{1} << 30),
           "an axis transform's strides must be positive");
  n_ = n;
  batch_ = outer;
  owns_tables_ = true;
  passes_.clear();
  const int m = ilog2(n);
  FftPass p;
  p.log2L = m;
  // Sequences a workgroup owns: as many neighbours as LDS holds (at most 16, 256 bytes
  // of each row), so that every load and store moves contiguous runs of G points.
  int g = 0;
  const std::size_t cap = std::min(max_points, static_cast<std::size_t>(kMaxElems));
  while (g < 4 && (std::size_t{1} << (m + g + 1)) <= cap && (std::size_t{1} << (g + 1)) <= inner &&
         inner % (std::size_t{1} << (g + 1)) == 0)
    ++g;
  p.log2G = g;
  p.in_sj = static_cast<long>(inner);
  p.in_ss = 1;
  p.out_sk = static_cast<long>(inner);
  p.out_ss = 1;
  // Rows are the outer blocks; a single block needs no row stride.
  p.batch_in = p.batch_out = outer > 1 ? static_cast<long>(n * inner) : 0;
  p.n_seq = static_cast<int>(inner);
  passes_.push_back(p);
  SP_CHECK(kernel_for(p.log2L, p.log2G) != nullptr, "no FFT pass kernel for this shape");
  auto upload = [&](std::size_t len, cplx** dst) {
    const auto host = unit_roots(len);
    SP_HIP(hipMalloc(dst, len * sizeof(cplx)));
    SP_HIP(hipMemcpy(*dst, host.data(), len * sizeof(cplx), hipMemcpyHostToDevice));
    bytes_ += len * sizeof(cplx);
  };
  upload(n, &roots_n_);
  upload(n, &roots_l_[0]);
}

std::size_t FftPlan::table_points(std::size_t n) {
  if (n <= static_cast<std::size_t>(kMaxElems)) return 2 * n;
  const int m = ilog2(n);
  const int m1 = m / 2, m2 = m - m1;
  return n + (std::size_t{1} << m1) + (std::size_t{1} << m2);
}

void FftPlan::init(std::size_t n, std::size_t batch) {
  layout(n, batch);
  owns_tables_ = true;
  auto upload = [&](std::size_t len, cplx** dst) {
    const auto host = unit_roots(len);
    SP_HIP(hipMalloc(dst, len * sizeof(cplx)));
    SP_HIP(hipMemcpy(*dst, host.data(), len * sizeof(cplx), hipMemcpyHostToDevice));
    bytes_ += len * sizeof(cplx);
  };
  upload(n, &roots_n_);
  upload(static_cast<std::size_t>(1) << passes_[0].log2L, &roots_l_[0]);
  if (passes_.size() == 2)
    upload(static_cast<std::size_t>(1) << passes_[1].log2L, &roots_l_[1]);
}

void FftPlan::init_external(std::size_t n, std::size_t batch, cplx* tables) {
  SP_CHECK(tables != nullptr, "external FFT tables need device memory");
  layout(n, batch);
  owns_tables_ = false;
  roots_n_ = tables;
  roots_l_[0] = tables + n;
  roots_l_[1] = passes_.size() == 2 ? roots_l_[0] + (std::size_t{1} << passes_[0].log2L)
                                    : nullptr;
}

void FftPlan::upload_tables(hipStream_t stream) const {
  std::vector<cplx> host = unit_roots(n_);
  const auto local = [&](int pass) {
    const auto roots = unit_roots(std::size_t{1} << passes_[pass].log2L);
    host.insert(host.end(), roots.begin(), roots.end());
  };
  local(0);
  if (passes_.size() == 2) local(1);
  SP_HIP(hipMemcpyAsync(roots_n_, host.data(), host.size() * sizeof(cplx),
                        hipMemcpyHostToDevice, stream));
  // The host rows must outlive the copy.
  SP_HIP(hipStreamSynchronize(stream));
}

void FftPlan::launch(const FftPass& p, const cplx* in, cplx* out, int sign,
                     hipStream_t stream) const {
  const int log2g = p.log2G;
  KernelFn fn = kernel_for(p.log2L, log2g);
  const int g = 1 << log2g;
  PassArgs a{p.in_sj, p.in_ss, p.out_sk, p.out_ss, p.batch_in, p.batch_out,
             p.global_twiddle ? 1 : 0, static_cast<int>((1L << p.log2N) - 1),
             nullptr, nullptr, 0, 0.0, static_cast<long>(n_)};
  if (spec_) {
    a.spec = spec_->spectrum;
    a.spec_roots = spec_->roots;
    a.spec_n = spec_->n_bins;
    a.spec_scale = spec_->scale;
  }
  const unsigned groups = static_cast<unsigned>(p.n_seq / g);
  const unsigned rows = p.batch_in == 0 && p.batch_out == 0 ? 1u : static_cast<unsigned>(batch_);
  const cplx* tw_l = (&p == &passes_[0]) ? roots_l_[0] : roots_l_[1];
  hipLaunchKernelGGL(fn, dim3(groups, rows), dim3(threads_for(p.log2L, log2g)), 0, stream,
                     in, out, tw_l, roots_n_, a, sign);
  SP_HIP(hipGetLastError());
}

std::vector<FftJob> FftPlan::jobs(cplx* data, cplx* scratch, const PackedSpectrum* packed,
                                  const int* gate) const {
  std::vector<FftJob> out;
  for (std::size_t i = 0; i < passes_.size(); ++i) {
    const FftPass& p = passes_[i];
    const bool last = i + 1 == passes_.size();
    FftJob j{};
    j.in = i == 0 ? data : scratch;
    j.out = last ? data : scratch;
    j.tw_l = i == 0 ? roots_l_[0] : roots_l_[1];
    j.tw_n = roots_n_;
    j.a = PassArgs{p.in_sj, p.in_ss, p.out_sk, p.out_ss, p.batch_in, p.batch_out,
                   p.global_twiddle ? 1 : 0, static_cast<int>((1L << p.log2N) - 1),
                   nullptr, nullptr, 0, 0.0, static_cast<long>(n_)};
    if (i == 0 && packed) {
      j.a.spec = packed->spectrum;
      j.a.spec_roots = packed->roots;
      j.a.spec_n = static_cast<int>(packed->n_bins);
      j.a.spec_scale = packed->scale;
    }
    j.groups = p.n_seq >> p.log2G;
    j.rows = p.batch_in == 0 && p.batch_out == 0 ? 1 : static_cast<int>(batch_);
    j.gate = gate;
    out.push_back(j);
  }
  return out;
}

void launch_fft_grouped(int log2L, int log2G, const FftJob* jobs, const int* first, int n_jobs,
                        int blocks, int sign, hipStream_t stream) {
  GroupedFn fn = grouped_for(log2L, log2G);
  SP_CHECK(fn != nullptr, "no grouped FFT pass kernel for this shape");
  if (blocks <= 0) return;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads_for(log2L, log2G)), 0, stream, jobs, first,
                     n_jobs, sign);
  SP_HIP(hipGetLastError());
}

bool fft_grouped_available(int log2L, int log2G) { return grouped_for(log2L, log2G) != nullptr; }

void FftPlan::exec_packed(const PackedSpectrum& spectrum, cplx* data, cplx* scratch, int sign,
                          hipStream_t stream) const {
  SP_CHECK(batch_ >= 1 && spectrum.spectrum && spectrum.roots &&
               spectrum.n_bins >= 1 && spectrum.n_bins <= static_cast<long>(n_) + 1,
           "a packed spectrum of at most M + 1 bins a row");
  // Only the first pass loads: a row of M points is the whole input of each row.
  spec_ = &spectrum;
  if (passes_.size() == 1) {
    launch(passes_[0], data, data, sign, stream);
  } else {
    launch(passes_[0], data, scratch, sign, stream);
    spec_ = nullptr;
    launch(passes_[1], scratch, data, sign, stream);
  }
  spec_ = nullptr;
}

void FftPlan::exec(cplx* data, cplx* scratch, int sign, hipStream_t stream) const {
  if (passes_.size() == 1) {
    launch(passes_[0], data, data, sign, stream);
    return;
  }
  launch(passes_[0], data, scratch, sign, stream);
  launch(passes_[1], scratch, data, sign, stream);
}

}  // namespace sp
