// The learned surrogate's kernels (see fno.hpp): BF16 matrix-core products with fused
// epilogues, the spectral mixing, the head, the loss, reductions and Adam.
#include "fno.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace sp::fno {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

__host__ __device__ inline bf16 to_bf16(float f) {
  std::uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                       // round to nearest even
  return static_cast<bf16>(u >> 16);
}

__device__ inline float from_bf16(bf16 h) {
  return __uint_as_float(static_cast<std::uint32_t>(h) << 16);
}

__device__ inline float gelu(float x) {
  return 0.5f * x * (1.0f + tanhf(kGeluC * (x + kGeluA * x * x * x)));
}

__device__ inline float gelu_grad(float x) {
  const float t = tanhf(kGeluC * (x + kGeluA * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * x * x);
}

// ------------------------------------------------------------------------- GEMM
//
// C[b] (M x N, FP32) = alpha A[b] (M x K) . B[b] (K x N), BF16 operands at any element
// strides (so transposes are free), one 64 x 64 tile a workgroup of four waves, each wave
// a 32 x 32 quadrant of 2 x 2 MFMA tiles, K in steps of 32 staged in LDS (A as [m][k],
// B transposed as [n][k], so that each lane's four K values are contiguous). With
// `reduce_batch` the workgroup sums every batch into one C. Epilogue: + C (accumulate),
// + bias[m], then the FP32 result stored and, when act16 is given, its gelu (act = 1) or
// itself (act = 0) stored in BF16.
struct GemmArgs {
  const bf16* A;
  long sAm, sAk, sAb;
  const bf16* B;
  long sBk, sBn, sBb;
  float* C;
  long ldc, sCb;
  int M, N, K, batch;
  int reduce_batch, accumulate;
  const float* bias;
  int act;
  bf16* act16;
  long ld16, s16b;
  float alpha;
  // Split K (reduced-batch or single-batch products only): workgroup z sums its share of
  // the (batch, K tile) iterations into partial[z] (M x N, raw); a second kernel adds the
  // shares in split order and applies the epilogue.
  int splits;
  float* partial;
};

constexpr int kBM = 64, kBN = 64, kBK = 32, kPad = 8;

__global__ void __launch_bounds__(256) gemm_bf16(GemmArgs g) {
  __shared__ bf16 As[kBM][kBK + kPad];
  __shared__ bf16 Bs[kBN][kBK + kPad];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  v4f acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  // The (batch, K tile) iterations this workgroup sums, in order.
  const int k_tiles = (g.K + kBK - 1) / kBK;
  const int total = (g.reduce_batch ? g.batch : 1) * k_tiles;
  int it0 = 0, it1 = total, b_base = g.reduce_batch ? 0 : blockIdx.z;
  if (g.splits > 1) {
    const int per = (total + g.splits - 1) / g.splits;
    it0 = blockIdx.z * per;
    it1 = min(total, it0 + per);
    b_base = 0;
  }
  for (int it = it0; it < it1; ++it) {
    const int b = b_base + it / k_tiles;
    const int k0 = (it % k_tiles) * kBK;
    const bf16* A = g.A + b * g.sAb;
    const bf16* B = g.B + b * g.sBb;
    {
      // A tile: 64 x 32, eight elements a thread along the unit-stride axis.
      if (g.sAm == 1 && g.sAk != 1) {
        const int k = tid >> 3, mc = (tid & 7) * 8;
        for (int j = 0; j < 8; ++j) {
          const int m = m0 + mc + j, kk = k0 + k;
          As[mc + j][k] = (m < g.M && kk < g.K) ? A[m + kk * g.sAk] : bf16(0);
        }
      } else {
        const int m = tid >> 2, kc = (tid & 3) * 8;
        for (int j = 0; j < 8; ++j) {
          const int mm = m0 + m, kk = k0 + kc + j;
          As[m][kc + j] = (mm < g.M && kk < g.K) ? A[mm * g.sAm + kk * g.sAk] : bf16(0);
        }
      }
      // B tile: 32 x 64 into Bs[n][k].
      if (g.sBn == 1 && g.sBk != 1) {
        const int k = tid >> 3, nc = (tid & 7) * 8;
        for (int j = 0; j < 8; ++j) {
          const int n = n0 + nc + j, kk = k0 + k;
          Bs[nc + j][k] = (n < g.N && kk < g.K) ? B[kk * g.sBk + n] : bf16(0);
        }
      } else {
        const int n = tid >> 2, kc = (tid & 3) * 8;
        for (int j = 0; j < 8; ++j) {
          const int nn = n0 + n, kk = k0 + kc + j;
          Bs[n][kc + j] = (nn < g.N && kk < g.K) ? B[kk * g.sBk + nn * g.sBn] : bf16(0);
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < kBK; ks += 16) {
        const int kq = ks + 4 * (lane >> 4);
        v4s a[2], bb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int r = wm + 16 * i + (lane & 15);
          a[i] = v4s{static_cast<short>(As[r][kq]), static_cast<short>(As[r][kq + 1]),
                     static_cast<short>(As[r][kq + 2]), static_cast<short>(As[r][kq + 3])};
          const int c = wn + 16 * i + (lane & 15);
          bb[i] = v4s{static_cast<short>(Bs[c][kq]), static_cast<short>(Bs[c][kq + 1]),
                      static_cast<short>(Bs[c][kq + 2]), static_cast<short>(Bs[c][kq + 3])};
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], bb[j], acc[i][j], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  // Epilogue: lane holds rows 4 (lane / 16) + r of column lane % 16 of each 16 x 16 tile.
  if (g.splits > 1) {
    float* part = g.partial + static_cast<long>(blockIdx.z) * g.M * g.N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
          const int n = n0 + wn + 16 * j + (lane & 15);
          if (m < g.M && n < g.N) part[static_cast<long>(m) * g.N + n] = acc[i][j][r];
        }
    return;
  }
  const int bc = g.reduce_batch ? 0 : blockIdx.z;
  float* C = g.C + bc * g.sCb;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = g.alpha * acc[i][j][r];
        float* at = C + static_cast<long>(m) * g.ldc + n;
        if (g.accumulate) v += *at;
        if (g.bias) v += g.bias[m];
        *at = v;
        if (g.act16)
          g.act16[bc * g.s16b + static_cast<long>(m) * g.ld16 + n] =
              to_bf16(g.act == 1 ? gelu(v) : v);
      }
}

// The split shares summed in split order, then the epilogue (one thread an element).
__global__ void splitk_reduce(GemmArgs g) {
  const long e = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<long>(g.M) * g.N) return;
  const int m = static_cast<int>(e / g.N), n = static_cast<int>(e % g.N);
  float acc = 0.f;
  for (int z = 0; z < g.splits; ++z) acc += g.partial[static_cast<long>(z) * g.M * g.N + e];
  float v = g.alpha * acc;
  float* at = g.C + static_cast<long>(m) * g.ldc + n;
  if (g.accumulate) v += *at;
  if (g.bias) v += g.bias[m];
  *at = v;
  if (g.act16) g.act16[static_cast<long>(m) * g.ld16 + n] = to_bf16(g.act == 1 ? gelu(v) : v);
}

// Launch one product; a reduced-batch or single-batch product with too few output tiles
// to fill the chip splits its K iterations over up to `max_splits` workgroups a tile
// (`scratch` holds max_splits x M x N floats at most `scratch_floats`).
void launch_gemm(GemmArgs g, hipStream_t s, float* scratch = nullptr, long scratch_floats = 0) {
  SP_CHECK(g.M > 0 && g.N > 0 && g.K > 0 && g.batch > 0, "GEMM dimensions must be positive");
  const int gx = (g.N + kBN - 1) / kBN, gy = (g.M + kBM - 1) / kBM;
  g.splits = 1;
  if (scratch && (g.reduce_batch || g.batch == 1)) {
    const long tiles = static_cast<long>(gx) * gy;
    const long iters = static_cast<long>(g.reduce_batch ? g.batch : 1) * ((g.K + kBK - 1) / kBK);
    long want = (512 + tiles - 1) / tiles;
    want = std::min(want, iters / 4);
    want = std::min(want, scratch_floats / (static_cast<long>(g.M) * g.N));
    want = std::min<long>(want, 1024);
    if (want > 1) {
      g.splits = static_cast<int>(want);
      g.partial = scratch;
    }
  }
  const dim3 grid(gx, gy, g.splits > 1 ? g.splits : (g.reduce_batch ? 1 : g.batch));
  hipLaunchKernelGGL(gemm_bf16, grid, dim3(256), 0, s, g);
  SP_HIP(hipGetLastError());
  if (g.splits > 1) {
    const long count = static_cast<long>(g.M) * g.N;
    hipLaunchKernelGGL(splitk_reduce, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                       0, s, g);
    SP_HIP(hipGetLastError());
  }
}

GemmArgs gemm(const bf16* A, long sAm, long sAk, long sAb, const bf16* B, long sBk, long sBn,
              long sBb, float* C, long ldc, long sCb, int M, int N, int K, int batch) {
  GemmArgs g{};
  g.A = A; g.sAm = sAm; g.sAk = sAk; g.sAb = sAb;
  g.B = B; g.sBk = sBk; g.sBn = sBn; g.sBb = sBb;
  g.C = C; g.ldc = ldc; g.sCb = sCb;
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.alpha = 1.0f;
  return g;
}

// ------------------------------------------------------------------ elementwise

// v0[b, c, s] = P[c, 0] spec[b, s] + sum_j P[c, 1 + j] prm[b, j] + P[c, 1 + np] s / n + p[c]
__global__ void lift_forward(const float* __restrict__ spec, const float* __restrict__ prm,
                             const float* __restrict__ pw, const float* __restrict__ pb,
                             bf16* __restrict__ v16, int W, int n, int np) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y, b = blockIdx.z;
  if (s >= n) return;
  const int in_ch = np + 2;
  float v = pw[c * in_ch] * spec[static_cast<long>(b) * n + s] + pb[c];
  for (int j = 0; j < np; ++j) v += pw[c * in_ch + 1 + j] * prm[b * np + j];
  v += pw[c * in_ch + 1 + np] * (static_cast<float>(s) / static_cast<float>(n));
  v16[(static_cast<long>(b) * W + c) * n + s] = to_bf16(v);
}

// Block (b, k), thread o: y[b, o, k] = sum_i vhat[b, i, k] R[k, i, o] (complex).
__global__ void mix_forward(const float* __restrict__ vhat, const float* __restrict__ rre,
                            const float* __restrict__ rim, bf16* __restrict__ y16, int W, int M) {
  const int k = blockIdx.x, b = blockIdx.y, o = threadIdx.x;
  if (o >= W) return;
  const float* vb = vhat + static_cast<long>(b) * W * 2 * M;
  float yr = 0.f, yi = 0.f;
  for (int i = 0; i < W; ++i) {
    const float vr = vb[i * 2 * M + k], vi = vb[i * 2 * M + M + k];
    const long at = (static_cast<long>(k) * W + i) * W + o;
    const float rr = rre[at], ri = rim[at];
    yr += vr * rr - vi * ri;
    yi += vr * ri + vi * rr;
  }
  bf16* yb = y16 + (static_cast<long>(b) * W + o) * 2 * M;
  yb[k] = to_bf16(yr);
  yb[M + k] = to_bf16(yi);
}

// Block (b, k), thread i: dvhat[b, i, k] from dy[b, o, k] and R[k, i, o].
__global__ void mix_backward_input(const float* __restrict__ dy, const float* __restrict__ rre,
                                   const float* __restrict__ rim, bf16* __restrict__ dv16, int W,
                                   int M) {
  const int k = blockIdx.x, b = blockIdx.y, i = threadIdx.x;
  if (i >= W) return;
  const float* db = dy + static_cast<long>(b) * W * 2 * M;
  float gr = 0.f, gi = 0.f;
  for (int o = 0; o < W; ++o) {
    const float yr = db[o * 2 * M + k], yi = db[o * 2 * M + M + k];
    const long at = (static_cast<long>(k) * W + i) * W + o;
    const float rr = rre[at], ri = rim[at];
    gr += yr * rr + yi * ri;
    gi += -yr * ri + yi * rr;
  }
  bf16* out = dv16 + (static_cast<long>(b) * W + i) * 2 * M;
  out[k] = to_bf16(gr);
  out[M + k] = to_bf16(gi);
}

// Block (k, i), thread o: dR[k, i, o] summed over the batch.
__global__ void mix_backward_weight(const float* __restrict__ vhat, const float* __restrict__ dy,
                                    float* __restrict__ gre, float* __restrict__ gim, int W,
                                    int M, int batch) {
  const int k = blockIdx.x, i = blockIdx.y, o = threadIdx.x;
  if (o >= W) return;
  float ar = 0.f, ai = 0.f;
  for (int b = 0; b < batch; ++b) {
    const float* vb = vhat + (static_cast<long>(b) * W + i) * 2 * M;
    const float* db = dy + (static_cast<long>(b) * W + o) * 2 * M;
    const float vr = vb[k], vi = vb[M + k], yr = db[k], yi = db[M + k];
    ar += vr * yr + vi * yi;
    ai += vr * yi - vi * yr;
  }
  const long at = (static_cast<long>(k) * W + i) * W + o;
  gre[at] = ar;
  gim[at] = ai;
}

// out[b, s] = sum_h q2[h] h16[b, h, s] + q2b.
__global__ void head_out(const bf16* __restrict__ h16, const float* __restrict__ q2,
                         const float* __restrict__ q2b, float* __restrict__ out, int H, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (s >= n) return;
  const bf16* hb = h16 + static_cast<long>(b) * H * n + s;
  float v = q2b[0];
  for (int h = 0; h < H; ++h) v += q2[h] * from_bf16(hb[static_cast<long>(h) * n]);
  out[static_cast<long>(b) * n + s] = v;
}

// dout = 2 (out - target) / count, with per-block partial sums of the squared error.
__global__ void loss_grad(const float* __restrict__ out, const float* __restrict__ tgt,
                          float* __restrict__ dout, float* __restrict__ partials, long count) {
  __shared__ float red[256];
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (i < count) {
    const float d = out[i] - tgt[i];
    dout[i] = 2.0f * d / static_cast<float>(count);
    e = d * d;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void loss_total(const float* __restrict__ partials, int n, long count,
                           float* __restrict__ loss) {
  __shared__ float red[256];
  float e = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) e += partials[i];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / static_cast<float>(count);
}

// dh1[b, h, s] = q2[h] dout[b, s] gelu'(h1[b, h, s]), FP32 and BF16.
__global__ void head_backward(const float* __restrict__ dout, const float* __restrict__ q2,
                              const float* __restrict__ h1, float* __restrict__ dh1,
                              bf16* __restrict__ dh116, int H, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x, h = blockIdx.y, b = blockIdx.z;
  if (s >= n) return;
  const long at = (static_cast<long>(b) * H + h) * n + s;
  const float d = q2[h] * dout[static_cast<long>(b) * n + s] * gelu_grad(h1[at]);
  dh1[at] = d;
  dh116[at] = to_bf16(d);
}

// dz = dv gelu'(z) (or dv for the linear last layer), FP32 and BF16.
__global__ void act_backward(const float* __restrict__ dv, const float* __restrict__ z,
                             float* __restrict__ dz, bf16* __restrict__ dz16, long count,
                             int linear) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float d = linear ? dv[i] : dv[i] * gelu_grad(z[i]);
  dz[i] = d;
  dz16[i] = to_bf16(d);
}

// out[c] = sum over (b, s) of x[b, c, s] (one workgroup a channel, fixed order).
__global__ void channel_sum(const float* __restrict__ x, float* __restrict__ out, int C, int n,
                            int batch) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float acc = 0.f;
  for (int b = 0; b < batch; ++b) {
    const float* row = x + (static_cast<long>(b) * C + c) * n;
    for (int s = threadIdx.x; s < n; s += blockDim.x) acc += row[s];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// dq2[h] = sum over (b, s) of dout[b, s] h16[b, h, s]; block H computes dq2b.
__global__ void head_weight_grad(const float* __restrict__ dout, const bf16* __restrict__ h16,
                                 float* __restrict__ dq2, float* __restrict__ dq2b, int H, int n,
                                 int batch) {
  __shared__ float red[256];
  const int h = blockIdx.x;
  float acc = 0.f;
  for (int b = 0; b < batch; ++b)
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
      const float d = dout[static_cast<long>(b) * n + s];
      acc += h < H ? d * from_bf16(h16[(static_cast<long>(b) * H + h) * n + s]) : d;
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (h < H)
      dq2[h] = red[0];
    else
      *dq2b = red[0];
  }
}

// The lift's gradient, one workgroup a channel c: dP[c, 0] = sum dv spec, dP[c, 1 + j] =
// sum_b prm[b, j] sum_s dv, dP[c, 1 + np] = sum dv s / n, dp[c] = sum dv.
__global__ void lift_backward(const float* __restrict__ dv, const float* __restrict__ spec,
                              const float* __restrict__ prm, float* __restrict__ dpw,
                              float* __restrict__ dpb, int W, int n, int np, int batch) {
  __shared__ float red[4][256];
  const int c = blockIdx.x;
  const int in_ch = np + 2;
  float a_spec = 0.f, a_grid = 0.f, a_sum = 0.f;
  float a_prm[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int b = 0; b < batch; ++b) {
    const float* row = dv + (static_cast<long>(b) * W + c) * n;
    float row_sum = 0.f;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
      const float d = row[s];
      a_spec += d * spec[static_cast<long>(b) * n + s];
      a_grid += d * (static_cast<float>(s) / static_cast<float>(n));
      row_sum += d;
    }
    a_sum += row_sum;
    for (int j = 0; j < np && j < 8; ++j) a_prm[j] += row_sum * prm[b * np + j];
  }
  auto reduce = [&](float v, int slot) {
    red[slot][threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[slot][threadIdx.x] += red[slot][threadIdx.x + w];
      __syncthreads();
    }
    const float r = red[slot][0];
    __syncthreads();
    return r;
  };
  const float t_spec = reduce(a_spec, 0);
  const float t_grid = reduce(a_grid, 1);
  const float t_sum = reduce(a_sum, 2);
  if (threadIdx.x == 0) {
    dpw[c * in_ch] = t_spec;
    dpw[c * in_ch + 1 + np] = t_grid;
    dpb[c] = t_sum;
  }
  for (int j = 0; j < np && j < 8; ++j) {
    const float t = reduce(a_prm[j], 3);
    if (threadIdx.x == 0) dpw[c * in_ch + 1 + j] = t;
  }
}

// Adam (Kingma and Ba 2015) with the bias corrections folded into the step size, and
// the BF16 mirror refreshed.
__global__ void adam(float* __restrict__ p, bf16* __restrict__ p16, const float* __restrict__ g,
                     float* __restrict__ m, float* __restrict__ v, long count, float step,
                     float beta1, float beta2, float eps) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - step * mi / (sqrtf(vi) + eps);
  p[i] = pi;
  p16[i] = to_bf16(pi);
}

__global__ void to_bf16_kernel(const float* __restrict__ x, bf16* __restrict__ y, long count) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) y[i] = to_bf16(x[i]);
}

// Rows `idx[r]` of the dataset into the batch buffers.
__global__ void gather(const int* __restrict__ idx, const float* __restrict__ spec,
                       const float* __restrict__ prm, const float* __restrict__ tgt,
                       float* __restrict__ bspec, float* __restrict__ bprm,
                       float* __restrict__ btgt, int n, int np) {
  const int r = blockIdx.y;
  const long row = idx[r];
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    bspec[static_cast<long>(r) * n + s] = spec[row * n + s];
    btgt[static_cast<long>(r) * n + s] = tgt[row * n + s];
  }
  if (blockIdx.x == 0 && threadIdx.x < np) bprm[r * np + threadIdx.x] = prm[row * np + threadIdx.x];
}

unsigned blocks(long n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

}  // namespace

Offsets offsets(const Shape& s) {
  Offsets o{};
  long at = 0;
  auto take = [&](long n) {
    const long here = at;
    at += n;
    return here;
  };
  const long W = s.width, M = s.modes;
  o.lift_w = take(W * s.in_ch());
  o.lift_b = take(W);
  for (int l = 0; l < s.layers; ++l) {
    o.r_re.push_back(take(M * W * W));
    o.r_im.push_back(take(M * W * W));
    o.w.push_back(take(W * W));
    o.b.push_back(take(W));
  }
  o.q1_w = take(static_cast<long>(s.hidden) * W);
  o.q1_b = take(s.hidden);
  o.q2_w = take(s.hidden);
  o.q2_b = take(1);
  o.total = at;
  return o;
}

template <typename T>
T* Network::alloc(std::size_t n) {
  void* p = nullptr;
  SP_HIP(hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(T)));
  owned_.push_back(p);
  SP_HIP(hipMemsetAsync(p, 0, std::max<std::size_t>(n, 1) * sizeof(T), stream_));
  return static_cast<T*>(p);
}

Network::Network(const Shape& s, int max_batch) : s_(s), off_(offsets(s)), max_batch_(max_batch) {
  SP_CHECK(s.n_params >= 0 && s.n_params <= 8, "a surrogate takes at most 8 scan parameters");
  SP_CHECK(s.width >= 1 && s.width <= 256, "width must lie in 1..256 (one workgroup a mode)");
  SP_CHECK(s.modes >= 1 && 2 * s.modes <= s.n + 2 && s.modes <= s.n / 2 + 1,
           "the retained modes must fit the grid's half spectrum");
  SP_CHECK(s.layers >= 1 && s.hidden >= 1 && s.n >= 4, "layers, hidden and n must be positive");
  SP_CHECK(max_batch >= 1 && max_batch <= 4096, "max_batch must lie in 1..4096");
  try {
    SP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const long P = off_.total;
    param_ = alloc<float>(P);
    grad_ = alloc<float>(P);
    m_ = alloc<float>(P);
    v_ = alloc<float>(P);
    param16_ = alloc<bf16>(P);
    const long n = s.n, M = s.modes, W = s.width, H = s.hidden, B = max_batch;
    // Transforms, rounded to BF16 on the host from FP64 angles reduced exactly.
    std::vector<bf16> f(n * 2 * M), g(2 * M * n);
    const double two_pi = 6.283185307179586;
    for (long t = 0; t < n; ++t)
      for (long k = 0; k < M; ++k) {
        const double a = two_pi * static_cast<double>((t * k) % n) / static_cast<double>(n);
        const double c = std::cos(a), sn = std::sin(a);
        const double w = (k == 0 ? 1.0 : 2.0) / static_cast<double>(n);
        f[t * 2 * M + k] = to_bf16(static_cast<float>(c));
        f[t * 2 * M + M + k] = to_bf16(static_cast<float>(-sn));
        g[k * n + t] = to_bf16(static_cast<float>(w * c));
        g[(M + k) * n + t] = to_bf16(static_cast<float>(-w * sn));
      }
    f16_ = alloc<bf16>(f.size());
    g16_ = alloc<bf16>(g.size());
    SP_HIP(hipMemcpyAsync(f16_, f.data(), f.size() * 2, hipMemcpyHostToDevice, stream_));
    SP_HIP(hipMemcpyAsync(g16_, g.data(), g.size() * 2, hipMemcpyHostToDevice, stream_));
    spec_ = alloc<float>(B * n);
    prm_ = alloc<float>(B * std::max(1, s.n_params));
    tgt_ = alloc<float>(B * n);
    out_ = alloc<float>(B * n);
    for (int l = 0; l <= s.layers; ++l) v16_.push_back(alloc<bf16>(B * W * n));
    for (int l = 0; l < s.layers; ++l) {
      z_.push_back(alloc<float>(B * W * n));
      vhat_.push_back(alloc<float>(B * W * 2 * M));
    }
    yhat16_ = alloc<bf16>(B * W * 2 * M);
    h1_ = alloc<float>(B * H * n);
    h16_ = alloc<bf16>(B * H * n);
    dout_ = alloc<float>(B * n);
    dout16_ = alloc<bf16>(B * n);
    dv_ = alloc<float>(B * W * n);
    dz_ = alloc<float>(B * W * n);
    dz16_ = alloc<bf16>(B * W * n);
    dyhat_ = alloc<float>(B * W * 2 * M);
    dvhat16_ = alloc<bf16>(B * W * 2 * M);
    dh1_ = alloc<float>(B * H * n);
    dh116_ = alloc<bf16>(B * H * n);
    partials_ = alloc<float>(blocks(B * n, 256) + 1);  // the partial sums, then the loss
    split_ = alloc<float>(kSplitFloats);
    SP_HIP(hipStreamSynchronize(stream_));
  } catch (...) {
    for (void* p : owned_) (void)hipFree(p);
    owned_.clear();
    if (stream_) (void)hipStreamDestroy(stream_);
    throw;
  }
}

Network::~Network() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (void* p : owned_) (void)hipFree(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Network::set_parameters(const float* host) {
  SP_CHECK(host != nullptr, "parameters needed");
  SP_HIP(hipMemcpyAsync(param_, host, off_.total * sizeof(float), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(off_.total, 256)), dim3(256), 0, stream_, param_,
                     param16_, off_.total);
  SP_HIP(hipGetLastError());
  SP_HIP(hipStreamSynchronize(stream_));
}

void Network::get_parameters(float* host) {
  SP_HIP(hipMemcpyAsync(host, param_, off_.total * sizeof(float), hipMemcpyDeviceToHost, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
}

void Network::reset_optimizer() {
  SP_HIP(hipMemsetAsync(m_, 0, off_.total * sizeof(float), stream_));
  SP_HIP(hipMemsetAsync(v_, 0, off_.total * sizeof(float), stream_));
  adam_t_ = 0;
}

void Network::upload_batch(const float* spectra, const float* params, const float* target,
                           int batch) {
  SP_CHECK(spectra && (params || s_.n_params == 0), "inputs needed");
  SP_CHECK(batch >= 1 && batch <= max_batch_, "batch must lie in 1..max_batch");
  const long n = s_.n;
  SP_HIP(hipMemcpyAsync(spec_, spectra, batch * n * sizeof(float), hipMemcpyHostToDevice, stream_));
  if (s_.n_params)
    SP_HIP(hipMemcpyAsync(prm_, params, batch * s_.n_params * sizeof(float),
                          hipMemcpyHostToDevice, stream_));
  if (target)
    SP_HIP(hipMemcpyAsync(tgt_, target, batch * n * sizeof(float), hipMemcpyHostToDevice, stream_));
}

void Network::run_forward(int B) {
  const int n = s_.n, M = s_.modes, W = s_.width, H = s_.hidden;
  hipLaunchKernelGGL(lift_forward, dim3(blocks(n, 256), W, B), dim3(256), 0, stream_, spec_, prm_,
                     param_ + off_.lift_w, param_ + off_.lift_b, v16_[0], W, n, s_.n_params);
  SP_HIP(hipGetLastError());
  for (int l = 0; l < s_.layers; ++l) {
    // vhat = v F over all (b, channel) rows: (B W, n) x (n, 2M).
    launch_gemm(gemm(v16_[l], n, 1, 0, f16_, 2 * M, 1, 0, vhat_[l], 2 * M, 0, B * W, 2 * M, n, 1),
                stream_, split_, kSplitFloats);
    hipLaunchKernelGGL(mix_forward, dim3(M, B), dim3(W), 0, stream_, vhat_[l],
                       param_ + off_.r_re[l], param_ + off_.r_im[l], yhat16_, W, M);
    SP_HIP(hipGetLastError());
    // z = yhat G: (B W, 2M) x (2M, n).
    launch_gemm(gemm(yhat16_, 2 * M, 1, 0, g16_, n, 1, 0, z_[l], n, 0, B * W, n, 2 * M, 1), stream_, split_, kSplitFloats);
    // z += W v + b per sample; the activation (gelu, or none on the last layer) in BF16.
    GemmArgs g = gemm(param16_ + off_.w[l], W, 1, 0, v16_[l], n, 1, static_cast<long>(W) * n,
                      z_[l], n, static_cast<long>(W) * n, W, n, W, B);
    g.accumulate = 1;
    g.bias = param_ + off_.b[l];
    g.act = l < s_.layers - 1 ? 1 : 0;
    g.act16 = v16_[l + 1];
    g.ld16 = n;
    g.s16b = static_cast<long>(W) * n;
    launch_gemm(g, stream_, split_, kSplitFloats);
  }
  // Head: h1 = Q1 v_L + q1 (gelu in BF16), out = q2 . h + q2b.
  GemmArgs g = gemm(param16_ + off_.q1_w, W, 1, 0, v16_[s_.layers], n, 1, static_cast<long>(W) * n,
                    h1_, n, static_cast<long>(H) * n, H, n, W, B);
  g.bias = param_ + off_.q1_b;
  g.act = 1;
  g.act16 = h16_;
  g.ld16 = n;
  g.s16b = static_cast<long>(H) * n;
  launch_gemm(g, stream_, split_, kSplitFloats);
  hipLaunchKernelGGL(head_out, dim3(blocks(n, 256), B), dim3(256), 0, stream_, h16_,
                     param_ + off_.q2_w, param_ + off_.q2_b, out_, H, n);
  SP_HIP(hipGetLastError());
}

void Network::run_backward(int B) {
  const int n = s_.n, M = s_.modes, W = s_.width, H = s_.hidden;
  const long Wn = static_cast<long>(W) * n;
  // Head.
  hipLaunchKernelGGL(head_weight_grad, dim3(H + 1), dim3(256), 0, stream_, dout_, h16_,
                     grad_ + off_.q2_w, grad_ + off_.q2_b, H, n, B);
  SP_HIP(hipGetLastError());
  hipLaunchKernelGGL(head_backward, dim3(blocks(n, 256), H, B), dim3(256), 0, stream_, dout_,
                     param_ + off_.q2_w, h1_, dh1_, dh116_, H, n);
  SP_HIP(hipGetLastError());
  // dQ1 = sum_b dh1_b v_L,b^T: (H, n) x (n, W), the batch reduced.
  GemmArgs g = gemm(dh116_, n, 1, static_cast<long>(H) * n, v16_[s_.layers], 1, n, Wn,
                    grad_ + off_.q1_w, W, 0, H, W, n, B);
  g.reduce_batch = 1;
  launch_gemm(g, stream_, split_, kSplitFloats);
  hipLaunchKernelGGL(channel_sum, dim3(H), dim3(256), 0, stream_, dh1_, grad_ + off_.q1_b, H, n, B);
  SP_HIP(hipGetLastError());
  // dv_L = Q1^T dh1: (W, H) x (H, n) per sample.
  launch_gemm(gemm(param16_ + off_.q1_w, 1, W, 0, dh116_, n, 1, static_cast<long>(H) * n, dv_, n,
                   Wn, W, n, H, B),
              stream_, split_, kSplitFloats);
  for (int l = s_.layers - 1; l >= 0; --l) {
    const long count = static_cast<long>(B) * Wn;
    hipLaunchKernelGGL(act_backward, dim3(blocks(count, 256)), dim3(256), 0, stream_, dv_, z_[l],
                       dz_, dz16_, count, l == s_.layers - 1 ? 1 : 0);
    SP_HIP(hipGetLastError());
    // dW = sum_b dz_b v_b^T, db = sum dz.
    GemmArgs gw = gemm(dz16_, n, 1, Wn, v16_[l], 1, n, Wn, grad_ + off_.w[l], W, 0, W, W, n, B);
    gw.reduce_batch = 1;
    launch_gemm(gw, stream_, split_, kSplitFloats);
    hipLaunchKernelGGL(channel_sum, dim3(W), dim3(256), 0, stream_, dz_, grad_ + off_.b[l], W, n, B);
    SP_HIP(hipGetLastError());
    // dyhat = dz G^T: (B W, n) x (n, 2M).
    launch_gemm(gemm(dz16_, n, 1, 0, g16_, 1, n, 0, dyhat_, 2 * M, 0, B * W, 2 * M, n, 1), stream_, split_, kSplitFloats);
    hipLaunchKernelGGL(mix_backward_weight, dim3(M, W), dim3(W), 0, stream_, vhat_[l], dyhat_,
                       grad_ + off_.r_re[l], grad_ + off_.r_im[l], W, M, B);
    SP_HIP(hipGetLastError());
    hipLaunchKernelGGL(mix_backward_input, dim3(M, B), dim3(W), 0, stream_, dyhat_,
                       param_ + off_.r_re[l], param_ + off_.r_im[l], dvhat16_, W, M);
    SP_HIP(hipGetLastError());
    // dv = W^T dz (per sample), then += dvhat F^T: (B W, 2M) x (2M, n).
    launch_gemm(gemm(param16_ + off_.w[l], 1, W, 0, dz16_, n, 1, Wn, dv_, n, Wn, W, n, W, B),
                stream_, split_, kSplitFloats);
    GemmArgs gf = gemm(dvhat16_, 2 * M, 1, 0, f16_, 1, 2 * M, 0, dv_, n, 0, B * W, n, 2 * M, 1);
    gf.accumulate = 1;
    launch_gemm(gf, stream_, split_, kSplitFloats);
  }
  hipLaunchKernelGGL(lift_backward, dim3(W), dim3(256), 0, stream_, dv_, spec_, prm_,
                     grad_ + off_.lift_w, grad_ + off_.lift_b, W, n, s_.n_params, B);
  SP_HIP(hipGetLastError());
}

void Network::forward(const float* spectra, const float* params, int batch, float* out) {
  SP_CHECK(out != nullptr, "an output is needed");
  upload_batch(spectra, params, nullptr, batch);
  run_forward(batch);
  SP_HIP(hipMemcpyAsync(out, out_, static_cast<long>(batch) * s_.n * sizeof(float),
                        hipMemcpyDeviceToHost, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
}

double Network::gradient(const float* spectra, const float* params, const float* target,
                         int batch, float* grad) {
  SP_CHECK(target && grad, "a target and a gradient output are needed");
  upload_batch(spectra, params, target, batch);
  run_forward(batch);
  const long count = static_cast<long>(batch) * s_.n;
  const unsigned nb = blocks(count, 256);
  hipLaunchKernelGGL(loss_grad, dim3(nb), dim3(256), 0, stream_, out_, tgt_, dout_, partials_, count);
  hipLaunchKernelGGL(loss_total, dim3(1), dim3(256), 0, stream_, partials_, static_cast<int>(nb),
                     count, partials_ + nb);
  SP_HIP(hipGetLastError());
  run_backward(batch);
  float loss = 0.f;
  SP_HIP(hipMemcpyAsync(&loss, partials_ + nb, sizeof(float), hipMemcpyDeviceToHost, stream_));
  SP_HIP(hipMemcpyAsync(grad, grad_, off_.total * sizeof(float), hipMemcpyDeviceToHost, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
  return loss;
}

void Network::load_dataset(const float* spectra, const float* params, const float* targets,
                           long n) {
  SP_CHECK(spectra && targets && (params || s_.n_params == 0) && n >= 1, "a dataset is needed");
  SP_HIP(hipStreamSynchronize(stream_));
  for (float* p : {ds_spec_, ds_prm_, ds_tgt_})
    if (p) {
      (void)hipFree(p);
      owned_.erase(std::remove(owned_.begin(), owned_.end(), static_cast<void*>(p)), owned_.end());
    }
  const long bins = s_.n;
  ds_spec_ = alloc<float>(n * bins);
  ds_prm_ = alloc<float>(n * std::max(1, s_.n_params));
  ds_tgt_ = alloc<float>(n * bins);
  SP_HIP(hipMemcpyAsync(ds_spec_, spectra, n * bins * sizeof(float), hipMemcpyHostToDevice, stream_));
  if (s_.n_params)
    SP_HIP(hipMemcpyAsync(ds_prm_, params, n * s_.n_params * sizeof(float),
                          hipMemcpyHostToDevice, stream_));
  SP_HIP(hipMemcpyAsync(ds_tgt_, targets, n * bins * sizeof(float), hipMemcpyHostToDevice, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
  ds_n_ = n;
}

void Network::train(const int* indices, long steps, int batch, float lr, float* losses) {
  SP_CHECK(ds_n_ > 0, "load a dataset before training");
  SP_CHECK(indices && losses && steps >= 1, "indices and a loss output are needed");
  SP_CHECK(batch >= 1 && batch <= max_batch_, "batch must lie in 1..max_batch");
  for (long i = 0; i < steps * batch; ++i)
    SP_CHECK(indices[i] >= 0 && indices[i] < ds_n_, "a training index lies outside the dataset");
  if (steps > losses_cap_) {
    losses_ = alloc<float>(steps);
    losses_cap_ = steps;
  }
  if (steps * batch > indices_cap_) {
    indices_ = alloc<int>(steps * batch);
    indices_cap_ = steps * batch;
  }
  SP_HIP(hipMemcpyAsync(indices_, indices, steps * batch * sizeof(int), hipMemcpyHostToDevice,
                        stream_));
  const long count = static_cast<long>(batch) * s_.n;
  const unsigned nb = blocks(count, 256);
  const auto t0 = std::chrono::steady_clock::now();
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  for (long step = 0; step < steps; ++step) {
    hipLaunchKernelGGL(gather, dim3(blocks(s_.n, 256), batch), dim3(256), 0, stream_,
                       indices_ + step * batch, ds_spec_, ds_prm_, ds_tgt_, spec_, prm_, tgt_, s_.n,
                       s_.n_params);
    SP_HIP(hipGetLastError());
    run_forward(batch);
    hipLaunchKernelGGL(loss_grad, dim3(nb), dim3(256), 0, stream_, out_, tgt_, dout_, partials_,
                       count);
    hipLaunchKernelGGL(loss_total, dim3(1), dim3(256), 0, stream_, partials_, static_cast<int>(nb),
                       count, losses_ + step);
    SP_HIP(hipGetLastError());
    run_backward(batch);
    adam_t_ += 1;
    const double c1 = 1.0 - std::pow(0.9, static_cast<double>(adam_t_));
    const double c2 = 1.0 - std::pow(0.999, static_cast<double>(adam_t_));
    const float step_size = static_cast<float>(lr * std::sqrt(c2) / c1);
    hipLaunchKernelGGL(adam, dim3(blocks(off_.total, 256)), dim3(256), 0, stream_, param_, param16_,
                       grad_, m_, v_, off_.total, step_size, beta1, beta2, eps);
    SP_HIP(hipGetLastError());
  }
  SP_HIP(hipMemcpyAsync(losses, losses_, steps * sizeof(float), hipMemcpyDeviceToHost, stream_));
  SP_HIP(hipStreamSynchronize(stream_));
  train_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void testing_gemm_bf16(int M, int N, int K, const float* A, const float* B, float* C) {
  SP_CHECK(A && B && C && M > 0 && N > 0 && K > 0, "GEMM operands needed");
  SP_CHECK(static_cast<long>(M) * K < (1L << 26) && static_cast<long>(K) * N < (1L << 26) &&
               static_cast<long>(M) * N < (1L << 26),
           "testing GEMM operands are limited to 2^26 elements");
  std::vector<bf16> a(static_cast<std::size_t>(M) * K), b(static_cast<std::size_t>(K) * N);
  for (std::size_t i = 0; i < a.size(); ++i) a[i] = to_bf16(A[i]);
  for (std::size_t i = 0; i < b.size(); ++i) b[i] = to_bf16(B[i]);
  hipStream_t s = nullptr;
  bf16 *da = nullptr, *db = nullptr;
  float* dc = nullptr;
  auto release = [&] {
    if (s) (void)hipStreamSynchronize(s);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dc) (void)hipFree(dc);
    if (s) (void)hipStreamDestroy(s);
  };
  try {
    SP_HIP(hipStreamCreate(&s));
    SP_HIP(hipMalloc(&da, a.size() * 2));
    SP_HIP(hipMalloc(&db, b.size() * 2));
    SP_HIP(hipMalloc(&dc, static_cast<std::size_t>(M) * N * 4));
    SP_HIP(hipMemcpy(da, a.data(), a.size() * 2, hipMemcpyHostToDevice));
    SP_HIP(hipMemcpy(db, b.data(), b.size() * 2, hipMemcpyHostToDevice));
    launch_gemm(gemm(da, K, 1, 0, db, N, 1, 0, dc, N, 0, M, N, K, 1), s);
    SP_HIP(hipStreamSynchronize(s));
    SP_HIP(hipMemcpy(C, dc, static_cast<std::size_t>(M) * N * 4, hipMemcpyDeviceToHost));
  } catch (...) {
    release();
    throw;
  }
  release();
}

}  // namespace sp::fno
