// FP64 GEMM with v_mfma_f64_16x16x4_f64. A workgroup of four waves (2 x 2) computes a
// TM x TN tile of C; each wave a (TM/2) x (TN/2) quarter as (TM/32) x (TN/32) MFMA
// tiles. K advances in 16-deep stages held in two LDS buffers: the next stage's
// global loads are in flight in registers while the current one feeds the matrix
// cores, so each stage costs one barrier.
//
// Operand layouts of the 16x16x4 FP64 MFMA (one double per lane for A and B, four for
// C): lane l holds A[l % 16][l / 16] of the 16 x 4 A block, B[l / 16][l % 16] of the
// 4 x 16 B block, and C[l / 16 + 4 r][l % 16] in accumulator register r.
//
// Workgroups are numbered along N first: the M tiles that share one B panel are
// gridDim.x apart, and with gridDim.x a multiple of 8 they land on one XCD, whose L2
// then serves that panel to all of them.
#include "gemm.hpp"

namespace sp {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kPad = 2;  // LDS row padding against bank conflicts

// An epilogue names, per output column n, where that column's values go (`column`), then
// writes row m of it (`put`).

// C[m][n] = v.
struct Store {
  double* C;
  long ldc;
  struct Column {
    double* p;
  };
  __device__ Column column(int n) const { return {C + n}; }
  __device__ void put(const Column& c, int m, double v) const { c.p[m * ldc] = v; }
};

// The time collars applied to the projected node streams (see gemm.hpp): a polarisation
// column scaled by g[t] into stream 0, a current column by g[t] into stream 1 and by
// g'[t] into stream 2. The factors are loaded once per column.
struct Collar {
  double* streams;
  const double* g;
  const double* gd;
  int C, T;
  struct Column {
    double* p0;
    double* p1;
    double f0, f1;
  };
  __device__ Column column(int n) const {
    if (n < T) return {streams + n, nullptr, g[n], 0.0};
    const int t = n - T;
    return {streams + static_cast<long>(C) * T + t, streams + 2L * C * T + t, g[t], gd[t]};
  }
  __device__ void put(const Column& c, int m, double v) const {
    const long at = static_cast<long>(m) * T;
    c.p0[at] = c.f0 * v;
    if (c.p1) c.p1[at] = c.f1 * v;
  }
};

// Two waves per SIMD: the register budget keeps the accumulators and the staged stage
// within 256 registers a lane.
template <int TM, int TN, class Out>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
    dgemm_kernel(int M, int N, int K, const double* __restrict__ A, long lda,
                 const double* __restrict__ B, long ldb, Out out) {
  constexpr int FM = TM / 32, FN = TN / 32;        // MFMA tiles per wave
  constexpr int LA = TM * kGemmTileK / kThreads;   // A doubles each thread stages
  constexpr int LB = TN * kGemmTileK / kThreads;   // B doubles each thread stages
  __shared__ double As[2][kGemmTileK][TM + kPad];  // As[buf][k][m]
  __shared__ double Bs[2][kGemmTileK][TN + kPad];  // Bs[buf][k][n]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);

  double ra[LA], rb[LB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = tid + i * kThreads;
      const int m = e / kGemmTileK, k = e % kGemmTileK;
      const int gm = m0 + m, gk = k0 + k;
      ra[i] = gm < M && gk < K ? A[gm * lda + gk] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = tid + i * kThreads;
      const int k = e / TN, n = e % TN;
      const int gk = k0 + k, gn = n0 + n;
      rb[i] = gk < K && gn < N ? B[gk * ldb + gn] : 0.0;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = tid + i * kThreads;
      As[buf][e % kGemmTileK][e / kGemmTileK] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = tid + i * kThreads;
      Bs[buf][e / TN][e % TN] = rb[i];
    }
  };

  double4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  fetch(0);
  stage(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += kGemmTileK) {
    const bool more = k0 + kGemmTileK < K;
    if (more) fetch(k0 + kGemmTileK);
#pragma unroll
    for (int kk = 0; kk < kGemmTileK; kk += 4) {
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = As[buf][kk + lane / 16][wm + 16 * i + lane % 16];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = Bs[buf][kk + lane / 16][wn + 16 * j + lane % 16];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // The other buffer was last read before the previous barrier.
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int gn = n0 + wn + 16 * j + lane % 16;
    if (gn < N) {
      const typename Out::Column column = out.column(gn);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm + 16 * i + lane / 16 + 4 * r;
          if (gm < M) out.put(column, gm, acc[i][j][r]);
        }
    }
  }
}

// Workgroups of a tile size over an M x N product.
template <int TM, int TN>
long tile_count(int M, int N) {
  return static_cast<long>((M + TM - 1) / TM) * ((N + TN - 1) / TN);
}

template <class Out>
void launch_any(int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                Out out, hipStream_t s) {
  SP_CHECK(M >= 1 && N >= 1 && K >= 1, "GEMM dimensions must be positive");
  // The large tiles once they alone give every CU a workgroup.
  if (tile_count<128, 128>(M, N) >= 256) {
    const dim3 grid((N + 127) / 128, (M + 127) / 128);
    hipLaunchKernelGGL((dgemm_kernel<128, 128, Out>), grid, dim3(kThreads), 0, s, M, N, K, A,
                       lda, B, ldb, out);
  } else {
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL((dgemm_kernel<64, 64, Out>), grid, dim3(kThreads), 0, s, M, N, K, A,
                       lda, B, ldb, out);
  }
  SP_HIP(hipGetLastError());
}

}  // namespace

void launch_dgemm(int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t s) {
  launch_any(M, N, K, A, lda, B, ldb, Store{C, ldc}, s);
}

void launch_collar_projection(int C, int T, int Q, const double* A, const double* B,
                              bool current, const double* g, const double* g_deriv,
                              double* streams, hipStream_t s) {
  const int width = current ? 2 * T : T;
  launch_any(C, width, Q, A, static_cast<long>(Q), B, static_cast<long>(width),
             Collar{streams, g, g_deriv, C, T}, s);
}

}  // namespace sp
