// FP64 GEMM with v_mfma_f64_16x16x4_f64: a workgroup of four waves computes a 64 x 64
// tile of C, each wave a 32 x 32 quarter as 2 x 2 MFMA tiles, staging 16-deep slices of
// A and B through LDS.
//
// Operand layouts of the 16x16x4 FP64 MFMA (one double per lane for A and B, four for
// C): lane l holds A[l % 16][l / 16] of the 16 x 4 A block, B[l / 16][l % 16] of the
// 4 x 16 B block, and C[l / 16 + 4 r][l % 16] in accumulator register r.
#include "gemm.hpp"

namespace sp {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kPad = 1;  // LDS row padding against bank conflicts

__global__ void __launch_bounds__(kThreads)
    dgemm_kernel(int M, int N, int K, const double* __restrict__ A, long lda,
                 const double* __restrict__ B, long ldb, double* __restrict__ C, long ldc) {
  __shared__ double As[kGemmTileK][kGemmTileM + kPad];  // As[k][m]
  __shared__ double Bs[kGemmTileK][kGemmTileN + kPad];  // Bs[k][n]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kGemmTileM, n0 = blockIdx.x * kGemmTileN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  double4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < K; k0 += kGemmTileK) {
    // Stage A (64 x 16) and B (16 x 64): 1024 doubles each, four per thread.
#pragma unroll
    for (int e = tid; e < kGemmTileM * kGemmTileK; e += kThreads) {
      const int m = e / kGemmTileK, k = e % kGemmTileK;
      const int gm = m0 + m, gk = k0 + k;
      As[k][m] = gm < M && gk < K ? A[gm * lda + gk] : 0.0;
    }
#pragma unroll
    for (int e = tid; e < kGemmTileK * kGemmTileN; e += kThreads) {
      const int k = e / kGemmTileN, n = e % kGemmTileN;
      const int gk = k0 + k, gn = n0 + n;
      Bs[k][n] = gk < K && gn < N ? B[gk * ldb + gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGemmTileK; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + lane / 16][wm + 16 * i + lane % 16];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + lane / 16][wn + 16 * j + lane % 16];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * i + lane / 16 + 4 * r;
        const int gn = n0 + wn + 16 * j + lane % 16;
        if (gm < M && gn < N) C[gm * ldc + gn] = acc[i][j][r];
      }
}

}  // namespace

void launch_dgemm(int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t s) {
  SP_CHECK(M >= 1 && N >= 1 && K >= 1, "GEMM dimensions must be positive");
  const dim3 grid((N + kGemmTileN - 1) / kGemmTileN, (M + kGemmTileM - 1) / kGemmTileM);
  hipLaunchKernelGGL(dgemm_kernel, grid, dim3(kThreads), 0, s, M, N, K, A, lda, B, ldb, C,
                     ldc);
  SP_HIP(hipGetLastError());
}

}  // namespace sp
