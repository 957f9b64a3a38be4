// Dense FP64 GEMM on the gfx950 matrix cores, for the radial route's Hankel transforms:
// C (M x N) = A (M x K) . B (K x N), all row-major with leading dimensions.
#pragma once

#include "common.hpp"

namespace sp {

constexpr int kGemmTileM = 64;  // rows of C per workgroup
constexpr int kGemmTileN = 64;  // columns of C per workgroup
constexpr int kGemmTileK = 16;  // K per LDS stage

// Any M, N, K >= 1; edges are handled by zero padding in LDS.
void launch_dgemm(int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t s);

}  // namespace sp
