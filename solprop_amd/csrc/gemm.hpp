// Dense FP64 GEMM on the gfx950 matrix cores, for the radial route's Hankel transforms:
// C (M x N) = A (M x K) . B (K x N), all row-major with leading dimensions.
#pragma once

#include "common.hpp"

namespace sp {

constexpr int kGemmTileK = 16;  // K per LDS stage

// Any M, N, K >= 1; edges are handled by zero padding in LDS. Large products run on
// 128 x 128 tiles, small ones on 64 x 64 tiles so that the chip still gets workgroups.
void launch_dgemm(int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t s);

// The projection of the node streams of the dense field route with the time collars
// applied on the way out: with X = A (C x Q) . B (Q x 2T), B holding each node's
// un-gated polarisation (columns 0..T-1) and current (T..2T-1),
//   streams[0][c][t] = g[t] X[c][t],
//   streams[1][c][t] = g[t] X[c][T + t],  streams[2][c][t] = g'[t] X[c][T + t],
// as (3, C, T) doubles. With `current` false B holds the polarisation alone (Q x T) and
// only streams[0] is written.
void launch_collar_projection(int C, int T, int Q, const double* A, const double* B,
                              bool current, const double* g, const double* g_deriv,
                              double* streams, hipStream_t s);

}  // namespace sp
