// The in-LDS part of the FFT passes, shared by the library's pass kernels and the fused
// kernels of the routes that transform in place inside a larger kernel: the workgroup
// shape of a pass of 2^LOG2L points over G sequences, and the radix-4/2 Stockham stages
// on G sequences staged in LDS (`buf`, G rows of STRIDE points).
#pragma once

#include "common.hpp"

namespace sp {

constexpr int kMaxElems = 4096;  // complex doubles a workgroup holds in LDS (64 KiB)

template <int LOG2L, int G>
struct Shape {
  static constexpr int L = 1 << LOG2L;
  static constexpr int E = L * G;
  static constexpr int NT = (E / 4 >= 256) ? 256 : (E / 4 >= 1 ? E / 4 : 1);
  static constexpr int STRIDE = L + (G > 1 ? 1 : 0);  // pad rows against bank conflicts
  static constexpr int PER4 = (E / 4 + NT - 1) / NT;
  static constexpr int PER2 = (E / 2 + NT - 1) / NT;
};

__device__ inline cplx twiddle(const cplx* table, int index, int sign) {
  const cplx w = table[index];
  return sign > 0 ? make(w.x, -w.y) : w;
}

// The transform of the G staged sequences in place (twiddles W_L from `tw_l`, conjugated
// for sign > 0). The caller synchronises the workgroup after staging `buf`; on return the
// workgroup is synchronised again and every thread may read any of it.
template <int LOG2L, int G>
__device__ __forceinline__ void lds_transform(cplx* buf, const cplx* __restrict__ tw_l,
                                              int sign) {
  using S = Shape<LOG2L, G>;
  constexpr int L = S::L, E = S::E, NT = S::NT, STRIDE = S::STRIDE;
  const int tid = threadIdx.x;
  int p = 1;
  // Radix-4 Stockham stages: inputs at i + m L/4, outputs at 4(i-k) + k + m p.
  for (; 4 * p <= L; p *= 4) {
    cplx v[S::PER4][4];
    const int step = L / (4 * p);
#pragma unroll
    for (int r = 0; r < S::PER4; ++r) {
      const int b = tid + r * NT;
      if (b < E / 4) {
        const int s = b / (L / 4), i = b % (L / 4);
        const int k = i & (p - 1);
        const cplx* x = buf + s * STRIDE;
        cplx u0 = x[i], u1 = x[i + L / 4], u2 = x[i + L / 2], u3 = x[i + 3 * L / 4];
        if (p > 1) {
          u1 = mul(u1, twiddle(tw_l, k * step, sign));
          u2 = mul(u2, twiddle(tw_l, 2 * k * step, sign));
          u3 = mul(u3, twiddle(tw_l, 3 * k * step, sign));
        }
        const cplx a0 = add(u0, u2), a1 = sub(u0, u2), a2 = add(u1, u3);
        const cplx a3 = rot(sub(u1, u3), sign);
        v[r][0] = add(a0, a2);
        v[r][1] = add(a1, a3);
        v[r][2] = sub(a0, a2);
        v[r][3] = sub(a1, a3);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < S::PER4; ++r) {
      const int b = tid + r * NT;
      if (b < E / 4) {
        const int s = b / (L / 4), i = b % (L / 4);
        const int k = i & (p - 1);
        cplx* y = buf + s * STRIDE + ((i - k) << 2) + k;
        y[0] = v[r][0];
        y[p] = v[r][1];
        y[2 * p] = v[r][2];
        y[3 * p] = v[r][3];
      }
    }
    __syncthreads();
  }
  if (p < L) {  // one radix-2 stage with p = L/2: output j = i
    cplx v[S::PER2][2];
#pragma unroll
    for (int r = 0; r < S::PER2; ++r) {
      const int b = tid + r * NT;
      if (b < E / 2) {
        const int s = b / (L / 2), i = b % (L / 2);
        const cplx* x = buf + s * STRIDE;
        const cplx u0 = x[i];
        const cplx u1 = mul(x[i + L / 2], twiddle(tw_l, i, sign));
        v[r][0] = add(u0, u1);
        v[r][1] = sub(u0, u1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < S::PER2; ++r) {
      const int b = tid + r * NT;
      if (b < E / 2) {
        const int s = b / (L / 2), i = b % (L / 2);
        cplx* y = buf + s * STRIDE;
        y[i] = v[r][0];
        y[i + L / 2] = v[r][1];
      }
    }
    __syncthreads();
  }

}

}  // namespace sp
