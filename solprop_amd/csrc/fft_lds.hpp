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


// ------------------------------------------------------------- radix-16 stages
//
// A pass kernel over sequences of at least 1024 points runs radix-16 Stockham stages
// instead (fft.hip): each thread takes whole 16-point butterflies through registers (a
// 4 x 4 decomposition with the W16 factors as constants), then one radix-2, 4 or 8 stage
// for what remains of L; the first stage reads HBM and the last writes it, so a
// 4096-point sequence crosses LDS twice rather than six times. Rows are padded by one
// point in 16 (`padded`): a stage stores each butterfly's outputs 16 points apart, which
// would otherwise put every lane of a wave on the same LDS banks.

__host__ __device__ constexpr int padded(int x) { return x + (x >> 4); }

template <int LOG2L, int G>
struct Shape16 {
  static constexpr int L = 1 << LOG2L;
  static constexpr int E = L * G;
  // The radix-16 stages serve sequences of at least 1024 points, whatever the sequences
  // a workgroup holds: the stages a sequence takes depend on its length alone, so that a
  // tuned launch shape changes no result.
  static constexpr bool use = L >= 1024;
  static constexpr int NT = E / 16 >= 256 ? 256 : E / 16;
  static constexpr int STRIDE = padded(L);
};

// cos and sin of 2 pi m / 16 (correctly rounded).
constexpr double kCos16[4] = {1.0, 0.92387953251128673848, 0.70710678118654752440,
                              0.38268343236508978178};

// v times exp(sign 2 pi i m / 16) for a constant m: exact for multiples of 4.
template <int M>
__device__ __forceinline__ cplx w16_mul(cplx v, int sign) {
  constexpr int m = M & 15;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (m == 4) {
    return rot(v, sign);
  } else if constexpr (m == 8) {
    return make(-v.x, -v.y);
  } else if constexpr (m == 12) {
    return rot(v, -sign);
  } else {
    // cos and sin of 2 pi m / 16 from the first quadrant's table.
    constexpr int q = m / 4, r = m % 4;
    constexpr double c0 = kCos16[r], s0 = r == 0 ? 0.0 : kCos16[4 - r];
    constexpr double c = q == 0 ? c0 : (q == 1 ? -s0 : (q == 2 ? -c0 : s0));
    constexpr double s = q == 0 ? s0 : (q == 1 ? c0 : (q == 2 ? -s0 : -c0));
    return mul(v, make(c, sign * s));
  }
}

// In-place DFTs of R values (exp(sign 2 pi i n k / R) kernels), natural order in; the
// output X[k] is left at position out_at<R>(k).
template <int R>
__host__ __device__ constexpr int out_at(int k) {
  return R == 16 ? 4 * (k % 4) + k / 4 : (R == 8 ? 4 * (k % 2) + k / 2 : k);
}

__device__ __forceinline__ void dft4_at(cplx* v, int a, int b, int c, int d, int sign) {
  const cplx a0 = add(v[a], v[c]), a1 = sub(v[a], v[c]), a2 = add(v[b], v[d]);
  const cplx a3 = rot(sub(v[b], v[d]), sign);
  v[a] = add(a0, a2);
  v[b] = add(a1, a3);
  v[c] = sub(a0, a2);
  v[d] = sub(a1, a3);
}

template <int R>
__device__ __forceinline__ void dft(cplx* v, int sign) {
  if constexpr (R == 2) {
    const cplx a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  } else if constexpr (R == 4) {
    dft4_at(v, 0, 1, 2, 3, sign);
  } else if constexpr (R == 8) {
    // n = 4 n1 + n2: two-point transforms over n1, the W8^{n2 k1} factors, then
    // four-point transforms over n2; X[k1 + 2 k2] lands at 4 k1 + k2.
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      const cplx a = v[n2], b = v[n2 + 4];
      v[n2] = add(a, b);
      v[n2 + 4] = sub(a, b);
    }
    v[5] = w16_mul<2>(v[5], sign);
    v[6] = w16_mul<4>(v[6], sign);
    v[7] = w16_mul<6>(v[7], sign);
    dft4_at(v, 0, 1, 2, 3, sign);
    dft4_at(v, 4, 5, 6, 7, sign);
  } else {
    // n = 4 n1 + n2: four-point transforms over n1, the W16^{n2 k1} factors, then over
    // n2; X[k1 + 4 k2] lands at 4 k1 + k2.
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4_at(v, n2, n2 + 4, n2 + 8, n2 + 12, sign);
    v[5] = w16_mul<1>(v[5], sign);
    v[6] = w16_mul<2>(v[6], sign);
    v[7] = w16_mul<3>(v[7], sign);
    v[9] = w16_mul<2>(v[9], sign);
    v[10] = w16_mul<4>(v[10], sign);
    v[11] = w16_mul<6>(v[11], sign);
    v[13] = w16_mul<3>(v[13], sign);
    v[14] = w16_mul<6>(v[14], sign);
    v[15] = w16_mul<9>(v[15], sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4_at(v, 4 * k1, 4 * k1 + 1, 4 * k1 + 2, 4 * k1 + 3, sign);
  }
}

}  // namespace sp
