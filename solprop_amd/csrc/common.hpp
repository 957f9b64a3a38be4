// Shared definitions of the MI355X propagation library: complex FP64 arithmetic on
// double2, error handling, and launch helpers. Device code is written for gfx950
// (64-wide wavefronts); every kernel here works in FP64 throughout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace sp {

using cplx = double2;

__host__ __device__ inline cplx make(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline cplx add(cplx a, cplx b) { return make(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx sub(cplx a, cplx b) { return make(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx mul(cplx a, cplx b) {
  return make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline cplx scale(double s, cplx a) { return make(s * a.x, s * a.y); }
__host__ __device__ inline cplx conj(cplx a) { return make(a.x, -a.y); }
__host__ __device__ inline double norm2(cplx a) { return a.x * a.x + a.y * a.y; }
// Multiply by -i (sign = -1) or +i (sign = +1).
__host__ __device__ inline cplx rot(cplx a, int sign) {
  return sign < 0 ? make(a.y, -a.x) : make(-a.y, a.x);
}

// A real inverse transform of 2M points as a complex one of M: the pair of half-spectrum
// bins (k, M - k) packed as Z = (Ya + conj Yb) + i conj(W) (Ya - conj Yb), W = W_2M^k
// (twice E + iO of the even and odd samples' transforms).
__host__ __device__ inline cplx c2r_pre(cplx ya, cplx yb, cplx w) {
  const cplx s = add(ya, conj(yb));
  const cplx d = sub(ya, conj(yb));
  return add(s, rot(mul(conj(w), d), +1));
}

// exp(z) in real form, as the reference requires (no complex exp): exp(re)(cos, sin).
__device__ inline cplx cexp(cplx z) {
  double s, c;
  sincos(z.y, &s, &c);
  const double m = exp(z.x);
  return make(m * c, m * s);
}

// exp(z) and exp(-z) sharing one sincos: (exp(a)(cos b, sin b), exp(-a)(cos b, -sin b)).
__device__ inline void cexp_pair(cplx z, cplx* forward, cplx* backward) {
  double s, c;
  sincos(z.y, &s, &c);
  const double m = exp(z.x);
  const double r = exp(-z.x);
  *forward = make(m * c, m * s);
  *backward = make(r * c, -(r * s));
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kValue = 1,       // invalid argument or shape (a refusal before any allocation)
  kBackend = 2,     // a HIP runtime failure
  kConvergence = 3, // the step floor was reached
  kState = 4,       // a call out of lifecycle order
};

#define SP_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      throw ::sp::Error(::sp::kBackend, std::string(#expr " failed: ") +              \
                                            hipGetErrorString(_e) + " at " __FILE__); \
  } while (0)

#define SP_CHECK(cond, msg)                                  \
  do {                                                       \
    if (!(cond)) throw ::sp::Error(::sp::kValue, (msg));     \
  } while (0)

// The owner of workgroup `block` in a grouped launch whose owners hold consecutive ranges:
// owner i holds first[i] .. first[i + 1] - 1 (first[0] = 0, first[n] = the grid). A binary
// search, the same in every lane of the workgroup.
__device__ inline int owner_of(const int* __restrict__ first, int n, int block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= block)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

inline unsigned blocks_for(std::size_t n, unsigned threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

}  // namespace sp
