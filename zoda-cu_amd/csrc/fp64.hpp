// Exact BabyBear arithmetic on integer-valued doubles, shared by the gfx950
// kernels and their CPU emulation.
//
// gfx950 issues FP64 add/mul/FMA at full rate, the same rate as 32-bit integer
// ops.  A field element held as an exact integer in a double needs no modular
// correction after an addition or subtraction (53-bit headroom), so a radix-2
// butterfly is one twiddle multiplication (five instructions, below) plus two
// plain FP64 add/sub, against nine or more instructions for canonical u32.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define ZHD __host__ __device__ __forceinline__
#else
#define ZHD inline
#endif

namespace zoda {

constexpr double PD = 2013265921.0;
constexpr double KD = 2013265920.0;                 // p - 1 = 15 * 2^27 (4 significant bits)
constexpr double ROUND_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
constexpr double PINV = 1.0 / PD;

// A constant multiplier w in [0, p) with its quotient w/p.
struct Tw {
  double w;
  double wq;
};

ZHD double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// For an integer-valued x with |x| < 2^45, returns an integer r == x*w (mod p)
// with |r| <= p/2 + p/256 < 2^30:
//   q = round(x * w/p)        (FMA onto 1.5*2^52 rounds to an integer; the
//                              error of w/p adds at most |x|*2^-53 < 2^-8)
//   t = fma(x, w, -q*(p-1))   the true value x*w - q*(p-1) = r + q is an
//                             integer below 2^46, so the FMA is exact, and
//                             q*(p-1) is exact since p-1 = 15*2^27
//   r = t - q                 = x*w - q*p, exact.
ZHD double mulmod(double x, double w, double wq) {
  double q = fmad(x, wq, ROUND_MAGIC) - ROUND_MAGIC;
  double t = fmad(x, w, -(q * KD));
  return t - q;
}

ZHD double mulmod(double x, Tw t) { return mulmod(x, t.w, t.wq); }

// |x| < 2^51 -> r == x (mod p) with |r| <= p/2 + 1.
ZHD double reduce(double x) {
  double q = fmad(x, PINV, ROUND_MAGIC) - ROUND_MAGIC;
  return fmad(-q, PD, x);
}

// |x| < 2^36 (an integer) -> canonical value in [0, p), by floor division:
//   y = x/p + 2^-40         one FMA; with |x/p| < 2^6 the error of 1/p adds
//                           below 2^-47 and the rounding of y at most 2^-48
//   q = floor(y)            = floor(x/p): x/p = n + j/p with 0 <= j < p, so
//                           y is within 2^-46 of a value between n + 2^-40 and
//                           n + 1 - 1/p + 2^-40, and 1/p > 2^-31
//   r = x - q p             in [0, p), exact.
constexpr double FLOOR_EPS = 0x1p-40;
ZHD uint32_t canonical(double x) {
  const double q = __builtin_floor(fmad(x, PINV, FLOOR_EPS));
  return (uint32_t)fmad(-q, PD, x);
}

}  // namespace zoda
