// Exact BabyBear products on integer-valued doubles in the four-instruction
// form, shared by every FP64 kernel (the slice, column, drop-in, decoder and
// per-side encode kernels) and their CPU emulation: the register constants
// of the form (MulK), the product as a double (mul_reg) or as the int32
// that goes back to LDS (mul_mem), the canonical reduction, and the
// compile-time roots of unity w_128^(+-e).
#pragma once

#include <stdint.h>

#include "field.hpp"
#include "fp64.hpp"

namespace zoda {
namespace slice {

// Twiddles as stored for the kernels: 16-byte (w, w/p) pairs (one ds_read_b128).
struct alignas(16) Tw2 {
  double w;
  double wq;
};

// ---- compile-time roots: w_128^e and w_128^-e --------------------------------------

struct RootTable {
  double w[2][128];
  double wq[2][128];
};

constexpr RootTable make_root_table() {
  RootTable t{};
  const uint32_t r = root_of_unity(7);
  const uint32_t ri = inv_mod_host(r);
  uint64_t f = 1, b = 1;
  for (int e = 0; e < 128; e++) {
    t.w[0][e] = (double)f;
    t.wq[0][e] = (double)f / PD;
    t.w[1][e] = (double)b;
    t.wq[1][e] = (double)b / PD;
    f = f * r % P;
    b = b * ri % P;
  }
  return t;
}

inline constexpr RootTable kRoots = make_root_table();

// Register constants of the four-instruction multiplications.  gfx9 VALU
// instructions read at most one scalar operand, so the constants that meet a
// scalar twiddle or p - 1 in one instruction live in vector registers; the
// empty asm hides their values, so the compiler keeps them there instead of
// rematerialising them before every use.
struct MulK {
  double m;    // 1.5 * 2^52
  double kr;   // -M (p - 1) - M
  double eps;  // 2^-40 (canonical reduction)
  double kd;   // p - 1, in scalar registers: as a literal it turns the FMA
               // into v_fmac, whose tied addend then costs a copy
};

ZHD MulK make_mulk() {
  MulK k{ROUND_MAGIC, -ROUND_MAGIC * KD - ROUND_MAGIC, FLOOR_EPS, KD};
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(k.m), "+v"(k.kr), "+v"(k.eps), "+s"(k.kd));
#endif
  return k;
}

// x * w mod p for integer-valued |x| < 2^45, as an integer-valued double with
// |r| <= p/2 + p/256:
//   y = x w/p + M         rounds to M + q, q = round(x w/p)
//   c = y (p-1) - M(p-1) - M = q (p-1) - M   (exact: q (p-1) has <= 49
//                         significant bits from bit 27 up)
//   t = x w - c           = r + q + M, exact (|r + q| < 2^46)
//   r = t - y
ZHD double mul_reg(double x, double w, double wq, const MulK& k) {
  const double y = fmad(x, wq, k.m);
  const double c = fmad(y, k.kd, k.kr);
  const double t = fmad(x, w, -c);
  return t - y;
}

// The same product as an int32, for values that leave the thread (LDS or
// held registers).  t = M + r + q and y = M + q both lie in [2^52, 2^53),
// where a double's low 32 bits are its integer value minus 2^52 modulo 2^32,
// so lo(t) - lo(y) = r (mod 2^32) and |r| < 2^31 makes that the int32 of r.
// One integer subtraction replaces the last FP64 one and yields a fresh
// 32-bit value (the conversion costs nothing, and no 64-bit register stays
// live behind the stored word).
ZHD int32_t mul_mem(double x, double w, double wq, const MulK& k) {
  const double y = fmad(x, wq, k.m);
  const double c = fmad(y, k.kd, k.kr);
  const double t = fmad(x, w, -c);
  const uint32_t lt = (uint32_t)__builtin_bit_cast(uint64_t, t);
  const uint32_t ly = (uint32_t)__builtin_bit_cast(uint64_t, y);
  return (int32_t)(lt - ly);
}

ZHD int32_t mul_mem(double x, const Tw2& t, const MulK& k) { return mul_mem(x, t.w, t.wq, k); }

// Canonical value of an integer-valued |x| < 2^36 (floor division, fp64.hpp).
ZHD uint32_t canon(double x, const MulK& k) {
  const double q = __builtin_floor(fmad(x, PINV, k.eps));
  return (uint32_t)fmad(-q, PD, x);
}

}  // namespace slice
}  // namespace zoda
