// BabyBear field arithmetic for the gfx950 device code and the host library.
//
// p = 15 * 2^27 + 1 = 2013265921.  Two device representations are used:
//
//  * canonical u32 in [0, p) with Montgomery multiplication by constants that
//    are stored pre-scaled by R = 2^32 (so mont_mul(x, w*R) = x*w, canonical in
//    and out, no conversion passes); used by the general batched NTT;
//  * integer-valued doubles (fp64.hpp), used by the encode kernels.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define ZHC __host__ __device__ constexpr
#else
#define ZHC constexpr
#endif

namespace zoda {

constexpr uint32_t P = 2013265921u;
constexpr uint32_t P_NEG_INV = 0x77FFFFFFu;  // -p^-1 mod 2^32
constexpr uint32_t GENERATOR = 31u;          // smallest generator of F_p^*

// ---- helpers on canonical values (host, device and constant evaluation) -------

ZHC uint32_t mul_mod_host(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) % P); }

ZHC uint32_t pow_mod_host(uint32_t b, uint64_t e) {
  uint32_t r = 1;
  b %= P;
  while (e) {
    if (e & 1) r = mul_mod_host(r, b);
    b = mul_mod_host(b, b);
    e >>= 1;
  }
  return r;
}

ZHC uint32_t inv_mod_host(uint32_t a) { return pow_mod_host(a, P - 2); }

// w * 2^32 mod p: the Montgomery form of a constant.
ZHC uint32_t to_mont_host(uint32_t w) { return (uint32_t)((((uint64_t)w) << 32) % P); }

// Primitive 2^log_n-th root of unity of the reference, 31^((p-1)/2^log_n)
// (src/ntt/cpu.rs:97-110 with the generator of src/ntt/cpu.rs:89-94).
ZHC uint32_t root_of_unity(uint32_t log_n) {
  return pow_mod_host(GENERATOR, (uint64_t)(P - 1) >> log_n);
}

#if defined(__HIPCC__)

// ---- canonical u32 device arithmetic -------------------------------------------

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b) {
  uint32_t s = a + b;
  return min(s, s - P);
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b) {
  uint32_t d = a - b;
  return min(d, d + P);
}

// a * b * 2^-32 mod p for a, b < p; canonical result.
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b) {
  uint64_t prod = (uint64_t)a * b;
  uint32_t m = (uint32_t)prod * P_NEG_INV;
  uint64_t t = (uint64_t)m * P + prod;  // < 2^64: prod < 2^62, m*p < 2^63
  uint32_t r = (uint32_t)(t >> 32);
  return min(r, r - P);
}

#endif

}  // namespace zoda
