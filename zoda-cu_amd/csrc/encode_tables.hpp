// Host construction of the constant tables the fused encode reads (twiddles of
// each digit, the twist, the radix-32 roots), laid out in one buffer.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "field.hpp"
#include "fp64.hpp"

namespace zoda {

struct EncodeTabOffsets {
  size_t inv1 = 0, inv2 = 0, fwd1 = 0, fwd2 = 0, twist = 0, roots = 0, total = 0;
};

inline Tw make_tw(uint32_t w) { return Tw{(double)w, (double)w / PD}; }

// Digit of R bits at bit LO of a size-2^(LO+R) sub-transform, value
// w_(2^(LO+R))^(+-low*k).  The inverse tables are read by the INTT pass of the
// digit itself, which needs all k for one `low`: laid out [low][k].  The
// forward tables are read by the pass below, which needs consecutive `low`
// for one k: laid out [k][low].
inline void push_digit_table(std::vector<Tw>& v, int lo, int r, bool inverse) {
  uint32_t w = root_of_unity((uint32_t)(lo + r));
  if (inverse) w = inv_mod_host(w);
  const size_t base = v.size();
  v.resize(base + ((size_t)1 << (lo + r)));
  for (uint32_t k = 0; k < (1u << r); k++) {
    const uint32_t step = pow_mod_host(w, k);  // w^k
    uint32_t cur = 1;
    for (uint32_t low = 0; low < (1u << lo); low++) {
      const size_t idx = inverse ? ((size_t)low << r) | k : ((size_t)k << lo) | low;
      v[base + idx] = make_tw(cur);
      cur = mul_mod_host(cur, step);
    }
  }
}

// Frequency index held at storage row i after the INTT passes: the top digit's
// output is the least significant part.
inline uint32_t digit_reverse(uint32_t i, int r1, int r2, int rd) {
  const int lo1 = rd + r2, lo2 = rd;
  const uint32_t d1 = r1 ? (i >> lo1) & ((1u << r1) - 1u) : 0u;
  const uint32_t d2 = r2 ? (i >> lo2) & ((1u << r2) - 1u) : 0u;
  const uint32_t dd = i & ((1u << rd) - 1u);
  return d1 | (d2 << r1) | (dd << (r1 + r2));
}

inline std::vector<Tw> build_encode_tables(int L, int r1, int r2, int rd, EncodeTabOffsets* off) {
  std::vector<Tw> v;
  const int lo1 = rd + r2, lo2 = rd;
  off->inv1 = v.size();
  if (r1) push_digit_table(v, lo1, r1, true);
  off->inv2 = v.size();
  if (r2) push_digit_table(v, lo2, r2, true);
  off->fwd1 = v.size();
  if (r1) push_digit_table(v, lo1, r1, false);
  off->fwd2 = v.size();
  if (r2) push_digit_table(v, lo2, r2, false);
  off->twist = v.size();
  const uint32_t s = 1u << L;
  const uint32_t s_inv = inv_mod_host(s);
  const uint32_t w2s = root_of_unity((uint32_t)L + 1);
  for (uint32_t i = 0; i < s; i++)
    v.push_back(make_tw(mul_mod_host(s_inv, pow_mod_host(w2s, digit_reverse(i, r1, r2, rd)))));
  off->roots = v.size();
  const uint32_t w32 = root_of_unity(5);
  for (int inv = 0; inv < 2; inv++) {
    const uint32_t w = inv ? inv_mod_host(w32) : w32;
    uint32_t cur = 1;
    for (int e = 0; e < 32; e++) {
      v.push_back(make_tw(cur));
      cur = mul_mod_host(cur, w);
    }
  }
  off->total = v.size();
  return v;
}

}  // namespace zoda
