// Host construction of the constant tables the fused encode reads (twiddles of
// each digit, the twist, the radix-64 roots), laid out in one buffer.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "encode_core.hpp"
#include "field.hpp"
#include "fp64.hpp"

namespace zoda {

struct EncodeTabOffsets {
  size_t inv1 = 0, inv2 = 0, fwd1 = 0, fwd2 = 0, twist = 0, roots = 0, total = 0;
};

// The same tables (without the roots) as canonical u32, in one block, with
// offsets in words; plans that keep twiddles in LDS copy this block.
struct PackedTables {
  std::vector<uint32_t> words;
  uint32_t inv1 = 0, inv2 = 0, fwd1 = 0, fwd2 = 0, twist = 0;
};

inline PackedTables pack_u32(const std::vector<Tw>& v, const EncodeTabOffsets& off) {
  PackedTables p;
  p.inv1 = 0;
  p.inv2 = (uint32_t)(off.inv2 - off.inv1);
  p.fwd1 = (uint32_t)(off.fwd1 - off.inv1);
  p.fwd2 = (uint32_t)(off.fwd2 - off.inv1);
  p.twist = (uint32_t)(off.twist - off.inv1);
  for (size_t i = off.inv1; i < off.roots; i++) p.words.push_back((uint32_t)v[i].w);
  return p;
}

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

// The last INTT pass before the middle one (digit of R bits at bit LO, with
// `hb` row bits above it holding the outputs of the digits before).  The twist
// of row i is t = s^-1 * w_2s^f, f the frequency index at i: f = fa + A*k +
// B*m with fa the frequency part of the digits above (A = 2^hb), k this
// digit's output and m the middle digit's output (B = A * 2^R).  The factor
// s^-1 * w_2s^(fa + A*k) is constant over a middle-pass group, so it is
// multiplied into this pass's twiddle (the middle pass's inverse DFT is
// linear); w_2s^(B*m) = w_(2^(RD+1))^m stays a per-value constant.
// Entry [high << (LO+R) | low << R | k].
inline void push_folded_table(std::vector<Tw>& v, int L, int hb, int lo, int r) {
  const uint32_t w = inv_mod_host(root_of_unity((uint32_t)(lo + r)));
  const uint32_t s_inv = inv_mod_host(1u << L);
  const uint32_t w2s = root_of_unity((uint32_t)L + 1);
  const size_t base = v.size();
  v.resize(base + ((size_t)1 << (hb + lo + r)));
  for (uint32_t high = 0; high < (1u << hb); high++) {
    // frequency part of the digits above: one digit at most (the top digit),
    // whose output is the least significant part of f
    const uint32_t fa = high;
    for (uint32_t k = 0; k < (1u << r); k++) {
      const uint32_t c = mul_mod_host(s_inv, pow_mod_host(w2s, fa + ((uint64_t)k << hb)));
      const uint32_t step = pow_mod_host(w, k);
      uint32_t cur = 1;
      for (uint32_t low = 0; low < (1u << lo); low++) {
        const size_t idx = ((size_t)high << (lo + r)) | ((size_t)low << r) | k;
        v[base + idx] = make_tw(mul_mod_host(cur, c));
        cur = mul_mod_host(cur, step);
      }
    }
  }
}

inline std::vector<Tw> build_encode_tables(int L, int r1, int r2, int rd, EncodeTabOffsets* off) {
  std::vector<Tw> v;
  const int lo1 = rd + r2, lo2 = rd;
  const int D = 1 + (r1 > 0) + (r2 > 0);
  off->inv1 = v.size();
  if (D == 2) push_folded_table(v, L, 0, lo1, r1);
  else if (D == 3) push_digit_table(v, lo1, r1, true);
  off->inv2 = v.size();
  if (D == 3) push_folded_table(v, L, r1, lo2, r2);
  off->fwd1 = v.size();
  if (r1) push_digit_table(v, lo1, r1, false);
  off->fwd2 = v.size();
  if (r2) push_digit_table(v, lo2, r2, false);
  off->twist = v.size();
  if (D == 1) {
    const uint32_t s = 1u << L;
    const uint32_t s_inv = inv_mod_host(s);
    const uint32_t w2s = root_of_unity((uint32_t)L + 1);
    for (uint32_t i = 0; i < s; i++) v.push_back(make_tw(mul_mod_host(s_inv, pow_mod_host(w2s, i))));
  }
  off->roots = v.size();
  const uint32_t w64 = root_of_unity(6);
  for (int inv = 0; inv < 2; inv++) {
    const uint32_t w = inv ? inv_mod_host(w64) : w64;
    uint32_t cur = 1;
    for (int e = 0; e < 64; e++) {
      v.push_back(make_tw(cur));
      cur = mul_mod_host(cur, w);
    }
  }
  off->total = v.size();
  return v;
}

// EncodeTabs over a table buffer `base` (laid out as `off`) and its packed
// u32 copy `packed` (laid out as `pk`).
inline EncodeTabs make_encode_tabs(const Tw* base, const EncodeTabOffsets& off, const uint32_t* packed,
                                   const PackedTables& pk) {
  EncodeTabs T;
  T.inv1 = base + off.inv1;
  T.inv2 = base + off.inv2;
  T.fwd1 = base + off.fwd1;
  T.fwd2 = base + off.fwd2;
  T.twist = base + off.twist;
  T.roots = base + off.roots;
  T.packed = packed;
  T.o_inv1 = pk.inv1;
  T.o_inv2 = pk.inv2;
  T.o_fwd1 = pk.fwd1;
  T.o_fwd2 = pk.fwd2;
  T.o_twist = pk.twist;
  T.packed_words = (uint32_t)pk.words.size();
  return T;
}

}  // namespace zoda
