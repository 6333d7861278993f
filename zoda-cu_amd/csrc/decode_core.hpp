// Per-thread work of the erasure decoder's column transforms, shared by the
// gfx950 kernel (decode.hip) and a CPU emulation used by the tests.
//
// A work-group holds W columns of n = 2^L values (y_r = out[r] * Z(w^r) on
// the known rows, 0 on the erased ones: the values of g = f Z at every w^r)
// in LDS as int32, and runs one transform pair on them:
//
//     v = NTT_n( f (.) INTT'_n(y) ),   (f (.) G)_f = f * G_f
//
// (INTT' without its 1/n): the coefficients of x g'(x) times n, evaluated, so
// v_r = n w^r g'(w^r).  Where row e is erased, Z(w^e) = 0 and
// g'(w^e) = f(w^e) Z'(w^e), so the data value there is v_e * s_e with
// s_e = 1 / (n w^e Z'(w^e)) (a per-row table); the known rows need no
// arithmetic.  Only the even rows (the data) leave the pair.
//
// The pair runs as five passes over three digits of the row index -- R1
// (top), R2 and RD (bottom) bits -- in the arrangement of the fused encode
// (encode_core.hpp): INTT passes on R1 and R2, a middle pass on RD (inverse
// DFT, the multiplication by f, forward DFT), NTT passes on R2 and R1.  The
// INTT passes leave their outputs in digit-reversed order, so the frequency
// index of the value at a middle-pass position is f = k1 + 2^R1 k2 +
// 2^(R1+R2) m.  The middle pass's forward DFT output index is the row's low
// RD bits; its even outputs are a DFT of half the size of the folded input
// (z_m + z_(m+H)), and from there on only the groups of even rows are
// transformed.
//
// Bounds: LDS holds |v| < 2^31 (canonical inputs, then values reduced below
// 2^30 by the constant multiplication or reduction that ends every pass); a
// register DFT of at most 32 points keeps them below 2^36, inside the 2^45
// of the products (slice::mul_reg, slice::mul_mem), and the product by
// f < 2^14 below 2^50, inside reduce's 2^51.
#pragma once

#include <stdint.h>

#include <vector>

#include "encode_core.hpp"
#include "encode_tables.hpp"
#include "slice_core.hpp"

namespace zoda {

struct DecodeTabs {
  // canonical u32, copied into LDS: inv1 | fwd1 | fwd2 | inv2
  const uint32_t* packed;
  uint32_t o_inv1, o_fwd1, o_fwd2, o_inv2, packed_words;
};

template <int L_, int W_, int NT_, int R1_, int R2_, int RD_, int MINW_ = 4>
struct DecodePlan {
  static constexpr int L = L_, W = W_, NT = NT_, R1 = R1_, R2 = R2_, RD = RD_;
  static constexpr int MINW = MINW_;  // waves per SIMD the kernel is compiled for
  static_assert(R1 > 0 && R2 > 0 && RD > 0 && R1 <= 5 && R2 <= 5 && RD <= 5, "three register digits");
  static_assert(R1 + R2 + RD == L, "digits must cover the row index");
  static constexpr int S = 1 << L;  // n
  static constexpr int LO1 = RD + R2, LO2 = RD;
  // LDS word of (row, column): one pad word every 2^RD rows, so the middle
  // pass's lanes (2^RD rows apart) fall into different banks.  (Pads chosen
  // to clear the remaining two-way conflicts of the NTT passes as well cost
  // more in address arithmetic than they saved: 8.02-8.10 against 7.70 ms
  // per 2048^2 decode.)
  static constexpr int TILE_VALUES = W * S;
  static constexpr int TILE_WORDS = W * S + (S >> RD);
  // The top digit's twiddles w^(+-low k), low < 2^LO1: one table of n
  // entries, or (FT, n >= 2^13) two factors of about sqrt(n) entries each,
  // low = lh 2^FA + ll: [k << FA | ll] then [k << (LO1 - FA) | lh].
  static constexpr bool FT = L >= 13;
  static constexpr int FA = FT ? (LO1 + 1) / 2 : LO1;
  static constexpr int TOP_WORDS = FT ? (1 << (R1 + FA)) + (1 << (R1 + LO1 - FA)) : S;
  // inv1, fwd1 (TOP_WORDS each), fwd2, inv2 (2^(LO2+R2) each)
  static constexpr int TW_WORDS = 2 * TOP_WORDS + 2 * (1 << (LO2 + R2));
  static constexpr int LDS_WORDS = TILE_WORDS + TW_WORDS;
};

template <class PL>
ZHD uint32_t lds_at(uint32_t row, uint32_t c) {
  return row * PL::W + c + (row >> PL::RD);
}

// A u32 twiddle from an LDS table as the (w, w/p) pair of the multiplication.
ZHD slice::Tw2 dec_tw(const uint32_t* __restrict__ t, uint32_t i) {
  const double w = (double)t[i];
  return slice::Tw2{w, w * PINV};
}

// The top digit's twiddle w^(+-low k) of row group `low`, digit value `k`,
// from its table(s) `t` (see DecodePlan::FT).  With FT it is two factors:
// top_first multiplies by [k << FA | ll] (mul_reg), and top_last returns
// [k << (LO1 - FA) | lh] for the multiplication that ends the pass; without,
// top_first does nothing and top_last returns the one table's entry.
template <class PL>
ZHD double top_first(double x, const uint32_t* __restrict__ t, uint32_t k, uint32_t low, const slice::MulK& mk) {
  if constexpr (PL::FT) {
    const slice::Tw2 a = dec_tw(t, (k << PL::FA) | (low & ((1u << PL::FA) - 1u)));
    return slice::mul_reg(x, a.w, a.wq, mk);
  } else {
    return x;
  }
}

template <class PL>
ZHD slice::Tw2 top_last(const uint32_t* __restrict__ t, uint32_t k, uint32_t low) {
  if constexpr (PL::FT) {
    constexpr int FA = PL::FA, HB = PL::LO1 - PL::FA;
    return dec_tw(t, (1u << (PL::R1 + FA)) + ((k << HB) | (low >> FA)));
  } else {
    return dec_tw(t, (k << PL::LO1) | low);
  }
}

// INTT pass on the digit of R bits at bit LO: inverse DFT over the digit,
// then the inter-digit twiddle: the top digit's from its table(s) `tl` (see
// DecodePlan::FT), the second digit's from the LDS table `tl` ([k << LO |
// low]: neighbouring lanes read neighbouring words).  The multiplication
// that ends the pass (times 1 for k = 0) yields the int32 LDS word directly
// (slice::mul_mem).
template <class PL, int LO, int R>
ZHD void dec_intt(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ tl, const slice::MulK& mk) {
  constexpr int RR = 1 << R;
  constexpr bool TOP = LO + R == PL::L;
  constexpr int G = (PL::W << (PL::L - R)) / PL::NT;
  static_assert((PL::W << (PL::L - R)) % PL::NT == 0, "groups must divide evenly over the threads");
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    group_fence();
    const uint32_t g = (uint32_t)(gi * PL::NT + tid);
    const uint32_t c = g % PL::W, u = g / PL::W;
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = (double)lds[lds_at<PL>(row_of<LO, R>(u, m), c)];
    dft4<R, true>(x, mk);
    const uint32_t low = u & ((1u << LO) - 1u);
    if constexpr (TOP) {
#pragma unroll
      for (int k = 1; k < RR; k++) x[k] = top_first<PL>(x[k], tl, (uint32_t)k, low, mk);
    }
    lds[lds_at<PL>(row_of<LO, R>(u, 0), c)] = slice::mul_mem(x[0], 1.0, PINV, mk);
#pragma unroll
    for (int k = 1; k < RR; k++) {
      const slice::Tw2 t = TOP ? top_last<PL>(tl, (uint32_t)k, low) : dec_tw(tl, ((uint32_t)k << LO) | low);
      lds[lds_at<PL>(row_of<LO, R>(u, k), c)] = slice::mul_mem(x[k], t, mk);
    }
  }
}

// Middle pass on the bottom digit: inverse DFT, the multiplication of every
// value by its frequency index f = k1 + 2^R1 k2 + 2^(R1+R2) m (exact below
// 2^50, then reduced), and the forward DFT's even outputs -- those of a DFT of
// half the size of the folded values z_m + z_(m+H), X[2q] = sum_{m < H}
// (z_m + z_(m+H)) (w^2)^(qm) -- each times the forward twiddle of the digit
// above (LDS table `ta`), written to the even rows.
template <class PL>
ZHD void dec_mid(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ ta, const slice::MulK& mk) {
  constexpr int R = PL::RD, RR = 1 << R, H = RR / 2, LOA = PL::LO2, RA = PL::R2;
  constexpr int G = (PL::W << (PL::L - R)) / PL::NT;
  static_assert((PL::W << (PL::L - R)) % PL::NT == 0, "groups must divide evenly over the threads");
  constexpr double FSTEP = (double)(1u << (PL::R1 + PL::R2));  // f step per m
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    group_fence();
    const uint32_t g = (uint32_t)(gi * PL::NT + tid);
    const uint32_t c = g % PL::W, u = g / PL::W;  // u = k1 << R2 | k2
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = (double)lds[lds_at<PL>(row_of<0, R>(u, m), c)];
    dft4<R, true>(x, mk);
    const double f0 = (double)((u >> PL::R2) + ((u & ((1u << PL::R2) - 1u)) << PL::R1));
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = reduce(x[m] * (f0 + m * FSTEP));
    double y[H];
#pragma unroll
    for (int q = 0; q < H; q++) y[q] = x[q] + x[q + H];
    dft4<R - 1, false>(y, mk);
#pragma unroll
    for (int q = 0; q < H; q++) {
      const uint32_t j = row_of<0, R>(u, 2u * (uint32_t)q);
      const uint32_t k = (j >> LOA) & ((1u << RA) - 1u);
      lds[lds_at<PL>(j, c)] = slice::mul_mem(y[q], dec_tw(ta, ((j & ((1u << LOA) - 1u)) << RA) | k), mk);
    }
  }
}

// NTT pass on the second digit: forward DFT, forward twiddle of the top
// digit; only the groups of even rows (the row's low bit is the group's).
template <class PL>
ZHD void dec_ntt2(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ ta, const slice::MulK& mk) {
  constexpr int R = PL::R2, RR = 1 << R, LO = PL::LO2, LOA = PL::LO1, RA = PL::R1;
  constexpr int NG = (PL::W << (PL::L - R)) >> 1;
  static_assert(NG % PL::NT == 0 || PL::NT % NG == 0, "groups must divide evenly over the threads");
  constexpr int G = NG >= PL::NT ? NG / PL::NT : 1;
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    group_fence();
    const uint32_t g = (uint32_t)(gi * PL::NT + tid);
    if (NG < PL::NT && g >= (uint32_t)NG) break;  // fewer groups than threads
    const uint32_t c = g % PL::W, u = (g / PL::W) << 1;
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = (double)lds[lds_at<PL>(row_of<LO, R>(u, m), c)];
    dft4<R, false>(x, mk);
#pragma unroll
    for (int m = 0; m < RR; m++) {
      const uint32_t j = row_of<LO, R>(u, m);
      const uint32_t kd = (j >> LOA) & ((1u << RA) - 1u), lw = j & ((1u << LOA) - 1u);
      lds[lds_at<PL>(j, c)] = slice::mul_mem(top_first<PL>(x[m], ta, kd, lw, mk), top_last<PL>(ta, kd, lw), mk);
    }
  }
}

// Last NTT pass (top digit), natural order out, only the groups of even rows:
// every value times the row's factor srow[r] (1 / (n w^r Z'(w^r)) on erased
// rows, 0 on known ones, canonical u32), reduced to canonical u32.
template <class PL>
ZHD void dec_ntt1(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ srow, const slice::MulK& mk) {
  constexpr int R = PL::R1, RR = 1 << R, LO = PL::LO1;
  constexpr int NG = (PL::W << (PL::L - R)) >> 1;
  static_assert(NG % PL::NT == 0 || PL::NT % NG == 0, "groups must divide evenly over the threads");
  constexpr int G = NG >= PL::NT ? NG / PL::NT : 1;
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    group_fence();
    const uint32_t g = (uint32_t)(gi * PL::NT + tid);
    if (NG < PL::NT && g >= (uint32_t)NG) break;  // fewer groups than threads
    const uint32_t c = g % PL::W, u = (g / PL::W) << 1;
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = (double)lds[lds_at<PL>(row_of<LO, R>(u, m), c)];
    dft4<R, false>(x, mk);
#pragma unroll
    for (int m = 0; m < RR; m++) {
      const uint32_t row = row_of<LO, R>(u, m);
      const slice::Tw2 t = dec_tw(srow, row);
      lds[lds_at<PL>(row, c)] = (int32_t)canonical(slice::mul_reg(x[m], t.w, t.wq, mk));
    }
  }
}

// 16-byte global accesses (the store with the wait states of GBuf::store4).
ZHD U4 dec_ld16(const uint32_t* p) { return *reinterpret_cast<const U4*>(p); }
ZHD void dec_st16(uint32_t* p, const U4& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  store16(p, v);
#else
  *reinterpret_cast<U4*>(p) = v;
#endif
}

// Tile `tile` (W columns) into LDS: the k packed known rows of each column
// (scratch entry c * k + q is rank q, at row rows[q]; the known rows in
// increasing order) and zeros at the erased rows (rows[k + i]), four
// ranks at a time: a 16-byte load of the values and one of their row
// numbers (`rows` 16-byte aligned, the same for every tile), scattered into
// LDS.  Columns past `cols` get zeros.
template <class PL>
ZHD void dec_load(int tid, uint32_t tile, uint32_t cols, const uint32_t* __restrict__ T,
                  const uint32_t* __restrict__ rows, int32_t* __restrict__ lds) {
  constexpr uint32_t K = PL::S / 2, KQ = K / 4;
  static_assert(K % 4 == 0, "ranks come four at a time");
  const uint32_t* src = T + (size_t)tile * PL::W * K;
  const uint32_t wv = cols - tile * PL::W < (uint32_t)PL::W ? cols - tile * PL::W : (uint32_t)PL::W;
  for (uint32_t i = (uint32_t)tid; i < (uint32_t)PL::W * KQ; i += PL::NT) {
    const uint32_t c = i / KQ, q = (i % KQ) * 4;
    const U4 r = dec_ld16(rows + q);
    U4 v{0, 0, 0, 0};
    if (c < wv) v = dec_ld16(src + (size_t)c * K + q);
    lds[lds_at<PL>(r.x, c)] = (int32_t)v.x;
    lds[lds_at<PL>(r.y, c)] = (int32_t)v.y;
    lds[lds_at<PL>(r.z, c)] = (int32_t)v.z;
    lds[lds_at<PL>(r.w, c)] = (int32_t)v.w;
  }
  for (uint32_t i = (uint32_t)tid; i < (uint32_t)PL::W * KQ; i += PL::NT) {
    const uint32_t c = i / KQ, q = K + (i % KQ) * 4;
    const U4 r = dec_ld16(rows + q);
    lds[lds_at<PL>(r.x, c)] = 0;
    lds[lds_at<PL>(r.y, c)] = 0;
    lds[lds_at<PL>(r.z, c)] = 0;
    lds[lds_at<PL>(r.w, c)] = 0;
  }
}

// The data rows (the even rows) of tile `tile` from LDS into the k entries of
// each scratch column, four rows to a 16-byte store.
template <class PL>
ZHD void dec_store(int tid, uint32_t tile, uint32_t cols, uint32_t* __restrict__ T, const int32_t* __restrict__ lds) {
  constexpr uint32_t K = PL::S / 2, KQ = K / 4;
  uint32_t* dst = T + (size_t)tile * PL::W * K;
  const uint32_t wv = cols - tile * PL::W < (uint32_t)PL::W ? cols - tile * PL::W : (uint32_t)PL::W;
  for (uint32_t i = (uint32_t)tid; i < wv * KQ; i += PL::NT) {
    const uint32_t c = i / KQ, j = (i % KQ) * 4;
    const U4 o{(uint32_t)lds[lds_at<PL>(2 * j, c)], (uint32_t)lds[lds_at<PL>(2 * j + 2, c)],
               (uint32_t)lds[lds_at<PL>(2 * j + 4, c)], (uint32_t)lds[lds_at<PL>(2 * j + 6, c)]};
    dec_st16(dst + (size_t)c * K + j, o);
  }
}

// The packed LDS tables behind the tile.
template <class PL>
ZHD void dec_load_tables(int tid, int32_t* __restrict__ lds, const DecodeTabs& T) {
  uint32_t* dst = reinterpret_cast<uint32_t*>(lds + PL::TILE_WORDS);
  for (uint32_t i = (uint32_t)tid; i < T.packed_words; i += PL::NT) dst[i] = T.packed[i];
}

// Step STEP (0 .. 4) of the transform pair on the tile in LDS; steps are
// separated by a work-group barrier on the GPU.
template <class PL, int STEP>
ZHD void dec_step(int tid, int32_t* __restrict__ lds, const DecodeTabs& T, const uint32_t* __restrict__ srow,
                  const slice::MulK& mk) {
  const uint32_t* sb = reinterpret_cast<const uint32_t*>(lds + PL::TILE_WORDS);
  if constexpr (STEP == 0)
    dec_intt<PL, PL::LO1, PL::R1>(tid, lds, sb + T.o_inv1, mk);
  else if constexpr (STEP == 1)
    dec_intt<PL, PL::LO2, PL::R2>(tid, lds, sb + T.o_inv2, mk);
  else if constexpr (STEP == 2)
    dec_mid<PL>(tid, lds, sb + T.o_fwd2, mk);
  else if constexpr (STEP == 3)
    dec_ntt2<PL>(tid, lds, sb + T.o_fwd1, mk);
  else
    dec_ntt1<PL>(tid, lds, srow, mk);
}

// ---- host construction of the tables ------------------------------------------

// Host image of the decoder's LDS tables for n = 2^L (canonical u32).
struct DecodeTableImage {
  std::vector<uint32_t> packed;
  uint32_t o_inv1 = 0, o_fwd1 = 0, o_fwd2 = 0, o_inv2 = 0;
};

// Forward twiddles of the digit of r bits at bit lo in [low << r | k]
// order: w^(low k), w the root of order 2^(lo+r).  The middle pass reads
// them with k running over the lanes (k << lo | low would put its 16 lanes'
// words 16 apart: eight-way bank conflicts).
inline void push_forward_low_major(std::vector<Tw>& v, int lo, int r) {
  const uint32_t w = root_of_unity((uint32_t)(lo + r));
  const size_t base = v.size();
  v.resize(base + ((size_t)1 << (lo + r)));
  for (uint32_t k = 0; k < (1u << r); k++) {
    const uint32_t step = pow_mod_host(w, k);
    uint32_t cur = 1;
    for (uint32_t low = 0; low < (1u << lo); low++) {
      v[base + (((size_t)low << r) | k)] = make_tw(cur);
      cur = mul_mod_host(cur, step);
    }
  }
}

// Inverse twiddles of the digit of r bits at bit lo in [k << lo | low]
// order: w^-(low k), w the root of order 2^(lo+r).
inline void push_inverse_digit(std::vector<Tw>& v, int lo, int r) {
  const uint32_t w = inv_mod_host(root_of_unity((uint32_t)(lo + r)));
  for (uint32_t k = 0; k < (1u << r); k++) {
    const uint32_t step = pow_mod_host(w, k);
    uint32_t cur = 1;
    for (uint32_t low = 0; low < (1u << lo); low++) {
      v.push_back(make_tw(cur));
      cur = mul_mod_host(cur, step);
    }
  }
}

// The top digit's twiddles w^(+-low k) as DecodePlan lays them out: one
// [k << lo | low] table, or (factored) the two tables of the low FA bits and
// the high bits of `low`.
inline void push_top_digit(std::vector<Tw>& v, int lo, int r, bool inverse, bool factored, int fa) {
  uint32_t w = root_of_unity((uint32_t)(lo + r));
  if (inverse) w = inv_mod_host(w);
  if (!factored) {
    for (uint32_t k = 0; k < (1u << r); k++)
      for (uint32_t low = 0; low < (1u << lo); low++) v.push_back(make_tw(pow_mod_host(w, (uint64_t)low * k)));
    return;
  }
  for (uint32_t k = 0; k < (1u << r); k++)
    for (uint32_t ll = 0; ll < (1u << fa); ll++) v.push_back(make_tw(pow_mod_host(w, (uint64_t)ll * k)));
  for (uint32_t k = 0; k < (1u << r); k++)
    for (uint32_t lh = 0; lh < (1u << (lo - fa)); lh++)
      v.push_back(make_tw(pow_mod_host(w, ((uint64_t)lh << fa) * k)));
}

template <class PL>
DecodeTableImage build_decode_tables() {
  DecodeTableImage im;
  constexpr int r1 = PL::R1, r2 = PL::R2, rd = PL::RD;
  const int lo1 = rd + r2, lo2 = rd;
  std::vector<Tw> t;
  push_top_digit(t, lo1, r1, true, PL::FT, PL::FA);
  im.o_inv1 = 0;
  im.o_fwd1 = (uint32_t)t.size();
  push_top_digit(t, lo1, r1, false, PL::FT, PL::FA);
  im.o_fwd2 = (uint32_t)t.size();
  push_forward_low_major(t, lo2, r2);
  im.o_inv2 = (uint32_t)t.size();
  push_inverse_digit(t, lo2, r2);
  for (const Tw& x : t) im.packed.push_back((uint32_t)x.w);
  if (im.packed.size() != (size_t)PL::TW_WORDS) im.packed.clear();  // a layout mismatch: no tables
  return im;
}

// ---- plans ----------------------------------------------------------------------

// Decoder plans: n = 2^L for 8 <= L <= 14, W columns so that every pass has
// at least one group per thread; up to n = 2^12 two 512-thread work-groups
// per CU, then the work-groups LDS allows (three at 2^13, one at 2^14), each
// compiled for at most that many waves per SIMD.
//                      L   W   NT  R1 R2 RD  MINW
using DecPlan8 = DecodePlan<8, 16, 512, 3, 3, 2>;
using DecPlan9 = DecodePlan<9, 8, 512, 3, 3, 3>;
using DecPlan10 = DecodePlan<10, 8, 512, 4, 3, 3>;
using DecPlan11 = DecodePlan<11, 4, 512, 4, 4, 3>;
using DecPlan12 = DecodePlan<12, 2, 512, 4, 4, 4>;
using DecPlan13 = DecodePlan<13, 1, 256, 5, 4, 4, 2>;
using DecPlan14 = DecodePlan<14, 1, 512, 5, 5, 4, 2>;
constexpr int DEC_FP64_MIN_LOG = 8, DEC_FP64_MAX_LOG = 14;

// DecodeTabs over a copy `packed` of the image.
inline DecodeTabs make_decode_tabs(const DecodeTableImage& im, const uint32_t* packed) {
  DecodeTabs T;
  T.packed = packed;
  T.o_inv1 = im.o_inv1;
  T.o_fwd1 = im.o_fwd1;
  T.o_fwd2 = im.o_fwd2;
  T.o_inv2 = im.o_inv2;
  T.packed_words = (uint32_t)im.packed.size();
  return T;
}

template <class F>
int with_decode_plan(int L, F&& f) {
  switch (L) {
    case 8: return f(DecPlan8{});
    case 9: return f(DecPlan9{});
    case 10: return f(DecPlan10{});
    case 11: return f(DecPlan11{});
    case 12: return f(DecPlan12{});
    case 13: return f(DecPlan13{});
    case 14: return f(DecPlan14{});
    default: return -1;
  }
}

}  // namespace zoda
