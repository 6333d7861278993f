// Column encode in the column-major layout (ZODA_LAYOUT_COL), shared by the
// gfx950 kernel (encode_cols.hip) and the CPU emulation of the tests.
//
// The input arrives transposed (a column-major scratch: column c's s values
// contiguous), so a work-group can take whole columns: tiles of TILE = 4096
// values (4096 / s columns) in padded LDS, as the drop-in transforms do
// (ntt_fp64_core.hpp).  Per column, the encode of encode_core.hpp,
//
//     out[c][2j] = x[j],   out[c][2j + 1] = NTT_s( t (.) INTT'_s(x) )[j],
//     t_f = s^-1 w_2s^f,
//
// runs as 2P - 1 passes over the digits of the index (P = ceil(L / 4), digits
// R_0 (top) .. R_(P-1) of nttf::Plan<L>, digit i at slot bit LO_i):
//
//   INTT passes, digits 0 .. P-2: inverse DFT over the digit, then the
//     twiddle w_(2^(LO+R))^-(low k) -- the last of them also times the part
//     of the twist that is constant over the bottom digit, s^-1 w_2s^f_low
//     with f_low the frequency bits found so far (a table by output slot).
//     Frequency f = sum k_i 2^KB_i sits in slot sum k_i 2^LO_i.
//   middle pass, digit P-1 (slot bit 0): inverse DFT, the rest of the twist
//     w_(2^(R+1))^k (constants), forward DFT, then the NTT's twiddle
//     w_s^(f_rest j0) (f_rest the frequency bits of the other digits; a table
//     by output slot).
//   NTT passes, digits P-2 .. 0: forward DFT over the digit, then (all but
//     the last) the twiddle w_(2^(KB+R))^(f_rest' j) of the frequency digits
//     still to come (a table by slot >> LO); the last reduces to canonical.
//
// Taking the NTT's frequency digits top first from slots laid out bottom
// first leaves its output in natural order: slot j holds the odd value of
// row j.  The store phase interleaves it with the input (kept in registers
// since the load) as 16-byte runs of (x_j, y_j, x_(j+1), y_(j+1)) along the
// column, so every output line is written whole.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "ntt_fp64_core.hpp"

namespace zoda {
namespace ecol {

using nttf::NT;
using nttf::phys;
using nttf::TILE;

template <int L_>
struct Plan {
  using D = nttf::Plan<L_>;  // the digits
  static constexpr int L = L_, S = 1 << L, P = D::P, COLS = TILE / S;
  static_assert(L >= 5 && L <= 12, "column plans cover sides 2^5..2^12");
  static constexpr int R(int i) { return D::R(i); }
  static constexpr int LO(int i) { return D::LO(i); }
  static constexpr int KB(int i) { return D::KB(i); }
  // tables (u32): INTT passes i < P - 2 as [k][low]; the last INTT pass (with
  // the twist) and the middle pass as [k][group]; NTT passes 1 .. P-2 as
  // [j][group >> LO] (a group's index within its column: see group_slot).
  // Lanes run through consecutive groups, so every table read of a wave is
  // consecutive words (no LDS bank conflicts).
  // At side 4096 the first INTT pass's [k][low] table (4096 entries) is
  // factored, w^(low k) = w^(lh 2^FA k) w^(ll k) with low = lh 2^FA + ll, into
  // [k][lh] and [k][ll] (256 entries each), at one more multiplication in
  // that pass: the tables then leave LDS for three work-groups per CU.
  static constexpr bool FT0 = L >= 12 && P > 2;
  static constexpr int FA = LO(0) / 2;
  static constexpr int INTT_WORDS(int i) {
    return (i == 0 && FT0) ? (1 << (R(0) + LO(0) - FA)) + (1 << (R(0) + FA)) : 1 << (LO(i) + R(i));
  }
  static constexpr int INTT_OFF(int i) {
    int o = 0;
    for (int j = 0; j < i; j++) o += INTT_WORDS(j);
    return o;
  }
  static constexpr int TWIST_OFF = INTT_OFF(P - 2);  // last INTT pass, S entries
  static constexpr int MID_OFF = TWIST_OFF + S;      // middle pass, S entries
  static constexpr int NTT_OFF(int d) {             // NTT pass over digit d, 1 <= d <= P-2
    int o = MID_OFF + S;
    for (int j = P - 2; j > d; j--) o += 1 << (L - LO(j));
    return o;
  }
  static constexpr int TW_WORDS = NTT_OFF(0);
  static constexpr int LDS_WORDS = nttf::LDS_DATA + TW_WORDS;
  // Slot within a column of value k of group gl (0 <= gl < 2^(L - R)) of a
  // pass over digit i: the group's bits below the digit stay at the bottom,
  // the rest go above it.
  static constexpr uint32_t group_slot(int i, uint32_t gl, uint32_t k) {
    return (gl & ((1u << LO(i)) - 1)) | ((gl >> LO(i)) << (LO(i) + R(i))) | (k << LO(i));
  }
};

// Frequency index of the INTT output in slot `slot` (digits below digit
// `upto` of the slot not yet counted: only k_0 .. k_upto contribute).
template <class PL>
constexpr uint32_t freq_of_slot(uint32_t slot, int upto) {
  uint32_t f = 0;
  for (int i = 0; i <= upto; i++) f |= ((slot >> PL::LO(i)) & ((1u << PL::R(i)) - 1)) << PL::KB(i);
  return f;
}

// Host-built tables (u32 canonical).
template <class PL>
inline void build_col_tables(uint32_t* tw) {
  const uint32_t s_inv = inv_mod_host((uint32_t)PL::S), w2s = root_of_unity(PL::L + 1);
  // INTT passes i < P - 2: [k][low] = w_(2^(LO+R))^-(low k); factored (FT0):
  // [k][lh] = w^-(lh 2^FA k), then [k][ll] = w^-(ll k)
  for (int i = 0; i + 2 < PL::P; i++) {
    const uint32_t wi = inv_mod_host(root_of_unity(PL::LO(i) + PL::R(i)));
    const int lows = 1 << PL::LO(i), ks = 1 << PL::R(i);
    if (i == 0 && PL::FT0) {
      const int nlh = 1 << (PL::LO(0) - PL::FA), nll = 1 << PL::FA;
      for (int k = 0; k < ks; k++) {
        for (int lh = 0; lh < nlh; lh++)
          tw[PL::INTT_OFF(0) + k * nlh + lh] = pow_mod_host(wi, (uint64_t)lh * nll * k);
        for (int ll = 0; ll < nll; ll++)
          tw[PL::INTT_OFF(0) + ks * nlh + k * nll + ll] = pow_mod_host(wi, (uint64_t)ll * k);
      }
      continue;
    }
    for (int k = 0; k < ks; k++)
      for (int low = 0; low < lows; low++)
        tw[PL::INTT_OFF(i) + k * lows + low] = pow_mod_host(wi, (uint64_t)low * k);
  }
  // the last INTT pass, [k][group]: w^-(low k) s^-1 w_2s^f_low at the slot
  // the group's value k goes to
  {
    constexpr int i = PL::P - 2, RI = PL::R(i);
    const uint32_t wi = inv_mod_host(root_of_unity(PL::LO(i) + RI));
    for (uint32_t gl = 0; gl < (1u << (PL::L - RI)); gl++)
      for (uint32_t k = 0; k < (1u << RI); k++) {
        const uint32_t slot = PL::group_slot(i, gl, k);
        const uint32_t low = slot & ((1u << PL::LO(i)) - 1);
        const uint32_t f_low = freq_of_slot<PL>(slot, i);
        tw[PL::TWIST_OFF + (k << (PL::L - RI)) + gl] =
            mul_mod_host(mul_mod_host(pow_mod_host(wi, (uint64_t)low * k), s_inv), pow_mod_host(w2s, f_low));
      }
  }
  // the middle pass, [j][group]: w_s^(f_rest j), j the slot's bottom digit
  {
    const uint32_t ws = root_of_unity(PL::L);
    constexpr int RB = PL::R(PL::P - 1);
    for (uint32_t gl = 0; gl < (1u << (PL::L - RB)); gl++)
      for (uint32_t j = 0; j < (1u << RB); j++) {
        const uint32_t slot = PL::group_slot(PL::P - 1, gl, j);
        const uint32_t f_rest = freq_of_slot<PL>(slot, PL::P - 2);
        tw[PL::MID_OFF + (j << (PL::L - RB)) + gl] = pow_mod_host(ws, (uint64_t)f_rest * j);
      }
  }
  // NTT passes over digits d = P-2 .. 1, [j][group >> LO_d]:
  // w_(2^(KB_d + R_d))^(f_rest' j), f_rest' the frequency digits above d
  for (int d = PL::P - 2; d >= 1; d--) {
    const uint32_t wd = root_of_unity(PL::KB(d) + PL::R(d));
    const int hi = PL::L - PL::LO(d) - PL::R(d);  // bits of group >> LO_d
    for (uint32_t gh = 0; gh < (1u << hi); gh++)
      for (uint32_t j = 0; j < (1u << PL::R(d)); j++) {
        const uint32_t slot = PL::group_slot(d, gh << PL::LO(d), j);
        const uint32_t f_rest = freq_of_slot<PL>(slot, d - 1);
        tw[PL::NTT_OFF(d) + (j << hi) + gh] = pow_mod_host(wd, (uint64_t)f_rest * j);
      }
  }
}

// The groups of a pass over digit I: thread tid, group q; base slot (digit I
// zero) within the tile and the slot bits below the digit.
template <class PL, int I>
ZHD void group_of(int tid, int q, int& base, int& low) {
  constexpr int R = PL::R(I), LO = PL::LO(I);
  const int g = tid + q * NT;
  low = g & ((1 << LO) - 1);
  const int rest = g >> LO;
  base = low | (rest << (LO + R));
}

// INTT pass over digit I (0 <= I <= P-2).
template <class PL, int I>
ZHD void intt_pass(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ tw, const slice::MulK& k) {
  constexpr int R = PL::R(I), RR = 1 << R, LO = PL::LO(I);
  constexpr int G = TILE / RR / NT;
  constexpr bool TWIST = I == PL::P - 2;
#pragma unroll
  for (int q = 0; q < G; q++) {
    int base, low;
    group_of<PL, I>(tid, q, base, low);
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) {
      const int32_t v = lds[phys(base + (m << LO))];
      x[m] = I == 0 ? (double)(uint32_t)v : (double)v;
    }
    slice::dft<R, true>(x, k);
    int32_t r[RR];
    if constexpr (TWIST) {
      const int gl = (tid + q * NT) & ((1 << (PL::L - R)) - 1);  // group within its column
      slice::mul_mem_batch(x, [&](int kk) -> slice::Tw2 {
        const double w = (double)tw[PL::TWIST_OFF + (kk << (PL::L - R)) + gl];
        return slice::Tw2{w, w * PINV};
      }, r, k);
    } else if constexpr (I == 0 && PL::FT0) {
      constexpr int NLH = 1 << (LO - PL::FA), NLL = 1 << PL::FA;
      const int lh = low >> PL::FA, ll = low & (NLL - 1);
#pragma unroll
      for (int kk = 1; kk < RR; kk++) {
        const double w = (double)tw[PL::INTT_OFF(0) + kk * NLH + lh];
        x[kk] = slice::mul_reg(x[kk], w, w * PINV, k);
      }
      slice::mul_mem_batch(x, [&](int kk) -> slice::Tw2 {
        const double w = (double)tw[PL::INTT_OFF(0) + RR * NLH + kk * NLL + ll];
        return slice::Tw2{w, w * PINV};
      }, r, k);
    } else {
      slice::mul_mem_batch(x, [&](int kk) -> slice::Tw2 {
        const double w = (double)tw[PL::INTT_OFF(I) + (kk << LO) + low];
        return slice::Tw2{w, w * PINV};
      }, r, k);
    }
#pragma unroll
    for (int kk = 0; kk < RR; kk++) lds[phys(base + (kk << LO))] = r[kk];
  }
}

// The middle pass over the bottom digit (slot bit 0).
template <class PL>
ZHD void mid_pass(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ tw, const slice::MulK& k) {
  constexpr int I = PL::P - 1, R = PL::R(I), RR = 1 << R;
  static_assert(PL::LO(I) == 0, "the bottom digit sits at slot bit 0");
  constexpr int G = TILE / RR / NT;
  constexpr int TS = 64 >> R;  // w_(2^(R+1))^k = w_128^(k TS)
#pragma unroll
  for (int q = 0; q < G; q++) {
    int base, low;
    group_of<PL, I>(tid, q, base, low);
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = (double)lds[phys(base + m)];
    slice::dft<R, true>(x, k);
#pragma unroll
    for (int kk = 1; kk < RR; kk++)
      x[kk] = slice::mul_reg(x[kk], slice::kRoots.w[0][kk * TS], slice::kRoots.wq[0][kk * TS], k);
    slice::dft<R, false>(x, k);
    const int gl = (tid + q * NT) & ((1 << (PL::L - R)) - 1);
    int32_t r[RR];
    slice::mul_mem_batch(x, [&](int j) -> slice::Tw2 {
      const double w = (double)tw[PL::MID_OFF + (j << (PL::L - R)) + gl];
      return slice::Tw2{w, w * PINV};
    }, r, k);
#pragma unroll
    for (int j = 0; j < RR; j++) lds[phys(base + j)] = r[j];
  }
}

// NTT pass over digit D (P-2 .. 0): forward DFT, then the twiddle of the
// frequency digits still to come, or (D = 0) the canonical reduction.
template <class PL, int D>
ZHD void ntt_pass(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ tw, const slice::MulK& k) {
  constexpr int R = PL::R(D), RR = 1 << R, LO = PL::LO(D);
  constexpr int G = TILE / RR / NT;
#pragma unroll
  for (int q = 0; q < G; q++) {
    int base, low;
    group_of<PL, D>(tid, q, base, low);
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) x[m] = (double)lds[phys(base + (m << LO))];
    slice::dft<R, false>(x, k);
    if constexpr (D > 0) {
      constexpr int HI = PL::L - LO - R;
      const int gh = ((tid + q * NT) >> LO) & ((1 << HI) - 1);  // group within its column, >> LO
      int32_t r[RR];
      slice::mul_mem_batch(x, [&](int j) -> slice::Tw2 {
        const double w = (double)tw[PL::NTT_OFF(D) + (j << HI) + gh];
        return slice::Tw2{w, w * PINV};
      }, r, k);
#pragma unroll
      for (int j = 0; j < RR; j++) lds[phys(base + (j << LO))] = r[j];
    } else {
#pragma unroll
      for (int j = 0; j < RR; j++) lds[phys(base + (j << LO))] = (int32_t)slice::canon(x[j], k);
    }
  }
}

// Every pass of a tile, in order, with `sync` between them (the work-group
// barrier on the device; nothing in the phase-by-phase emulation, which runs
// each pass for every thread before the next).
template <class PL, class Sync, class Each>
ZHD void all_passes(const Each& each, const Sync& sync) {
  each([](int tid, int32_t* lds, const uint32_t* tw, const slice::MulK& k) { intt_pass<PL, 0>(tid, lds, tw, k); });
  sync();
  if constexpr (PL::P > 2) {
    each([](int tid, int32_t* lds, const uint32_t* tw, const slice::MulK& k) { intt_pass<PL, 1>(tid, lds, tw, k); });
    sync();
  }
  each([](int tid, int32_t* lds, const uint32_t* tw, const slice::MulK& k) { mid_pass<PL>(tid, lds, tw, k); });
  sync();
  if constexpr (PL::P > 2) {
    each([](int tid, int32_t* lds, const uint32_t* tw, const slice::MulK& k) { ntt_pass<PL, 1>(tid, lds, tw, k); });
    sync();
  }
  each([](int tid, int32_t* lds, const uint32_t* tw, const slice::MulK& k) { ntt_pass<PL, 0>(tid, lds, tw, k); });
  sync();
}

// The transpose of the row-major input into the column-major scratch, in
// blocks of 64 rows x 64 columns through LDS (64 x 65 words, so a block
// column is read from 32 different banks).  Block b of a chunk covers rows
// 64 (b % (k / 64)) .. +63 and columns 64 (b / (k / 64)) .. +63 (the blocks in
// flight at once read whole rows of a band of columns and write one
// contiguous stretch of the scratch).  Thread tid's run q (q < 4): in, row
// id / 16 of the block, columns 4 (id % 16) .. +3; out, column id / 16, rows
// 4 (id % 16) .. +3 (id = tid + q NT).
constexpr int TB = 64, TB_PITCH = 65, TB_RUNS = TB * TB / 4 / NT;
static_assert(TB * TB_PITCH <= nttf::LDS_DATA, "a transpose block fits the tile's LDS");

struct TBlock {
  uint32_t r0, c0;  // first row and column of the block within the chunk
};

ZHD TBlock tblock(uint32_t b, uint32_t k) {
  const uint32_t tiles_r = k / TB;
  return TBlock{(b % tiles_r) * TB, (b / tiles_r) * TB};
}

// run q of thread tid: offset of its 16 input bytes from the block's first
// input value (pitch u32 a row), and its LDS words
ZHD size_t tb_in_offset(int tid, int q, uint32_t pitch) {
  const int id = tid + q * NT;
  return (size_t)(id >> 4) * pitch + 4 * (id & 15);
}

ZHD void tb_put(int tid, int q, uint32_t* __restrict__ sh, const U4& v) {
  const int id = tid + q * NT;
  uint32_t* d = sh + (id >> 4) * TB_PITCH + 4 * (id & 15);
  d[0] = v.x;
  d[1] = v.y;
  d[2] = v.z;
  d[3] = v.w;
}

// run q of thread tid on the way out: the four values and their offset from
// the block's first scratch value (k u32 a column)
ZHD U4 tb_get(int tid, int q, const uint32_t* __restrict__ sh) {
  const int id = tid + q * NT;
  const uint32_t* s = sh + 4 * (id & 15) * TB_PITCH + (id >> 4);
  return U4{s[0], s[TB_PITCH], s[2 * TB_PITCH], s[3 * TB_PITCH]};
}

ZHD size_t tb_out_offset(int tid, int q, uint32_t k) {
  const int id = tid + q * NT;
  return (size_t)(id >> 4) * k + 4 * (id & 15);
}

// A thread's share of a tile: the value pairs (rows j, j + 1 of one column)
// at tile index 2 (tid + q NT), q < PAIRS.  They are loaded as 8-byte runs of
// the column-major input, put into LDS, and kept in registers for the store
// phase, which writes each pair beside its odd values.
struct alignas(8) U2 {
  uint32_t x, y;
};

constexpr int PAIRS = TILE / 2 / NT;

template <class PL>
ZHD void pair_at(int tid, int q, int& cl, int& j) {
  const int v = 2 * (tid + q * NT);
  cl = v >> PL::L;
  j = v & (PL::S - 1);
}

// Pair q of thread tid of the tile whose first column is col0 (zeros past
// the live columns).
template <class PL, class Load2>
ZHD U2 tile_pair(int tid, int q, uint32_t live_cols, size_t col0, const Load2& load2) {
  int cl, j;
  pair_at<PL>(tid, q, cl, j);
  U2 x{0, 0};
  if ((uint32_t)cl < live_cols) x = load2(col0 + (size_t)cl, (uint32_t)j);
  return x;
}

ZHD void put_pair(int tid, int q, int32_t* __restrict__ lds, const U2& x) {
  const int v = 2 * (tid + q * NT);
  lds[phys(v)] = (int32_t)x.x;
  lds[phys(v + 1)] = (int32_t)x.y;
}

// Store phase: pair q (held(q) gives its input values) becomes output rows
// 2j .. 2j + 3 of its column, one 16-byte store (x_j, y_j, x_(j+1), y_(j+1)),
// handed over as store4(q, column, row, values).  Within a whole tile, pair
// q's input is value 2 (tid + q NT) of the tile's column-major values and its
// output starts at value 4 (tid + q NT) of the tile's output.
template <class PL, class Held, class Store4>
ZHD void store_pairs(int tid, uint32_t live_cols, size_t col0, const int32_t* __restrict__ lds, const Held& held,
                     const Store4& store4) {
#pragma unroll
  for (int q = 0; q < PAIRS; q++) {
    int cl, j;
    pair_at<PL>(tid, q, cl, j);
    if ((uint32_t)cl < live_cols) {
      const int v = 2 * (tid + q * NT);
      const U2 x = held(q);
      U4 o;
      o.x = x.x;
      o.y = (uint32_t)lds[phys(v)];
      o.z = x.y;
      o.w = (uint32_t)lds[phys(v + 1)];
      store4(q, col0 + (size_t)cl, (uint32_t)(2 * j), o);
    }
  }
}

}  // namespace ecol
}  // namespace zoda
