// Batched poly-major NTT / INTT of size n = 2^L (2 <= L <= 12) on FP64 digit
// passes, shared by the gfx950 kernel (ntt_poly.hip) and the CPU emulation of
// the tests.  The same transform as the reference's cuda_ntt / cuda_intt with
// the standard roots (cuda/ntt_kernel.cu:106-208; src/ntt/cpu.rs:1-60 for the
// arithmetic it must equal): poly c is the n values at values[c * stride],
//
//     NTT:  X[k] = sum_j x[j] w^(j k),   INTT: x[j] = n^-1 sum_k X[k] w^-(j k),
//
// w = 31^((p - 1) / n), natural order in and out; values n..stride-1 untouched.
//
// A work-group takes tiles of TILE = 4096 values (4096 / n polys) into LDS and
// runs P = ceil(L / 4) passes over digits of the index, top digit first: the
// digit i (R_i bits at bit LO_i) is transformed by a radix-2^R_i DFT in
// registers with constant roots, and every pass but the last multiplies its
// output k by w_(2^(LO_i + R_i))^(low * k), low the index bits below LO_i.
// (Decimation by the top digit: the first pass's output digit becomes the
// bottom digit of the frequency index.)  The frequency of the value in slot
// s = sum k_i 2^LO_i is k = sum k_i 2^(R_0 + ... + R_(i-1)), so the store phase
// reads the slots in that order: there is no permutation pass.  The INTT's
// n^-1 is folded into the first pass's twiddles (or, with one pass, multiplied
// in before the canonical reduction).
//
// Arithmetic as in the encode (slice_core.hpp): integer-valued doubles, each
// twiddle multiplication ending in an int32 below 2^30 that goes back to LDS,
// the last pass reduced to canonical u32.  Twiddles are u32 in LDS, read with
// their quotient w/p formed by one multiply.
//
// LDS words are padded by one every 32 (slot s at s + s/32), so that the last
// pass -- whose groups are 2^R consecutive slots, 16 words apart from lane to
// lane -- reads from 32 different banks.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "slice_core.hpp"

namespace zoda {
namespace nttf {

constexpr int TILE = 4096;                 // values per tile
constexpr int NT = 256;                    // threads per work-group
constexpr int LDS_DATA = TILE + TILE / 32; // padded data words

ZHD constexpr int phys(int s) { return s + (s >> 5); }

// n^-1 as a double (the single-pass inverse multiplies it in)
template <int N>
inline constexpr double kNinvW = (double)inv_mod_host((uint32_t)N);

template <int L_>
struct Plan {
  static constexpr int L = L_, N = 1 << L;
  static_assert(L >= 2 && L <= 12, "FP64 poly plans cover 2^2..2^12");
  static constexpr int P = (L + 3) / 4;  // passes
  static constexpr int POLYS = TILE / N; // polys per tile
  // digit bits, top digit first, as even as possible
  static constexpr int R(int i) { return L / P + (i < L % P ? 1 : 0); }
  static constexpr int LO(int i) {
    int lo = 0;
    for (int j = i + 1; j < P; j++) lo += R(j);
    return lo;
  }
  // bit of the frequency index where digit i's output lands
  static constexpr int KB(int i) {
    int b = 0;
    for (int j = 0; j < i; j++) b += R(j);
    return b;
  }
  // twiddle table of pass i < P - 1: [k][low], 2^(LO + R) entries
  static constexpr int TW_OFF(int i) {
    int o = 0;
    for (int j = 0; j < i; j++) o += 1 << (LO(j) + R(j));
    return o;
  }
  static constexpr int TW_WORDS = P > 1 ? TW_OFF(P - 1) : 0;
  static constexpr int LDS_WORDS = LDS_DATA + TW_WORDS;
  // slot of frequency k (within a poly)
  static ZHD int slot_of(int k) {
    int s = 0;
#pragma unroll
    for (int i = 0; i < P; i++) s |= ((k >> KB(i)) & ((1 << R(i)) - 1)) << LO(i);
    return s;
  }
};

// Twiddle tables of size 2^L (host): pass i's [k][low] = w_(2^(LO+R))^(+-low k),
// times n^-1 in the first pass of the inverse.  u32 canonical.
template <class PL>
inline void build_ntt_fp64_tables(bool inverse, uint32_t* tw) {
  const uint32_t ninv = inv_mod_host((uint32_t)PL::N);
  for (int i = 0; i + 1 < PL::P; i++) {
    const int lb = PL::LO(i) + PL::R(i);
    uint32_t w = root_of_unity(lb);
    if (inverse) w = inv_mod_host(w);
    const int lows = 1 << PL::LO(i), ks = 1 << PL::R(i);
    for (int k = 0; k < ks; k++)
      for (int low = 0; low < lows; low++) {
        uint32_t v = pow_mod_host(w, (uint64_t)low * k);
        if (inverse && i == 0) v = mul_mod_host(v, ninv);
        tw[PL::TW_OFF(i) + k * lows + low] = v;
      }
  }
}

// A tile's polys: poly p of the tile is poly `poly0 + p` of the batch; the
// first `live` are real (the rest of the tile is zeros and is not stored).
// The load and store callables address a poly's values (the kernel: raw-buffer
// accesses; the emulation: host memory).
struct TileSpan {
  uint32_t live;    // real polys in this tile
  size_t poly0;     // index of the tile's first poly
};

// Phase 1: the tile's values into LDS (16-byte loads, four values a lane).
// Chunk q of thread tid is the four values at 4 (tid + q NT) of the tile.
template <class PL, class LoadF>
ZHD U4 tile_chunk(int tid, int q, const TileSpan& ts, const LoadF& load4) {
  const int v = 4 * (tid + q * NT), p = v >> PL::L, j = v & (PL::N - 1);
  U4 x{0, 0, 0, 0};
  if ((uint32_t)p < ts.live) x = load4(ts.poly0 + (size_t)p, (uint32_t)j);
  return x;
}

// u32 offset of tile value v from the tile's first poly: poly v >> L, index
// v mod n.  For v = 4 tid + 4 NT q it splits into a per-thread part
// (chunk_offset(4 tid)) and a per-q part (chunk_offset(4 NT q)): 4 tid < 4 NT
// = 1024, so either n <= 1024 and 4 NT q is whole polys, or n > 1024 and
// neither part leaves its poly.
template <class PL>
ZHD uint32_t chunk_offset(uint32_t v, uint32_t stride) {
  return (v >> PL::L) * stride + (v & (PL::N - 1));
}

ZHD void put_chunk(int tid, int q, int32_t* __restrict__ lds, const U4& x) {
  int32_t* d = lds + phys(4 * (tid + q * NT));
  d[0] = (int32_t)x.x;
  d[1] = (int32_t)x.y;
  d[2] = (int32_t)x.z;
  d[3] = (int32_t)x.w;
}

template <class PL, class LoadF>
ZHD void load_tile(int tid, const TileSpan& ts, int32_t* __restrict__ lds, const LoadF& load4) {
  constexpr int CHUNKS = TILE / 4 / NT;
#pragma unroll
  for (int q = 0; q < CHUNKS; q++) put_chunk(tid, q, lds, tile_chunk<PL>(tid, q, ts, load4));
}

// Pass I: the radix-2^R DFT over digit I of every group, then the twiddle
// (or, in the last pass, the canonical reduction).
template <class PL, int I, bool INV>
ZHD void pass(int tid, int32_t* __restrict__ lds, const uint32_t* __restrict__ tw, const slice::MulK& k) {
  constexpr int R = PL::R(I), RR = 1 << R, LO = PL::LO(I);
  constexpr int GROUPS = TILE / RR, G = GROUPS / NT;
  static_assert(GROUPS % NT == 0, "groups divide evenly over the threads");
  constexpr bool FIRST = I == 0, LAST = I + 1 == PL::P;
#pragma unroll
  for (int q = 0; q < G; q++) {
    const int g = tid + q * NT;
    const int low = g & ((1 << LO) - 1), rest = g >> LO;
    const int base = low | (rest << (LO + R));  // slot of m = 0 (poly bits above)
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) {
      const int32_t v = lds[phys(base + (m << LO))];
      x[m] = FIRST ? (double)(uint32_t)v : (double)v;
    }
    slice::dft<R, INV>(x, k);
    if constexpr (!LAST) {
      const uint32_t* t = tw + PL::TW_OFF(I) + low;
      int32_t r[RR];
      slice::mul_mem_batch(x, [&](int kk) -> slice::Tw2 {
        const double w = (double)t[kk << LO];
        return slice::Tw2{w, w * PINV};
      }, r, k);
#pragma unroll
      for (int kk = 0; kk < RR; kk++) lds[phys(base + (kk << LO))] = r[kk];
    } else {
#pragma unroll
      for (int kk = 0; kk < RR; kk++) {
        double y = x[kk];
        if constexpr (INV && PL::P == 1) y = slice::mul_reg(y, kNinvW<PL::N>, kNinvW<PL::N> * PINV, k);
        lds[phys(base + (kk << LO))] = (int32_t)slice::canon(y, k);
      }
    }
  }
}

// Phase 3: natural-order values out of their slots (16-byte stores).
template <class PL, class StoreF>
ZHD void store_tile(int tid, const TileSpan& ts, const int32_t* __restrict__ lds, const StoreF& store4) {
  constexpr int CHUNKS = TILE / 4 / NT;
#pragma unroll
  for (int q = 0; q < CHUNKS; q++) {
    const int c = tid + q * NT;
    const int v = 4 * c, p = v >> PL::L, kq = v & (PL::N - 1);
    if ((uint32_t)p >= ts.live) continue;
    const int pb = p << PL::L;
    U4 x;
    x.x = (uint32_t)lds[phys(pb + PL::slot_of(kq))];
    x.y = (uint32_t)lds[phys(pb + PL::slot_of(kq + 1))];
    x.z = (uint32_t)lds[phys(pb + PL::slot_of(kq + 2))];
    x.w = (uint32_t)lds[phys(pb + PL::slot_of(kq + 3))];
    store4(ts.poly0 + (size_t)p, (uint32_t)kq, x);
  }
}

}  // namespace nttf
}  // namespace zoda
