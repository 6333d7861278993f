// Per-thread work of the fused ZODA column encode, shared by the gfx950 kernel
// (encode.hip) and a CPU emulation used by the tests to check the index
// arithmetic without a GPU.
//
// The reference encode (src/benchmarks/utils.rs:100-146) transforms every
// column c of the k x cols input twice: INTT at size s = n_k, zero-pad to
// 2s, NTT at size 2s.  With a = INTT_s(column), the evaluation at the even
// points w_2s^(2j) = w_s^j gives back the data, and the odd points give
// NTT_s(a_i * w_2s^i).  So, per column,
//
//     out[2j]     = data[j]
//     out[2j + 1] = NTT_s( t (.) INTT'_s(data) )[j],  t_i = s^-1 * w_2s^i
//
// (INTT' without its 1/s, folded into t).  Both size-s transforms run as a
// four-step decomposition of the row index into digits of R1 (top), R2 and RD
// (bottom) bits:
//
//   INTT pass on a digit: DFT over the digit (radix 2^R, in registers, with
//                         constant roots), then the inter-digit twiddle
//                         w^-(low bits * output digit).  Output digits land in
//                         digit-reversed order.
//   middle pass:          inverse DFT over the bottom digit, the twist t, the
//                         forward DFT over the same digit (the transpose of
//                         the INTT's decomposition consumes the digit-reversed
//                         order directly, so no permutation pass exists), and
//                         the forward twiddle of the digit above.
//   NTT pass on a digit:  DFT over the digit, then the twiddle of the digit
//                         above; the top digit ends in natural order and is
//                         reduced to canonical u32 and stored as odd rows.
//
// A work-group owns a tile of W adjacent columns (adjacent lanes read adjacent
// columns, so every row segment is one coalesced W*4-byte access).  The first
// pass reads the tile straight from the row-major input and writes the even
// output rows from the same registers; passes exchange data through LDS as
// int32 (every value leaving a pass is reduced below 2^30); the last pass
// writes the odd rows.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "field.hpp"
#include "fp64.hpp"

namespace zoda {

// Device (or host, for the emulation) pointers into one table buffer.
struct EncodeTabs {
  const Tw* inv1;   // INTT twiddles of the top digit, [low << R1 | k]
  const Tw* inv2;   // INTT twiddles of the second digit, [low << R2 | k]
  const Tw* fwd1;   // NTT twiddles of the top digit, [k << LO1 | low]
  const Tw* fwd2;   // NTT twiddles of the second digit, [k << LO2 | low]
  const Tw* twist;  // s entries in storage order: s^-1 * w_2s^(digitrev(i))
  const Tw* roots;  // 64 entries: w_32^e for e < 32, then w_32^-e
};

template <int L_, int W_, int NT_, int R1_, int R2_, int RD_>
struct EncodePlan {
  static constexpr int L = L_, W = W_, NT = NT_, R1 = R1_, R2 = R2_, RD = RD_;
  static_assert(R1 + R2 + RD == L, "digits must cover the row index");
  static_assert(R2 == 0 || R1 > 0, "digits fill from the top");
  static_assert(RD <= 5 && R1 <= 5 && R2 <= 5, "register DFTs are at most radix 32");
  static constexpr int S = 1 << L;
  static constexpr int D = 1 + (R1 > 0) + (R2 > 0);
  static constexpr int LO2 = RD;
  static constexpr int LO1 = RD + R2;
  static constexpr int LDS_WORDS = D > 1 ? W * S : 0;
  static constexpr int STEPS = 2 * D - 1;
};

ZHD constexpr int brev_bits(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

// Natural-order DFT of 2^R integer-valued doubles: X[k] = sum_m x[m] w^(m k)
// with w = rt[32 >> R] (a primitive 2^R-th root; rt[e] = w_32^(+-e)).
// Radix-2 decimation in frequency; the bit-reversed result is renamed back to
// natural order at no cost.  |x| < B on entry gives |X| < 2^R * B.
template <int R>
ZHD void dft_reg(double (&x)[1 << R], const Tw* __restrict__ rt) {
  constexpr int N = 1 << R;
#pragma unroll
  for (int hs = R - 1; hs >= 0; hs--) {
    const int half = 1 << hs;
#pragma unroll
    for (int b = 0; b < N; b += 2 * half) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        double u = x[b + j], v = x[b + j + half];
        x[b + j] = u + v;
        double d = u - v;
        if (j == 0) {
          x[b + j + half] = d;
        } else {
          const Tw t = rt[j << (4 - hs)];  // w_(2 half)^j = w_32^(j * 32 / (2 half))
          x[b + j + half] = mulmod(d, t.w, t.wq);
        }
      }
    }
  }
  double y[N];
#pragma unroll
  for (int k = 0; k < N; k++) y[k] = x[brev_bits(k, R)];
#pragma unroll
  for (int k = 0; k < N; k++) x[k] = y[k];
}

enum PassKind { PASS_INTT = 0, PASS_MID = 1, PASS_NTT = 2 };

// Row index of digit value m in group u for a digit of R bits at bit LO.
template <int LO, int R>
ZHD uint32_t row_of(uint32_t u, uint32_t m) {
  return (u & ((1u << LO) - 1u)) | (m << LO) | ((u >> LO) << (LO + R));
}

ZHD size_t out_index(uint32_t row, uint32_t cidx, uint32_t cols, uint32_t out_rows, int layout) {
  return layout == 0 ? (size_t)row * cols + cidx : (size_t)cidx * out_rows + row;
}

// One pass for thread `tid` of the work-group encoding tile `tile`.
//   LO, R    the digit this pass transforms
//   KIND     PASS_INTT / PASS_MID / PASS_NTT
//   FIRST    read the tile from global memory (and write the even rows)
//   LAST     write canonical odd rows to global memory
//   LOA, RA  digit above, whose forward twiddle ends MID/NTT passes (RA = 0: none)
template <class PL, int LO, int R, int KIND, bool FIRST, bool LAST, int LOA, int RA, bool FULL>
ZHD void encode_pass(int tid, uint32_t tile, const uint32_t* __restrict__ in,
                     uint32_t* __restrict__ out, uint32_t krows, uint32_t cols, int layout,
                     int32_t* __restrict__ lds, const EncodeTabs& T, const Tw* __restrict__ own,
                     const Tw* __restrict__ above) {
  constexpr int RR = 1 << R;
  constexpr int NGROUPS = PL::W << (PL::L - R);
  static_assert(NGROUPS % PL::NT == 0, "groups must divide evenly over the threads");
  constexpr int G = NGROUPS / PL::NT;
  constexpr uint32_t OUT_ROWS = 2u * PL::S;
  const Tw* rinv = T.roots + 32;
  const Tw* rfwd = T.roots;

  // The first pass loads up to 32 values per thread before computing, in
  // chunks of PB groups, to keep loads in flight without a large register set.
  constexpr int PB = FIRST ? (G * RR <= 32 ? G : (RR >= 32 ? 1 : 32 / RR)) : G;
  static_assert(G % PB == 0, "preload chunks must divide the groups");
#pragma unroll
  for (int gc = 0; gc < G; gc += PB) {
    uint32_t raw[FIRST ? PB : 1][FIRST ? RR : 1];
    if constexpr (FIRST) {
#pragma unroll
      for (int gi = 0; gi < PB; gi++) {
        const uint32_t g = (uint32_t)((gc + gi) * PL::NT + tid);
        const uint32_t c = g % PL::W, u = g / PL::W;
        const uint32_t cidx = tile * PL::W + c;
#pragma unroll
        for (int m = 0; m < RR; m++) {
          const uint32_t j = row_of<LO, R>(u, m);
          raw[gi][m] = (FULL || (cidx < cols && j < krows)) ? in[(size_t)j * cols + cidx] : 0u;
        }
      }
#pragma unroll
      for (int gi = 0; gi < PB; gi++) {
        const uint32_t g = (uint32_t)((gc + gi) * PL::NT + tid);
        const uint32_t c = g % PL::W, u = g / PL::W;
        const uint32_t cidx = tile * PL::W + c;
        if (FULL || cidx < cols) {
#pragma unroll
          for (int m = 0; m < RR; m++) {
            const uint32_t j = row_of<LO, R>(u, m);
            out[out_index(2 * j, cidx, cols, OUT_ROWS, layout)] = raw[gi][m];
          }
        }
      }
    }
#pragma unroll
  for (int gi = 0; gi < PB; gi++) {
    const uint32_t g = (uint32_t)((gc + gi) * PL::NT + tid);
    const uint32_t c = g % PL::W, u = g / PL::W;
    double x[RR];
#pragma unroll
    for (int m = 0; m < RR; m++) {
      if constexpr (FIRST) {
        x[m] = (double)raw[gi][m];
      } else {
        x[m] = (double)lds[row_of<LO, R>(u, m) * PL::W + c];
      }
    }
    if constexpr (KIND == PASS_INTT) {
      dft_reg<R>(x, rinv);
      const Tw* tw = own + ((u & ((1u << LO) - 1u)) << R);
      x[0] = reduce(x[0]);  // twiddle w^0 = 1
#pragma unroll
      for (int k = 1; k < RR; k++) x[k] = mulmod(x[k], tw[k]);
    } else {
      if constexpr (KIND == PASS_MID) {
        dft_reg<R>(x, rinv);
#pragma unroll
        for (int m = 0; m < RR; m++) x[m] = mulmod(x[m], T.twist[row_of<LO, R>(u, m)]);
      }
      dft_reg<R>(x, rfwd);
      if constexpr (!LAST) {
        static_assert(RA > 0, "a non-final pass applies the twiddle of the digit above");
#pragma unroll
        for (int m = 0; m < RR; m++) {
          const uint32_t j = row_of<LO, R>(u, m);
          const uint32_t k = (j >> LOA) & ((1u << RA) - 1u);
          x[m] = mulmod(x[m], above[(k << LOA) | (j & ((1u << LOA) - 1u))]);
        }
      }
    }
    if constexpr (LAST) {
      const uint32_t cidx = tile * PL::W + c;
      if (FULL || cidx < cols) {
#pragma unroll
        for (int m = 0; m < RR; m++) {
          const uint32_t j = row_of<LO, R>(u, m);
          out[out_index(2 * j + 1, cidx, cols, OUT_ROWS, layout)] = canonical(x[m]);
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < RR; m++) lds[row_of<LO, R>(u, m) * PL::W + c] = (int32_t)x[m];
    }
  }
  }
}

// Step STEP (0 .. 2D-2) of the encode of one tile; steps are separated by a
// work-group barrier on the GPU.  FULL: every column of the tile exists and
// k = 2^L, so no bounds checks are emitted.
template <class PL, int STEP, bool FULL>
ZHD void encode_step(int tid, uint32_t tile, const uint32_t* __restrict__ in,
                     uint32_t* __restrict__ out, uint32_t krows, uint32_t cols, int layout,
                     int32_t* __restrict__ lds, const EncodeTabs& T) {
  constexpr int L1 = PL::LO1, L2 = PL::LO2;
  if constexpr (PL::D == 1) {
    encode_pass<PL, 0, PL::RD, PASS_MID, true, true, 0, 0, FULL>(tid, tile, in, out, krows, cols, layout,
                                                           lds, T, nullptr, nullptr);
  } else if constexpr (PL::D == 2) {
    if constexpr (STEP == 0)
      encode_pass<PL, L1, PL::R1, PASS_INTT, true, false, 0, 0, FULL>(tid, tile, in, out, krows, cols,
                                                                layout, lds, T, T.inv1, nullptr);
    else if constexpr (STEP == 1)
      encode_pass<PL, 0, PL::RD, PASS_MID, false, false, L1, PL::R1, FULL>(
          tid, tile, in, out, krows, cols, layout, lds, T, nullptr, T.fwd1);
    else
      encode_pass<PL, L1, PL::R1, PASS_NTT, false, true, 0, 0, FULL>(tid, tile, in, out, krows, cols,
                                                               layout, lds, T, nullptr, nullptr);
  } else {
    if constexpr (STEP == 0)
      encode_pass<PL, L1, PL::R1, PASS_INTT, true, false, 0, 0, FULL>(tid, tile, in, out, krows, cols,
                                                                layout, lds, T, T.inv1, nullptr);
    else if constexpr (STEP == 1)
      encode_pass<PL, L2, PL::R2, PASS_INTT, false, false, 0, 0, FULL>(tid, tile, in, out, krows, cols,
                                                                 layout, lds, T, T.inv2, nullptr);
    else if constexpr (STEP == 2)
      encode_pass<PL, 0, PL::RD, PASS_MID, false, false, L2, PL::R2, FULL>(
          tid, tile, in, out, krows, cols, layout, lds, T, nullptr, T.fwd2);
    else if constexpr (STEP == 3)
      encode_pass<PL, L2, PL::R2, PASS_NTT, false, false, L1, PL::R1, FULL>(
          tid, tile, in, out, krows, cols, layout, lds, T, nullptr, T.fwd1);
    else
      encode_pass<PL, L1, PL::R1, PASS_NTT, false, true, 0, 0, FULL>(tid, tile, in, out, krows, cols,
                                                               layout, lds, T, nullptr, nullptr);
  }
}

}  // namespace zoda
