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

#include <type_traits>

#include "field.hpp"
#include "fp64.hpp"
#include "fp64_ops.hpp"

namespace zoda {

// Device (or host, for the emulation) pointers into one table buffer.  Each
// twiddle table exists twice: as (w, w/p) double pairs, read from global
// memory, and as canonical u32 in one packed block that plans with TWL copy
// into LDS (w/p is then recomputed with one multiply).
struct EncodeTabs {
  // INTT twiddles, [high << (LO+R) | low << R | k] for the digit's row bits
  // `low` below it and `high` above it.  The last INTT pass before the middle
  // one also carries the part of the twist that is constant over a middle
  // group (see build_encode_tables), so its table has one entry per row.
  const Tw* inv1;
  const Tw* inv2;
  const Tw* fwd1;   // NTT twiddles of the top digit, [k << LO1 | low]
  const Tw* fwd2;   // NTT twiddles of the second digit, [k << LO2 | low]
  const Tw* twist;  // one-pass plans only: s entries, s^-1 * w_2s^i
  const Tw* roots;  // 128 entries: w_64^e for e < 64, then w_64^-e
  const uint32_t* packed;  // u32 copies: inv1 | inv2 | fwd1 | fwd2 | twist
  uint32_t o_inv1, o_inv2, o_fwd1, o_fwd2, o_twist, packed_words;
};

// One twiddle table as a pass reads it.
struct TwTab {
  const Tw* g;         // global (w, w/p)
  const uint32_t* s;   // LDS u32 (plans with TWL)
};

// A u32 twiddle as the (w, w/p) pair the multiplication takes.
ZHD Tw tw_u32(uint32_t v) {
  const double w = (double)v;
  return Tw{w, w * PINV};
}

template <bool LDS_TW>
ZHD Tw tw_get(const TwTab& t, uint32_t i) {
  if constexpr (LDS_TW) {
    return tw_u32(t.s[i]);
  } else {
    return t.g[i];
  }
}

template <int L_, int W_, int NT_, int R1_, int R2_, int RD_, bool TWL_ = false,
          bool PERSIST_ = false, bool FENCE_ = false>
struct EncodePlan {
  static constexpr int L = L_, W = W_, NT = NT_, R1 = R1_, R2 = R2_, RD = RD_;
  static constexpr bool TWL = TWL_;          // twiddle tables in LDS
  static constexpr bool PERSIST = PERSIST_;  // persistent kernel, first pass pipelined in registers
  static constexpr bool FENCE = FENCE_;      // one group's values live at a time
  static_assert(R1 + R2 + RD == L, "digits must cover the row index");
  static_assert(R2 == 0 || R1 > 0, "digits fill from the top");
  static_assert(RD <= 5 && R1 <= 5 && R2 <= 5, "register DFTs are at most radix 32");
  static constexpr int S = 1 << L;
  static constexpr int D = 1 + (R1 > 0) + (R2 > 0);
  static constexpr int LO2 = RD;
  static constexpr int LO1 = RD + R2;
  static constexpr int TILE_WORDS = D > 1 ? W * S : 0;
  // u32 words of the packed tables (see build_encode_tables)
  static constexpr int TW_WORDS = D == 1 ? S : D == 2 ? 2 * S : 3 * S + (1 << LO1);
  // Three-digit plans with LDS tables also keep the middle pass's forward
  // twiddles (2^LO1 entries) as (w, w/p) pairs behind the u32 tables, which
  // saves that pass the conversion and the multiply of every twiddle read.
  static constexpr bool MID_PAIRS = TWL && D == 3;
  static constexpr int PAIR_WORDS = MID_PAIRS ? 4 << LO1 : 0;
  static constexpr int LDS_WORDS = TILE_WORDS + (TWL ? TW_WORDS : 0) + PAIR_WORDS;
  static constexpr int STEPS = 2 * D - 1;
};

ZHD constexpr int brev_bits(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

// Natural-order DFT of 2^R integer-valued doubles: X[k] = sum_m x[m] w^(m k)
// with w = rt[64 >> R] (a primitive 2^R-th root; rt[e] = w_64^(+-e)).
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
          const Tw t = rt[j << (5 - hs)];  // w_(2 half)^j = w_64^(j * 64 / (2 half))
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

// The same DFT with the four-instruction products (slice::mul_reg) on the
// compile-time roots w_128^(+-e) (INV selects the inverse roots), unbatched:
// with several waves per SIMD the other waves hide each product's latency,
// and batches (slice::dft) need registers the decoder's kernels do not have.
// (The per-side encode kernels keep dft_reg: with the product's constants
// held in registers through a pass, their plans for sides 256 to 4096 grew
// by up to 60 registers a thread or spilled.)
template <int R, bool INV>
ZHD void dft4(double (&x)[1 << R], const slice::MulK& mk) {
  constexpr int N = 1 << R;
  constexpr int D = INV ? 1 : 0;
#pragma unroll
  for (int hs = R - 1; hs >= 0; hs--) {
    const int half = 1 << hs;
#pragma unroll
    for (int b = 0; b < N; b += 2 * half) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        const double u = x[b + j], v = x[b + j + half];
        x[b + j] = u + v;
        const double d = u - v;
        x[b + j + half] =
            j == 0 ? d : slice::mul_reg(d, slice::kRoots.w[D][j << (6 - hs)], slice::kRoots.wq[D][j << (6 - hs)], mk);
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

// Keeps the compiler from interleaving the work of consecutive groups, which
// bounds a pass's live registers to about one group's values.
ZHD void group_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// Row index of digit value m in group u for a digit of R bits at bit LO.
template <int LO, int R>
ZHD uint32_t row_of(uint32_t u, uint32_t m) {
  return (u & ((1u << LO) - 1u)) | (m << LO) | ((u >> LO) << (LO + R));
}

// Work-groups are handed to the eight XCDs round-robin (block b goes to XCD
// b % 8).  Give each XCD a contiguous run of tiles, in dispatch order, so that
// neighbouring tiles -- which share the 128-byte lines of every row -- run at
// the same time behind the same L2.
ZHD uint32_t xcd_tile(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, i = b >> 3;
  const uint32_t q = nb >> 3, r = nb & 7u;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

ZHD size_t out_index(uint32_t row, uint32_t cidx, uint32_t cols, uint32_t out_rows, int layout) {
  return layout == 0 ? (size_t)row * cols + cidx : (size_t)cidx * out_rows + row;
}

// One output value.  In the row layout, plans whose row segments are whole
// 128-byte lines (W >= 32) store non-temporally (measured against plain
// stores: 6-8 % at sides 128 and 256, 5 % over the batched mixed-side queue);
// narrower tiles write parts of lines that neighbouring tiles complete.
template <class PL>
ZHD void store_out(uint32_t* p, uint32_t v, int layout) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (PL::W >= 32) {
    if (layout == 0) {
      __builtin_nontemporal_store(v, p);
      return;
    }
  }
#endif
  *p = v;
}

// Called with the group index at the start of every group of a pass.
struct NoHook {
  ZHD void operator()(int) const {}
};

// Four adjacent u32 (one 16-byte access).
struct alignas(16) U4 {
  uint32_t x, y, z, w;
};

// Raw-buffer view of a global u32 array.  An access names a per-thread 32-bit
// byte offset and a uniform 32-bit byte offset, and the buffer instruction adds
// both to the base itself, so the vector ALU computes no 64-bit address per
// access (a global load or store costs one 64-bit add per value otherwise).
// The hardware drops any access whose last byte lies at or beyond 2^32 - 1
// bytes from the base (the range check of num_records = 2^32 - 1): callers keep
// every sum of the two offsets at most 2^32 - 8.
struct GBuf {
#if defined(__HIP_DEVICE_COMPILE__)
  __amdgpu_buffer_rsrc_t r;
  // dword 3 = 0x00020000: 32-bit data format, the gfx9 family's raw buffer
  __device__ explicit GBuf(const uint32_t* p)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(p), 0, -1, 0x00020000)) {}
  // live = false: zero records, so the hardware drops every store and every
  // load returns 0 -- memory instructions that need no branch around them
  // (the record count is pinned to a scalar register: a descriptor in vector
  // registers turns every access into a loop over its lanes)
  __device__ GBuf(const uint32_t* p, bool live) : r(make(p, live)) {}
  // `bytes` records: accesses at or past that many bytes from p load 0 and
  // store nothing (bytes uniform over the wave)
  __device__ GBuf(const uint32_t* p, uint32_t bytes) : r(make_n(p, bytes)) {}
  static __device__ __amdgpu_buffer_rsrc_t make_n(const uint32_t* p, uint32_t bytes) {
    int nr = __builtin_amdgcn_readfirstlane((int)bytes);
    asm volatile("" : "+s"(nr));
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(p), 0, nr, 0x00020000);
  }
  static __device__ __amdgpu_buffer_rsrc_t make(const uint32_t* p, bool live) {
    int nr = __builtin_amdgcn_readfirstlane(live ? -1 : 0);
    asm volatile("" : "+s"(nr));
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(p), 0, nr, 0x00020000);
  }
  __device__ uint32_t load(uint32_t voff, uint32_t soff) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  }
  // Stores are non-temporal (cache policy bit nt): no kernel here reads the
  // output again.  Measured against plain stores on one box: 1-2 % faster at
  // side 2048, 0.5-4 % at 1024, 7-16 % at 512.
  __device__ void store(uint32_t v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, 2);
  }
  // 16-byte forms (four adjacent columns per lane)
  using V4 = __attribute__((ext_vector_type(4))) unsigned int;
  __device__ U4 load4(uint32_t voff, uint32_t soff) const {
    return __builtin_bit_cast(U4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
  // A 16-byte store reads the upper half of its data registers after it
  // issues, and on gfx950 the compiler's hazard model (which sees no hazard
  // for a buffer store with a register offset) lets arithmetic overwrite those
  // registers first: the side-4096 kernel, with stores followed within four
  // instructions by FP64 writes of their registers, stored such values in
  // place of data (bytes 8-15 of some lanes, a few hundred values in every
  // 4096^2 encode; none in any check with one wait state after each store).
  // Each store is therefore followed by two wait states and a wait for the
  // export counter (which on the first GCN parts also counted store data not
  // yet read from registers); the asm also keeps the compiler from moving
  // arithmetic in between the stores of a group.
  __device__ void store4(const U4& v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), r, voff, soff, 2);
    asm volatile("s_waitcnt expcnt(0)\n\ts_nop 1" ::: "memory");
  }
  // the same with the default cache policy, for values read again soon
  __device__ void store4_cached(const U4& v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), r, voff, soff, 0);
    asm volatile("s_waitcnt expcnt(0)\n\ts_nop 1" ::: "memory");
  }
  // 8-byte load (two adjacent u32, the first in the low half)
  using V2 = __attribute__((ext_vector_type(2))) unsigned int;
  __device__ uint64_t load2(uint32_t voff, uint32_t soff) const {
    return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
#else
  char* b;
  ZHD explicit GBuf(const uint32_t* p) : b(reinterpret_cast<char*>(const_cast<uint32_t*>(p))) {}
  ZHD GBuf(const uint32_t* p, bool live)
      : b(live ? reinterpret_cast<char*>(const_cast<uint32_t*>(p)) : nullptr) {}
  ZHD uint32_t load(uint32_t voff, uint32_t soff) const {
    return b ? *reinterpret_cast<const uint32_t*>(b + ((size_t)voff + soff)) : 0u;
  }
  ZHD void store(uint32_t v, uint32_t voff, uint32_t soff) const {
    if (b) *reinterpret_cast<uint32_t*>(b + ((size_t)voff + soff)) = v;
  }
  ZHD U4 load4(uint32_t voff, uint32_t soff) const {
    U4 v{0, 0, 0, 0};
    if (b) v = *reinterpret_cast<const U4*>(b + ((size_t)voff + soff));
    return v;
  }
  ZHD void store4(const U4& v, uint32_t voff, uint32_t soff) const {
    if (b) *reinterpret_cast<U4*>(b + ((size_t)voff + soff)) = v;
  }
  ZHD void store4_cached(const U4& v, uint32_t voff, uint32_t soff) const { store4(v, voff, soff); }
  ZHD uint64_t load2(uint32_t voff, uint32_t soff) const {
    return b ? *reinterpret_cast<const uint64_t*>(b + ((size_t)voff + soff)) : 0u;
  }
#endif
};

#if defined(__HIPCC__)
// A 16-byte global store followed by the wait states of GBuf::store4 (the
// same gfx950 store-data hazard).  NT selects the non-temporal cache policy,
// for output that no kernel reads back.
template <bool NT = false>
__device__ __forceinline__ void store16(uint32_t* p, const U4& v) {
  using V4 = __attribute__((ext_vector_type(4))) unsigned int;
  if constexpr (NT)
    __builtin_nontemporal_store(__builtin_bit_cast(V4, v), reinterpret_cast<V4*>(p));
  else
    *reinterpret_cast<V4*>(p) = __builtin_bit_cast(V4, v);
  asm volatile("s_waitcnt expcnt(0)\n\ts_nop 1" ::: "memory");
}
#endif

// One pass for thread `tid` of the work-group encoding tile `tile`.
//   LO, R    the digit this pass transforms
//   KIND     PASS_INTT / PASS_MID / PASS_NTT
//   FIRST    read the tile from global memory and write the even output rows
//   LAST     write canonical odd rows to global memory
//   LOA, RA  digit above, whose forward twiddle ends MID/NTT passes (RA = 0: none)
//   FULL     whole tile and k = 2^L: no bounds checks
//   BUF      LAST writes through a raw buffer based at the tile's first
//            column (FULL tiles in the row layout with
//            ((2 * 2^L - 1) * cols + W) * 4 < 2^32)
//   hook     called with the group index before each group is transformed
template <class PL, int LO, int R, int KIND, bool FIRST, bool LAST, int LOA, int RA, bool FULL,
          bool BUF, class Hook>
ZHD void encode_pass(int tid, uint32_t tile, const uint32_t* __restrict__ in,
                     uint32_t* __restrict__ out, uint32_t krows, uint32_t cols, int layout,
                     int32_t* __restrict__ lds, const EncodeTabs& T, const TwTab own,
                     const TwTab above, const TwTab twist, const Hook& hook) {
  constexpr int RR = 1 << R;
  constexpr int NGROUPS = PL::W << (PL::L - R);
  static_assert(NGROUPS % PL::NT == 0, "groups must divide evenly over the threads");
  constexpr int G = NGROUPS / PL::NT;
  constexpr uint32_t OUT_ROWS = 2u * PL::S;
  const Tw* rinv = T.roots + 64;
  const Tw* rfwd = T.roots;

  // The first pass loads up to 32 values per thread before computing, in
  // chunks of PB groups, to keep loads in flight without a large register set.
  constexpr int PB = FIRST ? (G * RR <= 32 ? G : (RR >= 32 ? 1 : 32 / RR)) : G;
  static_assert(G % PB == 0, "preload chunks must divide the groups");
#pragma unroll
  for (int gc = 0; gc < G; gc += PB) {
    uint32_t raw[FIRST ? PB : 1][FIRST ? RR : 1];
    if constexpr (FIRST) {
      // Rows of one group are 2^LO apart: one base address per group and a
      // uniform stride per value.
      const size_t istep = (size_t)cols << LO;
      const size_t ostep = layout == 0 ? (size_t)cols << (LO + 1) : (size_t)2 << LO;
#pragma unroll
      for (int gi = 0; gi < PB; gi++) {
        const uint32_t g = (uint32_t)((gc + gi) * PL::NT + tid);
        const uint32_t c = g % PL::W, u = g / PL::W;
        const uint32_t cidx = tile * PL::W + c;
        const uint32_t j0 = row_of<LO, R>(u, 0);
        const uint32_t* src = in + (size_t)j0 * cols + cidx;
#pragma unroll
        for (int m = 0; m < RR; m++) {
          const uint32_t j = j0 + ((uint32_t)m << LO);
          raw[gi][m] = (FULL || (cidx < cols && j < krows)) ? src[(size_t)m * istep] : 0u;
        }
      }
#pragma unroll
      for (int gi = 0; gi < PB; gi++) {
        const uint32_t g = (uint32_t)((gc + gi) * PL::NT + tid);
        const uint32_t c = g % PL::W, u = g / PL::W;
        const uint32_t cidx = tile * PL::W + c;
        if (FULL || cidx < cols) {
          uint32_t* dst = out + out_index(2 * row_of<LO, R>(u, 0), cidx, cols, OUT_ROWS, layout);
#pragma unroll
          for (int m = 0; m < RR; m++) store_out<PL>(dst + (size_t)m * ostep, raw[gi][m], layout);
        }
      }
    }
#pragma unroll
    for (int gi = 0; gi < PB; gi++) {
      if constexpr (PL::FENCE) group_fence();
      hook(gc + gi);
      const uint32_t g = (uint32_t)((gc + gi) * PL::NT + tid);
      const uint32_t c = g % PL::W, u = g / PL::W;
      double x[RR];
#pragma unroll
      for (int m = 0; m < RR; m++) {
        if constexpr (FIRST)
          x[m] = (double)raw[gi][m];
        else
          x[m] = (double)lds[row_of<LO, R>(u, m) * PL::W + c];
      }
      // The group's u32 twiddles (LDS tables) are read before its DFT, so that
      // their LDS latency hides behind the DFT's arithmetic.
      constexpr bool HOIST = PL::TWL && (KIND == PASS_INTT || (KIND == PASS_NTT && !LAST));
      uint32_t twr[HOIST ? RR : 1];
      if constexpr (HOIST) {
        if constexpr (KIND == PASS_INTT) {
          const uint32_t hb = ((u >> LO) << (LO + R)) | ((u & ((1u << LO) - 1u)) << R);
#pragma unroll
          for (int k = 0; k < RR; k++) twr[k] = own.s[hb + k];
        } else {
#pragma unroll
          for (int m = 0; m < RR; m++) {
            const uint32_t j = row_of<LO, R>(u, m);
            const uint32_t k = (j >> LOA) & ((1u << RA) - 1u);
            twr[m] = above.s[(k << LOA) | (j & ((1u << LOA) - 1u))];
          }
        }
      }
      if constexpr (KIND == PASS_INTT) {
        dft_reg<R>(x, rinv);
        const uint32_t base = ((u >> LO) << (LO + R)) | ((u & ((1u << LO) - 1u)) << R);
        constexpr bool LAST_INTT = PL::D == 2 || (PL::D == 3 && LO == PL::LO2);
        if constexpr (LAST_INTT) {
          // the table folds in the twist's per-group factor
#pragma unroll
          for (int k = 0; k < RR; k++) x[k] = HOIST ? mulmod(x[k], tw_u32(twr[k])) : mulmod(x[k], tw_get<PL::TWL>(own, base + k));
        } else {
          x[0] = reduce(x[0]);  // twiddle w^0 = 1
#pragma unroll
          for (int k = 1; k < RR; k++) x[k] = HOIST ? mulmod(x[k], tw_u32(twr[k])) : mulmod(x[k], tw_get<PL::TWL>(own, base + k));
        }
      } else {
        if constexpr (KIND == PASS_MID) {
          dft_reg<R>(x, rinv);
          if constexpr (PL::D == 1) {
#pragma unroll
            for (int m = 0; m < RR; m++)
              x[m] = mulmod(x[m], tw_get<PL::TWL>(twist, row_of<LO, R>(u, m)));
          } else {
            // rest of the twist: w_(2^(RD+1))^m, a constant per value
#pragma unroll
            for (int m = 1; m < RR; m++) {
              const Tw t = rfwd[m << (5 - R)];
              x[m] = mulmod(x[m], t.w, t.wq);
            }
          }
        }
        dft_reg<R>(x, rfwd);
        if constexpr (!LAST) {
          static_assert(RA > 0, "a non-final pass applies the twiddle of the digit above");
#pragma unroll
          for (int m = 0; m < RR; m++) {
            const uint32_t j = row_of<LO, R>(u, m);
            const uint32_t k = (j >> LOA) & ((1u << RA) - 1u);
            constexpr bool U32 = PL::TWL && !(PL::MID_PAIRS && KIND == PASS_MID);
            if constexpr (HOIST)
              x[m] = mulmod(x[m], tw_u32(twr[m]));
            else
              x[m] = mulmod(x[m], tw_get<U32>(above, (k << LOA) | (j & ((1u << LOA) - 1u))));
          }
        }
      }
      if constexpr (LAST && BUF) {
        static_assert(FULL, "buffer stores address whole tiles");
        const GBuf ob(out + (size_t)tile * PL::W);
        const uint32_t voff = ((2 * row_of<LO, R>(u, 0) + 1) * cols + c) * 4u;
        const uint32_t ostep = (cols << (LO + 1)) * 4u;
#pragma unroll
        for (int m = 0; m < RR; m++) ob.store(canonical(x[m]), voff, (uint32_t)m * ostep);
      } else if constexpr (LAST) {
        const uint32_t cidx = tile * PL::W + c;
        if (FULL || cidx < cols) {
          const size_t ostep = layout == 0 ? (size_t)cols << (LO + 1) : (size_t)2 << LO;
          uint32_t* dst = out + out_index(2 * row_of<LO, R>(u, 0) + 1, cidx, cols, OUT_ROWS, layout);
#pragma unroll
          for (int m = 0; m < RR; m++) store_out<PL>(dst + (size_t)m * ostep, canonical(x[m]), layout);
        }
      } else {
#pragma unroll
        for (int m = 0; m < RR; m++) lds[row_of<LO, R>(u, m) * PL::W + c] = (int32_t)x[m];
      }
    }
  }
}

// Copies the packed u32 twiddle tables behind the tile in LDS (plans with
// TWL); the caller puts a work-group barrier between this and step 0.
template <class PL>
ZHD void encode_load_tables(int tid, int32_t* __restrict__ lds, const EncodeTabs& T) {
  if constexpr (PL::TWL) {
    uint32_t* dst = reinterpret_cast<uint32_t*>(lds + PL::TILE_WORDS);
    for (uint32_t i = (uint32_t)tid; i < T.packed_words; i += PL::NT) dst[i] = T.packed[i];
  }
  if constexpr (PL::MID_PAIRS) {
    Tw* pairs = reinterpret_cast<Tw*>(lds + PL::TILE_WORDS + PL::TW_WORDS);
    for (uint32_t i = (uint32_t)tid; i < (1u << PL::LO1); i += PL::NT) pairs[i] = T.fwd2[i];
  }
}

// Step STEP (0 .. 2D-2) of the encode of one tile; steps are separated by a
// work-group barrier on the GPU.  FULL: every column of the tile exists and
// k = 2^L, so no bounds checks are emitted.
template <class PL, int STEP, bool FULL, bool BUF = false, class Hook = NoHook>
ZHD void encode_step(int tid, uint32_t tile, const uint32_t* __restrict__ in,
                     uint32_t* __restrict__ out, uint32_t krows, uint32_t cols, int layout,
                     int32_t* __restrict__ lds, const EncodeTabs& T,
                     const Hook& hook = Hook()) {
  constexpr int L1 = PL::LO1, L2 = PL::LO2;
  const uint32_t* sb = PL::TWL ? reinterpret_cast<const uint32_t*>(lds + PL::TILE_WORDS) : nullptr;
  const TwTab none{nullptr, nullptr};
  const TwTab inv1{T.inv1, PL::TWL ? sb + T.o_inv1 : nullptr};
  const TwTab inv2{T.inv2, PL::TWL ? sb + T.o_inv2 : nullptr};
  const TwTab fwd1{T.fwd1, PL::TWL ? sb + T.o_fwd1 : nullptr};
  const Tw* mid_pairs =
      PL::MID_PAIRS ? reinterpret_cast<const Tw*>(lds + PL::TILE_WORDS + PL::TW_WORDS) : T.fwd2;
  const TwTab fwd2{mid_pairs, PL::TWL ? sb + T.o_fwd2 : nullptr};
  const TwTab twist{T.twist, PL::TWL ? sb + T.o_twist : nullptr};
  if constexpr (PL::D == 1) {
    encode_pass<PL, 0, PL::RD, PASS_MID, true, true, 0, 0, FULL, BUF>(
        tid, tile, in, out, krows, cols, layout, lds, T, none, none, twist, hook);
  } else if constexpr (PL::D == 2) {
    if constexpr (STEP == 0)
      encode_pass<PL, L1, PL::R1, PASS_INTT, true, false, 0, 0, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, inv1, none, none, hook);
    else if constexpr (STEP == 1)
      encode_pass<PL, 0, PL::RD, PASS_MID, false, false, L1, PL::R1, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, none, fwd1, twist, hook);
    else
      encode_pass<PL, L1, PL::R1, PASS_NTT, false, true, 0, 0, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, none, none, none, hook);
  } else {
    if constexpr (STEP == 0)
      encode_pass<PL, L1, PL::R1, PASS_INTT, true, false, 0, 0, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, inv1, none, none, hook);
    else if constexpr (STEP == 1)
      encode_pass<PL, L2, PL::R2, PASS_INTT, false, false, 0, 0, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, inv2, none, none, hook);
    else if constexpr (STEP == 2)
      encode_pass<PL, 0, PL::RD, PASS_MID, false, false, L2, PL::R2, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, none, fwd2, twist, hook);
    else if constexpr (STEP == 3)
      encode_pass<PL, L2, PL::R2, PASS_NTT, false, false, L1, PL::R1, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, none, fwd1, none, hook);
    else
      encode_pass<PL, L1, PL::R1, PASS_NTT, false, true, 0, 0, FULL, BUF>(
          tid, tile, in, out, krows, cols, layout, lds, T, none, none, none, hook);
  }
}

// Register pipeline of the first pass for whole tiles in the row layout
// (plans with D >= 2).  load_part() brings part of a tile's first-pass values
// into registers -- issued group by group while the previous tile is still
// being transformed, so the loads are in flight behind that work and never
// queue up in one burst -- first() transforms them into the LDS tile, and
// store_even() writes them to the even output rows, one group at a time as
// first() reaches it.  BUF: every access goes through a raw buffer (GBuf) based
// at the tile's first column, which needs ((2 * 2^L - 1) * cols + W) * 4 < 2^32.
template <class PL, bool BUF = false>
struct Pipe {
  static constexpr int R = PL::R1, RR = 1 << R, LO = PL::LO1;
  static constexpr int G = (PL::W << (PL::L - R)) / PL::NT;
  static_assert(PL::D >= 2 && (PL::W << (PL::L - R)) % PL::NT == 0, "pipeline shape");

  // Byte offset of group gi's first value in a row-major array of `cols`
  // columns (`stride` rows per top-digit row: 1 for the input, 2 for the even
  // output rows), from base(array, tile).  Rows of the top digit are 2^LO
  // apart, so one uniform base per digit value plus this 32-bit offset
  // addresses every value; the caller guarantees (cols << (LO + 1)) * 4 < 2^32.
  static ZHD uint32_t offset(int tid, int gi, uint32_t tile, uint32_t cols, uint32_t stride) {
    const uint32_t g = (uint32_t)(gi * PL::NT + tid);
    const uint32_t c = g % PL::W, u = g / PL::W;  // u < 2^LO: the row within the digit
    return (u * stride * cols + (BUF ? 0u : tile * PL::W) + c) * 4u;
  }

  template <class T>
  static ZHD T* base(T* p, uint32_t tile) {
    return BUF ? p + (size_t)tile * PL::W : p;
  }

  template <class T>
  static ZHD T* at(T* base, uint32_t byte_offset) {
    return reinterpret_cast<T*>(reinterpret_cast<
        typename std::conditional<std::is_const<T>::value, const char, char>::type*>(base) + byte_offset);
  }

  // Value m of group gi (rows of the top digit are 2^LO apart).
  static ZHD uint32_t load_one(const uint32_t* __restrict__ in, uint32_t cols, uint32_t off, int m) {
    if constexpr (BUF) {
      return GBuf(in).load(off, (uint32_t)m * ((cols << LO) * 4u));
    } else {
      return *at(in + (size_t)m * ((size_t)cols << LO), off);
    }
  }

  static ZHD void load(int tid, uint32_t tile, const uint32_t* __restrict__ in, uint32_t cols,
                       uint32_t (&raw)[G][RR]) {
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
      const uint32_t off = offset(tid, gi, tile, cols, 1);
#pragma unroll
      for (int m = 0; m < RR; m++) raw[gi][m] = load_one(base(in, tile), cols, off, m);
    }
  }

  // Part p of np of load(): value q = gi * RR + m is loaded by part
  // q * np / (G * RR), so np calls made from successive groups of later passes
  // spread the loads over that work.  After unrolling every index is a
  // constant and the conditions fold away.
  static ZHD void load_part(int tid, uint32_t tile, const uint32_t* __restrict__ in, uint32_t cols,
                            uint32_t (&raw)[G][RR], int p, int np) {
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
      const uint32_t off = offset(tid, gi, tile, cols, 1);
#pragma unroll
      for (int m = 0; m < RR; m++)
        if ((gi * RR + m) * np / (G * RR) == p) raw[gi][m] = load_one(base(in, tile), cols, off, m);
    }
  }

  static ZHD void store_even(int tid, int gi, uint32_t tile, uint32_t* __restrict__ out,
                             uint32_t cols, const uint32_t (&r)[RR]) {
    const uint32_t off = offset(tid, gi, tile, cols, 2);
    if constexpr (BUF) {
      const GBuf ob(base(out, tile));
      const uint32_t ostep = (cols << (LO + 1)) * 4u;
#pragma unroll
      for (int m = 0; m < RR; m++) ob.store(r[m], off, (uint32_t)m * ostep);
    } else {
      const size_t ostep = (size_t)cols << (LO + 1);
#pragma unroll
      for (int m = 0; m < RR; m++) *at(out + (size_t)m * ostep, off) = r[m];
    }
  }

  // The first pass (INTT of the top digit) from registers into the LDS tile;
  // hook(gi) is called before group gi is transformed.
  template <class Hook = NoHook>
  static ZHD void first(int tid, const uint32_t (&raw)[G][RR], int32_t* __restrict__ lds,
                        const EncodeTabs& T, const TwTab inv1, const Hook& hook = Hook()) {
    const Tw* rinv = T.roots + 64;
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
      group_fence();
      hook(gi);
      const uint32_t g = (uint32_t)(gi * PL::NT + tid);
      const uint32_t c = g % PL::W, u = g / PL::W;
      double x[RR];
#pragma unroll
      for (int m = 0; m < RR; m++) x[m] = (double)raw[gi][m];
      dft_reg<R>(x, rinv);
      const uint32_t base = u << R;  // u < 2^LO for the top digit
      if constexpr (PL::D == 2) {
#pragma unroll
        for (int k = 0; k < RR; k++) x[k] = mulmod(x[k], tw_get<PL::TWL>(inv1, base + k));
      } else {
        x[0] = reduce(x[0]);
#pragma unroll
        for (int k = 1; k < RR; k++) x[k] = mulmod(x[k], tw_get<PL::TWL>(inv1, base + k));
      }
#pragma unroll
      for (int m = 0; m < RR; m++) lds[row_of<LO, R>(u, m) * PL::W + c] = (int32_t)x[m];
    }
  }
};

}  // namespace zoda
