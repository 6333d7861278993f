// Slice-pipelined column encode (side 2048, 32-column tiles), shared by the
// gfx950 kernel (encode_slice.hip) and the CPU emulation of the tests.
//
// Same transform as encode_core.hpp, per column of s = RA*RB*RC rows:
//
//     out[2j]     = data[j]
//     out[2j + 1] = NTT_s( t (.) INTT'_s(data) )[j],  t_i = s^-1 * w_2s^i
//
// (the reference's INTT-pad-NTT of src/benchmarks/utils.rs:100-146 with the
// zero half never formed), as three digits of the row index n = c + RC nB +
// RC RB nA, but with the middle digit pair held in registers instead of LDS:
//
//   P1  (front) DFT over nA (rows 128 apart, read from global memory) for one
//       (nB, c), times w_(RA RB)^-(nB kA)                            -> LDS A
//   P2  (front) DFT over nB for one (kA, c), times
//       s^-1 w_2s^((1 - 2c)(kA + RA kB)): the INTT's last twiddle with the
//       part of the twist that is constant over c                 -> registers
//   mid         per (kA, kB): DFT over c, the rest of the twist w_(2 RC)^kC,
//       the forward DFT over kC -> j1, times w_(RB RC)^(kB j1)     -> registers
//   P2' (back)  DFT over kB for one (kA, j1), times w_s^(kA (j1 + RC j2))
//                                                                  -> LDS B
//   P1' (back)  DFT over kA for one (j1, j2) -> j3, canonical, odd output row
//       2 (j1 + RC j2 + RC RB j3) + 1                              -> global
//
// A work-group of 512 threads owns a tile of 32 adjacent columns (every row
// segment a whole 128-byte line).  Thread (wave w, half h, column col) holds,
// as owner of kA = 2w + h, the RB x RC = 128 values of its (column, kA) as
// int32 between P2 and P2' -- the whole tile (256 KiB) lives in registers at
// the middle of the transform, so LDS only carries one step's exchange.  A
// step handles two slices (c or j1 = 2 step + h): P1 and P1' run one
// (column, w, h) task per thread with w = nB (resp. j2), so the P1 twiddles
// are uniform per wave.  The front of a tile and the back of the previous one
// share each step, so the loads and stores stream while the registers change
// hands slot by slot: the back reads the slots of j1 = 2 step + h before the
// front writes those of c = 2 step + h.
//
// Arithmetic (fp64.hpp): integer-valued doubles, additions exact; every
// constant multiplication takes four instructions, with two opaque register
// constants (MulK).  A multiplication whose result leaves the thread (LDS or
// held registers) ends in an integer subtraction of low words that yields
// the int32 directly: the conversion is free.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "encode_core.hpp"
#include "field.hpp"
#include "fp64.hpp"
#include "fp64_ops.hpp"

namespace zoda {
namespace slice {


// Natural-order DFT of 2^R values with the roots w_(2^R)^(+-1) (INV selects
// the inverse roots), radix-2 decimation in frequency; the bit-reversed
// result is renamed to natural order.  |x| < B on entry gives |X| < 2^R B.
// Each stage's multiplications run as four batches of independent
// instructions (every y, then every c, t and r): a multiplication is a chain
// of four dependent FP64 instructions, and with two waves per SIMD the chains
// of one wave must overlap each other to keep the VALU busy.
struct NoStageHook {
  ZHD void operator()(int) const {}
};

template <int R, bool INV, class Hook = NoStageHook>
ZHD void dft(double (&x)[1 << R], const MulK& k, const Hook& hook = Hook()) {
  constexpr int N = 1 << R;
  constexpr int D = INV ? 1 : 0;
  constexpr int M = N / 2;   // butterflies per stage
  constexpr int CH = 4;      // multiplications per batch
#pragma unroll
  for (int hs = R - 1; hs >= 0; hs--) {
    const int half = 1 << hs;
    double d[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      const int b = (i / half) * 2 * half, j = i % half;
      const double u = x[b + j], v = x[b + j + half];
      x[b + j] = u + v;
      d[i] = u - v;
    }
#pragma unroll
    for (int i0 = 0; i0 < M; i0 += CH) {
      double y[CH], c[CH];
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int i = i0 + q, j = i % half;
        if (i < M && j) y[q] = fmad(d[i], kRoots.wq[D][j << (6 - hs)], k.m);
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int i = i0 + q, j = i % half;
        if (i < M && j) c[q] = fmad(y[q], k.kd, k.kr);
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int i = i0 + q, j = i % half;
        if (i < M && j) d[i] = fmad(d[i], kRoots.w[D][j << (6 - hs)], -c[q]);
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int i = i0 + q, j = i % half;
        if (i < M && j) d[i] = d[i] - y[q];
      }
    }
#pragma unroll
    for (int i = 0; i < M; i++) {
      const int b = (i / half) * 2 * half, j = i % half;
      x[b + j + half] = d[i];
    }
    hook(R - 1 - hs);
  }
  double z[N];
#pragma unroll
  for (int i = 0; i < N; i++) z[i] = x[brev_bits(i, R)];
#pragma unroll
  for (int i = 0; i < N; i++) x[i] = z[i];
}

// The first radix-2 stage's twiddles of a DFT of 2^R values split between two
// threads: x_n (n < 2^(R-1)) times w_(2^R)^(-+n) where hd = 1 (odd outputs),
// times 1 where hd = 0; the per-lane constant is 1 + hd (w - 1), exact, and
// its quotient by p one multiply.  In batches of four, as in dft().
template <int R, bool INV>
ZHD void split_input_twist(double (&x)[1 << (R - 1)], double hd, const MulK& k) {
  constexpr int H = 1 << (R - 1);
  constexpr int D = INV ? 1 : 0;
#pragma unroll
  for (int n0 = 1; n0 < H; n0 += 4) {
    double w[4], wq[4], y[4], c[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (n0 + q < H) {
        w[q] = fmad(hd, kRoots.w[D][(n0 + q) << (7 - R)] - 1.0, 1.0);
        wq[q] = w[q] * PINV;
      }
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (n0 + q < H) y[q] = fmad(x[n0 + q], wq[q], k.m);
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (n0 + q < H) c[q] = fmad(y[q], k.kd, k.kr);
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (n0 + q < H) x[n0 + q] = fmad(x[n0 + q], w[q], -c[q]) - y[q];
  }
}

// n values times n twiddles, as int32 (mul_mem), in batches of four.
template <int NV, class TW>
ZHD void mul_mem_batch(const double (&x)[NV], const TW& tw, int32_t (&r)[NV], const MulK& k) {
  constexpr int CH = 4;
#pragma unroll
  for (int i0 = 0; i0 < NV; i0 += CH) {
    double y[CH], c[CH];
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (i0 + q < NV) y[q] = fmad(x[i0 + q], tw(i0 + q).wq, k.m);
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (i0 + q < NV) c[q] = fmad(y[q], k.kd, k.kr);
#pragma unroll
    for (int q = 0; q < CH; q++) {
      if (i0 + q < NV) {
        const double t = fmad(x[i0 + q], tw(i0 + q).w, -c[q]);
        r[i0 + q] = (int32_t)((uint32_t)__builtin_bit_cast(uint64_t, t) -
                              (uint32_t)__builtin_bit_cast(uint64_t, y[q]));
      }
    }
  }
}

// ---- the plan ------------------------------------------------------------------------

// The wave index of thread tid, wave-uniform (a scalar register on the
// device).  Row offsets built from it stay scalar, so the raw-buffer stores
// take them as their uniform offset; from the plain shift the compiler cannot
// tell that the offset is the same on every lane and wraps each store in a
// loop over its distinct values.
ZHD int wave_of(int tid) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(tid >> 6);
#else
  return tid >> 6;
#endif
}


struct Plan2048 {
  static constexpr int LRA = 4, LRB = 3, LRC = 4;
  static constexpr bool SPLIT = false;  // P1 / P1' DFTs whole in one thread
  static constexpr bool WIDE = false;   // every output offset within 32 bits of the tile base
  using TwT = double;                   // P2 / P2' twiddles in LDS
  using T1T = Tw2;                      // P1 twiddles, global memory
  static constexpr int RA = 1 << LRA, RB = 1 << LRB, RC = 1 << LRC;
  static constexpr int L = LRA + LRB + LRC, S = 1 << L;
  static constexpr int W = 32, NT = 512, WAVES = NT / 64;
  static constexpr int SPS = 2;           // slices per step (one per half-wave)
  static constexpr int STEPS = RC / SPS;  // steps per tile
  static_assert(RB == WAVES, "P1/P1': the wave index is nB (resp. j2)");
  static_assert(RA == 2 * WAVES, "owners: kA = 2 w + h");
  static_assert(W == 32, "half-waves hold the 32 columns");
  // 16-byte global accesses: lane = 8 column quads x 8 rows; wave w, access i
  // < QI covers the rows of slice w / 4 with nA (or j3) = (4 w + i) mod 16
  static constexpr int QI = 4;
  static_assert(QI * 8 * WAVES == SPS * RA * RB, "one 16-byte access per lane covers a step");
  // Even output rows of step s: copied (read again) in middle group s when
  // bit s of EVEN_MID_MASK is set, else written from the staged input rows in
  // the steps.  The copies keep memory requests in flight through the middle
  // passes, which have no memory work of their own; with every other step's
  // even rows there, memory work and arithmetic are in about the same ratio
  // in the middle passes as in the steps (measured best of none, every
  // fourth, every other and all steps).
  static constexpr unsigned EVEN_MID_MASK = 0x55u;
  static constexpr ZHD bool even_in_mid(int s) { return (EVEN_MID_MASK >> s) & 1u; }
  static_assert(EVEN_MID_MASK == 0 || RB == STEPS, "one step of even rows per middle group");
  static constexpr int XWORDS = SPS * RB * RA * W;  // one LDS buffer (int32 / u32)
  static constexpr int T1_PAIRS = RB * RA;          // [nB][kA] (w, w/p) pairs, global memory
  static constexpr int T2_WORDS = RC * RA * RB;     // [c][kA][kB] doubles in LDS
  static constexpr int T3_WORDS = RC * RA * RB;     // [j1][kA][j2] doubles in LDS
  // LDS: exchange A (P1 -> P2), exchange B (P2' -> P1'), the step's input
  // rows (IN), the step's odd rows (ODD), then the two twiddle tables
  static constexpr int A_OFF = 0, B_OFF = XWORDS, IN_OFF = 2 * XWORDS, ODD_OFF = 3 * XWORDS;
  static constexpr int LDS_BYTES = 4 * XWORDS * 4 + (T2_WORDS + T3_WORDS) * 8;
  // exchange index of (slice s, digit d in {nB, j2}, kA, column)
  static ZHD int xi(int s, int d, int ka, int col) { return ((s * RB + d) * RA + ka) * W + col; }
  // staging index (IN / ODD) of (slice s, digit a in {nA, j3}, digit b in {nB, j2}, column)
  static ZHD int si(int s, int a, int b, int col) { return ((s * RA + a) * RB + b) * W + col; }
  // Threads: wave w; task (P1, P1') slice s = half-wave, column = lane mod 32,
  // nB (resp. j2) = w; owner (P2, middle, P2') the same column and kA = 2 w + s;
  // 16-byte access lane: column quad q, row g.
  static ZHD void decode(int tid, int& w, int& s, int& hh, int& col, int& ocol, int& oka, int& q, int& g) {
    w = wave_of(tid);
    s = (tid >> 5) & 1;
    hh = 0;
    col = tid & 31;
    ocol = col;
    oka = 2 * w + s;
    q = tid & 7;
    g = (tid >> 3) & 7;
  }
  // Access i of wave w in step sp covers input rows (2 sp + s) + RC g + RC RB a
  // (slice s = w / 4, a = (4 w + i) mod 16, nB = g), four columns 4q..4q+3.
  static ZHD int slice_of(int w) { return w / (WAVES / SPS); }
  static ZHD int a_of(int w, int i) { return (QI * w + i) % RA; }
  static ZHD int lane_row(int g) { return RC * g; }
  static ZHD int acc_row(int sp, int w, int i) { return SPS * sp + slice_of(w) + RC * RB * a_of(w, i); }
  static ZHD int acc_lds(int w, int i, int g, int q) { return si(slice_of(w), a_of(w, i), g, 4 * q); }
};

// Side 4096: 32 x 8 x 16 digits, 16-column tiles (64-byte row segments, as
// wide as a tile can be with 128 held values per thread).  The radix-32 P1 and
// P1' DFTs are split between two threads: a first radix-2 stage on all 32
// inputs, then one radix-16 DFT per thread for the even (hh = 0) or odd
// (hh = 1) outputs.  Outputs pass 32-bit offsets (16 GiB), so every access
// addresses from a per-access base.
struct Plan4096 {
  static constexpr int LRA = 5, LRB = 3, LRC = 4;
  static constexpr bool SPLIT = true;
  static constexpr bool WIDE = true;
  using TwT = uint32_t;  // u32 in LDS: as doubles the two tables would not fit
  struct T1T;            // defined below
  static constexpr int RA = 1 << LRA, RB = 1 << LRB, RC = 1 << LRC;
  static constexpr int L = LRA + LRB + LRC, S = 1 << L;
  static constexpr int W = 16, NT = 512, WAVES = NT / 64;
  static constexpr int SPS = 2;
  static constexpr int STEPS = RC / SPS;
  static_assert(RB == WAVES, "P1/P1': the wave index is nB (resp. j2)");
  static_assert(W * SPS * 2 == 64, "a wave holds column x slice x half tasks");
  static_assert(W * RA == NT, "owners (column, kA)");
  static constexpr int QI = 4;
  static_assert(QI * 16 * WAVES == SPS * RA * RB, "one 16-byte access per lane covers a step");
  // Every step's even rows are written from the staged input (no second
  // reads in the middle passes): at this side the extra reads cost more than
  // the memory work they give the middle passes (4096^2: 11.66-11.69 ms with
  // none against 11.98 with steps 0 and 4 copied and 12.33 with every other
  // step, alternating runs on one box).
  static constexpr unsigned EVEN_MID_MASK = 0x00u;
  static constexpr ZHD bool even_in_mid(int s) { return (EVEN_MID_MASK >> s) & 1u; }
  static_assert(EVEN_MID_MASK == 0 || RB == STEPS, "one step of even rows per middle group");
  static constexpr int XWORDS = SPS * RB * RA * W;
  static constexpr int T1_PAIRS = RB * RA / 2;      // [nB][m] (base, diff) pairs, global memory
  static constexpr int T2_WORDS = RC * RA * RB;     // [c][kA][kB] u32 in LDS
  static constexpr int T3_WORDS = RC * RA * RB;     // [j1][kA][j2] u32 in LDS
  static constexpr int A_OFF = 0, B_OFF = XWORDS, IN_OFF = 2 * XWORDS, ODD_OFF = 3 * XWORDS;
  static constexpr int LDS_BYTES = 4 * XWORDS * 4 + (T2_WORDS + T3_WORDS) * 4;
  static_assert(LDS_BYTES <= 160 * 1024, "LDS of one CU");
  static ZHD int xi(int s, int d, int ka, int col) { return ((s * RB + d) * RA + ka) * W + col; }
  static ZHD int si(int s, int a, int b, int col) { return ((s * RA + a) * RB + b) * W + col; }
  // Threads: wave w; task slice s, half hh, column (lane bits 0-3, 4, 5),
  // nB (resp. j2) = w; owner column = lane mod 16, kA = 4 w + lane / 16;
  // 16-byte access lane: column quad q (bits 0-1), row g (bits 2-5).
  static ZHD void decode(int tid, int& w, int& s, int& hh, int& col, int& ocol, int& oka, int& q, int& g) {
    w = wave_of(tid);
    s = (tid >> 4) & 1;
    hh = (tid >> 5) & 1;
    col = tid & 15;
    ocol = col;
    oka = 4 * w + ((tid >> 4) & 3);
    q = tid & 3;
    g = (tid >> 2) & 15;
  }
  // Access i of wave w in step sp covers input rows (2 sp + s) + RC nB +
  // RC RB nA with slice s = w / 4, nA = 2 a + g / 8 (a = (4 w + i) mod 16),
  // nB = g mod 8, four columns 4q..4q+3.
  static ZHD int slice_of(int w) { return w / (WAVES / SPS); }
  static ZHD int a_of(int w, int i) { return (QI * w + i) % (RA / 2); }
  static ZHD int lane_row(int g) { return RC * (g & 7) + RC * RB * (g >> 3); }
  static ZHD int acc_row(int sp, int w, int i) { return SPS * sp + slice_of(w) + 2 * RC * RB * a_of(w, i); }
  static ZHD int acc_lds(int w, int i, int g, int q) {
    return si(slice_of(w), 2 * a_of(w, i) + (g >> 3), g & 7, 4 * q);
  }
};

// P1 twiddles of a split plan for one wave (nB): output m of half hh is kA =
// 2m + hh, twiddle w_(RA RB)^-(nB kA) = base + hh * diff (exact integers).
struct alignas(16) Plan4096::T1T {
  double base;
  double diff;
};

// Host-built twiddle tables.
//   t1[nB][kA]     = w_(RA RB)^-(nB kA)                  (w, w/p) pairs; for a
//                    split plan [nB][m] = (T(2m), T(2m+1) - T(2m))
//   t2[c][kA][kB]  = s^-1 w_2s^((1 - 2c)(kA + RA kB))    w only
//   t3[j1][kA][j2] = w_s^(kA (j1 + RC j2))               w only
template <class PL>
inline void build_slice_tables(typename PL::T1T* t1, typename PL::TwT* t2, typename PL::TwT* t3) {
  const uint32_t ws = root_of_unity(PL::L), w2s = root_of_unity(PL::L + 1);
  const uint32_t wm = pow_mod_host(ws, PL::RC);  // w_(RA RB)
  const uint32_t wmi = inv_mod_host(wm);
  const uint32_t sinv = inv_mod_host((uint32_t)PL::S);
  auto t1v = [&](int nb, int ka) { return pow_mod_host(wmi, (uint64_t)nb * ka); };
  for (int nb = 0; nb < PL::RB; nb++) {
    if constexpr (PL::SPLIT) {
      for (int m = 0; m < PL::RA / 2; m++) {
        const double a = (double)t1v(nb, 2 * m), b = (double)t1v(nb, 2 * m + 1);
        t1[nb * (PL::RA / 2) + m] = typename PL::T1T{a, b - a};
      }
    } else {
      for (int ka = 0; ka < PL::RA; ka++) {
        const uint32_t v = t1v(nb, ka);
        t1[nb * PL::RA + ka] = Tw2{(double)v, (double)v / PD};
      }
    }
  }
  const int64_t ord2s = 2ll * PL::S;
  for (int c = 0; c < PL::RC; c++)
    for (int ka = 0; ka < PL::RA; ka++)
      for (int kb = 0; kb < PL::RB; kb++) {
        const int64_t e = (int64_t)(1 - 2 * c) * (ka + PL::RA * kb);
        const uint64_t em = (uint64_t)(((e % ord2s) + ord2s) % ord2s);
        t2[(c * PL::RA + ka) * PL::RB + kb] = (typename PL::TwT)mul_mod_host(sinv, pow_mod_host(w2s, em));
      }
  for (int j1 = 0; j1 < PL::RC; j1++)
    for (int ka = 0; ka < PL::RA; ka++)
      for (int j2 = 0; j2 < PL::RB; j2++)
        t3[(j1 * PL::RA + ka) * PL::RB + j2] =
            (typename PL::TwT)pow_mod_host(ws, (uint64_t)ka * (j1 + PL::RC * j2) % PL::S);
}

// A twiddle w read as a double with its quotient w/p formed by one multiply.
template <class T = double>
struct TwW {
  const T* p;
  ZHD Tw2 operator()(int i) const {
    const double w = (double)p[i];
    return Tw2{w, w * PINV};
  }
};

// Thread coordinates (the plan's decode): wave w; as a P1 / P1' task slice
// s, DFT half hh and column col; as an owner column ocol and kA oka; for the
// 16-byte accesses column quad q and row g.
template <class PL>
struct Tid {
  int w, s, hh, col, ocol, oka, q, g;
  ZHD explicit Tid(int tid) { PL::decode(tid, w, s, hh, col, ocol, oka, q, g); }
};

// A tile's columns in a row-major matrix, for plans whose offsets pass 32
// bits: each access builds its own buffer descriptor at its row.
struct TileRef {
  const uint32_t* p;
  bool live;
};

// 16-byte accesses at a per-lane byte offset voff plus the uniform rows
// base + off: a 32-bit soffset of a tile's buffer, or, for a TileRef, a
// descriptor at row `base` (the same for every access of a step, so one per
// step) and a 32-bit soffset of `off` rows.
ZHD U4 acc_load4(const GBuf& b, uint32_t voff, uint32_t base, uint32_t off, uint32_t c4) {
  return b.load4(voff, (base + off) * c4);
}
ZHD void acc_store4(const GBuf& b, const U4& v, uint32_t voff, uint32_t base, uint32_t off, uint32_t c4) {
  b.store4(v, voff, (base + off) * c4);
}
ZHD U4 acc_load4(const TileRef& b, uint32_t voff, uint32_t base, uint32_t off, uint32_t c4) {
  return GBuf(b.p + (size_t)base * (c4 / 4u), b.live).load4(voff, off * c4);
}
ZHD void acc_store4(const TileRef& b, const U4& v, uint32_t voff, uint32_t base, uint32_t off, uint32_t c4) {
  GBuf(const_cast<uint32_t*>(b.p) + (size_t)base * (c4 / 4u), b.live).store4(v, voff, off * c4);
}

// Byte offsets of the 16-byte accesses: a per-thread voffset from the tile's
// first column plus the uniform row of access i of step sp (PL::acc_row):
// input row acc_row + lane_row, even output row twice that, odd output row
// twice that plus one (the odd row of P1' value j1 + RC j2 + RC RB j3 has the
// same digits as input row c + RC nB + RC RB nA).
template <class PL>
struct Offs {
  uint32_t in, ev, od;
  ZHD Offs(const Tid<PL>& t, uint32_t cols) {
    const uint32_t r = (uint32_t)PL::lane_row(t.g);
    in = (r * cols + 4u * (uint32_t)t.q) * 4u;
    ev = (2 * r * cols + 4u * (uint32_t)t.q) * 4u;
    od = ((2 * r + 1) * cols + 4u * (uint32_t)t.q) * 4u;
  }
  // uniform rows of access i (c4 = cols * 4, the byte stride of a row, is
  // applied where the offsets are used: the kernel hides its value per tile,
  // so they are not hoisted out of the tile loop into scalar registers that
  // spill)
  static ZHD uint32_t r_in(int sp, int w, int i) { return (uint32_t)PL::acc_row(sp, w, i); }
  static ZHD uint32_t r_out(int sp, int w, int i) { return 2u * r_in(sp, w, i); }
  // the same split into the step's first access and the offset of access i
  static ZHD uint32_t b_in(int sp, int w) { return r_in(sp, w, 0); }
  static ZHD uint32_t o_in(int sp, int w, int i) { return r_in(sp, w, i) - r_in(sp, w, 0); }
  static ZHD uint32_t b_out(int sp, int w) { return 2u * b_in(sp, w); }
  static ZHD uint32_t o_out(int sp, int w, int i) { return 2u * o_in(sp, w, i); }
  // staging index (IN / ODD) of access i's four columns
  static ZHD int lds(const Tid<PL>& t, int i) { return PL::acc_lds(t.w, i, t.g, t.q); }
};

// The 16-byte loads of one step's input rows (B: GBuf or TileRef).
template <class PL, class B>
ZHD void load_step(const B& ib, const Offs<PL>& o, int w, int sp, uint32_t c4, U4 (&r)[PL::QI]) {
#pragma unroll
  for (int i = 0; i < PL::QI; i++) r[i] = acc_load4(ib, o.in, Offs<PL>::b_in(sp, w), Offs<PL>::o_in(sp, w, i), c4);
}

// A step's loaded rows into the IN staging and, with EVEN, to the even
// output rows.
template <class PL, bool EVEN = true, class B = GBuf>
ZHD void stage_step(const Tid<PL>& t, const B& ob, const Offs<PL>& o, int sp, uint32_t c4,
                    const U4 (&r)[PL::QI], int32_t* __restrict__ lds) {
#pragma unroll
  for (int i = 0; i < PL::QI; i++) {
    if constexpr (EVEN) acc_store4(ob, r[i], o.ev, Offs<PL>::b_out(sp, t.w), Offs<PL>::o_out(sp, t.w, i), c4);
    *reinterpret_cast<U4*>(lds + PL::IN_OFF + Offs<PL>::lds(t, i)) = r[i];
  }
}

// The even output rows of access i of step sp (copied from registers).
template <class PL, class B>
ZHD void store_even(const B& ob, const Offs<PL>& o, int w, int sp, uint32_t c4, const U4 (&r)[PL::QI]) {
#pragma unroll
  for (int i = 0; i < PL::QI; i++) acc_store4(ob, r[i], o.ev, Offs<PL>::b_out(sp, w), Offs<PL>::o_out(sp, w, i), c4);
}

// The odd rows P1' left in the ODD staging, as 16-byte stores.
template <class PL, class B>
ZHD void store_odd(const Tid<PL>& t, const B& ob, const Offs<PL>& o, int sp, uint32_t c4,
                   const int32_t* __restrict__ lds) {
  U4 v[PL::QI];
#pragma unroll
  for (int i = 0; i < PL::QI; i++) v[i] = *reinterpret_cast<const U4*>(lds + PL::ODD_OFF + Offs<PL>::lds(t, i));
#pragma unroll
  for (int i = 0; i < PL::QI; i++) acc_store4(ob, v[i], o.od, Offs<PL>::b_out(sp, t.w), Offs<PL>::o_out(sp, t.w, i), c4);
}

// P1: INTT over nA of the thread's (nB = w, c = 2 sp + s) values (IN) -> A.
template <class PL>
ZHD void p1(const Tid<PL>& t, int32_t* __restrict__ lds, const typename PL::T1T* __restrict__ t1w,
            const MulK& k) {
  group_fence();
  const uint32_t* in = reinterpret_cast<const uint32_t*>(lds + PL::IN_OFF);
  if constexpr (!PL::SPLIT) {
    double x[PL::RA];
#pragma unroll
    for (int a = 0; a < PL::RA; a++) x[a] = (double)in[PL::si(t.s, a, t.w, t.col)];
    dft<PL::LRA, true>(x, k);
    int32_t r[PL::RA];
    mul_mem_batch(x, [&](int i) -> const Tw2& { return t1w[i]; }, r, k);
#pragma unroll
    for (int ka = 0; ka < PL::RA; ka++) lds[PL::A_OFF + PL::xi(t.s, t.w, ka, t.col)] = r[ka];
  } else {
    // Split radix-2^LRA DFT: kA = 2m + hh.  First stage u_n = x_n + x_(n+H),
    // v_n = (x_n - x_(n+H)) w_RA^-n (hh selects u or v per lane), then a
    // radix-H DFT; the output twiddle w_(RA RB)^-(nB kA) per lane.
    constexpr int H = PL::RA / 2;
    const double hd = (double)t.hh, sgn = 1.0 - 2.0 * hd;
    double x[H];
#pragma unroll
    for (int n = 0; n < H; n++)
      x[n] = fmad((double)in[PL::si(t.s, n + H, t.w, t.col)], sgn, (double)in[PL::si(t.s, n, t.w, t.col)]);
    split_input_twist<PL::LRA, true>(x, hd, k);
    dft<PL::LRA - 1, true>(x, k);
    int32_t r[H];
    mul_mem_batch(x, [&](int m) -> Tw2 {
      const double w = fmad(hd, t1w[m].diff, t1w[m].base);
      return Tw2{w, w * PINV};
    }, r, k);
#pragma unroll
    for (int m = 0; m < H; m++) lds[PL::A_OFF + PL::xi(t.s, t.w, 2 * m + t.hh, t.col)] = r[m];
  }
}

// P2: INTT over nB for the owner's kA and the two slices c = 2 SP + s.
template <class PL, int SP>
ZHD void p2(const Tid<PL>& t, const int32_t* __restrict__ lds, const typename PL::TwT* __restrict__ t2,
            int32_t (&held)[PL::RB][PL::RC], const MulK& k) {
  const int ka = t.oka;
#pragma unroll
  for (int s = 0; s < PL::SPS; s++) {
    group_fence();
    constexpr int c0 = PL::SPS * SP;
    double x[PL::RB];
#pragma unroll
    for (int nb = 0; nb < PL::RB; nb++) x[nb] = (double)lds[PL::A_OFF + PL::xi(s, nb, ka, t.ocol)];
    dft<PL::LRB, true>(x, k);
    int32_t r[PL::RB];
    mul_mem_batch(x, TwW<typename PL::TwT>{t2 + ((c0 + s) * PL::RA + ka) * PL::RB}, r, k);
#pragma unroll
    for (int kb = 0; kb < PL::RB; kb++) held[kb][c0 + s] = r[kb];
  }
}

// The middle of one (kA, kB): INTT over c, twist, NTT over kC, twiddle.
template <class PL, int KB>
ZHD void mid_group(int32_t (&held)[PL::RB][PL::RC], const MulK& k) {
  group_fence();
  double x[PL::RC];
#pragma unroll
  for (int c = 0; c < PL::RC; c++) x[c] = (double)held[KB][c];
  dft<PL::LRC, true>(x, k);
  // twist w_(2 RC)^kC = w_128^(kC * 64 / RC), in batches of four as in dft()
  constexpr int TS = 64 / PL::RC;
#pragma unroll
  for (int k0 = 0; k0 < PL::RC; k0 += 4) {
    double y[4], c[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q) y[q] = fmad(x[k0 + q], kRoots.wq[0][(k0 + q) * TS], k.m);
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q) c[q] = fmad(y[q], k.kd, k.kr);
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q) x[k0 + q] = fmad(x[k0 + q], kRoots.w[0][(k0 + q) * TS], -c[q]);
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (k0 + q) x[k0 + q] = x[k0 + q] - y[q];
  }
  dft<PL::LRC, false>(x, k);
  // twiddle w_s^(RA kB j1) = w_(RB RC)^(kB j1) = w_128^(kB j1 * 128 / (RB RC))
  static_assert(PL::RB * PL::RC <= 128, "mid twiddles come from the w_128 table");
  constexpr int SC = 128 / (PL::RB * PL::RC);
  int32_t r[PL::RC];
  mul_mem_batch(x, [](int j1) -> Tw2 {
    const int e = (KB * j1 * SC) & 127;
    return Tw2{kRoots.w[0][e], kRoots.wq[0][e]};
  }, r, k);
#pragma unroll
  for (int j1 = 0; j1 < PL::RC; j1++) held[KB][j1] = r[j1];
}

// P2': NTT over kB for the owner's kA and the two slices j1 = 2 SP + s -> B.
template <class PL, int SP>
ZHD void p2b(const Tid<PL>& t, const int32_t (&held)[PL::RB][PL::RC], int32_t* __restrict__ lds,
             const typename PL::TwT* __restrict__ t3, const MulK& k) {
  const int ka = t.oka;
#pragma unroll
  for (int s = 0; s < PL::SPS; s++) {
    group_fence();
    constexpr int j0 = PL::SPS * SP;
    double x[PL::RB];
#pragma unroll
    for (int kb = 0; kb < PL::RB; kb++) x[kb] = (double)held[kb][j0 + s];
    dft<PL::LRB, false>(x, k);
    int32_t r[PL::RB];
    mul_mem_batch(x, TwW<typename PL::TwT>{t3 + ((j0 + s) * PL::RA + ka) * PL::RB}, r, k);
#pragma unroll
    for (int j2 = 0; j2 < PL::RB; j2++) lds[PL::B_OFF + PL::xi(s, j2, ka, t.ocol)] = r[j2];
  }
}

// P1': NTT over kA of the thread's (j2 = w, j1 = 2 sp + s) values (B),
// canonical -> the ODD staging (odd rows 2 (j1 + RC j2 + RC RB j3) + 1).
template <class PL>
ZHD void p1b(const Tid<PL>& t, int32_t* __restrict__ lds, const MulK& k) {
  group_fence();
  uint32_t* odd = reinterpret_cast<uint32_t*>(lds + PL::ODD_OFF);
  if constexpr (!PL::SPLIT) {
    double x[PL::RA];
#pragma unroll
    for (int ka = 0; ka < PL::RA; ka++) x[ka] = (double)lds[PL::B_OFF + PL::xi(t.s, t.w, ka, t.col)];
    dft<PL::LRA, false>(x, k);
#pragma unroll
    for (int j3 = 0; j3 < PL::RA; j3++) odd[PL::si(t.s, j3, t.w, t.col)] = canon(x[j3], k);
  } else {
    // split as in p1, forward: j3 = 2m + hh
    constexpr int H = PL::RA / 2;
    const double hd = (double)t.hh, sgn = 1.0 - 2.0 * hd;
    double x[H];
#pragma unroll
    for (int n = 0; n < H; n++)
      x[n] = fmad((double)lds[PL::B_OFF + PL::xi(t.s, t.w, n + H, t.col)], sgn,
                  (double)lds[PL::B_OFF + PL::xi(t.s, t.w, n, t.col)]);
    split_input_twist<PL::LRA, false>(x, hd, k);
    dft<PL::LRA - 1, false>(x, k);
#pragma unroll
    for (int m = 0; m < H; m++) odd[PL::si(t.s, 2 * m + t.hh, t.w, t.col)] = canon(x[m], k);
  }
}

}  // namespace slice
}  // namespace zoda
