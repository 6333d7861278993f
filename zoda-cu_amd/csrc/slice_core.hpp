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

struct Plan2048 {
  static constexpr int LRA = 4, LRB = 3, LRC = 4;
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
};

// Host-built twiddle tables.
//   t1[nB][kA]     = w_(RA RB)^-(nB kA)                  (w, w/p) pairs
//   t2[c][kA][kB]  = s^-1 w_2s^((1 - 2c)(kA + RA kB))    w only
//   t3[j1][kA][j2] = w_s^(kA (j1 + RC j2))               w only
template <class PL>
inline void build_slice_tables(Tw2* t1, double* t2, double* t3) {
  const uint32_t ws = root_of_unity(PL::L), w2s = root_of_unity(PL::L + 1);
  const uint32_t wm = pow_mod_host(ws, PL::RC);  // w_(RA RB)
  const uint32_t wmi = inv_mod_host(wm);
  const uint32_t sinv = inv_mod_host((uint32_t)PL::S);
  for (int nb = 0; nb < PL::RB; nb++)
    for (int ka = 0; ka < PL::RA; ka++) {
      const uint32_t v = pow_mod_host(wmi, (uint64_t)nb * ka);
      t1[nb * PL::RA + ka] = Tw2{(double)v, (double)v / PD};
    }
  const int64_t ord2s = 2ll * PL::S;
  for (int c = 0; c < PL::RC; c++)
    for (int ka = 0; ka < PL::RA; ka++)
      for (int kb = 0; kb < PL::RB; kb++) {
        const int64_t e = (int64_t)(1 - 2 * c) * (ka + PL::RA * kb);
        const uint64_t em = (uint64_t)(((e % ord2s) + ord2s) % ord2s);
        t2[(c * PL::RA + ka) * PL::RB + kb] = (double)mul_mod_host(sinv, pow_mod_host(w2s, em));
      }
  for (int j1 = 0; j1 < PL::RC; j1++)
    for (int ka = 0; ka < PL::RA; ka++)
      for (int j2 = 0; j2 < PL::RB; j2++)
        t3[(j1 * PL::RA + ka) * PL::RB + j2] =
            (double)pow_mod_host(ws, (uint64_t)ka * (j1 + PL::RC * j2) % PL::S);
}

// A twiddle w read as a double with its quotient w/p formed by one multiply.
struct TwW {
  const double* p;
  ZHD Tw2 operator()(int i) const {
    const double w = p[i];
    return Tw2{w, w * PINV};
  }
};

// Thread coordinates: wave w, half h, column col (owner kA = 2 w + h), and
// for the 16-byte accesses column quad q and row g.
struct Tid {
  int w, h, col, q, g;
  ZHD explicit Tid(int tid)
      : w(tid >> 6), h((tid >> 5) & 1), col(tid & 31), q(tid & 7), g((tid >> 3) & 7) {}
  ZHD int ka() const { return 2 * w + h; }
};

// Byte offsets of the 16-byte accesses (raw-buffer form: a per-thread
// voffset plus a uniform soffset, from the tile's first column).  Access i of
// step sp reads input row 2 sp + s + RC g + RC RB a (slice s = w / 4, a =
// (4 w + i) mod 16, nB = g), four columns 4q..4q+3; the even output row is
// twice that; the odd row for P1' value (j1 = 2 sp + s, j2 = g, j3 = a) is
// 2 (2 sp + s + RC g + RC RB a) + 1.
template <class PL>
struct Offs {
  uint32_t in, ev, od;
  ZHD Offs(const Tid& t, uint32_t cols) {
    const uint32_t r = (uint32_t)(PL::RC * t.g);
    in = (r * cols + 4u * (uint32_t)t.q) * 4u;
    ev = (2 * r * cols + 4u * (uint32_t)t.q) * 4u;
    od = ((2 * r + 1) * cols + 4u * (uint32_t)t.q) * 4u;
  }
  static ZHD int slice_of(int w) { return w / (PL::WAVES / PL::SPS); }
  static ZHD int a_of(int w, int i) { return (PL::QI * w + i) % PL::RA; }
  // c4 = cols * 4, the byte stride of a row (the kernel hides its value per
  // tile, so these uniform offsets are computed where they are used instead
  // of being hoisted out of the tile loop into scalar registers that spill)
  static ZHD uint32_t s_in(int sp, int w, int i, uint32_t c4) {
    return (uint32_t)(PL::SPS * sp + slice_of(w) + PL::RC * PL::RB * a_of(w, i)) * c4;
  }
  static ZHD uint32_t s_out(int sp, int w, int i, uint32_t c4) { return 2u * s_in(sp, w, i, c4); }
  // staging index (IN / ODD) of access i's four columns
  static ZHD int lds(const Tid& t, int i) { return PL::si(slice_of(t.w), a_of(t.w, i), t.g, 4 * t.q); }
};

// The 16-byte loads of one step's input rows.
template <class PL>
ZHD void load_step(const GBuf& ib, const Offs<PL>& o, int w, int sp, uint32_t c4, U4 (&r)[PL::QI]) {
#pragma unroll
  for (int i = 0; i < PL::QI; i++) r[i] = ib.load4(o.in, Offs<PL>::s_in(sp, w, i, c4));
}

// A step's loaded rows into the IN staging and, with EVEN, to the even
// output rows.
template <class PL, bool EVEN = true>
ZHD void stage_step(const Tid& t, const GBuf& ob, const Offs<PL>& o, int sp, uint32_t c4,
                    const U4 (&r)[PL::QI], int32_t* __restrict__ lds) {
#pragma unroll
  for (int i = 0; i < PL::QI; i++) {
    if constexpr (EVEN) ob.store4(r[i], o.ev, Offs<PL>::s_out(sp, t.w, i, c4));
    *reinterpret_cast<U4*>(lds + PL::IN_OFF + Offs<PL>::lds(t, i)) = r[i];
  }
}

// The even output rows of step sp: the input rows read again (16-byte
// accesses) and written back out.
template <class PL>
ZHD void copy_even(const GBuf& ib, const GBuf& ob, const Offs<PL>& o, int w, int sp, uint32_t c4) {
  U4 r[PL::QI];
#pragma unroll
  for (int i = 0; i < PL::QI; i++) r[i] = ib.load4(o.in, Offs<PL>::s_in(sp, w, i, c4));
#pragma unroll
  for (int i = 0; i < PL::QI; i++) ob.store4(r[i], o.ev, Offs<PL>::s_out(sp, w, i, c4));
}

// The odd rows P1' left in the ODD staging, as 16-byte stores.
template <class PL>
ZHD void store_odd(const Tid& t, const GBuf& ob, const Offs<PL>& o, int sp, uint32_t c4,
                   const int32_t* __restrict__ lds) {
  U4 v[PL::QI];
#pragma unroll
  for (int i = 0; i < PL::QI; i++) v[i] = *reinterpret_cast<const U4*>(lds + PL::ODD_OFF + Offs<PL>::lds(t, i));
#pragma unroll
  for (int i = 0; i < PL::QI; i++) ob.store4(v[i], o.od, Offs<PL>::s_out(sp, t.w, i, c4));
}

// P1: INTT over nA of the thread's (nB = w, c = 2 sp + h) values (IN) -> A.
template <class PL>
ZHD void p1(const Tid& t, int32_t* __restrict__ lds, const Tw2* __restrict__ t1w, const MulK& k) {
  group_fence();
  const uint32_t* in = reinterpret_cast<const uint32_t*>(lds + PL::IN_OFF);
  double x[PL::RA];
#pragma unroll
  for (int a = 0; a < PL::RA; a++) x[a] = (double)in[PL::si(t.h, a, t.w, t.col)];
  dft<PL::LRA, true>(x, k);
  int32_t r[PL::RA];
  mul_mem_batch(x, [&](int i) -> const Tw2& { return t1w[i]; }, r, k);
#pragma unroll
  for (int ka = 0; ka < PL::RA; ka++) lds[PL::A_OFF + PL::xi(t.h, t.w, ka, t.col)] = r[ka];
}

// P2: INTT over nB for the owner's kA and the two slices c = 2 SP + s.
template <class PL, int SP>
ZHD void p2(const Tid& t, const int32_t* __restrict__ lds, const double* __restrict__ t2,
            int32_t (&held)[PL::RB][PL::RC], const MulK& k) {
  const int ka = t.ka();
#pragma unroll
  for (int s = 0; s < PL::SPS; s++) {
    group_fence();
    constexpr int c0 = PL::SPS * SP;
    double x[PL::RB];
#pragma unroll
    for (int nb = 0; nb < PL::RB; nb++) x[nb] = (double)lds[PL::A_OFF + PL::xi(s, nb, ka, t.col)];
    dft<PL::LRB, true>(x, k);
    int32_t r[PL::RB];
    mul_mem_batch(x, TwW{t2 + ((c0 + s) * PL::RA + ka) * PL::RB}, r, k);
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
ZHD void p2b(const Tid& t, const int32_t (&held)[PL::RB][PL::RC], int32_t* __restrict__ lds,
             const double* __restrict__ t3, const MulK& k) {
  const int ka = t.ka();
#pragma unroll
  for (int s = 0; s < PL::SPS; s++) {
    group_fence();
    constexpr int j0 = PL::SPS * SP;
    double x[PL::RB];
#pragma unroll
    for (int kb = 0; kb < PL::RB; kb++) x[kb] = (double)held[kb][j0 + s];
    dft<PL::LRB, false>(x, k);
    int32_t r[PL::RB];
    mul_mem_batch(x, TwW{t3 + ((j0 + s) * PL::RA + ka) * PL::RB}, r, k);
#pragma unroll
    for (int j2 = 0; j2 < PL::RB; j2++) lds[PL::B_OFF + PL::xi(s, j2, ka, t.col)] = r[j2];
  }
}

// P1': NTT over kA of the thread's (j2 = w, j1 = 2 sp + h) values (B),
// canonical -> the ODD staging (odd rows 2 (j1 + RC j2 + RC RB j3) + 1).
template <class PL>
ZHD void p1b(const Tid& t, int32_t* __restrict__ lds, const MulK& k) {
  group_fence();
  double x[PL::RA];
#pragma unroll
  for (int ka = 0; ka < PL::RA; ka++) x[ka] = (double)lds[PL::B_OFF + PL::xi(t.h, t.w, ka, t.col)];
  dft<PL::LRA, false>(x, k);
  uint32_t* odd = reinterpret_cast<uint32_t*>(lds + PL::ODD_OFF);
#pragma unroll
  for (int j3 = 0; j3 < PL::RA; j3++) odd[PL::si(t.h, j3, t.w, t.col)] = canon(x[j3], k);
}

}  // namespace slice
}  // namespace zoda
