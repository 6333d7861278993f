// CPU emulation of the fused encode kernel -- TEST INFRASTRUCTURE ONLY.
//
// Runs the exact per-thread code of zoda-cu_amd/csrc/encode_core.hpp on the
// host: for every tile, each step is executed for every thread id before the
// next step starts (the work-group barrier), with LDS modelled by a vector.
// The tests compare its output with the oracle to check the kernel's index and
// twiddle arithmetic on the CPU, before and besides the GPU parity tests.  The
// product library does not contain or call this code.
#include <stdint.h>

#include <type_traits>
#include <vector>

#include "../../zoda-cu_amd/csrc/encode_plans.hpp"
#include "../../zoda-cu_amd/csrc/sha256.hpp"
#include "../../zoda-cu_amd/csrc/encode_tables.hpp"
#include "../../zoda-cu_amd/csrc/decode_core.hpp"
#include "../../zoda-cu_amd/csrc/slice_core.hpp"
#include "../../zoda-cu_amd/csrc/ntt_fp64_core.hpp"
#include "../../zoda-cu_amd/csrc/encode_cols_core.hpp"

#include <array>

using namespace zoda;

template <class PL, int STEP, bool FULL>
static void run_steps(uint32_t tile, const uint32_t* in, uint32_t* out, uint32_t krows,
                      uint32_t cols, int layout, int32_t* lds, const EncodeTabs& T) {
  if constexpr (STEP == 0)
    for (int tid = 0; tid < PL::NT; tid++) encode_load_tables<PL>(tid, lds, T);
  for (int tid = 0; tid < PL::NT; tid++)
    encode_step<PL, STEP, FULL>(tid, tile, in, out, krows, cols, layout, lds, T);
  if constexpr (STEP + 1 < PL::STEPS)
    run_steps<PL, STEP + 1, FULL>(tile, in, out, krows, cols, layout, lds, T);
}

template <class PL>
static int sim_plan(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out, int layout) {
  EncodeTabOffsets off;
  std::vector<Tw> tabs = build_encode_tables(PL::L, PL::R1, PL::R2, PL::RD, &off);
  PackedTables pk = pack_u32(tabs, off);
  EncodeTabs T = make_encode_tabs(tabs.data(), off, pk.words.data(), pk);
  if (PL::TWL && pk.words.size() > (size_t)PL::TW_WORDS) return -2;
  std::vector<int32_t> lds(PL::LDS_WORDS ? PL::LDS_WORDS : 1);
  const uint32_t ntiles = (cols + PL::W - 1) / PL::W;
  const bool full = cols % PL::W == 0 && k == (uint32_t)PL::S;
  for (uint32_t tile = 0; tile < ntiles; tile++) {
    if (full)
      run_steps<PL, 0, true>(tile, in, out, k, cols, layout, lds.data(), T);
    else
      run_steps<PL, 0, false>(tile, in, out, k, cols, layout, lds.data(), T);
  }
  return 0;
}

template <class PL, bool BUF, int STEP>
static void run_late_steps(uint32_t tile, const uint32_t* in, uint32_t* out, uint32_t cols,
                           int32_t* lds, const EncodeTabs& T) {
  for (int tid = 0; tid < PL::NT; tid++)
    encode_step<PL, STEP, true, BUF>(tid, tile, in, out, (uint32_t)PL::S, cols, 0, lds, T);
  if constexpr (STEP + 1 < PL::STEPS) run_late_steps<PL, BUF, STEP + 1>(tile, in, out, cols, lds, T);
}

// The pipelined persistent kernel's order of work on one work-group
// (encode_pipelined): the first pass from the registers loaded ahead with the
// even-row stores of those registers, the middle passes with the loads of the
// next tile into the same registers in parts, the last pass.
template <class PL, bool BUF, int STEP>
static void run_step_loading(uint32_t tile, const uint32_t* in, uint32_t* out, uint32_t cols, int32_t* lds,
                             const EncodeTabs& T, bool more, uint32_t next, int part0, int np,
                             std::vector<uint32_t>* raw, int words) {
  using PP = Pipe<PL, BUF>;
  for (int tid = 0; tid < PL::NT; tid++)
    encode_step<PL, STEP, true, BUF>(tid, tile, in, out, (uint32_t)PL::S, cols, 0, lds, T, [&](int gi) {
      if (more)
        PP::load_part(tid, next, in, cols,
                      *reinterpret_cast<uint32_t(*)[PP::G][PP::RR]>(raw->data() + (size_t)tid * words), part0 + gi,
                      np);
    });
}

// BUF as the launch chooses it: raw-buffer accesses when the output is below 4 GiB.
template <class PL, bool BUF>
static int sim_pipelined(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out) {
  using PP = Pipe<PL, BUF>;
  if (cols % PL::W != 0 || k != (uint32_t)PL::S) return -3;
  EncodeTabOffsets off;
  std::vector<Tw> tabs = build_encode_tables(PL::L, PL::R1, PL::R2, PL::RD, &off);
  PackedTables pk = pack_u32(tabs, off);
  EncodeTabs T = make_encode_tabs(tabs.data(), off, pk.words.data(), pk);
  if (pk.words.size() > (size_t)PL::TW_WORDS) return -2;
  std::vector<int32_t> lds(PL::LDS_WORDS);
  for (int tid = 0; tid < PL::NT; tid++) encode_load_tables<PL>(tid, lds.data(), T);
  const uint32_t* sb = reinterpret_cast<const uint32_t*>(lds.data() + PL::TILE_WORDS);
  const TwTab inv1{T.inv1, sb + T.o_inv1};
  constexpr int WORDS = PP::G * PP::RR;
  using Raw = uint32_t[PP::G][PP::RR];
  std::vector<uint32_t> regs((size_t)PL::NT * WORDS);
  auto raw = [&](int tid) -> Raw& { return *reinterpret_cast<Raw*>(regs.data() + (size_t)tid * WORDS); };
  const uint32_t ntiles = cols / PL::W;
  for (int tid = 0; tid < PL::NT; tid++) PP::load(tid, 0, in, cols, raw(tid));
  static_assert(PL::D >= 2, "the pipelined kernel has two or three digits");
  // load parts over the middle passes, as the kernel counts them
  constexpr int G1 = (PL::W << (PL::L - (PL::D == 2 ? PL::RD : PL::R2))) / PL::NT;
  constexpr int G2 = PL::D == 3 ? (PL::W << (PL::L - PL::RD)) / PL::NT : 0;
  constexpr int NP = PL::D == 3 ? 2 * G1 + G2 : G1;
  for (uint32_t tile = 0; tile < ntiles; tile++) {
    const bool more = tile + 1 < ntiles;
    for (int tid = 0; tid < PL::NT; tid++)
      PP::first(tid, raw(tid), lds.data(), T, inv1,
                [&](int gi) { PP::store_even(tid, gi, tile, out, cols, raw(tid)[gi]); });
    run_step_loading<PL, BUF, 1>(tile, in, out, cols, lds.data(), T, more, tile + 1, 0, NP, &regs, WORDS);
    if constexpr (PL::D == 3) {
      run_step_loading<PL, BUF, 2>(tile, in, out, cols, lds.data(), T, more, tile + 1, G1, NP, &regs, WORDS);
      run_step_loading<PL, BUF, 3>(tile, in, out, cols, lds.data(), T, more, tile + 1, G1 + G2, NP, &regs, WORDS);
      run_late_steps<PL, BUF, 4>(tile, in, out, cols, lds.data(), T);
    } else {
      run_late_steps<PL, BUF, 2>(tile, in, out, cols, lds.data(), T);
    }
  }
  return 0;
}

// Persistent (pipelined) form of the single-square plan of side k (whole
// tiles only).
extern "C" int zoda_sim_encode_persistent(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out) {
  uint32_t nk = 1, L = 0;
  while (nk < k) { nk <<= 1; L++; }
  return with_encode_plan((int)L, [&](auto plan) -> int {
    using PL = decltype(plan);
    if constexpr (PL::PERSIST) {
      if (((uint64_t)(2 * PL::S - 1) * cols + PL::W) * 4 < (1ull << 32)) return sim_pipelined<PL, true>(in, k, cols, out);
      return sim_pipelined<PL, false>(in, k, cols, out);
    }
    else return -4;
  });
}

// The slice kernel (encode_slice) on one work-group: tile `it` runs its
// front steps (P1, P2) beside the back steps (P2', P1') of tile it - 1, step
// by step with the same order of the two phases as the kernel, so the owner
// registers change hands slot by slot and the LDS staging of the input rows
// and the odd rows is read and written exactly as on the device; the middle
// passes run between.  16-byte accesses, zero-record buffers for inactive
// parts, and the loads one step ahead in per-thread registers, as on the GPU.
namespace {
template <class PL>
struct SliceSim {
  using Held = int32_t[PL::RB][PL::RC];
  using B = std::conditional_t<PL::WIDE, slice::TileRef, GBuf>;
  using Tid = slice::Tid<PL>;
  const uint32_t* in;
  uint32_t* out;
  uint32_t cols;
  std::vector<typename PL::T1T> t1;
  std::vector<typename PL::TwT> t2, t3;
  std::vector<int32_t> lds;
  std::vector<std::array<int32_t, PL::RB * PL::RC>> held;
  std::vector<std::array<U4, PL::QI>> ld;
  slice::MulK k;
  Held& h(int tid) { return *reinterpret_cast<Held*>(held[tid].data()); }
  U4 (&l(int tid))[PL::QI] { return *reinterpret_cast<U4(*)[PL::QI]>(ld[tid].data()); }
  const typename PL::T1T* t1w(int w) const { return t1.data() + w * (PL::T1_PAIRS / PL::RB); }

  struct Bufs {
    B ibf, obf, ibn, obn, obb, obp;
  };
  static constexpr bool kLastOddInMid = PL::EVEN_MID_MASK != 0;

  template <int SP>
  void step(const Bufs& b, bool front, bool back) {
    const uint32_t c4 = cols * 4;
    for (int tid = 0; tid < PL::NT; tid++) {
      const Tid t(tid);
      const slice::Offs<PL> o(t, cols);
      // the previous iteration's last odd rows: here, or (plans with middle-pass
      // copies) in middle group 1, as on the device
      if (SP > 0)
        slice::store_odd<PL>(t, b.obb, o, SP - 1, c4, lds.data());
      else if (!kLastOddInMid)
        slice::store_odd<PL>(t, b.obp, o, PL::STEPS - 1, c4, lds.data());
      if (front) slice::p1<PL>(t, lds.data(), t1w(t.w), k);
      if (back) slice::p2b<PL, SP>(t, h(tid), lds.data(), t3.data(), k);
    }
    for (int tid = 0; tid < PL::NT; tid++) {
      const Tid t(tid);
      const slice::Offs<PL> o(t, cols);
      if (front) slice::p2<PL, SP>(t, lds.data(), t2.data(), h(tid), k);
      if (back) slice::p1b<PL>(t, lds.data(), k);
      if (SP + 1 < PL::STEPS)
        slice::stage_step<PL, !PL::even_in_mid(SP + 1)>(t, b.obf, o, SP + 1, c4, l(tid), lds.data());
      else
        slice::stage_step<PL, !PL::even_in_mid(0)>(t, b.obn, o, 0, c4, l(tid), lds.data());
      if (SP + 2 < PL::STEPS)
        slice::load_step<PL>(b.ibf, o, t.w, SP + 2, c4, l(tid));
      else
        slice::load_step<PL>(b.ibn, o, t.w, SP + 2 - PL::STEPS, c4, l(tid));
    }
    if constexpr (SP + 1 < PL::STEPS) step<SP + 1>(b, front, back);
  }

  // the middle passes; group KB loads step KB's rows when its bit of
  // EVEN_MID_MASK is set and group KB + 1 stores them as even rows, as on the
  // device
  template <int KB>
  void mid(int tid, const Bufs& b, U4 (&ev)[PL::QI]) {
    const Tid t(tid);
    const slice::Offs<PL> o(t, cols);
    if constexpr (PL::even_in_mid(KB)) slice::load_step<PL>(b.ibf, o, t.w, KB, cols * 4, ev);
    slice::mid_group<PL, KB>(h(tid), k);
    if constexpr (KB > 0 && PL::even_in_mid(KB - 1)) slice::store_even<PL>(b.obf, o, t.w, KB - 1, cols * 4, ev);
    if constexpr (KB == 1 && kLastOddInMid) slice::store_odd<PL>(t, b.obb, o, PL::STEPS - 1, cols * 4, lds.data());
    if constexpr (KB + 1 < PL::RB) mid<KB + 1>(tid, b, ev);
  }

  B gb(const uint32_t* base, uint32_t tile, bool live) const {
    if constexpr (PL::WIDE)
      return slice::TileRef{base + (size_t)tile * PL::W, live};
    else
      return GBuf(base + (size_t)tile * PL::W, live);
  }

  void run() {
    const uint32_t ntiles = cols / PL::W, c4 = cols * 4;
    for (int tid = 0; tid < PL::NT; tid++) {
      const Tid t(tid);
      const slice::Offs<PL> o(t, cols);
      slice::load_step<PL>(gb(in, 0, true), o, t.w, 0, c4, l(tid));
      slice::stage_step<PL, !PL::even_in_mid(0)>(t, gb(out, 0, true), o, 0, c4, l(tid), lds.data());
      slice::load_step<PL>(gb(in, 0, true), o, t.w, 1, c4, l(tid));
    }
    uint32_t tp = 0;
    for (uint32_t it = 0; it <= ntiles; it++) {
      const bool front = it < ntiles, back = it > 0, next = it + 1 < ntiles, prev = it > 1;
      const uint32_t tf = front ? it : 0, tb = back ? it - 1 : 0, tn = next ? it + 1 : 0;
      const Bufs b{gb(in, tf, front), gb(out, tf, front), gb(in, tn, next), gb(out, tn, next),
                   gb(out, tb, back), gb(out, tp, prev)};
      step<0>(b, front, back);
      if (front)
        for (int tid = 0; tid < PL::NT; tid++) {
          U4 ev[PL::QI];
          mid<0>(tid, b, ev);
        }
      tp = tb;
    }
    for (int tid = 0; tid < PL::NT; tid++) {
      const Tid t(tid);
      slice::store_odd<PL>(t, gb(out, tp, true), slice::Offs<PL>(t, cols), PL::STEPS - 1, c4, lds.data());
    }
  }
};

template <class PL>
int sim_slice(const uint32_t* in, uint32_t cols, uint32_t* out) {
  if (cols == 0 || cols % PL::W != 0) return -3;
  if (!PL::WIDE && ((uint64_t)(2 * PL::S - 1) * cols + PL::W) * 4 >= (1ull << 32)) return -3;
  SliceSim<PL> sim;
  sim.in = in;
  sim.out = out;
  sim.cols = cols;
  sim.t1.resize(PL::T1_PAIRS);
  sim.t2.resize(PL::T2_WORDS);
  sim.t3.resize(PL::T3_WORDS);
  slice::build_slice_tables<PL>(sim.t1.data(), sim.t2.data(), sim.t3.data());
  sim.lds.assign(4 * PL::XWORDS, 0);
  sim.held.assign(PL::NT, {});
  sim.ld.assign(PL::NT, {});
  sim.k = slice::make_mulk();
  sim.run();
  return 0;
}
}  // namespace

// The slice kernel of side k (2048 or 4096) run on the host.
extern "C" int zoda_sim_encode_slice(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out) {
  if (k == (uint32_t)slice::Plan2048::S) return sim_slice<slice::Plan2048>(in, cols, out);
  if (k == (uint32_t)slice::Plan4096::S) return sim_slice<slice::Plan4096>(in, cols, out);
  return -3;
}

// The FP64 poly transform (ntt_poly_fp64) run on the host: every tile's load,
// passes and store for all 256 threads, phase by phase, with the kernel's LDS
// layout (padded data, then the twiddle tables).
template <int L, bool INV>
static int sim_ntt_fp64_plan(uint32_t* values, uint32_t polys, uint32_t stride) {
  using PL = nttf::Plan<L>;
  std::vector<int32_t> lds(PL::LDS_WORDS, 0);
  uint32_t* tw = reinterpret_cast<uint32_t*>(lds.data() + nttf::LDS_DATA);
  std::vector<uint32_t> tab(std::max(1, PL::TW_WORDS));
  nttf::build_ntt_fp64_tables<PL>(INV, tab.data());
  for (int i = 0; i < PL::TW_WORDS; i++) tw[i] = tab[i];
  const slice::MulK k = slice::make_mulk();
  const uint32_t ntiles = (polys + PL::POLYS - 1) / PL::POLYS;
  for (uint32_t t = 0; t < ntiles; t++) {
    const uint32_t first = t * (uint32_t)PL::POLYS;
    const nttf::TileSpan ts{std::min((uint32_t)PL::POLYS, polys - first), (size_t)first};
    for (int tid = 0; tid < nttf::NT; tid++)
      nttf::load_tile<PL>(tid, ts, lds.data(), [&](size_t poly, uint32_t j) -> U4 {
        const uint32_t* q = values + poly * stride + j;
        return U4{q[0], q[1], q[2], q[3]};
      });
    for (int tid = 0; tid < nttf::NT; tid++) nttf::pass<PL, 0, INV>(tid, lds.data(), tw, k);
    if constexpr (PL::P > 1)
      for (int tid = 0; tid < nttf::NT; tid++) nttf::pass<PL, 1, INV>(tid, lds.data(), tw, k);
    if constexpr (PL::P > 2)
      for (int tid = 0; tid < nttf::NT; tid++) nttf::pass<PL, 2, INV>(tid, lds.data(), tw, k);
    for (int tid = 0; tid < nttf::NT; tid++)
      nttf::store_tile<PL>(tid, ts, lds.data(), [&](size_t poly, uint32_t kq, const U4& v) {
        uint32_t* q = values + poly * stride + kq;
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
        q[3] = v.w;
      });
  }
  return 0;
}

template <int L>
static int sim_ntt_fp64_size(uint32_t* values, uint32_t log_n, uint32_t polys, uint32_t stride, int inverse) {
  if constexpr (L <= 12) {
    if (log_n == (uint32_t)L)
      return inverse ? sim_ntt_fp64_plan<L, true>(values, polys, stride)
                     : sim_ntt_fp64_plan<L, false>(values, polys, stride);
    return sim_ntt_fp64_size<L + 1>(values, log_n, polys, stride, inverse);
  } else {
    return -1;
  }
}

// In place on host memory, the layout of the drop-in transforms (poly c at
// values[c * stride], n = 2^log_n of it transformed); log_n in 2..12, stride a
// multiple of 4.
extern "C" int zoda_sim_ntt_fp64(uint32_t* values, uint32_t log_n, uint32_t polys, uint32_t stride,
                                 int inverse) {
  if (log_n < 2 || log_n > 12 || stride % 4 != 0 || stride < (1u << log_n)) return -1;
  return sim_ntt_fp64_size<2>(values, log_n, polys, stride, inverse);
}

// ---- column-major encode (encode_cols_core.hpp), phase by phase ----------------

// The device's schedule (encode_cols.hip): chunks of `chunk` columns, launch i
// transposing chunk i block by block into one buffer and encoding chunk
// i - 1 tile by tile from the other.
template <int L>
static int sim_encode_cols_plan(const uint32_t* in, uint32_t cols, uint32_t* out, uint32_t chunk) {
  using PL = ecol::Plan<L>;
  std::vector<int32_t> lds(PL::LDS_WORDS, 0);
  uint32_t* tw = reinterpret_cast<uint32_t*>(lds.data() + nttf::LDS_DATA);
  uint32_t* sh = reinterpret_cast<uint32_t*>(lds.data());
  std::vector<uint32_t> tab(PL::TW_WORDS);
  ecol::build_col_tables<PL>(tab.data());
  for (int i = 0; i < PL::TW_WORDS; i++) tw[i] = tab[i];
  const slice::MulK k = slice::make_mulk();
  const uint32_t nchunks = (cols + chunk - 1) / chunk;
  std::vector<uint32_t> buf[2] = {std::vector<uint32_t>((size_t)PL::S * chunk),
                                  std::vector<uint32_t>((size_t)PL::S * chunk)};
  std::vector<ecol::U2> held((size_t)nttf::NT * ecol::PAIRS);
  for (uint32_t i = 0; i <= nchunks; i++) {
    // the transpose of chunk i
    if (i < nchunks) {
      const uint32_t c0 = i * chunk, tcols = std::min(chunk, cols - c0);
      if (tcols % ecol::TB) return -1;
      const uint32_t nblocks = (PL::S / ecol::TB) * (tcols / ecol::TB);
      std::vector<uint32_t>& t = buf[i % 2];
      for (uint32_t bi = 0; bi < nblocks; bi++) {
        const ecol::TBlock tb = ecol::tblock(bi, PL::S);
        const uint32_t* src = in + (size_t)tb.r0 * cols + c0 + tb.c0;
        for (int tid = 0; tid < nttf::NT; tid++)
          for (int q = 0; q < ecol::TB_RUNS; q++) {
            const uint32_t* p = src + ecol::tb_in_offset(tid, q, cols);
            ecol::tb_put(tid, q, sh, U4{p[0], p[1], p[2], p[3]});
          }
        uint32_t* dst = t.data() + (size_t)tb.c0 * PL::S + tb.r0;
        for (int tid = 0; tid < nttf::NT; tid++)
          for (int q = 0; q < ecol::TB_RUNS; q++) {
            const U4 v = ecol::tb_get(tid, q, sh);
            uint32_t* p = dst + ecol::tb_out_offset(tid, q, PL::S);
            p[0] = v.x;
            p[1] = v.y;
            p[2] = v.z;
            p[3] = v.w;
          }
      }
    }
    // the column tiles of chunk i - 1
    if (i > 0) {
      const uint32_t c0 = (i - 1) * chunk, ccols = std::min(chunk, cols - c0);
      const std::vector<uint32_t>& t = buf[(i - 1) % 2];
      uint32_t* o = out + (size_t)c0 * 2 * PL::S;
      const uint32_t ntiles = (ccols + PL::COLS - 1) / PL::COLS;
      for (uint32_t tile = 0; tile < ntiles; tile++) {
        const size_t col0 = (size_t)tile * PL::COLS;
        const uint32_t live = std::min((uint32_t)PL::COLS, ccols - (uint32_t)col0);
        for (int tid = 0; tid < nttf::NT; tid++)
          for (int q = 0; q < ecol::PAIRS; q++) {
            const ecol::U2 x = ecol::tile_pair<PL>(tid, q, live, col0, [&](size_t c, uint32_t j) -> ecol::U2 {
              const uint32_t* p = t.data() + c * PL::S + j;
              return ecol::U2{p[0], p[1]};
            });
            held[(size_t)tid * ecol::PAIRS + q] = x;
            ecol::put_pair(tid, q, lds.data(), x);
          }
        ecol::all_passes<PL>(
            [&](auto pass) {
              for (int tid = 0; tid < nttf::NT; tid++) pass(tid, lds.data(), tw, k);
            },
            [] {});
        for (int tid = 0; tid < nttf::NT; tid++) {
          auto x = [&](int q) { return held[(size_t)tid * ecol::PAIRS + q]; };
          ecol::store_pairs<PL>(tid, live, col0, lds.data(), x, [&](int, size_t c, uint32_t r, const U4& v) {
            uint32_t* p = o + c * 2 * PL::S + r;
            p[0] = v.x;
            p[1] = v.y;
            p[2] = v.z;
            p[3] = v.w;
          });
        }
      }
    }
  }
  return 0;
}

template <int L>
static int sim_encode_cols_size(const uint32_t* in, uint32_t log_k, uint32_t cols, uint32_t* out, uint32_t chunk) {
  if constexpr (L <= 12) {
    if (log_k == (uint32_t)L) return sim_encode_cols_plan<L>(in, cols, out, chunk);
    return sim_encode_cols_size<L + 1>(in, log_k, cols, out, chunk);
  } else {
    return -1;
  }
}

// The column-major encode of a whole k x cols square (k = 2^log_k, 6..12,
// cols a multiple of 64): out[c * 2k + r], as the device's transposes and
// column kernel compute it, in chunks of `chunk` columns (a multiple of 64).
extern "C" int zoda_sim_encode_cols(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out, uint32_t chunk) {
  uint32_t L = 0;
  while ((1u << L) < k) L++;
  if ((1u << L) != k || L < 6 || L > 12 || cols == 0 || cols % ecol::TB || chunk == 0 || chunk % ecol::TB) return -1;
  return sim_encode_cols_size<6>(in, L, cols, out, chunk);
}

// variant < 0: the single-square plan of the side; 0..3: the batched plans.
extern "C" int zoda_sim_encode_variant(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t n,
                                       uint32_t* out, int layout, int variant) {
  uint32_t nk = 1, L = 0;
  while (nk < k) { nk <<= 1; L++; }
  if (k == 0 || cols == 0 || k + n <= nk || k + n > 2 * nk) return -1;
  if (variant < 0)
    return with_encode_plan((int)L, [&](auto plan) -> int {
      return sim_plan<decltype(plan)>(in, k, cols, out, layout);
    });
  return with_batch_plan((int)L, variant, [&](auto plan) -> int {
    return sim_plan<decltype(plan)>(in, k, cols, out, layout);
  });
}

extern "C" int zoda_sim_encode(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t n,
                               uint32_t* out, int layout) {
  return zoda_sim_encode_variant(in, k, cols, n, out, layout, -1);
}

// The device code's constant multiplication, run on the host: out[i] =
// mulmod(x[i], w[i], wq) with wq = w/p as the tables hold it (tables != 0) or
// w * (1/p) as the LDS twiddle reads recompute it.
extern "C" void zoda_sim_mulmod(const double* x, const double* w, size_t count, int tables, double* out) {
  for (size_t i = 0; i < count; i++) out[i] = mulmod(x[i], w[i], tables ? w[i] / PD : w[i] * PINV);
}

// The device code's reduction to canonical form, run on the host, for each of
// `count` integer-valued doubles.
extern "C" void zoda_sim_canonical(const double* x, size_t count, uint32_t* out) {
  for (size_t i = 0; i < count; i++) out[i] = canonical(x[i]);
}

// Host runs of the device code's SHA-256 and RLC coefficient derivation.
extern "C" void zoda_sim_sha256(const uint8_t* data, size_t len, uint8_t* out) {
  Sha256 h;
  h.init();
  h.update(data, len);
  h.final(out);
}

// As the device derives it: roots of whole 64-byte blocks from the shared
// state after them (one compression per column), others hashed whole.
extern "C" void zoda_sim_rlc_coefficient(const uint8_t* root, uint32_t len, uint64_t col, uint32_t* out) {
  if (len % 64 == 0) {
    Sha256 h;
    h.init();
    h.update(root, len);
    rlc_coefficient_mid(h.h, (uint64_t)len * 8 + 64, col, out);
  } else {
    rlc_coefficient(root, len, col, out);
  }
}

// ---- erasure decoder (decode_core.hpp) ------------------------------------------

template <class PL, int STEP>
static void dec_run_steps(int32_t* lds, const DecodeTabs& T, const uint32_t* srow) {
  const slice::MulK mk = slice::make_mulk();
  for (int tid = 0; tid < PL::NT; tid++) dec_step<PL, STEP>(tid, lds, T, srow, mk);
  if constexpr (STEP < 4) dec_run_steps<PL, STEP + 1>(lds, T, srow);
}

// The decoder's column kernel on the host: `enc` is the 2k x cols extended
// square (row-major), the first k distinct of `known_rows` are used, `out`
// receives the k x cols data.  The vanishing polynomial's values and the
// erased rows' factors are computed here directly from their definitions;
// the scratch holds the known rows by rank, as the GPU's gather leaves it,
// and the known data rows are copied as the gather copies them.  Returns -1
// for sides without an FP64 plan.
extern "C" int zoda_sim_decode(const uint32_t* enc, uint32_t k, uint32_t cols, const uint32_t* known_rows,
                               uint32_t nknown, uint32_t* out) {
  const uint32_t n = 2 * k;
  uint32_t L = 0;
  while ((1u << L) < n) L++;
  if ((1u << L) != n) return -1;
  std::vector<uint8_t> known(n, 0);
  uint32_t have = 0;
  for (uint32_t i = 0; i < nknown && have < k; i++)
    if (known_rows[i] < n && !known[known_rows[i]]) {
      known[known_rows[i]] = 1;
      have++;
    }
  if (have < k) return -2;
  const uint32_t w = root_of_unity(L);
  std::vector<uint32_t> x(n), erased;
  for (uint32_t r = 0, cur = 1; r < n; r++, cur = mul_mod_host(cur, w)) {
    x[r] = cur;
    if (!known[r]) erased.push_back(cur);
  }
  auto Z = [&](uint32_t pt) {
    uint32_t acc = 1;
    for (uint32_t e : erased) acc = mul_mod_host(acc, (pt + P - e) % P);
    return acc;
  };
  // srow[e] = 1 / (n w^e Z'(w^e)) on the erased even rows, Z'(w^e) the product
  // over the other erasures of (w^e - w^e'); 0 elsewhere.
  std::vector<uint32_t> zk(n, 0), srow(n, 0);
  for (uint32_t r = 0; r < n; r++) {
    if (known[r]) {
      zk[r] = Z(x[r]);
    } else if (r % 2 == 0) {
      uint32_t acc = mul_mod_host(n, x[r]);
      for (uint32_t e : erased)
        if (e != x[r]) acc = mul_mod_host(acc, (x[r] + P - e) % P);
      srow[r] = inv_mod_host(acc);
    }
  }
  // The rank-packed scratch: per column the k known rows (in increasing
  // order); `rows` lists them, then the erased rows.
  std::vector<uint32_t> rows;
  for (uint32_t r = 0; r < n; r++)
    if (known[r]) rows.push_back(r);
  for (uint32_t r = 0; r < n; r++)
    if (!known[r]) rows.push_back(r);
  std::vector<uint32_t> T((size_t)cols * k);
  for (uint32_t i = 0; i < k; i++)
    for (uint32_t c = 0; c < cols; c++)
      T[(size_t)c * k + i] = mul_mod_host(enc[(size_t)rows[i] * cols + c], zk[rows[i]]);
  const int rc = with_decode_plan((int)L, [&](auto plan) -> int {
    using PL = decltype(plan);
    const DecodeTableImage im = build_decode_tables<PL>();
    if (im.packed.empty()) return -3;
    const DecodeTabs tabs = make_decode_tabs(im, im.packed.data());
    std::vector<int32_t> lds(PL::LDS_WORDS);
    for (int tid = 0; tid < PL::NT; tid++) dec_load_tables<PL>(tid, lds.data(), tabs);
    const uint32_t ntiles = (cols + PL::W - 1) / PL::W;
    for (uint32_t tile = 0; tile < ntiles; tile++) {
      for (int tid = 0; tid < PL::NT; tid++) dec_load<PL>(tid, tile, cols, T.data(), rows.data(), lds.data());
      dec_run_steps<PL, 0>(lds.data(), tabs, srow.data());
      for (int tid = 0; tid < PL::NT; tid++) dec_store<PL>(tid, tile, cols, T.data(), lds.data());
    }
    return 0;
  });
  if (rc != 0) return rc;
  for (uint32_t j = 0; j < k; j++)
    for (uint32_t c = 0; c < cols; c++)
      out[(size_t)j * cols + c] = known[2 * j] ? enc[(size_t)2 * j * cols + c] : T[(size_t)c * k + j];
  return 0;
}
