// Slice-pipelined encode of sides 2048 and 4096 on gfx950 (see slice_core.hpp
// for the transform and the two plans): one persistent 512-thread work-group
// per CU, 32-column (side 2048) or 16-column (side 4096) tiles, the middle of
// each tile's transform held in registers.
//
// Per CU the tiles run as a pipeline: eight steps in which the front of tile
// i (P1, P2) and the back of tile i - 1 (P2', P1') share each step's LDS
// exchanges, then the middle passes of tile i, which touch no memory.  Every
// global access is a 16-byte raw-buffer access (eight lanes per 128-byte row
// segment): the step's input rows arrive in registers one step ahead, go out
// again unchanged as the even output rows and into an LDS staging buffer that
// P1 reads; P1' leaves its odd rows in a second staging buffer that the next
// step stores.  A dword access per lane would take four times as many
// vector-memory instructions, and the CU's address unit, busy with every
// row segment, then stalls the issuing waves.
#include <type_traits>

#include "common.hpp"
#include "slice_core.hpp"

namespace zoda {

namespace {

using slice::MulK;
using slice::Tw2;

// Every wave's LDS accesses are complete, then the work-group meets.  Only the
// LDS counter is waited for: global loads and stores stay in flight across
// the barrier (a __syncthreads() fence would wait for every store as well).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// A wave raises its issue priority while it hands its loads and stores to the
// address unit, so that the SIMD's arbiter lets it through ahead of the other
// wave's arithmetic and its memory work starts early (1 % over equal
// priorities).
__device__ __forceinline__ void mem_prio(bool high) {
  if (high)
    __builtin_amdgcn_s_setprio(3);
  else
    __builtin_amdgcn_s_setprio(0);
}

template <class PL>
struct Ctx {
  slice::Tid<PL> t;
  slice::Offs<PL> o;
  int wu;  // the wave index, uniform
  int32_t* lds;
  const typename PL::T1T* t1w;
  const typename PL::TwT* t2;
  const typename PL::TwT* t3;
  MulK k;
};

// Buffers of one step's memory work.  Every global load and store is
// unconditional: an inactive part (no front tile in the last iteration, no
// back tile in the first, no tile after the last) addresses a zero-record
// buffer instead (GBuf(p, false)), so no memory instruction sits behind a
// branch and the compiler's counts of outstanding memory operations stay
// exact -- a wave waits for the loads it needs, never for the stores issued
// after them (loads and stores share one in-order counter on gfx9).
//
// A plan whose offsets pass 32 bits (WIDE) keeps tile pointers instead, and
// each access builds its descriptor at its row (slice::TileRef).
template <class PL>
using BufT = std::conditional_t<PL::WIDE, slice::TileRef, GBuf>;

template <class PL>
__device__ __forceinline__ BufT<PL> tile_buf(const uint32_t* p, bool live) {
  if constexpr (PL::WIDE)
    return slice::TileRef{p, live};
  else
    return GBuf(p, live);
}

template <class PL>
struct Bufs {
  BufT<PL> ibf, obf;  // front tile: input, output
  BufT<PL> ibn, obn;  // the next front tile
  BufT<PL> obb;       // back tile: output
  BufT<PL> obp;       // the tile of the odd rows left in the staging by the previous step
};

// Step SP (two phases):
//   phase 1: P1 (IN -> A), P2' (held -> B), the previous step's odd rows out
//   phase 2: P2 (A -> held), P1' (B -> ODD); the next step's input rows (in
//            registers since one step) to IN and to the even output rows,
//            and the loads of the step after
// MF (memory first) puts each phase's memory work before its arithmetic
// instead of after it.  The two waves that share a SIMD (w and w + 4) take
// opposite orders, so that while one of them waits for the address unit to
// take its loads and stores, the other keeps the SIMD busy.
// The odd rows of a back tile's last step wait in the ODD staging through the
// middle passes of the front tile.  Plans whose middle passes copy even rows
// (EVM != 0) store them there, in middle group 1 (2048^2: 2.06-2.08 ms
// against 2.09-2.10 ms, alternating runs); the others in step 0 of the next
// iteration (at side 4096 the middle-pass store measured 0.5 % slower).
template <unsigned EVM>
constexpr bool kLastOddInMid = EVM != 0;

template <class PL, unsigned EVM, int SP, bool MF>
__device__ __forceinline__ void step(const Ctx<PL>& c, uint32_t c4, bool front, bool back, const Bufs<PL>& b,
                                     U4 (&ld)[PL::QI], int32_t (&held)[PL::RB][PL::RC]) {
  // odd rows of step SP - 1 (of the previous iteration's last step for SP = 0,
  // unless the middle passes stored them)
  auto odd = [&] {
    if constexpr (SP > 0)
      slice::store_odd<PL>(c.t, b.obb, c.o, SP - 1, c4, c.lds);
    else if constexpr (!kLastOddInMid<EVM>)
      slice::store_odd<PL>(c.t, b.obp, c.o, PL::STEPS - 1, c4, c.lds);
  };
  // rows of step SP + 1 (the next tile's step 0 after the last step) staged,
  // then the loads of step SP + 2
  auto stage = [&] {
    if constexpr (SP + 1 < PL::STEPS)
      slice::stage_step<PL, !((EVM >> (SP + 1)) & 1u)>(c.t, b.obf, c.o, SP + 1, c4, ld, c.lds);
    else
      slice::stage_step<PL, !(EVM & 1u)>(c.t, b.obn, c.o, 0, c4, ld, c.lds);
    if constexpr (SP + 2 < PL::STEPS)
      slice::load_step<PL>(b.ibf, c.o, c.wu, SP + 2, c4, ld);
    else
      slice::load_step<PL>(b.ibn, c.o, c.wu, SP + 2 - PL::STEPS, c4, ld);
  };
  if constexpr (MF) {
    mem_prio(true);
    odd();
    group_fence();
    mem_prio(false);
  }
  if (front) slice::p1<PL>(c.t, c.lds, c.t1w, c.k);
  if (back) slice::p2b<PL, SP>(c.t, held, c.lds, c.t3, c.k);
  if constexpr (!MF) {
    group_fence();
    mem_prio(true);
    odd();
    mem_prio(false);
  }
  lds_barrier();
  if constexpr (MF) {
    mem_prio(true);
    stage();
    group_fence();
    mem_prio(false);
  }
  if (front) slice::p2<PL, SP>(c.t, c.lds, c.t2, held, c.k);
  if (back) slice::p1b<PL>(c.t, c.lds, c.k);
  if constexpr (!MF) {
    group_fence();
    mem_prio(true);
    stage();
    mem_prio(false);
  }
  lds_barrier();
  if constexpr (SP + 1 < PL::STEPS) step<PL, EVM, SP + 1, MF>(c, c4, front, back, b, ld, held);
}

// The middle passes of the front tile.  Group KB loads step KB's input rows
// (16 bytes a lane) when bit KB of EVM is set, and group KB + 1 stores them as
// the even output rows, so that the CU keeps memory requests in flight
// through the passes that have no memory work of their own, and each load has
// two groups' arithmetic to arrive before its data is needed (1 % over
// storing in the same group).  With kLastOddInMid, group 1 also stores the
// back tile's last odd rows from the ODD staging (written by the last step's
// P1' before its barrier; the next write to ODD follows step 0's first
// barrier, which waits for these reads).
template <class PL, unsigned EVM, int KB>
__device__ __forceinline__ void mid(const Ctx<PL>& c, uint32_t c4, const Bufs<PL>& b,
                                    int32_t (&held)[PL::RB][PL::RC], U4 (&ev)[PL::QI]) {
  static_assert((EVM & (EVM >> 1)) == 0 && !(EVM >> (PL::RB - 1)), "copies one group apart, none in the last");
  constexpr bool LOAD = (EVM >> KB) & 1u;
  constexpr bool STORE = KB > 0 && ((EVM >> (KB - 1)) & 1u);
  if constexpr (LOAD) slice::load_step<PL>(b.ibf, c.o, c.wu, KB, c4, ev);
  slice::mid_group<PL, KB>(held, c.k);
  if constexpr (STORE) {
    group_fence();
    slice::store_even<PL>(b.obf, c.o, c.wu, KB - 1, c4, ev);
  }
  if constexpr (KB == 1 && kLastOddInMid<EVM>) {
    group_fence();
    slice::store_odd<PL>(c.t, b.obb, c.o, PL::STEPS - 1, c4, c.lds);
  }
  if constexpr (KB + 1 < PL::RB) mid<PL, EVM, KB + 1>(c, c4, b, held, ev);
}

// The waves of the upper half of the work-group run their steps memory first.
constexpr bool kSplitOrder = true;

// One wave's whole run over its work-group's tiles in step order MF: the
// prologue, the pipelined tiles and the epilogue.  Both orders are complete
// functions, so no register state crosses the branch between them (at such a
// join the two orders' register assignments would have to meet, and spill).
template <class PL, unsigned EVM, bool MF>
__device__ __forceinline__ void run_tiles(const uint32_t* in, uint32_t* out, uint32_t cols,
                                          uint32_t ntiles, const Ctx<PL>& c, int tid, int32_t* lds) {
  const uint32_t b = blockIdx.x, g = gridDim.x;
  const uint32_t mine = b < ntiles ? (ntiles - 1 - b) / g + 1 : 0;
  auto tile = [&](uint32_t i) { return xcd_tile(b + i * g, ntiles); };
  int32_t held[PL::RB][PL::RC];
#pragma unroll
  for (int i = 0; i < PL::RB; i++) {
#pragma unroll
    for (int j = 0; j < PL::RC; j++) held[i][j] = 0;
  }
  // prologue: step 0 of the first tile into IN and its even rows; step 1's loads
  U4 ld[PL::QI];
  {
    const bool live = mine > 0;
    const uint32_t t0 = live ? tile(0) : 0u;
    const BufT<PL> ib = tile_buf<PL>(in + (size_t)t0 * PL::W, live);
    const BufT<PL> ob = tile_buf<PL>(out + (size_t)t0 * PL::W, live);
    slice::load_step<PL>(ib, c.o, c.wu, 0, cols * 4u, ld);
    slice::stage_step<PL, !(EVM & 1u)>(c.t, ob, c.o, 0, cols * 4u, ld, lds);
    slice::load_step<PL>(ib, c.o, c.wu, 1, cols * 4u, ld);
  }
  lds_barrier();  // the LDS tables and the first IN
  for (uint32_t it = 0; it <= mine; it++) {
    const bool front = it < mine, back = it > 0, next = it + 1 < mine, prev = it > 1;
    const uint32_t tf = front ? tile(it) : 0u;
    const uint32_t tb = back ? tile(it - 1) : 0u;
    const uint32_t tn = next ? tile(it + 1) : 0u;
    // the tile whose odd rows the previous step left in the staging, computed
    // afresh rather than carried: a value carried around the loop can settle
    // in vector registers, and a descriptor built from it wraps its stores in
    // a loop over lanes
    const uint32_t tp = prev ? tile(it - 2) : 0u;
    const Bufs<PL> bf{tile_buf<PL>(in + (size_t)tf * PL::W, front), tile_buf<PL>(out + (size_t)tf * PL::W, front),
                      tile_buf<PL>(in + (size_t)tn * PL::W, next),  tile_buf<PL>(out + (size_t)tn * PL::W, next),
                      tile_buf<PL>(out + (size_t)tb * PL::W, back), tile_buf<PL>(out + (size_t)tp * PL::W, prev)};
    uint32_t c4 = cols * 4u;
    asm volatile("" : "+s"(c4));
    // the thread's coordinates and addresses are rebuilt from an opaque copy of
    // its index in every iteration: hoisted out of the loop they outnumber the
    // free registers and spill
    int ti = tid;
    asm volatile("" : "+v"(ti));
    const slice::Tid<PL> tt(ti);
    const Ctx<PL> ci{tt, slice::Offs<PL>(tt, cols), c.wu, lds, c.t1w, c.t2, c.t3, c.k};
    step<PL, EVM, 0, MF>(ci, c4, front, back, bf, ld, held);
    if (front) {
      U4 ev[PL::QI];
      mid<PL, EVM, 0>(ci, c4, bf, held, ev);
    }
  }
  // the odd rows of the last tile's last step
  const uint32_t tl = mine > 0 ? tile(mine - 1) : 0u;
  const BufT<PL> ob = tile_buf<PL>(out + (size_t)tl * PL::W, mine > 0);
  slice::store_odd<PL>(c.t, ob, c.o, PL::STEPS - 1, cols * 4u, lds);
}

template <class PL, unsigned EVM>
__global__ __launch_bounds__(PL::NT) void encode_slice(const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, uint32_t cols,
                                                       uint32_t ntiles,
                                                       const typename PL::T1T* __restrict__ t1,
                                                       const typename PL::TwT* __restrict__ t2g,
                                                       const typename PL::TwT* __restrict__ t3g) {
  using TwT = typename PL::TwT;
  extern __shared__ int32_t lds[];
  const int tid = (int)threadIdx.x;
  TwT* t2 = reinterpret_cast<TwT*>(lds + 4 * PL::XWORDS);
  TwT* t3 = t2 + PL::T2_WORDS;
  for (int i = tid; i < PL::T2_WORDS; i += PL::NT) t2[i] = t2g[i];
  for (int i = tid; i < PL::T3_WORDS; i += PL::NT) t3[i] = t3g[i];
  const slice::Tid<PL> t(tid);
  const int wu = __builtin_amdgcn_readfirstlane(t.w);
  constexpr int T1_PER_WAVE = PL::T1_PAIRS / PL::RB;  // the P1 twiddles of nB = wave
  const Ctx<PL> c{t, slice::Offs<PL>(t, cols), wu, lds, t1 + wu * T1_PER_WAVE, t2, t3, slice::make_mulk()};
  if (kSplitOrder && wu >= PL::NT / 128)
    run_tiles<PL, EVM, true>(in, out, cols, ntiles, c, tid, lds);
  else
    run_tiles<PL, EVM, false>(in, out, cols, ntiles, c, tid, lds);
}

// Largest per-lane byte offset of a plan's accesses (the odd output row of the
// last lane row, its last column quad) below 2^32; a narrow plan's uniform
// offsets too.
template <class PL>
bool slice_plan_fits(uint32_t cols) {
  if (cols == 0 || cols % PL::W != 0) return false;
  int max_lane_row = 0;
  for (int g = 0; g < 64; g++) max_lane_row = std::max(max_lane_row, PL::lane_row(g & 15));
  // a step's accesses share a base row; the offsets of its last access
  const uint32_t last_off = slice::Offs<PL>::o_out(0, 0, PL::QI - 1);
  const uint64_t lane = ((uint64_t)(last_off + 2 * max_lane_row + 1) * cols + PL::W) * 4;
  if (PL::WIDE) return lane < (1ull << 32);
  return ((uint64_t)(2 * PL::S - 1) * cols + PL::W) * 4 < (1ull << 32);
}

}  // namespace

bool slice_fits(uint32_t k, uint32_t cols, int layout) {
  if (layout != ZODA_LAYOUT_ROW) return false;
  if (k == (uint32_t)slice::Plan2048::S) return slice_plan_fits<slice::Plan2048>(cols);
  if (k == (uint32_t)slice::Plan4096::S) return slice_plan_fits<slice::Plan4096>(cols);
  return false;
}

namespace {

template <class PL>
int get_slice_tables(DeviceState* st, int slot, const typename PL::T1T** t1, const typename PL::TwT** t2,
                     const typename PL::TwT** t3) {
  using T1T = typename PL::T1T;
  using TwT = typename PL::TwT;
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  constexpr size_t bytes = PL::T1_PAIRS * sizeof(T1T) + (PL::T2_WORDS + PL::T3_WORDS) * sizeof(TwT);
  if (!st->slice_tables[slot]) {
    std::vector<char> host(bytes);
    T1T* h1 = reinterpret_cast<T1T*>(host.data());
    TwT* h2 = reinterpret_cast<TwT*>(h1 + PL::T1_PAIRS);
    slice::build_slice_tables<PL>(h1, h2, h2 + PL::T2_WORDS);
    void* d = nullptr;
    if (const hipError_t ea = hipMalloc(&d, bytes)) return alloc_fail(ea, "slice tables");
    hipError_t e = hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail((int)e, "copying slice tables failed: %s", hipGetErrorString(e));
    }
    st->slice_tables[slot] = d;
  }
  *t1 = static_cast<const T1T*>(st->slice_tables[slot]);
  *t2 = reinterpret_cast<const TwT*>(*t1 + PL::T1_PAIRS);
  *t3 = *t2 + PL::T2_WORDS;
  return 0;
}

template <class PL>
int launch_slice_plan(DeviceState* st, int slot, const uint32_t* d_in, uint32_t cols, uint32_t* d_out,
                      hipStream_t stream, bool dry) {
  const typename PL::T1T* t1 = nullptr;
  const typename PL::TwT *t2 = nullptr, *t3 = nullptr;
  ZODA_TRY(get_slice_tables<PL>(st, slot, &t1, &t2, &t3));
  const void* fn = (const void*)encode_slice<PL, PL::EVEN_MID_MASK>;
  uint32_t grid = 0;
  ZODA_TRY(kernel_setup(st, fn, PL::NT, PL::LDS_BYTES, true, &grid));
  if (dry) return 0;
  uint32_t ntiles = cols / PL::W;
  if (grid > ntiles) grid = ntiles;
  void* args[] = {(void*)&d_in, (void*)&d_out, (void*)&cols, (void*)&ntiles,
                  (void*)&t1,   (void*)&t2,    (void*)&t3};
  ZODA_HIP(hipLaunchKernel(fn, dim3(grid), dim3(PL::NT), args, PL::LDS_BYTES, stream));
  return 0;
}

}  // namespace

int launch_slice(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t* d_out,
                 hipStream_t stream, bool dry) {
  if (k == (uint32_t)slice::Plan2048::S)
    return launch_slice_plan<slice::Plan2048>(st, 0, d_in, cols, d_out, stream, dry);
  if (k == (uint32_t)slice::Plan4096::S)
    return launch_slice_plan<slice::Plan4096>(st, 1, d_in, cols, d_out, stream, dry);
  return fail(ZODA_EINVAL, "no slice plan for side %u", k);
}

}  // namespace zoda
