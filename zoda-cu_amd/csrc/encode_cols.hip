// Encode of whole squares into the column-major layout (ZODA_LAYOUT_COL,
// out[c * 2k + r]), the reference's PinnedSquare layout (src/ntt/cuda.rs:31-67)
// that encode_gpu fills with transpose + pad + INTT + NTT
// (src/benchmarks/utils.rs:100-146).
//
// The columns go in chunks.  A chunk is first transposed from the row-major
// k x cols input into a column-major scratch (column c's k values
// contiguous); the column kernel then runs, per column, INTT_k, twist,
// NTT_k as FP64 digit passes over tiles of 4096 values in LDS
// (encode_cols_core.hpp) and writes each column's 2k output values as one
// contiguous run.  One persistent kernel does both kinds of work: launch i
// encodes the columns of chunk i - 1 from one scratch buffer while it
// transposes chunk i into the other, one 64 x 64 block per column tile (the
// counts match: a chunk of C columns has C / (4096 / k) tiles and
// (k / 64) (C / 64) blocks), so the memory-bound transpose runs beside the
// arithmetic instead of between its launches.
//
// A chunk's scratch is 64 MiB, a buffer pair half of MI355X's 256 MB Infinity
// Cache, and the chunks reuse it: the transposed values go to the cache and
// come back from it instead of making a round trip to HBM.
//
// The scratch comes from the library's stream-ordered pool (pool_alloc), so
// encodes on different streams never share it.
#include "common.hpp"
#include "encode_cols_core.hpp"

#include <algorithm>

namespace zoda {

constexpr uint32_t COLS_SCRATCH_BYTES = 64u << 20;  // one chunk's transposed columns

struct ColsWork {
  const uint32_t* t;  // column-major scratch of the chunk to encode (null: none)
  uint32_t cols;      // its columns
  uint32_t* out;      // its output columns
  const uint32_t* in; // row-major input of the chunk to transpose (null: none)
  uint32_t pitch;     // u32 between input rows
  uint32_t tcols;     // its columns (a multiple of 64)
  uint32_t* t_next;   // its column-major scratch
};

// Persistent: each work-group copies the twiddle tables into LDS once, then
// walks work items b: the transpose block b of the next chunk (through LDS
// to the other scratch buffer), then the column tile b of this one (load --
// each thread's sixteen values, kept in registers for the store phase --
// passes, store).  A block's input is loaded one item ahead, before the
// previous tile's passes, so its latency hides behind their arithmetic.  Loading the next tile into registers during the passes instead
// (as the drop-in transforms do) takes the kernel from 164 to 208 VGPRs, two
// waves per SIMD instead of three: 2.85 against 2.77 ms for the whole
// side-2048 encode, measured with separate transpose launches.
// MODE: COLS_BOTH (encode the tiles of one chunk and transpose the blocks of
// the next; the next chunk never has more blocks than this one has tiles),
// COLS_TRANSPOSE (the first launch) or COLS_ENCODE (the last).
enum { COLS_BOTH = 0, COLS_TRANSPOSE = 1, COLS_ENCODE = 2 };

template <int L, int MODE>
__global__ __launch_bounds__(nttf::NT) void encode_cols_kernel(ColsWork w, const uint32_t* __restrict__ twg) {
  using PL = ecol::Plan<L>;
  static_assert(ecol::TB % PL::COLS == 0, "chunks of whole transpose blocks are whole tiles");
  constexpr bool TR = MODE != COLS_ENCODE, COL = MODE != COLS_TRANSPOSE;
  extern __shared__ int32_t lds[];
  uint32_t* tw = reinterpret_cast<uint32_t*>(lds + nttf::LDS_DATA);
  uint32_t* sh = reinterpret_cast<uint32_t*>(lds);  // a transpose block, between tiles
  const int tid = (int)threadIdx.x;
  if constexpr (COL)
    for (int i = tid; i < PL::TW_WORDS; i += nttf::NT) tw[i] = twg[i];
  const slice::MulK k = slice::make_mulk();
  const uint32_t ntiles = COL ? w.cols / PL::COLS : 0u;
  const uint32_t nblocks = TR ? (uint32_t)(PL::S / ecol::TB) * (w.tcols / ecol::TB) : 0u;
  const uint32_t nitems = COL ? ntiles : nblocks;
  // Raw-buffer accesses: a per-lane 32-bit byte offset and a uniform one
  // (every byte offset stays below 2^32: a chunk's scratch is 64 MiB, its
  // output 128 MiB, a transpose block's input rows at most 63 pitches apart
  // with pitch < 2^24), so no 64-bit address stays live across the loop.
  // Every memory access in the loop is unconditional: an item past the end
  // of its kind goes through a buffer of zero records (loads return 0,
  // stores are dropped).  A branch around one would leave the compiler's
  // count of outstanding accesses conservative after it, and the next wait
  // would then cover loads issued to be consumed much later.
  const GBuf tin(w.t), tout(w.out);
  const uint32_t lane_in = 8u * tid, lane_out = 16u * tid;
  const uint32_t tb_lane_in = 4u * (uint32_t)ecol::tb_in_offset(tid, 0, w.pitch);
  const uint32_t tb_lane_out = 4u * (uint32_t)ecol::tb_out_offset(tid, 0, PL::S);
  // transpose block b's input, loaded one item ahead
  U4 tv[ecol::TB_RUNS];
  auto fetch_block = [&](uint32_t b) {
    const ecol::TBlock tb = ecol::tblock(b < nblocks ? b : 0u, PL::S);
    const GBuf src(w.in + (size_t)tb.r0 * w.pitch + tb.c0, b < nblocks);
#pragma unroll
    for (int q = 0; q < ecol::TB_RUNS; q++) tv[q] = src.load4(tb_lane_in, 4u * q * (nttf::NT / 16) * w.pitch);
  };
  if constexpr (TR) fetch_block(blockIdx.x);
  if constexpr (TR && COL) {
    // Stores into a zero-record buffer (they store nothing): entering the
    // loop, as around it, the block's loads are then followed by a tile's
    // stores before the next pair loads, so the wait for the block leaves the
    // previous tile's stores outstanding (the compiler takes the smaller count
    // where the two paths meet).
    const GBuf none(w.out, 0u);
#pragma unroll
    for (int q = 0; q < ecol::PAIRS; q++) none.store4(U4{0, 0, 0, 0}, lane_out, 16u * q * nttf::NT);
  }
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    // Each pair stays one 64-bit register pair from its load to the store
    // phase: split earlier, the compiler moves the second value into its
    // place in the 16-byte store's operand right after the load, which makes
    // the wave wait for the load there.  (An array of U2 structs is worse: the
    // compiler leaves it in scratch memory, and every store then waits on all
    // outstanding loads.)
    uint64_t pr[ecol::PAIRS];
    const uint32_t tbase = it * (uint32_t)(nttf::TILE * 4);  // the tile's first scratch byte
    if constexpr (COL) {
#pragma unroll
      for (int q = 0; q < ecol::PAIRS; q++) pr[q] = tin.load2(lane_in, tbase + 8u * q * nttf::NT);
    }
    if constexpr (TR) {
      const ecol::TBlock tb = ecol::tblock(it < nblocks ? it : 0u, PL::S);
      const GBuf dst(w.t_next + (size_t)tb.c0 * PL::S + tb.r0, it < nblocks);
      __syncthreads();  // the previous tile's last reads of LDS
#pragma unroll
      for (int q = 0; q < ecol::TB_RUNS; q++) ecol::tb_put(tid, q, sh, tv[q]);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < ecol::TB_RUNS; q++)
        dst.store4_cached(ecol::tb_get(tid, q, sh), tb_lane_out, 4u * q * (nttf::NT / 16) * PL::S);
      fetch_block(it + gridDim.x);  // arrives during this tile's passes
    }
    if constexpr (COL) {
      __syncthreads();  // the tables, the block's or the previous tile's last reads of LDS
#pragma unroll
      for (int q = 0; q < ecol::PAIRS; q++) ecol::put_pair(tid, q, lds, ecol::U2{(uint32_t)pr[q], (uint32_t)(pr[q] >> 32)});
      __syncthreads();
      ecol::all_passes<PL>([&](auto pass) { pass(tid, lds, tw, k); }, [] { __syncthreads(); });
      ecol::store_pairs<PL>(
          tid, (uint32_t)PL::COLS, (size_t)it * PL::COLS, lds,
          [&](int q) {
            asm volatile("" : "+v"(pr[q]));
            return ecol::U2{(uint32_t)pr[q], (uint32_t)(pr[q] >> 32)};
          },
          [&](int q, size_t, uint32_t, const U4& v) { tout.store4(v, lane_out, 2u * tbase + 16u * q * nttf::NT); });
    }
  }
}

template <int L>
static int launch_cols_plan(DeviceState* st, const uint32_t* d_in, uint32_t cols, uint32_t* d_out,
                            hipStream_t stream) {
  using PL = ecol::Plan<L>;
  const uint32_t key = 0x80000000u | (uint32_t)L;  // beside the poly transforms' keys
  const uint32_t* twg = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lock(st->mu);
    auto it = st->ntt_fp64.find(key);
    if (it == st->ntt_fp64.end()) {
      std::vector<uint32_t> host(PL::TW_WORDS);
      ecol::build_col_tables<PL>(host.data());
      void* d = nullptr;
      if (const hipError_t ea = hipMalloc(&d, host.size() * 4)) return alloc_fail(ea, "column encode tables");
      const hipError_t e = hipMemcpy(d, host.data(), host.size() * 4, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(d);
        return fail((int)e, "copying column encode tables failed: %s", hipGetErrorString(e));
      }
      it = st->ntt_fp64.emplace(key, static_cast<uint32_t*>(d)).first;
    }
    twg = it->second;
  }
  const void* fns[3] = {(const void*)encode_cols_kernel<L, COLS_BOTH>, (const void*)encode_cols_kernel<L, COLS_TRANSPOSE>,
                        (const void*)encode_cols_kernel<L, COLS_ENCODE>};
  const size_t lds_bytes = (size_t)PL::LDS_WORDS * sizeof(int32_t);
  uint32_t grid = 0;
  for (const void* f : fns) ZODA_TRY(kernel_setup(st, f, nttf::NT, lds_bytes, true, &grid));

  const uint32_t k = PL::S;
  const uint32_t chunk = std::min<uint32_t>(cols, COLS_SCRATCH_BYTES / ((uint32_t)sizeof(uint32_t) * k));
  const uint32_t nchunks = (cols + chunk - 1) / chunk;
  uint32_t* scratch = nullptr;
  const size_t buf = (size_t)k * chunk;  // u32 per buffer
  ZODA_TRY(pool_alloc(st, 2 * buf * sizeof(uint32_t), stream, (void**)&scratch));
  hipError_t e = hipSuccess;
  // launch i: encode chunk i - 1 (buffer (i - 1) % 2), transpose chunk i (buffer i % 2)
  for (uint32_t i = 0; i <= nchunks && e == hipSuccess; i++) {
    ColsWork w{};
    uint32_t items = 0;
    if (i > 0) {
      const uint32_t c0 = (i - 1) * chunk;
      w.t = scratch + ((i - 1) % 2) * buf;
      w.cols = std::min(chunk, cols - c0);
      w.out = d_out + (size_t)c0 * 2 * k;
      items = w.cols / PL::COLS;
    }
    if (i < nchunks) {
      const uint32_t c0 = i * chunk;
      w.in = d_in + c0;
      w.pitch = cols;
      w.tcols = std::min(chunk, cols - c0);
      w.t_next = scratch + (i % 2) * buf;
      if (i == 0) items = (k / ecol::TB) * (w.tcols / ecol::TB);
    }
    const void* fn = fns[i == 0 ? COLS_TRANSPOSE : i == nchunks ? COLS_ENCODE : COLS_BOTH];
    const uint32_t g = std::min(grid, items);
    void* args[] = {(void*)&w, (void*)&twg};
    e = hipLaunchKernel(fn, dim3(g), dim3(nttf::NT), args, lds_bytes, stream);
  }
  const hipError_t ef = hipFreeAsync(scratch, stream);
  if (e != hipSuccess) return fail((int)e, "column encode launch failed: %s", hipGetErrorString(e));
  if (ef != hipSuccess) return fail((int)ef, "hipFreeAsync failed: %s", hipGetErrorString(ef));
  return 0;
}

template <int L>
static int launch_cols_size(DeviceState* st, const uint32_t* d_in, uint32_t log_k, uint32_t cols,
                            uint32_t* d_out, hipStream_t stream) {
  if constexpr (L <= 12) {
    if (log_k == (uint32_t)L) return launch_cols_plan<L>(st, d_in, cols, d_out, stream);
    return launch_cols_size<L + 1>(st, d_in, log_k, cols, d_out, stream);
  } else {
    return fail(ZODA_EINVAL, "no column encode plan for k = 2^%u", log_k);
  }
}

bool cols_fits(const uint32_t* d_in, uint32_t k, uint32_t cols, const uint32_t* d_out, hipStream_t stream) {
  // sides 256 to 4096 (below 256 one launch of encode_kernel is faster: 11.8
  // against 29-33 us at side 64, 27.5 against 31-45 us at 128, 128 s columns)
  if (k < 256 || k > 4096 || (k & (k - 1)) || cols == 0 || cols % ecol::TB || cols >= (1u << 24)) return false;
  if (((uintptr_t)d_in & 15) || ((uintptr_t)d_out & 15)) return false;
  // the stream-ordered scratch stays out of graph capture
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  return true;
}

int launch_cols(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t* d_out,
                hipStream_t stream) {
  return launch_cols_size<8>(st, d_in, log2_ceil(k), cols, d_out, stream);
}

}  // namespace zoda
