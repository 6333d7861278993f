// Encode of whole squares into the column-major layout (ZODA_LAYOUT_COL,
// out[c * 2k + r]), the reference's PinnedSquare layout (src/ntt/cuda.rs:31-67)
// that encode_gpu fills with transpose + pad + INTT + NTT
// (src/benchmarks/utils.rs:100-146).  Per chunk of columns, two kernels on
// the caller's stream:
//
//   1. transpose_to_cols: the chunk of the row-major k x cols input into a
//      column-major scratch (column c's k values contiguous), 64 x 64 tiles
//      through LDS;
//   2. encode_cols_kernel: per column, INTT_k, twist, NTT_k as FP64 digit
//      passes over tiles of 4096 values in LDS (encode_cols_core.hpp), each
//      column's 2k output values written as one contiguous run.
//
// A chunk's scratch is 128 MiB, half of MI355X's 256 MB Infinity Cache, and
// every chunk reuses it: the transposed values go to the cache and come back
// from it instead of making a round trip to HBM.  At side 2048 (262144
// columns) chunks of 16384 columns take the encode from 2.81 to 2.51 ms
// against one transpose of the whole square (alternating runs); 8192 and
// 32768 columns came out at 2.72 and 2.69 ms.  The scratch comes from the
// stream-ordered allocator, so encodes on different streams never share it.
#include "common.hpp"
#include "encode_cols_core.hpp"

#include <algorithm>

namespace zoda {

constexpr int TR_SIDE = 64;    // transpose tile: 64 rows x 64 columns
constexpr int TR_PITCH = 65;   // LDS words per tile row (conflict-free columns)
constexpr int TR_THREADS = 256;
constexpr uint32_t COLS_SCRATCH_BYTES = 128u << 20;  // one chunk's transposed columns

// `pitch` u32 between input rows; the grid's tiles cover the chunk.  Tile b
// covers rows 64 (b % tiles_r) .. +63 and columns 64 (b / tiles_r) .. +63:
// the work-groups in flight at once read whole rows of a band of columns and
// write one contiguous stretch of the scratch.  Each thread moves four
// 16-byte runs in (16 lanes a tile row) and four out (16 lanes a tile
// column).
__global__ __launch_bounds__(TR_THREADS) void transpose_to_cols(const uint32_t* __restrict__ in, uint32_t rows,
                                                                uint32_t pitch, uint32_t* __restrict__ t) {
  __shared__ uint32_t sh[TR_SIDE * TR_PITCH];
  const uint32_t tiles_r = rows / TR_SIDE;
  const size_t r0 = (size_t)(blockIdx.x % tiles_r) * TR_SIDE;
  const size_t c0 = (size_t)(blockIdx.x / tiles_r) * TR_SIDE;
  const int tid = (int)threadIdx.x;
  U4 v[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int id = tid + q * TR_THREADS, r = id >> 4, cc = id & 15;
    v[q] = *reinterpret_cast<const U4*>(in + (r0 + r) * pitch + c0 + 4 * cc);
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int id = tid + q * TR_THREADS, r = id >> 4, cc = id & 15;
    uint32_t* d = sh + r * TR_PITCH + 4 * cc;
    d[0] = v[q].x;
    d[1] = v[q].y;
    d[2] = v[q].z;
    d[3] = v[q].w;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int id = tid + q * TR_THREADS, c = id >> 4, rc = id & 15;
    const uint32_t* s = sh + 4 * rc * TR_PITCH + c;
    const U4 o{s[0], s[TR_PITCH], s[2 * TR_PITCH], s[3 * TR_PITCH]};
    store16(t + (c0 + c) * rows + r0 + 4 * rc, o);
  }
}

// Persistent: each work-group copies the twiddle tables into LDS once, then
// walks tiles of 4096 / k columns: load (each thread's sixteen values, kept
// in registers for the store phase), passes, store.  Loading the next tile
// into registers during the passes instead (as the drop-in transforms do)
// measured 2.85 against 2.77 ms for the whole side-2048 encode: it takes the
// kernel from 164 to 208 VGPRs, two waves per SIMD instead of three.
template <int L>
__global__ __launch_bounds__(nttf::NT) void encode_cols_kernel(const uint32_t* __restrict__ t, uint32_t cols,
                                                               uint32_t* __restrict__ out,
                                                               const uint32_t* __restrict__ twg) {
  using PL = ecol::Plan<L>;
  extern __shared__ int32_t lds[];
  uint32_t* tw = reinterpret_cast<uint32_t*>(lds + nttf::LDS_DATA);
  const int tid = (int)threadIdx.x;
  for (int i = tid; i < PL::TW_WORDS; i += nttf::NT) tw[i] = twg[i];
  const slice::MulK k = slice::make_mulk();
  const uint32_t ntiles = (cols + PL::COLS - 1) / PL::COLS;
  auto load2 = [&](size_t c, uint32_t j) -> ecol::U2 {
    return *reinterpret_cast<const ecol::U2*>(t + c * PL::S + j);
  };
  // plain u32 arrays: an array of U2 is left in scratch memory by the
  // compiler, and every store then waits on all outstanding loads
  uint32_t cx[ecol::PAIRS], cy[ecol::PAIRS];
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const size_t col0 = (size_t)tile * PL::COLS;
    const uint32_t live = min((uint32_t)PL::COLS, cols - (uint32_t)col0);
#pragma unroll
    for (int q = 0; q < ecol::PAIRS; q++) {
      const ecol::U2 x = ecol::tile_pair<PL>(tid, q, live, col0, load2);
      cx[q] = x.x;
      cy[q] = x.y;
    }
    __syncthreads();  // the tables, or the previous tile's last reads of LDS
#pragma unroll
    for (int q = 0; q < ecol::PAIRS; q++) ecol::put_pair(tid, q, lds, ecol::U2{cx[q], cy[q]});
    __syncthreads();
    ecol::all_passes<PL>([&](auto pass) { pass(tid, lds, tw, k); }, [] { __syncthreads(); });
    ecol::store_pairs<PL>(
        tid, live, col0, lds, [&](int q) { return ecol::U2{cx[q], cy[q]}; },
        [&](size_t c, uint32_t r, const U4& v) { store16<true>(out + c * (2 * PL::S) + r, v); });
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
  const void* fn = (const void*)encode_cols_kernel<L>;
  const size_t lds_bytes = (size_t)PL::LDS_WORDS * sizeof(int32_t);
  uint32_t grid = 0;
  ZODA_TRY(kernel_setup(st, fn, nttf::NT, lds_bytes, true, &grid));

  const uint32_t k = PL::S;
  const uint32_t chunk = std::min<uint32_t>(cols, COLS_SCRATCH_BYTES / ((uint32_t)sizeof(uint32_t) * k));
  uint32_t* scratch = nullptr;
  ZODA_HIP(hipMallocAsync((void**)&scratch, (size_t)k * chunk * sizeof(uint32_t), stream));
  hipError_t e = hipSuccess;
  for (uint32_t c0 = 0; c0 < cols && e == hipSuccess; c0 += chunk) {
    const uint32_t n = std::min(chunk, cols - c0);
    const unsigned tr_blocks = (k / TR_SIDE) * (n / TR_SIDE);
    hipLaunchKernelGGL(transpose_to_cols, dim3(tr_blocks), dim3(TR_THREADS), 0, stream, d_in + c0, k, cols, scratch);
    e = hipGetLastError();
    if (e == hipSuccess) {
      const uint32_t* cs = scratch;
      uint32_t* o = d_out + (size_t)c0 * 2 * k;
      const uint32_t g = std::min(grid, (n + PL::COLS - 1) / PL::COLS);
      void* args[] = {(void*)&cs, (void*)&n, (void*)&o, (void*)&twg};
      e = hipLaunchKernel(fn, dim3(g), dim3(nttf::NT), args, lds_bytes, stream);
    }
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
  if (k < 2048 || k > 4096 || (k & (k - 1)) || cols == 0 || cols % TR_SIDE) return false;
  if (((uintptr_t)d_in & 15) || ((uintptr_t)d_out & 15)) return false;
  // the stream-ordered scratch stays out of graph capture
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  return true;
}

int launch_cols(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t* d_out,
                hipStream_t stream) {
  return launch_cols_size<11>(st, d_in, log2_ceil(k), cols, d_out, stream);
}

}  // namespace zoda
