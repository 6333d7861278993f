// Fused ZODA column encode for gfx950: one launch turns the row-major k x cols
// input into the 2*n_k x cols extended square (see encode_core.hpp for the
// algorithm and encode_plans.hpp for the per-side tiling).
#include "common.hpp"
#include "encode_plans.hpp"
#include "encode_tables.hpp"

namespace zoda {

// LAYOUT is a template parameter so each variant addresses its output with
// one base and a constant stride per value.
template <class PL, bool FULL, int LAYOUT>
__global__ __launch_bounds__(PL::NT) void encode_kernel(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out,
                                                        uint32_t krows, uint32_t cols,
                                                        EncodeTabs T) {
  constexpr int layout = LAYOUT;
  extern __shared__ int32_t lds[];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tid = (int)threadIdx.x;
  if constexpr (PL::TWL) {
    encode_load_tables<PL>(tid, lds, T);
    __syncthreads();
  }
  encode_step<PL, 0, FULL>(tid, tile, in, out, krows, cols, layout, lds, T);
  if constexpr (PL::STEPS > 1) {
    __syncthreads();
    encode_step<PL, 1, FULL>(tid, tile, in, out, krows, cols, layout, lds, T);
    __syncthreads();
    encode_step<PL, 2, FULL>(tid, tile, in, out, krows, cols, layout, lds, T);
  }
  if constexpr (PL::STEPS > 3) {
    __syncthreads();
    encode_step<PL, 3, FULL>(tid, tile, in, out, krows, cols, layout, lds, T);
    __syncthreads();
    encode_step<PL, 4, FULL>(tid, tile, in, out, krows, cols, layout, lds, T);
  }
}

// Persistent form with the first pass pipelined through registers (Pipe).
// Every tile moves 4 bytes in and 8 out per value and a work-group fills its
// CU, so the memory work is spread over the passes rather than issued in
// bursts the CU's queues cannot absorb (issuing the next tile's 32 loads per
// thread at once stalls every wave until most of them have returned): the
// even-row stores go out during the first pass, straight from the registers
// it reads; the next tile's loads are issued in parts, one per group of the
// middle passes; the
// odd-row stores leave the last pass as it produces them.  BUF: every global
// access goes through a raw buffer with 32-bit offsets from the tile's first
// column (outputs below 4 GiB).
template <class PL, bool BUF>
__global__ __launch_bounds__(PL::NT) void encode_pipelined(const uint32_t* __restrict__ in,
                                                           uint32_t* __restrict__ out,
                                                           uint32_t cols, uint32_t ntiles,
                                                           EncodeTabs T) {
  using PP = Pipe<PL, BUF>;
  // groups per thread of the passes between the first and the last
  constexpr int G1 = (PL::W << (PL::L - (PL::D == 2 ? PL::RD : PL::R2))) / PL::NT;
  constexpr int G2 = PL::D == 3 ? (PL::W << (PL::L - PL::RD)) / PL::NT : 0;
  constexpr int G3 = PL::D == 3 ? G1 : 0;
  constexpr int NP = G1 + G2 + G3;  // load parts
  extern __shared__ int32_t lds[];
  encode_load_tables<PL>((int)threadIdx.x, lds, T);
  const uint32_t* sb = reinterpret_cast<const uint32_t*>(lds + PL::TILE_WORDS);
  const TwTab inv1{T.inv1, PL::TWL ? sb + T.o_inv1 : nullptr};
  const uint32_t k = (uint32_t)PL::S;
  uint32_t raw[PP::G][PP::RR];
  uint32_t l = blockIdx.x;
  if (l < ntiles) PP::load((int)threadIdx.x, xcd_tile(l, ntiles), in, cols, raw);
  __syncthreads();
  for (; l < ntiles; l += gridDim.x) {
    const uint32_t tile = xcd_tile(l, ntiles);
    const uint32_t next = l + gridDim.x;
    const bool more = next < ntiles;
    const uint32_t ntile = more ? xcd_tile(next, ntiles) : 0u;
    const int tid = (int)threadIdx.x;
    PP::first(tid, raw, lds, T, inv1, [&](int gi) { PP::store_even(tid, gi, tile, out, cols, raw[gi]); });
    __syncthreads();
    encode_step<PL, 1, true, BUF>(tid, tile, in, out, k, cols, 0, lds, T, [&](int gi) {
      if (more) PP::load_part(tid, ntile, in, cols, raw, gi, NP);
    });
    __syncthreads();
    if constexpr (PL::D == 3) {
      encode_step<PL, 2, true, BUF>(tid, tile, in, out, k, cols, 0, lds, T, [&](int gi) {
        if (more) PP::load_part(tid, ntile, in, cols, raw, G1 + gi, NP);
      });
      __syncthreads();
      encode_step<PL, 3, true, BUF>(tid, tile, in, out, k, cols, 0, lds, T, [&](int gi) {
        if (more) PP::load_part(tid, ntile, in, cols, raw, G1 + G2 + gi, NP);
      });
      __syncthreads();
      encode_step<PL, 4, true, BUF>(tid, tile, in, out, k, cols, 0, lds, T);
    } else {
      encode_step<PL, 2, true, BUF>(tid, tile, in, out, k, cols, 0, lds, T);
    }
    __syncthreads();  // the next tile's first pass writes LDS
  }
}

int get_encode_tables(DeviceState* st, int L, int r1, int r2, int rd, const EncodeTableSet** out) {
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  const uint32_t key = encode_table_key(L, r1, r2, rd);
  auto it = st->encode.find(key);
  if (it == st->encode.end()) {
    EncodeTabOffsets off;
    std::vector<Tw> host = build_encode_tables(L, r1, r2, rd, &off);
    PackedTables pk = pack_u32(host, off);
    EncodeTableSet set;
    hipError_t ea = hipMalloc((void**)&set.base, host.size() * sizeof(Tw));
    if (ea == hipSuccess) ea = hipMalloc((void**)&set.packed, pk.words.size() * sizeof(uint32_t));
    if (ea != hipSuccess) {
      (void)hipFree(set.base);
      return alloc_fail(ea, "encode table allocation");
    }
    hipError_t e = hipMemcpy(set.base, host.data(), host.size() * sizeof(Tw), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(set.packed, pk.words.data(), pk.words.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(set.base);
      (void)hipFree(set.packed);
      return fail((int)e, "copying encode tables failed: %s", hipGetErrorString(e));
    }
    set.tabs = new EncodeTabs(make_encode_tabs(set.base, off, set.packed, pk));
    it = st->encode.emplace(key, set).first;
  }
  *out = &it->second;
  return 0;
}

// Once per device and kernel: raise the dynamic LDS limit when the kernel
// needs more than the default 64 KiB and, for a persistent kernel, size its
// grid to fill every CU (*grid; 0 otherwise).
int kernel_setup(DeviceState* st, const void* fn, int nt, size_t lds_bytes, bool persistent,
                 uint32_t* grid) {
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  auto it = st->kernel_setup.find(fn);
  if (it != st->kernel_setup.end()) {
    *grid = it->second;
    return 0;
  }
  if (lds_bytes > 64 * 1024)
    ZODA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  uint32_t g = 0;
  if (persistent) {
    int per_cu = 0, cus = 0;
    ZODA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, lds_bytes));
    ZODA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, st->device));
    g = (uint32_t)std::max(1, per_cu * cus);
  }
  st->kernel_setup.emplace(fn, g);
  *grid = g;
  return 0;
}

template <class PL>
static int launch_plan(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols,
                       uint32_t* d_out, int layout, hipStream_t stream, bool dry) {
  const EncodeTableSet* set = nullptr;
  ZODA_TRY(get_encode_tables(st, PL::L, PL::R1, PL::R2, PL::RD, &set));
  const EncodeTabs& T = *set->tabs;
  if (PL::TWL && T.packed_words > (uint32_t)PL::TW_WORDS)
    return fail(ZODA_ESTATE, "encode tables (%u words) exceed the plan's LDS share", T.packed_words);
  const size_t lds_bytes = (size_t)PL::LDS_WORDS * sizeof(int32_t);
  const bool full = cols % PL::W == 0 && k == (uint32_t)PL::S;
  const bool row = layout == ZODA_LAYOUT_ROW;
  // whole squares into the column-major layout: transpose + column kernel
  if (!row && !dry && k == (uint32_t)PL::S && cols_fits(d_in, k, cols, d_out, stream))
    return launch_cols(st, d_in, k, cols, d_out, stream);
  // sides 2048 and 4096: the slice-pipelined kernel where it fits
  if constexpr (PL::L == 11 || PL::L == 12) {
    if (k == (uint32_t)PL::S && slice_fits(k, cols, layout))
      return launch_slice(st, d_in, k, cols, d_out, stream, dry);
  }
  if constexpr (PL::PERSIST) {
    // Pipe addresses rows within a digit with 32-bit byte offsets; whole
    // arrays of at most 4 GiB also take the raw-buffer form.
    if (full && row && ((uint64_t)cols << (PL::LO1 + 1)) * 4 < (1ull << 32)) {
      const bool buf = ((uint64_t)(2 * PL::S - 1) * cols + PL::W) * 4 < (1ull << 32);
      const void* pfn = buf ? (const void*)encode_pipelined<PL, true> : (const void*)encode_pipelined<PL, false>;
      uint32_t grid = 0;
      ZODA_TRY(kernel_setup(st, pfn, PL::NT, lds_bytes, true, &grid));
      if (dry) return 0;
      uint32_t ntiles = cols / PL::W;
      if (grid > ntiles) grid = ntiles;
      void* pargs[] = {(void*)&d_in, (void*)&d_out, (void*)&cols, (void*)&ntiles, (void*)&T};
      ZODA_HIP(hipLaunchKernel(pfn, dim3(grid), dim3(PL::NT), pargs, lds_bytes, stream));
      return 0;
    }
  }
  const void* fn = full ? (row ? (const void*)encode_kernel<PL, true, 0> : (const void*)encode_kernel<PL, true, 1>)
                        : (row ? (const void*)encode_kernel<PL, false, 0> : (const void*)encode_kernel<PL, false, 1>);
  uint32_t unused = 0;
  ZODA_TRY(kernel_setup(st, fn, PL::NT, lds_bytes, false, &unused));
  if (dry) return 0;
  const uint32_t ntiles = (cols + PL::W - 1) / PL::W;
  void* args[] = {(void*)&d_in, (void*)&d_out, (void*)&k, (void*)&cols, (void*)&T};
  ZODA_HIP(hipLaunchKernel(fn, dim3(ntiles), dim3(PL::NT), args, lds_bytes, stream));
  ZODA_HIP(hipGetLastError());
  return 0;
}

// Column-major staging used for shapes the fused kernel does not cover
// (n_kn != 2 * n_k, or n_k > 2^12): transpose with zero padding, two batched
// transforms, transpose back.
__global__ void gather_columns(const uint32_t* __restrict__ in, uint32_t k, uint32_t cols,
                               uint32_t n_kn, uint32_t* __restrict__ tmp) {
  const size_t total = (size_t)cols * n_kn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)(i / n_kn), r = (uint32_t)(i % n_kn);
    tmp[i] = r < k ? in[(size_t)r * cols + c] : 0u;
  }
}

__global__ void scatter_rows(const uint32_t* __restrict__ tmp, uint32_t cols, uint32_t n_kn,
                             uint32_t* __restrict__ out) {
  const size_t total = (size_t)cols * n_kn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)(i / cols), c = (uint32_t)(i % cols);
    out[i] = tmp[(size_t)c * n_kn + r];
  }
}

static int encode_general(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols,
                          uint32_t log_nk, uint32_t log_nkn, uint32_t* d_out, int layout,
                          hipStream_t stream) {
  const uint32_t n_kn = 1u << log_nkn;
  const size_t total = (size_t)cols * n_kn;
  uint32_t* tmp = d_out;
  if (layout == ZODA_LAYOUT_ROW) ZODA_TRY(pool_alloc(st, total * 4, stream, (void**)&tmp));
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 1u << 16);
  hipLaunchKernelGGL(gather_columns, dim3(blocks), dim3(256), 0, stream, d_in, k, cols, n_kn, tmp);
  ZODA_HIP(hipGetLastError());
  MontTables tk, tkn;
  int rc = get_mont_tables(st, log_nk, &tk);
  if (!rc) rc = get_mont_tables(st, log_nkn, &tkn);
  if (!rc) rc = launch_ntt(st, tmp, log_nk, cols, n_kn, 1, tk, true, stream);
  if (!rc) rc = launch_ntt(st, tmp, log_nkn, cols, n_kn, 0, tkn, true, stream);
  if (!rc && layout == ZODA_LAYOUT_ROW) {
    hipLaunchKernelGGL(scatter_rows, dim3(blocks), dim3(256), 0, stream, tmp, cols, n_kn, d_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail((int)e, "scatter_rows launch failed: %s", hipGetErrorString(e));
  }
  if (layout == ZODA_LAYOUT_ROW) {
    hipError_t e = hipFreeAsync(tmp, stream);
    if (!rc && e != hipSuccess) rc = fail((int)e, "hipFreeAsync failed: %s", hipGetErrorString(e));
  }
  return rc;
}

int launch_encode(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t n,
                  uint32_t* d_out, int layout, hipStream_t stream, bool dry) {
  const uint32_t log_nk = log2_ceil(k);
  const uint32_t log_nkn = log2_ceil((uint64_t)k + n);
  if (log_nkn == log_nk + 1 && (int)log_nk <= ENCODE_MAX_LOG) {
    return with_encode_plan((int)log_nk, [&](auto plan) -> int {
      return launch_plan<decltype(plan)>(st, d_in, k, cols, d_out, layout, stream, dry);
    });
  }
  if (dry) return 0;
  return encode_general(st, d_in, k, cols, log_nk, log_nkn, d_out, layout, stream);
}

}  // namespace zoda
