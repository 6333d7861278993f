// Batched NTT / INTT over polynomials laid end to end (poly-major), the layout
// of the reference's cuda_ntt / cuda_intt (cuda/ntt_kernel.cu:106-208): poly c
// is the n values at values[c * stride]; values n..stride-1 are untouched.
//
// Canonical u32 values and Montgomery-form twiddles (mont_mul(x, w*2^32) =
// x*w), radix-2 decimation in time after a bit-reversed load, exactly the
// arithmetic of the reference's CPU ntt (src/ntt/cpu.rs:1-32) so any roots
// the caller passes give the reference's answer.  Sizes up to 2^12 run
// LDS-resident, one launch for all stages; larger sizes fall back to one
// launch per stage on device memory.  With the standard roots, sizes 2^2 to
// 2^12 run the FP64 digit-pass kernel instead (ntt_fp64_core.hpp).
#include "common.hpp"
#include "ntt_fp64_core.hpp"

namespace zoda {

constexpr int NTT_LDS_WORDS = 4096;
constexpr int NTT_THREADS = 256;

template <bool INV>
__global__ __launch_bounds__(NTT_THREADS) void ntt_poly_lds(uint32_t* __restrict__ data,
                                                            uint32_t log_n, uint32_t log_ppw,
                                                            uint32_t polys, uint32_t stride,
                                                            const uint32_t* __restrict__ tw,
                                                            uint32_t scale_mont) {
  __shared__ uint32_t lds[NTT_LDS_WORDS];
  const uint32_t n = 1u << log_n, total = n << log_ppw;
  const size_t p0 = (size_t)blockIdx.x << log_ppw;
  for (uint32_t idx = threadIdx.x; idx < total; idx += NTT_THREADS) {
    const uint32_t p = idx >> log_n, j = idx & (n - 1);
    const uint32_t r = log_n ? (__brev(j) >> (32 - log_n)) : 0u;
    lds[(p << log_n) | r] = (p0 + p < polys) ? data[(p0 + p) * stride + j] : 0u;
  }
  __syncthreads();
  for (uint32_t lh = 0; lh < log_n; lh++) {
    const uint32_t h = 1u << lh;
    for (uint32_t b = threadIdx.x; b < total / 2; b += NTT_THREADS) {
      const uint32_t p = b >> (log_n - 1);
      const uint32_t i = b & ((n >> 1) - 1);
      const uint32_t j = i & (h - 1);
      const uint32_t base = (p << log_n) + ((i >> lh) << (lh + 1)) + j;
      const uint32_t u = lds[base];
      const uint32_t v = mont_mul(lds[base + h], tw[h + j]);
      lds[base] = add_mod(u, v);
      lds[base + h] = sub_mod(u, v);
    }
    __syncthreads();
  }
  for (uint32_t idx = threadIdx.x; idx < total; idx += NTT_THREADS) {
    const uint32_t p = idx >> log_n, j = idx & (n - 1);
    if (p0 + p < polys) {
      uint32_t v = lds[idx];
      if (INV) v = mont_mul(v, scale_mont);
      data[(p0 + p) * stride + j] = v;
    }
  }
}

__global__ void bitrev_global(uint32_t* __restrict__ data, uint32_t log_n, uint32_t polys,
                              uint32_t stride) {
  const size_t n = (size_t)1 << log_n, total = n * polys;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t p = t >> log_n;
    const uint32_t i = (uint32_t)(t & (n - 1));
    const uint32_t r = __brev(i) >> (32 - log_n);
    if (i < r) {
      uint32_t* poly = data + p * stride;
      const uint32_t a = poly[i];
      poly[i] = poly[r];
      poly[r] = a;
    }
  }
}

__global__ void stage_global(uint32_t* __restrict__ data, uint32_t log_n, uint32_t lh,
                             uint32_t polys, uint32_t stride, const uint32_t* __restrict__ tw) {
  const size_t half_n = (size_t)1 << (log_n - 1), total = half_n * polys;
  const uint32_t h = 1u << lh;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t p = t >> (log_n - 1);
    const uint32_t i = (uint32_t)(t & (half_n - 1));
    const uint32_t j = i & (h - 1);
    uint32_t* poly = data + p * stride;
    const uint32_t base = ((i >> lh) << (lh + 1)) + j;
    const uint32_t u = poly[base];
    const uint32_t v = mont_mul(poly[base + h], tw[h + j]);
    poly[base] = add_mod(u, v);
    poly[base + h] = sub_mod(u, v);
  }
}

__global__ void scale_global(uint32_t* __restrict__ data, uint32_t log_n, uint32_t polys,
                             uint32_t stride, uint32_t scale_mont) {
  const size_t n = (size_t)1 << log_n, total = n * polys;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    uint32_t* v = data + (t >> log_n) * stride + (t & (n - 1));
    *v = mont_mul(*v, scale_mont);
  }
}

// ---- FP64 digit passes (standard roots, n = 2^2..2^12) ---------------------------

// Persistent: each work-group copies the plan's twiddle tables into LDS once,
// then walks tiles of 4096 values (4096 / n polys): load, P passes, store.
// From n = 512 the next tile's values are loaded into registers while the
// passes run (sixteen a thread): 3-6 % faster at n = 1024 and 4096 (2^30
// values, alternating runs), within noise or slower at n = 256.
// Loads and stores are raw-buffer accesses through a descriptor per tile
// whose record count ends at its last live poly: the lanes of a ragged last
// tile past it load zeros and store nothing, with no branch around the
// access (a branch leaves the compiler's count of outstanding accesses
// conservative, and each tile then waited for the previous one's stores
// before its prefetched values could go to LDS).  Every byte offset stays
// below 2^32 (ntt_fp64_fits).
template <int L, bool INV>
__global__ __launch_bounds__(nttf::NT) void ntt_poly_fp64(uint32_t* __restrict__ data, uint32_t polys,
                                                          uint32_t stride, const uint32_t* __restrict__ twg) {
  using PL = nttf::Plan<L>;
  __shared__ int32_t lds[PL::LDS_WORDS];
  uint32_t* tw = reinterpret_cast<uint32_t*>(lds + nttf::LDS_DATA);
  const int tid = (int)threadIdx.x;
  for (int i = tid; i < PL::TW_WORDS; i += nttf::NT) tw[i] = twg[i];
  const slice::MulK k = slice::make_mulk();
  const uint32_t ntiles = (polys + PL::POLYS - 1) / PL::POLYS;
  constexpr bool PREFETCH = L >= 9;
  constexpr int CH = nttf::TILE / 4 / nttf::NT;
  const uint32_t lane = 4u * nttf::chunk_offset<PL>(4u * tid, stride);
  // the tile's buffer: its first poly, records up to the end of its last live one
  auto tile_buf = [&](uint32_t t) -> GBuf {
    const uint32_t first = t < ntiles ? t * (uint32_t)PL::POLYS : 0u;
    const uint32_t live = t < ntiles ? min((uint32_t)PL::POLYS, polys - first) : 0u;
    return GBuf(data + (size_t)first * stride, live ? 4u * ((live - 1) * stride + (uint32_t)PL::N) : 0u);
  };
  auto qoff = [&](int q) -> uint32_t { return 4u * nttf::chunk_offset<PL>(4u * nttf::NT * q, stride); };
  U4 pre[CH];
  auto fetch = [&](uint32_t t) {
    const GBuf src = tile_buf(t);
#pragma unroll
    for (int q = 0; q < CH; q++) pre[q] = src.load4(lane, qoff(q));
  };
  if constexpr (PREFETCH) {
    fetch(blockIdx.x);
    // Stores into a zero-record buffer (they store nothing): entering the
    // loop, as around it, the prefetch is then followed by CH stores, so the
    // wait before put_chunk can leave the previous tile's stores outstanding
    // (the compiler takes the smaller count where the two paths meet).
    const GBuf none(data, 0u);
#pragma unroll
    for (int q = 0; q < CH; q++) none.store4_cached(U4{0, 0, 0, 0}, lane, qoff(q));
  }
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GBuf buf = tile_buf(t);
    __syncthreads();  // the tables, or the previous tile's last reads of LDS
    if constexpr (PREFETCH) {
#pragma unroll
      for (int q = 0; q < CH; q++) nttf::put_chunk(tid, q, lds, pre[q]);
    } else {
#pragma unroll
      for (int q = 0; q < CH; q++) nttf::put_chunk(tid, q, lds, buf.load4(lane, qoff(q)));
    }
    __syncthreads();
    if constexpr (PREFETCH) fetch(t + gridDim.x);
    nttf::pass<PL, 0, INV>(tid, lds, tw, k);
    __syncthreads();
    if constexpr (PL::P > 1) {
      nttf::pass<PL, 1, INV>(tid, lds, tw, k);
      __syncthreads();
    }
    if constexpr (PL::P > 2) {
      nttf::pass<PL, 2, INV>(tid, lds, tw, k);
      __syncthreads();
    }
    // the natural-order values of chunk q out of their slots (store_tile's
    // order), every lane storing
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int v = 4 * (tid + q * nttf::NT), pb = (v >> PL::L) << PL::L, kq = v & (PL::N - 1);
      const U4 x{(uint32_t)lds[nttf::phys(pb + PL::slot_of(kq))], (uint32_t)lds[nttf::phys(pb + PL::slot_of(kq + 1))],
                 (uint32_t)lds[nttf::phys(pb + PL::slot_of(kq + 2))], (uint32_t)lds[nttf::phys(pb + PL::slot_of(kq + 3))]};
      buf.store4_cached(x, lane, qoff(q));
    }
  }
}

template <int L, bool INV>
static int launch_fp64_plan(DeviceState* st, uint32_t* d_values, uint32_t polys, uint32_t stride,
                            hipStream_t stream) {
  using PL = nttf::Plan<L>;
  const uint32_t key = ((uint32_t)L << 1) | (INV ? 1u : 0u);
  const uint32_t* twg = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lock(st->mu);
    auto it = st->ntt_fp64.find(key);
    if (it == st->ntt_fp64.end()) {
      std::vector<uint32_t> host(std::max(1, PL::TW_WORDS));
      nttf::build_ntt_fp64_tables<PL>(INV, host.data());
      void* d = nullptr;
      if (const hipError_t ea = hipMalloc(&d, host.size() * 4)) return alloc_fail(ea, "FP64 NTT tables");
      const hipError_t e = hipMemcpy(d, host.data(), host.size() * 4, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(d);
        return fail((int)e, "copying FP64 NTT tables failed: %s", hipGetErrorString(e));
      }
      it = st->ntt_fp64.emplace(key, static_cast<uint32_t*>(d)).first;
    }
    twg = it->second;
  }
  const void* fn = (const void*)ntt_poly_fp64<L, INV>;
  uint32_t grid = 0;
  ZODA_TRY(kernel_setup(st, fn, nttf::NT, 0, true, &grid));
  const uint32_t ntiles = (polys + PL::POLYS - 1) / PL::POLYS;
  if (grid > ntiles) grid = ntiles;
  void* args[] = {(void*)&d_values, (void*)&polys, (void*)&stride, (void*)&twg};
  ZODA_HIP(hipLaunchKernel(fn, dim3(grid), dim3(nttf::NT), args, 0, stream));
  return 0;
}

template <int L>
static int launch_fp64_size(DeviceState* st, uint32_t* d, uint32_t log_n, uint32_t polys, uint32_t stride,
                            int inverse, hipStream_t stream) {
  if constexpr (L <= 12) {
    if (log_n == (uint32_t)L)
      return inverse ? launch_fp64_plan<L, true>(st, d, polys, stride, stream)
                     : launch_fp64_plan<L, false>(st, d, polys, stride, stream);
    return launch_fp64_size<L + 1>(st, d, log_n, polys, stride, inverse, stream);
  } else {
    return fail(ZODA_EINVAL, "no FP64 plan for 2^%u", log_n);
  }
}

bool ntt_fp64_fits(const uint32_t* d_values, uint32_t log_n, uint32_t stride) {
  // a tile's byte offsets (4096 / n polys of `stride` u32) below 2^32
  return log_n >= 2 && log_n <= 12 && stride % 4 == 0 && ((uintptr_t)d_values & 15) == 0 &&
         ((uint64_t)(nttf::TILE >> log_n) * stride * 4u) < (1ull << 32);
}

int launch_ntt(DeviceState* st, uint32_t* d_values, uint32_t log_n, uint32_t polys, uint32_t stride,
               int inverse, const MontTables& tabs, bool standard_roots, hipStream_t stream) {
  if (polys == 0) return 0;
  if (standard_roots && ntt_fp64_fits(d_values, log_n, stride))
    return launch_fp64_size<2>(st, d_values, log_n, polys, stride, inverse, stream);
  return launch_ntt_poly(d_values, log_n, polys, stride, inverse, tabs, stream);
}

static unsigned grid_for(size_t work) {
  return (unsigned)std::min<size_t>((work + NTT_THREADS - 1) / NTT_THREADS, 1u << 16);
}

int launch_ntt_poly(uint32_t* d_values, uint32_t log_n, uint32_t polys, uint32_t stride,
                    int inverse, const MontTables& tabs, hipStream_t stream) {
  if (polys == 0) return 0;
  const uint32_t* tw = inverse ? tabs.inv : tabs.fwd;
  // n^-1 = (2^-1)^log_n, as cuda_intt computes it (cuda/ntt_kernel.cu:201-203)
  const uint32_t scale = to_mont_host(pow_mod_host((P + 1) / 2, log_n));
  if ((1u << log_n) <= (uint32_t)NTT_LDS_WORDS) {
    const uint32_t log_ppw = log2_ceil(NTT_LDS_WORDS) - log_n;
    const size_t blocks = (((size_t)polys) + (1u << log_ppw) - 1) >> log_ppw;
    if (inverse)
      hipLaunchKernelGGL(ntt_poly_lds<true>, dim3((unsigned)blocks), dim3(NTT_THREADS), 0, stream,
                         d_values, log_n, log_ppw, polys, stride, tw, scale);
    else
      hipLaunchKernelGGL(ntt_poly_lds<false>, dim3((unsigned)blocks), dim3(NTT_THREADS), 0, stream,
                         d_values, log_n, log_ppw, polys, stride, tw, scale);
    ZODA_HIP(hipGetLastError());
    return 0;
  }
  const size_t n = (size_t)1 << log_n;
  hipLaunchKernelGGL(bitrev_global, dim3(grid_for(n * polys)), dim3(NTT_THREADS), 0, stream,
                     d_values, log_n, polys, stride);
  ZODA_HIP(hipGetLastError());
  for (uint32_t lh = 0; lh < log_n; lh++) {
    hipLaunchKernelGGL(stage_global, dim3(grid_for(n / 2 * polys)), dim3(NTT_THREADS), 0, stream,
                       d_values, log_n, lh, polys, stride, tw);
    ZODA_HIP(hipGetLastError());
  }
  if (inverse) {
    hipLaunchKernelGGL(scale_global, dim3(grid_for(n * polys)), dim3(NTT_THREADS), 0, stream,
                       d_values, log_n, polys, stride, scale);
    ZODA_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace zoda
