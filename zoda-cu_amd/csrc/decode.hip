// Erasure decoder: the data square back from any k of the 2k rows of its
// extended square.  The reference has no decoder; what it relies on is that
// every column is a Reed-Solomon codeword -- the evaluations at the 2k-th
// roots of unity w^r of a polynomial f of degree < k, with the data at the
// even rows (src/zoda/babybear.rs:195-211).
//
// Per column, with K the k known rows and E the others (k of them), n = 2k:
//   Z(x) = prod_{e in E} (x - w^e)                  vanishes on the erasures
//   y_r  = out[r] * Z(w^r) on K, 0 on E              = g(w^r), g = f * Z, deg g < n
//
// For 2^8 <= n <= 2^14 (dec_columns, FP64, decode_core.hpp), from the formal
// derivative: g' = f' Z + f Z', and Z(w^e) = 0, so at an erased row
//   f(w^e) = w^e g'(w^e) / (w^e Z'(w^e)),   w^r g'(w^r) = NTT_n(i * G_i)_r,  G = INTT_n(y)
// -- one transform pair, of which only the even rows are computed; the known
// data rows are copies of the input.  For other n (dec_fused), on a coset
// where Z has no zero (h = 31):
//   g(h w^r) = NTT_n(G_i h^i),  f(h w^r) = g(h w^r) / Z(h w^r)
//   f_i  = INTT_n(f(h w^r))_i * h^-i, i < k          (the upper k are zero)
//   data[j] = f(w^(2j)) = NTT_k(f)_j
//
// Device work, all in canonical u32 with Montgomery multiplications:
//   dec_prep    Z(w^r) on K, and 1/(n w^e Z'(w^e)) on the erased even rows
//               (FP64 path) or 1/Z(h w^r) for every r (n threads, k factors
//               each): the only O(k^2) part, once per call.
//   dec_gather_ranked  (FP64 path) the known rows packed by rank into a
//               column-major scratch, times Z(w^r); the known even rows also
//               straight into the data.
//   dec_gather  (other n) row-major square -> full-height column-major
//               scratch, times Z(w^r) (erased rows are never read).
//   dec_fused   one work-group per column, in LDS: INTT_n (decimation in time,
//               bit-reversed in, natural out), * h^i/n, NTT_n (decimation in
//               frequency, natural in, bit-reversed out), * 1/Z(h w^r) in the
//               same bit-reversed order, INTT_n, * h^-i/n, NTT_k -- no
//               permutation passes, one read and one write of the scratch.
//   dec_scatter column-major -> row-major data (tiled transpose; on the FP64
//               path only the erased even rows).
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "decode_core.hpp"

namespace zoda {

namespace {

constexpr uint32_t DEC_COSET = 31;  // a generator: h w^r is never a 2k-th root of unity
constexpr uint32_t DEC_MAX_K = 16384;  // n = 2k values of one column in LDS: <= 128 KiB
constexpr int TT = 32;              // transpose tile
constexpr int DEC_NT = 256;
constexpr int GATHER_NY = 2;  // dec_gather_ranked blocks are TT x GATHER_NY threads
constexpr int SCATTER_NY = 2;  // dec_scatter blocks are TT x SCATTER_NY threads

constexpr uint32_t MONT_ONE = (uint32_t)((1ull << 32) % P);  // 1 in Montgomery form

__device__ __forceinline__ uint32_t brev_n(uint32_t i, uint32_t log_n) {
  return log_n ? (__brev(i) >> (32 - log_n)) : 0u;
}

// Montgomery form a^e.
__device__ uint32_t mont_pow(uint32_t a, uint32_t e) {
  uint32_t r = MONT_ONE;
  while (e) {
    if (e & 1) r = mont_mul(r, a);
    a = mont_mul(a, a);
    e >>= 1;
  }
  return r;
}

// prod_j (x - ep_m[j]) in Montgomery form over the ne erased points (SKIP:
// leaving out a point equal to x).  Sixteen independent partial products: one
// chain of ne dependent Montgomery products per thread left this kernel
// latency-bound (164 us at n = 4096 with 2048 erasures).
template <bool SKIP>
__device__ uint32_t prod_diffs(uint32_t x, const uint32_t* __restrict__ ep_m, uint32_t ne) {
  constexpr int NA = 16;
  uint32_t a[NA];
#pragma unroll
  for (int q = 0; q < NA; q++) a[q] = MONT_ONE;
  auto term = [&](uint32_t e) { return (SKIP && e == x) ? MONT_ONE : sub_mod(x, e); };
  uint32_t j = 0;
  for (; j + NA <= ne; j += NA) {
#pragma unroll
    for (int q = 0; q < NA; q++) a[q] = mont_mul(a[q], term(ep_m[j + q]));
  }
  for (; j < ne; j++) a[0] = mont_mul(a[0], term(ep_m[j]));
#pragma unroll
  for (int q = 1; q < NA; q++) a[0] = mont_mul(a[0], a[q]);
  return a[0];
}

// x_m[r] = w^r in Montgomery form; ep_m: the erased points, Montgomery form.
// zk_m[r] = Z(w^r) (0 on erased rows).  FP64 path (`deriv`): srow[r] =
// 1 / (n w^r Z'(w^r)) on the erased even rows, Z'(w^e) = prod over the other
// erasures of (w^e - w^e'), and 0 on the other rows (canonical, n_m = n in
// Montgomery form).  Otherwise srow[i] = 1 / Z(h w^(brev(i))) in Montgomery
// form (the coset path).
__global__ void dec_prep(const uint32_t* __restrict__ x_m, const uint8_t* __restrict__ known,
                         const uint32_t* __restrict__ ep_m, uint32_t ne, uint32_t log_n, uint32_t h_m,
                         int deriv, uint32_t n_m, uint32_t* __restrict__ zk_m, uint32_t* __restrict__ srow) {
  const uint32_t n = 1u << log_n;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n) return;
  if (t < n) {  // Z on the known rows
    if (!known[t]) {
      zk_m[t] = 0;
      return;
    }
    zk_m[t] = prod_diffs<false>(x_m[t], ep_m, ne);
    return;
  }
  if (deriv) {
    const uint32_t r = t - n;
    if (known[r] || (r & 1u)) {
      srow[r] = 0;
      return;
    }
    const uint32_t x = x_m[r];
    const uint32_t acc = mont_mul(mont_mul(n_m, x), prod_diffs<true>(x, ep_m, ne));  // n w^r Z'(w^r)
    srow[r] = mont_mul(mont_pow(acc, P - 2), 1u);  // mont_mul by 1 leaves Montgomery form
    return;
  }
  const uint32_t r = brev_n(t - n, log_n);
  const uint32_t x = mont_mul(h_m, x_m[r]);
  srow[t - n] = mont_pow(prod_diffs<false>(x, ep_m, ne), P - 2);
}

// T[c * n + r] = enc[r * cols + c] * Z(w^r) (a 32 x 32 tile per block).  The
// erased rows are never read and are written as 0 (whole 128-byte segments
// write faster than the known rows alone).
__global__ void dec_gather(const uint32_t* __restrict__ enc, uint32_t cols, uint32_t n,
                           const uint32_t* __restrict__ zk_m, uint32_t* __restrict__ T) {
  __shared__ uint32_t tile[TT][TT + 1];
  const uint32_t c0 = blockIdx.x * TT, r0 = blockIdx.y * TT;
  const uint32_t tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (uint32_t i = ty; i < TT; i += 8) {
    const uint32_t r = r0 + i, c = c0 + tx;
    uint32_t v = 0;
    if (r < n && c < cols) {
      const uint32_t z = zk_m[r];
      v = z ? mont_mul(enc[(size_t)r * cols + c], z) : 0u;  // erased rows are not read
    }
    tile[i][tx] = v;
  }
  __syncthreads();
  for (uint32_t i = ty; i < TT; i += 8) {
    const uint32_t c = c0 + i, r = r0 + tx;
    if (c < cols && r < n) T[(size_t)c * n + r] = tile[tx][i];
  }
}

// The known rows only, packed by rank: T[c * k + i] = enc[rows[i] * cols + c]
// * Z(w^rows[i]) for the k known rows rows[0..k) (a 32 x 32 tile of (rank,
// column) per block: whole 128-byte segments both ways); a known even row 2j
// is also the data row j, written as read.  The block's row
// numbers and factors are read once into LDS, so every value costs one global
// load, and each thread moves 32 / GATHER_NY values (at 2048^2: 1.73 ms with
// per-value row and factor loads and 4 values per thread, 1.09 ms so).
__global__ void dec_gather_ranked(const uint32_t* __restrict__ enc, uint32_t cols, uint32_t k,
                                  const uint32_t* __restrict__ rows, const uint32_t* __restrict__ zk_m,
                                  uint32_t* __restrict__ T, uint32_t* __restrict__ data) {
  __shared__ uint32_t tile[TT][TT + 1];
  __shared__ uint32_t rid[TT], zr[TT];  // the block's 32 row numbers and their Z(w^r)
  const uint32_t c0 = blockIdx.x * TT, i0 = blockIdx.y * TT;
  const uint32_t tx = threadIdx.x, ty = threadIdx.y;  // 32 x GATHER_NY
  if (ty == 0) {
    const uint32_t i = i0 + tx;
    const uint32_t r = i < k ? rows[i] : 0u;
    rid[tx] = r;
    zr[tx] = i < k ? zk_m[r] : 0u;
  }
  __syncthreads();
  for (uint32_t q = ty; q < TT; q += GATHER_NY) {
    const uint32_t i = i0 + q, c = c0 + tx;
    uint32_t v = 0;
    if (i < k && c < cols) {
      const uint32_t r = rid[q], e = enc[(size_t)r * cols + c];
      v = mont_mul(e, zr[q]);
      if (!(r & 1u)) data[(size_t)(r >> 1) * cols + c] = e;  // a known data row
    }
    tile[q][tx] = v;
  }
  __syncthreads();
  for (uint32_t q = ty; q < TT; q += GATHER_NY) {
    const uint32_t c = c0 + q, i = i0 + tx;
    if (c < cols && i < k) T[(size_t)c * k + i] = tile[tx][q];
  }
}

// Radix-2 decimation in time over v[0, 1 << log_n): bit-reversed in, natural
// out; tw[h + j] = (root of order 2h)^j in Montgomery form.
__device__ __forceinline__ void dit(uint32_t* v, uint32_t log_n, const uint32_t* __restrict__ tw) {
  const uint32_t half_n = (1u << log_n) >> 1;
  for (uint32_t lh = 0; lh < log_n; lh++) {
    const uint32_t h = 1u << lh;
    for (uint32_t b = threadIdx.x; b < half_n; b += DEC_NT) {
      const uint32_t j = b & (h - 1), base = ((b >> lh) << (lh + 1)) + j;
      const uint32_t u = v[base], w = mont_mul(v[base + h], tw[h + j]);
      v[base] = add_mod(u, w);
      v[base + h] = sub_mod(u, w);
    }
    __syncthreads();
  }
}

// Radix-2 decimation in frequency: natural in, bit-reversed out.
__device__ __forceinline__ void dif(uint32_t* v, uint32_t log_n, const uint32_t* __restrict__ tw) {
  const uint32_t half_n = (1u << log_n) >> 1;
  for (uint32_t lh = log_n; lh-- > 0;) {
    const uint32_t h = 1u << lh;
    for (uint32_t b = threadIdx.x; b < half_n; b += DEC_NT) {
      const uint32_t j = b & (h - 1), base = ((b >> lh) << (lh + 1)) + j;
      const uint32_t u = v[base], w = v[base + h];
      v[base] = add_mod(u, w);
      v[base + h] = mont_mul(sub_mod(u, w), tw[h + j]);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void scale(uint32_t* v, uint32_t len, const uint32_t* __restrict__ m) {
  for (uint32_t i = threadIdx.x; i < len; i += DEC_NT) v[i] = mont_mul(v[i], m[i]);
  __syncthreads();
}

// One work-group per column of T (n values); leaves the k data values of the
// column, in order, at T[c * n + j].
__global__ __launch_bounds__(DEC_NT) void dec_fused(uint32_t* __restrict__ T, uint32_t log_n,
                                                    const uint32_t* __restrict__ fwd_n,
                                                    const uint32_t* __restrict__ inv_n,
                                                    const uint32_t* __restrict__ fwd_k,
                                                    const uint32_t* __restrict__ hpow_m,
                                                    const uint32_t* __restrict__ zinv_br_m,
                                                    const uint32_t* __restrict__ hinv_m) {
  extern __shared__ uint32_t v[];
  const uint32_t n = 1u << log_n, k = n >> 1;
  uint32_t* col = T + (size_t)blockIdx.x * n;
  for (uint32_t i = threadIdx.x; i < n; i += DEC_NT) v[brev_n(i, log_n)] = col[i];
  __syncthreads();
  dit(v, log_n, inv_n);            // n * G, natural order
  scale(v, n, hpow_m);             // * h^i / n
  dif(v, log_n, fwd_n);            // g(h w^r) at position brev(r)
  scale(v, n, zinv_br_m);          // f(h w^r), same positions
  dit(v, log_n, inv_n);            // n * f_i h^i, natural order
  scale(v, k, hinv_m);             // f_i, i < k
  dif(v, log_n - 1, fwd_k);        // f(w_k^j) at position brev_k(j)
  for (uint32_t j = threadIdx.x; j < k; j += DEC_NT) col[j] = v[brev_n(j, log_n - 1)];
}

// Persistent form for 2^8 <= n <= 2^14: tiles of W columns of the
// rank-packed scratch (k values each) through the FP64 transform pair in LDS
// (decode_core.hpp); the k even-row values of each column (the data on the
// erased rows, 0 on the known ones) replace its entries.
template <class PL>
__global__ __launch_bounds__(PL::NT, PL::MINW) void dec_columns(uint32_t* __restrict__ T, uint32_t cols, uint32_t ntiles,
                                                      const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ srow, DecodeTabs tabs) {
  extern __shared__ int32_t lds[];
  dec_load_tables<PL>((int)threadIdx.x, lds, tabs);
  const slice::MulK mk = slice::make_mulk();
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    // A thread's addresses are the same for every tile.  Making its index
    // depend on the tile (through an asm the optimizer cannot see through,
    // not volatile: no side effects) keeps that arithmetic in the loop
    // instead of hoisted out of it and held -- or spilled -- for the whole
    // kernel.
    int tid = (int)threadIdx.x;
    asm("" : "+v"(tid) : "s"(tile));
    dec_load<PL>(tid, tile, cols, T, rows, lds);
    __syncthreads();  // (on the first tile also the tables)
    dec_step<PL, 0>(tid, lds, tabs, srow, mk);
    __syncthreads();
    dec_step<PL, 1>(tid, lds, tabs, srow, mk);
    __syncthreads();
    dec_step<PL, 2>(tid, lds, tabs, srow, mk);
    __syncthreads();
    dec_step<PL, 3>(tid, lds, tabs, srow, mk);
    __syncthreads();
    dec_step<PL, 4>(tid, lds, tabs, srow, mk);
    __syncthreads();
    dec_store<PL>(tid, tile, cols, T, lds);
    __syncthreads();  // the next tile's load overwrites LDS
  }
}

// ---- chunked form of the FP64 path (whole 64-column blocks) -------------------
//
// The columns go in chunks through three scratch buffers; launch i gathers
// chunk i, runs the column kernel on chunk i - 1 and scatters chunk i - 2, so
// the memory-bound transposes run beside the VALU-bound column work instead
// of in launches of their own.  A gather block is 64 known ranks x 64
// columns, a scatter block 64 data rows x 64 columns, each through LDS (the
// tile's words, 64 x 65); item b of a launch is gather block b, column tile b
// and scatter block b, whichever exist.

constexpr int DB = 64, DB_PITCH = 65;  // transpose block of the chunked path

struct DecChunks {
  const uint32_t* enc;  // extended square, row-major (cols a row)
  uint32_t* data;       // data square, row-major
  uint32_t cols, k;
  const uint32_t* rows; // known rows by rank, then the erased rows
  const uint32_t* zk;   // Z(w^r), Montgomery form
  const uint8_t* known; // known[r]
  uint32_t* tg;         // gather: scratch of chunk g (null: none), its columns and first column
  uint32_t gcols, gc0;
  uint32_t* tc;         // column kernel: scratch of chunk c (null: none), its columns
  uint32_t ccols;
  const uint32_t* ts;   // scatter: scratch of chunk s (null: none), its columns and first column
  uint32_t scols, sc0;
};

template <class PL>
__global__ __launch_bounds__(PL::NT, PL::MINW) void dec_chunks(DecChunks a, const uint32_t* __restrict__ srow,
                                                               DecodeTabs tabs) {
  static_assert(DB * DB_PITCH <= PL::TILE_WORDS, "a transpose block fits the tile's LDS");
  constexpr uint32_t K = PL::S / 2, KB = K / DB;  // rank (or data-row) blocks per column band
  constexpr int GI = (DB * DB / 4 + PL::NT - 1) / PL::NT;  // 16-byte runs of a block per thread
  extern __shared__ int32_t lds[];
  uint32_t* sh = reinterpret_cast<uint32_t*>(lds);
  __shared__ uint32_t rid[2][DB], zr[2][DB];  // a gather block's row numbers and Z(w^r), two items
  const int tid = (int)threadIdx.x;
  dec_load_tables<PL>(tid, lds, tabs);
  const slice::MulK mk = slice::make_mulk();
  const uint32_t ng = a.tg ? KB * (a.gcols / DB) : 0u;
  const uint32_t ntile = a.tc ? a.ccols / PL::W : 0u;
  const uint32_t ns = a.ts ? KB * (a.scols / DB) : 0u;
  const uint32_t nitems = max(ng, max(ntile, ns));
  // Item b's gather and scatter loads are issued one item ahead (before the
  // passes of the column tile in between), so their latency hides behind
  // arithmetic; an item past the end of its kind loads harmlessly (the first
  // row of the square; a zero-record buffer) instead of branching.
  U4 gv[GI], sv[GI];
  auto stage = [&](uint32_t b, int buf) {  // the block's row numbers into LDS
    if (tid < DB) {
      const bool live = b < ng;
      const uint32_t r = live ? a.rows[(b % KB) * DB + tid] : 0u;
      rid[buf][tid] = r;
      zr[buf][tid] = live ? mont_mul(a.zk[r], 1u) : 0u;  // Z(w^r) out of Montgomery form
    }
  };
  auto fetch = [&](uint32_t b, int buf) {  // after a barrier that follows stage(b, buf)
    const bool gl = b < ng;
    const size_t gcol = gl ? (size_t)a.gc0 + (b / KB) * DB : 0u;
#pragma unroll
    for (int g = 0; g < GI; g++) {
      const int idx = tid + g * PL::NT;
      const uint32_t q = (idx / 16) % DB, c4 = gl ? 4 * (idx % 16) : 0u;
      gv[g] = *reinterpret_cast<const U4*>(a.enc + (size_t)rid[buf][q] * a.cols + gcol + c4);
    }
    const GBuf src(a.ts, b < ns);
    const uint32_t j0 = (b % KB) * DB, cb = (b / KB) * DB;
#pragma unroll
    for (int g = 0; g < GI; g++) {
      const int idx = tid + g * PL::NT;
      const uint32_t c = (idx / 16) % DB, j4 = 4 * (idx % 16);
      sv[g] = src.load4(4u * ((cb + c) * K + j0 + j4), 0u);
    }
  };
  stage(blockIdx.x, 0);
  __syncthreads();
  fetch(blockIdx.x, 0);
  int buf = 0;
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x, buf ^= 1) {
    const uint32_t next = it + gridDim.x;
    if (it < ng) {  // gather: ranks i0.., columns cb.. of chunk g
      const uint32_t i0 = (it % KB) * DB, cb = (it / KB) * DB;
      __syncthreads();  // the previous item's last reads of LDS
#pragma unroll
      for (int g = 0; g < GI; g++) {
        const int idx = tid + g * PL::NT;
        if (idx >= DB * DB / 4) break;
        const uint32_t q = idx / 16, c4 = 4 * (idx % 16), r = rid[buf][q];
        const U4 e = gv[g];
        if (!(r & 1u)) store16(a.data + (size_t)(r >> 1) * a.cols + a.gc0 + cb + c4, e);  // a known data row
        // y_r = value * Z(w^r) as the int32 the column kernel's LDS takes (|y| < 2^30,
        // not canonical): the four-instruction FP64 product in place of a
        // Montgomery product (four 32-bit multiplies and the 64-bit additions
        // and the final correction around them)
        const double zw = (double)zr[buf][q], zq = zw * PINV;
        uint32_t* d = sh + q * DB_PITCH + c4;
        d[0] = (uint32_t)slice::mul_mem((double)e.x, zw, zq, mk);
        d[1] = (uint32_t)slice::mul_mem((double)e.y, zw, zq, mk);
        d[2] = (uint32_t)slice::mul_mem((double)e.z, zw, zq, mk);
        d[3] = (uint32_t)slice::mul_mem((double)e.w, zw, zq, mk);
      }
      __syncthreads();
      for (int idx = tid; idx < DB * DB / 4; idx += PL::NT) {
        const uint32_t c = idx / 16, q4 = 4 * (idx % 16);
        const uint32_t* s = sh + q4 * DB_PITCH + c;
        store16(a.tg + (size_t)(cb + c) * K + i0 + q4, U4{s[0], s[DB_PITCH], s[2 * DB_PITCH], s[3 * DB_PITCH]});
      }
    }
    if (it < ns) {  // scatter: data rows j0.., columns cb.. of chunk s (the erased ones)
      const uint32_t j0 = (it % KB) * DB, cb = (it / KB) * DB;
      __syncthreads();
#pragma unroll
      for (int g = 0; g < GI; g++) {
        const int idx = tid + g * PL::NT;
        if (idx >= DB * DB / 4) break;
        const uint32_t c = idx / 16, j4 = 4 * (idx % 16);
        uint32_t* d = sh + j4 * DB_PITCH + c;
        d[0] = sv[g].x;
        d[DB_PITCH] = sv[g].y;
        d[2 * DB_PITCH] = sv[g].z;
        d[3 * DB_PITCH] = sv[g].w;
      }
      __syncthreads();
      for (int idx = tid; idx < DB * DB / 4; idx += PL::NT) {
        const uint32_t j = idx / 16, c4 = 4 * (idx % 16);
        if (a.known[2 * (j0 + j)]) continue;  // written by the gather
        const uint32_t* s = sh + j * DB_PITCH + c4;
        store16(a.data + (size_t)(j0 + j) * a.cols + a.sc0 + cb + c4, U4{s[0], s[1], s[2], s[3]});
      }
    }
    stage(next, buf ^ 1);  // read after the next barrier
    if (it < ntile) {  // the column kernel on tile it of chunk c
      int tid2 = tid;
      asm("" : "+v"(tid2) : "s"(it));  // keep the per-thread address arithmetic in the loop
      __syncthreads();  // the tables, or the previous reads of LDS
      dec_load<PL>(tid2, it, a.ccols, a.tc, a.rows, lds);
      __syncthreads();
      fetch(next, buf ^ 1);  // arrives during the passes
      dec_step<PL, 0>(tid2, lds, tabs, srow, mk);
      __syncthreads();
      dec_step<PL, 1>(tid2, lds, tabs, srow, mk);
      __syncthreads();
      dec_step<PL, 2>(tid2, lds, tabs, srow, mk);
      __syncthreads();
      dec_step<PL, 3>(tid2, lds, tabs, srow, mk);
      __syncthreads();
      dec_step<PL, 4>(tid2, lds, tabs, srow, mk);
      __syncthreads();
      dec_store<PL>(tid2, it, a.ccols, a.tc, lds);
    } else {
      __syncthreads();
      fetch(next, buf ^ 1);
    }
  }
}

// Tables of the FP64 decoder plan PL, built and uploaded once per device.
template <class PL>
int get_decode_tables(DeviceState* st, const DecodeTabs** out) {
  constexpr int L = PL::L;
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  auto it = st->decode.find((uint32_t)L);
  if (it == st->decode.end()) {
    const DecodeTableImage im = build_decode_tables<PL>();
    if (im.packed.empty()) return fail(ZODA_ESTATE, "decoder table layout mismatch for n = 2^%d", L);
    DecodeTableSet set;
    if (const hipError_t ea = hipMalloc((void**)&set.packed, im.packed.size() * sizeof(uint32_t)))
      return alloc_fail(ea, "decoder table allocation");
    const hipError_t e =
        hipMemcpy(set.packed, im.packed.data(), im.packed.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(set.packed);
      return fail((int)e, "copying decoder tables failed: %s", hipGetErrorString(e));
    }
    set.tabs = new DecodeTabs(make_decode_tabs(im, set.packed));
    it = st->decode.emplace((uint32_t)L, set).first;
  }
  *out = it->second.tabs;
  return 0;
}

// Once per device: the LDS limit of dec_columns<PL> and its persistent grid
// (every CU filled).
template <class PL>
int decode_grid(DeviceState* st, uint32_t* grid) {
  const void* fn = (const void*)dec_columns<PL>;
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  auto it = st->kernel_setup.find(fn);
  if (it == st->kernel_setup.end()) {
    constexpr size_t lds = PL::LDS_WORDS * sizeof(int32_t);
    if (lds > 64 * 1024) ZODA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, cus = 0;
    ZODA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, PL::NT, PL::LDS_WORDS * sizeof(int32_t)));
    ZODA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, st->device));
    it = st->kernel_setup.emplace(fn, (uint32_t)std::max(1, per_cu * cus)).first;
  }
  *grid = it->second;
  return 0;
}

// Once per device: the LDS limit of dec_chunks<PL> and its persistent grid.
template <class PL>
int decode_chunks_grid(DeviceState* st, uint32_t* grid) {
  const void* fn = (const void*)dec_chunks<PL>;
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  auto it = st->kernel_setup.find(fn);
  if (it == st->kernel_setup.end()) {
    constexpr size_t lds = PL::LDS_WORDS * sizeof(int32_t);
    if (lds > 64 * 1024) ZODA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, cus = 0;
    ZODA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, PL::NT, lds));
    ZODA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, st->device));
    it = st->kernel_setup.emplace(fn, (uint32_t)std::max(1, per_cu * cus)).first;
  }
  *grid = it->second;
  return 0;
}

// The chunked FP64 decode (dec_chunks): chunks of DEC_CHUNK_BYTES of scratch
// each, three buffers, nchunks + 2 launches.
constexpr size_t DEC_CHUNK_BYTES = 64u << 20;

template <class PL>
int decode_chunked(DeviceState* st, DecChunks base, const uint32_t* srow, const DecodeTabs& tabs, hipStream_t s) {
  constexpr uint32_t K = PL::S / 2;
  uint32_t grid = 0;
  ZODA_TRY(decode_chunks_grid<PL>(st, &grid));
  const uint32_t cols = base.cols;
  uint32_t chunk = (uint32_t)std::min<size_t>(cols, DEC_CHUNK_BYTES / (sizeof(uint32_t) * K));
  chunk -= chunk % DB;
  const uint32_t nchunks = (cols + chunk - 1) / chunk;
  const size_t buf = (size_t)K * chunk;
  uint32_t* T = nullptr;
  ZODA_TRY(pool_alloc(st, 3 * buf * sizeof(uint32_t), s, (void**)&T));
  int rc = 0;
  for (uint32_t i = 0; i < nchunks + 2 && !rc; i++) {
    DecChunks a = base;
    uint32_t items = 0;
    if (i < nchunks) {
      a.tg = T + (i % 3) * buf;
      a.gc0 = i * chunk;
      a.gcols = std::min(chunk, cols - a.gc0);
      items = std::max(items, (K / DB) * (a.gcols / DB));
    }
    if (i >= 1 && i <= nchunks) {
      a.tc = T + ((i - 1) % 3) * buf;
      a.ccols = std::min(chunk, cols - (i - 1) * chunk);
      items = std::max(items, a.ccols / (uint32_t)PL::W);
    }
    if (i >= 2) {
      a.ts = T + ((i - 2) % 3) * buf;
      a.sc0 = (i - 2) * chunk;
      a.scols = std::min(chunk, cols - a.sc0);
      items = std::max(items, (K / DB) * (a.scols / DB));
    }
    hipLaunchKernelGGL(dec_chunks<PL>, dim3(std::min(grid, items)), dim3(PL::NT), PL::LDS_WORDS * sizeof(int32_t), s,
                       a, srow, tabs);
    hip_note(rc, hipGetLastError(), "decode launch");
  }
  hip_note(rc, hipFreeAsync(T, s), "hipFreeAsync");
  return rc;
}

// data[j * cols + c] = T[c * stride + j] for j < k, except the rows j whose
// row 2j is known (`known` non-null: the FP64 path's gather wrote them).
__global__ void dec_scatter(const uint32_t* __restrict__ T, uint32_t stride, uint32_t k, uint32_t cols,
                            const uint8_t* __restrict__ known, uint32_t* __restrict__ data) {
  __shared__ uint32_t tile[TT][TT + 1];
  __shared__ uint8_t skip[TT];  // data row j0 + i already written
  const uint32_t j0 = blockIdx.x * TT, c0 = blockIdx.y * TT;
  const uint32_t tx = threadIdx.x, ty = threadIdx.y;
  if (ty == 0) skip[tx] = known && j0 + tx < k ? known[2 * (j0 + tx)] : 0u;
  __syncthreads();
  for (uint32_t i = ty; i < TT; i += SCATTER_NY) {
    const uint32_t c = c0 + i, j = j0 + tx;
    tile[i][tx] = (c < cols && j < k && !skip[tx]) ? T[(size_t)c * stride + j] : 0u;
  }
  __syncthreads();
  for (uint32_t i = ty; i < TT; i += SCATTER_NY) {
    const uint32_t j = j0 + i, c = c0 + tx;
    if (j < k && c < cols && !skip[i]) data[(size_t)j * cols + c] = tile[tx][i];
  }
}

}  // namespace

}  // namespace zoda

using namespace zoda;

extern "C" int zoda_hip_decode_device(const uint32_t* d_enc, uint32_t k, uint32_t cols, const uint32_t* known_rows,
                                      uint32_t nknown, uint32_t* d_data, void* stream) {
  clear_error();
  if (!d_enc || !d_data || !known_rows || cols == 0 || !is_pow2(k) || k > DEC_MAX_K)
    return fail(ZODA_EINVAL, "bad arguments (k must be a power of two <= %u)", DEC_MAX_K);
  const uint32_t n = 2 * k, log_n = log2_ceil(n);
  std::vector<uint8_t> known(n, 0);
  uint32_t have = 0;
  for (uint32_t i = 0; i < nknown && have < k; i++) {
    const uint32_t r = known_rows[i];
    if (r >= n) return fail(ZODA_EINVAL, "known row %u is outside the %u rows", r, n);
    if (!known[r]) {
      known[r] = 1;
      have++;
    }
  }
  if (have < k) return fail(ZODA_EINVAL, "%u distinct known rows, %u needed", have, k);
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  // 2^8 <= n <= 2^14: the FP64 column kernel; other sizes: the Montgomery one.
  const bool fp64 = (int)log_n >= DEC_FP64_MIN_LOG && (int)log_n <= DEC_FP64_MAX_LOG;
  MontTables tn, tk;
  const DecodeTabs* dtabs = nullptr;
  uint32_t dgrid = 0;
  if (fp64) {
    ZODA_TRY(with_decode_plan((int)log_n, [&](auto plan) -> int {
      using PL = decltype(plan);
      ZODA_TRY(get_decode_tables<PL>(st, &dtabs));
      return decode_grid<PL>(st, &dgrid);
    }));
  } else {
    ZODA_TRY(get_mont_tables(st, log_n, &tn));
    ZODA_TRY(get_mont_tables(st, log_n - 1, &tk));
  }

  // O(n) host vectors, Montgomery form: the points w^r, the erased points,
  // h^i / n and h^-i / n.
  const uint32_t w = root_of_unity(log_n);
  const uint32_t n_inv = inv_mod_host(n), h_inv = inv_mod_host(DEC_COSET);
  std::vector<uint32_t> x_m(n), ep_m, hpow_m(n), hinv_m(k);
  uint32_t x = 1, hp = n_inv, hq = n_inv;
  for (uint32_t r = 0; r < n; r++) {
    x_m[r] = to_mont_host(x);
    if (!known[r]) ep_m.push_back(x_m[r]);
    hpow_m[r] = to_mont_host(hp);
    if (r < k) hinv_m[r] = to_mont_host(hq);
    x = mul_mod_host(x, w);
    hp = mul_mod_host(hp, DEC_COSET);
    hq = mul_mod_host(hq, h_inv);
  }
  const uint32_t ne = (uint32_t)ep_m.size();  // = k
  std::vector<uint32_t> rows;  // the known rows, then the erased ones, each increasing
  rows.reserve(n);
  for (uint32_t r = 0; r < n; r++)
    if (known[r]) rows.push_back(r);
  for (uint32_t r = 0; r < n; r++)
    if (!known[r]) rows.push_back(r);

  // One device block: rows | x_m | hpow_m | hinv_m | ep_m | zk_m | zinv | known
  // (rows first: dec_columns reads it 16 bytes at a time).
  hipStream_t s = (hipStream_t)stream;
  const size_t up_words = (size_t)n + n + n + k + ne;
  const size_t words = up_words + n + n;
  uint32_t* d_blk = nullptr;
  uint32_t* T = nullptr;
  // stream-ordered, from the library's pool: no device-wide synchronisation
  // for the allocation or its release
  ZODA_TRY(pool_alloc(st, words * 4 + n, s, (void**)&d_blk));
  uint32_t* d_rows = d_blk;
  uint32_t* d_x = d_rows + n;
  uint32_t* d_hpow = d_x + n;
  uint32_t* d_hinv = d_hpow + n;
  uint32_t* d_ep = d_hinv + k;
  uint32_t* d_zk = d_ep + ne;
  uint32_t* d_zinv = d_zk + n;
  uint8_t* d_known = reinterpret_cast<uint8_t*>(d_zinv + n);
  std::vector<uint32_t> host;
  host.reserve(up_words);
  host.insert(host.end(), rows.begin(), rows.end());
  host.insert(host.end(), x_m.begin(), x_m.end());
  host.insert(host.end(), hpow_m.begin(), hpow_m.end());
  host.insert(host.end(), hinv_m.begin(), hinv_m.end());
  host.insert(host.end(), ep_m.begin(), ep_m.end());
  int rc = 0;
  // queued on the stream (the host vectors live until the synchronisation below)
  hip_note(rc, hipMemcpyAsync(d_blk, host.data(), up_words * 4, hipMemcpyHostToDevice, s), "decode vectors upload");
  if (!rc)
    hip_note(rc, hipMemcpyAsync(d_known, known.data(), n, hipMemcpyHostToDevice, s), "decode vectors upload");
  // Whole 64-column blocks of 16-byte aligned squares take the chunked FP64
  // path (dec_chunks, its own scratch); otherwise the scratch holds the k
  // known rows per column (FP64 path) or all n rows.
  const bool chunked = fp64 && cols % DB == 0 && !(((uintptr_t)d_enc | (uintptr_t)d_data) & 15);
  const uint32_t t_rows = fp64 ? k : n;
  if (!rc && !chunked) {
    rc = pool_alloc(st, (size_t)cols * t_rows * sizeof(uint32_t), s, (void**)&T);
  }
  const size_t lds = (size_t)n * sizeof(uint32_t);
  if (!rc && !fp64 && lds > 65536)
    hip_note(rc, hipFuncSetAttribute((const void*)dec_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
             "raising the decoder's LDS limit");
  if (!rc) {
    hipLaunchKernelGGL(dec_prep, dim3((2 * n + 255) / 256), dim3(256), 0, s, d_x, d_known, d_ep, ne, log_n,
                       to_mont_host(DEC_COSET), fp64 ? 1 : 0, to_mont_host(n), d_zk, d_zinv);
    if (chunked) {
      const DecChunks base{d_enc, d_data, cols, k, d_rows, d_zk, d_known, nullptr, 0, 0, nullptr, 0, nullptr, 0, 0};
      hip_note(rc, hipGetLastError(), "decode launch");
      if (!rc)
        rc = with_decode_plan((int)log_n, [&](auto plan) -> int {
          return decode_chunked<decltype(plan)>(st, base, d_zinv, *dtabs, s);
        });
    } else if (fp64) {
      hipLaunchKernelGGL(dec_gather_ranked, dim3((cols + TT - 1) / TT, (k + TT - 1) / TT), dim3(TT, GATHER_NY), 0, s, d_enc,
                         cols, k, d_rows, d_zk, T, d_data);
      (void)with_decode_plan((int)log_n, [&](auto plan) -> int {
        using PL = decltype(plan);
        const uint32_t ntiles = (cols + PL::W - 1) / PL::W;
        hipLaunchKernelGGL(dec_columns<PL>, dim3(std::min(dgrid, ntiles)), dim3(PL::NT),
                           PL::LDS_WORDS * sizeof(int32_t), s, T, cols, ntiles, d_rows, d_zinv, *dtabs);
        return 0;
      });
    } else {
      hipLaunchKernelGGL(dec_gather, dim3((cols + TT - 1) / TT, (n + TT - 1) / TT), dim3(TT, 8), 0, s, d_enc, cols,
                         n, d_zk, T);
      hipLaunchKernelGGL(dec_fused, dim3(cols), dim3(DEC_NT), lds, s, T, log_n, tn.fwd, tn.inv, tk.fwd, d_hpow,
                         d_zinv, d_hinv);
    }
    if (!chunked) {
      hipLaunchKernelGGL(dec_scatter, dim3((k + TT - 1) / TT, (cols + TT - 1) / TT), dim3(TT, SCATTER_NY), 0, s, T, t_rows,
                         k, cols, fp64 ? d_known : nullptr, d_data);
      hip_note(rc, hipGetLastError(), "decode launch");
    }
  }
  if (T) hip_note(rc, hipFreeAsync(T, s), "hipFreeAsync");
  hip_note(rc, hipFreeAsync(d_blk, s), "hipFreeAsync");  // after the kernels that read it
  hip_note(rc, hipStreamSynchronize(s), "decode");
  return rc;
}
