// Random linear combinations of the encoded rows on the device, and the
// commitment values they are checked against (src/zoda/rlc.rs,
// src/benchmarks/utils.rs:51-59,221-272, src/zoda/babybear.rs:164-192).
//
//   rlc(row)[l] = sum over columns c of row[c] * coeff[c][l]  (mod p), l < 4
//
// The row kernel streams the row-major square once.  A work-group owns a
// block of RLC_COLS columns and RLC_ROWS rows; the products run on the int8
// matrix cores over signed base-256 digits of the values and coefficients
// (the kernel's comment has the decomposition), so the vector ALUs only split
// digits, and each wave folds its int32 sums into four canonical limbs per
// row once per block.  A second kernel adds the column blocks' partial sums.
#include <string.h>

#include <vector>

#include "common.hpp"
#include "sha256.hpp"

namespace zoda {

namespace {

constexpr int RLC_NT = 256;
constexpr uint32_t RLC_COLS = 2048;  // columns per work-group (32 KiB of coefficient digits)
constexpr uint32_t RLC_ROWS = 64;    // rows per work-group (16 per wave)
constexpr int RLC_STEPS = RLC_COLS / 32;
constexpr uint64_t TWO32_MOD_P = 4294967296ull - 2ull * 2013265921ull;

// x < 2^64 -> x' = x (mod p) with x' < 2^61.
__device__ __forceinline__ uint64_t fold(uint64_t x) {
  return (x >> 32) * TWO32_MOD_P + (x & 0xffffffffull);
}

__device__ __forceinline__ uint32_t mod_p(uint64_t x) {
  x = fold(fold(x));  // < 2^57
  return (uint32_t)(x % 2013265921ull);
}

// SHA-256 state after a data root of whole 64-byte blocks (the same for every
// column).
struct ShaMid {
  uint32_t h[8];
};

// Roots of whole blocks: one compression per column from the shared state.
__global__ void rlc_coeff_mid_kernel(ShaMid mid, uint64_t total_bits, uint32_t cols, uint32_t* __restrict__ coeffs) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  uint32_t limb[4];
  rlc_coefficient_mid(mid.h, total_bits, c, limb);
#pragma unroll
  for (int l = 0; l < 4; l++) coeffs[(size_t)c * 4 + l] = limb[l];
}

// Other root lengths: the whole hash per column.
__global__ void rlc_coeff_kernel(const uint8_t* __restrict__ root, uint32_t root_len, uint32_t cols,
                                 uint32_t* __restrict__ coeffs) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  uint32_t limb[4];
  rlc_coefficient(root, root_len, c, limb);
#pragma unroll
  for (int l = 0; l < 4; l++) coeffs[(size_t)c * 4 + l] = limb[l];
}

// Any u32 -> its canonical value (every u32 is below 3p): two conditional
// subtractions, each keeping the smaller of v and v - kp (which wraps above v
// when v < kp).
__device__ __forceinline__ uint32_t canon_u32(uint32_t v) {
  v = min(v, v - 2u * 2013265921u);
  return min(v, v - 2013265921u);
}

// x < p -> a word whose bytes, read as int8, are the signed base-256 digits
// d0..d3 of x (x = sum d_j 256^j, -128 <= d_j < 128): adding 128 to each of the
// three low digits lets the carries settle, and flipping the top bit of those
// bytes takes the 128 off again in two's complement.  The top digit stays below
// 0x79 because x + 0x808080 < p + 0x808080 = 0x78808081.
__device__ __forceinline__ uint32_t signed_digits(uint32_t x) { return (x + 0x808080u) ^ 0x808080u; }

// o[j] = byte j of x0, x1, x2, x3 (in that order), by byte permutes: the
// selector picks bytes 0-3 from the second operand and 4-7 from the first.
__device__ __forceinline__ void byte_transpose(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3,
                                               uint32_t (&o)[4]) {
  const uint32_t t0 = __builtin_amdgcn_perm(x1, x0, 0x06020400u);  // x0.0 x1.0 x0.2 x1.2
  const uint32_t t1 = __builtin_amdgcn_perm(x1, x0, 0x07030501u);  // x0.1 x1.1 x0.3 x1.3
  const uint32_t t2 = __builtin_amdgcn_perm(x3, x2, 0x06020400u);
  const uint32_t t3 = __builtin_amdgcn_perm(x3, x2, 0x07030501u);
  o[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
  o[1] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
  o[2] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
  o[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

typedef int rlc_v4i __attribute__((ext_vector_type(4)));

// The matrix-core sums of one lane over a block's RLC_STEPS k-steps: src
// points at the lane's first column (c_first) of its row; FAST blocks are
// whole, with 16-byte aligned rows.  U steps' values are loaded before any of
// them is split, so that 8 U 16-byte loads per lane are in flight.
template <bool FAST>
__device__ __forceinline__ void rlc_block_sums(const uint32_t* __restrict__ src, uint32_t c_first,
                                               uint32_t cols, const uint2* __restrict__ bt, int lane,
                                               rlc_v4i (&acc)[4]) {
  constexpr int U = 4;
  for (int s0 = 0; s0 < RLC_STEPS; s0 += U) {
    uint32_t v[U][8];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int s = s0 + u;
      if constexpr (FAST) {
        const uint4 a = *reinterpret_cast<const uint4*>(src + 32 * s);
        const uint4 b = *reinterpret_cast<const uint4*>(src + 32 * s + 4);
        v[u][0] = a.x, v[u][1] = a.y, v[u][2] = a.z, v[u][3] = a.w;
        v[u][4] = b.x, v[u][5] = b.y, v[u][6] = b.z, v[u][7] = b.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; e++) v[u][e] = c_first + 32u * s + e < cols ? src[32 * s + e] : 0u;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
      for (int e = 0; e < 8; e++) v[u][e] = signed_digits(canon_u32(v[u][e]));
      uint32_t lo[4], hi[4];
      byte_transpose(v[u][0], v[u][1], v[u][2], v[u][3], lo);
      byte_transpose(v[u][4], v[u][5], v[u][6], v[u][7], hi);
      const uint2 cw = bt[(s0 + u) * 64 + lane];
      const long cb = (long)(((uint64_t)cw.y << 32) | cw.x);
#pragma unroll
      for (int i = 0; i < 4; i++)
        acc[i] = __builtin_amdgcn_mfma_i32_16x16x32_i8((long)(((uint64_t)hi[i] << 32) | lo[i]), cb, acc[i], 0, 0,
                                                       0);
    }
  }
}

// partial[(i * nblk + blk) * 4 + l] for the i-th requested row (row rowidx[i],
// or i without an index); a requested row outside the square contributes 0.
//
// The sum over the block's columns runs on the int8 matrix cores.  Values and
// coefficients are split into signed base-256 digits a_i and d_j, so that
//   sum_c v_c * coeff_c[l] = sum_{i,j} 256^(i+j) * sum_c a_i(v_c) * d_j(coeff_c[l])
// and the inner sums are the entries of four 16 x 16 int32 products per
// 16 rows: D_i[row][4l + j] = sum_c a_i[row][c] * d_j[c][l] (one per value
// digit i; |a d| <= 2^14, so 2048 columns stay below 2^25).  A k-step of the
// 16x16x32 instruction takes 32 columns: lane (m, g) = (lane % 16, lane / 16)
// holds digit i of columns 8g..8g+7 of row m (two 16-byte loads give the
// values, byte permutes the digit planes), and the coefficient operand holds
// digit j of coefficient l of the same 8 columns for n = 4l + j, built once
// per work-group into LDS in lane order.  A wave owns 16 rows, a work-group
// RLC_ROWS rows and RLC_COLS columns.
__global__ __launch_bounds__(RLC_NT) void rlc_rows_kernel(const uint32_t* __restrict__ sq, uint32_t sq_rows,
                                                          uint32_t cols, const uint32_t* __restrict__ rowidx,
                                                          uint32_t rows, const uint4* __restrict__ coeffs,
                                                          uint32_t* __restrict__ partial) {
  __shared__ uint2 bt[RLC_STEPS * 64];
  const uint32_t blk = blockIdx.x, nblk = gridDim.x;
  const uint32_t c0 = blk * RLC_COLS;
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    // Thread t builds the coefficient operands of columns c0 + 8t .. 8t + 7:
    // k-step t / 4, lane group t % 4.
    const uint32_t cb = c0 + 8u * (uint32_t)tid;
    uint32_t d[4][8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const uint4 cf = cb + e < cols ? coeffs[cb + e] : make_uint4(0u, 0u, 0u, 0u);
      d[0][e] = signed_digits(canon_u32(cf.x));
      d[1][e] = signed_digits(canon_u32(cf.y));
      d[2][e] = signed_digits(canon_u32(cf.z));
      d[3][e] = signed_digits(canon_u32(cf.w));
    }
    uint2* dst = bt + (tid >> 2) * 64 + (tid & 3) * 16;
#pragma unroll
    for (int l = 0; l < 4; l++) {
      uint32_t lo[4], hi[4];
      byte_transpose(d[l][0], d[l][1], d[l][2], d[l][3], lo);
      byte_transpose(d[l][4], d[l][5], d[l][6], d[l][7], hi);
#pragma unroll
      for (int j = 0; j < 4; j++) dst[4 * l + j] = make_uint2(lo[j], hi[j]);
    }
  }
  __syncthreads();

  const int m = lane & 15, g = lane >> 4;
  const uint32_t slot0 = blockIdx.y * RLC_ROWS + (uint32_t)wave * 16u;
  // Rows that are not requested or lie outside the square read row 0 (their
  // products are discarded: row m of a product depends on row m of the values
  // only).
  uint32_t r = 0;
  if (slot0 + m < rows) {
    r = rowidx ? rowidx[slot0 + m] : slot0 + m;
    if (r >= sq_rows) r = 0;
  }
  const uint32_t* src = sq + (size_t)r * cols + c0 + 8u * (uint32_t)g;
  const bool fast = c0 + RLC_COLS <= cols && (cols & 3u) == 0 && ((uintptr_t)sq & 15u) == 0;
  rlc_v4i acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  if (fast)
    rlc_block_sums<true>(src, c0 + 8u * (uint32_t)g, cols, bt, lane, acc);
  else
    rlc_block_sums<false>(src, c0 + 8u * (uint32_t)g, cols, bt, lane, acc);

  // Output layout of the 16x16 product: lane L, register q holds
  // D[4 * (L / 16) + q][L % 16].  Column n = 4l + j: weight 256^j, then the
  // four j lanes of a limb are added.
  const int n = lane & 15, l = n >> 2, j = n & 3;
  const double wj = (double)(1u << (8 * j));
#pragma unroll
  for (int q = 0; q < 4; q++) {
    // exact: |D| < 2^25, so every term and partial sum stays below 2^50
    double x = (double)acc[0][q] + 256.0 * (double)acc[1][q] + 65536.0 * (double)acc[2][q] +
               16777216.0 * (double)acc[3][q];
    x = mulmod(reduce(x), wj, wj * PINV);
    x += __shfl_xor(x, 1);
    x += __shfl_xor(x, 2);
    const uint32_t slot = slot0 + 4u * (uint32_t)g + (uint32_t)q;
    if (j == 0 && slot < rows) {
      const uint32_t rr = rowidx ? rowidx[slot] : slot;
      partial[((size_t)slot * nblk + blk) * 4 + l] = rr < sq_rows ? canonical(x) : 0u;
    }
  }
}

__global__ void rlc_sum_kernel(const uint32_t* __restrict__ partial, uint32_t rows, uint32_t nblk,
                               uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;  // (row, limb)
  if (t >= rows * 4) return;
  const uint32_t r = t >> 2, l = t & 3;
  uint64_t x = 0;
  for (uint32_t b = 0; b < nblk; b++) x = fold(x + partial[((size_t)r * nblk + b) * 4 + l]);
  out[t] = mod_p(x);
}

// Partial sums of zoda_hip_rlc_rows_noalloc: a static device array, so the
// entry point allocates nothing.  Each device holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (rlc_scratch_address).
constexpr size_t RLC_SCRATCH_WORDS = size_t(1) << 22;  // 16 MiB: count * column blocks <= 2^20
static __device__ __align__(256) uint32_t scratch[RLC_SCRATCH_WORDS];

// Both kernels of a row RLC: the column blocks' sums into `partial` (every
// entry the second kernel reads is written, nothing has to start at zero),
// then their sum per row and limb.
int launch_rlc_rows(const uint32_t* d_square, uint32_t rows, uint32_t cols, const uint32_t* d_row_index,
                    uint32_t count, const uint32_t* d_coeffs, uint32_t* d_out, uint32_t* partial,
                    uint32_t nblk, uint32_t nrb, hipStream_t s) {
  hipLaunchKernelGGL(rlc_rows_kernel, dim3(nblk, nrb), dim3(RLC_NT), 0, s, d_square, rows, cols, d_row_index,
                     count, reinterpret_cast<const uint4*>(d_coeffs), partial);
  int rc = 0;
  hip_note(rc, hipGetLastError(), "RLC kernel launch");
  if (!rc) {
    hipLaunchKernelGGL(rlc_sum_kernel, dim3((count * 4 + 255) / 256), dim3(256), 0, s, partial, count, nblk,
                       d_out);
    hip_note(rc, hipGetLastError(), "RLC reduction launch");
  }
  return rc;
}

// Argument checks shared by both row entry points; *nblk and *nrb get the grid.
int check_rlc_rows(const uint32_t* d_square, uint32_t rows, uint32_t cols, const uint32_t* d_row_index,
                   uint32_t count, const uint32_t* d_coeffs, uint32_t* d_out, uint32_t* nblk, uint32_t* nrb) {
  if (!d_square || !d_coeffs || !d_out || rows == 0 || cols == 0 || count == 0)
    return fail(ZODA_EINVAL, "bad arguments");
  if (!d_row_index && count > rows) return fail(ZODA_EINVAL, "count %u exceeds rows %u", count, rows);
  *nblk = (cols + RLC_COLS - 1) / RLC_COLS;
  *nrb = (count + RLC_ROWS - 1) / RLC_ROWS;
  if (*nrb > 65535) return fail(ZODA_EINVAL, "too many rows (%u)", count);
  return 0;
}

// The current device's address of `scratch`, looked up once per device (with
// that device current) and kept in its state.
int rlc_scratch_address(DeviceState* st, uint32_t** out) {
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  if (!st->rlc_scratch) {
    void* p = nullptr;
    ZODA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ZODA_ESTATE, "no device address for the RLC scratch array");
    st->rlc_scratch = static_cast<uint32_t*>(p);
  }
  *out = st->rlc_scratch;
  return 0;
}

}  // namespace

}  // namespace zoda

using namespace zoda;

extern "C" int zoda_hip_rlc_coefficients(const char* data_root, uint32_t root_len, uint32_t cols,
                                         uint32_t* d_coeffs, void* stream) {
  clear_error();
  if (!data_root || !d_coeffs || cols == 0 || root_len > 4096) return fail(ZODA_EINVAL, "bad arguments");
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  hipStream_t s = (hipStream_t)stream;
  if (root_len % 64 == 0) {  // e.g. the 64-character hex data root
    Sha256 h;
    h.init();
    h.update(reinterpret_cast<const uint8_t*>(data_root), root_len);
    ShaMid mid;
    for (int i = 0; i < 8; i++) mid.h[i] = h.h[i];
    hipLaunchKernelGGL(rlc_coeff_mid_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, mid,
                       (uint64_t)root_len * 8 + 64, cols, d_coeffs);
    int rc = 0;
    hip_note(rc, hipGetLastError(), "coefficient kernel launch");
    hip_note(rc, hipStreamSynchronize(s), "coefficient derivation");
    return rc;
  }
  uint8_t* d_root = nullptr;
  ZODA_TRY(pool_alloc(st, root_len ? root_len : 1, s, (void**)&d_root));
  int rc = 0;
  if (root_len) hip_note(rc, hipMemcpyAsync(d_root, data_root, root_len, hipMemcpyHostToDevice, s), "copying the data root");
  if (!rc) {
    hipLaunchKernelGGL(rlc_coeff_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, d_root, root_len, cols,
                       d_coeffs);
    hip_note(rc, hipGetLastError(), "coefficient kernel launch");
  }
  // The copy reads the caller's (pageable) string: finish before returning.
  hip_note(rc, hipFreeAsync(d_root, s), "hipFreeAsync");
  hip_note(rc, hipStreamSynchronize(s), "coefficient derivation");
  return rc;
}

extern "C" int zoda_hip_rlc_rows(const uint32_t* d_square, uint32_t rows, uint32_t cols,
                                 const uint32_t* d_row_index, uint32_t count, const uint32_t* d_coeffs,
                                 uint32_t* d_out, void* stream) {
  clear_error();
  uint32_t nblk = 0, nrb = 0;
  ZODA_TRY(check_rlc_rows(d_square, rows, cols, d_row_index, count, d_coeffs, d_out, &nblk, &nrb));
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  hipStream_t s = (hipStream_t)stream;
  uint32_t* partial = nullptr;
  ZODA_TRY(pool_alloc(st, (size_t)count * nblk * 4 * sizeof(uint32_t), s, (void**)&partial));
  int rc = launch_rlc_rows(d_square, rows, cols, d_row_index, count, d_coeffs, d_out, partial, nblk, nrb, s);
  hip_note(rc, hipFreeAsync(partial, s), "hipFreeAsync");
  return rc;
}

extern "C" int zoda_hip_rlc_rows_noalloc(const uint32_t* d_square, uint32_t rows, uint32_t cols,
                                         const uint32_t* d_row_index, uint32_t count, const uint32_t* d_coeffs,
                                         uint32_t* d_out, void* stream) {
  clear_error();
  uint32_t nblk = 0, nrb = 0;
  ZODA_TRY(check_rlc_rows(d_square, rows, cols, d_row_index, count, d_coeffs, d_out, &nblk, &nrb));
  if ((size_t)count * nblk * 4 > RLC_SCRATCH_WORDS)
    return fail(ZODA_EINVAL,
                "%u rows x %u column blocks of %u columns need %zu scratch words, above the limit of %zu "
                "(rows x blocks <= %zu)",
                count, nblk, RLC_COLS, (size_t)count * nblk * 4, RLC_SCRATCH_WORDS, RLC_SCRATCH_WORDS / 4);
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  uint32_t* partial = nullptr;
  ZODA_TRY(rlc_scratch_address(st, &partial));
  return launch_rlc_rows(d_square, rows, cols, d_row_index, count, d_coeffs, d_out, partial, nblk, nrb,
                         (hipStream_t)stream);
}

// SHA-256 of the rows' values as little-endian u32 in row order, as lowercase
// hex (src/benchmarks/utils.rs:51-59).  Host memory, on the host: the hash is
// one sequential chain.
extern "C" int zoda_data_root(const uint32_t* values, size_t count, char* hex_out) {
  clear_error();
  if ((!values && count) || !hex_out) return fail(ZODA_EINVAL, "bad arguments");
  Sha256 h;
  h.init();
  // The bytes of each u32 in little-endian order are its memory image on
  // this (little-endian) host.
  h.update(reinterpret_cast<const uint8_t*>(values), count * sizeof(uint32_t));
  uint8_t d[32];
  h.final(d);
  static const char* hexd = "0123456789abcdef";
  for (int i = 0; i < 32; i++) {
    hex_out[2 * i] = hexd[d[i] >> 4];
    hex_out[2 * i + 1] = hexd[d[i] & 15];
  }
  hex_out[64] = '\0';
  return 0;
}
