// Random linear combinations of the encoded rows on the device, and the
// commitment values they are checked against (src/zoda/rlc.rs,
// src/benchmarks/utils.rs:51-59,221-272, src/zoda/babybear.rs:164-192).
//
//   rlc(row)[l] = sum over columns c of row[c] * coeff[c][l]  (mod p), l < 4
//
// The row kernel streams the row-major square once.  A work-group owns a
// block of RLC_COLS columns, keeps their coefficients in LDS, and each of its
// waves takes whole rows of the block: lane l reads columns l, l + 64, ...
// (every load instruction covers 256 contiguous bytes), accumulates the
// 31 x 31-bit products in 64-bit integers -- folding the high word back with
// 2^32 = 2^32 - 2p (mod p) every three products so the sum never overflows --
// and the wave adds its lanes once per row.  A second kernel adds the column
// blocks' partial sums.
#include <string.h>

#include <vector>

#include "common.hpp"
#include "sha256.hpp"

namespace zoda {

namespace {

constexpr int RLC_NT = 256;
constexpr uint32_t RLC_COLS = 2048;  // columns per work-group (32 KiB of coefficients)
constexpr uint32_t RLC_ROWS = 32;    // rows per work-group
constexpr int RLC_PER = RLC_COLS / 64;
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

// partial[(i * nblk + blk) * 4 + l] for the i-th requested row (row rowidx[i],
// or i without an index); a requested row outside the square contributes 0.
__global__ __launch_bounds__(RLC_NT) void rlc_rows_kernel(const uint32_t* __restrict__ sq, uint32_t sq_rows,
                                                          uint32_t cols, const uint32_t* __restrict__ rowidx,
                                                          uint32_t rows, const uint4* __restrict__ coeffs,
                                                          uint32_t* __restrict__ partial) {
  __shared__ uint4 cs[RLC_COLS];
  const uint32_t blk = blockIdx.x, nblk = gridDim.x;
  const uint32_t c0 = blk * RLC_COLS, r0 = blockIdx.y * RLC_ROWS;
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t i = (uint32_t)tid; i < RLC_COLS; i += RLC_NT)
    cs[i] = c0 + i < cols ? coeffs[c0 + i] : make_uint4(0u, 0u, 0u, 0u);
  __syncthreads();
  const bool full = c0 + RLC_COLS <= cols;
  for (uint32_t i = (uint32_t)wave; i < RLC_ROWS && r0 + i < rows; i += RLC_NT / 64) {
    const uint32_t r = rowidx ? rowidx[r0 + i] : r0 + i;
    // Two accumulator sets (even and odd j) halve the dependency chains.
    uint64_t acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    if (r < sq_rows) {
      const uint32_t* src = sq + (size_t)r * cols + c0 + lane;
#pragma unroll
      for (int jb = 0; jb < RLC_PER; jb += 16) {
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const uint32_t c = (uint32_t)(jb + j) * 64 + (uint32_t)lane;
          v[j] = (full || c0 + c < cols) ? src[(size_t)(jb + j) * 64] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const uint4 cf = cs[(jb + j) * 64 + lane];
          uint64_t* a = acc[j & 1];
          a[0] += (uint64_t)v[j] * cf.x;
          a[1] += (uint64_t)v[j] * cf.y;
          a[2] += (uint64_t)v[j] * cf.z;
          a[3] += (uint64_t)v[j] * cf.w;
          if ((j >> 1) == 2 || (j >> 1) == 5 || (j >> 1) == 7) {  // at most three products per set between folds
#pragma unroll
            for (int l = 0; l < 4; l++) a[l] = fold(a[l]);
          }
        }
      }
    }
    uint64_t accs[4];
#pragma unroll
    for (int l = 0; l < 4; l++) accs[l] = fold(acc[0][l]) + fold(acc[1][l]);
    // Across the wave: each step adds two values below 2^61 and folds.
#pragma unroll
    for (int l = 0; l < 4; l++) {
      uint64_t x = fold(accs[l]);
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) x = fold(x + __shfl_xor(x, off));
      accs[l] = x;
    }
    if (lane == 0) {
      uint4 o;
      o.x = mod_p(accs[0]);
      o.y = mod_p(accs[1]);
      o.z = mod_p(accs[2]);
      o.w = mod_p(accs[3]);
      reinterpret_cast<uint4*>(partial)[(size_t)(r0 + i) * nblk + blk] = o;
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
    if (hipGetLastError() != hipSuccess) return fail(ZODA_ESTATE, "coefficient kernel launch failed");
    if (hipStreamSynchronize(s) != hipSuccess) return fail(ZODA_ESTATE, "coefficient derivation failed");
    return 0;
  }
  uint8_t* d_root = nullptr;
  ZODA_HIP(hipMallocAsync((void**)&d_root, root_len ? root_len : 1, s));
  int rc = 0;
  if (root_len && hipMemcpyAsync(d_root, data_root, root_len, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = fail(ZODA_ESTATE, "copying the data root failed");
  if (!rc) {
    hipLaunchKernelGGL(rlc_coeff_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, d_root, root_len, cols,
                       d_coeffs);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "coefficient kernel launch failed");
  }
  // The copy reads the caller's (pageable) string: finish before returning.
  if (hipFreeAsync(d_root, s) != hipSuccess && !rc) rc = fail(ZODA_ESTATE, "hipFreeAsync failed");
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(ZODA_ESTATE, "coefficient derivation failed");
  return rc;
}

extern "C" int zoda_hip_rlc_rows(const uint32_t* d_square, uint32_t rows, uint32_t cols,
                                 const uint32_t* d_row_index, uint32_t count, const uint32_t* d_coeffs,
                                 uint32_t* d_out, void* stream) {
  clear_error();
  if (!d_square || !d_coeffs || !d_out || rows == 0 || cols == 0 || count == 0)
    return fail(ZODA_EINVAL, "bad arguments");
  if (!d_row_index && count > rows) return fail(ZODA_EINVAL, "count %u exceeds rows %u", count, rows);
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  hipStream_t s = (hipStream_t)stream;
  const uint32_t nblk = (cols + RLC_COLS - 1) / RLC_COLS;
  const uint32_t nrb = (count + RLC_ROWS - 1) / RLC_ROWS;
  if (nrb > 65535) return fail(ZODA_EINVAL, "too many rows (%u)", count);
  uint32_t* partial = nullptr;
  ZODA_HIP(hipMallocAsync((void**)&partial, (size_t)count * nblk * 4 * sizeof(uint32_t), s));
  hipLaunchKernelGGL(rlc_rows_kernel, dim3(nblk, nrb), dim3(RLC_NT), 0, s, d_square, rows, cols, d_row_index,
                     count, reinterpret_cast<const uint4*>(d_coeffs), partial);
  int rc = 0;
  if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "RLC kernel launch failed");
  if (!rc) {
    hipLaunchKernelGGL(rlc_sum_kernel, dim3((count * 4 + 255) / 256), dim3(256), 0, s, partial, count, nblk,
                       d_out);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "RLC reduction launch failed");
  }
  if (hipFreeAsync(partial, s) != hipSuccess && !rc) rc = fail(ZODA_ESTATE, "hipFreeAsync failed");
  return rc;
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
