// The reference's generic transforms over any modulus below 2^31 with the
// caller's roots (src/ntt/cpu.rs:1-32 `ntt`, :45-60 `intt`), on the GPU.  The
// arithmetic is the reference's exactly: bit-reversed input, radix-2
// decimation in time with roots[j * n / len] at each stage,
// b = hi * w mod m (64-bit), lo' = (a + b) mod m, hi' = (a + m - b) mod m, and
// for the inverse the roots order reversed after roots[0] and a final scaling
// by n^(m-2) mod m.  One work-group transforms one polynomial in LDS.
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace zoda {

namespace {

constexpr uint32_t MOD_MAX_N = 8192;  // 32 KiB of LDS per polynomial
constexpr int MOD_NT = 256;

uint32_t pow_mod_m(uint32_t b, uint32_t e, uint32_t m) {
  uint64_t r = 1 % m, x = b % m;
  while (e) {
    if (e & 1) r = r * x % m;
    x = x * x % m;
    e >>= 1;
  }
  return (uint32_t)r;
}

// roots: n entries (already reversed after roots[0] for the inverse).
__global__ __launch_bounds__(MOD_NT) void ntt_mod_kernel(uint32_t* __restrict__ values, const uint32_t* __restrict__ roots,
                                                         uint32_t log_n, uint32_t stride, uint32_t m,
                                                         uint32_t n_inv, int scale) {
  extern __shared__ uint32_t v[];
  const uint32_t n = 1u << log_n;
  uint32_t* poly = values + (size_t)blockIdx.x * stride;
  for (uint32_t i = threadIdx.x; i < n; i += MOD_NT) {
    const uint32_t r = log_n ? (__brev(i) >> (32 - log_n)) : 0u;
    v[r] = poly[i];
  }
  __syncthreads();
  for (uint32_t len = 2; len <= n; len <<= 1) {
    const uint32_t half = len >> 1, rstride = n / len;
    for (uint32_t t = threadIdx.x; t < n / 2; t += MOD_NT) {
      const uint32_t blk = t / half, j = t % half;
      const uint32_t lo = blk * len + j, hi = lo + half;
      const uint64_t w = roots[(size_t)j * rstride];
      const uint32_t a = v[lo];
      const uint32_t b = (uint32_t)((uint64_t)v[hi] * w % m);
      v[lo] = (a + b) % m;
      v[hi] = (a + m - b) % m;
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < n; i += MOD_NT)
    poly[i] = scale ? (uint32_t)((uint64_t)v[i] * n_inv % m) : v[i];
}

}  // namespace

}  // namespace zoda

using namespace zoda;

extern "C" int zoda_hip_ntt_mod(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys,
                                uint32_t stride, uint32_t modulus, int inverse) {
  clear_error();
  if (!values || !roots) return fail(ZODA_EINVAL, "values or roots is NULL");
  if (n == 0 || !is_pow2(n) || n > MOD_MAX_N) return fail(ZODA_EINVAL, "n = %u must be a power of two <= %u", n, MOD_MAX_N);
  if (stride < n) return fail(ZODA_EINVAL, "stride %u < n %u", stride, n);
  if (modulus < 2 || modulus >= (1u << 31)) return fail(ZODA_EINVAL, "modulus %u must be in [2, 2^31)", modulus);
  if (polys == 0) return 0;
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::vector<uint32_t> r(roots, roots + n);
  if (inverse) std::reverse(r.begin() + 1, r.end());
  const uint32_t n_inv = pow_mod_m(n % modulus, modulus - 2, modulus);
  hipStream_t s = nullptr;
  ZODA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  uint32_t *d_v = nullptr, *d_r = nullptr;
  int rc = 0;
  const size_t pitch = (size_t)stride * 4, row = (size_t)n * 4;
  if (const hipError_t ea = hipMalloc((void**)&d_v, (size_t)polys * row)) rc = alloc_fail(ea, "device buffer of the values");
  if (!rc)
    if (const hipError_t ea = hipMalloc((void**)&d_r, row)) rc = alloc_fail(ea, "device buffer of the roots");
  if (!rc) hip_note(rc, hipMemcpy(d_r, r.data(), row, hipMemcpyHostToDevice), "copying the roots");
  if (!rc)
    hip_note(rc, hipMemcpy2DAsync(d_v, row, values, pitch, row, polys, hipMemcpyHostToDevice, s), "copying the values");
  if (!rc) {
    uint32_t done = 0;
    while (!rc && done < polys) {  // grid x is limited: launch in slices
      const uint32_t w = std::min<uint32_t>(polys - done, 1u << 20);
      hipLaunchKernelGGL(ntt_mod_kernel, dim3(w), dim3(MOD_NT), row, s, d_v + (size_t)done * n, d_r,
                         log2_ceil(n), n, modulus, n_inv, inverse ? 1 : 0);
      hip_note(rc, hipGetLastError(), "transform launch");
      done += w;
    }
  }
  if (!rc)
    hip_note(rc, hipMemcpy2DAsync(values, pitch, d_v, row, row, polys, hipMemcpyDeviceToHost, s),
             "copying the values back");
  hip_note(rc, hipStreamSynchronize(s), "transform");
  (void)hipFree(d_v);
  (void)hipFree(d_r);
  (void)hipStreamDestroy(s);
  return rc;
}
