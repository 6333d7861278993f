// Erasure decoder: the data square back from any k of the 2k rows of its
// extended square.  The reference has no decoder; what it relies on is that
// every column is a Reed-Solomon codeword -- the evaluations at the 2k-th
// roots of unity w^r of a polynomial f of degree < k, with the data at the
// even rows (src/zoda/babybear.rs:195-211).
//
// Per column, with K the k known rows and E the others (k of them):
//   Z(x) = prod_{e in E} (x - w^e)                  vanishes on the erasures
//   y_r  = out[r] * Z(w^r) on K, 0 on E              = g(w^r), g = f * Z, deg g < 2k
//   G    = INTT_2k(y)                                coefficients of g
//   g(h w^r) = NTT_2k(G_i h^i)                       on a coset where Z != 0 (h = 31)
//   f(h w^r) = g(h w^r) / Z(h w^r)
//   f_i  = INTT_2k(f(h w^r))_i * h^-i, i < k         (the upper k are zero)
//   data[j] = f(w^(2j)) = NTT_k(f)_j
// Z on K and 1/Z on the coset depend only on which rows are known: they are
// computed once per call on the host (O(k^2)).  On the device the square goes
// column-major into a scratch buffer (tiled transpose, scaled on the way),
// through the four batched transforms of ntt_poly.hip with pointwise scalings
// between them, and back row-major.
#include <algorithm>
#include <thread>
#include <vector>

#include "common.hpp"

namespace zoda {

namespace {

constexpr uint32_t DEC_COSET = 31;  // a generator: h w^r is never a 2k-th root of unity
constexpr int TT = 32;               // transpose tile

uint32_t mulp(uint32_t a, uint32_t b) { return (uint32_t)((uint64_t)a * b % P); }

// Multiplier w with its Shoup companion floor(w * 2^32 / p).
struct Shoup {
  uint32_t w, wp;
};

Shoup shoup(uint32_t w) { return Shoup{w, (uint32_t)(((uint64_t)w << 32) / P)}; }

// a * s.w mod p, canonical, for any 32-bit a (Shoup's method: the quotient
// estimate mulhi(a, wp) is exact or one short).
__device__ __forceinline__ uint32_t mul_shoup(uint32_t a, Shoup s) {
  const uint32_t q = __umulhi(a, s.wp);
  const uint32_t r = a * s.w - q * P;  // in [0, 2p), mod 2^32
  return r >= P ? r - P : r;
}

// T[c * n + r] = enc[r * cols + c] * z[r] for r < n (a 32 x 32 tile per block).
__global__ void dec_gather(const uint32_t* __restrict__ enc, uint32_t cols, uint32_t n,
                           const Shoup* __restrict__ z, uint32_t* __restrict__ T) {
  __shared__ uint32_t tile[TT][TT + 1];
  const uint32_t c0 = blockIdx.x * TT, r0 = blockIdx.y * TT;
  const uint32_t tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (uint32_t i = ty; i < TT; i += 8) {
    const uint32_t r = r0 + i, c = c0 + tx;
    uint32_t v = 0;
    if (r < n && c < cols) {
      const Shoup s = z[r];
      v = s.w ? mul_shoup(enc[(size_t)r * cols + c], s) : 0u;  // erased rows are not read
    }
    tile[i][tx] = v;
  }
  __syncthreads();
  for (uint32_t i = ty; i < TT; i += 8) {
    const uint32_t c = c0 + i, r = r0 + tx;
    if (c < cols && r < n) T[(size_t)c * n + r] = tile[tx][i];
  }
}

// T[c * stride + i] *= w[i] for i < len.
__global__ void dec_scale(uint32_t* __restrict__ T, uint32_t stride, uint32_t len, uint32_t polys,
                          const Shoup* __restrict__ w) {
  const size_t total = (size_t)polys * len;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)(t / len), i = (uint32_t)(t % len);
    uint32_t* p = T + (size_t)c * stride + i;
    *p = mul_shoup(*p, w[i]);
  }
}

// data[j * cols + c] = T[c * stride + j] for j < k.
__global__ void dec_scatter(const uint32_t* __restrict__ T, uint32_t stride, uint32_t k, uint32_t cols,
                            uint32_t* __restrict__ data) {
  __shared__ uint32_t tile[TT][TT + 1];
  const uint32_t j0 = blockIdx.x * TT, c0 = blockIdx.y * TT;
  const uint32_t tx = threadIdx.x, ty = threadIdx.y;
  for (uint32_t i = ty; i < TT; i += 8) {
    const uint32_t c = c0 + i, j = j0 + tx;
    tile[i][tx] = (c < cols && j < k) ? T[(size_t)c * stride + j] : 0u;
  }
  __syncthreads();
  for (uint32_t i = ty; i < TT; i += 8) {
    const uint32_t j = j0 + i, c = c0 + tx;
    if (j < k && c < cols) data[(size_t)j * cols + c] = tile[tx][i];
  }
}

// The per-row vectors of one decode (host, O(n k), split over threads).
struct DecodeVectors {
  std::vector<Shoup> zk;    // Z(w^r) on known rows, 0 on erased ones
  std::vector<Shoup> zinv;  // 1 / Z(h w^r)
  std::vector<Shoup> hpow;  // h^i, i < n
  std::vector<Shoup> hinv;  // h^-i, i < k
};

void decode_vectors(uint32_t k, const std::vector<uint8_t>& known, DecodeVectors* v) {
  const uint32_t n = 2 * k;
  const uint32_t w = root_of_unity(log2_ceil(n));
  std::vector<uint32_t> x(n);
  x[0] = 1;
  for (uint32_t r = 1; r < n; r++) x[r] = mulp(x[r - 1], w);
  std::vector<uint32_t> erased;
  for (uint32_t r = 0; r < n; r++)
    if (!known[r]) erased.push_back(x[r]);
  std::vector<uint32_t> zk(n), zc(n);
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; t++)
    th.emplace_back([&, t] {
      for (uint32_t r = t; r < n; r += nth) {
        const uint32_t hx = mulp(DEC_COSET, x[r]);
        uint32_t a = 1, b = 1;
        for (uint32_t e : erased) {
          if (known[r]) a = mulp(a, x[r] + P - e);
          b = mulp(b, hx + P - e);
        }
        zk[r] = known[r] ? a : 0;
        zc[r] = b;
      }
    });
  for (auto& t : th) t.join();
  v->zk.resize(n);
  v->zinv.resize(n);
  v->hpow.resize(n);
  v->hinv.resize(k);
  const uint32_t hi = inv_mod_host(DEC_COSET);
  uint32_t hp = 1, hq = 1;
  for (uint32_t r = 0; r < n; r++) {
    v->zk[r] = shoup(zk[r]);
    v->zinv[r] = shoup(inv_mod_host(zc[r]));
    v->hpow[r] = shoup(hp);
    hp = mulp(hp, DEC_COSET);
    if (r < k) {
      v->hinv[r] = shoup(hq);
      hq = mulp(hq, hi);
    }
  }
}

unsigned dec_grid(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 1u << 16); }

}  // namespace

}  // namespace zoda

using namespace zoda;

extern "C" int zoda_hip_decode_device(const uint32_t* d_enc, uint32_t k, uint32_t cols, const uint32_t* known_rows,
                                      uint32_t nknown, uint32_t* d_data, void* stream) {
  clear_error();
  if (!d_enc || !d_data || !known_rows || cols == 0 || k == 0 || !is_pow2(k) || k > (1u << 15))
    return fail(ZODA_EINVAL, "bad arguments (k must be a power of two <= 2^15)");
  const uint32_t n = 2 * k;
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
  DecodeVectors v;
  decode_vectors(k, known, &v);
  MontTables tn, tk;
  ZODA_TRY(get_mont_tables(st, log2_ceil(n), &tn));
  if (k > 1) ZODA_TRY(get_mont_tables(st, log2_ceil(k), &tk));

  hipStream_t s = (hipStream_t)stream;
  Shoup* d_vec = nullptr;
  uint32_t* T = nullptr;
  const size_t nvec = 3 * (size_t)n + k;
  ZODA_HIP(hipMalloc((void**)&d_vec, nvec * sizeof(Shoup)));
  std::vector<Shoup> all;
  all.reserve(nvec);
  all.insert(all.end(), v.zk.begin(), v.zk.end());
  all.insert(all.end(), v.hpow.begin(), v.hpow.end());
  all.insert(all.end(), v.zinv.begin(), v.zinv.end());
  all.insert(all.end(), v.hinv.begin(), v.hinv.end());
  int rc = 0;
  hipError_t e = hipMemcpy(d_vec, all.data(), nvec * sizeof(Shoup), hipMemcpyHostToDevice);
  if (e != hipSuccess) rc = fail((int)e, "decode vectors upload failed: %s", hipGetErrorString(e));
  if (!rc) {
    e = hipMallocAsync((void**)&T, (size_t)cols * n * sizeof(uint32_t), s);
    if (e != hipSuccess) rc = fail((int)e, "decode scratch allocation failed: %s", hipGetErrorString(e));
  }
  const Shoup* zk = d_vec;
  const Shoup* hpow = d_vec + n;
  const Shoup* zinv = d_vec + 2 * n;
  const Shoup* hinv = d_vec + 3 * n;
  if (!rc) {
    hipLaunchKernelGGL(dec_gather, dim3((cols + TT - 1) / TT, (n + TT - 1) / TT), dim3(TT, 8), 0, s, d_enc, cols,
                       n, zk, T);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "decode gather launch failed");
  }
  if (!rc) rc = launch_ntt_poly(T, log2_ceil(n), cols, n, 1, tn, s);
  if (!rc) {
    hipLaunchKernelGGL(dec_scale, dim3(dec_grid((size_t)cols * n)), dim3(256), 0, s, T, n, n, cols, hpow);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "decode scale launch failed");
  }
  if (!rc) rc = launch_ntt_poly(T, log2_ceil(n), cols, n, 0, tn, s);
  if (!rc) {
    hipLaunchKernelGGL(dec_scale, dim3(dec_grid((size_t)cols * n)), dim3(256), 0, s, T, n, n, cols, zinv);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "decode scale launch failed");
  }
  if (!rc) rc = launch_ntt_poly(T, log2_ceil(n), cols, n, 1, tn, s);
  if (!rc) {
    hipLaunchKernelGGL(dec_scale, dim3(dec_grid((size_t)cols * k)), dim3(256), 0, s, T, n, k, cols, hinv);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "decode scale launch failed");
  }
  if (!rc && k > 1) rc = launch_ntt_poly(T, log2_ceil(k), cols, n, 0, tk, s);
  if (!rc) {
    hipLaunchKernelGGL(dec_scatter, dim3((k + TT - 1) / TT, (cols + TT - 1) / TT), dim3(TT, 8), 0, s, T, n, k, cols,
                       d_data);
    if (hipGetLastError() != hipSuccess) rc = fail(ZODA_ESTATE, "decode scatter launch failed");
  }
  if (T && hipFreeAsync(T, s) != hipSuccess && !rc) rc = fail(ZODA_ESTATE, "hipFreeAsync failed");
  // d_vec is read by the kernels just queued: free it once they are done.
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(ZODA_ESTATE, "decode failed");
  (void)hipFree(d_vec);
  return rc;
}
