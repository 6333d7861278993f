// C ABI of libzoda_hip (include/zoda_hip.h): argument checking, the per-thread
// last error, per-device state, the drop-in host-memory transforms and the
// streamed end-to-end encode.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"

namespace zoda {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

constexpr int MAX_DEVICES = 64;
static DeviceState* g_states[MAX_DEVICES];
static std::mutex g_states_mu;

int current_state(DeviceState** out) {
  int dev = 0;
  ZODA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(ZODA_ENODEV, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_states_mu);
  if (!g_states[dev]) {
    g_states[dev] = new DeviceState();
    g_states[dev]->device = dev;
  }
  *out = g_states[dev];
  return 0;
}

int get_buffer(DeviceState* st, int slot, size_t bytes, void** out) {
  if ((int)st->buffers.size() <= slot) {
    st->buffers.resize(slot + 1, nullptr);
    st->buffer_bytes.resize(slot + 1, 0);
  }
  if (st->buffer_bytes[slot] < bytes) {
    if (st->buffers[slot]) ZODA_HIP(hipFree(st->buffers[slot]));
    st->buffers[slot] = nullptr;
    st->buffer_bytes[slot] = 0;
    if (const hipError_t ea = hipMalloc(&st->buffers[slot], bytes)) {
      st->buffers[slot] = nullptr;
      return alloc_fail(ea, "device staging buffer");
    }
    st->buffer_bytes[slot] = bytes;
  }
  *out = st->buffers[slot];
  return 0;
}

int pool_alloc(DeviceState* st, size_t bytes, hipStream_t stream, void** out) {
  {
    std::lock_guard<std::recursive_mutex> lock(st->mu);
    if (!st->pool) {
      hipMemPoolProps props{};
      props.allocType = hipMemAllocationTypePinned;
      props.handleTypes = hipMemHandleTypeNone;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = st->device;
      hipMemPool_t pool = nullptr;
      ZODA_HIP(hipMemPoolCreate(&pool, &props));
      uint64_t keep = POOL_KEEP_BYTES;
      if (const hipError_t e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep)) {
        (void)hipMemPoolDestroy(pool);
        return fail((int)e, "setting the scratch pool's release threshold failed: %s", hipGetErrorString(e));
      }
      st->pool = pool;
    }
  }
  if (const hipError_t ea = hipMallocFromPoolAsync(out, bytes, st->pool, stream)) {
    *out = nullptr;
    return alloc_fail(ea, "scratch allocation");
  }
  return 0;
}

int get_streams(DeviceState* st, int count, hipStream_t** out) {
  while ((int)st->streams.size() < count) {
    hipStream_t s;
    ZODA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    st->streams.push_back(s);
  }
  *out = st->streams.data();
  return 0;
}

static int upload_mont(const std::vector<uint32_t>& fwd, const std::vector<uint32_t>& inv,
                       MontTables* out) {
  MontTables t;
  const size_t bytes = fwd.size() * sizeof(uint32_t);
  if (const hipError_t ea = hipMalloc((void**)&t.fwd, bytes)) return alloc_fail(ea, "twiddle table allocation");
  if (const hipError_t ea = hipMalloc((void**)&t.inv, bytes)) {
    (void)hipFree(t.fwd);
    return alloc_fail(ea, "twiddle table allocation");
  }
  ZODA_HIP(hipMemcpy(t.fwd, fwd.data(), bytes, hipMemcpyHostToDevice));
  ZODA_HIP(hipMemcpy(t.inv, inv.data(), bytes, hipMemcpyHostToDevice));
  *out = t;
  return 0;
}

// tw[h + j] = roots[j * n/(2h)] * 2^32 mod p; the inverse table reads the
// reversed roots inv[i] = roots[(n - i) % n] (cuda/ntt_kernel.cu:198-200).
static void mont_from_roots(const uint32_t* roots, uint32_t n, std::vector<uint32_t>* fwd,
                            std::vector<uint32_t>* inv) {
  fwd->assign(std::max<uint32_t>(n, 2), 0);
  inv->assign(std::max<uint32_t>(n, 2), 0);
  for (uint32_t h = 1; h < n; h <<= 1)
    for (uint32_t j = 0; j < h; j++) {
      const uint32_t e = j * (n / (2 * h));
      (*fwd)[h + j] = to_mont_host(roots[e]);
      (*inv)[h + j] = to_mont_host(roots[(n - e) % n]);
    }
}

static std::vector<uint32_t> standard_roots(uint32_t log_n) {
  const uint32_t n = 1u << log_n, w = root_of_unity(log_n);
  std::vector<uint32_t> r(n);
  uint32_t cur = 1;
  for (uint32_t i = 0; i < n; i++) {
    r[i] = cur;
    cur = mul_mod_host(cur, w);
  }
  return r;
}

int get_mont_tables(DeviceState* st, uint32_t log_n, MontTables* out) {
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  auto it = st->mont.find(log_n);
  if (it == st->mont.end()) {
    std::vector<uint32_t> roots = standard_roots(log_n), fwd, inv;
    mont_from_roots(roots.data(), 1u << log_n, &fwd, &inv);
    MontTables t;
    ZODA_TRY(upload_mont(fwd, inv, &t));
    it = st->mont.emplace(log_n, t).first;
  }
  *out = it->second;
  return 0;
}

int get_user_mont_tables(DeviceState* st, const uint32_t* roots, uint32_t log_n, MontTables* out) {
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  const uint32_t n = 1u << log_n;
  auto it = st->user_roots.find(log_n);
  if (it != st->user_roots.end() && std::memcmp(it->second.data(), roots, n * 4) == 0) {
    *out = st->user_mont[log_n];
    return 0;
  }
  std::vector<uint32_t> fwd, inv;
  mont_from_roots(roots, n, &fwd, &inv);
  auto old = st->user_mont.find(log_n);
  if (old != st->user_mont.end()) {
    ZODA_HIP(hipDeviceSynchronize());  // earlier calls may still read the old tables
    (void)hipFree(old->second.fwd);
    (void)hipFree(old->second.inv);
    st->user_mont.erase(old);
  }
  MontTables t;
  ZODA_TRY(upload_mont(fwd, inv, &t));
  st->user_mont[log_n] = t;
  st->user_roots[log_n].assign(roots, roots + n);
  st->user_standard.erase(log_n);  // decided again for the new roots
  *out = t;
  return 0;
}

// Called after get_user_mont_tables with the same roots: the answer is cached
// beside the roots that table was built from (and dropped when they change).
bool roots_are_standard(DeviceState* st, const uint32_t* roots, uint32_t log_n) {
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  const size_t n = (size_t)1 << log_n;
  const auto ur = st->user_roots.find(log_n);
  const bool cached_roots = ur != st->user_roots.end() && std::memcmp(ur->second.data(), roots, n * 4) == 0;
  if (cached_roots) {
    const auto us = st->user_standard.find(log_n);
    if (us != st->user_standard.end()) return us->second;
  }
  const std::vector<uint32_t> std_roots = standard_roots(log_n);
  const bool same = std::memcmp(std_roots.data(), roots, n * 4) == 0;
  if (cached_roots) st->user_standard[log_n] = same;
  return same;
}

// ---- seeded uniform generator -------------------------------------------------

__device__ __forceinline__ uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Element i: the first of z_a = mix(seed + (64 i + a + 1) * gamma) >> 33,
// a = 0, 1, ..., below p.  A value needs another attempt with probability
// 1/16, so 64 attempts fail with probability 2^-256; the cap only bounds the
// loop.
__global__ void fill_uniform_kernel(uint32_t* __restrict__ out, size_t count, uint64_t seed) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x) {
    uint32_t v = 0;
    for (uint64_t a = 0; a < 64; a++) {
      v = (uint32_t)(splitmix64_mix(seed + (64ull * i + a + 1) * 0x9E3779B97F4A7C15ull) >> 33);
      if (v < P) break;
    }
    out[i] = v < P ? v : v - P;
  }
}

int launch_fill_uniform(uint32_t* d_out, size_t count, uint64_t seed, hipStream_t stream) {
  if (count == 0) return 0;
  const unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 1u << 16);
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(blocks), dim3(256), 0, stream, d_out, count, seed);
  ZODA_HIP(hipGetLastError());
  return 0;
}

}  // namespace zoda

using namespace zoda;

// ---- C ABI ----------------------------------------------------------------------

extern "C" const char* zoda_hip_last_error(void) { return g_last_error.c_str(); }

extern "C" int zoda_hip_device_count(int* count) {
  clear_error();
  if (!count) return fail(ZODA_EINVAL, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    c = 0;
  }
  *count = c;
  return 0;
}

extern "C" uint32_t* zoda_hip_host_alloc(size_t count) {
  clear_error();
  if (count == 0 || count > (SIZE_MAX / 4)) {
    fail(ZODA_EINVAL, "bad pinned allocation size %zu", count);
    return nullptr;
  }
  void* p = nullptr;
  if (hipHostMalloc(&p, count * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    fail(ZODA_ENOMEM, "pinned allocation of %zu u32 failed", count);
    return nullptr;
  }
  return (uint32_t*)p;
}

extern "C" void zoda_hip_host_free(uint32_t* p) {
  if (p) (void)hipHostFree(p);
}

extern "C" int zoda_hip_roots_of_unity(uint32_t n, uint32_t* roots) {
  clear_error();
  if (!roots) return fail(ZODA_EINVAL, "roots is NULL");
  if (!is_pow2(n) || n > (1u << 27)) return fail(ZODA_EINVAL, "n = %u must be a power of two <= 2^27", n);
  const uint32_t w = root_of_unity(log2_ceil(n));
  uint32_t cur = 1;
  for (uint32_t i = 0; i < n; i++) {
    roots[i] = cur;
    cur = mul_mod_host(cur, w);
  }
  return 0;
}

static int check_transform_args(const void* values, uint32_t n, uint32_t stride) {
  if (!values) return fail(ZODA_EINVAL, "values is NULL");
  if (!is_pow2(n) || n > (1u << 27)) return fail(ZODA_EINVAL, "NTT size %u must be a power of two <= 2^27", n);
  if (stride < n) return fail(ZODA_EINVAL, "stride %u must be at least n = %u", stride, n);
  return 0;
}

// cuda/ntt_kernel.cu:106-180 contract: host values, polys chunks streamed
// through device buffers with 2D copies; device memory is reused across calls.
static int host_transform(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys,
                          uint32_t stride, int inverse) {
  clear_error();
  ZODA_TRY(check_transform_args(values, n, stride));
  if (!roots) return fail(ZODA_EINVAL, "roots is NULL");
  if (polys == 0) return 0;
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  const uint32_t log_n = log2_ceil(n);
  MontTables tabs;
  ZODA_TRY(get_user_mont_tables(st, roots, log_n, &tabs));
  const bool standard = roots_are_standard(st, roots, log_n);
  constexpr int NBUF = 3;
  const size_t chunk_bytes = 128u << 20;
  const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(polys, chunk_bytes / ((size_t)n * 4)));
  hipStream_t* streams = nullptr;
  ZODA_TRY(get_streams(st, NBUF, &streams));
  uint32_t* bufs[NBUF];
  for (int b = 0; b < NBUF; b++) ZODA_TRY(get_buffer(st, b, (size_t)chunk * n * 4, (void**)&bufs[b]));
  const size_t poly_bytes = (size_t)n * 4, pitch = (size_t)stride * 4;
  int ci = 0;
  for (uint32_t p0 = 0; p0 < polys; p0 += chunk, ci++) {
    const uint32_t w = std::min(chunk, polys - p0);
    hipStream_t s = streams[ci % NBUF];
    uint32_t* d = bufs[ci % NBUF];
    uint32_t* h = values + (size_t)p0 * stride;
    ZODA_HIP(hipMemcpy2DAsync(d, poly_bytes, h, pitch, poly_bytes, w, hipMemcpyHostToDevice, s));
    ZODA_TRY(launch_ntt(st, d, log_n, w, n, inverse, tabs, standard, s));
    ZODA_HIP(hipMemcpy2DAsync(h, pitch, d, poly_bytes, poly_bytes, w, hipMemcpyDeviceToHost, s));
  }
  for (int b = 0; b < NBUF; b++) ZODA_HIP(hipStreamSynchronize(streams[b]));
  return 0;
}

extern "C" int zoda_hip_ntt(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys,
                            uint32_t stride) {
  return host_transform(values, roots, n, polys, stride, 0);
}

extern "C" int zoda_hip_intt(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys,
                             uint32_t stride) {
  return host_transform(values, roots, n, polys, stride, 1);
}

extern "C" int zoda_hip_ntt_device(uint32_t* d_values, uint32_t n, uint32_t polys, uint32_t stride,
                                   int inverse, void* stream) {
  clear_error();
  ZODA_TRY(check_transform_args(d_values, n, stride));
  if (polys == 0) return 0;
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  MontTables tabs;
  ZODA_TRY(get_mont_tables(st, log2_ceil(n), &tabs));
  return launch_ntt(st, d_values, log2_ceil(n), polys, stride, inverse ? 1 : 0, tabs, true,
                    (hipStream_t)stream);
}

static int check_encode_args(const void* in, uint32_t k, uint32_t cols, uint32_t n, const void* out,
                             int layout) {
  if (!in || !out) return fail(ZODA_EINVAL, "input or output is NULL");
  if (k == 0 || cols == 0) return fail(ZODA_EINVAL, "k = %u and cols = %u must be positive", k, cols);
  if ((uint64_t)k + n > (1u << 26)) return fail(ZODA_EINVAL, "k + n = %llu exceeds 2^26", (unsigned long long)k + n);
  if (layout != ZODA_LAYOUT_ROW && layout != ZODA_LAYOUT_COL) return fail(ZODA_EINVAL, "unknown layout %d", layout);
  const uint64_t n_kn = 1ull << log2_ceil((uint64_t)k + n);
  if (n_kn * cols > (1ull << 40)) return fail(ZODA_EINVAL, "output of %llu values is too large", (unsigned long long)(n_kn * cols));
  return 0;
}

extern "C" int zoda_hip_encode_device(const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t n,
                                      uint32_t* d_out, int layout, void* stream) {
  clear_error();
  ZODA_TRY(check_encode_args(d_in, k, cols, n, d_out, layout));
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  return launch_encode(st, d_in, k, cols, n, d_out, layout, (hipStream_t)stream);
}

// End-to-end encode from host memory: column slabs of `slab` columns go up
// with a 2D copy (no host transpose), are encoded, and come back with a 2D
// copy into their place in h_out; three streams overlap the copies of one slab
// with the kernels of another.
extern "C" int zoda_hip_encode_host(const uint32_t* h_in, uint32_t k, uint32_t cols, uint32_t n,
                                    uint32_t* h_out, int layout) {
  clear_error();
  ZODA_TRY(check_encode_args(h_in, k, cols, n, h_out, layout));
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  const uint64_t n_kn = 1ull << log2_ceil((uint64_t)k + n);
  constexpr int NBUF = 3;
  const size_t slab_target = 96u << 20;  // bytes of input + output per slab
  uint64_t slab = slab_target / ((k + n_kn) * 4);
  slab = std::max<uint64_t>(64, slab & ~63ull);
  slab = std::min<uint64_t>(slab, cols);
  hipStream_t* streams = nullptr;
  ZODA_TRY(get_streams(st, NBUF, &streams));
  uint32_t *din[NBUF], *dout[NBUF];
  for (int b = 0; b < NBUF; b++) {
    ZODA_TRY(get_buffer(st, 2 * b + 0 + NBUF, (size_t)k * slab * 4, (void**)&din[b]));
    ZODA_TRY(get_buffer(st, 2 * b + 1 + NBUF, (size_t)n_kn * slab * 4, (void**)&dout[b]));
  }
  int si = 0;
  for (uint64_t c0 = 0; c0 < cols; c0 += slab, si++) {
    const uint32_t w = (uint32_t)std::min<uint64_t>(slab, cols - c0);
    hipStream_t s = streams[si % NBUF];
    uint32_t* di = din[si % NBUF];
    uint32_t* dq = dout[si % NBUF];
    ZODA_HIP(hipMemcpy2DAsync(di, (size_t)w * 4, h_in + c0, (size_t)cols * 4, (size_t)w * 4, k,
                              hipMemcpyHostToDevice, s));
    ZODA_TRY(launch_encode(st, di, k, w, n, dq, layout, s));
    if (layout == ZODA_LAYOUT_ROW) {
      ZODA_HIP(hipMemcpy2DAsync(h_out + c0, (size_t)cols * 4, dq, (size_t)w * 4, (size_t)w * 4,
                                n_kn, hipMemcpyDeviceToHost, s));
    } else {
      ZODA_HIP(hipMemcpyAsync(h_out + c0 * n_kn, dq, (size_t)w * n_kn * 4, hipMemcpyDeviceToHost, s));
    }
  }
  for (int b = 0; b < NBUF; b++) ZODA_HIP(hipStreamSynchronize(streams[b]));
  return 0;
}

extern "C" int zoda_hip_fill_uniform(uint32_t* d_out, size_t count, uint64_t seed, void* stream) {
  clear_error();
  if (!d_out && count) return fail(ZODA_EINVAL, "output is NULL");
  return launch_fill_uniform(d_out, count, seed, (hipStream_t)stream);
}
