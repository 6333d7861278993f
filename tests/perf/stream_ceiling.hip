// Streaming ceiling of the encode's memory traffic on one GPU: the bytes of a
// 2048 x 262144 encode (4 B read and 8 B written per input element: input row
// r goes to output rows 2r and 2r + 1) moved with no arithmetic, in the
// access shapes the encode kernels use -- tiles of W adjacent columns (row
// segments of 4W bytes), a persistent grid, tiles handed to the XCDs in
// contiguous runs.  The encode's kernel time is judged against these lines:
// no arrangement of the arithmetic can beat its own access shape.
//
//   dword:  one 4-byte access per lane, NT / W rows per instruction
//   x4:     16-byte accesses, eight lanes per 128-byte row segment (W = 32)
//   modes:  read + write (the encode), write only, read only
//   stores: plain or non-temporal
//
// Build and run (on a GPU):  make -C tests/perf && tests/perf/stream_ceiling
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#define CK(x)                                                                 \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);   \
      exit(1);                                                                \
    }                                                                         \
  } while (0)

constexpr uint32_t S = 2048, C = 262144;

// Block b of nb goes to XCD b % 8; each XCD gets a contiguous run of tiles.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, i = b >> 3;
  const uint32_t q = nb >> 3, r = nb & 7u;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <bool NTS, class T>
__device__ __forceinline__ void put(T* p, T v) {
  if constexpr (NTS)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

enum { RW = 0, WO = 1, RO = 2 };

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Dword shape: thread t covers column t % W of the tile, rows t / W + k NT / W.
template <int W, int NT, int MODE, bool NTS>
__global__ __launch_bounds__(NT) void stream_dword(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   uint32_t* sink) {
  constexpr uint32_t RPS = NT / W;
  static_assert(S % (8 * RPS) == 0, "whole row batches");
  const uint32_t ntiles = C / W, col = threadIdx.x % W, r0 = threadIdx.x / W;
  uint32_t acc = 0;
  for (uint32_t l = blockIdx.x; l < ntiles; l += gridDim.x) {
    const uint32_t c = xcd_tile(l, ntiles) * W + col;
    for (uint32_t r = r0; r < S; r += 8 * RPS) {
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = MODE != WO ? in[(size_t)(r + u * RPS) * C + c] : r ^ c;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const uint32_t rr = r + u * RPS;
        if constexpr (MODE == RO) {
          acc += v[u];
        } else {
          put<NTS>(out + (size_t)(2 * rr) * C + c, v[u]);
          put<NTS>(out + (size_t)(2 * rr + 1) * C + c, v[u] + 1u);
        }
      }
    }
  }
  if (MODE == RO && acc == 0x12345678u) sink[threadIdx.x] = acc;
}

// 16-byte shape of a W-column tile: lane q = t % (W / 4) covers columns
// 4q..4q+3, thread t covers row t / (W / 4) + k NT / (W / 4).
template <int W, int NT, int MODE, bool NTS>
__global__ __launch_bounds__(NT) void stream_x4(const v4u* __restrict__ in, v4u* __restrict__ out,
                                                uint32_t* sink) {
  constexpr uint32_t QW = W / 4, RPS = NT / QW, CQ = C / 4;
  static_assert(S % (4 * RPS) == 0, "whole row batches");
  const uint32_t ntiles = C / W, q = threadIdx.x % QW, r0 = threadIdx.x / QW;
  uint32_t acc = 0;
  for (uint32_t l = blockIdx.x; l < ntiles; l += gridDim.x) {
    const uint32_t c = xcd_tile(l, ntiles) * QW + q;
    for (uint32_t r = r0; r < S; r += 4 * RPS) {
      v4u v[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        v[u] = MODE != WO ? in[(size_t)(r + u * RPS) * CQ + c] : v4u{r, c, (uint32_t)u, 0u};
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t rr = r + u * RPS;
        if constexpr (MODE == RO) {
          acc += v[u].x ^ v[u].w;
        } else {
          put<NTS>(out + (size_t)(2 * rr) * CQ + c, v[u]);
          put<NTS>(out + (size_t)(2 * rr + 1) * CQ + c, v[u]);
        }
      }
    }
  }
  if (MODE == RO && acc == 0x12345678u) sink[threadIdx.x] = acc;
}

template <class K, class... A>
float time_kernel(K k, int grid, int nt, A... args) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  std::vector<float> t;
  for (int i = 0; i < 6; i++) {  // one warm-up, median of five
    CK(hipEventRecord(a, 0));
    hipLaunchKernelGGL(k, dim3(grid), dim3(nt), 0, 0, args...);
    CK(hipGetLastError());
    CK(hipEventRecord(b, 0));
    CK(hipEventSynchronize(b));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, a, b));
    if (i) t.push_back(ms);
  }
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

const char* mode_name(int m) { return m == RW ? "read+write" : m == WO ? "write-only" : "read-only "; }
double mode_bytes(int m) { return (m == RW ? 12.0 : m == WO ? 8.0 : 4.0) * S * C; }

void report(const char* shape, int W, int nt, int per_cu, int mode, bool nts, float ms) {
  printf("| %-5s | %2d | %4d | %d | %s | %-5s | %.3f | %.1f |\n", shape, W, nt, per_cu, mode_name(mode),
         nts ? "nt" : "plain", ms, mode_bytes(mode) / ms / 1e6);
  fflush(stdout);
}

template <int W, int NT, int MODE, bool NTS>
void dword_line(const uint32_t* in, uint32_t* out, uint32_t* sink, int cus, int per_cu) {
  const float ms = time_kernel(stream_dword<W, NT, MODE, NTS>, cus * per_cu, NT, in, out, sink);
  report("dword", W, NT, per_cu, MODE, NTS, ms);
}

template <int W, int NT, int MODE, bool NTS>
void x4_line(const uint32_t* in, uint32_t* out, uint32_t* sink, int cus, int per_cu) {
  const float ms = time_kernel(stream_x4<W, NT, MODE, NTS>, cus * per_cu, NT, reinterpret_cast<const v4u*>(in),
                               reinterpret_cast<v4u*>(out), sink);
  report("x4", W, NT, per_cu, MODE, NTS, ms);
}

int main() {
  int cus = 0;
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  uint32_t *in = nullptr, *out = nullptr, *sink = nullptr;
  CK(hipMalloc(&in, (size_t)S * C * 4));
  CK(hipMalloc(&out, (size_t)2 * S * C * 4));
  CK(hipMalloc(&sink, 4096));
  CK(hipMemset(in, 1, (size_t)S * C * 4));
  CK(hipMemset(out, 0, (size_t)2 * S * C * 4));
  CK(hipDeviceSynchronize());
  {  // a plain device copy of the input's 2 GiB, for scale
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    std::vector<float> t;
    for (int i = 0; i < 6; i++) {
      CK(hipEventRecord(a, 0));
      CK(hipMemcpyAsync(out, in, (size_t)S * C * 4, hipMemcpyDeviceToDevice, 0));
      CK(hipEventRecord(b, 0));
      CK(hipEventSynchronize(b));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, a, b));
      if (i) t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    printf("device copy of 2 GiB: %.3f ms, %.1f GB/s (read + write)\n", t[2], 2.0 * S * C * 4 / t[2] / 1e6);
  }
  printf("\n%u x %u u32 in, %u x %u out; GB/s = bytes moved (12 / 8 / 4 per element) / median time\n\n", S, C,
         2 * S, C);
  printf("| shape | W | threads | WG/CU | traffic | stores | ms | GB/s |\n|---|---|---|---|---|---|---|---|\n");
  dword_line<8, 512, RW, false>(in, out, sink, cus, 2);
  dword_line<8, 512, RW, true>(in, out, sink, cus, 2);
  dword_line<16, 512, RW, false>(in, out, sink, cus, 2);
  dword_line<16, 512, RW, true>(in, out, sink, cus, 2);
  dword_line<32, 512, RW, false>(in, out, sink, cus, 2);
  dword_line<32, 512, RW, true>(in, out, sink, cus, 2);
  dword_line<64, 512, RW, false>(in, out, sink, cus, 2);
  dword_line<64, 512, RW, true>(in, out, sink, cus, 2);
  dword_line<16, 1024, RW, true>(in, out, sink, cus, 1);
  x4_line<32, 512, RW, false>(in, out, sink, cus, 1);
  x4_line<32, 512, RW, true>(in, out, sink, cus, 1);
  x4_line<32, 512, RW, true>(in, out, sink, cus, 2);
  x4_line<16, 512, RW, false>(in, out, sink, cus, 1);
  x4_line<16, 512, RW, true>(in, out, sink, cus, 1);
  x4_line<8, 512, RW, false>(in, out, sink, cus, 1);
  x4_line<64, 512, RW, true>(in, out, sink, cus, 1);
  dword_line<8, 512, WO, true>(in, out, sink, cus, 2);
  dword_line<16, 512, WO, true>(in, out, sink, cus, 2);
  dword_line<32, 512, WO, true>(in, out, sink, cus, 2);
  dword_line<64, 512, WO, true>(in, out, sink, cus, 2);
  x4_line<32, 512, WO, true>(in, out, sink, cus, 1);
  dword_line<8, 512, RO, false>(in, out, sink, cus, 2);
  dword_line<16, 512, RO, false>(in, out, sink, cus, 2);
  dword_line<32, 512, RO, false>(in, out, sink, cus, 2);
  dword_line<64, 512, RO, false>(in, out, sink, cus, 2);
  x4_line<32, 512, RO, false>(in, out, sink, cus, 1);
  CK(hipFree(in));
  CK(hipFree(out));
  CK(hipFree(sink));
  printf("\ndone\n");
  return 0;
}
