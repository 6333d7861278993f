// Shared host-side plumbing of libzoda_hip: status codes, the per-thread last
// error, HIP call checking and the per-device state (cached constant tables,
// reusable device buffers and streams).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/zoda_hip.h"
#include "field.hpp"
#include "fp64.hpp"

namespace zoda {

// Records a formatted message as this thread's last error and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Records a failed HIP call in rc as its hipError_t (positive), unless rc
// already holds an earlier error.  The call is evaluated either way, so
// clean-up calls still run after a failure.
inline void hip_note(int& rc, hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  (void)hipGetLastError();
  if (rc == 0) rc = fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

// Status of a failed allocation: ZODA_ENOMEM when the device is out of
// memory, the positive hipError_t for any other cause (no device, ...).
inline int alloc_fail(hipError_t e, const char* what) {
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return fail(ZODA_ENOMEM, "%s: out of device memory", what);
  return fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

// Evaluates a HIP call; on failure records the error and returns it (positive).
#define ZODA_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::zoda::fail((int)e_, "%s failed: %s (%s:%d)", #call,                \
                          hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

// Propagates a non-zero status.
#define ZODA_TRY(expr)         \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != 0) return rc_;  \
  } while (0)

inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline uint32_t log2_ceil(uint64_t x) {
  uint32_t l = 0;
  while ((1ull << l) < x) l++;
  return l;
}

// Montgomery-form radix-2 twiddles of one transform size on one device:
// tw[h + j] = w_(2h)^(+-j) * 2^32 mod p for h = 1, 2, 4, ..., n/2 and j < h.
struct MontTables {
  uint32_t* fwd = nullptr;
  uint32_t* inv = nullptr;
};

struct EncodeTabs;
// Device copy of the encode tables of one (side, digit split): the (w, w/p)
// buffer, its packed u32 copy, and the EncodeTabs pointing into both.
struct EncodeTableSet {
  Tw* base = nullptr;
  uint32_t* packed = nullptr;
  EncodeTabs* tabs = nullptr;  // host-side struct (owned)
};

struct DecodeTabs;
// Device copy of the erasure decoder's LDS tables for one n and the
// DecodeTabs pointing into it.
struct DecodeTableSet {
  uint32_t* packed = nullptr;
  DecodeTabs* tabs = nullptr;  // host-side struct (owned)
};

struct DeviceState {
  int device = 0;
  std::recursive_mutex mu;
  std::map<uint32_t, MontTables> mont;           // standard roots, keyed by log n
  std::map<uint32_t, EncodeTableSet> encode;     // keyed by encode_table_key()
  std::map<uint32_t, DecodeTableSet> decode;     // keyed by log n
  // Autotuned batched plan per log side: the work-queue plan variant, and
  // whether the side's squares run as their own fused launches instead.
  int batch_variant[16] = {0};
  bool batch_direct[16] = {false, false, false, false, false, false, false, false,
                           true,  true,  true,  true,  true,  true,  true,  true};
  std::map<uint32_t, uint32_t> batch_grid;       // batched kernels' persistent grid by LDS bytes
  // Drop-in entry points: tables built from the caller's roots.
  std::map<uint32_t, std::vector<uint32_t>> user_roots;
  std::map<uint32_t, MontTables> user_mont;
  // Grow-only scratch buffers and streams for the host-memory entry points.
  std::vector<void*> buffers;
  std::vector<size_t> buffer_bytes;
  std::vector<hipStream_t> streams;
  // Per-kernel launch set-up done once: persistent grid size (0: the LDS
  // limit is raised, no grid needed).
  std::map<const void*, uint32_t> kernel_setup;
  // Twiddle tables of the slice-pipelined 2048 encode (encode_slice.hip).
  void* slice_tables[2] = {nullptr, nullptr};  // sides 2048, 4096
  // This device's address of the RLC's static scratch array (rlc.hip).
  uint32_t* rlc_scratch = nullptr;
  // Twiddle tables of the FP64 poly transforms (ntt_poly.hip), keyed by
  // log n << 1 | inverse; whether a size's caller roots are the standard ones.
  std::map<uint32_t, uint32_t*> ntt_fp64;
  std::map<uint32_t, bool> user_standard;
  // The library's own stream-ordered pool for kernel scratch (pool_alloc).
  hipMemPool_t pool = nullptr;
};

// State of the calling thread's current device (created on first use).
int current_state(DeviceState** out);
int get_buffer(DeviceState* st, int slot, size_t bytes, void** out);
int get_streams(DeviceState* st, int count, hipStream_t** out);
// Stream-ordered scratch from the library's pool (freed with hipFreeAsync on
// the same stream).  The pool keeps up to POOL_KEEP_BYTES between calls, so a
// caller that synchronizes after every call does not pay an allocation from
// the system each time (the default pool returns everything at each
// synchronization).
constexpr uint64_t POOL_KEEP_BYTES = 1ull << 30;
int pool_alloc(DeviceState* st, size_t bytes, hipStream_t stream, void** out);

// Standard-root Montgomery tables for size 2^log_n.
int get_mont_tables(DeviceState* st, uint32_t log_n, MontTables* out);
// Montgomery tables from caller roots (roots[i] = w^i, i < n).
int get_user_mont_tables(DeviceState* st, const uint32_t* roots, uint32_t log_n, MontTables* out);

// Batched poly-major transform on device memory (ntt_poly.hip): radix-2
// Montgomery stages with the tables' roots.
int launch_ntt_poly(uint32_t* d_values, uint32_t log_n, uint32_t polys, uint32_t stride,
                    int inverse, const MontTables& tabs, hipStream_t stream);
// The same transform, on FP64 digit passes when the roots are the standard
// ones and the size, stride and alignment allow (else launch_ntt_poly).
int launch_ntt(DeviceState* st, uint32_t* d_values, uint32_t log_n, uint32_t polys, uint32_t stride,
               int inverse, const MontTables& tabs, bool standard_roots, hipStream_t stream);
// Whether caller roots of size 2^log_n are the standard roots (cached).
bool roots_are_standard(DeviceState* st, const uint32_t* roots, uint32_t log_n);

inline uint32_t encode_table_key(int L, int r1, int r2, int rd) {
  return (uint32_t)L | ((uint32_t)r1 << 8) | ((uint32_t)r2 << 16) | ((uint32_t)rd << 24);
}
// Constant tables of the fused encode for side 2^L and digit split (r1, r2, rd).
int get_encode_tables(DeviceState* st, int L, int r1, int r2, int rd, const EncodeTableSet** out);

// Fused column encode on device memory (encode.hip).  With dry, only the
// one-time set-up (tables, kernel attributes) is done and nothing launches --
// the set-up calls are not allowed inside a stream capture.
int launch_encode(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols,
                  uint32_t n, uint32_t* d_out, int layout, hipStream_t stream, bool dry = false);

// Once per device and kernel: raise the dynamic LDS limit when the kernel
// needs more than 64 KiB and, for a persistent kernel, size its grid to fill
// every CU (*grid; 0 otherwise).
int kernel_setup(DeviceState* st, const void* fn, int nt, size_t lds_bytes, bool persistent,
                 uint32_t* grid);

// The slice-pipelined encode of sides 2048 and 4096 (whole 32- resp.
// 16-column tiles, row layout, offsets within its addressing limits): true
// when it takes the shape.
bool slice_fits(uint32_t k, uint32_t cols, int layout);
int launch_slice(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t* d_out,
                 hipStream_t stream, bool dry);

// Column-major encode of a whole square by transpose + column kernel
// (encode_cols.hip): cols_fits says whether it takes the shape on `stream`
// (not while the stream is being captured: it allocates its scratch).
bool cols_fits(const uint32_t* d_in, uint32_t k, uint32_t cols, const uint32_t* d_out, hipStream_t stream);
int launch_cols(DeviceState* st, const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t* d_out,
                hipStream_t stream);

// Before tuning, squares of side >= 2^DIRECT_MIN_LOG run as their own fused
// launches in batched and served work (their kernel fills the device), next
// to the persistent work queue of the smaller squares; the autotuner decides
// per side (DeviceState::batch_direct).
constexpr uint32_t DIRECT_MIN_LOG = 8;

// Seeded uniform generator (api.hip).
int launch_fill_uniform(uint32_t* d_out, size_t count, uint64_t seed, hipStream_t stream);

}  // namespace zoda
