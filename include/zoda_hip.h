/*
 * zoda_hip.h -- C ABI of the MI355X (gfx950) ZODA column encoder.
 *
 * All arithmetic is exact over BabyBear, p = 15 * 2^27 + 1 = 2013265921.
 * Values cross this boundary as canonical u32 in [0, p).  Every entry point
 * returns an int status: 0 on success, a negative ZODA_E* code for an argument
 * or resource error, or a positive hipError_t from the HIP runtime.  No entry
 * point aborts the process; the message of the last failure on the calling
 * thread is available from zoda_hip_last_error().
 *
 * Reference interfaces replaced (jonas089/zoda-cu):
 *   src/ntt/cuda.rs:15-22          extern "C" block of the Rust crate
 *   cuda/ntt_kernel.cu:184-208     cuda_host_alloc / cuda_host_free /
 *                                  cuda_ntt / cuda_intt
 *   src/benchmarks/utils.rs:100-146 encode_gpu (transpose + pad + INTT + NTT)
 *   src/ntt/cpu.rs:97-110          roots_of_unity (host side of run(),
 *                                  src/ntt/cuda.rs:112)
 */
#ifndef ZODA_HIP_H
#define ZODA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZODA_P 2013265921u

#define ZODA_OK 0
#define ZODA_EINVAL (-1)  /* bad argument: n not a power of two, stride < n, ... */
#define ZODA_ENOMEM (-2)  /* device or pinned host allocation failed */
#define ZODA_ENODEV (-3)  /* no HIP device visible */
#define ZODA_EBUDGET (-4) /* request larger than the configured memory budget */
#define ZODA_ESTATE (-5)  /* object used in the wrong state */

/* Output layouts of the encode entry points. */
#define ZODA_LAYOUT_ROW 0 /* out[r * cols + c]: the extended square, row-major */
#define ZODA_LAYOUT_COL 1 /* out[c * n_kn + r]: PinnedSquare layout, src/ntt/cuda.rs:31-67 */

/* Message of the last failed call on this thread ("" if none). */
const char* zoda_hip_last_error(void);

/* ---- drop-in for the reference FFI --------------------------------------- */

/* Replaces the direct cudaGetDeviceCount binding, src/ntt/cuda.rs:21,25-29.
 * Writes the number of visible HIP devices (0 when none) and returns 0. */
int zoda_hip_device_count(int* count);

/* Replaces cuda_host_alloc, cuda/ntt_kernel.cu:184-187: `count` u32 of pinned
 * host memory, NULL on failure (and for count == 0). */
uint32_t* zoda_hip_host_alloc(size_t count);

/* Replaces cuda_host_free, cuda/ntt_kernel.cu:188.  NULL is a no-op. */
void zoda_hip_host_free(uint32_t* p);

/* Replaces cuda_ntt, cuda/ntt_kernel.cu:190-193 / src/ntt/cuda.rs:17.
 * `values` is host memory (pinned or pageable) holding `polys` polynomials
 * `stride` u32 apart; the first `n` values of each are replaced by their
 * natural-order NTT, X[k] = sum_j x_j roots[(j*k) mod n]; values n..stride-1
 * of each polynomial are left untouched.  `roots` holds n canonical values
 * roots[i] = w^i for a primitive n-th root of unity w (as produced by
 * zoda_hip_roots_of_unity); the transform uses exactly the roots given.
 * Synchronous. */
int zoda_hip_ntt(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys,
                 uint32_t stride);

/* Replaces cuda_intt, cuda/ntt_kernel.cu:196-208 / src/ntt/cuda.rs:18.  Same
 * contract; computes x_j = n^-1 sum_k X_k roots[(-j*k) mod n]. */
int zoda_hip_intt(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys,
                  uint32_t stride);

/* The n roots of unity of src/ntt/cpu.rs:97-110: w = 31^((p-1)/n) (31 is the
 * smallest generator of F_p^*), roots[i] = w^i.  n must be a power of two
 * dividing p - 1 (n <= 2^27). */
int zoda_hip_roots_of_unity(uint32_t n, uint32_t* roots);

/* ---- device-resident transforms and encode -------------------------------- */

/* Batched NTT (inverse = 0) or INTT (inverse = 1) on device memory, same
 * layout and result as zoda_hip_ntt / zoda_hip_intt with the standard roots.
 * Asynchronous on `stream` (a hipStream_t, NULL for the null stream). */
int zoda_hip_ntt_device(uint32_t* d_values, uint32_t n, uint32_t polys, uint32_t stride,
                        int inverse, void* stream);

/* ZODA column encode of encode_gpu (src/benchmarks/utils.rs:100-146) on device
 * memory.  d_in is k rows of `cols` canonical values, row-major.  Each column
 * is zero-padded to n_k = next_pow2(k), inverse-transformed at size n_k,
 * zero-padded to n_kn = next_pow2(k + n) and transformed at size n_kn.  d_out
 * receives n_kn x cols values in `layout`.  d_in and d_out must not overlap.
 * Asynchronous on `stream`. */
int zoda_hip_encode_device(const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t n,
                           uint32_t* d_out, int layout, void* stream);

/* The same encode end to end from host memory: h_in row-major k x cols, h_out
 * n_kn x cols in `layout`.  Column slabs are streamed through the device with
 * copies overlapping the kernels; no host-side transpose.  Pinned buffers
 * (zoda_hip_host_alloc) give full PCIe bandwidth.  Synchronous. */
int zoda_hip_encode_host(const uint32_t* h_in, uint32_t k, uint32_t cols, uint32_t n,
                         uint32_t* h_out, int layout);

/* Fills d_out[i] for i < count with the seeded uniform BabyBear generator:
 * for attempt a = 0, 1, ...: z = splitmix64(seed + (64*i + a + 1) * 0x9E3779B97F4A7C15),
 * v = z >> 33; the first v < p is the value.  Asynchronous on `stream`. */
int zoda_hip_fill_uniform(uint32_t* d_out, size_t count, uint64_t seed, void* stream);

/* ---- batched encode (no counterpart in the reference) --------------------- */

/* One square of a batch: device buffers, row-major in and out (n_kn x cols
 * out, n_kn = next_pow2(k + n) = 2 * next_pow2(k)); k <= 4096. */
typedef struct zoda_square {
  const uint32_t* in;
  uint32_t* out;
  uint32_t k;
  uint32_t cols;
  uint32_t n;
  uint32_t reserved; /* 0 */
} zoda_square;

/* Encodes `count` squares of any mix of sides: squares of side < 256 in one
 * persistent launch that drains a device work queue of (square, column tile)
 * items, largest side first; squares of side >= 256 (whose own fused kernel
 * fills the device) as their own launches on a parallel branch.  Synchronous
 * on `stream`.  Results equal zoda_hip_encode_device of each square alone.
 * No counterpart in the reference (a batching pattern over its encode). */
int zoda_hip_encode_batch(const zoda_square* squares, uint32_t count, void* stream);

/* A batch captured once into a hipGraph ([queue reset, persistent kernel])
 * and replayed by zoda_hip_batch_plan_launch (asynchronous on `stream`).  The
 * squares' buffers must outlive the plan; their contents may change between
 * replays. */
typedef struct zoda_batch_plan zoda_batch_plan;
int zoda_hip_batch_plan_create(const zoda_square* squares, uint32_t count, zoda_batch_plan** plan);
int zoda_hip_batch_plan_launch(zoda_batch_plan* plan, void* stream);
int zoda_hip_batch_plan_destroy(zoda_batch_plan* plan);

/* Autotuner of batched work: for each side (a power of two <= 4096) it times
 * ZODA_TUNE_CANDIDATES candidates on a synthetic side x 128*side square --
 * the ZODA_BATCH_VARIANTS plans of the persistent work-queue kernel (tile
 * width, digit split, twiddles in LDS or global memory), and the side's own
 * fused launch (ZODA_BATCH_DIRECT) -- and records the fastest in this
 * device's table, which zoda_hip_encode_batch, batch plans and the server
 * follow.  times_us, if not NULL, receives count * ZODA_TUNE_CANDIDATES
 * microsecond timings (candidate order: queue variants 0..3, direct).
 * get/set read and override the entry of log2(side) (0..3: the queue with
 * that plan; ZODA_BATCH_DIRECT: own launches).  info[6] = {W, threads, R1,
 * R2, RD, twiddles in LDS} of a candidate.  save/load persist the table as
 * text, one line per side keyed by device name and CU count (load takes the
 * lines of this device; *entries = how many). */
#define ZODA_BATCH_VARIANTS 4
#define ZODA_BATCH_DIRECT 4
#define ZODA_TUNE_CANDIDATES 5
int zoda_hip_autotune(const uint32_t* sides, uint32_t count, int iters, float* times_us);
/* The same candidates tuned for one batch as it runs: for each side the
 * batch holds (largest first) each candidate is timed as the whole batch with
 * the other sides at their current entries, and the fastest kept (one sweep
 * of coordinate descent from the current table).  Writes the squares' outputs
 * as zoda_hip_encode_batch would.  times_us (optional): 13 *
 * ZODA_TUNE_CANDIDATES floats, [log side][candidate], 0 for absent sides. */
int zoda_hip_autotune_batch(const zoda_square* squares, uint32_t count, int iters, float* times_us);
int zoda_hip_get_variant(uint32_t log_side, int* variant);
int zoda_hip_set_variant(uint32_t log_side, int variant);
int zoda_hip_batch_variant_info(uint32_t log_side, int variant, int* info);
int zoda_hip_tune_save(const char* path);
int zoda_hip_tune_load(const char* path, uint32_t* entries);

/* ---- serving (no counterpart in the reference) ----------------------------- */

typedef struct zoda_serve_stats {
  double makespan_us;       /* first arrival to last completion */
  uint32_t batches;         /* launches: batched launches (server) or requests (plain loop) */
  uint32_t max_batch;       /* most requests in one launch */
  uint64_t peak_bytes;      /* largest footprint (inputs + outputs) admitted at once */
  uint32_t rejected;        /* requests larger than the whole budget */
  uint64_t pool_peak_bytes; /* most device memory the output pool held (<= budget) */
  uint64_t pool_held_bytes; /* device memory the output pool holds now (in use + cached) */
} zoda_serve_stats;

/* Continuous-batching encode server on the current device.  Requests are
 * device-resident squares (zoda_square); a request whose `out` is NULL gets
 * its output from the server's device pool at admission (all pool memory,
 * cached or in use, stays within hbm_budget; the request waits while its
 * block does not fit).  Each of `slots` launch slots
 * (1..16) owns a stream.  Whenever one is free it takes the next launch: a
 * large request (side >= 256) at the head of the queue alone through the
 * fused single-square kernel (at most slots - 1 at once, in arrival order),
 * or else every queued small request that fits `hbm_budget` next to the
 * requests on the GPU as one persistent batched launch, overlapping the
 * launches in flight.  A scheduler thread runs until zoda_server_destroy.  A
 * request larger than the whole budget is rejected with ZODA_EBUDGET (its
 * ticket reads state 3).  No counterpart in the reference (a serving pattern
 * over its encode). */
typedef struct zoda_server zoda_server;
int zoda_server_create(uint64_t hbm_budget, uint32_t slots, zoda_server** server);
int zoda_server_submit(zoda_server* server, const zoda_square* request, uint64_t* ticket);
/* state: 0 queued, 1 on the GPU, 2 done (latency_us = completion - arrival), 3 rejected */
int zoda_server_poll(zoda_server* server, uint64_t ticket, int* state, double* latency_us);
/* The pool output of a finished request (state 2) whose `out` was NULL: the
 * device pointer stays valid until zoda_server_release returns it to the
 * pool, which may hand it to the next request at once: the caller's reads of
 * it must be complete (their stream synchronized) before the release.  A finished request is forgotten once it has been polled and holds
 * no pool output; its ticket is then unknown (ZODA_EINVAL). */
int zoda_server_output(zoda_server* server, uint64_t ticket, uint32_t** d_out);
int zoda_server_release(zoda_server* server, uint64_t ticket);
int zoda_server_wait_all(zoda_server* server);
int zoda_server_stats(zoda_server* server, zoda_serve_stats* stats);
int zoda_server_destroy(zoda_server* server);

/* Replays a request trace in real time on this thread: request i arrives
 * arrival_us[i] microseconds after the start (non-decreasing).  mode 0 runs
 * the continuous-batching server above (outputs of requests with a NULL `out`
 * come from its pool and go back at completion); mode 1 the plain loop, one
 * zoda_hip_encode_device per request in arrival order on one stream
 * (requests with a NULL `out` write to one block per output size).
 * latency_us[i] receives completion minus arrival (-1 if rejected). */
int zoda_hip_serve_trace(const zoda_square* requests, const double* arrival_us, uint32_t count,
                         uint64_t hbm_budget, uint32_t slots, int mode, double* latency_us,
                         zoda_serve_stats* stats);

/* ---- Random linear combinations of encoded rows (commitment checks) ----
 *
 * Replaces the CPU loops of src/zoda/rlc.rs:32-66 and the data root of
 * src/benchmarks/utils.rs:51-59, as used by the validation of utils.rs:221-272
 * and the encoder check of src/zoda/babybear.rs:164-192. */

/* Coefficients of columns 0..cols-1 derived from the data root text
 * (generate_deterministic_coefficients, rlc.rs:48-66): SHA-256 of the text
 * followed by the column index as 8 little-endian bytes; limb l is the
 * big-endian u64 in digest bytes [8l, 8l+8) mod p.  d_coeffs[c*4 + l].
 * Synchronous on `stream`. */
int zoda_hip_rlc_coefficients(const char* data_root, uint32_t root_len, uint32_t cols, uint32_t* d_coeffs,
                              void* stream);

/* rlc_row (rlc.rs:32-40) of `count` rows of the row-major rows x cols device
 * matrix d_square: row d_row_index[i] (device array; NULL means rows
 * 0..count-1, and then count <= rows), d_out[i*4 + l] = sum over columns of
 * row[c] * d_coeffs[c*4 + l] mod p (d_coeffs 16-byte aligned).  A row index
 * >= rows gives zeros.
 * Asynchronous on `stream`. */
int zoda_hip_rlc_rows(const uint32_t* d_square, uint32_t rows, uint32_t cols, const uint32_t* d_row_index,
                      uint32_t count, const uint32_t* d_coeffs, uint32_t* d_out, void* stream);

/* zoda_hip_rlc_rows without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The partial sums
 * live in a static 16 MiB device array instead of a per-call allocation:
 * count * ceil(cols / 2048) must be <= 2^20 (e.g. 8192 rows of 262144
 * columns), larger calls return ZODA_EINVAL.  Same arguments, results and
 * stream semantics as zoda_hip_rlc_rows.  Every call on a device shares that
 * device's array: calls that overlap on one GPU (concurrent streams, or a
 * second call queued before the first has finished on another stream) race;
 * calls ordered on one stream do not. */
int zoda_hip_rlc_rows_noalloc(const uint32_t* d_square, uint32_t rows, uint32_t cols,
                              const uint32_t* d_row_index, uint32_t count, const uint32_t* d_coeffs,
                              uint32_t* d_out, void* stream);

/* compute_data_root (utils.rs:51-59): SHA-256 of `count` host u32 values as
 * little-endian bytes in order, written as 64 lowercase hex digits and a NUL
 * to hex_out (65 bytes).  Runs on the host (one sequential hash chain). */
int zoda_data_root(const uint32_t* values, size_t count, char* hex_out);

/* Generic-modulus transforms: ntt / intt of src/ntt/cpu.rs:1-32,45-60 over
 * any modulus in [2, 2^31) with the caller's n roots (roots[i] = w^i), with
 * the reference's exact arithmetic (including its n^(m-2) inverse of n).
 * Host memory, `polys` polynomials `stride` u32 apart, the first n (a power
 * of two <= 8192) of each transformed in place; inverse != 0 for intt. */
int zoda_hip_ntt_mod(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t polys, uint32_t stride,
                     uint32_t modulus, int inverse);

/* ---- Erasure decoding ----
 *
 * The data square back from any k of the 2k rows of its extended square
 * (device memory, row-major).  No counterpart in the reference, which only
 * checks that every column is a Reed-Solomon codeword with the data at the
 * even rows (src/zoda/babybear.rs:195-211); this inverts that map.
 * d_enc: 2k x cols, rows not listed are never read.  known_rows: host array of
 * row indices (the first k distinct ones are used).  d_data: k x cols.  k a
 * power of two <= 16384 (one column of 2k values is transformed in LDS).
 * Scratch: 4k * cols bytes of device memory for 128 <= k <= 8192, 8k * cols
 * otherwise.  Synchronous on `stream`. */
int zoda_hip_decode_device(const uint32_t* d_enc, uint32_t k, uint32_t cols, const uint32_t* known_rows,
                           uint32_t nknown, uint32_t* d_data, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ZODA_HIP_H */
