// Batched encode of many squares of uneven sides in one persistent launch.
//
// Work items are (square, column tile) pairs, ordered largest side first.  A
// grid sized to fill every CU drains them through one atomic counter: each
// work-group takes an item, runs all passes of that tile with the plan of the
// square's side (every pass group of a tile is fused, the tile being
// LDS-resident), and takes the next.  The counter is reset by a one-thread
// kernel, so [reset, persistent kernel] can be captured once into a hipGraph
// and replayed (zoda_hip_batch_plan_*).  The plan variant of each side comes
// from the per-device autotune table (zoda_hip_autotune).
#include <stdio.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "encode_plans.hpp"
#include "encode_tables.hpp"

namespace zoda {

// Two instantiations of each batched kernel: one for every side, and one for
// queues of small squares only (below 2^DIRECT_MIN_LOG, as when the large
// squares run as their own launches).  A kernel's register budget is its largest
// plan's, so the small one fits several times more work-groups on a CU.
constexpr int BATCH_SMALL_LOG = DIRECT_MIN_LOG - 1;
constexpr int BATCH_BRANCHES = 2;
static_assert(ZODA_BATCH_VARIANTS == BATCH_VARIANTS && ZODA_BATCH_DIRECT == BATCH_DIRECT &&
                  ZODA_TUNE_CANDIDATES == BATCH_VARIANTS + 1,
              "the header's tuning constants match the plans");  // side branches of the batch's direct (large-square) launches

template <class PL, bool FULL>
__device__ __forceinline__ void batch_tile_body(int tid, uint32_t tile, const SquareDesc& d, int32_t* lds,
                                                const EncodeTabs& T, bool load) {
  if constexpr (PL::TWL) {
    if (load) {  // uniform: the work-group changed plan since its last tile
      encode_load_tables<PL>(tid, lds, T);
      __syncthreads();
    }
  }
  encode_step<PL, 0, FULL>(tid, tile, d.in, d.out, d.k, d.cols, 0, lds, T);
  if constexpr (PL::STEPS > 1) {
    __syncthreads();
    encode_step<PL, 1, FULL>(tid, tile, d.in, d.out, d.k, d.cols, 0, lds, T);
    __syncthreads();
    encode_step<PL, 2, FULL>(tid, tile, d.in, d.out, d.k, d.cols, 0, lds, T);
  }
  if constexpr (PL::STEPS > 3) {
    __syncthreads();
    encode_step<PL, 3, FULL>(tid, tile, d.in, d.out, d.k, d.cols, 0, lds, T);
    __syncthreads();
    encode_step<PL, 4, FULL>(tid, tile, d.in, d.out, d.k, d.cols, 0, lds, T);
  }
}

// Not inlined: each plan is register-allocated on its own, so a kernel's VGPR
// budget is its largest plan's rather than a merge of all of them.  MAXL
// keeps the instances each kernel calls apart, so the small-side kernel's
// occupancy bound reaches its plans.
template <class PL, bool FULL, int MAXL>
__device__ __attribute__((noinline)) void batch_tile_steps(int tid, uint32_t tile, const SquareDesc& d,
                                                           int32_t* lds, const EncodeTabs& T, bool load) {
  batch_tile_body<PL, FULL>(tid, tile, d, lds, T, load);
}

__global__ void batch_reset(uint32_t* counter) { *counter = 0u; }

// All passes of column tile `tile` of square d with the batched plan of its
// side (sides up to 2^MAXL; the launcher picks the kernel).  *loaded names the
// plan whose twiddle tables are in LDS (-1: none); they are reloaded when the
// plan changes.  The caller's barrier after the previous tile separates its
// last reads of LDS from the reload.
template <int MAXL>
__device__ __forceinline__ void run_batch_tile(int tid, const SquareDesc& d, uint32_t tile,
                                               int32_t* lds, const EncodeTabs& T, int* loaded) {
  const int key = (int)(d.L * BATCH_VARIANTS + d.variant);
  const bool load = key != *loaded;
  *loaded = key;
  with_batch_plan((int)d.L, (int)d.variant, [&](auto plan) -> int {
    using PL = decltype(plan);
    if constexpr (PL::L <= MAXL) {
      if (d.full)
        batch_tile_steps<PL, true, MAXL>(tid, tile, d, lds, T, load);
      else
        batch_tile_steps<PL, false, MAXL>(tid, tile, d, lds, T, load);
    }
    return 0;
  });
}

// Next work item of the work-group (one atomic per work-group, broadcast
// through LDS); uniform across the work-group.
__device__ __forceinline__ uint32_t take_item(int tid, uint32_t* counter, uint32_t base) {
  __shared__ uint32_t s_item;
  if (tid == 0) s_item = atomicAdd(counter, 1u) - base;
  __syncthreads();
  const uint32_t it = s_item;
  __syncthreads();
  return it;
}

// Waves per SIMD each instantiation is compiled for.
template <int MAXL>
constexpr int batch_min_waves() { return MAXL <= BATCH_SMALL_LOG ? 4 : 1; }

template <int MAXL>
__global__ __launch_bounds__(BATCH_NT, batch_min_waves<MAXL>()) void batch_kernel(const SquareDesc* __restrict__ squares,
                                                         const BatchItem* __restrict__ items,
                                                         uint32_t nitems, uint32_t* counter,
                                                         const EncodeTabs* __restrict__ tabs) {
  extern __shared__ int32_t lds[];
  const int tid = (int)threadIdx.x;
  int loaded = -1;
  for (;;) {
    const uint32_t it = take_item(tid, counter, 0);
    if (it >= nitems) break;
    const BatchItem bi = items[it];
    const SquareDesc d = squares[bi.square];
    run_batch_tile<MAXL>(tid, d, bi.tile, lds, tabs[d.L * BATCH_VARIANTS + d.variant], &loaded);
  }
}

template <int MAXL>
__global__ __launch_bounds__(BATCH_NT, batch_min_waves<MAXL>()) void serve_batch_kernel(const ServeBatch b) {
  extern __shared__ int32_t lds[];
  const int tid = (int)threadIdx.x;
  int loaded = -1;
  for (;;) {
    const uint32_t it = take_item(tid, b.counter, b.base);
    if (it >= b.items) break;
    uint32_t q = 0;
    while (q + 1 < b.n && b.first[q + 1] <= it) q++;
    const SquareDesc d = b.sq[q];
    run_batch_tile<MAXL>(tid, d, it - b.first[q], lds, b.tabs[d.L * BATCH_VARIANTS + d.variant], &loaded);
  }
}

uint32_t batch_lds_bytes(uint32_t L, uint32_t variant) {
  const int words = with_batch_plan((int)L, (int)variant, [](auto plan) -> int {
    return decltype(plan)::LDS_WORDS;
  });
  return (uint32_t)std::max(words, 1) * 4;
}

int batch_grid(DeviceState* st, uint32_t lds_bytes, uint32_t max_log, uint32_t* grid) {
  const bool small = (int)max_log <= BATCH_SMALL_LOG;
  const uint32_t key = lds_bytes | (small ? 0x80000000u : 0u);
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  auto it = st->batch_grid.find(key);
  if (it == st->batch_grid.end()) {
    int per_cu = 0, cus = 0;
    // The largest plans take more than the default 64 KiB of dynamic LDS.
    const void* fns[4] = {(const void*)batch_kernel<BATCH_MAX_LOG>, (const void*)serve_batch_kernel<BATCH_MAX_LOG>,
                          (const void*)batch_kernel<BATCH_SMALL_LOG>,
                          (const void*)serve_batch_kernel<BATCH_SMALL_LOG>};
    for (const void* fn : fns)
      ZODA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, BATCH_LDS_WORDS * 4));
    ZODA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, small ? fns[2] : fns[0], BATCH_NT, lds_bytes));
    ZODA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, st->device));
    it = st->batch_grid.emplace(key, (uint32_t)std::max(1, per_cu * cus)).first;
  }
  *grid = it->second;
  return 0;
}

int launch_batch_items(const SquareDesc* d_squares, const BatchItem* d_items, uint32_t nitems,
                       uint32_t* d_counter, const EncodeTabs* d_tabs, uint32_t grid, uint32_t lds_bytes,
                       uint32_t max_log, hipStream_t stream) {
  hipLaunchKernelGGL(batch_reset, dim3(1), dim3(1), 0, stream, d_counter);
  ZODA_HIP(hipGetLastError());
  if ((int)max_log <= BATCH_SMALL_LOG)
    hipLaunchKernelGGL(batch_kernel<BATCH_SMALL_LOG>, dim3(std::min(grid, nitems)), dim3(BATCH_NT), lds_bytes,
                       stream, d_squares, d_items, nitems, d_counter, d_tabs);
  else
    hipLaunchKernelGGL(batch_kernel<BATCH_MAX_LOG>, dim3(std::min(grid, nitems)), dim3(BATCH_NT), lds_bytes,
                       stream, d_squares, d_items, nitems, d_counter, d_tabs);
  ZODA_HIP(hipGetLastError());
  return 0;
}

int launch_serve_batch(const ServeBatch& b, uint32_t grid, uint32_t lds_bytes, uint32_t max_log,
                       hipStream_t stream, uint32_t* advance) {
  const uint32_t g = std::min(grid, b.items);
  if ((int)max_log <= BATCH_SMALL_LOG)
    hipLaunchKernelGGL(serve_batch_kernel<BATCH_SMALL_LOG>, dim3(g), dim3(BATCH_NT), lds_bytes, stream, b);
  else
    hipLaunchKernelGGL(serve_batch_kernel<BATCH_MAX_LOG>, dim3(g), dim3(BATCH_NT), lds_bytes, stream, b);
  ZODA_HIP(hipGetLastError());
  *advance = b.items + g;
  return 0;
}

int describe_square(DeviceState* st, const zoda_square& s, SquareDesc* d, EncodeTabs* tabs_by_slot,
                    bool* have_slot) {
  d->in = s.in;
  d->out = s.out;
  d->k = s.k;
  d->cols = s.cols;
  d->L = log2_ceil(s.k);
  d->variant = (uint32_t)st->batch_variant[d->L];
  return with_batch_plan((int)d->L, (int)d->variant, [&](auto plan) -> int {
    using PL = decltype(plan);
    d->tiles = (s.cols + PL::W - 1) / PL::W;
    d->full = (s.cols % PL::W == 0 && s.k == (uint32_t)PL::S) ? 1u : 0u;
    const size_t slot = (size_t)d->L * BATCH_VARIANTS + d->variant;
    if (!have_slot[slot]) {
      const EncodeTableSet* set = nullptr;
      ZODA_TRY(get_encode_tables(st, PL::L, PL::R1, PL::R2, PL::RD, &set));
      tabs_by_slot[slot] = *set->tabs;
      have_slot[slot] = true;
    }
    return 0;
  });
}

int check_batch_square(const zoda_square& s, uint32_t i) {
  if (!s.in || !s.out) return fail(ZODA_EINVAL, "square %u: NULL buffer", i);
  if (s.k == 0 || s.cols == 0) return fail(ZODA_EINVAL, "square %u: k and cols must be positive", i);
  const uint32_t L = log2_ceil(s.k);
  if ((int)L > BATCH_MAX_LOG)
    return fail(ZODA_EINVAL, "square %u: k = %u exceeds the batched maximum 2^%d", i, s.k, BATCH_MAX_LOG);
  if (log2_ceil((uint64_t)s.k + s.n) != L + 1)
    return fail(ZODA_EINVAL, "square %u: next_pow2(k + n) must be 2 * next_pow2(k)", i);
  return 0;
}

}  // namespace zoda

using namespace zoda;

struct zoda_batch_plan {
  int device = 0;
  SquareDesc* d_squares = nullptr;
  BatchItem* d_items = nullptr;
  EncodeTabs* d_tabs = nullptr;
  uint32_t* d_counter = nullptr;
  uint32_t nitems = 0, grid = 0, lds_bytes = 4, max_log = 0;
  // Squares that run as their own fused launches, largest first, dealt in
  // turn to BATCH_BRANCHES branches (side streams, fork/join events) parallel
  // to the work-queue kernel.
  std::vector<zoda_square> direct;
  hipStream_t side[BATCH_BRANCHES] = {};
  hipEvent_t fork = nullptr, join[BATCH_BRANCHES] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void free_plan(zoda_batch_plan* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->fork) (void)hipEventDestroy(p->fork);
  for (int b = 0; b < BATCH_BRANCHES; b++) {
    if (p->join[b]) (void)hipEventDestroy(p->join[b]);
    if (p->side[b]) (void)hipStreamDestroy(p->side[b]);
  }
  (void)hipFree(p->d_squares);
  (void)hipFree(p->d_items);
  (void)hipFree(p->d_tabs);
  (void)hipFree(p->d_counter);
  delete p;
}

// [queue reset, persistent kernel] on `stream`, the direct squares' fused
// kernels dealt in turn to the side branches, joined back into `stream`.
// Each fused kernel fills the GPU, but consecutive ones on different
// branches overlap one's last, partly idle wave of tiles with the next one's
// first (96-square batch: 1.61-1.62 ms on one branch, 1.46-1.48 ms on two).
static int launch_batch_direct(zoda_batch_plan* p, hipStream_t stream) {
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  const bool fork = !p->direct.empty();
  if (fork) {
    ZODA_HIP(hipEventRecord(p->fork, stream));
    for (int b = 0; b < BATCH_BRANCHES; b++) ZODA_HIP(hipStreamWaitEvent(p->side[b], p->fork, 0));
  }
  for (size_t i = 0; i < p->direct.size(); i++) {
    const zoda_square& q = p->direct[i];
    ZODA_TRY(launch_encode(st, q.in, q.k, q.cols, q.n, q.out, ZODA_LAYOUT_ROW, p->side[i % BATCH_BRANCHES]));
  }
  if (p->nitems > 0)
    ZODA_TRY(launch_batch_items(p->d_squares, p->d_items, p->nitems, p->d_counter, p->d_tabs, p->grid,
                                p->lds_bytes, p->max_log, stream));
  if (fork) {
    for (int b = 0; b < BATCH_BRANCHES; b++) {
      ZODA_HIP(hipEventRecord(p->join[b], p->side[b]));
      ZODA_HIP(hipStreamWaitEvent(stream, p->join[b], 0));
    }
  }
  return 0;
}

// Builds descriptors, the largest-first item list and the per-(side, variant)
// table pointers on the device; with `graph`, also captures the launch.  With
// allow_direct, squares of the sides the device's table marks direct run as
// their own fused launches beside the queue (the autotuner times the queue
// plans alone).
static int build_plan(const zoda_square* squares, uint32_t count, bool graph, bool allow_direct,
                      zoda_batch_plan** out) {
  clear_error();
  if (!out) return fail(ZODA_EINVAL, "plan pointer is NULL");
  *out = nullptr;
  if (!squares || count == 0) return fail(ZODA_EINVAL, "no squares");
  for (uint32_t i = 0; i < count; i++) ZODA_TRY(check_batch_square(squares[i], i));
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::lock_guard<std::recursive_mutex> lock(st->mu);

  std::vector<zoda_square> direct, queued;
  for (uint32_t i = 0; i < count; i++)
    (allow_direct && st->batch_direct[log2_ceil(squares[i].k)] ? direct : queued).push_back(squares[i]);
  std::stable_sort(direct.begin(), direct.end(),
                   [](const zoda_square& a, const zoda_square& b) { return a.k > b.k; });
  std::vector<SquareDesc> descs(queued.size());
  std::vector<EncodeTabs> tabs((BATCH_MAX_LOG + 1) * BATCH_VARIANTS, EncodeTabs{});
  std::unique_ptr<bool[]> have(new bool[tabs.size()]());
  uint64_t nitems = 0;
  for (size_t i = 0; i < queued.size(); i++) {
    ZODA_TRY(describe_square(st, queued[i], &descs[i], tabs.data(), have.get()));
    nitems += descs[i].tiles;
  }
  if (nitems >= (1ull << 32)) return fail(ZODA_EINVAL, "batch of %llu tiles is too large", (unsigned long long)nitems);
  // Largest sides first so the long tiles start early and the short ones fill the tail.
  std::vector<uint32_t> order(queued.size());
  for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return descs[a].L > descs[b].L; });
  std::vector<BatchItem> items;
  items.reserve(nitems);
  for (uint32_t i : order)
    for (uint32_t t = 0; t < descs[i].tiles; t++) items.push_back(BatchItem{i, t});
  // One-time set-up of the direct squares' kernels now: not allowed while
  // capturing.
  for (const zoda_square& q : direct)
    ZODA_TRY(launch_encode(st, q.in, q.k, q.cols, q.n, q.out, ZODA_LAYOUT_ROW, nullptr, true));

  zoda_batch_plan* p = new zoda_batch_plan();
  p->device = st->device;
  p->nitems = (uint32_t)items.size();
  p->direct = std::move(direct);
  auto bail = [&](int rc) {
    free_plan(p);
    return rc;
  };
  int made = 0;
  hip_note(made, hipEventCreateWithFlags(&p->fork, hipEventDisableTiming), "batch plan event creation");
  for (int b = 0; b < BATCH_BRANCHES && !made; b++) {
    hip_note(made, hipStreamCreateWithFlags(&p->side[b], hipStreamNonBlocking), "batch plan stream creation");
    if (!made) hip_note(made, hipEventCreateWithFlags(&p->join[b], hipEventDisableTiming), "batch plan event creation");
  }
  if (made) return bail(made);
  // Zero-sized arrays still get an allocation so the kernel arguments are valid.
  {
    hipError_t ea = hipMalloc((void**)&p->d_squares, std::max<size_t>(1, descs.size()) * sizeof(SquareDesc));
    if (ea == hipSuccess) ea = hipMalloc((void**)&p->d_items, std::max<size_t>(1, items.size()) * sizeof(BatchItem));
    if (ea == hipSuccess) ea = hipMalloc((void**)&p->d_tabs, tabs.size() * sizeof(EncodeTabs));
    if (ea == hipSuccess) ea = hipMalloc((void**)&p->d_counter, sizeof(uint32_t));
    if (ea != hipSuccess) return bail(alloc_fail(ea, "batch plan allocation"));
  }
  hipError_t e = hipSuccess;
  if (!descs.empty())
    e = hipMemcpy(p->d_squares, descs.data(), descs.size() * sizeof(SquareDesc), hipMemcpyHostToDevice);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(p->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_tabs, tabs.data(), tabs.size() * sizeof(EncodeTabs), hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(fail((int)e, "batch plan upload failed: %s", hipGetErrorString(e)));

  for (const SquareDesc& d : descs) {
    p->lds_bytes = std::max(p->lds_bytes, batch_lds_bytes(d.L, d.variant));
    p->max_log = std::max(p->max_log, d.L);
  }
  int rc = batch_grid(st, p->lds_bytes, p->max_log, &p->grid);
  if (rc) return bail(rc);
  p->grid = std::max<uint32_t>(1, std::min<uint32_t>(p->grid, p->nitems));

  if (graph) {
    hipStream_t cap;
    e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(fail((int)e, "stream creation failed: %s", hipGetErrorString(e)));
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    rc = e == hipSuccess ? launch_batch_direct(p, cap) : fail((int)e, "begin capture failed");
    hipError_t e2 = hipStreamEndCapture(cap, &p->graph);
    (void)hipStreamDestroy(cap);
    if (rc) return bail(rc);
    if (e2 != hipSuccess) return bail(fail((int)e2, "end capture failed: %s", hipGetErrorString(e2)));
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(fail((int)e, "graph instantiation failed: %s", hipGetErrorString(e)));
  }
  *out = p;
  return 0;
}

extern "C" int zoda_hip_batch_plan_create(const zoda_square* squares, uint32_t count,
                                          zoda_batch_plan** plan) {
  return build_plan(squares, count, true, true, plan);
}

extern "C" int zoda_hip_batch_plan_launch(zoda_batch_plan* plan, void* stream) {
  clear_error();
  if (!plan || !plan->exec) return fail(ZODA_EINVAL, "invalid batch plan");
  ZODA_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream));
  return 0;
}

extern "C" int zoda_hip_batch_plan_destroy(zoda_batch_plan* plan) {
  clear_error();
  if (!plan) return 0;
  ZODA_HIP(hipDeviceSynchronize());  // a replay may still be reading the plan
  free_plan(plan);
  return 0;
}

extern "C" int zoda_hip_encode_batch(const zoda_square* squares, uint32_t count, void* stream) {
  zoda_batch_plan* p = nullptr;
  ZODA_TRY(build_plan(squares, count, false, true, &p));
  int rc = launch_batch_direct(p, (hipStream_t)stream);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  free_plan(p);
  if (rc) return rc;
  if (e != hipSuccess) return fail((int)e, "batch encode failed: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int zoda_hip_get_variant(uint32_t log_side, int* variant) {
  clear_error();
  if (!variant || (int)log_side > BATCH_MAX_LOG) return fail(ZODA_EINVAL, "bad arguments");
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  *variant = st->batch_direct[log_side] ? BATCH_DIRECT : st->batch_variant[log_side];
  return 0;
}

extern "C" int zoda_hip_set_variant(uint32_t log_side, int variant) {
  clear_error();
  if ((int)log_side > BATCH_MAX_LOG || variant < 0 || variant > BATCH_DIRECT)
    return fail(ZODA_EINVAL, "bad arguments");
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  st->batch_direct[log_side] = variant == BATCH_DIRECT;
  if (variant < BATCH_VARIANTS) st->batch_variant[log_side] = variant;
  return 0;
}

extern "C" int zoda_hip_batch_variant_info(uint32_t log_side, int variant, int* info) {
  clear_error();
  if (!info || (int)log_side > BATCH_MAX_LOG || variant < 0 || variant > BATCH_DIRECT)
    return fail(ZODA_EINVAL, "bad arguments");
  auto put = [&](auto plan) -> int {
    using PL = decltype(plan);
    info[0] = PL::W;
    info[1] = PL::NT;
    info[2] = PL::R1;
    info[3] = PL::R2;
    info[4] = PL::RD;
    info[5] = PL::TWL ? 1 : 0;
    return 0;
  };
  if (variant == BATCH_DIRECT) {
    if (with_encode_plan((int)log_side, put) != 0) return fail(ZODA_EINVAL, "no plan for 2^%u", log_side);
    if (log_side == 11) {  // the slice-pipelined kernel (32-column tiles, 512 threads, LDS twiddles)
      info[0] = 32;
      info[1] = 512;
      info[2] = 4;
      info[3] = 3;
      info[4] = 4;
    }
  } else {
    with_batch_plan((int)log_side, variant, put);
  }
  return 0;
}

// For each listed side, times every candidate on one synthetic square of that
// side (side rows x 128*side columns, device-resident, the best of `iters`
// launches after one warm-up): the BATCH_VARIANTS work-queue plans, as a
// one-square batch, and the side's own fused launch.  The fastest queue plan
// becomes the side's variant, and the side runs direct when its own launch
// beat it.  times_us (optional) receives count * ZODA_TUNE_CANDIDATES values.
extern "C" int zoda_hip_autotune(const uint32_t* sides, uint32_t count, int iters, float* times_us) {
  clear_error();
  if (!sides || count == 0 || iters < 1) return fail(ZODA_EINVAL, "bad arguments");
  for (uint32_t i = 0; i < count; i++)
    if (!is_pow2(sides[i]) || (int)log2_ceil(sides[i]) > BATCH_MAX_LOG)
      return fail(ZODA_EINVAL, "side %u must be a power of two <= 2^%d", sides[i], BATCH_MAX_LOG);
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  hipStream_t s;
  ZODA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  hip_note(rc, hipEventCreate(&e0), "autotune event");
  if (!rc) hip_note(rc, hipEventCreate(&e1), "autotune event");
  // best of `iters` timed runs of fn after one warm-up, in microseconds
  auto time_it = [&](auto&& fn, float* best) {
    *best = 1e30f;
    for (int it = 0; it <= iters && !rc; it++) {
      hip_note(rc, hipEventRecord(e0, s), "autotune event");
      if (!rc) rc = fn();
      if (!rc) hip_note(rc, hipEventRecord(e1, s), "autotune event");
      if (!rc) hip_note(rc, hipEventSynchronize(e1), "autotune launch");
      float ms = 0;
      if (!rc) hip_note(rc, hipEventElapsedTime(&ms, e0, e1), "autotune timing");
      if (!rc && it > 0) *best = std::min(*best, ms * 1000.0f);
    }
  };
  for (uint32_t i = 0; i < count && !rc; i++) {
    const uint32_t side = sides[i];
    const uint32_t L = log2_ceil(side);
    const uint64_t cols = 128ull * side;
    uint32_t *din = nullptr, *dout = nullptr;
    hipError_t ea = hipMalloc((void**)&din, side * cols * 4);
    if (ea == hipSuccess) ea = hipMalloc((void**)&dout, 2 * side * cols * 4);
    if (ea != hipSuccess) {
      (void)hipFree(din);
      rc = alloc_fail(ea, "autotune buffers");
      break;
    }
    rc = launch_fill_uniform(din, side * cols, side, s);
    const int keep_variant = st->batch_variant[L];
    float best = 1e30f;
    int best_v = 0;
    for (int v = 0; v < BATCH_VARIANTS && !rc; v++) {
      st->batch_variant[L] = v;
      zoda_square sq{din, dout, side, (uint32_t)cols, side, 0};
      zoda_batch_plan* p = nullptr;
      rc = build_plan(&sq, 1, false, false, &p);
      if (rc) break;
      float t = 0;
      time_it([&] { return launch_batch_direct(p, s); }, &t);
      free_plan(p);
      if (times_us) times_us[i * ZODA_TUNE_CANDIDATES + v] = t;
      if (t < best) {
        best = t;
        best_v = v;
      }
    }
    float t_direct = 1e30f;
    if (!rc) {
      time_it([&] { return launch_encode(st, din, side, (uint32_t)cols, side, dout, ZODA_LAYOUT_ROW, s); },
              &t_direct);
      if (times_us) times_us[i * ZODA_TUNE_CANDIDATES + BATCH_DIRECT] = t_direct;
    }
    {
      std::lock_guard<std::recursive_mutex> lock(st->mu);
      st->batch_variant[L] = rc ? keep_variant : best_v;
      if (!rc) st->batch_direct[L] = t_direct < best;
    }
    (void)hipFree(din);
    (void)hipFree(dout);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return rc;
}

// Tunes the device's table for one batch as it will run: for each side the
// batch holds (largest first), every candidate of that side is timed as the
// whole batch (best of `iters` launches after a warm-up), with the other
// sides at their current entries, and the fastest is kept (one sweep of
// coordinate descent from the current table, so the result is never slower
// than the entries it started from, as measured).  The squares' outputs are
// written as by zoda_hip_encode_batch.  times_us (optional, 13 *
// ZODA_TUNE_CANDIDATES floats) receives [log side][candidate] timings, 0 for
// sides the batch lacks.
extern "C" int zoda_hip_autotune_batch(const zoda_square* squares, uint32_t count, int iters, float* times_us) {
  clear_error();
  if (!squares || count == 0 || iters < 1) return fail(ZODA_EINVAL, "bad arguments");
  for (uint32_t i = 0; i < count; i++) ZODA_TRY(check_batch_square(squares[i], i));
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  if (times_us)
    for (int i = 0; i < (BATCH_MAX_LOG + 1) * ZODA_TUNE_CANDIDATES; i++) times_us[i] = 0.0f;
  bool present[BATCH_MAX_LOG + 1] = {};
  for (uint32_t i = 0; i < count; i++) present[log2_ceil(squares[i].k)] = true;
  hipStream_t s;
  ZODA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  hip_note(rc, hipEventCreate(&e0), "autotune event");
  if (!rc) hip_note(rc, hipEventCreate(&e1), "autotune event");
  auto set_choice = [&](int L, int v) {
    std::lock_guard<std::recursive_mutex> lock(st->mu);
    st->batch_direct[L] = v == BATCH_DIRECT;
    if (v < BATCH_VARIANTS) st->batch_variant[L] = v;
  };
  for (int L = BATCH_MAX_LOG; L >= BATCH_MIN_LOG && !rc; L--) {
    if (!present[L]) continue;
    const int start_v = st->batch_direct[L] ? BATCH_DIRECT : st->batch_variant[L];
    const int start_q = st->batch_variant[L];
    float best = 1e30f;
    int best_v = start_v;
    for (int v = 0; v < ZODA_TUNE_CANDIDATES && !rc; v++) {
      set_choice(L, v);
      if (v == BATCH_DIRECT) {  // the queue plan the side keeps while it runs direct
        std::lock_guard<std::recursive_mutex> lock(st->mu);
        st->batch_variant[L] = start_q;
      }
      zoda_batch_plan* p = nullptr;
      rc = build_plan(squares, count, false, true, &p);
      if (rc) break;
      float t = 1e30f;
      for (int it = 0; it <= iters && !rc; it++) {
        hip_note(rc, hipEventRecord(e0, s), "autotune event");
        if (!rc) rc = launch_batch_direct(p, s);
        if (!rc) hip_note(rc, hipEventRecord(e1, s), "autotune event");
        if (!rc) hip_note(rc, hipEventSynchronize(e1), "autotune launch");
        float ms = 0;
        if (!rc) hip_note(rc, hipEventElapsedTime(&ms, e0, e1), "autotune timing");
        if (!rc && it > 0) t = std::min(t, ms * 1000.0f);
      }
      free_plan(p);
      if (times_us) times_us[L * ZODA_TUNE_CANDIDATES + v] = t;
      if (t < best) {
        best = t;
        best_v = v;
      }
    }
    // the queue plan stays the best queue candidate's (when the side runs direct too)
    int best_q = start_q;
    if (times_us) {
      float bq = 1e30f;
      for (int v = 0; v < BATCH_VARIANTS; v++)
        if (times_us[L * ZODA_TUNE_CANDIDATES + v] < bq) {
          bq = times_us[L * ZODA_TUNE_CANDIDATES + v];
          best_q = v;
        }
    }
    set_choice(L, rc ? start_v : best_v);
    if (!rc && best_v == BATCH_DIRECT && times_us) {
      std::lock_guard<std::recursive_mutex> lock(st->mu);
      st->batch_variant[L] = best_q;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return rc;
}

// The device's name and CU count, the key of a tuning table on disk.
static int device_key(DeviceState* st, std::string* key) {
  hipDeviceProp_t prop;
  ZODA_HIP(hipGetDeviceProperties(&prop, st->device));
  char buf[300];
  snprintf(buf, sizeof buf, "%s/%dcu", prop.gcnArchName, prop.multiProcessorCount);
  for (char* c = buf; *c; c++)
    if (*c == ' ') *c = '_';
  *key = buf;
  return 0;
}

// Tuning table persistence: one line per log side, "<device> <L> <variant>"
// (variant ZODA_BATCH_DIRECT for the side's own fused launch).  Saving
// replaces the file; loading takes the lines of this device's name and CU
// count and skips the others.
extern "C" int zoda_hip_tune_save(const char* path) {
  clear_error();
  if (!path) return fail(ZODA_EINVAL, "bad arguments");
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::string key;
  ZODA_TRY(device_key(st, &key));
  FILE* f = fopen(path, "w");
  if (!f) return fail(ZODA_EINVAL, "cannot write %s", path);
  {
    std::lock_guard<std::recursive_mutex> lock(st->mu);
    for (int L = BATCH_MIN_LOG; L <= BATCH_MAX_LOG; L++)
      fprintf(f, "%s %d %d\n", key.c_str(), L, st->batch_direct[L] ? BATCH_DIRECT : st->batch_variant[L]);
  }
  if (fclose(f) != 0) return fail(ZODA_EINVAL, "writing %s failed", path);
  return 0;
}

extern "C" int zoda_hip_tune_load(const char* path, uint32_t* entries) {
  clear_error();
  if (!path) return fail(ZODA_EINVAL, "bad arguments");
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  std::string key;
  ZODA_TRY(device_key(st, &key));
  FILE* f = fopen(path, "r");
  if (!f) return fail(ZODA_EINVAL, "cannot read %s", path);
  char dev[300];
  int L = 0, v = 0;
  uint32_t n = 0;
  std::lock_guard<std::recursive_mutex> lock(st->mu);
  while (fscanf(f, "%299s %d %d", dev, &L, &v) == 3) {
    if (key != dev || L < BATCH_MIN_LOG || L > BATCH_MAX_LOG || v < 0 || v > BATCH_DIRECT) continue;
    st->batch_direct[L] = v == BATCH_DIRECT;
    if (v < BATCH_VARIANTS) st->batch_variant[L] = v;
    n++;
  }
  fclose(f);
  if (entries) *entries = n;
  return 0;
}
