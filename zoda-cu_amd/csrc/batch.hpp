// Work-queue types and launchers shared by the batched encode (batch.hip) and
// the serving loop (serve.hip).
#pragma once

#include "common.hpp"
#include "encode_core.hpp"

namespace zoda {

// One square as the persistent kernel sees it.
struct SquareDesc {
  const uint32_t* in;
  uint32_t* out;
  uint32_t k, cols, L, variant, full, tiles;
};

// One work item: column tile `tile` of square `square`.
struct BatchItem {
  uint32_t square, tile;
};

// A batch whose descriptors travel in the kernel arguments (serving loop):
// square q owns work items [first[q], first[q + 1]).  The work counter is
// never reset: a launch takes items counter - base, and the host advances
// `base` by items + work-groups after each launch on the same counter, which
// is exactly how far that launch moves it (every work-group ends on one take
// past the last item).
constexpr uint32_t SERVE_MAX_BATCH = 48;
struct ServeBatch {
  SquareDesc sq[SERVE_MAX_BATCH];
  uint32_t first[SERVE_MAX_BATCH + 1];
  uint32_t n, items, base;
  uint32_t* counter;
  const EncodeTabs* tabs;
};

int check_batch_square(const zoda_square& s, uint32_t i);
// Fills *d for square s with the device's autotuned variant; builds the
// (side, variant) table pointers into tabs_by_slot[L * 2 + variant] once.
int describe_square(DeviceState* st, const zoda_square& s, SquareDesc* d, EncodeTabs* tabs_by_slot,
                    bool* have_slot);
// Dynamic LDS bytes of the batched plan of side 2^L, variant.  A launch takes
// the largest over its squares: small sides need far less than the largest
// plan, and more of their work-groups then fit a CU.
uint32_t batch_lds_bytes(uint32_t L, uint32_t variant);
// Persistent grid that fills every CU with work-groups of `lds_bytes`, for
// the kernel of squares of side up to 2^max_log.
int batch_grid(DeviceState* st, uint32_t lds_bytes, uint32_t max_log, uint32_t* grid);
// [counter reset, persistent kernel] on `stream`.
int launch_batch_items(const SquareDesc* d_squares, const BatchItem* d_items, uint32_t nitems,
                       uint32_t* d_counter, const EncodeTabs* d_tabs, uint32_t grid, uint32_t lds_bytes,
                       uint32_t max_log, hipStream_t stream);

// Persistent kernel over a ServeBatch on `stream`, on min(grid, items)
// work-groups; *advance receives how far it moves the counter.
int launch_serve_batch(const ServeBatch& b, uint32_t grid, uint32_t lds_bytes, uint32_t max_log,
                       hipStream_t stream, uint32_t* advance);

}  // namespace zoda
