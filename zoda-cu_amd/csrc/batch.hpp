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

int check_batch_square(const zoda_square& s, uint32_t i);
// Fills *d for square s with the device's autotuned variant; builds the
// (side, variant) table pointers into tabs_by_slot[L * 2 + variant] once.
int describe_square(DeviceState* st, const zoda_square& s, SquareDesc* d, EncodeTabs* tabs_by_slot,
                    bool* have_slot);
// Persistent grid that fills every CU.
int batch_grid(DeviceState* st, uint32_t* grid);
// [counter reset, persistent kernel] on `stream`.
int launch_batch_items(const SquareDesc* d_squares, const BatchItem* d_items, uint32_t nitems,
                       uint32_t* d_counter, const EncodeTabs* d_tabs, uint32_t grid,
                       hipStream_t stream);

}  // namespace zoda
