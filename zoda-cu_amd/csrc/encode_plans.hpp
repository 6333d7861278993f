// Tile width, work-group size and digit split of the fused encode for each
// side s = 2^L.  Tiles are LDS-resident as int32 (W * s * 4 bytes <= 128 KiB);
// sides up to 32 keep a whole column in one thread's registers (no LDS).  The
// flags keep the twiddle tables in LDS as u32 next to the tile, and run whole
// tiles in a persistent kernel whose first pass reads registers loaded while
// the previous tile was transformed.
#pragma once

#include "encode_core.hpp"

namespace zoda {

//                     L   W     NT   R1 R2 RD
using Plan0 = EncodePlan<0, 256, 256, 0, 0, 0>;
using Plan1 = EncodePlan<1, 256, 256, 0, 0, 1>;
using Plan2 = EncodePlan<2, 256, 256, 0, 0, 2>;
using Plan3 = EncodePlan<3, 256, 256, 0, 0, 3>;
using Plan4 = EncodePlan<4, 256, 256, 0, 0, 4>;
using Plan5 = EncodePlan<5, 256, 256, 0, 0, 5>;
using Plan6 = EncodePlan<6, 64, 256, 3, 0, 3, true>;
using Plan7 = EncodePlan<7, 64, 256, 4, 0, 3, true>;
using Plan8 = EncodePlan<8, 64, 512, 4, 0, 4, true, true>;
using Plan9 = EncodePlan<9, 32, 512, 3, 3, 3, true, true>;
using Plan10 = EncodePlan<10, 32, 1024, 4, 3, 3, true, true>;
using Plan11 = EncodePlan<11, 16, 1024, 4, 4, 3, true, true>;
using Plan12 = EncodePlan<12, 8, 1024, 4, 4, 4, false, true>;

constexpr int ENCODE_MAX_LOG = 12;

// Calls f(PlanL{}) for 0 <= L <= ENCODE_MAX_LOG; returns -1 for other L.
template <class F>
int with_encode_plan(int L, F&& f) {
  switch (L) {
    case 0: return f(Plan0{});
    case 1: return f(Plan1{});
    case 2: return f(Plan2{});
    case 3: return f(Plan3{});
    case 4: return f(Plan4{});
    case 5: return f(Plan5{});
    case 6: return f(Plan6{});
    case 7: return f(Plan7{});
    case 8: return f(Plan8{});
    case 9: return f(Plan9{});
    case 10: return f(Plan10{});
    case 11: return f(Plan11{});
    case 12: return f(Plan12{});
    default: return -1;
  }
}

}  // namespace zoda

namespace zoda {

// Plans of the batched persistent kernel: one block size (256) for every
// side, at most 80 KiB of LDS so two work-groups share a CU, and two variants
// per side (tile width W, digit split = radix and stages per register pass)
// for the autotuner to choose from.  Up to side 1024 the twiddle tables sit
// in LDS behind the tile (loaded when a work-group changes plan); the two
// largest sides read them from global memory.
constexpr int BATCH_NT = 256;
constexpr int BATCH_MIN_LOG = 0;
constexpr int BATCH_MAX_LOG = 12;
constexpr int BATCH_VARIANTS = 2;
// 80 KiB less 64 bytes for the kernel's static LDS (its work-item slot).
constexpr int BATCH_LDS_WORDS = 20480 - 16;

//                          L   W    NT        R1 R2 RD  LDS twiddles
using B0a = EncodePlan<0, 256, BATCH_NT, 0, 0, 0, true>;
using B1a = EncodePlan<1, 256, BATCH_NT, 0, 0, 1, true>;
using B2a = EncodePlan<2, 256, BATCH_NT, 0, 0, 2, true>;
using B3a = EncodePlan<3, 256, BATCH_NT, 0, 0, 3, true>;
using B4a = EncodePlan<4, 256, BATCH_NT, 0, 0, 4, true>;
using B4b = EncodePlan<4, 64, BATCH_NT, 2, 0, 2, true>;
using B5b = EncodePlan<5, 64, BATCH_NT, 3, 0, 2, true>;
using B6a = EncodePlan<6, 64, BATCH_NT, 3, 0, 3, true>;
using B6b = EncodePlan<6, 32, BATCH_NT, 3, 0, 3, true>;
using B7a = EncodePlan<7, 64, BATCH_NT, 4, 0, 3, true>;
using B7b = EncodePlan<7, 32, BATCH_NT, 4, 0, 3, true>;
using B8a = EncodePlan<8, 64, BATCH_NT, 4, 0, 4, true>;
using B8b = EncodePlan<8, 32, BATCH_NT, 3, 2, 3, true>;
using B9a = EncodePlan<9, 32, BATCH_NT, 3, 3, 3, true>;
using B9b = EncodePlan<9, 32, BATCH_NT, 5, 0, 4, true>;
using B10a = EncodePlan<10, 16, BATCH_NT, 4, 3, 3, true>;
using B10b = EncodePlan<10, 16, BATCH_NT, 5, 0, 5, true>;
using B11a = EncodePlan<11, 8, BATCH_NT, 4, 4, 3>;
using B12a = EncodePlan<12, 4, BATCH_NT, 4, 4, 4>;

// Calls f(Plan{}) for batched plan (L, variant); variants a side lacks fall
// back to variant 0.  Returns -1 for L out of range.
template <class PL>
constexpr bool fits_batch() {
  return PL::NT == BATCH_NT && PL::LDS_WORDS <= BATCH_LDS_WORDS;
}
static_assert(fits_batch<B8a>() && fits_batch<B9a>() && fits_batch<B9b>() && fits_batch<B10a>() &&
                  fits_batch<B10b>() && fits_batch<B11a>() && fits_batch<B12a>(),
              "batched plans must fit the batched kernel's LDS");

template <class F>
ZHD int with_batch_plan(int L, int variant, F&& f) {
  switch (L * BATCH_VARIANTS + (variant & 1)) {
    case 0: case 1: return f(B0a{});
    case 2: case 3: return f(B1a{});
    case 4: case 5: return f(B2a{});
    case 6: case 7: return f(B3a{});
    case 8: return f(B4a{});
    case 9: return f(B4b{});
    case 10: case 11: return f(B5b{});
    case 12: return f(B6a{});
    case 13: return f(B6b{});
    case 14: return f(B7a{});
    case 15: return f(B7b{});
    case 16: return f(B8a{});
    case 17: return f(B8b{});
    case 18: return f(B9a{});
    case 19: return f(B9b{});
    case 20: return f(B10a{});
    case 21: return f(B10b{});
    case 22: case 23: return f(B11a{});
    case 24: case 25: return f(B12a{});
    default: return -1;
  }
}

}  // namespace zoda
