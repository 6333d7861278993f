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
// side, at most 80 KiB of LDS so two work-groups share a CU, and four
// variants per side for the autotuner to choose from (besides running the
// side's squares as their own fused launches): tile width W, digit split
// (radix and stages per register pass) and twiddle tables in LDS behind the
// tile (loaded when a work-group changes plan) or read from global memory.
// Variant 0 is the default before tuning.  The two largest sides have one
// plan each (their tiles leave no room for variants under the LDS limit).
constexpr int BATCH_NT = 256;
constexpr int BATCH_MIN_LOG = 0;
constexpr int BATCH_MAX_LOG = 12;
constexpr int BATCH_VARIANTS = 4;
constexpr int BATCH_DIRECT = BATCH_VARIANTS;  // tuning choice: the side's own fused launch
// 80 KiB less 64 bytes for the kernel's static LDS (its work-item slot).
constexpr int BATCH_LDS_WORDS = 20480 - 16;

//                          L   W    NT        R1 R2 RD  LDS twiddles
using B0a = EncodePlan<0, 256, BATCH_NT, 0, 0, 0, true>;
using B1a = EncodePlan<1, 256, BATCH_NT, 0, 0, 1, true>;
using B2a = EncodePlan<2, 256, BATCH_NT, 0, 0, 2, true>;
using B3a = EncodePlan<3, 256, BATCH_NT, 0, 0, 3, true>;
using B4a = EncodePlan<4, 256, BATCH_NT, 0, 0, 4, true>;
using B4b = EncodePlan<4, 64, BATCH_NT, 2, 0, 2, true>;
using B4c = EncodePlan<4, 128, BATCH_NT, 2, 0, 2, true>;
using B4d = EncodePlan<4, 64, BATCH_NT, 2, 0, 2, false>;
using B5a = EncodePlan<5, 64, BATCH_NT, 3, 0, 2, true>;
using B5b = EncodePlan<5, 128, BATCH_NT, 3, 0, 2, true>;
using B5c = EncodePlan<5, 64, BATCH_NT, 2, 0, 3, true>;
using B5d = EncodePlan<5, 64, BATCH_NT, 3, 0, 2, false>;
using B6a = EncodePlan<6, 64, BATCH_NT, 3, 0, 3, true>;
using B6b = EncodePlan<6, 32, BATCH_NT, 3, 0, 3, true>;
using B6c = EncodePlan<6, 128, BATCH_NT, 3, 0, 3, true>;
using B6d = EncodePlan<6, 64, BATCH_NT, 3, 0, 3, false>;
using B7a = EncodePlan<7, 64, BATCH_NT, 4, 0, 3, true>;
using B7b = EncodePlan<7, 32, BATCH_NT, 4, 0, 3, true>;
using B7c = EncodePlan<7, 64, BATCH_NT, 3, 0, 4, true>;
using B7d = EncodePlan<7, 64, BATCH_NT, 4, 0, 3, false>;
using B8a = EncodePlan<8, 64, BATCH_NT, 4, 0, 4, true>;
using B8b = EncodePlan<8, 32, BATCH_NT, 3, 2, 3, true>;
using B8c = EncodePlan<8, 32, BATCH_NT, 4, 0, 4, true>;
using B8d = EncodePlan<8, 64, BATCH_NT, 4, 0, 4, false>;
using B9a = EncodePlan<9, 32, BATCH_NT, 3, 3, 3, true>;
using B9b = EncodePlan<9, 32, BATCH_NT, 5, 0, 4, true>;
using B9c = EncodePlan<9, 16, BATCH_NT, 3, 3, 3, true>;
using B9d = EncodePlan<9, 32, BATCH_NT, 3, 3, 3, false>;
using B10a = EncodePlan<10, 16, BATCH_NT, 4, 3, 3, true>;
using B10b = EncodePlan<10, 16, BATCH_NT, 5, 0, 5, true>;
using B10c = EncodePlan<10, 8, BATCH_NT, 4, 3, 3, true>;
using B10d = EncodePlan<10, 16, BATCH_NT, 4, 3, 3, false>;
using B11a = EncodePlan<11, 8, BATCH_NT, 4, 4, 3>;
using B12a = EncodePlan<12, 4, BATCH_NT, 4, 4, 4>;

// Calls f(Plan{}) for batched plan (L, variant); sides with fewer plans map
// the other variants onto theirs.  Returns -1 for L out of range.
template <class PL>
constexpr bool fits_batch() {
  return PL::NT == BATCH_NT && PL::LDS_WORDS <= BATCH_LDS_WORDS;
}

template <class F>
ZHD int with_batch_plan(int L, int variant, F&& f) {
  const int v = variant & (BATCH_VARIANTS - 1);
  switch (L) {
    case 0: return f(B0a{});
    case 1: return f(B1a{});
    case 2: return f(B2a{});
    case 3: return f(B3a{});
    case 4: return v == 0 ? f(B4a{}) : v == 1 ? f(B4b{}) : v == 2 ? f(B4c{}) : f(B4d{});
    case 5: return v == 0 ? f(B5a{}) : v == 1 ? f(B5b{}) : v == 2 ? f(B5c{}) : f(B5d{});
    case 6: return v == 0 ? f(B6a{}) : v == 1 ? f(B6b{}) : v == 2 ? f(B6c{}) : f(B6d{});
    case 7: return v == 0 ? f(B7a{}) : v == 1 ? f(B7b{}) : v == 2 ? f(B7c{}) : f(B7d{});
    case 8: return v == 0 ? f(B8a{}) : v == 1 ? f(B8b{}) : v == 2 ? f(B8c{}) : f(B8d{});
    case 9: return v == 0 ? f(B9a{}) : v == 1 ? f(B9b{}) : v == 2 ? f(B9c{}) : f(B9d{});
    case 10: return v == 0 ? f(B10a{}) : v == 1 ? f(B10b{}) : v == 2 ? f(B10c{}) : f(B10d{});
    case 11: return f(B11a{});
    case 12: return f(B12a{});
    default: return -1;
  }
}

template <class... P>
constexpr bool all_fit_batch() {
  return (fits_batch<P>() && ...);
}
static_assert(all_fit_batch<B0a, B1a, B2a, B3a, B4a, B4b, B4c, B4d, B5a, B5b, B5c, B5d, B6a, B6b, B6c, B6d, B7a,
                            B7b, B7c, B7d, B8a, B8b, B8c, B8d, B9a, B9b, B9c, B9d, B10a, B10b, B10c, B10d, B11a,
                            B12a>(),
              "batched plans must fit the batched kernel's LDS");

}  // namespace zoda
