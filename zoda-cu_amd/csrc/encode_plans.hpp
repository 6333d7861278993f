// Tile width, work-group size and digit split of the fused encode for each
// side s = 2^L.  Tiles are LDS-resident as int32 (W * s * 4 bytes <= 128 KiB);
// sides up to 32 keep a whole column in one thread's registers (no LDS).
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
using Plan6 = EncodePlan<6, 64, 256, 3, 0, 3>;
using Plan7 = EncodePlan<7, 64, 256, 4, 0, 3>;
using Plan8 = EncodePlan<8, 64, 512, 4, 0, 4>;
using Plan9 = EncodePlan<9, 32, 512, 3, 3, 3>;
using Plan10 = EncodePlan<10, 32, 1024, 4, 3, 3>;
using Plan11 = EncodePlan<11, 16, 1024, 4, 4, 3>;
using Plan12 = EncodePlan<12, 8, 1024, 4, 4, 4>;

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
