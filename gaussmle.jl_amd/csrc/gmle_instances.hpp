// gmle_instances.hpp -- the (box size, lanes-per-ROI) grid of kernel
// specialisations built for one (model, camera, precision) triple, and the
// lookup that maps run-time arguments onto them.
//
// Every box size 3..16 is built.  Lanes per ROI: 1, 2 and 4 up to 12 px;
// 2 and 4 above, where one lane cannot keep a whole ROI resident in VGPRs.
#pragma once

#include "gmle_fit_kernel.hpp"

namespace gmle {

constexpr bool lanes_built(int box, int lanes) {
  return (box >= kMinBox && box <= kMaxBox) &&
         ((lanes == 1 && box <= 12) || lanes == 2 || lanes == 4);
}

template <int MODEL, int CAM, int PREC, int N>
LaunchFn pick_lanes(int lanes) {
  if constexpr (lanes_built(N, 1)) {
    if (lanes == 1) return &launch_fit<MODEL, CAM, N, 1, PREC>;
  }
  if (lanes == 2) return &launch_fit<MODEL, CAM, N, 2, PREC>;
  if (lanes == 4) return &launch_fit<MODEL, CAM, N, 4, PREC>;
  return nullptr;
}

template <int MODEL, int CAM, int PREC>
LaunchFn pick(int box, int lanes) {
  switch (box) {
    case 3: return pick_lanes<MODEL, CAM, PREC, 3>(lanes);
    case 4: return pick_lanes<MODEL, CAM, PREC, 4>(lanes);
    case 5: return pick_lanes<MODEL, CAM, PREC, 5>(lanes);
    case 6: return pick_lanes<MODEL, CAM, PREC, 6>(lanes);
    case 7: return pick_lanes<MODEL, CAM, PREC, 7>(lanes);
    case 8: return pick_lanes<MODEL, CAM, PREC, 8>(lanes);
    case 9: return pick_lanes<MODEL, CAM, PREC, 9>(lanes);
    case 10: return pick_lanes<MODEL, CAM, PREC, 10>(lanes);
    case 11: return pick_lanes<MODEL, CAM, PREC, 11>(lanes);
    case 12: return pick_lanes<MODEL, CAM, PREC, 12>(lanes);
    case 13: return pick_lanes<MODEL, CAM, PREC, 13>(lanes);
    case 14: return pick_lanes<MODEL, CAM, PREC, 14>(lanes);
    case 15: return pick_lanes<MODEL, CAM, PREC, 15>(lanes);
    case 16: return pick_lanes<MODEL, CAM, PREC, 16>(lanes);
  }
  return nullptr;
}

}  // namespace gmle

// One translation unit per (model, camera, precision) defines
//   LaunchFn find_launcher_<MODEL>_<CAM>_<PREC>(int box, int lanes)
#define GMLE_DEFINE_FINDER(MODEL, CAM, PREC)                                     \
  namespace gmle {                                                               \
  LaunchFn find_launcher_##MODEL##_##CAM##_##PREC(int box, int lanes) {          \
    return pick<MODEL, CAM, PREC>(box, lanes);                                   \
  }                                                                              \
  }
