// gmle_instances.hpp -- the (box size, lanes-per-ROI) grid of kernel
// specialisations built for one (model, camera, precision) triple, and the
// lookup that maps run-time arguments onto them.
//
// Every box size 3..16 is built.  Lanes per ROI: 1, 2, 4, 8 and 16 up to
// 12 px; 2 to 16 above, where one lane cannot keep a whole ROI resident in
// VGPRs.  8 and 16 lanes serve small batches: shorter waves, more of them.
#pragma once

#include "gmle_fit_kernel.hpp"

namespace gmle {

// The launcher (Describe = false) or the kernel description (true) of one
// specialisation, as a function pointer.
template <int MODEL, int CAM, int PREC, int N, int L, bool Describe>
constexpr auto entry() {
  if constexpr (Describe) {
    return &describe_fit<MODEL, CAM, N, L, PREC>;
  } else {
    return &launch_fit<MODEL, CAM, N, L, PREC>;
  }
}

template <int MODEL, int CAM, int PREC, int N, bool Describe, class Fn>
Fn pick_lanes(int lanes) {
  if constexpr (lanes_built(N, 1)) {
    if (lanes == 1) return entry<MODEL, CAM, PREC, N, 1, Describe>();
  }
  if (lanes == 2) return entry<MODEL, CAM, PREC, N, 2, Describe>();
  if (lanes == 4) return entry<MODEL, CAM, PREC, N, 4, Describe>();
  if (lanes == 8) return entry<MODEL, CAM, PREC, N, 8, Describe>();
  if (lanes == 16) return entry<MODEL, CAM, PREC, N, 16, Describe>();
  return nullptr;
}

template <int MODEL, int CAM, int PREC, bool Describe = false, class Fn = LaunchFn>
Fn pick(int box, int lanes) {
  switch (box) {
    case 3: return pick_lanes<MODEL, CAM, PREC, 3, Describe, Fn>(lanes);
    case 4: return pick_lanes<MODEL, CAM, PREC, 4, Describe, Fn>(lanes);
    case 5: return pick_lanes<MODEL, CAM, PREC, 5, Describe, Fn>(lanes);
    case 6: return pick_lanes<MODEL, CAM, PREC, 6, Describe, Fn>(lanes);
    case 7: return pick_lanes<MODEL, CAM, PREC, 7, Describe, Fn>(lanes);
    case 8: return pick_lanes<MODEL, CAM, PREC, 8, Describe, Fn>(lanes);
    case 9: return pick_lanes<MODEL, CAM, PREC, 9, Describe, Fn>(lanes);
    case 10: return pick_lanes<MODEL, CAM, PREC, 10, Describe, Fn>(lanes);
    case 11: return pick_lanes<MODEL, CAM, PREC, 11, Describe, Fn>(lanes);
    case 12: return pick_lanes<MODEL, CAM, PREC, 12, Describe, Fn>(lanes);
    case 13: return pick_lanes<MODEL, CAM, PREC, 13, Describe, Fn>(lanes);
    case 14: return pick_lanes<MODEL, CAM, PREC, 14, Describe, Fn>(lanes);
    case 15: return pick_lanes<MODEL, CAM, PREC, 15, Describe, Fn>(lanes);
    case 16: return pick_lanes<MODEL, CAM, PREC, 16, Describe, Fn>(lanes);
  }
  return nullptr;
}

}  // namespace gmle

// One translation unit per (model, camera, precision) defines
//   LaunchFn   find_launcher_<MODEL>_<CAM>_<PREC>(int box, int lanes)
//   DescribeFn find_describe_<MODEL>_<CAM>_<PREC>(int box, int lanes)
#define GMLE_DEFINE_FINDER(MODEL, CAM, PREC)                                     \
  namespace gmle {                                                               \
  LaunchFn find_launcher_##MODEL##_##CAM##_##PREC(int box, int lanes) {          \
    return pick<MODEL, CAM, PREC>(box, lanes);                                   \
  }                                                                              \
  DescribeFn find_describe_##MODEL##_##CAM##_##PREC(int box, int lanes) {        \
    return pick<MODEL, CAM, PREC, true, DescribeFn>(box, lanes);                 \
  }                                                                              \
  }
