// gmle_instances.hpp -- the (box size, lanes-per-ROI) grid of kernel
// specialisations built for one (model, camera, precision) triple, and the
// lookup that maps run-time arguments onto them.
//
// Every box size 3..32 is built (lanes_built).  Lanes per ROI: 1, 2, 4, 8
// and 16 up to 12 px; 2 to 16 up to 16 px, where one lane cannot keep a whole
// ROI resident in VGPRs; 8 and 16 up to 24 px and 16 above (large boxes, rows
// in the outer loop).  8 and 16 lanes also serve small batches: shorter
// waves, more of them.
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
  if constexpr (lanes_built(N, 2)) {
    if (lanes == 2) return entry<MODEL, CAM, PREC, N, 2, Describe>();
  }
  if constexpr (lanes_built(N, 4)) {
    if (lanes == 4) return entry<MODEL, CAM, PREC, N, 4, Describe>();
  }
  if constexpr (lanes_built(N, 8)) {
    if (lanes == 8) return entry<MODEL, CAM, PREC, N, 8, Describe>();
  }
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
    case 17: return pick_lanes<MODEL, CAM, PREC, 17, Describe, Fn>(lanes);
    case 18: return pick_lanes<MODEL, CAM, PREC, 18, Describe, Fn>(lanes);
    case 19: return pick_lanes<MODEL, CAM, PREC, 19, Describe, Fn>(lanes);
    case 20: return pick_lanes<MODEL, CAM, PREC, 20, Describe, Fn>(lanes);
    case 21: return pick_lanes<MODEL, CAM, PREC, 21, Describe, Fn>(lanes);
    case 22: return pick_lanes<MODEL, CAM, PREC, 22, Describe, Fn>(lanes);
    case 23: return pick_lanes<MODEL, CAM, PREC, 23, Describe, Fn>(lanes);
    case 24: return pick_lanes<MODEL, CAM, PREC, 24, Describe, Fn>(lanes);
    case 25: return pick_lanes<MODEL, CAM, PREC, 25, Describe, Fn>(lanes);
    case 26: return pick_lanes<MODEL, CAM, PREC, 26, Describe, Fn>(lanes);
    case 27: return pick_lanes<MODEL, CAM, PREC, 27, Describe, Fn>(lanes);
    case 28: return pick_lanes<MODEL, CAM, PREC, 28, Describe, Fn>(lanes);
    case 29: return pick_lanes<MODEL, CAM, PREC, 29, Describe, Fn>(lanes);
    case 30: return pick_lanes<MODEL, CAM, PREC, 30, Describe, Fn>(lanes);
    case 31: return pick_lanes<MODEL, CAM, PREC, 31, Describe, Fn>(lanes);
    case 32: return pick_lanes<MODEL, CAM, PREC, 32, Describe, Fn>(lanes);
  }
  return nullptr;
}

}  // namespace gmle

// Specialisations compiled in a translation unit of their own, under another
// machine scheduler: that unit instantiates the launcher and the description
// (GMLE_INSTANTIATE_FIT), and the finder's unit declares them extern
// (GMLE_EXTERN_FIT, before GMLE_DEFINE_FINDER) so that it builds neither them
// nor their kernel.  Float32 only: a mixed-precision launcher also launches
// the Float64 final pass, which its finder's unit builds as well.
#define GMLE_EXTERN_FIT(MODEL, CAM, PREC, N, L)                                                 \
  namespace gmle {                                                                             \
  static_assert(PREC == PREC_FP32, "only Float32 launchers are built apart");                  \
  extern template hipError_t launch_fit<MODEL, CAM, N, L, PREC>(const FitArgs&, hipStream_t);   \
  extern template bool describe_fit<MODEL, CAM, N, L, PREC>(KernelInfo*);                       \
  }
#define GMLE_INSTANTIATE_FIT(MODEL, CAM, PREC, N, L)                                            \
  namespace gmle {                                                                             \
  static_assert(PREC == PREC_FP32, "only Float32 launchers are built apart");                  \
  template hipError_t launch_fit<MODEL, CAM, N, L, PREC>(const FitArgs&, hipStream_t);          \
  template bool describe_fit<MODEL, CAM, N, L, PREC>(KernelInfo*);                              \
  }

// The same for the iteration kernel alone, for mixed precision: the launcher
// and the Float64 final pass stay in the finder's unit.
#define GMLE_EXTERN_KERNEL(MODEL, CAM, PREC, N, L)                                              \
  namespace gmle {                                                                             \
  extern template __global__ void fit_kernel<MODEL, CAM, N, L, PREC>(FitArgs);                  \
  }
#define GMLE_INSTANTIATE_KERNEL(MODEL, CAM, PREC, N, L)                                         \
  namespace gmle {                                                                             \
  template __global__ void fit_kernel<MODEL, CAM, N, L, PREC>(FitArgs);                         \
  }

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
