// Kernel specialisations built with the max-ILP machine scheduler
// (-amdgpu-sched-strategy=max-ilp, set for this file in build.py).
//
// Measured on gfx950, Float32, 1,000,000 ROIs, against the default scheduler:
//   XYNBS 21 px, ideal camera, 8 lanes:  8.94 -> 8.73 ms per launch
// and at 10,000 ROIs per launch, where each SIMD holds one wave:
//   XYNB 11 px, ideal camera, 4 lanes:    2.02-2.03e8 -> 2.07e8 fits/s
// Under the same scheduler the one-lane XYNB 11 px kernel did not change
// (1.402 against 1.405-1.409 ms), and the Float64 final passes of XYNBS 13 px
// sCMOS and the astigmatic 15 px model did not gain, so they stay where they
// are.  The finders declare these specialisations extern.
#include "gmle_instances.hpp"
GMLE_INSTANTIATE_FIT(XYNBS, IDEAL, PREC_FP32, 21, 8)
GMLE_INSTANTIATE_FIT(XYNB, IDEAL, PREC_FP32, 11, 4)
