// Kernel specialisations built with the max-ILP machine scheduler
// (-amdgpu-sched-strategy=max-ilp, set for this file in build.py).
//
// Measured on gfx950, Float32, 1,000,000 ROIs, against the default scheduler:
//   XYNBS 21 px, ideal camera, 8 lanes:  8.94 -> 8.73 ms per launch
// Under the same scheduler the one-lane XYNB 11 px kernel did not change
// (1.402 against 1.405-1.409 ms), so it stays where it is.  The finder
// declares these specialisations extern.
#include "gmle_instances.hpp"
GMLE_INSTANTIATE_FIT(XYNBS, IDEAL, PREC_FP32, 21, 8)
