// Kernel specialisations built with the iterative ILP machine scheduler
// (-amdgpu-sched-strategy=iterative-ilp, set for this file in build.py).
//
// Measured on gfx950 at full size, Float32, against the default scheduler:
//   XYNBS 13 px, sCMOS, 4 lanes, 1,000,000 ROIs:  3.313 -> 3.261 ms
//   ASTIG 15 px, ideal, 4 lanes, 4,000,000 ROIs:  24.84 -> 24.04 ms
// and the mixed-precision iteration kernel of the first (its launcher and
// Float64 final pass stay in inst_xynbs_scmos_mixed.hip):
//   fit + final pass, 1,000,000 ROIs:             3.521 -> 3.436 ms
// (The astigmatic mixed-precision kernel ran slower under it: 27.02 -> 27.60.)
// The same scheduler on whole translation units slowed the one-lane XYNB
// 11 px kernel by 1.5 % and made many one- and two-lane kernels spill, so
// only these specialisations use it.  Their finders declare them extern.
#include "gmle_instances.hpp"
GMLE_INSTANTIATE_FIT(XYNBS, SCMOS, PREC_FP32, 13, 4)
GMLE_INSTANTIATE_FIT(ASTIG, IDEAL, PREC_FP32, 15, 4)
GMLE_INSTANTIATE_KERNEL(XYNBS, SCMOS, PREC_MIXED, 13, 4)
