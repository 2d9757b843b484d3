// gmle_host.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmle_args.hpp"

namespace gmle {

// Records the calling thread's error message (printf format) and returns code.
int fail(int code, const char* fmt, ...);
// GMLE_ERR_NOMEM / GMLE_ERR_HIP with the HIP error string.
int hip_fail(hipError_t e, const char* what);
// A kernel with `lanes` lanes per ROI exists for this box size.
bool lanes_ok(int box, int lanes);
// Lanes per ROI chosen when a caller passes 0.
int default_lanes(int model, int camera, int box, int precision, int64_t n_rois);
// FitArgs with the model's psf / constraints and sizes; pointers left null.
void fill_args(FitArgs& a, int model, const float* psf, const float* lower, const float* upper,
               const float* max_step, int64_t n_rois, int ny, int nx, int iterations);

}  // namespace gmle
