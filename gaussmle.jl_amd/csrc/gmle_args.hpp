// gmle_args.hpp -- kernel argument block and launch plumbing shared by the
// kernel translation units and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmle {

enum Model { XYNB = 0, XYNBS = 1, XYNBSXSY = 2, ASTIG = 3 };
enum Camera { IDEAL = 0, SCMOS = 1 };
enum Precision { PREC_FP32 = 0, PREC_MIXED = 1 };

constexpr int kMinBox = 3;
constexpr int kMaxBox = 32;        // plain fits (gmle_fit, gmle_fit_device)
constexpr int kMaxSmallBox = 16;   // above: the large-box kernels (8 or 16 lanes per ROI, rows outer)
constexpr int kMaxBatchBox = 16;   // mixed batches, the service, ROI extraction
constexpr int kWave = 64;

__host__ __device__ constexpr int n_params(int model) {
  return model == XYNB ? 4 : model == XYNBS ? 5 : model == XYNBSXSY ? 6 : 5;
}

// The fit kernels built: every box size kMinBox..kMaxSmallBox with 2, 4, 8 or
// 16 lanes per ROI, and with 1 lane up to 12 px, where one lane can still keep
// a whole ROI resident in VGPRs; above kMaxSmallBox (the reference accepts any
// box, src/interface.jl:22; its SMITE heritage's largest window is 21 px,
// src/constants.jl:5) 8 or 16 lanes up to 24 px and 16 lanes up to kMaxBox,
// so that a lane keeps at most 3 columns.
__host__ __device__ constexpr bool lanes_built(int box, int lanes) {
  if (box < kMinBox || box > kMaxBox) return false;
  if (box > kMaxSmallBox) return lanes == 16 || (lanes == 8 && box <= 24);
  return (lanes == 1 && box <= 12) || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16;
}

// Largest box size whose kernel with `lanes` lanes per ROI still fits 256
// VGPRs without spilling, i.e. runs two or more waves per SIMD, with a few
// registers to spare for the multi-shape kernel's call.  A lane keeps its
// ceil(box / lanes) pixel columns (and, for sCMOS, their variances) in VGPRs
// for the whole fit.  From the gfx950 build's register counts
// (kernel-resource-usage): 4 lanes fit every box except astigmatic sCMOS ROIs
// of 14 px and up, 8 lanes every box but 16 px outside XYNB, 16 lanes all.
__host__ __device__ constexpr int max_box_within_256(int model, int camera, int lanes) {
  const bool sc = camera == SCMOS;
  if (lanes == 1)
    return model == XYNB ? (sc ? 6 : 11) : model == XYNBS ? (sc ? 5 : 6) : model == XYNBSXSY ? (sc ? 5 : 7)
                                                                                             : (sc ? 5 : 8);
  if (lanes == 2)
    return model == XYNB ? (sc ? 11 : 14) : model == XYNBS ? (sc ? 9 : 11) : model == XYNBSXSY ? (sc ? 9 : 11)
                                                                                               : (sc ? 8 : 10);
  if (lanes == 4) return (model == ASTIG && sc) ? 13 : kMaxSmallBox;
  if (lanes == 8) return model == XYNB ? kMaxSmallBox : 15;
  return kMaxSmallBox;
}

// Lanes per ROI of a large box (above kMaxSmallBox) in a large batch: 8 where
// that kernel fits 256 VGPRs (two waves per SIMD), else 16.  From the gfx950
// build's register counts: 8 lanes fit up to 24 px (216-256 VGPRs) except
// astigmatic sCMOS ROIs above 20 px; 16 lanes fit up to 32 px (161-245).
__host__ __device__ constexpr int large_box_lanes(int model, int camera, int box) {
  return (box <= 24 && !(model == ASTIG && camera == SCMOS && box > 20)) ? 8 : 16;
}

// Passed by value to every fitting kernel.  Arrays follow the reference's
// Julia layout: data (box, box, n) column-major with ROI k at offset box*box*k;
// params / crlb (P, n) with ROI k's P values contiguous; var_map (ny, nx)
// column-major indexed [row = y, col = x]; corners 1-based.
struct FitArgs {
  const float* data;
  const float* var_map;
  const int32_t* x_corners;
  const int32_t* y_corners;
  float* params;
  float* crlb;
  float* llr;
  double* crlb64;   // optional (mixed precision); may be null
  double* llr64;    // optional (mixed precision); may be null
  int64_t n_rois;
  int ny, nx;
  int iterations;
  float psf[8];     // XYNB: [sigma]; ASTIG: [sx0, sy0, Ax, Ay, Bx, By, gamma, d] (pixel units)
  float lower[6];
  float upper[6];
  float max_step[6];
  // Mixed-shape batches.  When `order` is set, this launch fits the ROIs
  // order[first], ..., order[first + n_rois - 1] of a concatenated batch:
  // ROI k's pixels start at data + offsets[k], its corner is
  // (x_corners[k], y_corners[k]) and its results fill slot k of params / crlb
  // (out_stride values per slot, NaN past P) and of llr.  Negative order
  // entries are skipped.  Without `order`, ROI k sits at data + box*box*k.
  const int32_t* order;
  const int64_t* offsets;
  int64_t first;
  int out_stride;   // 0: P
};

// Host-side launcher of one (model, camera, box, lanes-per-ROI, precision)
// specialisation; returns hipSuccess or the launch error.
typedef hipError_t (*LaunchFn)(const FitArgs&, hipStream_t);

// Lookup implemented by the per-model translation units.
LaunchFn find_launcher(int model, int camera, int box, int lanes, int precision);

// The kernels behind a launcher, for callers that build HIP graphs: the fit
// kernel (grid: ceil(n / rois_per_block) blocks of kWave threads) and, for
// mixed precision, the Float64 final pass (ceil(n / final_rois_per_block)
// blocks); both take one FitArgs by value.
struct KernelInfo {
  const void* fit;
  const void* final_pass;   // null for PREC_FP32
  int rois_per_block;
  int final_rois_per_block; // the final pass's grid: ceil(n / final_rois_per_block) blocks
};
typedef bool (*DescribeFn)(KernelInfo*);
bool find_kernels(int model, int camera, int box, int lanes, int precision, KernelInfo* out);

}  // namespace gmle
