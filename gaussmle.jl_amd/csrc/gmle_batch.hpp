// gmle_batch.hpp -- planning kernels of mixed-shape batches.
//
// A batch mixes box sizes and PSF models.  VGPR allocation is fixed per
// kernel, so one kernel holding every shape's fit would run all of them at the
// occupancy (and spills) of the heaviest one; instead every shape keeps its
// own fit kernel (the plain one, in gather mode: FitArgs::order) and the batch
// is grouped by shape on the device first:
//   batch_init     order[] = -1, per-shape cursors = 0
//   batch_scatter  ROI indices grouped by shape, into ranges planned on the
//                  host from the shape counts; per-block LDS ranks, one global
//                  atomic per (block, shape)
//   batch_count    shape histogram, for callers without host-side counts
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmle_args.hpp"

namespace gmle {

constexpr int kNumBoxes = kMaxBatchBox - kMinBox + 1;
constexpr int kNumShapes = 4 * kNumBoxes;
constexpr int kBatchStride = 6;   // params / crlb values per ROI slot

__host__ __device__ constexpr int shape_index(int model, int box) {
  return (model < XYNB || model > ASTIG || box < kMinBox || box > kMaxBatchBox) ? -1
                                                                             : model * kNumBoxes + (box - kMinBox);
}

struct BatchPlan {
  const int8_t* models;
  const int8_t* boxes;
  int64_t n_rois;
  int32_t* order;    // n_rois slots, grouped by shape
  int32_t* cursor;   // kNumShapes
  float* params;     // (kBatchStride, n) -- NaN for ROIs that are not fitted
  float* crlb;
  float* llr;
  double* crlb64;    // optional
  double* llr64;     // optional
  int32_t start[kNumShapes];   // first slot of each shape
  int32_t count[kNumShapes];   // slots of each shape
};

// Workspace: order (n) | cursor (kNumShapes) | counts (kNumShapes), 256-byte aligned pieces.
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
inline size_t batch_workspace(int64_t n) {
  return align256((size_t)n * sizeof(int32_t)) + 2 * align256(kNumShapes * sizeof(int32_t));
}

constexpr int kBatchBlock = 256;   // threads per block of the grouping kernels

// The grouping kernels themselves (one BatchPlan by value; grid
// ceil(n_rois / kBatchBlock) blocks of kBatchBlock threads), for HIP graphs.
const void* batch_init_kernel();
const void* batch_scatter_kernel();

// Enqueue batch_init + batch_scatter.
hipError_t enqueue_batch_plan(const BatchPlan& p, hipStream_t stream);
// Enqueue batch_count into counts[kNumShapes] (device).
hipError_t enqueue_batch_count(const int8_t* models, const int8_t* boxes, int64_t n, int32_t* counts,
                               hipStream_t stream);

}  // namespace gmle
