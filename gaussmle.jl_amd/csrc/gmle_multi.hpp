// gmle_multi.hpp -- every shape of a mixed batch in one launch.
//
// multi_fit<CAM> runs tiles of 64/L ROIs (L lanes per ROI) of any
// (model, box) shape: block t finds its shape in the launch's tile table
// (wave-uniform) and calls that shape's tile function.  The tile functions
// are separate, non-inlined functions, so each is register-allocated on its
// own: the kernel needs the largest of them (at most 256 VGPRs, two waves per
// SIMD) instead of what one body holding every shape would need.  Each shape
// offers its 4-lane tile and, where they fit 256 VGPRs, 1-, 2-, 8- and
// 16-lane tiles (8 and 16 lanes give small batches more, shorter waves).
// The per-ROI code is fit_one, as in the per-shape kernels, so results are
// bit-identical to a per-shape launch with the same lanes per ROI.
#pragma once

#include "gmle_batch.hpp"
#include "gmle_fit_kernel.hpp"

namespace gmle {

// Shapes and lanes the multi-shape kernel holds: those whose kernel stays
// within 256 VGPRs (two waves per SIMD).  Shapes without a 4-lane tile here
// (astigmatic sCMOS ROIs of 14 px and up) keep their own launch.
__host__ __device__ constexpr bool multi_has(int model, int camera, int box, int lanes) {
  return (lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16) && box >= kMinBox &&
         box <= max_box_within_256(model, camera, lanes);
}
__host__ __device__ constexpr bool multi_eligible(int model, int camera, int box) {
  return multi_has(model, camera, box, 4);
}
// Lanes used for a shape in the multi-shape kernel when `want` is requested.
__host__ __device__ constexpr int multi_lanes(int model, int camera, int box, int want) {
  return multi_has(model, camera, box, want) ? want : 4;
}

struct MultiArgs {
  FitArgs fa[4];                  // per model: psf / constraints; shared data, order and outputs
  int32_t n_entries;
  int32_t shape[kNumShapes];      // entry q: shape index
  int32_t lanes[kNumShapes];      //          lanes per ROI (multi_has)
  int32_t first[kNumShapes];      //          its first slot in order[]
  int32_t count[kNumShapes];      //          its slots
  int32_t block_start[kNumShapes + 1];   // first block (= tile) of entry q; total at n_entries
};

// Tile t of one shape.  One tile per call and per block: a loop over tiles,
// in here or around the call, keeps state live across fit_one and costs
// registers (the kernel would drop to one wave per SIMD).
template <int MODEL, int CAM, int N, int L>
__device__ __attribute__((noinline)) void multi_tile(const FitArgs& a, int32_t first, int32_t count, int32_t t) {
  constexpr int RPT = kWave / L;   // ROIs per tile
  const int lane = threadIdx.x;
  const int r = (L == 1) ? 0 : (lane & (L - 1));
  // An entry whose first slot is empty holds no ROI of this shape.
  const int32_t k0 = __builtin_amdgcn_readfirstlane(a.order[first]);
  if (k0 < 0) return;
  const int32_t j = t * RPT + lane / L;
  const bool in = j < count;
  int32_t k = a.order[first + (in ? j : 0)];
  const bool active = in && k >= 0;
  if (k < 0) k = k0;
  RoiRef ref;
  ref.data = a.data + a.offsets[k];
  ref.out = k;
  ref.stride = kBatchStride;
  ref.cx = ref.cy = 1;
  if constexpr (CAM == SCMOS) {
    ref.cx = a.x_corners[k];
    ref.cy = a.y_corners[k];
  }
  fit_one<MODEL, CAM, N, L, PREC_FP32>(a, ref, active, r);
}

template <int CAM, int MODEL, int N>
__device__ __forceinline__ void multi_case(const MultiArgs& m, int q, int t) {
  const int L = m.lanes[q];
  if constexpr (multi_has(MODEL, CAM, N, 1)) {
    if (L == 1) { multi_tile<MODEL, CAM, N, 1>(m.fa[MODEL], m.first[q], m.count[q], t); return; }
  }
  if constexpr (multi_has(MODEL, CAM, N, 2)) {
    if (L == 2) { multi_tile<MODEL, CAM, N, 2>(m.fa[MODEL], m.first[q], m.count[q], t); return; }
  }
  if constexpr (multi_has(MODEL, CAM, N, 8)) {
    if (L == 8) { multi_tile<MODEL, CAM, N, 8>(m.fa[MODEL], m.first[q], m.count[q], t); return; }
  }
  if constexpr (multi_has(MODEL, CAM, N, 16)) {
    if (L == 16) { multi_tile<MODEL, CAM, N, 16>(m.fa[MODEL], m.first[q], m.count[q], t); return; }
  }
  if constexpr (multi_has(MODEL, CAM, N, 4)) {
    if (L == 4) multi_tile<MODEL, CAM, N, 4>(m.fa[MODEL], m.first[q], m.count[q], t);
  }
}

template <int CAM, int MODEL>
__device__ __forceinline__ void multi_box(const MultiArgs& m, int box, int q, int t) {
  switch (box) {
    case 3: multi_case<CAM, MODEL, 3>(m, q, t); break;
    case 4: multi_case<CAM, MODEL, 4>(m, q, t); break;
    case 5: multi_case<CAM, MODEL, 5>(m, q, t); break;
    case 6: multi_case<CAM, MODEL, 6>(m, q, t); break;
    case 7: multi_case<CAM, MODEL, 7>(m, q, t); break;
    case 8: multi_case<CAM, MODEL, 8>(m, q, t); break;
    case 9: multi_case<CAM, MODEL, 9>(m, q, t); break;
    case 10: multi_case<CAM, MODEL, 10>(m, q, t); break;
    case 11: multi_case<CAM, MODEL, 11>(m, q, t); break;
    case 12: multi_case<CAM, MODEL, 12>(m, q, t); break;
    case 13: multi_case<CAM, MODEL, 13>(m, q, t); break;
    case 14: multi_case<CAM, MODEL, 14>(m, q, t); break;
    case 15: multi_case<CAM, MODEL, 15>(m, q, t); break;
    case 16: multi_case<CAM, MODEL, 16>(m, q, t); break;
    default: break;
  }
}

// One block (one wave) per tile; tiles of the heaviest shapes come first.
template <int CAM>
__global__ __launch_bounds__(kWave) void multi_fit(MultiArgs m) {
  const int32_t blk = blockIdx.x;
  if (blk >= m.block_start[m.n_entries]) return;
  int q = 0;
  while (q + 1 < m.n_entries && m.block_start[q + 1] <= blk) ++q;
  const int s = m.shape[q];
  const int model = s / kNumBoxes, box = kMinBox + s % kNumBoxes;
  const int tt = blk - m.block_start[q];
  switch (model) {
    case XYNB: multi_box<CAM, XYNB>(m, box, q, tt); break;
    case XYNBS: multi_box<CAM, XYNBS>(m, box, q, tt); break;
    case XYNBSXSY: multi_box<CAM, XYNBSXSY>(m, box, q, tt); break;
    default: multi_box<CAM, ASTIG>(m, box, q, tt); break;
  }
}

// Fills block_start from the entries' counts and lanes (one tile per block).
inline void multi_plan_blocks(MultiArgs& m) {
  int32_t blocks = 0;
  for (int q = 0; q < m.n_entries; ++q) {
    m.block_start[q] = blocks;
    blocks += (m.count[q] + kWave / m.lanes[q] - 1) / (kWave / m.lanes[q]);
  }
  m.block_start[m.n_entries] = blocks;
}

// Launcher and kernel address of one camera's multi-shape kernel.
hipError_t launch_multi_ideal(const MultiArgs& m, hipStream_t stream);
hipError_t launch_multi_scmos(const MultiArgs& m, hipStream_t stream);
const void* multi_kernel_ideal();
const void* multi_kernel_scmos();

}  // namespace gmle
