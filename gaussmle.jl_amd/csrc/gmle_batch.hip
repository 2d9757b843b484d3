// gmle_batch.hip -- device-side grouping of a mixed-shape batch by shape.
#include "gmle_batch.hpp"

namespace gmle {
namespace {

constexpr int kBlock = kBatchBlock;

__device__ void nan_slot(const BatchPlan& p, int64_t i) {
  const float qf = __builtin_nanf("");
#pragma unroll
  for (int k = 0; k < kBatchStride; ++k) {
    p.params[i * kBatchStride + k] = qf;
    p.crlb[i * kBatchStride + k] = qf;
  }
  p.llr[i] = qf;
  if (p.crlb64) {
#pragma unroll
    for (int k = 0; k < kBatchStride; ++k) p.crlb64[i * kBatchStride + k] = __builtin_nan("");
  }
  if (p.llr64) p.llr64[i] = __builtin_nan("");
}

__global__ __launch_bounds__(kBlock) void batch_init(BatchPlan p) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < p.n_rois) p.order[i] = -1;
  if (blockIdx.x == 0 && threadIdx.x < kNumShapes) p.cursor[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void batch_scatter(BatchPlan p) {
  __shared__ int hist[kNumShapes];
  __shared__ int base[kNumShapes];
  const int t = threadIdx.x;
  if (t < kNumShapes) hist[t] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBlock + t;
  int key = -1, rank = 0;
  if (i < p.n_rois) {
    key = shape_index(p.models[i], p.boxes[i]);
    if (key >= 0 && p.count[key] == 0) key = -1;
    if (key >= 0) rank = atomicAdd(&hist[key], 1);
  }
  __syncthreads();
  if (t < kNumShapes && hist[t] > 0) base[t] = atomicAdd(&p.cursor[t], hist[t]);
  __syncthreads();
  if (i >= p.n_rois) return;
  if (key >= 0) {
    const int pos = base[key] + rank;
    if (pos < p.count[key]) {
      p.order[p.start[key] + pos] = (int32_t)i;
      return;
    }
  }
  nan_slot(p, i);   // unknown shape, or more ROIs of a shape than its count
}

__global__ __launch_bounds__(kBlock) void batch_zero(int32_t* counts) {
  if (threadIdx.x < kNumShapes) counts[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void batch_count(const int8_t* models, const int8_t* boxes, int64_t n,
                                                      int32_t* counts) {
  __shared__ int hist[kNumShapes];
  const int t = threadIdx.x;
  if (t < kNumShapes) hist[t] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBlock + t;
  if (i < n) {
    const int key = shape_index(models[i], boxes[i]);
    if (key >= 0) atomicAdd(&hist[key], 1);
  }
  __syncthreads();
  if (t < kNumShapes && hist[t] > 0) atomicAdd(&counts[t], hist[t]);
}

}  // namespace

const void* batch_init_kernel() { return reinterpret_cast<const void*>(&batch_init); }
const void* batch_scatter_kernel() { return reinterpret_cast<const void*>(&batch_scatter); }

hipError_t enqueue_batch_plan(const BatchPlan& p, hipStream_t stream) {
  if (p.n_rois <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((p.n_rois + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(batch_init, dim3(blocks), dim3(kBlock), 0, stream, p);
  hipLaunchKernelGGL(batch_scatter, dim3(blocks), dim3(kBlock), 0, stream, p);
  return hipGetLastError();
}

hipError_t enqueue_batch_count(const int8_t* models, const int8_t* boxes, int64_t n, int32_t* counts,
                               hipStream_t stream) {
  hipLaunchKernelGGL(batch_zero, dim3(1), dim3(kBlock), 0, stream, counts);
  if (n > 0) {
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(batch_count, dim3(blocks), dim3(kBlock), 0, stream, models, boxes, n, counts);
  }
  return hipGetLastError();
}

}  // namespace gmle
