// Multi-shape kernel (every shape, 4 lanes per ROI, Float32), camera IDEAL.
#include "gmle_multi.hpp"
namespace gmle {
hipError_t launch_multi_ideal(const MultiArgs& m, hipStream_t stream) {
  const int32_t blocks = m.block_start[m.n_entries];
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((multi_fit<IDEAL>), dim3((unsigned)blocks), dim3(kWave), 0, stream, m);
  return hipGetLastError();
}
const void* multi_kernel_ideal() { return reinterpret_cast<const void*>(&multi_fit<IDEAL>); }
}  // namespace gmle
