// Multi-shape kernel (every shape, 4 lanes per ROI, Float32), camera SCMOS.
#include "gmle_multi.hpp"
namespace gmle {
hipError_t launch_multi_scmos(const MultiArgs& m, hipStream_t stream) {
  const int32_t tiles = m.tile_start[m.n_entries];
  if (tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL((multi_fit<SCMOS>), dim3((unsigned)tiles), dim3(kWave), 0, stream, m);
  return hipGetLastError();
}
const void* multi_kernel_scmos() { return reinterpret_cast<const void*>(&multi_fit<SCMOS>); }
}  // namespace gmle
