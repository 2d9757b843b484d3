// gmle_noalloc.hip -- the mixed-shape batch on device arrays without any
// allocation, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// gmle_fit_batch_device takes its grouping workspace (the shape-ordered ROI
// table, per-shape cursors and counts) from the caller.  This entry point
// takes it from `scratch`, a file-scope device array of kScratchBytes that
// every GPU the library runs on holds until the process ends, and otherwise
// runs the same code: same kernels, bit-identical results.
//
// `scratch` names a host-side stand-in in host code, not a device address:
// the device address is looked up once per device with hipGetSymbolAddress
// (with that device current) and only that address is ever passed to
// kernels.  Every call on a GPU shares its array and overwrites it (the
// grouping kernels clear what must start at zero), so calls that overlap on
// one GPU race: serialise them, e.g. on one stream.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gaussmle.h"
#include "gmle_batch.hpp"
#include "gmle_host.hpp"

namespace gmle {
namespace {

constexpr size_t kScratchBytes = (size_t)GMLE_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address on each device, looked up on the first call made
// with that device current.
struct ScratchSlot {
  std::mutex mu;
  void* ptr = nullptr;
};
ScratchSlot g_slots[kMaxScratchDevices];

int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(GMLE_ERR_UNSUPPORTED, "device %d beyond the scratch table", dev);
  ScratchSlot& s = g_slots[dev];
  std::lock_guard<std::mutex> lock(s.mu);
  if (!s.ptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(GMLE_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    s.ptr = p;
  }
  *out = s.ptr;
  return GMLE_OK;
}

// Largest batch whose grouping workspace fits the array.
int64_t max_rois() {
  int64_t lo = 0, hi = (int64_t)(kScratchBytes / sizeof(int32_t));
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (batch_workspace(mid) <= kScratchBytes) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace
}  // namespace gmle

using namespace gmle;

extern "C" {

int gmle_fit_batch_device_noalloc(int64_t n_rois, const int8_t* d_models, const int8_t* d_boxes,
                                  const int64_t* d_offsets, const float* d_data, const float* psf,
                                  const float* lower, const float* upper, const float* max_step, int camera,
                                  const float* d_var_map, int ny, int nx, const int32_t* d_x_corners,
                                  const int32_t* d_y_corners, int iterations, int precision,
                                  const int32_t* lanes, int streams, const int64_t* shape_counts,
                                  float* d_params, float* d_crlb, float* d_llr, double* d_crlb64,
                                  double* d_llr64, void* stream) {
  const int64_t cap = max_rois();
  if (n_rois > cap)
    return fail(GMLE_ERR_ARG, "the allocation-free batch holds at most %lld ROIs (a %zu-byte static workspace), got %lld",
                (long long)cap, kScratchBytes, (long long)n_rois);
  if (streams != 0 && streams != 1)
    return fail(GMLE_ERR_ARG, "the allocation-free batch runs on the caller's stream only: streams must be 0 or 1, got %d",
                streams);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != GMLE_OK) return rc;
  return gmle_fit_batch_device(n_rois, d_models, d_boxes, d_offsets, d_data, psf, lower, upper, max_step, camera,
                               d_var_map, ny, nx, d_x_corners, d_y_corners, iterations, precision, lanes, streams,
                               shape_counts, d_params, d_crlb, d_llr, d_crlb64, d_llr64, ws, kScratchBytes, stream);
}

}  // extern "C"
