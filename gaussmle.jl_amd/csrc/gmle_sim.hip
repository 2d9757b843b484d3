// gmle_sim.hip -- the simulator on the device: noisy ROI stacks for tests and
// benchmarks, as GaussMLE.jl's generate_roi_batch produces them
// (src/simulator.jl:35-345).
//
// Per pixel (i = row, j = column, 1-based), in Float32 as the reference:
//   expected = bg + photons * Phi(j; x, sigma_x) * Phi(i; y, sigma_y)
//              (src/simulator.jl:312-344; Phi as src/gausslib/baselibrary.jl:8-14,
//              astigmatic widths sigma0 sqrt(alpha((z -+ gamma) / d)))
//   Ideal camera: Poisson(expected) if expected > 0, else 0    (:259-275)
//   sCMOS: electrons = Poisson(expected * qe) + N(0, 1) * readnoise,
//          ADU = electrons / gain + offset                      (:281-307)
// Sensor maps are indexed the way the fitting kernel reads the variance map,
// [row = y_corner + i - 1, column = x_corner + j - 1] (src/unified_kernel.jl:
// 279-281); the reference simulator indexes [i + x_corner - 1, j + y_corner -
// 1] instead, which transposes them relative to the kernel.
// Random numbers come from Philox4x32-10 keyed by the seed and counted by
// (ROI, pixel, purpose), so the output depends on the seed only.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/gaussmle.h"
#include "gmle_args.hpp"
#include "gmle_host.hpp"
#include "gmle_rng.hpp"

namespace gmle {
namespace {

constexpr int kSimBlock = 256;
constexpr int kSimMaxBlocks = 16384;

struct SimArgs {
  int model;
  float psf[8];
  int box;
  int64_t n;
  const float* params;   // per ROI: `stride` values, the model's parameters first
  int stride;
  int camera;
  const int32_t* xc;
  const int32_t* yc;
  const float* qe_map;   // (ny, nx) column-major, or null for the scalar
  const float* rn_map;
  const float* gain_map;
  const float* offset_map;
  int ny, nx;
  float qe, readnoise, gain, offset;
  uint64_t seed;
  float* out;            // (box, box, n)
};

// integral_gaussian_1d (src/gausslib/baselibrary.jl:8-14) in Float32.
__host__ __device__ inline float pixel_integral(int k, float mu, float sigma) {
#pragma clang fp contract(off)
  const float s = sqrtf(0.5f / (sigma * sigma));
  return 0.5f * (erff(((float)k - mu + 0.5f) * s) - erff(((float)k - mu - 0.5f) * s));
}

__host__ __device__ inline float astig_alpha(float u, float A, float B) {
#pragma clang fp contract(off)
  return 1.0f + u * u + A * (u * u * u) + B * (u * u * u * u);
}

// Element e = k box^2 + p of the output: ROI k, pixel p (row-fastest).
__host__ __device__ inline float sim_pixel(const SimArgs& a, int64_t e) {
#pragma clang fp contract(off)   // Float32 products and sums as written, like the reference
  const int B = a.box, BB = a.box * a.box;
  const int64_t k = e / BB;
  const int p = (int)(e - k * BB);
  const int i = p % B + 1, j = p / B + 1;   // row, column (1-based)
  // The ROI's parameters as values first (the model then selects among
  // values, not among pointers into the row).
  const float* th = a.params + k * a.stride;
  const int np = a.stride < 6 ? a.stride : 6;
  float t[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) t[q] = q < np ? th[q] : 0.0f;
  const float x = t[0], y = t[1];
  const bool astig = a.model == ASTIG;
  const float N = astig ? t[3] : t[2];
  const float bg = astig ? t[4] : t[3];
  float sx, sy;
  if (a.model == XYNB) {
    sx = sy = a.psf[0];
  } else if (a.model == XYNBS) {
    sx = sy = t[4];
  } else if (a.model == XYNBSXSY) {
    sx = t[4];
    sy = t[5];
  } else {
    const float z = t[2];
    sx = a.psf[0] * sqrtf(astig_alpha((z - a.psf[6]) / a.psf[7], a.psf[2], a.psf[4]));
    sy = a.psf[1] * sqrtf(astig_alpha((z + a.psf[6]) / a.psf[7], a.psf[3], a.psf[5]));
  }
  const float expected = bg + N * pixel_integral(j, x, sx) * pixel_integral(i, y, sy);
  Stream shot(a.seed, k, (uint32_t)p, 0u);
  if (a.camera == IDEAL) return expected > 0.0f ? (float)sample_poisson(shot, (double)expected) : 0.0f;
  const int cy = a.yc[k], cx = a.xc[k];
  if (cy < 1 || cx < 1 || cy > a.ny - B + 1 || cx > a.nx - B + 1) return __builtin_nanf("");   // off the sensor
  const int64_t m = (int64_t)(cy - 1 + i - 1) + (int64_t)a.ny * (int64_t)(cx - 1 + j - 1);
  const float qe = a.qe_map ? a.qe_map[m] : a.qe;
  const float rn = a.rn_map ? a.rn_map[m] : a.readnoise;
  const float gain = a.gain_map ? a.gain_map[m] : a.gain;
  const float offset = a.offset_map ? a.offset_map[m] : a.offset;
  const float el = expected * qe;
  const float sig = el > 0.0f ? (float)sample_poisson(shot, (double)el) : 0.0f;
  Stream read(a.seed, k, (uint32_t)p, 1u);
  const float total_e = sig + (float)sample_normal(read) * rn;
  return total_e / gain + offset;
}

__global__ __launch_bounds__(kSimBlock) void sim_kernel(SimArgs a) {
  const int64_t total = a.n * a.box * a.box;
  for (int64_t e = (int64_t)blockIdx.x * kSimBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kSimBlock)
    a.out[e] = sim_pixel(a, e);
}

int check_sim(int model, const float* psf, int box, int64_t n, int stride, int camera, int ny, int nx) {
  if (n <= 0) return fail(GMLE_ERR_ARG, "n_rois must be > 0");
  if (model < XYNB || model > ASTIG) return fail(GMLE_ERR_ARG, "unknown PSF model %d", model);
  if ((model == XYNB || model == ASTIG) && !psf) return fail(GMLE_ERR_ARG, "psf must not be NULL for this model");
  if (box < kMinBox || box > kMaxBox) return fail(GMLE_ERR_ARG, "box size %d out of range", box);
  if (stride < n_params(model)) return fail(GMLE_ERR_ARG, "stride %d < %d parameters", stride, n_params(model));
  if (camera != IDEAL && camera != SCMOS) return fail(GMLE_ERR_ARG, "camera must be 0 (ideal) or 1 (sCMOS)");
  if (camera == SCMOS && (ny < box || nx < box)) return fail(GMLE_ERR_ARG, "sensor %d x %d smaller than the box", ny, nx);
  return GMLE_OK;
}

}  // namespace
}  // namespace gmle

using namespace gmle;

extern "C" {

void gmle_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out) {
  const U32x4 r = philox4x32_10(U32x4{{counter[0], counter[1], counter[2], counter[3]}}, key[0], key[1]);
  for (int q = 0; q < 4; ++q) out[q] = r.v[q];
}

int gmle_simulate_device(int model, const float* psf, int box, int64_t n_rois, const float* d_params, int stride,
                         int camera, const int32_t* d_x_corners, const int32_t* d_y_corners, const float* d_qe,
                         const float* d_readnoise, const float* d_gain, const float* d_offset, int ny, int nx,
                         float qe, float readnoise, float gain, float offset, uint64_t seed, float* d_out,
                         void* stream) {
  int rc = check_sim(model, psf, box, n_rois, stride, camera, ny, nx);
  if (rc != GMLE_OK) return rc;
  if (!d_params || !d_out) return fail(GMLE_ERR_ARG, "params and out must not be NULL");
  if (camera == SCMOS && (!d_x_corners || !d_y_corners)) return fail(GMLE_ERR_ARG, "sCMOS needs ROI corners");
  if (camera == SCMOS && (!(gain != 0.0f) && !d_gain)) return fail(GMLE_ERR_ARG, "gain must be non-zero");
  SimArgs a{};
  a.model = model;
  // XYNB reads one float (sigma), ASTIG eight, the others none, as gmle_fit.
  if (model == XYNB) a.psf[0] = psf[0];
  if (model == ASTIG)
    for (int q = 0; q < 8; ++q) a.psf[q] = psf[q];
  a.box = box;
  a.n = n_rois;
  a.params = d_params;
  a.stride = stride;
  a.camera = camera;
  a.xc = d_x_corners;
  a.yc = d_y_corners;
  a.qe_map = d_qe;
  a.rn_map = d_readnoise;
  a.gain_map = d_gain;
  a.offset_map = d_offset;
  a.ny = ny;
  a.nx = nx;
  a.qe = qe;
  a.readnoise = readnoise;
  a.gain = gain;
  a.offset = offset;
  a.seed = seed;
  a.out = d_out;
  const int64_t total = n_rois * box * box;
  const unsigned blocks = (unsigned)std::min<int64_t>(kSimMaxBlocks, (total + kSimBlock - 1) / kSimBlock);
  hipLaunchKernelGGL(sim_kernel, dim3(blocks), dim3(kSimBlock), 0, (hipStream_t)stream, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? GMLE_OK : hip_fail(e, "kernel launch");
}

}  // extern "C"
