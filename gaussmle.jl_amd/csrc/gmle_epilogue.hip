// gmle_epilogue.hip -- per-emitter result conversion on the device.
//
// GaussMLE.jl finishes every fit on the host, one emitter at a time:
//   * goodness of fit: chi2 = -2 LLR with box^2 - P degrees of freedom,
//     p = 1 - Float32(cdf(Chisq(df), chi2))            (src/interface.jl:265-275)
//   * ROI -> camera coordinates: x_corner + x_roi - 1  (src/roi_batch.jl:29-34)
//   * camera -> microns: (x_camera - 1) * pixel_size, uncertainties and fitted
//     widths times the pixel size, z and its uncertainty unscaled
//                                                      (src/roi_batch.jl:125-302)
// gmle_epilogue_device does all of it for a whole batch in one launch, in the
// same arithmetic types (Float32 values, Float64 pixel size, Float64 CDF).
//
// The chi-square CDF is the regularized lower incomplete gamma function
// P(df/2, chi2/2), evaluated in Float64 by the series
//   P(a, x) = x^a e^-x / Gamma(a + 1) * sum_n x^n / ((a + 1) ... (a + n))
// for x < a + 1 and otherwise as 1 - Q(a, x) with Q from its continued
// fraction (Legendre), evaluated by the modified Lentz method (Numerical
// Recipes, 3rd ed., sec. 6.2).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/gaussmle.h"
#include "gmle_args.hpp"
#include "gmle_host.hpp"

namespace gmle {
namespace {

// Relative stopping tolerance of both expansions: a few Float64 ulp.  (A
// tolerance below one ulp would never be met by the continued fraction, whose
// ratio settles within an ulp of 1, and would run it to the iteration cap.)
constexpr double kTol = 1e-15;

// x <= 0 gives 0 (a chi^2 below zero, from an LLR rounded just above zero,
// has CDF 0 as in Distributions.jl, whose gamma CDF clamps x at 0); NaN stays NaN.
__host__ __device__ inline double gamma_p(double a, double x) {
  if (x != x) return x;
  if (!(x > 0.0)) return 0.0;
  if (isinf(x)) return 1.0;
  const double lg = lgamma(a);
  const double lead = exp(-x + a * log(x) - lg);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int n = 0; n < 1000; ++n) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * kTol) break;
    }
    return sum * lead;
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < kTol) break;
  }
  return 1.0 - lead * h;
}

struct EpilogueArgs {
  int64_t n;
  int model, box;                 // used when models / boxes are null
  const int8_t* models;
  const int8_t* boxes;
  int stride;                     // params / crlb values per ROI
  const float* params;
  const float* crlb;
  const float* llr;
  const int32_t* xc;              // null: corners of fit(Array), x = 1 + k box, y = 1
  const int32_t* yc;
  double psx, psy;
  float* out;                     // (kEpilogueOut, n)
};

constexpr int kOut = 10;

__global__ __launch_bounds__(256) void epilogue_kernel(EpilogueArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const int model = a.models ? a.models[k] : a.model;
  const int box = a.boxes ? a.boxes[k] : a.box;
  float* o = a.out + k * kOut;
  const float qnan = __builtin_nanf("");
  if (model < XYNB || model > ASTIG || box < kMinBox || box > kMaxBox) {
    for (int q = 0; q < kOut; ++q) o[q] = qnan;
    return;
  }
  const float* p = a.params + k * a.stride;
  const float* u = a.crlb + k * a.stride;
  const int32_t xc = a.xc ? a.xc[k] : (int32_t)(1 + k * box);
  const int32_t yc = a.yc ? a.yc[k] : 1;
  // roi_to_camera_coords: Int32 corner + Float32 ROI coordinate - 1, in Float32
  const float xcam = ((float)xc + p[0]) - 1.0f;
  const float ycam = ((float)yc + p[1]) - 1.0f;
  o[0] = (float)((double)(xcam - 1.0f) * a.psx);
  o[1] = (float)((double)(ycam - 1.0f) * a.psy);
  o[2] = (float)((double)u[0] * a.psx);
  o[3] = (float)((double)u[1] * a.psy);
  const int P = n_params(model);
  const float chi2 = -2.0f * a.llr[k];
  const double cdf = gamma_p(0.5 * (double)(box * box - P), 0.5 * (double)chi2);
  o[4] = 1.0f - (float)cdf;
  o[5] = o[6] = o[7] = o[8] = qnan;
  if (model == XYNBS) {   // one width, in x pixels (src/roi_batch.jl, Emitter2DFitSigma)
    o[5] = o[6] = (float)((double)p[4] * a.psx);
    o[7] = o[8] = (float)((double)u[4] * a.psx);
  } else if (model == XYNBSXSY) {
    o[5] = (float)((double)p[4] * a.psx);
    o[6] = (float)((double)p[5] * a.psy);
    o[7] = (float)((double)u[4] * a.psx);
    o[8] = (float)((double)u[5] * a.psy);
  }
  o[9] = chi2;
}

}  // namespace
}  // namespace gmle

using namespace gmle;

extern "C" {

int gmle_epilogue_device(int64_t n_rois, int model, int box, const int8_t* d_models, const int8_t* d_boxes,
                         int stride, const float* d_params, const float* d_crlb, const float* d_llr,
                         const int32_t* d_x_corners, const int32_t* d_y_corners, double pixel_size_x,
                         double pixel_size_y, float* d_out, void* stream) {
  if (n_rois <= 0) return fail(GMLE_ERR_ARG, "Input data array is empty");
  if (!d_params || !d_crlb || !d_llr || !d_out) return fail(GMLE_ERR_ARG, "device pointers must not be NULL");
  if (!d_models && (model < XYNB || model > ASTIG)) return fail(GMLE_ERR_ARG, "unknown PSF model %d", model);
  if (!d_boxes && (box < kMinBox || box > kMaxBox)) return fail(GMLE_ERR_ARG, "box size %d out of range", box);
  if ((d_x_corners == nullptr) != (d_y_corners == nullptr))
    return fail(GMLE_ERR_ARG, "pass both corner arrays or neither");
  const int pmax = d_models ? 6 : n_params(model);
  if (stride < pmax) return fail(GMLE_ERR_ARG, "stride %d is smaller than the parameter count %d", stride, pmax);
  if (!(pixel_size_x > 0.0) || !(pixel_size_y > 0.0)) return fail(GMLE_ERR_ARG, "pixel sizes must be > 0");
  EpilogueArgs a{n_rois, model, box, d_models, d_boxes, stride, d_params, d_crlb, d_llr, d_x_corners, d_y_corners,
                 pixel_size_x, pixel_size_y, d_out};
  const unsigned blocks = (unsigned)((n_rois + 255) / 256);
  hipLaunchKernelGGL(epilogue_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? GMLE_OK : hip_fail(e, "kernel launch");
}

// The chi-square survival function as the epilogue computes it, on the host
// (1 - Float32(P(df/2, chi2/2))): for callers and tests without a device.
float gmle_chi2_pvalue(float chi2, int df) {
  if (df <= 0) return __builtin_nanf("");
  return 1.0f - (float)gamma_p(0.5 * df, 0.5 * (double)chi2);
}

}  // extern "C"
