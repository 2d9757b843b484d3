// gmle_erf.hpp -- Float32 pixel integrals of the Gaussian that reuse exp(-u^2).
//
// The separable model needs, at every pixel edge, erf and the Gaussian
// exp(-u^2) (u = t / (sigma sqrt 2)); the Gaussian is needed anyway for the
// derivative tables, so the erf side costs one reciprocal and one polynomial.
// The kernels work in the scaled variable
//   v = k u,  k = sqrt(log2 e),  so that  exp(-u^2) = exp2(-v^2)
// (one multiply and one v_exp_f32 per edge), and evaluate
//   erfc(|u|) = exp2(-v^2) R(t),   t = 1 - 2 c / (|v| + c),   R(t) ~ erfcx(|u|)
// The map t takes |v| in [0, inf) onto [-1, 1) and turns erfcx's slow 1/u
// tail into a smooth function of t, so a degree-10 polynomial R (fitted by
// gen_erf_coeffs.py: Lawson-reweighted least squares on Chebyshev nodes, over
// |u| <= 10, beyond which exp2(-v^2) underflows) has the accuracy a degree-16
// polynomial in |v| had.  Measured Float32 error of R: 3.7e-7 relative with a
// correctly rounded reciprocal, 5.8e-7 with the reciprocal one ulp off (what
// v_rcp_f32 may return).  Per edge: one add, one v_rcp_f32, one FMA for t and
// ten FMAs, against a clamp, a subtraction and sixteen FMAs before.
//
// Pixel integrals.  (erf(b) - erf(a)) / 2 of a pixel in the tails subtracts
// two values close to +-1 and keeps only their rounding: an absolute error of
// ~1e-7 that N times the integral turns into a model error comparable to a
// low background.  The kernels therefore carry each edge as its signed
// complement se = sign(u) erfc(|u|) and its half-sign hs = sign(u) / 2, since
// erf(u) = 2 hs - se, and form
//   Phi = (erf(b) - erf(a)) / 2 = (se_a - se_b) / 2 + (hs_b - hs_a)
// where the second term is 0, or 1 for the pixel that straddles the centre.
// Every pixel integral then carries R's relative accuracy.  Callable from
// host code too, for testing (the reciprocal is then correctly rounded).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace gmle {

constexpr float kErfK = 1.2011224031448364f;   // sqrt(log2 e): v = k u
constexpr float kErfC = 3.0028059482574463f;   // c = k * 2.5: t = 0 at |u| = 2.5
constexpr float kErfC2 = 6.005611896514893f;   // 2 c (exact)

// 1/x: v_rcp_f32 on the device (1 ulp), a division in host code.
__host__ __device__ __forceinline__ float erf_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// R(t), t = 1 - 2c / (|v| + c): erfcx(|v| / k).  v = +-inf gives t = 1
// (finite), NaN propagates.
__host__ __device__ __forceinline__ float erfcx_poly(float v) {
  const float t = fmaf(-kErfC2, erf_rcp(fabsf(v) + kErfC), 1.0f);
  float r = -3.776571247726679e-05f;
  r = fmaf(r, t, -0.00011116439418401569f);
  r = fmaf(r, t, 0.00023663679894525558f);
  r = fmaf(r, t, 0.0008690135437063873f);
  r = fmaf(r, t, -0.0026459149084985256f);
  r = fmaf(r, t, -0.0039923484437167645f);
  r = fmaf(r, t, 0.039923105388879776f);
  r = fmaf(r, t, -0.1250331848859787f);
  r = fmaf(r, t, 0.2517131567001343f);
  r = fmaf(r, t, -0.3717367351055145f);
  r = fmaf(r, t, 0.21080636978149414f);
  return r;
}

// One pixel edge for the pixel integrals, from v = k u and g = exp(-u^2):
// se = sign(u) erfc(|u|), hs = sign(u) / 2.
__host__ __device__ __forceinline__ void erf_edge(float v, float g, float& se, float& hs) {
  se = copysignf(g * erfcx_poly(v), v);
  hs = copysignf(0.5f, v);
}

// erf(u) from v = k u and g = exp(-u^2) (for tests: the kernels use erf_edge).
__host__ __device__ __forceinline__ float erf_with_gauss(float v, float g) {
  return copysignf(fmaf(-g, erfcx_poly(v), 1.0f), v);
}

// (erf(hi) - erf(lo)) / 2 from the two edges' (se, hs), without cancellation.
template <class T>
__host__ __device__ __forceinline__ T pixel_integral(T se_lo, T hs_lo, T se_hi, T hs_hi) {
  return T(0.5) * (se_lo - se_hi) + (hs_hi - hs_lo);
}

}  // namespace gmle
