// gmle_erf.hpp -- Float32 pixel integrals of the Gaussian that reuse exp(-u^2).
//
// The separable model needs, at every pixel edge, erf and the Gaussian
// exp(-u^2) (u = t / (sigma sqrt 2)); the Gaussian is needed anyway for the
// derivative tables, so the erf side costs one polynomial and no
// transcendental of its own.  The kernels work in the scaled variable
//   v = k u,  k = sqrt(log2 e),  so that  exp(-u^2) = exp2(-v^2)
// (one multiply and one v_exp_f32 per edge), and evaluate
//   erfc(|u|) = exp2(-v^2) R(min(|v|, kErfVMax) - kErfC),   R(v) ~ erfcx(v / k)
// with R a near-minimax polynomial of degree 16 fitted by gen_erf_coeffs.py
// (Lawson-reweighted least squares on Chebyshev nodes) over |u| <= 5.6, where
// erfc(5.6) = 1.5e-15 lies below any tail a fit can see.  Measured Float32
// FMA-Horner error: 3.6e-7 relative on erfcx.
//
// Pixel integrals.  (erf(b) - erf(a)) / 2 of a pixel in the tails subtracts
// two values close to +-1 and keeps only their rounding: an absolute error of
// ~1e-7 that N times the integral turns into a model error comparable to a
// low background.  The kernels therefore carry each edge as its signed
// complement se = sign(u) erfc(|u|) and its half-sign hs = sign(u) / 2, since
// erf(u) = 2 hs - se, and form
//   Phi = (erf(b) - erf(a)) / 2 = (se_a - se_b) / 2 + (hs_b - hs_a)
// where the second term is 0, or 1 for the pixel that straddles the centre.
// Every pixel integral then carries R's relative accuracy.  Beyond |u| = 5.6
// R is held at its end value (erfc overestimated by 16 % at u = 6.5, where
// erfc < 1e-19).  Callable from host code too, for testing.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace gmle {

constexpr float kErfK = 1.2011224031448364f;     // sqrt(log2 e): v = k u
constexpr float kErfVMax = 6.726285457611084f;   // k * 5.6
constexpr float kErfC = 3.363142728805542f;      // kErfVMax / 2

// R(min(|v|, kErfVMax) - kErfC) ~ erfcx(|v| / k).
__host__ __device__ __forceinline__ float erfcx_poly(float v) {
  const float x = fminf(fabsf(v), kErfVMax) - kErfC;
  float r = 1.5094743163746038e-11f;
  r = fmaf(r, x, -9.99696297965258e-11f);
  r = fmaf(r, x, -1.3795589670628772e-10f);
  r = fmaf(r, x, 1.419491746368351e-09f);
  r = fmaf(r, x, 5.186261198275588e-09f);
  r = fmaf(r, x, -3.8145707037529064e-08f);
  r = fmaf(r, x, 6.890451231811312e-08f);
  r = fmaf(r, x, -2.4120461716847785e-07f);
  r = fmaf(r, x, 1.846569261942932e-06f);
  r = fmaf(r, x, -9.162412425212096e-06f);
  r = fmaf(r, x, 3.9917653339216486e-05f);
  r = fmaf(r, x, -0.00017699696763884276f);
  r = fmaf(r, x, 0.0007703695446252823f);
  r = fmaf(r, x, -0.003233563620597124f);
  r = fmaf(r, x, 0.013096202164888382f);
  r = fmaf(r, x, -0.051039982587099075f);
  r = fmaf(r, x, 0.1905488669872284f);
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
