// gmle_erf.hpp -- Float32 erf that reuses exp(-u^2).
//
// The separable model needs erf and the Gaussian exp(-u^2) at every pixel edge
// (u = t / (sigma sqrt 2)); the Gaussian is already computed for the derivative
// tables, so erf is evaluated without a transcendental of its own, as one
// piece over the whole range:
//   erf(u) = sign(u) (1 - exp(-u^2) R(min(|u|, kErfUMax) - kErfC)),  R ~ erfcx
// (erfcx(u) = exp(u^2) erfc(u)).  Beyond kErfUMax the product is below 3e-8
// and the result rounds to 1 by itself.  R is a near-minimax polynomial of
// degree 15 fitted by gen_erf_coeffs.py (Lawson-reweighted least squares on
// Chebyshev nodes).  Measured Float32 FMA-Horner error: 3.0e-7 relative on
// erfcx, 2.2e-7 absolute on erf; pixel integrals (erf(b) - erf(a)) / 2 of one
// pixel are within 1.2e-7 absolute, 7.4e-7 relative for the central ones.
// Near u = 0 the error is absolute, not relative, which is what the pixel
// integrals consume.  Callable from host code too, for testing.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace gmle {

constexpr float kErfUMax = 3.92f;
constexpr float kErfC = 1.9600000381469727f;
constexpr float kErfLog2e = 1.4426950408889634f;

__host__ __device__ __forceinline__ float erf_with_gauss(float u, float g) {
  const float x = fminf(fabsf(u), kErfUMax) - kErfC;
  float r = -1.6134004354739773e-08f;
  r = fmaf(r, x, 7.332921825309313e-08f);
  r = fmaf(r, x, -5.116217138834145e-08f);
  r = fmaf(r, x, 6.4967665025506e-08f);
  r = fmaf(r, x, -1.5845183725105016e-06f);
  r = fmaf(r, x, 6.3585248426534235e-06f);
  r = fmaf(r, x, -1.8177152014686726e-05f);
  r = fmaf(r, x, 5.988409975543618e-05f);
  r = fmaf(r, x, -0.0002034423960139975f);
  r = fmaf(r, x, 0.0006487831706181169f);
  r = fmaf(r, x, -0.0019805324263870716f);
  r = fmaf(r, x, 0.005820787977427244f);
  r = fmaf(r, x, -0.016361111775040627f);
  r = fmaf(r, x, 0.04371161758899689f);
  r = fmaf(r, x, -0.11021632701158524f);
  r = fmaf(r, x, 0.2597354054450989f);
  return copysignf(fmaf(-g, r, 1.0f), u);
}

}  // namespace gmle
