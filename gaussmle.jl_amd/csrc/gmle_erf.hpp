// gmle_erf.hpp -- Float32 erf that reuses exp(-u^2).
//
// The separable model needs erf and the Gaussian exp(-u^2) at every pixel edge
// (u = t / (sigma sqrt 2)); the Gaussian is already computed for the derivative
// tables, so erf is evaluated without a transcendental of its own:
//   |u| <  kErfT : erf(u) = c0 u + u s Q(s),        s = u^2, c0 = 2/sqrt(pi)
//   |u| >= kErfT : erf(u) = 1 - exp(-u^2) R(|u| - kErfC),   R ~ erfcx(u) = exp(u^2) erfc(u)
//   |u| >= kErfUMax : erf(u) = 1 (Float32 erf rounds to 1 there)
// Both pieces are always evaluated and one is selected (no divergent branch).
// Q and R are near-minimax polynomials fitted by gen_erf_coeffs.py (Lawson-
// reweighted least squares on Chebyshev nodes).  Measured Float32 FMA-Horner
// error: 1.5e-7 relative on the small piece, 1.6e-7 relative on erfcx (at most
// 5e-8 absolute on erf).  Callable from host code too, for testing.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace gmle {

constexpr float kErfT = 0.875f;
constexpr float kErfUMax = 3.92f;
constexpr float kErfC = 2.3975000381469727f;
constexpr float kErfC0 = 1.128379225730896f;
constexpr float kErfLog2e = 1.4426950408889634f;

__host__ __device__ __forceinline__ float erf_with_gauss(float u, float g) {
  const float a = fabsf(u);
  const float s = a * a;
  // small piece
  float q = 0.00016622350085526705f;
  q = fmaf(q, s, -0.0009764697169885039f);
  q = fmaf(q, s, 0.005313991568982601f);
  q = fmaf(q, s, -0.02689233049750328f);
  q = fmaf(q, s, 0.11284056305885315f);
  q = fmaf(q, s, -0.3761264383792877f);
  const float small = fmaf(a, kErfC0, (a * s) * q);
  // large piece
  const float x = fminf(a, kErfUMax) - kErfC;
  float r = 1.405045821911699e-07f;
  r = fmaf(r, x, -5.948114676357363e-07f);
  r = fmaf(r, x, 1.3180316500438494e-06f);
  r = fmaf(r, x, -4.802132934855763e-06f);
  r = fmaf(r, x, 2.0551316993078217e-05f);
  r = fmaf(r, x, -7.470209675375372e-05f);
  r = fmaf(r, x, 0.0002583200694061816f);
  r = fmaf(r, x, -0.0008767007384449244f);
  r = fmaf(r, x, 0.002878588391467929f);
  r = fmaf(r, x, -0.009095106273889542f);
  r = fmaf(r, x, 0.02756226807832718f);
  r = fmaf(r, x, -0.07972286641597748f);
  r = fmaf(r, x, 0.21869787573814392f);
  float large = fmaf(-g, r, 1.0f);
  large = a >= kErfUMax ? 1.0f : large;
  const float v = a < kErfT ? small : large;
  return copysignf(v, u);
}

}  // namespace gmle
