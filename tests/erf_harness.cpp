// Host build of the kernel's Gaussian pixel-integral pieces (gmle_erf.hpp) for
// the CPU accuracy test.  Inputs are u and g = exp(-u^2); the kernels' scaled
// variable v = k u is formed here as the kernels form it (one Float32 product).
#include "gmle_erf.hpp"

extern "C" void erf_with_gauss_batch(const float* u, const float* g, float* out, long n) {
  for (long k = 0; k < n; ++k) out[k] = gmle::erf_with_gauss(u[k] * gmle::kErfK, g[k]);
}

// sign(u) erfc(|u|) as the kernels carry it.
extern "C" void signed_erfc_batch(const float* u, const float* g, float* out, long n) {
  for (long k = 0; k < n; ++k) {
    float hs;
    gmle::erf_edge(u[k] * gmle::kErfK, g[k], out[k], hs);
  }
}

// (erf(hi) - erf(lo)) / 2 as the kernels form it, from each edge's u and
// g = exp(-u^2).
extern "C" void pixel_integral_batch(const float* u_lo, const float* g_lo, const float* u_hi, const float* g_hi,
                                     float* out, long n) {
  for (long k = 0; k < n; ++k) {
    float se_lo, hs_lo, se_hi, hs_hi;
    gmle::erf_edge(u_lo[k] * gmle::kErfK, g_lo[k], se_lo, hs_lo);
    gmle::erf_edge(u_hi[k] * gmle::kErfK, g_hi[k], se_hi, hs_hi);
    out[k] = gmle::pixel_integral<float>(se_lo, hs_lo, se_hi, hs_hi);
  }
}
