// Host build of the kernel's erf (gmle_erf.hpp) for the CPU accuracy test.
#include "gmle_erf.hpp"

extern "C" void erf_with_gauss_batch(const float* u, const float* g, float* out, long n) {
  for (long k = 0; k < n; ++k) out[k] = gmle::erf_with_gauss(u[k], g[k]);
}
