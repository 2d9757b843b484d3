// gmle_rng.hpp -- counter-based random numbers and the samplers of the
// on-device simulator.  Host- and device-callable.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): a keyed bijection of a 128-bit counter, so the numbers
// of every (ROI, pixel, draw) are fixed by the seed alone, whatever the launch
// shape or the order in which threads run.
//
// Samplers:
//   Poisson, mean < 10: multiplication of uniforms against exp(-mean) (Knuth,
//     TAOCP vol. 2, 3.4.1);
//   Poisson, mean >= 10: transformed rejection with squeeze, PTRS (Hoermann,
//     "The transformed rejection method for generating Poisson random
//     variables", Insurance: Mathematics and Economics 12, 1993);
//   normal: Box-Muller.
// Arithmetic is Float64 (the generator is not a hot path).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace gmle {

// Bound on sampler loops: every thread finishes, whatever the stream.
constexpr int kMaxDraws = 1000;

struct U32x4 {
  uint32_t v[4];
};

__host__ __device__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t& lo, uint32_t& hi) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  lo = (uint32_t)p;
  hi = (uint32_t)(p >> 32);
}

__host__ __device__ inline U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;   // round multipliers
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;   // key schedule (Weyl) increments
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0, hi0, lo1, hi1;
    mulhilo32(kM0, c.v[0], lo0, hi0);
    mulhilo32(kM1, c.v[2], lo1, hi1);
    const U32x4 n{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
    c = n;
    k0 += kW0;
    k1 += kW1;
  }
  return c;
}

// A stream of uniforms for one (ROI, pixel, purpose): counter words are
// (pixel, ROI low, ROI high, purpose << 24 | block), four uniforms per block.
struct Stream {
  uint32_t k0, k1;
  uint32_t c0, c1, c2, c3;
  uint32_t block = 0;
  int used = 4;
  U32x4 buf{};

  __host__ __device__ Stream(uint64_t seed, int64_t roi, uint32_t pixel, uint32_t purpose)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), c0(pixel), c1((uint32_t)(uint64_t)roi),
        c2((uint32_t)((uint64_t)roi >> 32)), c3(purpose << 24) {}

  // Uniform on (0, 1): 32 random bits, centred in their interval.
  __host__ __device__ double uniform() {
    if (used == 4) {
      buf = philox4x32_10(U32x4{{c0, c1, c2, c3 | block}}, k0, k1);
      ++block;
      used = 0;
    }
    return ((double)buf.v[used++] + 0.5) * (1.0 / 4294967296.0);
  }
};

__host__ __device__ inline double sample_normal(Stream& s) {
  const double u1 = s.uniform(), u2 = s.uniform();
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__host__ __device__ inline double sample_poisson(Stream& s, double mean) {
  if (!(mean > 0.0)) return 0.0;
  if (mean < 10.0) {
    const double limit = exp(-mean);
    double p = 1.0;
    int k = -1;
    do {
      ++k;
      p *= s.uniform();
    } while (p > limit && k < kMaxDraws);   // P(k >= kMaxDraws) is nil for mean < 10
    return (double)k;
  }
  const double slam = sqrt(mean), loglam = log(mean);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int attempt = 0; attempt < kMaxDraws; ++attempt) {   // each attempt accepts with probability > 0.9
    const double U = s.uniform() - 0.5;
    const double V = s.uniform();
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + mean + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -mean + k * loglam - lgamma(k + 1.0)) return k;
  }
  return floor(mean + 0.5);   // unreachable in practice; bounds the loop
}

}  // namespace gmle
