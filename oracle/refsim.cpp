// refsim.cpp -- CPU restatement of the on-device simulator (test
// infrastructure only: the checker for gmle_simulate_device).
//
// GaussMLE.jl's generate_roi_batch (src/simulator.jl:35-345) draws, per pixel,
// Poisson(bg + photons Phi_x Phi_y) for an ideal camera, and for sCMOS
// (Poisson(expected qe) + N(0,1) readnoise) / gain + offset, with Julia's
// global RNG and Distributions.jl's samplers -- neither reproducible here.
// The device simulator replaces the RNG with Philox4x32-10 (Salmon et al.,
// SC'11) counted by (ROI, pixel, purpose) and the samplers with published
// methods: multiplication (Knuth, TAOCP 3.4.1) below mean 10, PTRS (Hoermann
// 1993) above, Box-Muller for the read noise.  This file restates the same
// steps in plain C++; tests pin the generator to its published known-answer
// vectors and the samplers to their distributions (scipy), then require the
// device to reproduce this restatement.
#include <math.h>
#include <stdint.h>

namespace {

void philox(const uint32_t in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Uniforms {
  uint32_t key0, key1, ctr[4];
  uint32_t block = 0, buf[4] = {0, 0, 0, 0};
  int used = 4;
  Uniforms(uint64_t seed, int64_t roi, uint32_t pixel, uint32_t purpose)
      : key0((uint32_t)seed), key1((uint32_t)(seed >> 32)),
        ctr{pixel, (uint32_t)(uint64_t)roi, (uint32_t)((uint64_t)roi >> 32), purpose << 24} {}
  double next() {
    if (used == 4) {
      const uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3] | block};
      philox(c, key0, key1, buf);
      ++block;
      used = 0;
    }
    return ((double)buf[used++] + 0.5) / 4294967296.0;
  }
};

double poisson(Uniforms& u, double mean) {
  if (!(mean > 0.0)) return 0.0;
  if (mean < 10.0) {
    const double limit = exp(-mean);
    double p = 1.0;
    int k = -1;
    do {
      ++k;
      p *= u.next();
    } while (p > limit && k < 1000);
    return k;
  }
  const double slam = sqrt(mean), loglam = log(mean);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int t = 0; t < 1000; ++t) {
    const double U = u.next() - 0.5, V = u.next();
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + mean + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -mean + k * loglam - lgamma(k + 1.0)) return k;
  }
  return floor(mean + 0.5);
}

double normal(Uniforms& u) {
  const double u1 = u.next(), u2 = u.next();
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

float phi(int k, float mu, float sigma) {   // src/gausslib/baselibrary.jl:8-14
  const float s = sqrtf(0.5f / (sigma * sigma));
  return 0.5f * (erff(((float)k - mu + 0.5f) * s) - erff(((float)k - mu - 0.5f) * s));
}

}  // namespace

extern "C" {

void refsim_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out) {
  philox(counter, key[0], key[1], out);
}

// n draws of one sampler from the (roi = 0, pixel = i, purpose) streams: kind
// 0 = Poisson(mean), 1 = normal.
void refsim_draws(int kind, double mean, uint64_t seed, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    Uniforms u(seed, 0, (uint32_t)i, kind == 0 ? 0u : 1u);
    out[i] = kind == 0 ? poisson(u, mean) : normal(u);
  }
}

// The simulator (see gmle_simulate_device); maps may be null (scalars).
void refsim_simulate(int model, const float* psf, int box, int64_t n, const float* params, int stride, int camera,
                     const int32_t* xc, const int32_t* yc, const float* qe_map, const float* rn_map,
                     const float* gain_map, const float* offset_map, int ny, int nx, float qe, float readnoise,
                     float gain, float offset, uint64_t seed, float* out, float* expected_out) {
  for (int64_t k = 0; k < n; ++k) {
    const float* th = params + k * stride;
    float x = th[0], y = th[1], N, bg, sx, sy;
    if (model == 0) { N = th[2]; bg = th[3]; sx = sy = psf[0]; }
    else if (model == 1) { N = th[2]; bg = th[3]; sx = sy = th[4]; }
    else if (model == 2) { N = th[2]; bg = th[3]; sx = th[4]; sy = th[5]; }
    else {
      const float z = th[2];
      N = th[3];
      bg = th[4];
      const float ux = (z - psf[6]) / psf[7], uy = (z + psf[6]) / psf[7];
      sx = psf[0] * sqrtf(1.0f + ux * ux + psf[2] * (ux * ux * ux) + psf[4] * (ux * ux * ux * ux));
      sy = psf[1] * sqrtf(1.0f + uy * uy + psf[3] * (uy * uy * uy) + psf[5] * (uy * uy * uy * uy));
    }
    for (int p = 0; p < box * box; ++p) {
      const int i = p % box + 1, j = p / box + 1;
      const float e = bg + N * phi(j, x, sx) * phi(i, y, sy);
      if (expected_out) expected_out[k * box * box + p] = e;
      Uniforms shot(seed, k, (uint32_t)p, 0u);
      float v;
      if (camera == 0) {
        v = e > 0.0f ? (float)poisson(shot, (double)e) : 0.0f;
      } else {
        const int64_t m = (int64_t)(yc[k] - 1 + i - 1) + (int64_t)ny * (int64_t)(xc[k] - 1 + j - 1);
        const float q = qe_map ? qe_map[m] : qe, rn = rn_map ? rn_map[m] : readnoise;
        const float g = gain_map ? gain_map[m] : gain, o = offset_map ? offset_map[m] : offset;
        const float el = e * q;
        const float sig = el > 0.0f ? (float)poisson(shot, (double)el) : 0.0f;
        Uniforms read(seed, k, (uint32_t)p, 1u);
        v = (sig + (float)normal(read) * rn) / g + o;
      }
      out[k * box * box + p] = v;
    }
  }
}

}  // extern "C"
