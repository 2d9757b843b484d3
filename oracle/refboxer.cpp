// refboxer.cpp -- CPU restatement of ROI extraction (test infrastructure only:
// the checker for gmle_find_boxes, and the CPU baseline of its benchmark line).
//
// GaussMLE.jl hands fitting a ROIBatch cut by SMLMBoxer.getboxes
// (examples/full_pipeline_boxer.jl:70-80, examples/full_pipeline.jl:73-81).
// SMLMBoxer is a separate package, not vendored and listed without a version
// (examples/Project.toml:9), so its exact conventions cannot be checked here.
// This file restates its published steps -- difference-of-Gaussians filter,
// local maxima above minval, boxes that overlap by at most `overlap`, box cut
// -- with every convention fixed explicitly:
//
//   weights  w_k = exp(-k^2 / (2 sigma^2)) / S for |k| <= R = floor(3 sigma +
//            0.5), sigma a Float32 widened to Float64, S = g0 + 2 g1 + ... + 2 gR
//            summed in that order in Float64,
//            each w_k rounded to Float32 (scipy.ndimage's gaussian_filter1d
//            with truncate = 3 uses the same weights and radius);
//   filter   Float32, replicated edges; one pass v = w0 x0, then
//            v = fma(w_k, x_-k + x_k, v) for k = 1..R; along x (columns) first,
//            then along y (rows); D = small - large;
//   detect   D(p) > minval and no q in the frame with max(|dr|, |dc|) <= r
//            beats p: D(q) > D(p), or D(q) == D(p) and q is earlier in
//            column-major order; r = max(1, ceil(boxsize - overlap) - 1);
//   box      corner = p - (boxsize - 1) / 2 per axis, clamped into the frame;
//   order    frame, then column-major index.
// The plain definitions below (a brute-force window test at every pixel)
// are what the tests pin against scipy / numpy and brute force.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace {

int radius_of(double sigma) { return (int)floor(3.0 * sigma + 0.5); }

std::vector<float> half_kernel(double sigma) {
  const int R = radius_of(sigma);
  std::vector<double> g(R + 1);
  double s = 0.0;
  for (int k = 0; k <= R; ++k) {
    g[k] = exp(-(double)k * (double)k / (2.0 * sigma * sigma));
    s += (k == 0) ? g[k] : 2.0 * g[k];
  }
  std::vector<float> w(R + 1);
  for (int k = 0; k <= R; ++k) w[k] = (float)(g[k] / s);
  return w;
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One separable Gaussian of a frame (ny rows, nx columns, column-major).
void gaussian(const float* in, int ny, int nx, const std::vector<float>& w, float* out) {
  const int R = (int)w.size() - 1;
  std::vector<float> tmp((size_t)ny * nx);
  for (int c = 0; c < nx; ++c)
    for (int r = 0; r < ny; ++r) {
      float v = w[0] * in[(size_t)c * ny + r];
      for (int k = 1; k <= R; ++k) {
        const float a = in[(size_t)clampi(c - k, 0, nx - 1) * ny + r];
        const float b = in[(size_t)clampi(c + k, 0, nx - 1) * ny + r];
        v = fmaf(w[k], a + b, v);
      }
      tmp[(size_t)c * ny + r] = v;
    }
  for (int c = 0; c < nx; ++c)
    for (int r = 0; r < ny; ++r) {
      const float* col = tmp.data() + (size_t)c * ny;
      float v = w[0] * col[r];
      for (int k = 1; k <= R; ++k) v = fmaf(w[k], col[clampi(r - k, 0, ny - 1)] + col[clampi(r + k, 0, ny - 1)], v);
      out[(size_t)c * ny + r] = v;
    }
}

}  // namespace

extern "C" {

int refboxer_weights(float sigma, float* w, int capacity) {
  const std::vector<float> h = half_kernel(sigma);
  if ((int)h.size() > capacity) return -1;
  std::copy(h.begin(), h.end(), w);
  return (int)h.size() - 1;
}

// D for every frame.
void refboxer_dog(const float* frames, int ny, int nx, int nf, float sigma_small, float sigma_large, float* dog) {
  const std::vector<float> ws = half_kernel(sigma_small), wl = half_kernel(sigma_large);
  std::vector<float> a((size_t)ny * nx), b((size_t)ny * nx);
  for (int f = 0; f < nf; ++f) {
    const float* in = frames + (size_t)f * ny * nx;
    gaussian(in, ny, nx, ws, a.data());
    gaussian(in, ny, nx, wl, b.data());
    float* o = dog + (size_t)f * ny * nx;
    for (size_t i = 0; i < (size_t)ny * nx; ++i) o[i] = a[i] - b[i];
  }
}

// Detections and boxes; returns the number of detections (all of them; only
// the first max_boxes are written).  dog_out (optional) receives D.
int64_t refboxer_find(const float* frames, int ny, int nx, int nf, int boxsize, float overlap, float sigma_small,
                      float sigma_large, float minval, int64_t max_boxes, float* rois, int32_t* xc, int32_t* yc,
                      int32_t* frame, float* strength, float* dog_out) {
  std::vector<float> dog((size_t)ny * nx * nf);
  refboxer_dog(frames, ny, nx, nf, sigma_small, sigma_large, dog.data());
  if (dog_out) std::copy(dog.begin(), dog.end(), dog_out);
  const int win = std::max(1, (int)ceilf((float)boxsize - overlap) - 1);
  const int half = (boxsize - 1) / 2;
  int64_t n = 0;
  for (int f = 0; f < nf; ++f) {
    const float* D = dog.data() + (size_t)f * ny * nx;
    for (int c = 0; c < nx; ++c)
      for (int r = 0; r < ny; ++r) {
        const float v = D[(size_t)c * ny + r];
        if (!(v > minval)) continue;
        bool best = true;
        for (int qc = std::max(0, c - win); qc <= std::min(nx - 1, c + win) && best; ++qc)
          for (int qr = std::max(0, r - win); qr <= std::min(ny - 1, r + win); ++qr) {
            const float q = D[(size_t)qc * ny + qr];
            const bool earlier = (int64_t)qc * ny + qr < (int64_t)c * ny + r;
            if (q > v || (q == v && earlier)) { best = false; break; }
          }
        if (!best) continue;
        if (n < max_boxes) {
          const int r0 = clampi(r - half, 0, ny - boxsize), c0 = clampi(c - half, 0, nx - boxsize);
          xc[n] = c0 + 1;
          yc[n] = r0 + 1;
          frame[n] = f + 1;
          if (strength) strength[n] = v;
          float* o = rois + (size_t)n * boxsize * boxsize;
          const float* in = frames + (size_t)f * ny * nx;
          for (int bc = 0; bc < boxsize; ++bc)
            for (int br = 0; br < boxsize; ++br) o[bc * boxsize + br] = in[(size_t)(c0 + bc) * ny + r0 + br];
        }
        ++n;
      }
  }
  return n;
}

}  // extern "C"
