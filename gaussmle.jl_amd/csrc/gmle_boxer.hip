// gmle_boxer.hip -- candidate boxes cut from camera frames on the device.
//
// The step before the fit in GaussMLE.jl's pipeline is SMLMBoxer.getboxes
// (examples/full_pipeline_boxer.jl:70-80, examples/full_pipeline.jl:73-81):
// a difference-of-Gaussians filter, local maxima above `minval`, boxes of
// `boxsize` pixels around them that overlap by at most `overlap` pixels, cut
// from the raw frames into a ROIBatch.  Here every step runs on the device:
//
//   D = G(sigma_small) * I - G(sigma_large) * I
//     separable, replicated edges; 1-D weights exp(-k^2 / (2 sigma^2)) for
//     |k| <= floor(3 sigma + 0.5), normalised to sum 1 (computed on the host in
//     Float64: sum = g0 + 2 g1 + 2 g2 + ..., then each weight rounded to
//     Float32); each pass is v = w0 x0 + sum_k w_k (x_-k + x_k), accumulated
//     by fma in increasing k; columns (x) first, then rows (y).
//   detection: D(p) > minval and no pixel q of the frame within Chebyshev
//     distance r of p beats it, where q beats p if D(q) > D(p) or D(q) == D(p)
//     and q comes first in column-major order; r = ceil(boxsize - overlap) - 1
//     (at least 1), so kept detections are at least boxsize - overlap apart.
//   box: top-left corner (row, col) = detection - (boxsize - 1) / 2, clamped
//     into the frame; the ROI holds the raw input pixels.
//   order: by frame, then column-major pixel index (Julia's findall order).
//
// Kernels (all deterministic; no atomics):
//   1. boxer_dog: one tile of 62 rows x 30 columns per block, staged through
//      LDS with a halo of R + 1; each pass runs over register strips of 8
//      outputs (one LDS read per input value and strip); writes D and, per
//      tile column, a ballot word of the 3x3 maxima above minval (a necessary
//      condition).
//   2. boxer_suppress: one candidate word per lane; a wave tests its lanes'
//      candidates one at a time, reading each window with all 64 lanes;
//      per-block survivor counts.
//   3. boxer_scan: exclusive scan of the block counts; total detections.
//   4. boxer_place: each survivor's rank from the scan -> corners and frame.
//   5. boxer_cut: one wave per box copies its pixels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/gaussmle.h"
#include "gmle_args.hpp"
#include "gmle_host.hpp"

namespace gmle {
namespace {

constexpr int kDRows = 64;         // filtered region per tile: rows (the contiguous axis) ...
constexpr int kDCols = 32;         // ... and columns, with a 1-pixel halo around the interior
constexpr int kWordRows = kDRows - 2;   // 62 interior rows: one candidate word per tile column
constexpr int kInnerCols = kDCols - 2;  // 30 interior columns
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxRadius = 12;     // weights radius limit: sigma <= 4 px
constexpr int kStrip = 8;          // outputs per thread along the filter axis
constexpr int kScanThreads = 1024;
constexpr int kCutBlocks = 4096;

struct BoxerArgs {
  const float* frames;   // (ny, nx, nf) column-major
  int ny, nx, nf;
  int rs, rl;            // weight radii of the two kernels
  int halo;              // max(rs, rl)
  float ws[kMaxRadius + 1];
  float wl[kMaxRadius + 1];
  float minval;
  int win;               // suppression radius r
  int box;
  int wpc;               // candidate words per column: ceil(ny / 62)
  int64_t n_words;       // nf * nx * wpc
  int64_t max_boxes;
  float* dog;            // (ny, nx, nf)
  uint64_t* cand;        // (wpc, nx, nf); bit b of word w <-> row 62 w + b
  uint64_t* keep;        // (wpc, nx, nf)
  int64_t* block_base;   // per suppress / place block
  int64_t* count;        // total detections
  float* rois;           // (box, box, max_boxes)
  int32_t* xc;
  int32_t* yc;
  int32_t* frame;
  float* strength;       // D at the detection (may be null)
};

// LDS layout of a tile for radii up to RMAX.  Rows and columns are stored
// with an offset of RMAX, so every strip reads the same 8 + 2 RMAX values
// whatever the actual radius (the values beyond it are never used), and the
// row stride is odd, so lanes reading one value per column hit different banks.
template <int RMAX>
struct TileLayout {
  static constexpr int kRows = kDRows + 2 * RMAX;         // stored rows of `in`, `hs`, `hl`
  static constexpr int kStride = kRows | 1;
  static constexpr int kInCols = kDCols + 2 * RMAX;
  static constexpr int kDStride = kDRows | 1;
  static constexpr int kWeights = 2 * (kMaxRadius + 1);
  static constexpr int kFloats = kWeights + kInCols * kStride + 2 * kDCols * kStride;
  static constexpr int kStripLen = kStrip + 2 * RMAX;
};

// One Gaussian over a register strip: out[o] = w0 x[c+o] + sum_k w_k
// (x[c+o-k] + x[c+o+k]), fma in increasing k, c = RMAX.  R is uniform.
template <int RMAX>
__device__ __forceinline__ void strip_filter(const float (&x)[kStrip + 2 * RMAX], const float* w, int R,
                                             float (&out)[kStrip]) {
#pragma clang fp contract(off)
  const float w0 = w[0];
#pragma unroll
  for (int o = 0; o < kStrip; ++o) out[o] = w0 * x[RMAX + o];
#pragma unroll
  for (int k = 1; k <= RMAX; ++k) {
    if (k <= R) {
      const float wk = w[k];
#pragma unroll
      for (int o = 0; o < kStrip; ++o) out[o] = fmaf(wk, x[RMAX + o - k] + x[RMAX + o + k], out[o]);
    }
  }
}

template <int RMAX>
__device__ __forceinline__ void strip_load(const float* src, int step, float (&x)[kStrip + 2 * RMAX]) {
#pragma unroll
  for (int m = 0; m < kStrip + 2 * RMAX; ++m) x[m] = src[m * step];
}

// Tile (blockIdx.x, blockIdx.y) of frame blockIdx.z: interior rows
// [62 by, 62 by + 62), columns [30 bx, 30 bx + 30); the filtered region adds
// one pixel around it, the input region R more.
template <int RMAX>
__global__ __launch_bounds__(kBlock) void boxer_dog(BoxerArgs a) {
#pragma clang fp contract(off)
  using T = TileLayout<RMAX>;
  __shared__ float lds[T::kFloats];
  const int R = a.halo;
  const int f = blockIdx.z;
  const int r0 = blockIdx.y * kWordRows - 1, c0 = blockIdx.x * kInnerCols - 1;   // filtered region origin
  constexpr int S = T::kStride;
  float* wgt = lds;
  float* in = wgt + T::kWeights;                 // [kInCols][S]: column c, row r <-> (c0 + c - RMAX, r0 + r - RMAX)
  float* hs = in + T::kInCols * S;               // [kDCols][S]: column c, row r <-> (c0 + c, r0 + r - RMAX)
  float* hl = hs + kDCols * S;
  float* d = in;                                 // [kDCols][kDStride], reuses `in` after the column pass
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x <= kMaxRadius) {
    wgt[threadIdx.x] = a.ws[threadIdx.x];
    wgt[kMaxRadius + 1 + threadIdx.x] = a.wl[threadIdx.x];
  }
  const float* ws = wgt;
  const float* wl = wgt + kMaxRadius + 1;
  const size_t plane = (size_t)a.ny * a.nx;
  const float* src = a.frames + (size_t)f * plane;
  const int lo = RMAX - R, hi_r = RMAX + kDRows + R;
  // stage the input region (radius RMAX) with replicated edges: every load of
  // a thread is issued before the first LDS write, so they overlap in flight
  {
    constexpr int kIn = T::kInCols * T::kRows;
    constexpr int kPer = (kIn + kBlock - 1) / kBlock;
    float v[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = threadIdx.x + i * kBlock;
      const int c = e / T::kRows, r = e - c * T::kRows;
      if (i + 1 < kPer || e < kIn)
        v[i] = src[(size_t)min(max(c0 + c - RMAX, 0), a.nx - 1) * a.ny + min(max(r0 + r - RMAX, 0), a.ny - 1)];
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = threadIdx.x + i * kBlock;
      const int c = e / T::kRows, r = e - c * T::kRows;
      if (i + 1 < kPer || e < kIn) in[c * S + r] = v[i];
    }
  }
  __syncthreads();
  // along x, over rows [lo, hi_r) and 4 strips of 8 columns: first rows
  // lo..lo+63 (each wave one strip, each lane one row), then the 2R remaining
  // rows x 4 strips packed onto the fewest lanes (threads 0 .. 8R-1), so the
  // other waves leave their issue slots to other blocks
  {
    auto x_pass = [&](int r, int cb) {
      float x[T::kStripLen], vs[kStrip], vl[kStrip];
      strip_load<RMAX>(in + cb * S + r, S, x);
      strip_filter<RMAX>(x, ws, a.rs, vs);
      strip_filter<RMAX>(x, wl, a.rl, vl);
#pragma unroll
      for (int o = 0; o < kStrip; ++o) {
        hs[(cb + o) * S + r] = vs[o];
        hl[(cb + o) * S + r] = vl[o];
      }
    };
    x_pass(lo + lane, wave * kStrip);
    const int t = threadIdx.x;
    if (t < 4 * (hi_r - lo - 64)) x_pass(lo + 64 + (t >> 2), (t & 3) * kStrip);
  }
  __syncthreads();
  // along y: each thread takes 8 rows of one column (lanes across columns)
  {
    const int c = lane & 31, rb = (wave * 2 + (lane >> 5)) * kStrip;
    float x[T::kStripLen], vs[kStrip], vl[kStrip];
    strip_load<RMAX>(hs + c * S + rb, 1, x);
    strip_filter<RMAX>(x, ws, a.rs, vs);
    strip_load<RMAX>(hl + c * S + rb, 1, x);
    strip_filter<RMAX>(x, wl, a.rl, vl);
    // `d` overwrites `in`, which nothing reads after the column pass
#pragma unroll
    for (int o = 0; o < kStrip; ++o) d[c * T::kDStride + rb + o] = vs[o] - vl[o];
  }
  __syncthreads();
  // one wave per interior column, one lane per region row (rows 1..62 interior)
  constexpr int DS = T::kDStride;
  float* dst = a.dog + (size_t)f * plane;
  for (int cc = 1 + wave; cc <= kInnerCols; cc += kWaves) {
    const int gc = c0 + cc, gr = r0 + lane;
    const bool interior = lane >= 1 && lane <= kWordRows && gc < a.nx && gr < a.ny;
    bool cand = false;
    if (interior) {
      const float v = d[cc * DS + lane];
      dst[(size_t)gc * a.ny + gr] = v;
      if (v > a.minval) {
        cand = true;
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc) {
          if (gc + dc < 0 || gc + dc >= a.nx) continue;
#pragma unroll
          for (int dr = -1; dr <= 1; ++dr) {
            if (gr + dr < 0 || gr + dr >= a.ny || (dc == 0 && dr == 0)) continue;
            const float q = d[(cc + dc) * DS + lane + dr];
            const bool first = dc < 0 || (dc == 0 && dr < 0);
            if (q > v || (q == v && first)) cand = false;
          }
        }
      }
    }
    const uint64_t word = __ballot(cand) >> 1;
    if (lane == 0 && gc < a.nx) a.cand[((size_t)f * a.nx + gc) * a.wpc + blockIdx.y] = word;
  }
}

// Block-wide exclusive scan of one int64 per thread (inclusive scans within
// each wave by shuffles, then the wave totals through LDS); returns the prefix
// and leaves the block total in *total.  sh holds at least blockDim.x / 64.
__device__ int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int64_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t up = __shfl_up(incl, off);
    if (lane >= off) incl += up;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  int64_t before = 0, all = 0;
  for (int q = 0; q < nw; ++q) {
    const int64_t t = sh[q];
    before += q < wave ? t : 0;
    all += t;
  }
  *total = all;
  __syncthreads();
  return before + incl - v;
}

// One word per lane; the wave tests its lanes' candidates one at a time,
// all 64 lanes reading the candidate's window together.
__global__ __launch_bounds__(kBlock) void boxer_suppress(BoxerArgs a) {
  __shared__ int64_t sh[kBlock];
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t m0 = w < a.n_words ? a.cand[w] : 0;
  uint64_t todo = m0, keep = m0;
  const int side = 2 * a.win + 1, area = side * side;
  const int step_c = 64 / side, step_r = 64 - step_c * side;
  for (;;) {
    const uint64_t busy = __ballot(todo != 0);
    if (!busy) break;
    const int src = __builtin_ctzll(busy);
    // the candidate: lowest pending bit of lane `src`
    const uint64_t ms = __shfl(todo, src);
    const int64_t ws = __shfl(w, src);
    const int b = __builtin_ctzll(ms);
    const int wr = (int)(ws % a.wpc);
    const int64_t fc = ws / a.wpc;
    const int gc = (int)(fc % a.nx), f = (int)(fc / a.nx);
    const int gr = wr * kWordRows + b;
    const float* D = a.dog + (size_t)f * a.ny * a.nx;
    const float v = D[(size_t)gc * a.ny + gr];
    bool beaten = false;
    // window element e = lane + 64 i <-> (column e / side, row e % side), stepped
    int ec = lane / side, er = lane - ec * side;
    for (int e = lane; e < area; e += 64) {
      const int qc = gc - a.win + ec, qr = gr - a.win + er;
      if (qc >= 0 && qc < a.nx && qr >= 0 && qr < a.ny) {
        const float q = D[(size_t)qc * a.ny + qr];
        const bool first = qc < gc || (qc == gc && qr < gr);
        beaten |= q > v || (q == v && first);
      }
      ec += step_c;
      er += step_r;
      if (er >= side) {
        er -= side;
        ++ec;
      }
    }
    const bool lost = __ballot(beaten) != 0;
    if (lane == src) {
      todo &= todo - 1;
      if (lost) keep &= ~(1ull << b);
    }
  }
  if (w < a.n_words) a.keep[w] = keep;
  int64_t total;
  block_exclusive_scan(__popcll(keep), sh, &total);
  if (threadIdx.x == 0) a.block_base[blockIdx.x] = total;
}

// Exclusive scan of the per-block counts in place (one block); the total
// number of detections goes to *count.
__global__ __launch_bounds__(kScanThreads) void boxer_scan(int64_t* base, int64_t nb, int64_t* count) {
  __shared__ int64_t sh[kScanThreads];
  const int t = threadIdx.x;
  const int64_t per = (nb + kScanThreads - 1) / kScanThreads;
  const int64_t lo = min(nb, (int64_t)t * per), hi = min(nb, lo + per);
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += base[i];
  int64_t total;
  int64_t run = block_exclusive_scan(s, sh, &total);
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t v = base[i];
    base[i] = run;
    run += v;
  }
  if (t == 0) *count = total;
}

__global__ __launch_bounds__(kBlock) void boxer_place(BoxerArgs a) {
  __shared__ int64_t sh[kBlock];
  const int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  uint64_t keep = w < a.n_words ? a.keep[w] : 0;
  int64_t total;
  int64_t k = a.block_base[blockIdx.x] + block_exclusive_scan(__popcll(keep), sh, &total);
  if (!keep) return;
  const int wr = (int)(w % a.wpc);
  const int64_t fc = w / a.wpc;
  const int gc = (int)(fc % a.nx), f = (int)(fc / a.nx);
  const int half = (a.box - 1) / 2;
  const int c0 = min(max(gc - half, 0), a.nx - a.box);
  while (keep && k < a.max_boxes) {
    const int b = __builtin_ctzll(keep);
    keep &= keep - 1;
    const int gr = wr * kWordRows + b;
    const int r0 = min(max(gr - half, 0), a.ny - a.box);
    a.xc[k] = c0 + 1;
    a.yc[k] = r0 + 1;
    a.frame[k] = f + 1;
    if (a.strength) a.strength[k] = a.dog[((size_t)f * a.nx + gc) * a.ny + gr];
    ++k;
  }
}

__global__ __launch_bounds__(kBlock) void boxer_cut(BoxerArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t n = min(*a.count, a.max_boxes);
  const int B = a.box, BB = a.box * a.box;
  const size_t plane = (size_t)a.ny * a.nx;
  for (int64_t k = (int64_t)blockIdx.x * kWaves + wave; k < n; k += (int64_t)gridDim.x * kWaves) {
    const float* src = a.frames + (size_t)(a.frame[k] - 1) * plane + (size_t)(a.xc[k] - 1) * a.ny + (a.yc[k] - 1);
    float* dst = a.rois + (size_t)k * BB;
    for (int e = lane; e < BB; e += 64) {
      const int col = e / B, row = e - col * B;
      dst[e] = src[(size_t)col * a.ny + row];
    }
  }
}

int radius_of(double sigma) { return (int)floor(3.0 * sigma + 0.5); }

// Half of a normalised Gaussian kernel, w[0..R], in the documented order.
void gaussian_half(double sigma, int R, float* w) {
  double g[kMaxRadius + 1];
  double sum = 0.0;
  for (int k = 0; k <= R; ++k) {
    g[k] = exp(-(double)k * (double)k / (2.0 * sigma * sigma));
    sum += k == 0 ? g[k] : 2.0 * g[k];
  }
  for (int k = 0; k <= R; ++k) w[k] = (float)(g[k] / sum);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Layout {
  size_t dog, cand, keep, base, total;
  int wpc;
  int64_t n_words, n_blocks;
};

Layout layout(int ny, int nx, int nf) {
  Layout L;
  L.wpc = (ny + kWordRows - 1) / kWordRows;
  L.n_words = (int64_t)nf * nx * L.wpc;
  L.n_blocks = (L.n_words + kBlock - 1) / kBlock;
  L.dog = 0;
  L.cand = L.dog + align256((size_t)ny * nx * nf * sizeof(float));
  L.keep = L.cand + align256((size_t)L.n_words * sizeof(uint64_t));
  L.base = L.keep + align256((size_t)L.n_words * sizeof(uint64_t));
  L.total = L.base + align256((size_t)L.n_blocks * sizeof(int64_t));
  return L;
}

int check_params(int ny, int nx, int nf, int box, float overlap, float sigma_small, float sigma_large, float minval) {
  if (nf <= 0 || ny <= 0 || nx <= 0) return fail(GMLE_ERR_ARG, "frame stack %d x %d x %d is empty", ny, nx, nf);
  if (nf > 65535) return fail(GMLE_ERR_ARG, "at most 65535 frames per call, got %d", nf);
  if (box < kMinBox || box > kMaxBatchBox) return fail(GMLE_ERR_ARG, "boxsize %d out of range [%d, %d]", box, kMinBox, kMaxBatchBox);
  if (ny < box || nx < box) return fail(GMLE_ERR_ARG, "frames of %d x %d are smaller than the box %d", ny, nx, box);
  if (!(overlap >= 0.0f) || !(overlap < (float)box)) return fail(GMLE_ERR_ARG, "overlap must be in [0, boxsize)");
  if (!(sigma_small > 0.0f) || !(sigma_large > 0.0f)) return fail(GMLE_ERR_ARG, "sigmas must be > 0");
  if (radius_of(sigma_small) > kMaxRadius || radius_of(sigma_large) > kMaxRadius)
    return fail(GMLE_ERR_ARG, "sigmas above 4 px are not supported (kernel radius > %d)", kMaxRadius);
  if (isnan(minval)) return fail(GMLE_ERR_ARG, "minval is NaN");
  return GMLE_OK;
}

BoxerArgs make_args(const float* frames, int ny, int nx, int nf, int box, float overlap, float sigma_small,
                    float sigma_large, float minval, int64_t max_boxes, void* ws) {
  BoxerArgs a{};
  a.frames = frames;
  a.ny = ny;
  a.nx = nx;
  a.nf = nf;
  a.rs = radius_of(sigma_small);
  a.rl = radius_of(sigma_large);
  a.halo = std::max(a.rs, a.rl);
  gaussian_half(sigma_small, a.rs, a.ws);
  gaussian_half(sigma_large, a.rl, a.wl);
  a.minval = minval;
  a.win = std::max(1, (int)ceilf((float)box - overlap) - 1);
  a.box = box;
  const Layout L = layout(ny, nx, nf);
  a.wpc = L.wpc;
  a.n_words = L.n_words;
  a.max_boxes = max_boxes;
  char* w = (char*)ws;
  a.dog = (float*)(w + L.dog);
  a.cand = (uint64_t*)(w + L.cand);
  a.keep = (uint64_t*)(w + L.keep);
  a.block_base = (int64_t*)(w + L.base);
  return a;
}


int enqueue_boxer(BoxerArgs a, hipStream_t stream) {
  const Layout L = layout(a.ny, a.nx, a.nf);
  const dim3 grid((a.nx + kInnerCols - 1) / kInnerCols, L.wpc, a.nf);
  if (a.halo <= 6)
    hipLaunchKernelGGL(boxer_dog<6>, grid, dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL(boxer_dog<kMaxRadius>, grid, dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(boxer_suppress, dim3((unsigned)L.n_blocks), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(boxer_scan, dim3(1), dim3(kScanThreads), 0, stream, a.block_base, L.n_blocks, a.count);
  hipLaunchKernelGGL(boxer_place, dim3((unsigned)L.n_blocks), dim3(kBlock), 0, stream, a);
  const int64_t cut = std::min<int64_t>(kCutBlocks, std::max<int64_t>(1, (a.max_boxes + kWaves - 1) / kWaves));
  if (a.max_boxes > 0) hipLaunchKernelGGL(boxer_cut, dim3((unsigned)cut), dim3(kBlock), 0, stream, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? GMLE_OK : hip_fail(e, "kernel launch");
}

}  // namespace
}  // namespace gmle

using namespace gmle;

extern "C" {

int64_t gmle_boxer_workspace_bytes(int ny, int nx, int n_frames) {
  if (ny <= 0 || nx <= 0 || n_frames <= 0) return 0;
  return (int64_t)layout(ny, nx, n_frames).total;
}

int gmle_find_boxes_device(const float* d_frames, int ny, int nx, int n_frames, int boxsize, float overlap,
                           float sigma_small, float sigma_large, float minval, int64_t max_boxes, float* d_rois,
                           int32_t* d_x_corners, int32_t* d_y_corners, int32_t* d_frame_indices,
                           float* d_strength, int64_t* d_count, void* d_workspace, void* stream) {
  int rc = check_params(ny, nx, n_frames, boxsize, overlap, sigma_small, sigma_large, minval);
  if (rc != GMLE_OK) return rc;
  if (max_boxes < 0) return fail(GMLE_ERR_ARG, "max_boxes must be >= 0");
  if (!d_frames || !d_count || !d_workspace) return fail(GMLE_ERR_ARG, "frames, count and workspace must not be NULL");
  if (max_boxes > 0 && (!d_rois || !d_x_corners || !d_y_corners || !d_frame_indices))
    return fail(GMLE_ERR_ARG, "output arrays must not be NULL when max_boxes > 0");
  BoxerArgs a = make_args(d_frames, ny, nx, n_frames, boxsize, overlap, sigma_small, sigma_large, minval, max_boxes,
                          d_workspace);
  a.count = d_count;
  a.rois = d_rois;
  a.xc = d_x_corners;
  a.yc = d_y_corners;
  a.frame = d_frame_indices;
  a.strength = d_strength;
  return enqueue_boxer(a, (hipStream_t)stream);
}

int gmle_find_boxes(const float* frames, int ny, int nx, int n_frames, int boxsize, float overlap,
                    float sigma_small, float sigma_large, float minval, int64_t max_boxes, float* rois,
                    int32_t* x_corners, int32_t* y_corners, int32_t* frame_indices, float* strength,
                    int64_t* n_found) {
  int rc = check_params(ny, nx, n_frames, boxsize, overlap, sigma_small, sigma_large, minval);
  if (rc != GMLE_OK) return rc;
  if (max_boxes < 0) return fail(GMLE_ERR_ARG, "max_boxes must be >= 0");
  if (!frames || !n_found) return fail(GMLE_ERR_ARG, "frames and n_found must not be NULL");
  if (max_boxes > 0 && (!rois || !x_corners || !y_corners || !frame_indices))
    return fail(GMLE_ERR_ARG, "output arrays must not be NULL when max_boxes > 0");
  const size_t npix = (size_t)ny * nx * n_frames;
  const size_t BB = (size_t)boxsize * boxsize;
  const size_t ws = layout(ny, nx, n_frames).total;
  // one allocation: frames | workspace | count | outputs
  const size_t o_ws = align256(npix * sizeof(float));
  const size_t o_count = o_ws + align256(ws);
  const size_t o_rois = o_count + 256;
  const size_t o_xc = o_rois + align256((size_t)max_boxes * BB * sizeof(float));
  const size_t o_yc = o_xc + align256((size_t)max_boxes * sizeof(int32_t));
  const size_t o_fr = o_yc + align256((size_t)max_boxes * sizeof(int32_t));
  const size_t o_st = o_fr + align256((size_t)max_boxes * sizeof(int32_t));
  const size_t bytes = o_st + align256((size_t)max_boxes * sizeof(float));
  char* dev = nullptr;
  hipError_t e = hipMalloc((void**)&dev, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  hipStream_t s = nullptr;
  e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    return hip_fail(e, "hipStreamCreate");
  }
  auto done = [&](int code) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    (void)hipFree(dev);
    return code;
  };
  int64_t* d_count = (int64_t*)(dev + o_count);
  if ((e = hipMemcpyAsync(dev, frames, npix * sizeof(float), hipMemcpyHostToDevice, s)) != hipSuccess)
    return done(hip_fail(e, "hipMemcpy frames"));
  BoxerArgs a = make_args((const float*)dev, ny, nx, n_frames, boxsize, overlap, sigma_small, sigma_large, minval,
                          max_boxes, dev + o_ws);
  a.count = d_count;
  a.rois = (float*)(dev + o_rois);
  a.xc = (int32_t*)(dev + o_xc);
  a.yc = (int32_t*)(dev + o_yc);
  a.frame = (int32_t*)(dev + o_fr);
  a.strength = strength ? (float*)(dev + o_st) : nullptr;
  if ((rc = enqueue_boxer(a, s)) != GMLE_OK) return done(rc);
  int64_t n = 0;
  if ((e = hipMemcpyAsync(&n, d_count, sizeof(n), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return done(hip_fail(e, "hipMemcpy count"));
  *n_found = n;
  const size_t m = (size_t)std::min(n, max_boxes);
  if (m > 0) {
    if ((e = hipMemcpyAsync(rois, dev + o_rois, m * BB * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(x_corners, dev + o_xc, m * sizeof(int32_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(y_corners, dev + o_yc, m * sizeof(int32_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(frame_indices, dev + o_fr, m * sizeof(int32_t), hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (strength &&
         (e = hipMemcpyAsync(strength, dev + o_st, m * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess))
      return done(hip_fail(e, "hipMemcpy results"));
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return done(hip_fail(e, "hipStreamSynchronize"));
  return done(GMLE_OK);
}

// Weights the filter uses for one sigma (w[0..R], R = floor(3 sigma + 0.5));
// returns R, or -1 when R is above the supported radius.
int gmle_boxer_weights(float sigma, float* w, int capacity) {
  if (!(sigma > 0.0f)) return -1;
  const int R = radius_of(sigma);
  if (R > kMaxRadius || !w || capacity < R + 1) return -1;
  gaussian_half(sigma, R, w);
  return R;
}

}  // extern "C"
