// gmle_api.cpp -- the C ABI declared in include/gaussmle.h.
//
// Argument validation mirrors validate_fit_input (src/interface.jl:10-40) and
// default_constraints (src/constraints.jl:16-47) of GaussMLE.jl.  The host-
// pointer entry point replaces the GPU branch's per-batch allocate / copy /
// launch / synchronise loop (src/interface.jl:220-262): it uploads the whole
// call once, runs one launch per kernel shape and copies results back, keeping
// device buffers cached between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gaussmle.h"
#include "gmle_args.hpp"
#include "gmle_batch.hpp"
#include "gmle_host.hpp"
#include "gmle_hostpipe.hpp"
#include "gmle_multi.hpp"

namespace gmle {

#define GMLE_DECLARE_FINDER(MODEL, CAM, PREC)                             \
  LaunchFn find_launcher_##MODEL##_##CAM##_##PREC(int box, int lanes); \
  DescribeFn find_describe_##MODEL##_##CAM##_##PREC(int box, int lanes);
#define GMLE_DECLARE_ALL(MODEL)               \
  GMLE_DECLARE_FINDER(MODEL, IDEAL, PREC_FP32) \
  GMLE_DECLARE_FINDER(MODEL, IDEAL, PREC_MIXED) \
  GMLE_DECLARE_FINDER(MODEL, SCMOS, PREC_FP32) \
  GMLE_DECLARE_FINDER(MODEL, SCMOS, PREC_MIXED)
GMLE_DECLARE_ALL(XYNB)
GMLE_DECLARE_ALL(XYNBS)
GMLE_DECLARE_ALL(XYNBSXSY)
GMLE_DECLARE_ALL(ASTIG)

LaunchFn find_launcher(int model, int camera, int box, int lanes, int precision) {
#define GMLE_CASE(MODEL)                                                                   \
  case MODEL:                                                                              \
    if (camera == IDEAL)                                                                   \
      return precision == PREC_FP32 ? find_launcher_##MODEL##_IDEAL_PREC_FP32(box, lanes)  \
                                    : find_launcher_##MODEL##_IDEAL_PREC_MIXED(box, lanes); \
    return precision == PREC_FP32 ? find_launcher_##MODEL##_SCMOS_PREC_FP32(box, lanes)    \
                                  : find_launcher_##MODEL##_SCMOS_PREC_MIXED(box, lanes);
  switch (model) {
    GMLE_CASE(XYNB)
    GMLE_CASE(XYNBS)
    GMLE_CASE(XYNBSXSY)
    GMLE_CASE(ASTIG)
  }
#undef GMLE_CASE
  return nullptr;
}

bool find_kernels(int model, int camera, int box, int lanes, int precision, KernelInfo* out) {
  DescribeFn fn = nullptr;
#define GMLE_CASE(MODEL)                                                                   \
  case MODEL:                                                                              \
    if (camera == IDEAL)                                                                   \
      fn = precision == PREC_FP32 ? find_describe_##MODEL##_IDEAL_PREC_FP32(box, lanes)    \
                                  : find_describe_##MODEL##_IDEAL_PREC_MIXED(box, lanes);  \
    else                                                                                   \
      fn = precision == PREC_FP32 ? find_describe_##MODEL##_SCMOS_PREC_FP32(box, lanes)    \
                                  : find_describe_##MODEL##_SCMOS_PREC_MIXED(box, lanes);  \
    break;
  switch (model) {
    GMLE_CASE(XYNB)
    GMLE_CASE(XYNBS)
    GMLE_CASE(XYNBSXSY)
    GMLE_CASE(ASTIG)
  }
#undef GMLE_CASE
  return fn != nullptr && fn(out);
}

namespace {
// The calling thread's last message, in a fixed buffer (recording an error
// allocates nothing).
thread_local char g_err[512];
thread_local size_t g_err_len = 0;
}  // namespace

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int n = vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_err_len = n < 0 ? 0 : std::min<size_t>((size_t)n, sizeof(g_err) - 1);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? GMLE_ERR_NOMEM : GMLE_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

bool lanes_ok(int box, int lanes) { return lanes_built(box, lanes); }

// Lanes per ROI for a large batch: the fewest lanes whose kernel still runs
// two or more waves per SIMD without spilling (the batch autotuner measures).
int large_batch_lanes(int model, int camera, int box) {
  if (box > kMaxSmallBox) return large_box_lanes(model, camera, box);
  return box <= max_box_within_256(model, camera, 1) ? 1 : box <= max_box_within_256(model, camera, 2) ? 2 : 4;
}

// Lanes per ROI when the caller passes 0: the large-batch choice, raised to
// the most lanes that keep the launch within one wave per SIMD.  Below that
// the chip is latency-bound and more, shorter waves finish sooner; above it
// the extra waves share SIMDs and the replicated per-ROI work costs more than
// it saves.  Measured on XYNB 11 px and sCMOS XYNBS 13 px: 1K and 4K ROIs run
// fastest at 16 lanes, 10K at 4, 30K at 2.
int default_lanes(int model, int camera, int box, int precision, int64_t n_rois) {
  (void)precision;
  const int64_t fill = 256 * 4 * 64;   // CUs x SIMDs x lanes: one wave per SIMD
  int lanes = large_batch_lanes(model, camera, box);
  if (!lanes_ok(box, lanes)) lanes = box > kMaxSmallBox ? 16 : 2;
  for (int more = 2; more <= 16; more *= 2)
    if (lanes < more && n_rois * more <= fill && lanes_ok(box, more)) lanes = more;
  return lanes;
}

// Shared argument checks of both entry points.
int check_common(int model, const float* psf, int box, int64_t n_rois, int camera, int ny, int nx,
                 const float* lower, const float* upper, const float* max_step, int iterations, int precision,
                 int* lanes) {
  if (model < XYNB || model > ASTIG) return fail(GMLE_ERR_ARG, "unknown PSF model %d", model);
  if (camera != IDEAL && camera != SCMOS) return fail(GMLE_ERR_ARG, "unknown camera model %d", camera);
  if (precision != PREC_FP32 && precision != PREC_MIXED)
    return fail(GMLE_ERR_ARG, "unknown precision %d", precision);
  if (n_rois <= 0) return fail(GMLE_ERR_ARG, "Input data array is empty");
  if (box < kMinBox) return fail(GMLE_ERR_ARG, "ROI size must be at least 3x3, got %dx%d", box, box);
  if (box > kMaxBox) return fail(GMLE_ERR_UNSUPPORTED, "ROI size %d exceeds the supported maximum %d", box, kMaxBox);
  if (iterations < 0) return fail(GMLE_ERR_ARG, "iterations must be >= 0, got %d", iterations);
  if (!lower || !upper || !max_step) return fail(GMLE_ERR_ARG, "constraints arrays must not be NULL");
  if ((model == XYNB || model == ASTIG) && !psf) return fail(GMLE_ERR_ARG, "psf parameters must not be NULL");
  if (model == XYNB && !(psf[0] > 0.f)) return fail(GMLE_ERR_ARG, "PSF sigma must be > 0");
  if (model == ASTIG && !(psf[0] > 0.f && psf[1] > 0.f && psf[7] != 0.f && psf[6] != 0.f))
    return fail(GMLE_ERR_ARG, "astigmatic PSF needs sigma_x0, sigma_y0 > 0 and nonzero gamma, d");
  if (camera == SCMOS && (ny < box || nx < box))
    return fail(GMLE_ERR_ARG, "variance map %dx%d is smaller than one ROI", ny, nx);
  if (*lanes == 0) *lanes = default_lanes(model, camera, box, precision, n_rois);
  if (!lanes_ok(box, *lanes))
    return fail(GMLE_ERR_UNSUPPORTED, "no kernel with %d lanes per ROI for box size %d", *lanes, box);
  return GMLE_OK;
}

void fill_args(FitArgs& a, int model, const float* psf, const float* lower, const float* upper,
               const float* max_step, int64_t n_rois, int ny, int nx, int iterations) {
  std::memset(&a, 0, sizeof(a));
  const int P = n_params(model);
  for (int q = 0; q < 8; ++q) a.psf[q] = 0.f;
  if (model == XYNB) a.psf[0] = psf[0];
  if (model == ASTIG)
    for (int q = 0; q < 8; ++q) a.psf[q] = psf[q];
  for (int k = 0; k < P; ++k) {
    a.lower[k] = lower[k];
    a.upper[k] = upper[k];
    a.max_step[k] = max_step[k];
  }
  a.n_rois = n_rois;
  a.ny = ny;
  a.nx = nx;
  a.iterations = iterations;
}

namespace {

// Device buffers cached across host-pointer calls (grow only), one set per
// device, each with its own mutex and stream.
// Slots: 0 data, 1 params, 2 crlb, 3 llr, 4 var_map, 5 corners, 6 crlb64,
// 7 llr64, 8 models, 9 boxes, 10 offsets, 11 batch workspace.
constexpr int kCacheSlots = 12;
constexpr int kMaxDevices = 64;
struct Cache {
  std::mutex mu;
  void* buf[kCacheSlots] = {};
  size_t cap[kCacheSlots] = {};
  hipStream_t stream = nullptr;
  HostPipe pipe;   // pinned upload ring of the host-array entry points
};
Cache g_caches[kMaxDevices];

hipError_t ensure(Cache& c, int slot, size_t bytes) {
  if (c.cap[slot] >= bytes) return hipSuccess;
  if (c.buf[slot]) (void)hipFree(c.buf[slot]);
  c.buf[slot] = nullptr;
  c.cap[slot] = 0;
  hipError_t e = hipMalloc(&c.buf[slot], bytes);
  if (e == hipSuccess) c.cap[slot] = bytes;
  return e;
}

// The current device's cache.
int current_cache(Cache** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(GMLE_ERR_UNSUPPORTED, "device %d beyond the cache table", dev);
  *out = &g_caches[dev];
  return GMLE_OK;
}

// The cache's stream and upload pipeline, created on first use (caller holds c.mu).
int cache_stream(Cache& c) {
  if (!c.stream) {
    hipError_t e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  }
  const hipError_t e = c.pipe.init();
  if (e != hipSuccess) return hip_fail(e, "host pipeline set-up (pinned staging)");
  return GMLE_OK;
}

// Copies device outputs back through the pinned output buffer: the pieces
// (device source, host destination, bytes) are queued on `s`, waited for and
// copied to the caller's arrays by the worker pool.
struct OutPiece {
  const void* dev;
  void* host;
  size_t bytes;
};
// Offsets of the pieces in the pinned output buffer (256-byte aligned), and
// the buffer grown to hold them.
hipError_t out_layout(Cache& C, const OutPiece* pieces, int n, size_t* offs) {
  size_t total = 0;
  for (int k = 0; k < n; ++k) {
    offs[k] = total;
    total += (pieces[k].bytes + 255) & ~(size_t)255;
  }
  return C.pipe.ensure_out(total);
}
// After the stream's downloads into the pinned buffer: the caller's arrays.
int out_finish(Cache& C, hipStream_t s, const OutPiece* pieces, int n, const size_t* offs) {
  const hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (download)");
  for (int k = 0; k < n; ++k) copy_parallel(pieces[k].host, (char*)C.pipe.out + offs[k], pieces[k].bytes);
  return GMLE_OK;
}

int download(Cache& C, hipStream_t s, const OutPiece* pieces, int n) {
  size_t total = 0;
  for (int k = 0; k < n; ++k) total += (pieces[k].bytes + 255) & ~(size_t)255;
  hipError_t e = C.pipe.ensure_out(total);
  if (e != hipSuccess) return hip_fail(e, "hipHostMalloc (output staging)");
  size_t at = 0;
  for (int k = 0; k < n && e == hipSuccess; ++k) {
    e = hipMemcpyAsync((char*)C.pipe.out + at, pieces[k].dev, pieces[k].bytes, hipMemcpyDeviceToHost, s);
    at += (pieces[k].bytes + 255) & ~(size_t)255;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (download)");
  at = 0;
  for (int k = 0; k < n; ++k) {
    copy_parallel(pieces[k].host, (char*)C.pipe.out + at, pieces[k].bytes);
    at += (pieces[k].bytes + 255) & ~(size_t)255;
  }
  return GMLE_OK;
}

// Chunk ends (in floats) of an upload of n_rois ROIs whose ROI k ends at
// end(k) floats: each chunk holds whole ROIs and at most kStageBytes.
template <class EndFn>
std::vector<size_t> chunk_cuts(int64_t n_rois, EndFn end) {
  std::vector<size_t> cuts;
  const size_t cap = kStageBytes / sizeof(float);
  // the first chunks are smaller (1/8, 1/4, 1/2 of a buffer), so the first
  // upload starts early and the copies, uploads and fits fill in sooner
  auto cap_of = [&](size_t i) { return i < 3 ? cap >> (3 - i) : cap; };
  size_t start = 0;
  for (int64_t k = 0; k < n_rois; ++k) {
    const size_t e = end(k);
    if (e - start > cap_of(cuts.size()) && k > 0 && end(k - 1) > start) {
      cuts.push_back(end(k - 1));
      start = end(k - 1);
    }
  }
  cuts.push_back(end(n_rois - 1));
  return cuts;
}

// ---------------------------------------------------------------------------
// Mixed-shape batches.
// ---------------------------------------------------------------------------
constexpr int kMaxStreams = 4;

// Extra streams (default priority) and fork / join events per device; the
// caller's stream is the first of the set.  `init` guards the one-time
// creation (handles are made once and never replaced); `mu` is held by a
// call for as long as it uses the streams and events.
struct StreamSet {
  std::mutex init;
  std::mutex mu;
  bool ready = false;
  hipStream_t s[kMaxStreams - 1] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxStreams - 1] = {};
};
StreamSet g_streams[kMaxDevices];

int stream_set(StreamSet** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(GMLE_ERR_UNSUPPORTED, "device %d beyond the stream table", dev);
  StreamSet& ss = g_streams[dev];
  std::lock_guard<std::mutex> guard(ss.init);
  if (!ss.ready) {
    // Created into locals and published only when all exist, so a failure
    // leaves the set empty for the next call to retry.
    hipStream_t s[kMaxStreams - 1] = {};
    hipEvent_t join[kMaxStreams - 1] = {};
    hipEvent_t fork = nullptr;
    auto undo = [&]() {
      for (int k = 0; k < kMaxStreams - 1; ++k) {
        if (s[k]) (void)hipStreamDestroy(s[k]);
        if (join[k]) (void)hipEventDestroy(join[k]);
      }
      if (fork) (void)hipEventDestroy(fork);
    };
    for (int k = 0; k < kMaxStreams - 1; ++k) {
      if ((e = hipStreamCreateWithFlags(&s[k], hipStreamNonBlocking)) != hipSuccess) {
        undo();
        return hip_fail(e, "hipStreamCreate");
      }
      if ((e = hipEventCreateWithFlags(&join[k], hipEventDisableTiming)) != hipSuccess) {
        undo();
        return hip_fail(e, "hipEventCreate");
      }
    }
    if ((e = hipEventCreateWithFlags(&fork, hipEventDisableTiming)) != hipSuccess) {
      undo();
      return hip_fail(e, "hipEventCreate");
    }
    for (int k = 0; k < kMaxStreams - 1; ++k) {
      ss.s[k] = s[k];
      ss.join[k] = join[k];
    }
    ss.fork = fork;
    ss.ready = true;
  }
  *out = &ss;
  return GMLE_OK;
}

// Relative cost of one ROI of a shape: FP32 operations per Newton iteration
// of the separable formulation (per pixel plus per edge on both axes).
double shape_cost(int model, int box) {
  const double c_pix = model == XYNB ? 32 : model == XYNBSXSY ? 50 : 45;
  const double c_axis = model == XYNB ? 13 : 23;
  return box * box * c_pix + 2.0 * box * c_axis;
}

struct BatchSpec {
  int64_t n;
  const int8_t* models;
  const int8_t* boxes;
  const int64_t* offsets;
  const float* data;
  const float* psf;
  const float* lower;
  const float* upper;
  const float* max_step;
  int camera;
  const float* var_map;
  int ny, nx;
  const int32_t* xc;
  const int32_t* yc;
  int iterations, precision;
  const int32_t* lanes;
  int streams;
  const int64_t* counts;   // host, kNumShapes
  float* params;
  float* crlb;
  float* llr;
  double* crlb64;
  double* llr64;
  void* workspace;
};

// Checks shared by both batch entry points, once the shape counts are known.
int check_batch(const BatchSpec& b) {
  if (b.camera != IDEAL && b.camera != SCMOS) return fail(GMLE_ERR_ARG, "unknown camera model %d", b.camera);
  if (b.precision != PREC_FP32 && b.precision != PREC_MIXED)
    return fail(GMLE_ERR_ARG, "unknown precision %d", b.precision);
  if (b.iterations < 0) return fail(GMLE_ERR_ARG, "iterations must be >= 0, got %d", b.iterations);
  if (b.streams < 0 || b.streams > kMaxStreams)
    return fail(GMLE_ERR_ARG, "streams must be 0..%d, got %d", kMaxStreams, b.streams);
  if (b.streams == 0 && b.precision != PREC_FP32)
    return fail(GMLE_ERR_UNSUPPORTED, "the one-launch mode (streams = 0) runs Float32 only");
  if (!b.psf || !b.lower || !b.upper || !b.max_step)
    return fail(GMLE_ERR_ARG, "psf and constraints arrays must not be NULL");
  int64_t total = 0;
  for (int s = 0; s < kNumShapes; ++s) {
    if (b.counts[s] < 0) return fail(GMLE_ERR_ARG, "negative shape count");
    total += b.counts[s];
    if (b.counts[s] == 0) continue;
    const int model = s / kNumBoxes, box = kMinBox + s % kNumBoxes;
    const float* psf = b.psf + 8 * model;
    if (model == XYNB && !(psf[0] > 0.f)) return fail(GMLE_ERR_ARG, "PSF sigma must be > 0");
    if (model == ASTIG && !(psf[0] > 0.f && psf[1] > 0.f && psf[7] != 0.f && psf[6] != 0.f))
      return fail(GMLE_ERR_ARG, "astigmatic PSF needs sigma_x0, sigma_y0 > 0 and nonzero gamma, d");
    if (b.camera == SCMOS && (b.ny < box || b.nx < box))
      return fail(GMLE_ERR_ARG, "variance map %dx%d is smaller than one ROI", b.ny, b.nx);
    const int lanes = (b.lanes && b.lanes[s] > 0) ? b.lanes[s] : 0;
    if (lanes && !lanes_ok(box, lanes))
      return fail(GMLE_ERR_UNSUPPORTED, "no kernel with %d lanes per ROI for box size %d", lanes, box);
  }
  if (total > b.n) return fail(GMLE_ERR_ARG, "shape counts add up to %lld > n_rois %lld", (long long)total, (long long)b.n);
  return GMLE_OK;
}

// Enqueues the grouping kernels on `stream`, then one fit launch per shape
// present, spread over b.streams streams (largest estimated work first, each
// to the least loaded stream), joined back into `stream`.
int enqueue_batch(const BatchSpec& b, hipStream_t stream) {
  BatchPlan p;
  std::memset(&p, 0, sizeof(p));
  p.models = b.models;
  p.boxes = b.boxes;
  p.n_rois = b.n;
  char* ws = (char*)b.workspace;
  p.order = (int32_t*)ws;
  p.cursor = (int32_t*)(ws + align256((size_t)b.n * sizeof(int32_t)));
  p.params = b.params;
  p.crlb = b.crlb;
  p.llr = b.llr;
  p.crlb64 = b.precision == PREC_MIXED ? b.crlb64 : nullptr;
  p.llr64 = b.precision == PREC_MIXED ? b.llr64 : nullptr;
  // Shape lists in fixed arrays: enqueueing a batch allocates nothing.
  int64_t slot = 0;
  int shapes[kNumShapes];
  int n_shapes = 0;
  for (int s = 0; s < kNumShapes; ++s) {
    p.start[s] = (int32_t)slot;
    p.count[s] = (int32_t)b.counts[s];
    slot += b.counts[s];
    if (b.counts[s] > 0) shapes[n_shapes++] = s;
  }
  hipError_t e = enqueue_batch_plan(p, stream);
  if (e != hipSuccess) return hip_fail(e, "batch planning launch");

  if (b.streams == 0) {
    // One launch for every shape the multi-shape kernel holds, heaviest tiles
    // first, with the requested lanes where it has them (else 4); the others
    // follow on the same stream with their own kernels.
    int multi[kNumShapes], rest[kNumShapes];
    int n_multi = 0, n_rest = 0;
    for (int q = 0; q < n_shapes; ++q) {
      const int s = shapes[q];
      if (multi_eligible(s / kNumBoxes, b.camera, kMinBox + s % kNumBoxes))
        multi[n_multi++] = s;
      else
        rest[n_rest++] = s;
    }
    std::sort(multi, multi + n_multi, [&](int u, int v) {
      return shape_cost(u / kNumBoxes, kMinBox + u % kNumBoxes) > shape_cost(v / kNumBoxes, kMinBox + v % kNumBoxes);
    });
    MultiArgs m;
    std::memset(&m, 0, sizeof(m));
    for (int model = 0; model < 4; ++model) {
      FitArgs& a = m.fa[model];
      fill_args(a, model, b.psf + 8 * model, b.lower + 6 * model, b.upper + 6 * model, b.max_step + 6 * model, 0,
                b.ny, b.nx, b.iterations);
      a.data = b.data;
      a.var_map = b.camera == SCMOS ? b.var_map : nullptr;
      a.x_corners = b.camera == SCMOS ? b.xc : nullptr;
      a.y_corners = b.camera == SCMOS ? b.yc : nullptr;
      a.params = b.params;
      a.crlb = b.crlb;
      a.llr = b.llr;
      a.order = p.order;
      a.offsets = b.offsets;
      a.out_stride = kBatchStride;
    }
    for (int qm = 0; qm < n_multi; ++qm) {
      const int s = multi[qm];
      const int model = s / kNumBoxes, box = kMinBox + s % kNumBoxes;
      const int want = (b.lanes && b.lanes[s] > 0) ? b.lanes[s] : large_batch_lanes(model, b.camera, box);
      const int q = m.n_entries++;
      m.shape[q] = s;
      m.lanes[q] = multi_lanes(model, b.camera, box, want);
      m.first[q] = p.start[s];
      m.count[q] = p.count[s];
    }
    multi_plan_blocks(m);
    if (m.n_entries > 0) {
      e = b.camera == SCMOS ? launch_multi_scmos(m, stream) : launch_multi_ideal(m, stream);
      if (e != hipSuccess) return hip_fail(e, "kernel launch");
    }
    for (int q = 0; q < n_rest; ++q) shapes[q] = rest[q];
    n_shapes = n_rest;
  }
  std::sort(shapes, shapes + n_shapes, [&](int u, int v) {
    const double cu = b.counts[u] * shape_cost(u / kNumBoxes, kMinBox + u % kNumBoxes);
    const double cv = b.counts[v] * shape_cost(v / kNumBoxes, kMinBox + v % kNumBoxes);
    return cu > cv;
  });
  const int ns = std::min<int>(b.streams, std::max<int>(1, n_shapes));
  StreamSet* ss = nullptr;
  std::unique_lock<std::mutex> lock;
  hipStream_t strm[kMaxStreams] = {stream};
  if (ns > 1) {
    int rc = stream_set(&ss);
    if (rc != GMLE_OK) return rc;
    lock = std::unique_lock<std::mutex>(ss->mu);
    if ((e = hipEventRecord(ss->fork, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    for (int k = 1; k < ns; ++k) {
      strm[k] = ss->s[k - 1];
      if ((e = hipStreamWaitEvent(strm[k], ss->fork, 0)) != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
    }
  }
  double load[kMaxStreams] = {};
  for (int q = 0; q < n_shapes; ++q) {
    const int s = shapes[q];
    const int model = s / kNumBoxes, box = kMinBox + s % kNumBoxes;
    int lanes = (b.lanes && b.lanes[s] > 0) ? b.lanes[s] : default_lanes(model, b.camera, box, b.precision, b.counts[s]);
    LaunchFn fn = find_launcher(model, b.camera, box, lanes, b.precision);
    if (!fn) return fail(GMLE_ERR_UNSUPPORTED, "no kernel for model %d box %d lanes %d", model, box, lanes);
    FitArgs a;
    fill_args(a, model, b.psf + 8 * model, b.lower + 6 * model, b.upper + 6 * model, b.max_step + 6 * model,
              b.counts[s], b.ny, b.nx, b.iterations);
    a.data = b.data;
    a.var_map = b.camera == SCMOS ? b.var_map : nullptr;
    a.x_corners = b.camera == SCMOS ? b.xc : nullptr;
    a.y_corners = b.camera == SCMOS ? b.yc : nullptr;
    a.params = b.params;
    a.crlb = b.crlb;
    a.llr = b.llr;
    a.crlb64 = p.crlb64;
    a.llr64 = p.llr64;
    a.order = p.order;
    a.offsets = b.offsets;
    a.first = p.start[s];
    a.out_stride = kBatchStride;
    int k = 0;
    for (int q = 1; q < ns; ++q)
      if (load[q] < load[k]) k = q;
    load[k] += b.counts[s] * shape_cost(model, box);
    if ((e = fn(a, strm[k])) != hipSuccess) return hip_fail(e, "kernel launch");
  }
  for (int k = 1; k < ns; ++k) {
    if ((e = hipEventRecord(ss->join[k - 1], strm[k])) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = hipStreamWaitEvent(stream, ss->join[k - 1], 0)) != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
  }
  return GMLE_OK;
}

}  // namespace
}  // namespace gmle

using namespace gmle;

extern "C" {

const char* gmle_version(void) { return "gaussmle-hip 0.1.0 (gfx950)"; }

int gmle_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

size_t gmle_last_error(char* buf, size_t len) {
  if (buf && len) {
    const size_t n = g_err_len < len - 1 ? g_err_len : len - 1;
    std::memcpy(buf, g_err, n);
    buf[n] = '\0';
  }
  return g_err_len;
}

int gmle_n_params(int model) {
  if (model < XYNB || model > ASTIG) return fail(GMLE_ERR_ARG, "unknown PSF model %d", model);
  return n_params(model);
}

int gmle_default_constraints(int model, int box, float* lower, float* upper, float* max_step) {
  if (model < XYNB || model > ASTIG) return fail(GMLE_ERR_ARG, "unknown PSF model %d", model);
  if (!lower || !upper || !max_step) return fail(GMLE_ERR_ARG, "output arrays must not be NULL");
  const float inf = INFINITY;
  const float hi = (float)box + 2.f;
  // src/constraints.jl:16-47
  switch (model) {
    case XYNB: {
      const float lo[4] = {-2.f, -2.f, 1.f, 0.01f}, up[4] = {hi, hi, inf, inf}, ms[4] = {1.f, 1.f, 100.f, 2.f};
      for (int k = 0; k < 4; ++k) { lower[k] = lo[k]; upper[k] = up[k]; max_step[k] = ms[k]; }
      break;
    }
    case XYNBS: {
      const float lo[5] = {-2.f, -2.f, 1.f, 0.01f, 0.3f}, up[5] = {hi, hi, inf, inf, 5.f};
      const float ms[5] = {1.f, 1.f, 100.f, 2.f, 0.5f};
      for (int k = 0; k < 5; ++k) { lower[k] = lo[k]; upper[k] = up[k]; max_step[k] = ms[k]; }
      break;
    }
    case XYNBSXSY: {
      const float lo[6] = {-2.f, -2.f, 1.f, 0.01f, 0.3f, 0.3f}, up[6] = {hi, hi, inf, inf, 5.f, 5.f};
      const float ms[6] = {1.f, 1.f, 100.f, 2.f, 0.5f, 0.5f};
      for (int k = 0; k < 6; ++k) { lower[k] = lo[k]; upper[k] = up[k]; max_step[k] = ms[k]; }
      break;
    }
    default: {
      const float lo[5] = {-2.f, -2.f, -1.f, 1.f, 0.01f}, up[5] = {hi, hi, 1.f, inf, inf};
      const float ms[5] = {1.f, 1.f, 0.1f, 100.f, 2.f};
      for (int k = 0; k < 5; ++k) { lower[k] = lo[k]; upper[k] = up[k]; max_step[k] = ms[k]; }
      break;
    }
  }
  return GMLE_OK;
}

int gmle_max_box(void) { return kMaxBox; }

int gmle_lanes_supported(int box, int lanes) { return lanes_ok(box, lanes) ? 1 : 0; }

int gmle_default_lanes(int model, int camera, int box, int precision, int64_t n_rois) {
  return default_lanes(model, camera, box, precision, n_rois);
}

int gmle_fit_device(int model, const float* psf, int box, int64_t n_rois, const float* d_data, int camera,
                    const float* d_var_map, int ny, int nx, const int32_t* d_x_corners,
                    const int32_t* d_y_corners, const float* lower, const float* upper, const float* max_step,
                    int iterations, int precision, int lanes, float* d_params, float* d_crlb, float* d_llr,
                    double* d_crlb64, double* d_llr64, void* stream) {
  int rc = check_common(model, psf, box, n_rois, camera, ny, nx, lower, upper, max_step, iterations, precision,
                        &lanes);
  if (rc != GMLE_OK) return rc;
  if (!d_data || !d_params || !d_crlb || !d_llr) return fail(GMLE_ERR_ARG, "data / output pointers must not be NULL");
  if (camera == SCMOS && (!d_var_map || !d_x_corners || !d_y_corners))
    return fail(GMLE_ERR_ARG, "sCMOS fits need the variance map and ROI corners");
  LaunchFn fn = find_launcher(model, camera, box, lanes, precision);
  if (!fn) return fail(GMLE_ERR_UNSUPPORTED, "no kernel for model %d box %d lanes %d", model, box, lanes);
  FitArgs a;
  fill_args(a, model, psf, lower, upper, max_step, n_rois, ny, nx, iterations);
  a.data = d_data;
  a.var_map = d_var_map;
  a.x_corners = d_x_corners;
  a.y_corners = d_y_corners;
  a.params = d_params;
  a.crlb = d_crlb;
  a.llr = d_llr;
  a.crlb64 = precision == PREC_MIXED ? d_crlb64 : nullptr;
  a.llr64 = precision == PREC_MIXED ? d_llr64 : nullptr;
  const hipError_t e = fn(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return GMLE_OK;
}

int gmle_fit(int model, const float* psf, int box, int64_t n_rois, const float* data, int camera,
             const float* var_map, int ny, int nx, const int32_t* x_corners, const int32_t* y_corners,
             const float* lower, const float* upper, const float* max_step, int iterations, int precision,
             int lanes, float* params, float* crlb, float* llr, double* crlb64, double* llr64) {
  int rc = check_common(model, psf, box, n_rois, camera, ny, nx, lower, upper, max_step, iterations, precision,
                        &lanes);
  if (rc != GMLE_OK) return rc;
  if (!data || !params || !crlb || !llr) return fail(GMLE_ERR_ARG, "data / output pointers must not be NULL");
  if (camera == SCMOS) {
    if (!var_map || !x_corners || !y_corners)
      return fail(GMLE_ERR_ARG, "sCMOS fits need the variance map and ROI corners");
    for (int64_t k = 0; k < n_rois; ++k) {
      const int cx = x_corners[k], cy = y_corners[k];
      if (cx < 1 || cy < 1 || cx > nx - box + 1 || cy > ny - box + 1)
        return fail(GMLE_ERR_BOUNDS, "ROI %lld at corner (x=%d, y=%d) lies outside the %dx%d variance map",
                    (long long)(k + 1), cx, cy, ny, nx);
    }
  }
  // With a device the upload pipeline checks every value as it copies it; without one the data are
  // still validated first, as validate_fit_input (src/interface.jl:26-29) does before any fit.
  if (gmle_device_count() <= 0) {
    if (!all_finite(data, (size_t)n_rois * (size_t)box * (size_t)box))
      return fail(GMLE_ERR_DATA, "Input data contains NaN or Inf values");
    return fail(GMLE_ERR_NODEVICE, "no HIP device available");
  }
  LaunchFn fn = find_launcher(model, camera, box, lanes, precision);
  if (!fn) return fail(GMLE_ERR_UNSUPPORTED, "no kernel for model %d box %d lanes %d", model, box, lanes);

  Cache* cache = nullptr;
  if ((rc = current_cache(&cache)) != GMLE_OK) return rc;
  Cache& C = *cache;
  std::lock_guard<std::mutex> lock(C.mu);
  if ((rc = cache_stream(C)) != GMLE_OK) return rc;
  hipError_t e = hipSuccess;
  const int P = n_params(model);
  const size_t box2 = (size_t)box * (size_t)box;
  const size_t npix = (size_t)n_rois * box2;
  const size_t nvar = camera == SCMOS ? (size_t)ny * (size_t)nx : 0;
  const bool want64 = precision == PREC_MIXED && (crlb64 || llr64);
  if ((e = ensure(C, 0, npix * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc data");
  if ((e = ensure(C, 1, (size_t)n_rois * P * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc params");
  if ((e = ensure(C, 2, (size_t)n_rois * P * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc crlb");
  if ((e = ensure(C, 3, (size_t)n_rois * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc llr");
  if (camera == SCMOS) {
    if ((e = ensure(C, 4, nvar * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc var_map");
    if ((e = ensure(C, 5, (size_t)n_rois * 2 * sizeof(int32_t))) != hipSuccess) return hip_fail(e, "hipMalloc corners");
  }
  if (want64) {
    if ((e = ensure(C, 6, (size_t)n_rois * P * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc crlb64");
    if ((e = ensure(C, 7, (size_t)n_rois * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc llr64");
  }
  hipStream_t s = C.stream;
  float* d_data = (float*)C.buf[0];
  int32_t* d_corners = (int32_t*)C.buf[5];
  // the variance map and corners go first, on the work stream
  if (camera == SCMOS) {
    e = hipMemcpyAsync(C.buf[4], var_map, nvar * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_corners, x_corners, (size_t)n_rois * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_corners + n_rois, y_corners, (size_t)n_rois * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (upload)");
  }
  FitArgs base;
  fill_args(base, model, psf, lower, upper, max_step, n_rois, ny, nx, iterations);
  base.var_map = camera == SCMOS ? (const float*)C.buf[4] : nullptr;
  base.params = (float*)C.buf[1];
  base.crlb = (float*)C.buf[2];
  base.llr = (float*)C.buf[3];
  base.crlb64 = want64 ? (double*)C.buf[6] : nullptr;
  base.llr64 = want64 ? (double*)C.buf[7] : nullptr;
  // outputs: each chunk's results go to the pinned output buffer right after
  // its fit (downloads overlap the later chunks' uploads); the caller's arrays
  // are written once every chunk has passed the check
  OutPiece pieces[5] = {{base.params, params, (size_t)n_rois * P * sizeof(float)},
                        {base.crlb, crlb, (size_t)n_rois * P * sizeof(float)},
                        {base.llr, llr, (size_t)n_rois * sizeof(float)}};
  int np = 3;
  if (want64 && crlb64) pieces[np++] = {base.crlb64, crlb64, (size_t)n_rois * P * sizeof(double)};
  if (want64 && llr64) pieces[np++] = {base.llr64, llr64, (size_t)n_rois * sizeof(double)};
  size_t offs[5];
  if ((e = out_layout(C, pieces, np, offs)) != hipSuccess) return hip_fail(e, "hipHostMalloc (output staging)");
  // pixels stream in through the pinned ring; each chunk is fitted as soon as
  // it is on the device, with the lanes chosen for the whole call (a ROI's
  // result does not depend on which launch holds it)
  const std::vector<size_t> cuts = chunk_cuts(n_rois, [&](int64_t k) { return (size_t)(k + 1) * box2; });
  auto launch = [&](int k) -> int {
    const int64_t r0 = k == 0 ? 0 : (int64_t)(cuts[k - 1] / box2), r1 = (int64_t)(cuts[k] / box2);
    FitArgs a = base;
    a.n_rois = r1 - r0;
    a.data = d_data + (size_t)r0 * box2;
    if (camera == SCMOS) {
      a.x_corners = d_corners + r0;
      a.y_corners = d_corners + n_rois + r0;
    }
    a.params += r0 * P;
    a.crlb += r0 * P;
    a.llr += r0;
    if (a.crlb64) a.crlb64 += r0 * P;
    if (a.llr64) a.llr64 += r0;
    hipError_t le = fn(a, s);
    if (le != hipSuccess) return hip_fail(le, "kernel launch");
    for (int q = 0; q < np && le == hipSuccess; ++q) {
      const size_t per = pieces[q].bytes / (size_t)n_rois;   // bytes per ROI
      le = hipMemcpyAsync((char*)C.pipe.out + offs[q] + (size_t)r0 * per, (const char*)pieces[q].dev + r0 * per,
                          (size_t)(r1 - r0) * per, hipMemcpyDeviceToHost, s);
    }
    return le == hipSuccess ? GMLE_OK : hip_fail(le, "hipMemcpyAsync (download)");
  };
  rc = pipe_upload(C.pipe, data, d_data, cuts.data(), (int)cuts.size(), s, launch);
  if (rc != GMLE_OK) {
    (void)hipStreamSynchronize(s);
    return rc;
  }
  return out_finish(C, s, pieces, np, offs);
}

int gmle_shape_index(int model, int box) { return shape_index(model, box); }

int gmle_one_launch_lanes(int model, int camera, int box, int lanes) {
  if (shape_index(model, box) < 0 || (camera != IDEAL && camera != SCMOS)) return 0;
  if (!multi_eligible(model, camera, box)) return 0;
  return multi_lanes(model, camera, box, lanes > 0 ? lanes : large_batch_lanes(model, camera, box));
}

size_t gmle_batch_workspace_bytes(int64_t n_rois) { return batch_workspace(n_rois < 0 ? 0 : n_rois); }

int gmle_fit_batch_device(int64_t n_rois, const int8_t* d_models, const int8_t* d_boxes,
                          const int64_t* d_offsets, const float* d_data, const float* psf, const float* lower,
                          const float* upper, const float* max_step, int camera, const float* d_var_map, int ny,
                          int nx, const int32_t* d_x_corners, const int32_t* d_y_corners, int iterations,
                          int precision, const int32_t* lanes, int streams, const int64_t* shape_counts,
                          float* d_params, float* d_crlb, float* d_llr, double* d_crlb64, double* d_llr64,
                          void* d_workspace, size_t workspace_bytes, void* stream) {
  if (n_rois <= 0) return fail(GMLE_ERR_ARG, "Input data array is empty");
  if (n_rois > INT32_MAX) return fail(GMLE_ERR_UNSUPPORTED, "batches hold at most %d ROIs", INT32_MAX);
  if (!d_models || !d_boxes || !d_offsets || !d_data || !d_params || !d_crlb || !d_llr || !d_workspace)
    return fail(GMLE_ERR_ARG, "device pointers must not be NULL");
  if (workspace_bytes < batch_workspace(n_rois))
    return fail(GMLE_ERR_ARG, "workspace of %zu bytes is smaller than the %zu needed", workspace_bytes,
                batch_workspace(n_rois));
  if (camera == SCMOS && (!d_var_map || !d_x_corners || !d_y_corners))
    return fail(GMLE_ERR_ARG, "sCMOS fits need the variance map and ROI corners");
  hipStream_t s = (hipStream_t)stream;
  int64_t counts[kNumShapes];
  if (shape_counts) {
    for (int q = 0; q < kNumShapes; ++q) counts[q] = shape_counts[q];
  } else {
    int32_t* d_counts = (int32_t*)((char*)d_workspace + align256((size_t)n_rois * sizeof(int32_t)) +
                                   align256(kNumShapes * sizeof(int32_t)));
    int32_t h[kNumShapes];
    hipError_t e = enqueue_batch_count(d_models, d_boxes, n_rois, d_counts, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h, d_counts, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "batch shape count");
    for (int q = 0; q < kNumShapes; ++q) counts[q] = h[q];
  }
  BatchSpec b{n_rois, d_models, d_boxes, d_offsets, d_data, psf, lower, upper, max_step, camera, d_var_map, ny, nx,
              d_x_corners, d_y_corners, iterations, precision, lanes, streams, counts, d_params, d_crlb, d_llr,
              d_crlb64, d_llr64, d_workspace};
  int rc = check_batch(b);
  if (rc != GMLE_OK) return rc;
  return enqueue_batch(b, s);
}

int gmle_fit_batch(int64_t n_rois, const int8_t* models, const int8_t* boxes, const float* data,
                   const float* psf, const float* lower, const float* upper, const float* max_step,
                   int camera, const float* var_map, int ny, int nx, const int32_t* x_corners,
                   const int32_t* y_corners, int iterations, int precision, const int32_t* lanes, int streams,
                   float* params, float* crlb, float* llr, double* crlb64, double* llr64) {
  if (n_rois <= 0) return fail(GMLE_ERR_ARG, "Input data array is empty");
  if (n_rois > INT32_MAX) return fail(GMLE_ERR_UNSUPPORTED, "batches hold at most %d ROIs", INT32_MAX);
  if (!models || !boxes || !data || !params || !crlb || !llr)
    return fail(GMLE_ERR_ARG, "data / output pointers must not be NULL");
  if (camera == SCMOS && (!var_map || !x_corners || !y_corners))
    return fail(GMLE_ERR_ARG, "sCMOS fits need the variance map and ROI corners");
  std::vector<int64_t> offsets((size_t)n_rois);
  int64_t counts[kNumShapes] = {};
  int64_t npix = 0;
  for (int64_t k = 0; k < n_rois; ++k) {
    const int m = models[k], box = boxes[k];
    if (m < XYNB || m > ASTIG) return fail(GMLE_ERR_ARG, "ROI %lld: unknown PSF model %d", (long long)(k + 1), m);
    if (box < kMinBox) return fail(GMLE_ERR_ARG, "ROI size must be at least 3x3, got %dx%d", box, box);
    if (box > kMaxBatchBox)
      return fail(GMLE_ERR_UNSUPPORTED, "ROI size %d exceeds the mixed-batch maximum %d", box, kMaxBatchBox);
    if (camera == SCMOS) {
      const int cx = x_corners[k], cy = y_corners[k];
      if (cx < 1 || cy < 1 || cx > nx - box + 1 || cy > ny - box + 1)
        return fail(GMLE_ERR_BOUNDS, "ROI %lld at corner (x=%d, y=%d) lies outside the %dx%d variance map",
                    (long long)(k + 1), cx, cy, ny, nx);
    }
    offsets[(size_t)k] = npix;
    npix += (int64_t)box * box;
    ++counts[shape_index(m, box)];
  }
  BatchSpec b{n_rois, nullptr, nullptr, nullptr, nullptr, psf, lower, upper, max_step, camera, nullptr, ny, nx,
              nullptr, nullptr, iterations, precision, lanes, streams, counts, nullptr, nullptr, nullptr,
              nullptr, nullptr, nullptr};
  int rc = check_batch(b);
  if (rc != GMLE_OK) return rc;
  if (gmle_device_count() <= 0) {
    if (!all_finite(data, (size_t)npix)) return fail(GMLE_ERR_DATA, "Input data contains NaN or Inf values");
    return fail(GMLE_ERR_NODEVICE, "no HIP device available");
  }

  // chunks of whole ROIs through the pinned ring; each chunk is its own
  // batch (grouped and fitted on the device as soon as it has landed)
  const std::vector<size_t> cuts =
      chunk_cuts(n_rois, [&](int64_t k) { return (size_t)(offsets[(size_t)k] + (int64_t)boxes[k] * boxes[k]); });
  std::vector<int64_t> first(cuts.size() + 1, 0);
  {
    size_t c = 0;
    for (int64_t k = 0; k < n_rois; ++k) {
      const size_t e = (size_t)(offsets[(size_t)k] + (int64_t)boxes[k] * boxes[k]);
      if (e > cuts[c]) first[++c] = k;
    }
    first[cuts.size()] = n_rois;
  }
  int64_t max_chunk = 0;
  for (size_t c = 0; c < cuts.size(); ++c) max_chunk = std::max(max_chunk, first[c + 1] - first[c]);

  Cache* cache = nullptr;
  if ((rc = current_cache(&cache)) != GMLE_OK) return rc;
  Cache& C = *cache;
  std::lock_guard<std::mutex> lock(C.mu);
  if ((rc = cache_stream(C)) != GMLE_OK) return rc;
  const size_t n = (size_t)n_rois;
  const size_t nvar = camera == SCMOS ? (size_t)ny * (size_t)nx : 0;
  const bool want64 = precision == PREC_MIXED && (crlb64 || llr64);
  hipError_t e = hipSuccess;
  if ((e = ensure(C, 0, (size_t)npix * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc data");
  if ((e = ensure(C, 1, n * kBatchStride * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc params");
  if ((e = ensure(C, 2, n * kBatchStride * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc crlb");
  if ((e = ensure(C, 3, n * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc llr");
  if (camera == SCMOS) {
    if ((e = ensure(C, 4, nvar * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc var_map");
    if ((e = ensure(C, 5, n * 2 * sizeof(int32_t))) != hipSuccess) return hip_fail(e, "hipMalloc corners");
  }
  if (want64) {
    if ((e = ensure(C, 6, n * kBatchStride * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc crlb64");
    if ((e = ensure(C, 7, n * sizeof(double))) != hipSuccess) return hip_fail(e, "hipMalloc llr64");
  }
  if ((e = ensure(C, 8, n)) != hipSuccess) return hip_fail(e, "hipMalloc models");
  if ((e = ensure(C, 9, n)) != hipSuccess) return hip_fail(e, "hipMalloc boxes");
  if ((e = ensure(C, 10, n * sizeof(int64_t))) != hipSuccess) return hip_fail(e, "hipMalloc offsets");
  if ((e = ensure(C, 11, batch_workspace(max_chunk))) != hipSuccess) return hip_fail(e, "hipMalloc workspace");
  hipStream_t s = C.stream;
  int32_t* d_corners = (int32_t*)C.buf[5];
  e = hipMemcpyAsync(C.buf[8], models, n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(C.buf[9], boxes, n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(C.buf[10], offsets.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && camera == SCMOS) {
    e = hipMemcpyAsync(C.buf[4], var_map, nvar * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_corners, x_corners, n * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_corners + n, y_corners, n * sizeof(int32_t), hipMemcpyHostToDevice, s);
  }
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (upload)");
  b.models = (const int8_t*)C.buf[8];
  b.boxes = (const int8_t*)C.buf[9];
  b.offsets = (const int64_t*)C.buf[10];
  b.data = (const float*)C.buf[0];
  b.var_map = camera == SCMOS ? (const float*)C.buf[4] : nullptr;
  b.xc = camera == SCMOS ? d_corners : nullptr;
  b.yc = camera == SCMOS ? d_corners + n : nullptr;
  b.params = (float*)C.buf[1];
  b.crlb = (float*)C.buf[2];
  b.llr = (float*)C.buf[3];
  b.crlb64 = want64 ? (double*)C.buf[6] : nullptr;
  b.llr64 = want64 ? (double*)C.buf[7] : nullptr;
  b.workspace = C.buf[11];
  // lanes per shape fixed for the whole call (automatic choices from the
  // whole batch's shape counts), so that a ROI's result does not depend on
  // how the call is cut into chunks
  // (the one-launch mode's automatic choice does not depend on counts)
  int32_t call_lanes[kNumShapes];
  for (int q = 0; q < kNumShapes; ++q) {
    const int model = q / kNumBoxes, box = kMinBox + q % kNumBoxes;
    const bool one_launch = streams == 0 && multi_eligible(model, camera, box);
    call_lanes[q] = (lanes && lanes[q] > 0)      ? lanes[q]
                    : (counts[q] > 0 && !one_launch) ? default_lanes(model, camera, box, precision, counts[q])
                                                     : 0;
  }
  b.lanes = call_lanes;
  int64_t chunk_counts[kNumShapes];
  auto launch = [&](int c) -> int {
    const int64_t r0 = first[(size_t)c], r1 = first[(size_t)c + 1];
    std::fill(chunk_counts, chunk_counts + kNumShapes, (int64_t)0);
    for (int64_t k = r0; k < r1; ++k) ++chunk_counts[shape_index(models[k], boxes[k])];
    BatchSpec q = b;
    q.n = r1 - r0;
    q.models = b.models + r0;
    q.boxes = b.boxes + r0;
    q.offsets = b.offsets + r0;   // offsets stay global: ROI pixels at data + offsets[k]
    if (camera == SCMOS) {
      q.xc = b.xc + r0;
      q.yc = b.yc + r0;
    }
    q.params = b.params + r0 * kBatchStride;
    q.crlb = b.crlb + r0 * kBatchStride;
    q.llr = b.llr + r0;
    if (q.crlb64) q.crlb64 = b.crlb64 + r0 * kBatchStride;
    if (q.llr64) q.llr64 = b.llr64 + r0;
    q.counts = chunk_counts;
    return enqueue_batch(q, s);
  };
  rc = pipe_upload(C.pipe, data, (float*)C.buf[0], cuts.data(), (int)cuts.size(), s, launch);
  if (rc != GMLE_OK) {
    (void)hipStreamSynchronize(s);
    return rc;
  }
  OutPiece pieces[5] = {{b.params, params, n * kBatchStride * sizeof(float)},
                        {b.crlb, crlb, n * kBatchStride * sizeof(float)},
                        {b.llr, llr, n * sizeof(float)}};
  int np = 3;
  if (want64 && crlb64) pieces[np++] = {b.crlb64, crlb64, n * kBatchStride * sizeof(double)};
  if (want64 && llr64) pieces[np++] = {b.llr64, llr64, n * sizeof(double)};
  return download(C, s, pieces, np);
}

}  // extern "C"
