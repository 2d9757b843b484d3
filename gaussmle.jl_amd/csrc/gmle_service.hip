// gmle_service.hip -- resident fitting service with continuous batching.
//
// Pieces:
//   * pinned registry: gmle_host_alloc hands out mapped pinned memory and
//     remembers its device address, so request buffers can be read and written
//     by kernels directly (no staging copies);
//   * slots: a device arena (raw counts, pixels in electrons, per-ROI
//     metadata, grouping workspace, results) plus a HIP graph
//         ingest -> init -> scatter -> { fit[shape] (-> final[shape]) } -> results
//     whose kernel nodes are re-parameterised for every batch;
//   * a dispatcher thread: packs pending requests FIFO into a free slot while
//     they fit, brings their raw counts to the GPU (see "ingest modes" below),
//     writes the slot's request / chunk tables (pinned), updates the graph
//     nodes and launches it on the slot's stream;
//   * one waiter thread per slot: waits for the slot's completion event, marks
//     its requests done, frees the slot.
// Ingest modes (how a batch's raw counts reach the arena):
//   * direct: the ingest kernel reads them from the pinned request buffers
//     over PCIe (GPU-initiated loads; no copy, the lowest latency for small
//     batches, but such loads reach only ~25-35 GB/s);
//   * staged: worker threads pack the batch's counts back to back into the
//     slot's pinned staging buffer, and one DMA copies them into the arena at
//     the copy engine's rate (~57 GB/s);
//   * dma: one DMA per run of requests whose counts lie back to back in host
//     memory (requests whose buffers come from one pinned pool in submission
//     order form one run per batch; separate buffers cost one copy each);
//   * auto (the default): dma for runs of at least kDmaMinRunBytes, direct for
//     the rest of the batch.
// GMLE_SERVICE_INGEST=direct|staged|dma|auto overrides the default.  On
// 2048 x 2048 frames of 50-5000 ROIs from one pool, with the queue kept full,
// auto sustains ~1.4-1.5e8 fits/s against ~9e7 for direct.
// The sCMOS conversion electrons = (adu - offset) * gain follows
// src/camera_models.jl:110-113 with per-pixel maps gathered through the ROI
// corner, the indexing the kernel uses for the variance map
// (src/unified_kernel.jl:277-282).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gaussmle.h"
#include "gmle_args.hpp"
#include "gmle_batch.hpp"
#include "gmle_host.hpp"
#include "gmle_hostpipe.hpp"
#include "gmle_multi.hpp"

namespace gmle {
namespace {

constexpr int kChunk = 16;          // ROIs per ingest / result block
constexpr int kIngestBlock = 256;
constexpr int kMaxReqPerBatch = 4096;
// A batch stops growing at this many ROIs (a larger request still forms a
// batch of its own): past it the GPU is full anyway, and splitting lets one
// slot's PCIe copies overlap another slot's fits.
constexpr int64_t kTargetBatchRois = 1 << 16;
constexpr int64_t kXferBlocks = 256;
// Grid cap of the ingest kernel when a batch's counts are already in the
// arena (8 blocks per CU).
constexpr int64_t kDeviceIngestBlocks = 2048;
// Auto ingest: a run of requests whose raw counts lie back to back in host
// memory moves by one DMA from this many bytes up (one copy costs ~10-15 us
// of fixed overhead, so shorter runs are read in place).
constexpr int64_t kDmaMinRunBytes = (int64_t)1 << 20;
// Pieces of the host-side packing handed to one worker.
constexpr size_t kPackPiece = (size_t)256 << 10;
enum IngestMode { INGEST_AUTO = 0, INGEST_DIRECT = 1, INGEST_STAGED = 2, INGEST_DMA = 3 };
// LDS staging of one chunk's raw counts: 16-byte vectors, plus one for a
// start that is not 16-byte aligned.
constexpr int kRawVecs = (kChunk * kMaxBatchBox * kMaxBatchBox * 2) / 16 + 1;
// Completion polling: spin this long, then sleep between polls.
constexpr double kSpinSeconds = 2e-3;
// Pinned memory the GPU reads and writes directly: mapped, visible to every
// device, coherent (not cached on the GPU, so tables the host rewrites for
// every batch are never read stale).
constexpr unsigned kPinnedFlags = hipHostMallocMapped | hipHostMallocPortable | hipHostMallocCoherent;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Pinned registry
// ---------------------------------------------------------------------------
struct PinnedBlock {
  char* host;
  char* dev;
  size_t bytes;
};
std::mutex g_pin_mu;
std::map<uintptr_t, PinnedBlock> g_pinned;   // by host start address

// Device address of [h, h + bytes) if it lies inside one gmle_host_alloc block.
bool pinned_dev(const void* h, size_t bytes, const void** d) {
  std::lock_guard<std::mutex> lock(g_pin_mu);
  const uintptr_t a = (uintptr_t)h;
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  const PinnedBlock& b = it->second;
  if (a < (uintptr_t)b.host || a + bytes > (uintptr_t)b.host + b.bytes) return false;
  *d = b.dev + (a - (uintptr_t)b.host);
  return true;
}

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------
struct ReqDesc {              // one request of a batch (device-visible addresses)
  const int8_t* models;
  const int8_t* boxes;
  const int32_t* xc;
  const int32_t* yc;
  const uint16_t* adu;        // the request's raw counts, copied into the arena
  float* params;
  float* crlb;
  float* llr;
  int64_t roi_base;           // first arena ROI slot
  int64_t pix_base;           // first arena pixel
};

struct Chunk {                // up to kChunk consecutive ROIs of one request
  int32_t req;
  int32_t first;              // first ROI, within the request
  int32_t count;
  int32_t npix;               // pixels of these ROIs, as validated at submission
  int64_t pix0;               // pixel offset of that ROI within the request
};

// Slack after a slot's pixels: reads stay inside the arena even if a caller
// rewrote a request's box sizes after submitting it (results are then
// meaningless, but nothing is read or written out of bounds).
constexpr int64_t kPixSlack = (int64_t)kChunk * kMaxBatchBox * kMaxBatchBox + kMaxBatchBox * kMaxBatchBox;

struct IngestArgs {
  const ReqDesc* reqs;
  const Chunk* chunks;
  int64_t n_chunks;
  const float* offset_map;    // (ny, nx) column-major
  const float* gain_map;
  int ny;
  int64_t map_size;           // ny * nx
  float* data;
  int8_t* models;
  int8_t* boxes;
  int32_t* xc;
  int32_t* yc;
  int64_t* offsets;
};

struct ResultArgs {
  const ReqDesc* reqs;
  const Chunk* chunks;
  int64_t n_chunks;
  const float* params;
  const float* crlb;
  const float* llr;
};

// One block per chunk of kChunk ROIs: lanes t < count load the ROIs'
// metadata (box sizes clamped, so a request rewritten after submission can
// never move a write outside its slot), then the whole block converts the
// chunk's pixels, thread t taking pixels t, t + kIngestBlock, ..., so every
// thread's map gathers are in flight together.
__global__ __launch_bounds__(kIngestBlock) void service_ingest(IngestArgs a) {
  __shared__ int sbox[kChunk];
  __shared__ int sxc[kChunk];
  __shared__ int syc[kChunk];
  __shared__ int spre[kChunk + 1];
  __shared__ float sinv[kChunk];
  __shared__ uint4 sraw[kRawVecs];
  const int t = threadIdx.x;
  for (int64_t ci = blockIdx.x; ci < a.n_chunks; ci += gridDim.x) {
  const Chunk c = a.chunks[ci];
  const ReqDesc r = a.reqs[c.req];
  // The chunk's raw counts are contiguous: fetch them with aligned 16-byte
  // loads (over PCIe when they sit in pinned host memory, where wide requests
  // matter) into LDS.  The aligned span never leaves the pages (host) or the
  // arena and its slack (device) that hold the counts.
  const int npix = min(c.npix, kChunk * kMaxBatchBox * kMaxBatchBox);
  const uintptr_t beg = (uintptr_t)(r.adu + c.pix0);
  const uintptr_t b0 = beg & ~(uintptr_t)15;
  const int nvec = (int)((beg + 2 * (uintptr_t)npix - b0 + 15) >> 4);
  for (int v = t; v < nvec; v += kIngestBlock) sraw[v] = ((const uint4*)b0)[v];
  const uint16_t* raw = (const uint16_t*)sraw + ((beg - b0) >> 1);
  if (t < c.count) {
    const int k = c.first + t;
    const int64_t roi = r.roi_base + k;
    const int b = min(max((int)r.boxes[k], kMinBox), kMaxBatchBox);
    const int x = r.xc[k], y = r.yc[k];
    sbox[t] = b;
    sinv[t] = 1.0f / (float)b;
    sxc[t] = x;
    syc[t] = y;
    a.models[roi] = r.models[k];
    a.boxes[roi] = (int8_t)b;
    a.xc[roi] = x;
    a.yc[roi] = y;
  }
  __syncthreads();
  if (t == 0) {
    int sum = 0;
    for (int q = 0; q < c.count; ++q) {
      spre[q] = sum;
      sum += sbox[q] * sbox[q];
    }
    spre[c.count] = sum;
  }
  __syncthreads();
  if (t < c.count) a.offsets[r.roi_base + c.first + t] = r.pix_base + c.pix0 + spre[t];
  // Pixels past npix (the chunk's validated count) are never written: with
  // validated box sizes spre[count] == npix.
  const int tot = min(spre[c.count], npix);
  float* dst = a.data + r.pix_base + c.pix0;
  for (int p = t; p < tot; p += kIngestBlock) {
    // the pixel's ROI: the last q < count with spre[q] <= p
    int q = 0;
#pragma unroll
    for (int step = kChunk / 2; step > 0; step >>= 1)
      if (q + step < c.count && spre[q + step] <= p) q += step;
    const int b = sbox[q];
    const int pq = p - spre[q];
    // column j = floor(pq / b): (pq + 1/2) / b lies at least 1/(2b) >= 1/32
    // from an integer and pq < 256, so the Float32 product is exact enough
    const int j = (int)(((float)pq + 0.5f) * sinv[q]);
    const int i = pq - j * b;   // row inside the ROI
    int64_t m = (int64_t)(syc[q] - 1) + (int64_t)a.ny * (int64_t)(sxc[q] - 1 + j) + i;
    m = m < 0 ? 0 : (m >= a.map_size ? a.map_size - 1 : m);   // corners are validated; never read outside
    dst[p] = ((float)raw[p] - a.offset_map[m]) * a.gain_map[m];
  }
  __syncthreads();   // LDS is reused by the next chunk
  }
}

// Copies n floats from device memory to a host-visible buffer: staged in LDS,
// then written with 16-byte stores where the destination allows (posted
// writes over PCIe carry 16 bytes per lane instead of 4).
__device__ __forceinline__ void put_floats(float* dst, const float* src, int n, float* stage) {
  const int t = threadIdx.x;
  for (int e = t; e < n; e += kIngestBlock) stage[e] = src[e];
  __syncthreads();
  const int lead = (int)(((16 - ((uintptr_t)dst & 15)) & 15) >> 2);   // floats before a 16-byte boundary
  if (((uintptr_t)dst & 3) == 0 && lead < n) {
    const int nv = (n - lead) >> 2;
    for (int v = t; v < nv; v += kIngestBlock) {
      const float* s = stage + lead + 4 * v;
      *(float4*)(dst + lead + 4 * v) = make_float4(s[0], s[1], s[2], s[3]);
    }
    const int tail = lead + 4 * nv;
    if (t < lead) dst[t] = stage[t];
    if (t >= lead && t < lead + (n - tail)) dst[tail + t - lead] = stage[tail + t - lead];
  } else {
    for (int e = t; e < n; e += kIngestBlock) dst[e] = stage[e];
  }
  __syncthreads();
}

// Results of each chunk's ROIs into the request's output buffers (blocks stride
// over the chunks).
__global__ __launch_bounds__(kIngestBlock) void service_results(ResultArgs a) {
  __shared__ float stage[kChunk * kBatchStride];
  for (int64_t ci = blockIdx.x; ci < a.n_chunks; ci += gridDim.x) {
    const Chunk c = a.chunks[ci];
    const ReqDesc r = a.reqs[c.req];
    const int64_t s0 = r.roi_base + c.first;
    put_floats(r.params + (int64_t)c.first * kBatchStride, a.params + s0 * kBatchStride, c.count * kBatchStride,
               stage);
    put_floats(r.crlb + (int64_t)c.first * kBatchStride, a.crlb + s0 * kBatchStride, c.count * kBatchStride,
               stage);
    put_floats(r.llr + c.first, a.llr + s0, c.count, stage);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct Pending {
  int64_t ticket = 0;
  const int8_t* models = nullptr;   // device-visible addresses
  const int8_t* boxes = nullptr;
  const int32_t* xc = nullptr;
  const int32_t* yc = nullptr;
  const uint16_t* adu = nullptr;   // host address (pinned)
  const uint16_t* adu_dev = nullptr;   // its device address
  float* params = nullptr;
  float* crlb = nullptr;
  float* llr = nullptr;
  int64_t n = 0, npix = 0;
  int64_t counts[kNumShapes] = {};
  std::vector<int64_t> chunk_pix;   // pixel offset of each kChunk-ROI chunk
  double t_submit = 0, t_done = 0;
  bool done = false;
  int status = GMLE_OK;
};
using PendingP = std::shared_ptr<Pending>;

struct ShapeNode {
  int shape;
  int model, box;
  KernelInfo k;
  hipGraphNode_t fit = nullptr, fin = nullptr;
};

struct Slot {
  int index = 0;
  hipStream_t stream = nullptr;
  hipEvent_t start = nullptr;   // recorded before the graph (timing)
  hipEvent_t done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t n_ingest = nullptr, n_init = nullptr, n_scatter = nullptr, n_results = nullptr;
  hipGraphNode_t n_multi = nullptr;   // every shape served by the multi-shape kernel (Float32)
  std::vector<int> multi_shapes;
  std::vector<ShapeNode> shapes;      // shapes with a kernel of their own
  // device arena
  char* arena = nullptr;
  uint16_t* raw = nullptr;   // raw counts of the batch's requests, back to back
  float* data = nullptr;
  int8_t* models = nullptr;
  int8_t* boxes = nullptr;
  int32_t* xc = nullptr;
  int32_t* yc = nullptr;
  int64_t* offsets = nullptr;
  int32_t* order = nullptr;
  int32_t* cursor = nullptr;
  float* params = nullptr;
  float* crlb = nullptr;
  float* llr = nullptr;
  double* crlb64 = nullptr;
  double* llr64 = nullptr;
  // pinned staging of the batch's raw counts (staged ingest), grown on demand
  uint16_t* stage = nullptr;
  size_t stage_bytes = 0;
  // the batch's runs of requests with adjacent host counts ([first, end)) and
  // which requests arrive by DMA (dma ingest, or auto for long runs)
  std::vector<std::pair<size_t, size_t>> runs;
  std::vector<uint8_t> req_dma;
  // pinned tables (host and device addresses)
  ReqDesc* reqs_h = nullptr;
  ReqDesc* reqs_d = nullptr;
  Chunk* chunks_h = nullptr;
  Chunk* chunks_d = nullptr;
  int64_t max_chunks = 0;
  // batch in flight
  std::vector<PendingP> batch;
  bool busy = false;       // holds a batch (being launched or in flight)
  bool launched = false;   // its graph and completion event are enqueued
  double t_formed = 0, t_launched = 0;   // host clock: batch taken, graph launched
  int64_t batch_rois = 0;
  std::thread waiter;
};

}  // namespace
}  // namespace gmle

struct gmle_service {
  int device = 0;
  int precision = 0, iterations = 0;
  float psf[32], lower[24], upper[24], max_step[24];
  int32_t lanes[gmle::kNumShapes];
  size_t budget = 0;
  int64_t cap_rois = 0, cap_pix = 0;
  int max_box = 0;
  // calibration
  float* offset_map = nullptr;
  float* gain_map = nullptr;
  float* var_map = nullptr;
  int ny = 0, nx = 0;
  std::vector<std::unique_ptr<gmle::Slot>> slots;
  // queue
  std::mutex mu;
  std::condition_variable cv_dispatch;   // pending work or a free slot
  std::condition_variable cv_done;       // a request completed
  std::deque<gmle::PendingP> pending;
  std::unordered_map<int64_t, gmle::PendingP> tickets;
  int64_t next_ticket = 0;
  int64_t batches = 0;
  bool stop = false;
  std::thread dispatcher;
  // per-batch statistics (sums)
  double stat_batches = 0, stat_rois = 0, stat_reqs = 0, stat_host = 0, stat_gpu = 0, stat_seen = 0, stat_start = 0;
  double stat_gpu_max = 0, stat_seen_max = 0;
  int launch_error = GMLE_OK;
  int ingest_mode = gmle::INGEST_AUTO;     // how raw counts reach the GPU (IngestMode)
  int64_t batches_staged = 0;            // batches whose counts went through the staging buffer
  std::atomic<int64_t> raw_bytes{0}, raw_bytes_dma{0};   // raw counts ingested, and the part moved by DMA
  std::atomic<int64_t> dma_copies{0};                    // host-to-device copies of raw counts
  int64_t target_rois = gmle::kTargetBatchRois;   // a batch closes once it holds this many ROIs
  int64_t xfer_blocks = gmle::kXferBlocks;        // grid of the ingest / result kernels
  std::string error_msg;        // first failure of a batch (launch or completion)
};

namespace gmle {
namespace {

int64_t roi_bytes(int max_box, bool mixed) {
  // raw counts + pixels + models/boxes + corners + offsets + order + params/crlb + llr (+ FP64 outputs)
  return (int64_t)max_box * max_box * 6 + 2 + 8 + 8 + 4 + 2 * kBatchStride * 4 + 4 +
         (mixed ? kBatchStride * 8 + 8 : 0);
}

template <class T>
T* carve(char*& p, size_t count) {
  T* r = (T*)p;
  p += align256(count * sizeof(T));
  return r;
}

hipError_t add_kernel_node(hipGraphNode_t* node, hipGraph_t g, const std::vector<hipGraphNode_t>& deps,
                           const void* func, unsigned block, void* arg) {
  hipKernelNodeParams kp;
  std::memset(&kp, 0, sizeof(kp));
  void* args[1] = {arg};
  kp.func = const_cast<void*>(func);
  kp.gridDim = dim3(1);
  kp.blockDim = dim3(block);
  kp.kernelParams = args;
  return hipGraphAddKernelNode(node, g, deps.empty() ? nullptr : deps.data(), deps.size(), &kp);
}

hipError_t set_kernel_node(hipGraphExec_t e, hipGraphNode_t node, const void* func, unsigned grid, unsigned block,
                           void* arg) {
  hipKernelNodeParams kp;
  std::memset(&kp, 0, sizeof(kp));
  void* args[1] = {arg};
  kp.func = const_cast<void*>(func);
  kp.gridDim = dim3(grid);
  kp.blockDim = dim3(block);
  kp.kernelParams = args;
  return hipGraphExecKernelNodeSetParams(e, node, &kp);
}

FitArgs shape_args(const gmle_service* s, const Slot& sl, const ShapeNode& sh, int64_t first, int64_t count) {
  FitArgs a;
  fill_args(a, sh.model, s->psf + 8 * sh.model, s->lower + 6 * sh.model, s->upper + 6 * sh.model,
            s->max_step + 6 * sh.model, count, s->ny, s->nx, s->iterations);
  a.data = sl.data;
  a.var_map = s->var_map;
  a.x_corners = sl.xc;
  a.y_corners = sl.yc;
  a.params = sl.params;
  a.crlb = sl.crlb;
  a.llr = sl.llr;
  a.crlb64 = sl.crlb64;
  a.llr64 = sl.llr64;
  a.order = sl.order;
  a.offsets = sl.offsets;
  a.first = first;
  a.out_stride = kBatchStride;
  return a;
}

BatchPlan slot_plan(const Slot& sl, int64_t n) {
  BatchPlan p;
  std::memset(&p, 0, sizeof(p));
  p.models = sl.models;
  p.boxes = sl.boxes;
  p.n_rois = n;
  p.order = sl.order;
  p.cursor = sl.cursor;
  p.params = sl.params;
  p.crlb = sl.crlb;
  p.llr = sl.llr;
  p.crlb64 = sl.crlb64;
  p.llr64 = sl.llr64;
  return p;
}

int build_slot(gmle_service* s, Slot& sl, size_t slot_bytes) {
  hipError_t e;
  const bool mixed = s->precision == PREC_MIXED;
  const int64_t C = s->cap_rois;
  if ((e = hipMalloc(&sl.arena, slot_bytes)) != hipSuccess) return hip_fail(e, "hipMalloc service arena");
  char* p = sl.arena;
  sl.raw = carve<uint16_t>(p, (size_t)s->cap_pix);
  sl.data = carve<float>(p, (size_t)(s->cap_pix + kPixSlack));
  sl.models = carve<int8_t>(p, C);
  sl.boxes = carve<int8_t>(p, C);
  sl.xc = carve<int32_t>(p, C);
  sl.yc = carve<int32_t>(p, C);
  sl.offsets = carve<int64_t>(p, C);
  sl.order = carve<int32_t>(p, C);
  sl.cursor = carve<int32_t>(p, kNumShapes);
  sl.params = carve<float>(p, C * kBatchStride);
  sl.crlb = carve<float>(p, C * kBatchStride);
  sl.llr = carve<float>(p, C);
  if (mixed) {
    sl.crlb64 = carve<double>(p, C * kBatchStride);
    sl.llr64 = carve<double>(p, C);
  }
  if ((size_t)(p - sl.arena) > slot_bytes) return fail(GMLE_ERR_ARG, "service arena layout exceeds its budget");
  sl.max_chunks = (C + kChunk - 1) / kChunk + kMaxReqPerBatch;
  void* h = nullptr;
  void* d = nullptr;
  const unsigned fl = kPinnedFlags;
  if ((e = hipHostMalloc(&h, sizeof(ReqDesc) * kMaxReqPerBatch, fl)) != hipSuccess) return hip_fail(e, "hipHostMalloc");
  if ((e = hipHostGetDevicePointer(&d, h, 0)) != hipSuccess) return hip_fail(e, "hipHostGetDevicePointer");
  sl.reqs_h = (ReqDesc*)h;
  sl.reqs_d = (ReqDesc*)d;
  if ((e = hipHostMalloc(&h, sizeof(Chunk) * sl.max_chunks, fl)) != hipSuccess) return hip_fail(e, "hipHostMalloc");
  if ((e = hipHostGetDevicePointer(&d, h, 0)) != hipSuccess) return hip_fail(e, "hipHostGetDevicePointer");
  sl.chunks_h = (Chunk*)h;
  sl.chunks_d = (Chunk*)d;
  if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "hipStreamCreate");
  // The first host-to-device copy on a stream takes milliseconds of set-up:
  // pay it here (the pinned request table into the raw area) rather than in
  // the first batch whose counts arrive by DMA.
  const size_t warm = std::min(sizeof(ReqDesc) * (size_t)kMaxReqPerBatch, (size_t)s->cap_pix * sizeof(uint16_t));
  if ((e = hipMemcpyAsync(sl.raw, sl.reqs_h, warm, hipMemcpyHostToDevice, sl.stream)) != hipSuccess ||
      (e = hipStreamSynchronize(sl.stream)) != hipSuccess)
    return hip_fail(e, "first copy on the slot's stream");
  if ((e = hipEventCreate(&sl.start)) != hipSuccess) return hip_fail(e, "hipEventCreate");
  if ((e = hipEventCreate(&sl.done)) != hipSuccess) return hip_fail(e, "hipEventCreate");

  // The graph: ingest -> init -> scatter -> fits (-> final passes) -> results.
  if ((e = hipGraphCreate(&sl.graph, 0)) != hipSuccess) return hip_fail(e, "hipGraphCreate");
  IngestArgs ia;
  std::memset(&ia, 0, sizeof(ia));
  ResultArgs ra;
  std::memset(&ra, 0, sizeof(ra));
  BatchPlan bp = slot_plan(sl, 0);
  if ((e = add_kernel_node(&sl.n_ingest, sl.graph, {}, (const void*)&service_ingest, kIngestBlock, &ia)) != hipSuccess ||
      (e = add_kernel_node(&sl.n_init, sl.graph, {sl.n_ingest}, batch_init_kernel(), kBatchBlock, &bp)) != hipSuccess ||
      (e = add_kernel_node(&sl.n_scatter, sl.graph, {sl.n_init}, batch_scatter_kernel(), kBatchBlock, &bp)) != hipSuccess)
    return hip_fail(e, "hipGraphAddKernelNode");
  std::vector<hipGraphNode_t> tails;
  // Float32: one multi-shape node for every shape it can hold (with the
  // table's lanes where it has them, else 4); HIP runs a graph's kernel nodes
  // one after another, so one node for all shapes rather than one per shape.
  const bool use_multi = s->precision == PREC_FP32;
  if (use_multi) {
    MultiArgs m;
    std::memset(&m, 0, sizeof(m));
    if ((e = add_kernel_node(&sl.n_multi, sl.graph, {sl.n_scatter}, multi_kernel_scmos(), kWave, &m)) != hipSuccess)
      return hip_fail(e, "hipGraphAddKernelNode");
    tails.push_back(sl.n_multi);
  }
  for (int sh = 0; sh < kNumShapes; ++sh) {
    if (s->lanes[sh] <= 0) continue;
    if (use_multi && multi_eligible(sh / kNumBoxes, SCMOS, kMinBox + sh % kNumBoxes)) {
      sl.multi_shapes.push_back(sh);
      continue;
    }
    ShapeNode n;
    n.shape = sh;
    n.model = sh / kNumBoxes;
    n.box = kMinBox + sh % kNumBoxes;
    if (!find_kernels(n.model, SCMOS, n.box, s->lanes[sh], s->precision, &n.k))
      return fail(GMLE_ERR_UNSUPPORTED, "no kernel for model %d box %d lanes %d", n.model, n.box, s->lanes[sh]);
    FitArgs fa = shape_args(s, sl, n, 0, 0);
    if ((e = add_kernel_node(&n.fit, sl.graph, {sl.n_scatter}, n.k.fit, kWave, &fa)) != hipSuccess)
      return hip_fail(e, "hipGraphAddKernelNode");
    if (n.k.final_pass) {
      if ((e = add_kernel_node(&n.fin, sl.graph, {n.fit}, n.k.final_pass, kWave, &fa)) != hipSuccess)
        return hip_fail(e, "hipGraphAddKernelNode");
      tails.push_back(n.fin);
    } else {
      tails.push_back(n.fit);
    }
    sl.shapes.push_back(n);
  }
  std::sort(sl.multi_shapes.begin(), sl.multi_shapes.end(), [](int u, int v) {
    const int bu = kMinBox + u % kNumBoxes, bv = kMinBox + v % kNumBoxes;
    return bu * bu * (u / kNumBoxes + 4) > bv * bv * (v / kNumBoxes + 4);
  });
  if ((e = add_kernel_node(&sl.n_results, sl.graph, tails, (const void*)&service_results, kIngestBlock, &ra)) != hipSuccess)
    return hip_fail(e, "hipGraphAddKernelNode");
  if ((e = hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0)) != hipSuccess)
    return hip_fail(e, "hipGraphInstantiate");
  return GMLE_OK;
}

// Fills the slot's tables from its batch, updates the graph and launches it.
// Packs the batch's raw counts back to back into the slot's staging buffer
// (request r at its pix_base), split over the worker pool.  False when the
// buffer cannot be grown (the batch then goes direct).
bool stage_counts(Slot& sl, int64_t pix) {
  const size_t bytes = (size_t)pix * sizeof(uint16_t);
  if (bytes > sl.stage_bytes) {
    // the slot is free: its previous copy out of the buffer has completed
    if (sl.stage) (void)hipHostFree(sl.stage);
    sl.stage = nullptr;
    sl.stage_bytes = 0;
    const size_t cap = std::max(bytes, (size_t)32 << 20);
    if (hipHostMalloc((void**)&sl.stage, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      sl.stage = nullptr;
      return false;
    }
    sl.stage_bytes = cap;
  }
  struct Piece {
    char* dst;
    const char* src;
    size_t n;
  };
  std::vector<Piece> pieces;
  int64_t at = 0;
  for (const auto& qp : sl.batch) {
    const size_t n = (size_t)qp->npix * sizeof(uint16_t);
    char* dst = (char*)(sl.stage + at);
    const char* src = (const char*)qp->adu;
    for (size_t o = 0; o < n; o += kPackPiece) pieces.push_back({dst + o, src + o, std::min(kPackPiece, n - o)});
    at += qp->npix;
  }
  WorkerPool::get().parallel_for((int)pieces.size(), [&](int k) {
    std::memcpy(pieces[k].dst, pieces[k].src, pieces[k].n);
  });
  return true;
}

// Splits the batch into runs of requests whose raw counts follow each other in
// host memory (request r + 1's first count right after request r's last), and
// marks the requests whose run moves by DMA: every run in dma mode, runs of at
// least kDmaMinRunBytes in auto mode, none otherwise.  Requests are laid out
// back to back in the arena in batch order, so a run is also contiguous there
// and one copy moves it.  Returns the bytes marked for DMA.
int64_t plan_runs(Slot& sl, int mode) {
  const size_t n = sl.batch.size();
  sl.runs.clear();
  sl.req_dma.assign(n, 0);
  int64_t dma = 0;
  for (size_t r = 0; r < n;) {
    const char* end = (const char*)sl.batch[r]->adu + (size_t)sl.batch[r]->npix * sizeof(uint16_t);
    size_t r2 = r + 1;
    while (r2 < n && (const char*)sl.batch[r2]->adu == end) {
      end += (size_t)sl.batch[r2]->npix * sizeof(uint16_t);
      ++r2;
    }
    const int64_t bytes = (int64_t)(end - (const char*)sl.batch[r]->adu);
    if (mode == INGEST_DMA || (mode == INGEST_AUTO && bytes >= kDmaMinRunBytes)) {
      sl.runs.emplace_back(r, r2);
      for (size_t k = r; k < r2; ++k) sl.req_dma[k] = 1;
      dma += bytes;
    }
    r = r2;
  }
  return dma;
}

int launch_slot(gmle_service* s, Slot& sl) {
  hipError_t e;
  int64_t roi = 0, pix = 0, nch = 0;
  int64_t counts[kNumShapes] = {};
  int64_t batch_pix = 0;
  for (const auto& qp : sl.batch) batch_pix += qp->npix;
  int mode = s->ingest_mode;
  if (mode == INGEST_STAGED && !stage_counts(sl, batch_pix)) mode = INGEST_DIRECT;
  if (mode == INGEST_STAGED) ++s->batches_staged;
  const int64_t dma_bytes = plan_runs(sl, mode);
  s->raw_bytes += batch_pix * (int64_t)sizeof(uint16_t);
  s->raw_bytes_dma += mode == INGEST_STAGED ? batch_pix * (int64_t)sizeof(uint16_t) : dma_bytes;
  s->dma_copies += mode == INGEST_STAGED ? 1 : (int64_t)sl.runs.size();
  for (size_t r = 0; r < sl.batch.size(); ++r) {
    const Pending& q = *sl.batch[r];
    ReqDesc& d = sl.reqs_h[r];
    d.models = q.models;
    d.boxes = q.boxes;
    d.xc = q.xc;
    d.yc = q.yc;
    d.adu = (mode == INGEST_STAGED || sl.req_dma[r]) ? sl.raw + pix : q.adu_dev;
    d.params = q.params;
    d.crlb = q.crlb;
    d.llr = q.llr;
    d.roi_base = roi;
    d.pix_base = pix;
    for (size_t c = 0; c < q.chunk_pix.size(); ++c) {
      Chunk& ch = sl.chunks_h[nch++];
      ch.req = (int32_t)r;
      ch.first = (int32_t)(c * kChunk);
      ch.count = (int32_t)std::min<int64_t>(kChunk, q.n - (int64_t)c * kChunk);
      ch.pix0 = q.chunk_pix[c];
      ch.npix = (int32_t)((c + 1 < q.chunk_pix.size() ? q.chunk_pix[c + 1] : q.npix) - q.chunk_pix[c]);
    }
    for (int k = 0; k < kNumShapes; ++k) counts[k] += q.counts[k];
    roi += q.n;
    pix += q.npix;
  }
  IngestArgs ia{sl.reqs_d, sl.chunks_d, nch, s->offset_map, s->gain_map, s->ny, (int64_t)s->ny * s->nx, sl.data,
                sl.models, sl.boxes, sl.xc, sl.yc, sl.offsets};
  ResultArgs ra{sl.reqs_d, sl.chunks_d, nch, sl.params, sl.crlb, sl.llr};
  // Host-memory transfers wait on PCIe: a bounded grid keeps enough requests
  // in flight without parking every CU, so other slots' fits run beside them.
  // Counts already in the arena (staged, or every run by DMA) are converted
  // from HBM by a grid of up to one block per chunk.
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nch, s->xfer_blocks));
  const bool raw_on_device = mode == INGEST_STAGED || dma_bytes == batch_pix * (int64_t)sizeof(uint16_t);
  const unsigned gi =
      raw_on_device ? (unsigned)std::max<int64_t>(1, std::min<int64_t>(nch, kDeviceIngestBlocks)) : gx;
  BatchPlan bp = slot_plan(sl, roi);
  int64_t start = 0;
  for (int k = 0; k < kNumShapes; ++k) {
    bp.start[k] = (int32_t)start;
    bp.count[k] = (int32_t)counts[k];
    start += counts[k];
  }
  const unsigned gb = (unsigned)((roi + kBatchBlock - 1) / kBatchBlock);
  if ((e = set_kernel_node(sl.exec, sl.n_ingest, (const void*)&service_ingest, gi, kIngestBlock, &ia)) != hipSuccess ||
      (e = set_kernel_node(sl.exec, sl.n_init, batch_init_kernel(), gb, kBatchBlock, &bp)) != hipSuccess ||
      (e = set_kernel_node(sl.exec, sl.n_scatter, batch_scatter_kernel(), gb, kBatchBlock, &bp)) != hipSuccess ||
      (e = set_kernel_node(sl.exec, sl.n_results, (const void*)&service_results, gx, kIngestBlock, &ra)) != hipSuccess)
    return hip_fail(e, "hipGraphExecKernelNodeSetParams");
  if (sl.n_multi) {
    MultiArgs m;
    std::memset(&m, 0, sizeof(m));
    for (int model = 0; model < 4; ++model) {
      ShapeNode n;
      n.model = model;
      m.fa[model] = shape_args(s, sl, n, 0, 0);
    }
    for (int sh : sl.multi_shapes) {   // heaviest shapes' tiles first
      if (counts[sh] == 0) continue;
      const int q = m.n_entries++;
      m.shape[q] = sh;
      m.lanes[q] = multi_lanes(sh / kNumBoxes, SCMOS, kMinBox + sh % kNumBoxes, s->lanes[sh]);
      m.first[q] = bp.start[sh];
      m.count[q] = bp.count[sh];
    }
    multi_plan_blocks(m);
    const int32_t blocks = m.block_start[m.n_entries];
    if ((e = set_kernel_node(sl.exec, sl.n_multi, multi_kernel_scmos(), (unsigned)std::max(1, blocks), kWave, &m)) !=
        hipSuccess)
      return hip_fail(e, "hipGraphExecKernelNodeSetParams");
  }
  for (const ShapeNode& n : sl.shapes) {
    const int64_t cnt = counts[n.shape];
    FitArgs fa = shape_args(s, sl, n, bp.start[n.shape], cnt);
    const unsigned gf = (unsigned)std::max<int64_t>(1, (cnt + n.k.rois_per_block - 1) / n.k.rois_per_block);
    if ((e = set_kernel_node(sl.exec, n.fit, n.k.fit, gf, kWave, &fa)) != hipSuccess)
      return hip_fail(e, "hipGraphExecKernelNodeSetParams");
    if (n.fin) {
      const unsigned gp =
          (unsigned)std::max<int64_t>(1, (cnt + n.k.final_rois_per_block - 1) / n.k.final_rois_per_block);
      if ((e = set_kernel_node(sl.exec, n.fin, n.k.final_pass, gp, kWave, &fa)) != hipSuccess)
        return hip_fail(e, "hipGraphExecKernelNodeSetParams");
    }
  }
  if ((e = hipEventRecord(sl.start, sl.stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  // Raw counts over PCIe by DMA (unless ingest reads them in place), then the graph.
  if (mode == INGEST_STAGED) {
    if ((e = hipMemcpyAsync(sl.raw, sl.stage, (size_t)batch_pix * sizeof(uint16_t), hipMemcpyHostToDevice,
                            sl.stream)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync (staged counts)");
  }
  for (const auto& run : sl.runs) {
    const Pending& q0 = *sl.batch[run.first];
    const int64_t p0 = sl.reqs_h[run.first].pix_base;
    const int64_t p1 = sl.reqs_h[run.second - 1].pix_base + sl.batch[run.second - 1]->npix;
    if ((e = hipMemcpyAsync(sl.raw + p0, q0.adu, (size_t)(p1 - p0) * sizeof(uint16_t), hipMemcpyHostToDevice,
                            sl.stream)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync (request counts)");
  }
  if ((e = hipGraphLaunch(sl.exec, sl.stream)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  if ((e = hipEventRecord(sl.done, sl.stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return GMLE_OK;
}

void finish(gmle_service* s, std::vector<PendingP>& batch, int status) {
  const double t = now_s();
  std::lock_guard<std::mutex> lock(s->mu);
  for (auto& q : batch) {
    q->t_done = t;
    q->done = true;
    q->status = status;
  }
  batch.clear();
}

void dispatcher_main(gmle_service* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  for (;;) {
    Slot* free_slot = nullptr;
    s->cv_dispatch.wait(lock, [&] {
      if (s->stop) return true;
      if (s->pending.empty()) return false;
      for (auto& sl : s->slots)
        if (!sl->busy) { free_slot = sl.get(); return true; }
      return false;
    });
    if (s->stop) return;
    Slot& sl = *free_slot;
    int64_t rois = 0, pix = 0;
    while (!s->pending.empty() && (int)sl.batch.size() < kMaxReqPerBatch) {
      const Pending& q = *s->pending.front();
      if (rois + q.n > s->cap_rois || pix + q.npix > s->cap_pix) break;
      if (rois > 0 && rois + q.n > s->target_rois) break;
      rois += q.n;
      pix += q.npix;
      sl.batch.push_back(s->pending.front());
      s->pending.pop_front();
    }
    sl.busy = true;
    sl.t_formed = now_s();
    sl.batch_rois = rois;
    ++s->batches;
    lock.unlock();
    const int rc = launch_slot(s, sl);
    sl.t_launched = now_s();
    lock.lock();
    if (rc != GMLE_OK) {
      s->launch_error = rc;
      if (s->error_msg.empty()) {
        char buf[512];
        gmle_last_error(buf, sizeof(buf));
        s->error_msg = buf;
      }
      lock.unlock();
      finish(s, sl.batch, rc);
      lock.lock();
      sl.busy = false;
      s->cv_done.notify_all();
      continue;
    }
    sl.launched = true;
    s->cv_dispatch.notify_all();   // wakes the slot's waiter
  }
}

void waiter_main(gmle_service* s, Slot* sl) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  for (;;) {
    s->cv_dispatch.wait(lock, [&] { return s->stop || sl->launched; });
    if (!sl->launched) return;   // stopping, nothing in flight
    lock.unlock();
    // Poll rather than block: a blocking wait's wake-up costs more than a
    // small batch takes on the GPU.
    const double t_poll = now_s();
    auto poll = [&](hipEvent_t ev) {
      hipError_t r;
      while ((r = hipEventQuery(ev)) == hipErrorNotReady) {
        if (now_s() - t_poll < kSpinSeconds) {
          std::this_thread::yield();
        } else {
          std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
      }
      return r;
    };
    hipError_t e = poll(sl->start);
    const double t_started = now_s();
    if (e == hipSuccess) e = poll(sl->done);
    const double t_seen = now_s();
    float gpu_ms = 0.f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&gpu_ms, sl->start, sl->done);
    std::vector<PendingP> b;
    lock.lock();
    if (e != hipSuccess && s->error_msg.empty())
      s->error_msg = std::string("batch completion: ") + hipGetErrorString(e);
    b.swap(sl->batch);
    s->stat_batches += 1;
    s->stat_rois += (double)sl->batch_rois;
    s->stat_reqs += (double)b.size();
    s->stat_host += sl->t_launched - sl->t_formed;
    s->stat_gpu += gpu_ms * 1e-3;
    s->stat_seen += t_seen - sl->t_launched;
    s->stat_start += t_started - sl->t_launched;
    s->stat_gpu_max = std::max(s->stat_gpu_max, gpu_ms * 1e-3);
    s->stat_seen_max = std::max(s->stat_seen_max, t_seen - sl->t_launched);
    lock.unlock();
    finish(s, b, e == hipSuccess ? GMLE_OK : GMLE_ERR_HIP);
    lock.lock();
    sl->launched = false;
    sl->busy = false;
    s->cv_done.notify_all();
    s->cv_dispatch.notify_all();
  }
}

int make_pending(gmle_service* s, const gmle_request* r, PendingP& out) {
  if (!r) return fail(GMLE_ERR_ARG, "request must not be NULL");
  const int64_t n = r->n_rois;
  if (n <= 0) return fail(GMLE_ERR_ARG, "request holds no ROIs");
  if (n > s->cap_rois) return fail(GMLE_ERR_ARG, "request of %lld ROIs exceeds a slot's %lld", (long long)n,
                                   (long long)s->cap_rois);
  if (!s->var_map) return fail(GMLE_ERR_ARG, "service has no calibration maps (gmle_service_calibrate)");
  auto q = std::make_shared<Pending>();
  q->n = n;
  const void* d;
  if (!pinned_dev(r->models, n, &d)) return fail(GMLE_ERR_ARG, "request models are not in gmle_host_alloc memory");
  q->models = (const int8_t*)d;
  if (!pinned_dev(r->boxes, n, &d)) return fail(GMLE_ERR_ARG, "request boxes are not in gmle_host_alloc memory");
  q->boxes = (const int8_t*)d;
  if (!pinned_dev(r->x_corners, 4 * n, &d) ) return fail(GMLE_ERR_ARG, "request corners are not in gmle_host_alloc memory");
  q->xc = (const int32_t*)d;
  if (!pinned_dev(r->y_corners, 4 * n, &d)) return fail(GMLE_ERR_ARG, "request corners are not in gmle_host_alloc memory");
  q->yc = (const int32_t*)d;
  q->chunk_pix.reserve((size_t)((n + kChunk - 1) / kChunk));
  int64_t npix = 0;
  for (int64_t k = 0; k < n; ++k) {
    if (k % kChunk == 0) q->chunk_pix.push_back(npix);
    const int m = r->models[k], b = r->boxes[k];
    const int sh = shape_index(m, b);
    if (sh < 0 || s->lanes[sh] <= 0)
      return fail(GMLE_ERR_UNSUPPORTED, "ROI %lld: model %d box %d is not served", (long long)(k + 1), m, b);
    const int cx = r->x_corners[k], cy = r->y_corners[k];
    if (cx < 1 || cy < 1 || cx > s->nx - b + 1 || cy > s->ny - b + 1)
      return fail(GMLE_ERR_BOUNDS, "ROI %lld at corner (x=%d, y=%d) lies outside the %dx%d sensor", (long long)(k + 1),
                  cx, cy, s->ny, s->nx);
    ++q->counts[sh];
    npix += (int64_t)b * b;
  }
  if (npix > s->cap_pix) return fail(GMLE_ERR_ARG, "request pixels exceed a slot's capacity");
  q->npix = npix;
  if (!pinned_dev(r->adu, 2 * npix, &d)) return fail(GMLE_ERR_ARG, "request pixels are not in gmle_host_alloc memory");
  q->adu = r->adu;
  q->adu_dev = (const uint16_t*)d;
  if (!pinned_dev(r->params, 4 * kBatchStride * n, &d)) return fail(GMLE_ERR_ARG, "params not in gmle_host_alloc memory");
  q->params = (float*)d;
  if (!pinned_dev(r->crlb, 4 * kBatchStride * n, &d)) return fail(GMLE_ERR_ARG, "crlb not in gmle_host_alloc memory");
  q->crlb = (float*)d;
  if (!pinned_dev(r->llr, 4 * n, &d)) return fail(GMLE_ERR_ARG, "llr not in gmle_host_alloc memory");
  q->llr = (float*)d;
  out = q;
  return GMLE_OK;
}

int64_t submit_pending(gmle_service* s, PendingP q) {
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->stop) return fail(GMLE_ERR_ARG, "service is shutting down");
  q->ticket = s->next_ticket++;
  q->t_submit = now_s();
  s->tickets[q->ticket] = q;
  s->pending.push_back(q);
  s->cv_dispatch.notify_all();
  return q->ticket;
}

void teardown(gmle_service* s) {
  {
    std::unique_lock<std::mutex> lock(s->mu);
    // let queued and in-flight work finish
    s->cv_done.wait(lock, [&] {
      if (!s->pending.empty()) return false;
      for (auto& sl : s->slots)
        if (sl->busy) return false;
      return true;
    });
    s->stop = true;
    s->cv_dispatch.notify_all();
  }
  if (s->dispatcher.joinable()) s->dispatcher.join();
  for (auto& sl : s->slots)
    if (sl->waiter.joinable()) sl->waiter.join();
  (void)hipSetDevice(s->device);
  for (auto& sl : s->slots) {
    if (sl->exec) (void)hipGraphExecDestroy(sl->exec);
    if (sl->graph) (void)hipGraphDestroy(sl->graph);
    if (sl->done) (void)hipEventDestroy(sl->done);
    if (sl->start) (void)hipEventDestroy(sl->start);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
    if (sl->reqs_h) (void)hipHostFree(sl->reqs_h);
    if (sl->chunks_h) (void)hipHostFree(sl->chunks_h);
    if (sl->stage) (void)hipHostFree(sl->stage);
    if (sl->arena) (void)hipFree(sl->arena);
  }
  if (s->offset_map) (void)hipFree(s->offset_map);
  if (s->gain_map) (void)hipFree(s->gain_map);
  if (s->var_map) (void)hipFree(s->var_map);
}

}  // namespace
}  // namespace gmle

using namespace gmle;

extern "C" {

void* gmle_host_alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  void* h = nullptr;
  void* d = nullptr;
  if (hipHostMalloc(&h, bytes, kPinnedFlags) != hipSuccess) {
    fail(GMLE_ERR_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
    return nullptr;
  }
  if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
    (void)hipHostFree(h);
    fail(GMLE_ERR_HIP, "hipHostGetDevicePointer failed");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_pin_mu);
  g_pinned[(uintptr_t)h] = PinnedBlock{(char*)h, (char*)d, bytes};
  return h;
}

int gmle_host_free(void* p) {
  if (!p) return GMLE_OK;
  {
    std::lock_guard<std::mutex> lock(g_pin_mu);
    auto it = g_pinned.find((uintptr_t)p);
    if (it == g_pinned.end()) return fail(GMLE_ERR_ARG, "pointer was not returned by gmle_host_alloc");
    g_pinned.erase(it);
  }
  const hipError_t e = hipHostFree(p);
  return e == hipSuccess ? GMLE_OK : hip_fail(e, "hipHostFree");
}

int gmle_service_create(int device, size_t hbm_budget, int slots, int precision, int iterations,
                        const float* psf, const float* lower, const float* upper, const float* max_step,
                        const int32_t* lanes, gmle_service** out) {
  if (!out) return fail(GMLE_ERR_ARG, "out must not be NULL");
  *out = nullptr;
  if (!psf || !lower || !upper || !max_step || !lanes) return fail(GMLE_ERR_ARG, "NULL argument");
  if (slots < 1 || slots > 16) return fail(GMLE_ERR_ARG, "slots must be 1..16, got %d", slots);
  if (precision != PREC_FP32 && precision != PREC_MIXED) return fail(GMLE_ERR_ARG, "unknown precision %d", precision);
  if (iterations < 0) return fail(GMLE_ERR_ARG, "iterations must be >= 0");
  int max_box = 0, served = 0;
  for (int sh = 0; sh < kNumShapes; ++sh) {
    if (lanes[sh] <= 0) continue;
    const int m = sh / kNumBoxes, b = kMinBox + sh % kNumBoxes;
    if (!lanes_ok(b, lanes[sh]))
      return fail(GMLE_ERR_UNSUPPORTED, "no kernel with %d lanes per ROI for box size %d", lanes[sh], b);
    if (m == XYNB && !(psf[8 * m] > 0.f)) return fail(GMLE_ERR_ARG, "PSF sigma must be > 0");
    if (m == ASTIG && !(psf[24] > 0.f && psf[25] > 0.f && psf[31] != 0.f && psf[30] != 0.f))
      return fail(GMLE_ERR_ARG, "astigmatic PSF needs sigma_x0, sigma_y0 > 0 and nonzero gamma, d");
    max_box = std::max(max_box, b);
    ++served;
  }
  if (!served) return fail(GMLE_ERR_ARG, "the lanes table serves no shape");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(GMLE_ERR_NODEVICE, "no HIP device available");
  }
  if (device < 0 || device >= ndev) return fail(GMLE_ERR_ARG, "device %d out of range", device);
  std::unique_ptr<gmle_service> s(new gmle_service());
  s->device = device;
  s->precision = precision;
  s->iterations = iterations;
  std::memcpy(s->psf, psf, sizeof(s->psf));
  std::memcpy(s->lower, lower, sizeof(s->lower));
  std::memcpy(s->upper, upper, sizeof(s->upper));
  std::memcpy(s->max_step, max_step, sizeof(s->max_step));
  std::memcpy(s->lanes, lanes, sizeof(s->lanes));
  s->budget = hbm_budget;
  s->max_box = max_box;
  // How raw counts reach the GPU (ingest modes above): auto unless
  // GMLE_SERVICE_INGEST names another.
  if (const char* ingest = std::getenv("GMLE_SERVICE_INGEST")) {
    const std::string m(ingest);
    s->ingest_mode = m == "staged" ? INGEST_STAGED : m == "dma" ? INGEST_DMA : m == "direct" ? INGEST_DIRECT
                                                                                            : INGEST_AUTO;
  }
  // Batch size target (ROIs): smaller batches let one slot's ingest overlap
  // another's fit; larger ones amortize the per-batch latency floor.
  if (const char* tb = std::getenv("GMLE_SERVICE_BATCH_ROIS")) {
    const long long v = std::atoll(tb);
    if (v >= 256 && v <= (1LL << 24)) s->target_rois = v;
  }
  if (const char* xb = std::getenv("GMLE_SERVICE_XFER_BLOCKS")) {
    const long long v = std::atoll(xb);
    if (v >= 1 && v <= (1LL << 20)) s->xfer_blocks = v;
  }
  (void)hipSetDevice(device);
  s->slots.clear();
  for (int k = 0; k < slots; ++k) {
    s->slots.emplace_back(new Slot());
    s->slots.back()->index = k;
  }
  *out = s.release();
  return GMLE_OK;
}

int gmle_service_calibrate(gmle_service* s, const float* offset, const float* gain, const float* var, int ny, int nx) {
  if (!s || !offset || !gain || !var) return fail(GMLE_ERR_ARG, "NULL argument");
  if (ny < kMinBox || nx < kMinBox) return fail(GMLE_ERR_ARG, "sensor %dx%d is smaller than one ROI", ny, nx);
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->var_map) return fail(GMLE_ERR_ARG, "service is already calibrated");
  }
  const size_t map_bytes = align256((size_t)ny * nx * sizeof(float));
  const bool mixed = s->precision == PREC_MIXED;
  const int nslots = (int)s->slots.size();
  if (s->budget <= 3 * map_bytes) return fail(GMLE_ERR_NOMEM, "HBM budget does not hold the sensor maps");
  const size_t slot_bytes = (s->budget - 3 * map_bytes) / nslots;
  const int64_t per_roi = roi_bytes(s->max_box, mixed);
  // 256-byte alignment of 13 arena pieces, the grouping cursors and the pixel slack
  int64_t cap = ((int64_t)slot_bytes - 16 * 256 - kNumShapes * 4 - kPixSlack * 4) / per_roi;
  cap = std::min<int64_t>(cap, INT32_MAX / 2);
  if (cap < 1) return fail(GMLE_ERR_NOMEM, "HBM budget leaves no room for ROIs");
  s->cap_rois = cap;
  s->cap_pix = cap * s->max_box * s->max_box;
  s->ny = ny;
  s->nx = nx;
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  const size_t bytes = (size_t)ny * nx * sizeof(float);
  float* m[3] = {};
  const float* src[3] = {offset, gain, var};
  for (int k = 0; k < 3; ++k) {
    if ((e = hipMalloc(&m[k], map_bytes)) != hipSuccess) return hip_fail(e, "hipMalloc sensor map");
    if ((e = hipMemcpy(m[k], src[k], bytes, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "hipMemcpy");
  }
  s->offset_map = m[0];
  s->gain_map = m[1];
  s->var_map = m[2];
  for (auto& sl : s->slots) {
    int rc = build_slot(s, *sl, slot_bytes);
    if (rc != GMLE_OK) return rc;
  }
  s->dispatcher = std::thread(dispatcher_main, s);
  for (auto& sl : s->slots) sl->waiter = std::thread(waiter_main, s, sl.get());
  return GMLE_OK;
}

int64_t gmle_service_submit(gmle_service* s, const gmle_request* r) {
  if (!s) return fail(GMLE_ERR_ARG, "NULL service");
  PendingP q;
  int rc = make_pending(s, r, q);
  if (rc != GMLE_OK) return rc;
  return submit_pending(s, q);
}

int gmle_service_wait(gmle_service* s, int64_t ticket, double timeout_s) {
  if (!s) return fail(GMLE_ERR_ARG, "NULL service");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return fail(GMLE_ERR_ARG, "unknown ticket %lld", (long long)ticket);
  PendingP q = it->second;
  const auto ok = s->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), [&] { return q->done; });
  if (!ok) return 1;
  s->tickets.erase(ticket);
  if (q->status != GMLE_OK) return fail(q->status, "a batch failed: %s", s->error_msg.c_str());
  return GMLE_OK;
}

int gmle_service_replay(gmle_service* s, int64_t n, const gmle_request* reqs, const double* arrival_s,
                        double* latency_s, double* elapsed_s) {
  if (!s || !reqs || !arrival_s || !latency_s || n <= 0) return fail(GMLE_ERR_ARG, "bad replay arguments");
  std::vector<PendingP> qs((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    int rc = make_pending(s, &reqs[i], qs[(size_t)i]);
    if (rc != GMLE_OK) return rc;
    if (i && arrival_s[i] < arrival_s[i - 1]) return fail(GMLE_ERR_ARG, "arrival times must ascend");
  }
  const double t0 = now_s();
  for (int64_t i = 0; i < n; ++i) {
    const double due = t0 + arrival_s[i];
    // sleep through long gaps, spin the last stretch
    for (double left = due - now_s(); left > 3e-4; left = due - now_s())
      std::this_thread::sleep_for(std::chrono::duration<double>(left - 2e-4));
    while (now_s() < due) {
    }
    PendingP& q = qs[(size_t)i];
    std::lock_guard<std::mutex> lock(s->mu);
    q->ticket = s->next_ticket++;
    q->t_submit = due;
    s->pending.push_back(q);
    s->cv_dispatch.notify_all();
  }
  int status = GMLE_OK;
  {
    std::unique_lock<std::mutex> lock(s->mu);
    for (auto& q : qs) {
      s->cv_done.wait(lock, [&] { return q->done; });
      if (q->status != GMLE_OK) status = q->status;
    }
  }
  const double t1 = now_s();
  double last = t0;
  for (int64_t i = 0; i < n; ++i) {
    latency_s[i] = qs[(size_t)i]->t_done - qs[(size_t)i]->t_submit;
    last = std::max(last, qs[(size_t)i]->t_done);
  }
  if (elapsed_s) *elapsed_s = last - t0;
  (void)t1;
  if (status == GMLE_OK) return GMLE_OK;
  std::lock_guard<std::mutex> lock(s->mu);
  return fail(status, "a batch failed: %s", s->error_msg.c_str());
}

int gmle_service_info(gmle_service* s, int64_t* slot_rois, int64_t* slot_pixels, int* slots, int64_t* batches) {
  if (!s) return fail(GMLE_ERR_ARG, "NULL service");
  std::lock_guard<std::mutex> lock(s->mu);
  if (slot_rois) *slot_rois = s->cap_rois;
  if (slot_pixels) *slot_pixels = s->cap_pix;
  if (slots) *slots = (int)s->slots.size();
  if (batches) *batches = s->batches;
  return GMLE_OK;
}

int gmle_service_stats(gmle_service* s, double* out, int n) {
  if (!s || !out) return fail(GMLE_ERR_ARG, "NULL argument");
  std::lock_guard<std::mutex> lock(s->mu);
  const double b = s->stat_batches > 0 ? s->stat_batches : 1.0;
  const double raw = (double)s->raw_bytes.load();
  const double v[12] = {s->stat_batches, s->stat_rois / b, s->stat_reqs / b, s->stat_host / b * 1e6,
                        s->stat_gpu / b * 1e6, s->stat_seen / b * 1e6, s->stat_start / b * 1e6,
                        s->stat_gpu_max * 1e6, s->stat_seen_max * 1e6, (double)s->batches_staged,
                        raw > 0 ? (double)s->raw_bytes_dma.load() / raw : 0.0, (double)s->dma_copies.load() / b};
  for (int k = 0; k < n && k < 12; ++k) out[k] = v[k];
  return GMLE_OK;
}

int gmle_service_destroy(gmle_service* s) {
  if (!s) return GMLE_OK;
  teardown(s);
  delete s;
  return GMLE_OK;
}

}  // extern "C"
