// gmle_hostpipe.hpp -- the host-array entry points' upload pipeline.
//
// GaussMLE.jl's GPU branch copies each batch to the device, launches, waits
// and copies back, one batch after another (src/interface.jl:220-262), from
// pageable Julia arrays.  Here the caller's arrays are streamed in chunks of
// at most kStageBytes through a ring of kStages pinned buffers:
//   * worker threads copy chunk k from the caller's array into a free pinned
//     buffer, checking every value for NaN / Inf in the same pass (the
//     reference's validate_fit_input, src/interface.jl:26-29);
//   * the chunk's DMA upload runs on the pipeline's copy stream while the
//     workers fill the next buffer;
//   * the chunk's fit is enqueued on the work stream as soon as its upload
//     has landed, so uploads, host copies and fits overlap.
// Results come back through a pinned output buffer and reach the caller's
// arrays only after every chunk passed the check, so a rejected call leaves
// them untouched, as the reference's up-front validation does.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gaussmle.h"
#include "gmle_host.hpp"

namespace gmle {

constexpr size_t kStageBytes = (size_t)32 << 20;   // one pinned staging buffer
constexpr int kStages = 3;

// A small persistent pool: parallel_for(n, fn) runs fn(0..n-1) on the workers
// and the calling thread, and returns when all are done.  One parallel_for
// runs at a time.
class WorkerPool {
 public:
  static WorkerPool& get() {
    static WorkerPool pool;
    return pool;
  }
  int threads() const { return (int)workers_.size() + 1; }

  void parallel_for(int n, const std::function<void(int)>& fn) {
    std::lock_guard<std::mutex> serial(run_mu_);
    if (n <= 0) return;
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      pending_ = n;
      ++generation_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  WorkerPool() {
    int t = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("GMLE_HOST_THREADS")) t = std::atoi(e);
    t = std::max(1, std::min(t, 8));
    for (int k = 0; k + 1 < t; ++k) workers_.emplace_back([this] { loop(); });
  }
  ~WorkerPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  void drain() {
    for (;;) {
      const int i = next_.fetch_add(1);
      if (i >= n_) return;
      (*fn_)(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_cv_.notify_all();
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
        if (stop_) return;
        seen = generation_;
      }
      drain();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, pending_ = 0;
  uint64_t generation_ = 0;
  bool stop_ = false;
};

// dst[0..n) = src[0..n), split over the pool; false when any value is NaN or
// infinite.
inline bool copy_finite(float* dst, const float* src, size_t n) {
  WorkerPool& pool = WorkerPool::get();
  const int parts = (int)std::min<size_t>((size_t)pool.threads() * 2, std::max<size_t>(1, n / 65536));
  std::atomic<int> bad{0};
  pool.parallel_for(parts, [&](int p) {
    const size_t a = n * (size_t)p / (size_t)parts, b = n * (size_t)(p + 1) / (size_t)parts;
    int ok = 1;
    for (size_t i = a; i < b; ++i) {
      const float x = src[i];
      dst[i] = x;
      ok &= (std::fabs(x) <= FLT_MAX);
    }
    if (!ok) bad.store(1);
  });
  return bad.load() == 0;
}

// true when src[0..n) holds no NaN or infinite value (split over the pool).
inline bool all_finite(const float* src, size_t n) {
  WorkerPool& pool = WorkerPool::get();
  const int parts = (int)std::min<size_t>((size_t)pool.threads() * 2, std::max<size_t>(1, n / 65536));
  std::atomic<int> bad{0};
  pool.parallel_for(parts, [&](int p) {
    const size_t a = n * (size_t)p / (size_t)parts, b = n * (size_t)(p + 1) / (size_t)parts;
    int ok = 1;
    for (size_t i = a; i < b; ++i) ok &= (std::fabs(src[i]) <= FLT_MAX);
    if (!ok) bad.store(1);
  });
  return bad.load() == 0;
}

// memcpy split over the pool.
inline void copy_parallel(void* dst, const void* src, size_t bytes) {
  WorkerPool& pool = WorkerPool::get();
  const int parts = (int)std::min<size_t>((size_t)pool.threads() * 2, std::max<size_t>(1, bytes / (256 << 10)));
  pool.parallel_for(parts, [&](int p) {
    const size_t a = bytes * (size_t)p / (size_t)parts, b = bytes * (size_t)(p + 1) / (size_t)parts;
    std::memcpy((char*)dst + a, (const char*)src + a, b - a);
  });
}

// Per device: the pinned ring, its events, the copy stream and the pinned
// output buffer.  Used under the device's host-path mutex.
struct HostPipe {
  bool ready = false;
  void* stage[kStages] = {};
  hipEvent_t uploaded[kStages] = {};
  hipStream_t up = nullptr;
  void* out = nullptr;
  size_t out_cap = 0;

  hipError_t init() {
    if (ready) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking);
    for (int k = 0; k < kStages && e == hipSuccess; ++k) {
      e = hipHostMalloc(&stage[k], kStageBytes, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&uploaded[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) ready = true;
    return e;
  }
  hipError_t ensure_out(size_t bytes) {
    if (out_cap >= bytes) return hipSuccess;
    if (out) (void)hipHostFree(out);
    out = nullptr;
    out_cap = 0;
    const hipError_t e = hipHostMalloc(&out, bytes, hipHostMallocDefault);
    if (e == hipSuccess) out_cap = bytes;
    return e;
  }
};

// Streams src[0..cuts[nchunks-1]) to d_dst in the chunks [cuts[k-1], cuts[k])
// (floats; each chunk at most kStageBytes) through the pinned ring, checking
// every value; after chunk k is on the device, launch(k) enqueues its work on
// `work` and returns a GMLE code.  Returns GMLE_OK, GMLE_ERR_DATA when a value
// is not finite (nothing is launched for that chunk or after it), or the
// first error.  Waits for every upload it issued before returning.
inline int pipe_upload(HostPipe& P, const float* src, float* d_dst, const size_t* cuts, int nchunks,
                       hipStream_t work, const std::function<int(int)>& launch) {
  hipError_t e = hipSuccess;
  int rc = GMLE_OK;
  size_t a = 0;
  int issued = 0;
  for (int k = 0; k < nchunks && rc == GMLE_OK; ++k) {
    const int slot = k % kStages;
    const size_t b = cuts[k];
    if (k >= kStages && (e = hipEventSynchronize(P.uploaded[slot])) != hipSuccess) {
      rc = hip_fail(e, "hipEventSynchronize (upload ring)");
      break;
    }
    float* st = (float*)P.stage[slot];
    if (!copy_finite(st, src + a, b - a)) {
      rc = fail(GMLE_ERR_DATA, "Input data contains NaN or Inf values");
      break;
    }
    if ((e = hipMemcpyAsync(d_dst + a, st, (b - a) * sizeof(float), hipMemcpyHostToDevice, P.up)) != hipSuccess ||
        (e = hipEventRecord(P.uploaded[slot], P.up)) != hipSuccess) {
      rc = hip_fail(e, "hipMemcpyAsync (upload)");
      break;
    }
    ++issued;
    if ((e = hipStreamWaitEvent(work, P.uploaded[slot], 0)) != hipSuccess) {
      rc = hip_fail(e, "hipStreamWaitEvent");
      break;
    }
    rc = launch(k);
    a = b;
  }
  // the ring is reused by the next call: every issued upload must be done
  if (issued && (e = hipStreamSynchronize(P.up)) != hipSuccess && rc == GMLE_OK)
    rc = hip_fail(e, "hipStreamSynchronize (upload)");
  return rc;
}

}  // namespace gmle
