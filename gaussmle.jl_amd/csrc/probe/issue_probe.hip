// issue_probe.hip -- what one gfx950 SIMD issues per clock: FP32 FMA,
// v_rcp_f32, v_exp_f32, DPP moves and v_mfma_f32_4x4x1f32, alone and mixed,
// at 1-4 resident waves per SIMD.  The fit kernels are bound by VALU issue,
// so these rates are the ceiling their roofline fraction is read against.
//
// Method.  Every variant runs W waves per SIMD (a grid of W blocks of 256
// threads per CU: one wave per SIMD per block), each wave looping over C = 8
// independent chains, with the loop body repeated 8 times per trip so the
// loop's branch costs little.  Each wave reads s_memtime (shader clock) and
// s_memrealtime (100 MHz) around its loop and writes the two counts to
// cyc[wave] / rt[wave]; the median wave gives
//   clocks per SIMD instruction = ticks / (trips x bodies per trip x
//                                          instructions per body x W)
// and the clock rate itself.  Every buffer is sized by the largest grid; each
// thread writes out[gid] only.
//
//   gaussmle.jl_amd/lib/gmle_issue_probe      (built by build.py)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kC = 8;          // independent chains per wave
constexpr int kU = 8;          // loop body repeats per trip (keeps the branch out of the rates)
constexpr int kIters = 1000;   // loop trips per wave

enum Op { FMA, RCP, EXP2, RCP_FMA, DPP_FMA, PIXEL, MFMA, MFMA_FMA };

// VALU (or MFMA) instructions one repeat of the loop body issues per wave, as
// the compiler emits them for each op (checked against the ISA of this file).
constexpr int per_body(Op op) {
  switch (op) {
    case FMA: return kC;
    case RCP: return kC;
    case EXP2: return 2 * kC;       // v_exp_f32 + the subtraction that keeps it in range
    case RCP_FMA: return 2 * kC;    // v_rcp_f32 + the FMA that consumes it
    case DPP_FMA: return 3 * kC;    // v_mov_b32 (old value), v_mov_b32_dpp, v_fma_f32
    case PIXEL: return 13 * kC;     // v_rcp_f32, 10 FMAs, an add and an FMA
    case MFMA: return 6;            // six independent 4x4x1 accumulations
    case MFMA_FMA: return 2 + 10;   // two 4x4x1 accumulations and ten FMAs
  }
  return 1;
}

template <Op OP>
__global__ __launch_bounds__(256) void probe(float* out, uint64_t* cyc, uint64_t* rt, int iters, float s) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const float a = s * (float)(threadIdx.x & 7) + 1.0f;
  float v[kC], w[kC];
#pragma unroll
  for (int j = 0; j < kC; ++j) {
    v[j] = a + 0.125f * (float)j;
    w[j] = 0.5f + 0.01f * (float)j;
  }
  v4f c[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) c[j] = (v4f){0.f, 0.f, 0.f, 0.f};
  const uint64_t r0 = __builtin_amdgcn_s_memrealtime();
  const uint64_t t0 = __builtin_amdgcn_s_memtime();
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if constexpr (OP == FMA) {
#pragma unroll
        for (int j = 0; j < kC; ++j) v[j] = fmaf(v[j], 0.999f, a);
      } else if constexpr (OP == RCP) {
#pragma unroll
        for (int j = 0; j < kC; ++j) v[j] = __builtin_amdgcn_rcpf(v[j]);
      } else if constexpr (OP == EXP2) {
#pragma unroll
        for (int j = 0; j < kC; ++j) v[j] = __builtin_amdgcn_exp2f(v[j]) - 1.0f;
      } else if constexpr (OP == RCP_FMA) {
#pragma unroll
        for (int j = 0; j < kC; ++j) v[j] = fmaf(__builtin_amdgcn_rcpf(v[j]), 0.5f, 1.0f);
      } else if constexpr (OP == DPP_FMA) {
        // row_shr:1 within each row of 16 lanes
#pragma unroll
        for (int j = 0; j < kC; ++j)
          v[j] = fmaf(__builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v[j]), 0x111,
                                                                            0xf, 0xf, false)),
                      0.5f, a);
      } else if constexpr (OP == PIXEL) {
#pragma unroll
        for (int j = 0; j < kC; ++j) {
          const float r = __builtin_amdgcn_rcpf(v[j]);
          float t = w[j];
#pragma unroll
          for (int q = 0; q < 10; ++q) t = fmaf(t, r, 0.25f);
          w[j] = t;
          v[j] = fmaf(t, 1e-6f, v[j] + 1e-3f);
        }
      } else if constexpr (OP == MFMA) {
#pragma unroll
        for (int j = 0; j < 6; ++j) c[j] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, s, c[j], 0, 0, 0);
      } else if constexpr (OP == MFMA_FMA) {
        c[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, s, c[0], 0, 0, 0);
        c[1] = __builtin_amdgcn_mfma_f32_4x4x1f32(s, a, c[1], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 10; ++j) v[j % kC] = fmaf(v[j % kC], 0.999f, a);
      }
    }
  }
  const uint64_t t1 = __builtin_amdgcn_s_memtime();
  const uint64_t r1 = __builtin_amdgcn_s_memrealtime();
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < kC; ++j) acc += v[j] + w[j];
#pragma unroll
  for (int j = 0; j < 6; ++j) acc += c[j][0] + c[j][1] + c[j][2] + c[j][3];
  out[gid] = acc;
  if ((threadIdx.x & 63) == 0) {
    cyc[gid >> 6] = t1 - t0;
    rt[gid >> 6] = r1 - r0;
  }
}

struct Bufs {
  float* out = nullptr;
  uint64_t* cyc = nullptr;
  uint64_t* rt = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int cus = 0;
};

bool ok(hipError_t e, const char* what) {
  if (e != hipSuccess) std::fprintf(stderr, "issue_probe: %s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

// Clocks per SIMD instruction from the median wave (cpi) and from the
// launch's HIP-event time at the measured clock (cpi_launch: every SIMD's
// share of the grid, so uneven wave placement or staggered starts show up
// here and not in cpi), and the shader clock in MHz.
template <Op OP>
bool run(const Bufs& b, int waves_per_simd, double& cpi, double& cpi_launch, double& mhz) {
  const int blocks = b.cus * waves_per_simd;
  hipLaunchKernelGGL(probe<OP>, dim3(blocks), dim3(256), 0, 0, b.out, b.cyc, b.rt, kIters, 0.5f);   // warms the clock
  if (!ok(hipEventRecord(b.e0, 0), "event"))
    return false;
  hipLaunchKernelGGL(probe<OP>, dim3(blocks), dim3(256), 0, 0, b.out, b.cyc, b.rt, kIters, 0.5f);
  if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(b.e1, 0), "event") ||
      !ok(hipEventSynchronize(b.e1), "synchronize"))
    return false;
  float ms = 0.f;
  if (!ok(hipEventElapsedTime(&ms, b.e0, b.e1), "elapsed time")) return false;
  const int waves = blocks * 4;
  std::vector<uint64_t> t(waves), r(waves);
  if (!ok(hipMemcpy(t.data(), b.cyc, waves * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy") ||
      !ok(hipMemcpy(r.data(), b.rt, waves * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy"))
    return false;
  std::sort(t.begin(), t.end());
  std::sort(r.begin(), r.end());
  const double ticks = (double)t[waves / 2], real_us = (double)r[waves / 2] / 100.0;
  const double per_simd = (double)kIters * kU * per_body(OP) * waves_per_simd;
  cpi = ticks / per_simd;
  mhz = ticks / real_us;
  cpi_launch = (double)ms * 1e3 * mhz / per_simd;
  return true;
}

template <Op OP>
bool row(const Bufs& b, const char* name, const char* note) {
  std::printf("%-9s", name);
  double mhz_last = 0.0;
  for (int w = 1; w <= 4; ++w) {
    double cpi = 0.0, cpi_launch = 0.0, mhz = 0.0;
    if (!run<OP>(b, w, cpi, cpi_launch, mhz)) return false;
    std::printf(" | %5.2f %5.2f", cpi, cpi_launch);
    mhz_last = mhz;
  }
  std::printf(" | %5.0f | %s\n", mhz_last, note);
  return true;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  if (!ok(hipGetDeviceProperties(&prop, 0), "device properties")) return 1;
  Bufs b;
  b.cus = prop.multiProcessorCount;
  const size_t max_blocks = (size_t)b.cus * 4, max_waves = max_blocks * 4;
  if (!ok(hipMalloc(&b.out, max_blocks * 256 * sizeof(float)), "malloc") ||
      !ok(hipMalloc(&b.cyc, max_waves * sizeof(uint64_t)), "malloc") ||
      !ok(hipMalloc(&b.rt, max_waves * sizeof(uint64_t)), "malloc") || !ok(hipEventCreate(&b.e0), "event") ||
      !ok(hipEventCreate(&b.e1), "event"))
    return 1;
  std::printf("device %s, %d CUs; %d chains per wave, %d loop trips of %d bodies\n", prop.gcnArchName, b.cus, kC,
              kIters, kU);
  std::printf("clocks per SIMD instruction by resident waves per SIMD (W): median wave, then launch time\n");
  std::printf("%-9s | %11s | %11s | %11s | %11s | %5s | %s\n", "op", "W=1", "W=2", "W=3", "W=4", "MHz",
              "loop body per chain");
  bool good = row<FMA>(b, "fma", "v_fma_f32") && row<RCP>(b, "rcp", "v_rcp_f32") &&
              row<EXP2>(b, "exp2", "v_exp_f32 + v_add_f32 (per instruction)") &&
              row<RCP_FMA>(b, "rcp+fma", "v_rcp_f32 then the v_fma_f32 that reads it (per instruction)") &&
              row<DPP_FMA>(b, "dpp+fma", "v_mov_b32, v_mov_b32_dpp row_shr:1, v_fma_f32 (per instruction)") &&
              row<PIXEL>(b, "pixel", "v_rcp_f32 + 12 FMA/add (per instruction)") &&
              row<MFMA>(b, "mfma", "6 x v_mfma_f32_4x4x1f32 per wave and trip (per MFMA)") &&
              row<MFMA_FMA>(b, "mfma+fma", "2 x v_mfma_f32_4x4x1f32 + 10 v_fma_f32 (per instruction)");
  const bool freed = ok(hipFree(b.out), "free") && ok(hipFree(b.cyc), "free") && ok(hipFree(b.rt), "free") &&
                     ok(hipEventDestroy(b.e0), "event") && ok(hipEventDestroy(b.e1), "event");
  return good && freed ? 0 : 1;
}
