// gmle_device.hpp -- device-side building blocks of the gfx950 fitting kernels.
//
// Small helpers shared by every kernel instantiation: wave64 group reductions
// through DPP, fast transcendental wrappers, Julia-compatible clamp/min/max, and
// the Fisher-matrix inversion (regularised Cholesky, no-pivot LU fallback, Inf)
// that mirrors src/unified_kernel.jl:7-137,399-431 of GaussMLE.jl.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmle_args.hpp"

namespace gmle {

// ---------------------------------------------------------------------------
// Cross-lane sums inside a group of L (1, 2, 4, 8 or 16) lanes of one DPP row.
// quad_perm [1,0,3,2] swaps neighbours and [2,3,0,1] swaps pairs; after the
// quad sums, row_half_mirror (lane i <-> 7 - i within 8 lanes) meets the other
// quad of an 8-lane group and row_mirror (i <-> 15 - i within 16 lanes) the
// other half of a 16-lane group.  Every step adds two partial sums in either
// order, so every lane of the group ends with the bit-identical sum and
// replicated per-ROI arithmetic stays identical across the group.
// ---------------------------------------------------------------------------
constexpr int kDppSwap1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppSwap2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror
constexpr int kDppMirror = 0x140;      // row_mirror

// bound_ctrl is set: a lane whose source falls outside its row reads 0
// instead of an undefined value.  Only row_shr in group_next<8> has such lanes,
// and their values are never selected.  With it set, the compiler folds the
// move into the add or subtraction that reads it (v_add_f32_dpp): one VALU
// instruction fewer per reduction step.
template <int CTRL> __device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ double dpp_move(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ float dpp_swap1(float v) { return dpp_move<kDppSwap1>(v); }
__device__ __forceinline__ float dpp_swap2(float v) { return dpp_move<kDppSwap2>(v); }
__device__ __forceinline__ double dpp_swap1(double v) { return dpp_move<kDppSwap1>(v); }
__device__ __forceinline__ double dpp_swap2(double v) { return dpp_move<kDppSwap2>(v); }

template <int L, class T> __device__ __forceinline__ T group_sum(T v) {
  static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16, "lanes per ROI must be 1, 2, 4, 8 or 16");
  if constexpr (L >= 2) v = v + dpp_move<kDppSwap1>(v);
  if constexpr (L >= 4) v = v + dpp_move<kDppSwap2>(v);
  if constexpr (L >= 8) v = v + dpp_move<kDppHalfMirror>(v);
  if constexpr (L >= 16) v = v + dpp_move<kDppMirror>(v);
  return v;
}

template <int L> __device__ __forceinline__ float group_max(float v) {
  if constexpr (L >= 2) v = fmaxf(v, dpp_move<kDppSwap1>(v));
  if constexpr (L >= 4) v = fmaxf(v, dpp_move<kDppSwap2>(v));
  if constexpr (L >= 8) v = fmaxf(v, dpp_move<kDppHalfMirror>(v));
  if constexpr (L >= 16) v = fmaxf(v, dpp_move<kDppMirror>(v));
  return v;
}

// Moves inside a group of L lanes (lane r of a group is lane r of its ROI):
//   group_next<L>(v):     lane r receives lane (r + 1) mod L's v
//   group_bcast<L, O>(v): every lane receives lane O's v
// quad_perm patterns for 2 and 4 lanes, row_ror / row_newbcast for 16 (one
// DPP row), and two row moves plus a select for the half rows of 8.
constexpr int kDppRowShr = 0x110;       // row_shr:n = 0x110 + n, lane i <- i - n
constexpr int kDppRowRor = 0x120;       // row_ror:n = 0x120 + n, lane i <- (i - n) mod 16
constexpr int kDppRowNewBcast = 0x150;  // row_newbcast:n = 0x150 + n, lane i <- lane n of its row

template <int L> __device__ __forceinline__ float group_next(float v) {
  static_assert(L == 2 || L == 4 || L == 8 || L == 16, "groups of 2, 4, 8 or 16 lanes");
  if constexpr (L == 2) {
    return dpp_move<kDppSwap1>(v);
  } else if constexpr (L == 4) {
    return dpp_move<0x39>(v);   // quad_perm [1,2,3,0]
  } else if constexpr (L == 16) {
    return dpp_move<kDppRowRor + 15>(v);
  } else {
    const float a = dpp_move<kDppRowRor + 15>(v);   // lane i <- (i + 1) mod 16
    const float b = dpp_move<kDppRowShr + 7>(v);    // lanes 7 and 15 <- lanes 0 and 8
    return (threadIdx.x & 7) == 7 ? b : a;
  }
}

template <int L, int O> __device__ __forceinline__ float group_bcast(float v) {
  static_assert(L == 2 || L == 4 || L == 8 || L == 16, "groups of 2, 4, 8 or 16 lanes");
  static_assert(O >= 0 && O < L, "source lane inside the group");
  if constexpr (L == 2) {
    return dpp_move<O | (O << 2) | ((2 + O) << 4) | ((2 + O) << 6)>(v);
  } else if constexpr (L == 4) {
    return dpp_move<O * 0x55>(v);   // quad_perm [O,O,O,O]
  } else if constexpr (L == 16) {
    return dpp_move<kDppRowNewBcast + O>(v);
  } else {
    const float a = dpp_move<kDppRowNewBcast + O>(v);
    const float b = dpp_move<kDppRowNewBcast + 8 + O>(v);
    return (threadIdx.x & 8) ? b : a;
  }
}

// ---------------------------------------------------------------------------
// Transcendentals: exp2 is a single v_exp_f32, 1/x a single v_rcp_f32 (1 ulp).
// ---------------------------------------------------------------------------
constexpr float kInvSqrt2Pi = 0.39894228040143267794f;      // 1 / sqrt(2 pi)
constexpr double kInvSqrt2PiD = 0.39894228040143267794;
constexpr double kHalfD = 0.5;

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Julia semantics: clamp(x, lo, hi) = x > hi ? hi : (x < lo ? lo : x) (NaN passes);
// min / max propagate NaN.
template <class T> __device__ __forceinline__ T jclamp(T x, T lo, T hi) {
  return x > hi ? hi : (x < lo ? lo : x);
}
template <class T> __device__ __forceinline__ T jmin(T x, T y) {
  return (x != x || y != y) ? (x + y) : (y < x ? y : x);
}
template <class T> __device__ __forceinline__ T jmax(T x, T y) {
  return (x != x || y != y) ? (x + y) : (y > x ? y : x);
}

// ---------------------------------------------------------------------------
// Fisher inversion (src/unified_kernel.jl:7-137, 399-431).  F is P x P
// column-major, F[i + P*j].  Writes crlb[k] = sqrt(max(0, inv(F)[k,k])) and
// returns 0 (Cholesky), 1 (LU fallback) or 2 (singular -> Inf).
// ---------------------------------------------------------------------------
template <int P, class T> __device__ __forceinline__ int invert_fisher(T (&F)[P * P], T (&crlb)[P]) {
  T mx = T(0);
#pragma unroll
  for (int q = 0; q < P * P; ++q) mx = jmax(mx, F[q] < T(0) ? -F[q] : F[q]);
  const T reg = T(1e-10) * mx;
#pragma unroll
  for (int k = 0; k < P; ++k) F[k + P * k] += reg;

  // Cholesky A = L L^T on a copy.
  T Lm[P * P];
#pragma unroll
  for (int q = 0; q < P * P; ++q) Lm[q] = F[q];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < P; ++j) {
#pragma unroll
    for (int i = j; i < P; ++i) {
      T s = Lm[i + P * j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= Lm[i + P * k] * Lm[j + P * k];
      if (i == j) {
        ok = ok && (s > T(0));
        Lm[i + P * i] = sqrt(s > T(0) ? s : T(1));
      } else {
        Lm[i + P * j] = s / Lm[j + P * j];
      }
    }
  }
  if (ok) {
    // L^-1 (lower), then diag(A^-1)[k] = sum_{m>=k} Linv[m,k]^2.
    T Li[P * P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      Li[j + P * j] = T(1) / Lm[j + P * j];
#pragma unroll
      for (int i = j + 1; i < P; ++i) {
        T s = T(0);
#pragma unroll
        for (int k = j; k < i; ++k) s += Lm[i + P * k] * Li[k + P * j];
        Li[i + P * j] = -s / Lm[i + P * i];
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      T s = T(0);
#pragma unroll
      for (int m = k; m < P; ++m) s += Li[m + P * k] * Li[m + P * k];
      crlb[k] = sqrt(jmax(T(0), s));
    }
    return 0;
  }

  // LU without pivoting on the regularised matrix, relative pivot tolerance.
  T A[P * P];
  T amax = T(0);
#pragma unroll
  for (int q = 0; q < P * P; ++q) {
    A[q] = F[q];
    amax = jmax(amax, F[q] < T(0) ? -F[q] : F[q]);
  }
  const T tol = T(1e-10) * amax;
  bool lu_ok = true;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const T piv = A[k + P * k];
    lu_ok = lu_ok && !((piv < T(0) ? -piv : piv) < tol);
#pragma unroll
    for (int i = k + 1; i < P; ++i) {
      A[i + P * k] /= piv;
#pragma unroll
      for (int j = k + 1; j < P; ++j) A[i + P * j] -= A[i + P * k] * A[k + P * j];
    }
  }
  if (!lu_ok) {
#pragma unroll
    for (int k = 0; k < P; ++k) crlb[k] = T(__builtin_inf());
    return 2;
  }
  // Column k of the inverse gives the k-th diagonal entry.
#pragma unroll
  for (int col = 0; col < P; ++col) {
    T x[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      x[i] = (i == col) ? T(1) : T(0);
#pragma unroll
      for (int j = 0; j < i; ++j) x[i] -= A[i + P * j] * x[j];
    }
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
#pragma unroll
      for (int j = i + 1; j < P; ++j) x[i] -= A[i + P * j] * x[j];
      x[i] /= A[i + P * i];
    }
    crlb[col] = sqrt(jmax(T(0), x[col]));
  }
  return 1;
}

}  // namespace gmle
