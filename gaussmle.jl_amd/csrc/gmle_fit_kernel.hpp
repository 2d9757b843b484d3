// gmle_fit_kernel.hpp -- the per-ROI maximum-likelihood fitting kernel for gfx950.
//
// One kernel specialisation per (PSF model, camera, box size N, lanes per ROI L,
// precision).  It replaces GaussMLE.jl's unified_gaussian_mle_kernel!
// (src/unified_kernel.jl:285-434) together with everything it inlines:
// compute_pixel_derivatives (src/psf_derivatives.jl:6-187), the pixel-integrated
// Gaussian and its derivatives (src/gausslib/baselibrary.jl:8-67), the camera
// likelihood terms (src/camera_models.jl:19-96) and apply_constraints!
// (src/constraints.jl:172-177).
//
// Work mapping (wave64, one wave per workgroup):
//   * L lanes (1, 2, 4, 8 or 16, inside one DPP row) share one ROI; 64/L ROIs
//     per wave.
//   * Lane r of a group owns the contiguous column block [r*CB, r*CB + CB) with
//     CB = ceil(N/L); its CB*N pixels (and sCMOS variances) are loaded once and
//     stay in VGPRs for every Newton iteration and the final pass.
//   * The model is separable: value = bg + N*Phix[j]*Phiy[i].  Per iteration
//     each axis needs only N+1 edge evaluations (one erf and one exp each)
//     instead of the reference's 4 erf + 4-8 exp per pixel; every per-pixel
//     gradient / diagonal-Hessian term is a product of a column table and a row
//     table, so the per-pixel work is one reciprocal plus about ten FMAs that
//     feed per-column accumulators.  Column totals are folded into the P
//     gradient and P diagonal-Hessian sums, which are summed across the L lanes
//     with DPP permutes.
//   * The Newton update (exact division, step clamp, box clamp) is replicated
//     on the L lanes so they stay bit-identical.
//   * PREC_FP32 computes the final pass (LLR, full Fisher matrix, inversion,
//     CRLB) in Float32 like the reference; PREC_MIXED recomputes the model and
//     derivatives at the final estimate in Float64 and does the LLR, Fisher
//     accumulation and inversion in Float64.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "gmle_args.hpp"
#include "gmle_device.hpp"
#include "gmle_erf.hpp"

namespace gmle {

// ---------------------------------------------------------------------------
// Model traits.  EXTRA: 0 none (XYNB), 1 one sigma on both axes (XYNBS),
// 2 separate sigma per axis (XYNBSXSY), 3 astigmatic z (chain rule through
// sigma_x(z), sigma_y(z)).
// ---------------------------------------------------------------------------
template <int MODEL> struct MT;
template <> struct MT<XYNB> {
  static constexpr int P = 4, IX = 0, IY = 1, IN = 2, IBG = 3, IE1 = -1, IE2 = -1, EXTRA = 0;
};
template <> struct MT<XYNBS> {
  static constexpr int P = 5, IX = 0, IY = 1, IN = 2, IBG = 3, IE1 = 4, IE2 = -1, EXTRA = 1;
};
template <> struct MT<XYNBSXSY> {
  static constexpr int P = 6, IX = 0, IY = 1, IN = 2, IBG = 3, IE1 = 4, IE2 = 5, EXTRA = 2;
};
template <> struct MT<ASTIG> {
  static constexpr int P = 5, IX = 0, IY = 1, IN = 3, IBG = 4, IE1 = 2, IE2 = -1, EXTRA = 3;
};

// ---------------------------------------------------------------------------
// Scalar math per arithmetic type.
// ---------------------------------------------------------------------------
template <class T> struct DM;
template <> struct DM<float> {
  // Axis scales carry sqrt(log2 e): t * s is v = k u of gmle_erf.hpp, so the
  // Gaussian exp(-u^2) is exp2(-v^2).
  static __device__ __forceinline__ float erf_scale() { return kErfK; }
  // exp(-u^2), and the edge's signed erfc and half-sign for the pixel
  // integrals (gmle_erf.hpp), which reuse the Gaussian
  static __device__ __forceinline__ void erf_gauss(float t, float s, float& se, float& hs, float& g) {
    const float v = t * s;
    g = fast_exp2((-v) * v);
    erf_edge(v, g, se, hs);
  }
  static __device__ __forceinline__ float inv_sqrt_2pi() { return kInvSqrt2Pi; }
  // a / b for the per-iteration width and axis constants: one v_rcp_f32
  // (1 ulp) and a multiply instead of the ~10-instruction IEEE division.
  // The Newton step keeps the IEEE division.
  static __device__ __forceinline__ float div(float a, float b) { return a * fast_rcp(b); }
};
template <> struct DM<double> {
  static __device__ __forceinline__ double erf_scale() { return 1.0; }
  // Float64 keeps the reference's erf difference: its cancellation (~1e-16
  // absolute) is negligible, and tails that erf rounds to exactly +-1 stay
  // exactly 0 as in the reference.  se = -erf(u), hs = 0 in pixel_integral.
  static __device__ __forceinline__ void erf_gauss(double t, double s, double& se, double& hs, double& g) {
    const double u = t * s;
    g = exp(-(u * u));
    se = -erf(u);
    hs = 0.0;
  }
  static __device__ __forceinline__ double inv_sqrt_2pi() { return kInvSqrt2PiD; }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
};

// ---------------------------------------------------------------------------
// Per-axis tables for M consecutive pixels starting at 0-based index `first`
// (pixel k covers [k+1/2, k+3/2] in the reference's 1-based coordinates).  For
// pixel p with lower edge lo and upper edge hi (t = edge - mu):
//   phi = (erf(t_hi*s) - erf(t_lo*s)) / 2                  baselibrary.jl:8-14
//         (formed from signed complements, pixel_integral in gmle_erf.hpp)
//   D   = c (g_hi - g_lo),  c = -N/(sqrt(2pi) sigma)        baselibrary.jl:21-33
//   D2  = c/sigma^2 (t_hi g_hi - t_lo g_lo)
//   S   = c/sigma   (t_hi g_hi - t_lo g_lo)                 baselibrary.jl:40-52
//   SS  = -2/sigma S - N/(sqrt(2pi) sigma^5) (t_hi^3 g_hi - t_lo^3 g_lo)
// with g = exp(-t^2/(2 sigma^2)).  Pixel-level derivatives are these times the
// other axis' phi: dudt_x = D[j] phiy[i], d2udt2_x = D2[j] phiy[i], and the
// sigma derivative of one axis is S[j] phiy[i] (second: SS[j] phiy[i]).
// ---------------------------------------------------------------------------
template <class T, int M> struct AxisTab {
  T phi[M], D[M], D2[M], S[M], SS[M];
};

template <class T, int M, bool SIG>
__device__ __forceinline__ void axis_tables(T mu, T sigma, T nph, int first, AxisTab<T, M>& tb) {
  const T inv_s = DM<T>::div(T(1), sigma);
  const T s_erf = sqrt(DM<T>::div(T(0.5), sigma * sigma)) * DM<T>::erf_scale();
  const T c = -nph * DM<T>::inv_sqrt_2pi() * inv_s;
  const T c2 = c * inv_s * inv_s;
  const T cs = c * inv_s;
  const T css = c2 * inv_s * inv_s;
  const T m2s = T(-2) * inv_s;
  T se_prev = T(0), hs_prev = T(0), g_prev = T(0), tg_prev = T(0), t3g_prev = T(0);
#pragma unroll
  for (int e = 0; e <= M; ++e) {
    const T t = (T(first + e) + T(0.5)) - mu;
    T se, hs, g;
    DM<T>::erf_gauss(t, s_erf, se, hs, g);
    const T tg = t * g;
    const T t3g = SIG ? t * t * tg : T(0);
    if (e > 0) {
      const int p = e - 1;
      tb.phi[p] = pixel_integral<T>(se_prev, hs_prev, se, hs);
      tb.D[p] = c * (g - g_prev);
      const T w = tg - tg_prev;
      tb.D2[p] = c2 * w;
      if (SIG) {
        tb.S[p] = cs * w;
        tb.SS[p] = m2s * tb.S[p] + css * (t3g - t3g_prev);
      }
    }
    se_prev = se;
    hs_prev = hs;
    g_prev = g;
    tg_prev = tg;
    t3g_prev = t3g;
  }
}

// Axis constants of axis_tables() and single-edge evaluation, for computing
// one pixel's tables from its two edges on the fly (column side).
template <class T> struct AxisK {
  T mu, s_erf, c, c2, cs, css, m2s;
};
template <class T> struct Edge {
  T se, hs, g, tg, t3g;   // signed erfc, half-sign (gmle_erf.hpp), Gaussian, t g, t^3 g
};
template <class T>
__device__ __forceinline__ AxisK<T> axis_consts(T mu, T sigma, T nph) {
  AxisK<T> k;
  const T inv_s = DM<T>::div(T(1), sigma);
  k.mu = mu;
  k.s_erf = sqrt(DM<T>::div(T(0.5), sigma * sigma)) * DM<T>::erf_scale();
  k.c = -nph * DM<T>::inv_sqrt_2pi() * inv_s;
  k.c2 = k.c * inv_s * inv_s;
  k.cs = k.c * inv_s;
  k.css = k.c2 * inv_s * inv_s;
  k.m2s = T(-2) * inv_s;
  return k;
}
template <class T, bool SIG>
__device__ __forceinline__ Edge<T> edge_eval(const AxisK<T>& k, int e) {
  Edge<T> r;
  const T t = (T(e) + T(0.5)) - k.mu;
  DM<T>::erf_gauss(t, k.s_erf, r.se, r.hs, r.g);
  r.tg = t * r.g;
  r.t3g = SIG ? t * t * r.tg : T(0);
  return r;
}
// One pixel's tables (see axis_tables) from its lower / upper edges.
template <class T> struct Pix {
  T phi, D, D2, S, SS;
};
template <class T, bool SIG>
__device__ __forceinline__ Pix<T> pix_from_edges(const AxisK<T>& k, const Edge<T>& lo, const Edge<T>& hi) {
  Pix<T> p;
  p.phi = pixel_integral<T>(lo.se, lo.hs, hi.se, hi.hs);
  p.D = k.c * (hi.g - lo.g);
  const T w = hi.tg - lo.tg;
  p.D2 = k.c2 * w;
  p.S = SIG ? k.cs * w : T(0);
  p.SS = SIG ? k.m2s * p.S + k.css * (hi.t3g - lo.t3g) : T(0);
  return p;
}

// ---------------------------------------------------------------------------
// Row tables shared by the L lanes of an ROI.  Every lane needs every row's
// tables, but evaluating the N + 1 row edges (one erf and one exp each) on
// all L lanes repeats the same work L times.  Instead lane r evaluates edges
// r, r + L, r + 2L, ... (slot s holds edge r + L s) and builds the tables of
// rows r, r + L, ...; a row's tables reach the other lanes by a DPP
// broadcast from its owner where they are used.  Row r + L s's upper edge is
// lane r + 1's slot-s edge, or lane 0's slot-(s + 1) edge for the group's
// last lane: lane 0 offers its next edge, the others their current one, and
// every lane reads the next lane's offer.  Each value comes from the same
// expressions as in the replicated evaluation; the compiler's FMA contraction
// can still round a rare one differently, so a kernel and its replicated twin
// may differ in the last bit of an ROI's result.
// ---------------------------------------------------------------------------
// Where sharing is used (gfx950 register counts): sharing keeps a few more
// values live, so the tiles at the top of each two-wave range keep the
// replicated evaluation: with 4 lanes the largest box that fits 256 VGPRs
// (max_box_within_256), with 2 lanes the two largest, since the multi-shape
// kernel, which calls every 1-, 2- and 4-lane tile, has to stay at two waves
// per SIMD too.  From 8 lanes up every tile shares.
template <int MODEL, int CAM, int N, int L> constexpr bool share_rows() {
  if (L == 1) return false;
  if (L >= 8) return true;
  if (L == 2) return N + 1 < max_box_within_256(MODEL, CAM, L);
  return N != max_box_within_256(MODEL, CAM, L);
}

template <int N, int L> constexpr int row_slots() { return (N + L - 1) / L; }   // rows per lane
template <int N, int L> constexpr int edge_slots() { return (N + L) / L; }     // edges 0..N per lane

// ---------------------------------------------------------------------------
// Tail split.  With N = L q + 1 the plain column blocks give each lane
// ceil(N / L) = q + 1 columns, and the group computes L (q + 1) columns for N:
// almost a whole padding column per lane (13 px on 4 lanes: 52 pixel slots
// per lane for 169 pixels).  Split instead, each lane holds q full columns
// [r q, r q + q) and, of the last column N - 1, the rows r, r + L, r + 2L, ...
// -- exactly the rows whose tables the lane builds itself when rows are
// shared, so those pixels need no broadcast: q N + ceil(N / L) slots (13 px on
// 4 lanes: 43).  The last column's partial row sums are folded with that
// column's tables on every lane and summed across the group with the rest
// (the fold is linear in the sums).
// ---------------------------------------------------------------------------
template <int MODEL, int CAM, int N, int L> constexpr bool tail_split() {
  return L >= 2 && N > L && N % L == 1 && N <= kMaxSmallBox && share_rows<MODEL, CAM, N, L>();
}
// Register-resident full columns per lane.
template <int MODEL, int CAM, int N, int L> constexpr int block_cols() {
  return tail_split<MODEL, CAM, N, L>() ? N / L : (N + L - 1) / L;
}

template <int L, bool SIG>
__device__ __forceinline__ Pix<float> slot_row(const AxisK<float>& kr, int r, const Edge<float>& e_s,
                                               const Edge<float>& e_next) {
  const bool first = r == 0;
  Edge<float> hi;
  hi.se = group_next<L>(first ? e_next.se : e_s.se);
  hi.hs = group_next<L>(first ? e_next.hs : e_s.hs);
  hi.g = group_next<L>(first ? e_next.g : e_s.g);
  hi.tg = group_next<L>(first ? e_next.tg : e_s.tg);
  hi.t3g = SIG ? group_next<L>(first ? e_next.t3g : e_s.t3g) : 0.f;
  return pix_from_edges<float, SIG>(kr, e_s, hi);
}

// The edge of slot s + 1 when some lane has one (else unused).
template <int N, int L, bool SIG>
__device__ __forceinline__ Edge<float> next_slot_edge(const AxisK<float>& kr, int r, int s, const Edge<float>& e_s) {
  return (s + 1 < edge_slots<N, L>()) ? edge_eval<float, SIG>(kr, r + L * (s + 1)) : e_s;
}

// Row i's tables on every lane of the group, from lane o = i mod L.  o is a
// constant once the caller's row loop is unrolled, so the chain of tests
// folds to a single broadcast.
template <int L, bool SIG, int O = 0>
__device__ __forceinline__ Pix<float> bcast_pix(const Pix<float>& p, int o) {
  if constexpr (O + 1 < L) {
    if (o != O) return bcast_pix<L, SIG, O + 1>(p, o);
  }
  Pix<float> q;
  q.phi = group_bcast<L, O>(p.phi);
  q.D = group_bcast<L, O>(p.D);
  q.D2 = group_bcast<L, O>(p.D2);
  q.S = SIG ? group_bcast<L, O>(p.S) : 0.f;
  q.SS = SIG ? group_bcast<L, O>(p.SS) : 0.f;
  return q;
}

// All N rows' tables on every lane, as axis_tables() builds them; own[s]
// receives this lane's own row r + L s (the tail split uses them).
template <int N, int L, bool SIG>
__device__ __forceinline__ void shared_axis_tables(const AxisK<float>& kr, int r, AxisTab<float, N>& tb,
                                                   Pix<float>* own) {
  constexpr int RS = row_slots<N, L>();
  Edge<float> e_s = edge_eval<float, SIG>(kr, r);
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const Edge<float> e_next = next_slot_edge<N, L, SIG>(kr, r, s, e_s);
    own[s] = slot_row<L, SIG>(kr, r, e_s, e_next);
    e_s = e_next;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Pix<float> q = bcast_pix<L, SIG>(own[i / L], i % L);
    tb.phi[i] = q.phi;
    tb.D[i] = q.D;
    tb.D2[i] = q.D2;
    if constexpr (SIG) {
      tb.S[i] = q.S;
      tb.SS[i] = q.SS;
    }
  }
}

// Per-ROI width parameters for one evaluation of the model.
template <class T> struct Widths {
  T sx, sy;          // sigma per axis (pixels)
  T dsx, dsy;        // astig: d sigma / dz
  T d2sx, d2sy;      // astig: d2 sigma / dz2
};

// psf_derivatives.jl:115-163: sigma_x(z), sigma_y(z) and their z derivatives.
template <class T>
__device__ __forceinline__ void astig_widths(T z, const float* psf, Widths<T>& w) {
  const T sx0 = (T)psf[0], sy0 = (T)psf[1], Ax = (T)psf[2], Ay = (T)psf[3];
  const T Bx = (T)psf[4], By = (T)psf[5], gamma = (T)psf[6], dd = (T)psf[7];
  const T zdx = DM<T>::div(z - gamma, dd);
  const T zdy = DM<T>::div(z + gamma, dd);
  const T zdx2 = zdx * zdx, zdy2 = zdy * zdy;
  const T ax = T(1) + zdx2 + Ax * (zdx2 * zdx) + Bx * (zdx2 * zdx2);
  const T ay = T(1) + zdy2 + Ay * (zdy2 * zdy) + By * (zdy2 * zdy2);
  const T sqax = sqrt(ax), sqay = sqrt(ay);
  w.sx = sx0 * sqax;
  w.sy = sy0 * sqay;
  const T dax = DM<T>::div(T(2) * zdx + T(3) * Ax * zdx2 + T(4) * Bx * (zdx2 * zdx), dd);
  const T day = DM<T>::div(T(2) * zdy + T(3) * Ay * zdy2 + T(4) * By * (zdy2 * zdy), dd);
  const T hx = DM<T>::div(sx0, T(2) * sqax), hy = DM<T>::div(sy0, T(2) * sqay);
  w.dsx = hx * dax;
  w.dsy = hy * day;
  const T d2ax = DM<T>::div(T(2) + T(6) * Ax * zdx + T(12) * Bx * zdx2, dd * dd);
  const T d2ay = DM<T>::div(T(2) + T(6) * Ay * zdy + T(12) * By * zdy2, dd * dd);
  w.d2sx = -DM<T>::div(sx0, T(4) * (ax * sqax)) * (dax * dax) + hx * d2ax;
  w.d2sy = -DM<T>::div(sy0, T(4) * (ay * sqay)) * (day * day) + hy * d2ay;
}

template <int MODEL, class T>
__device__ __forceinline__ void model_widths(const T* th, const float* psf, Widths<T>& w) {
  w.dsx = w.dsy = w.d2sx = w.d2sy = T(0);
  if constexpr (MODEL == XYNB) {
    w.sx = w.sy = (T)psf[0];
  } else if constexpr (MODEL == XYNBS) {
    w.sx = w.sy = th[4];
  } else if constexpr (MODEL == XYNBSXSY) {
    w.sx = th[4];
    w.sy = th[5];
  } else {
    astig_widths<T>(th[2], psf, w);
  }
}

// Fold one column's row-weighted sums S[0..10] = S1..S11 (see accumulate)
// into the gradient / diagonal Hessian with the column's tables
// (unified_kernel.jl:330-333): grad_k += dudt_k cf, H_k += d2_k cf - dudt_k^2 df.
template <int MODEL>
__device__ __forceinline__ void fold_column(const Pix<float>& cx, const Widths<float>& w, const float* S, float* G,
                                            float* H) {
  using Tr = MT<MODEL>;
  constexpr bool SIG = Tr::EXTRA != 0;
  constexpr bool TWO = (Tr::EXTRA == 1 || Tr::EXTRA == 3);
  const float px = cx.phi, Dx = cx.D, D2x = cx.D2;
  G[Tr::IX] += Dx * S[0];
  H[Tr::IX] += D2x * S[0] - Dx * Dx * S[1];
  G[Tr::IY] += px * S[2];
  H[Tr::IY] += px * S[3] - px * px * S[4];
  G[Tr::IN] += px * S[0];
  H[Tr::IN] -= px * px * S[1];
  G[Tr::IBG] += S[5];
  H[Tr::IBG] -= S[6];
  if constexpr (SIG) {
    float ax, aax;
    if constexpr (MODEL == ASTIG) {
      ax = w.dsx * cx.S;
      aax = cx.SS * (w.dsx * w.dsx) + cx.S * w.d2sx;
    } else {
      ax = cx.S;
      aax = cx.SS;
    }
    if constexpr (TWO) {
      G[Tr::IE1] += ax * S[0] + px * S[7];
      H[Tr::IE1] += aax * S[0] + px * S[8] - ax * ax * S[1] - 2.f * ax * px * S[9] - px * px * S[10];
    } else {
      G[Tr::IE1] += ax * S[0];
      H[Tr::IE1] += aax * S[0] - ax * ax * S[1];
      G[Tr::IE2] += px * S[7];
      H[Tr::IE2] += px * S[8] - px * px * S[10];
    }
  }
}

// One pixel of the tail column (tail split): row ry of this lane's slot,
// column tables folded later.  ok is false for slots past the last row; their
// terms are zero.  For the Poisson fast path rs[] collects the row sums the q
// form subtracts (phi, D, D2, A, AA, rows).
template <int MODEL, int CAM, bool FAST>
__device__ __forceinline__ void tail_pixel(const Pix<float>& ry, const Widths<float>& w, float d, float var, bool ok,
                                           float ncx, float bg, float* St, float* rs) {
  using Tr = MT<MODEL>;
  constexpr bool SIG = Tr::EXTRA != 0;
  constexpr bool TWO = (Tr::EXTRA == 1 || Tr::EXTRA == 3);
  float ay = 0.f, aay = 0.f;
  if constexpr (SIG) {
    if constexpr (MODEL == ASTIG) {
      ay = w.dsy * ry.S;
      aay = ry.SS * (w.dsy * w.dsy) + ry.S * w.d2sy;
    } else {
      ay = ry.S;
      aay = ry.SS;
    }
  }
  const float py = ry.phi;
  const float mu = fmaf(ncx, py, bg);
  float cf, df;
  if constexpr (CAM == IDEAL) {
    const float rc = fast_rcp(mu);
    const float q = d * rc;
    if constexpr (FAST) {
      cf = q;
      df = q * rc;
    } else {
      const bool pos = mu > 0.01f;   // camera_models.jl:23-32
      cf = pos ? jmin(q - 1.f, 1.0e4f) : 0.f;
      df = pos ? jmin(q * rc, 1.0e4f) : 0.f;
    }
  } else {
    const float v = mu + var;   // camera_models.jl:36-42
    const float rc = fast_rcp(v);
    cf = (d - mu) * rc;
    df = rc;
  }
  cf = ok ? cf : 0.f;
  df = ok ? df : 0.f;
  St[0] = fmaf(py, cf, St[0]);
  St[1] = fmaf(py * py, df, St[1]);
  St[2] = fmaf(ry.D, cf, St[2]);
  St[3] = fmaf(ry.D2, cf, St[3]);
  St[4] = fmaf(ry.D * ry.D, df, St[4]);
  St[5] += cf;
  St[6] += df;
  if constexpr (SIG) {
    St[7] = fmaf(ay, cf, St[7]);
    St[8] = fmaf(aay, cf, St[8]);
    St[10] = fmaf(ay * ay, df, St[10]);
    if constexpr (TWO) St[9] = fmaf(py * ay, df, St[9]);
  }
  if constexpr (CAM == IDEAL && FAST) {
    rs[0] += ok ? py : 0.f;
    rs[1] += ok ? ry.D : 0.f;
    rs[2] += ok ? ry.D2 : 0.f;
    rs[3] += ok ? ay : 0.f;
    rs[4] += ok ? aay : 0.f;
    rs[5] += ok ? 1.f : 0.f;
  }
}

// The tail column's sums folded into G / H with its tables tl.
template <int MODEL, int CAM, bool FAST>
__device__ __forceinline__ void tail_fold(float* St, const float* rs, const Pix<float>& tl, const Widths<float>& w,
                                          float* G, float* H) {
  if constexpr (CAM == IDEAL && FAST) {
    St[0] -= rs[0];
    St[2] -= rs[1];
    St[3] -= rs[2];
    St[5] -= rs[5];
    if constexpr (MT<MODEL>::EXTRA != 0) {
      St[7] -= rs[3];
      St[8] -= rs[4];
    }
  }
  fold_column<MODEL>(tl, w, St, G, H);
}

// The same sums with rows in the outer loop: each row's tables are built on
// the fly from its two edges and the CB columns of this lane keep their S
// vectors live instead.  Cheaper in VGPRs than holding every row's tables
// when there are many row tables (sigma / z models) and few columns per lane.
template <int MODEL, int CAM, int N, int L, bool FAST, bool SHARE>
__device__ __forceinline__ void accumulate_rows(const float* dv, const float* vr, const float* dt, const float* vt,
                                                int c0, float nph, float bg, const AxisK<float>& kr, int r,
                                                const AxisK<float>& kc, const Widths<float>& w, float* G, float* H) {
  using Tr = MT<MODEL>;
  constexpr int P = Tr::P;
  constexpr bool TAIL = tail_split<MODEL, CAM, N, L>();
  constexpr int CB = block_cols<MODEL, CAM, N, L>();
  constexpr bool SIG = Tr::EXTRA != 0;
  constexpr bool TWO = (Tr::EXTRA == 1 || Tr::EXTRA == 3);
  constexpr int NS = SIG ? 11 : 7;
  static_assert(!TAIL || SHARE, "the tail split uses the lane's own row tables");
#pragma unroll
  for (int k = 0; k < P; ++k) { G[k] = 0.f; H[k] = 0.f; }
  // this lane's column tables (and the tail column's, from lane L - 1, whose
  // block ends at edge N - 1)
  Pix<float> cx[CB];
  float ncx[CB];
  Pix<float> tl{};
  {
    Edge<float> lo = edge_eval<float, SIG>(kc, c0);
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const Edge<float> hi = edge_eval<float, SIG>(kc, c0 + cc + 1);
      cx[cc] = pix_from_edges<float, SIG>(kc, lo, hi);
      lo = hi;
      ncx[cc] = nph * cx[cc].phi;
    }
    if constexpr (TAIL) {
      const Edge<float> hi = edge_eval<float, SIG>(kc, N);
      tl = bcast_pix<L, SIG>(pix_from_edges<float, SIG>(kc, lo, hi), L - 1);
    }
  }
  const float ncx_t = nph * tl.phi;
  float St[11], rs[6];
#pragma unroll
  for (int q = 0; q < 11; ++q) St[q] = 0.f;
#pragma unroll
  for (int q = 0; q < 6; ++q) rs[q] = 0.f;
  float S[CB][11];
#pragma unroll
  for (int cc = 0; cc < CB; ++cc)
#pragma unroll
    for (int q = 0; q < 11; ++q) S[cc][q] = 0.f;
  float sphi = 0.f, sD = 0.f, sD2 = 0.f, sA = 0.f, sAA = 0.f;
  // Without sharing, rlo is the running lower edge.  With it, rlo is this
  // lane's edge of the current slot, and `own` the tables of its row in that
  // slot, built when the group reaches the slot's first row.
  Edge<float> rlo = edge_eval<float, SIG>(kr, SHARE ? r : 0);
  Pix<float> own;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    // Keep each row's work together: without this the scheduler hoists later
    // rows' edge evaluations and every row's tables become live at once.
    __builtin_amdgcn_sched_barrier(0);
    Pix<float> ry;
    if constexpr (!SHARE) {
      const Edge<float> rhi = edge_eval<float, SIG>(kr, i + 1);
      ry = pix_from_edges<float, SIG>(kr, rlo, rhi);
      rlo = rhi;
    } else {
      if (i % L == 0) {
        const Edge<float> e_next = next_slot_edge<N, L, SIG>(kr, r, i / L, rlo);
        own = slot_row<L, SIG>(kr, r, rlo, e_next);
        rlo = e_next;
        // the tail pixel of this slot's row r + i, with the lane's own tables
        if constexpr (TAIL)
          tail_pixel<MODEL, CAM, FAST>(own, w, dt[i / L], CAM == SCMOS ? vt[i / L] : 0.f, r + i < N, ncx_t, bg, St,
                                       rs);
      }
      ry = bcast_pix<L, SIG>(own, i % L);
    }
    float ay = 0.f, aay = 0.f;
    if constexpr (SIG) {
      if constexpr (MODEL == ASTIG) {
        ay = w.dsy * ry.S;
        aay = ry.SS * (w.dsy * w.dsy) + ry.S * w.d2sy;
      } else {
        ay = ry.S;
        aay = ry.SS;
      }
    }
    if constexpr (CAM == IDEAL && FAST) {
      sphi += ry.phi;
      sD += ry.D;
      sD2 += ry.D2;
      if constexpr (SIG) { sA += ay; sAA += aay; }
    }
    const float py = ry.phi, py2 = py * py, Dy2 = ry.D * ry.D, a2 = ay * ay, pa = py * ay;
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const float d = dv[cc * N + i];
      const float mu = fmaf(ncx[cc], py, bg);
      float cf, df;
      if constexpr (CAM == IDEAL) {
        const float rc = fast_rcp(mu);
        const float q = d * rc;
        if constexpr (FAST) {
          cf = q;   // cf + 1; the constant is removed after the row sums
          df = q * rc;
        } else {
          const bool ok = mu > 0.01f;   // camera_models.jl:23-32
          cf = ok ? jmin(q - 1.f, 1.0e4f) : 0.f;
          df = ok ? jmin(q * rc, 1.0e4f) : 0.f;
        }
      } else {
        const float v = mu + vr[cc * N + i];   // camera_models.jl:36-42
        const float rc = fast_rcp(v);
        cf = (d - mu) * rc;
        df = rc;
      }
      S[cc][0] = fmaf(py, cf, S[cc][0]);
      S[cc][1] = fmaf(py2, df, S[cc][1]);
      S[cc][2] = fmaf(ry.D, cf, S[cc][2]);
      S[cc][3] = fmaf(ry.D2, cf, S[cc][3]);
      S[cc][4] = fmaf(Dy2, df, S[cc][4]);
      S[cc][5] += cf;
      S[cc][6] += df;
      if constexpr (SIG) {
        S[cc][7] = fmaf(ay, cf, S[cc][7]);
        S[cc][8] = fmaf(aay, cf, S[cc][8]);
        S[cc][10] = fmaf(a2, df, S[cc][10]);
        if constexpr (TWO) S[cc][9] = fmaf(pa, df, S[cc][9]);
      }
    }
    // With shared rows, also tie this row's sums to the end of the row, or
    // the pixel work of every row sinks below the last row's broadcast.
    if constexpr (SHARE) {
#pragma unroll
      for (int cc = 0; cc < CB; ++cc) {
        asm volatile("" : "+v"(S[cc][0]), "+v"(S[cc][1]), "+v"(S[cc][2]), "+v"(S[cc][3]), "+v"(S[cc][4]),
                     "+v"(S[cc][5]), "+v"(S[cc][6]));
        if constexpr (SIG) asm volatile("" : "+v"(S[cc][7]), "+v"(S[cc][8]), "+v"(S[cc][10]));
        if constexpr (TWO) asm volatile("" : "+v"(S[cc][9]));
      }
    }
  }
#pragma unroll
  for (int cc = 0; cc < CB; ++cc) {
    if constexpr (CAM == IDEAL && FAST) {
      S[cc][0] -= sphi;
      S[cc][2] -= sD;
      S[cc][3] -= sD2;
      S[cc][5] -= (float)N;
      if constexpr (SIG) { S[cc][7] -= sA; S[cc][8] -= sAA; }
    }
    // Columns past the ROI (last lane of a group when L does not divide N)
    // contribute nothing.  Selects, not a branch: a per-lane branch here
    // splits the block and costs ~100 VGPRs of live ranges.
    const bool cvalid = (L == 1) || TAIL || (c0 + cc < N);
#pragma unroll
    for (int q = 0; q < 11; ++q) S[cc][q] = cvalid ? S[cc][q] : 0.f;
    fold_column<MODEL>(cx[cc], w, S[cc], G, H);
  }
  if constexpr (TAIL) tail_fold<MODEL, CAM, FAST>(St, rs, tl, w, G, H);
  (void)NS;
}

// Loop order with the smaller live-register estimate: column-outer holds every
// row's tables (5 per row, 9 with sigma / z terms, 8 for separate sigmas);
// row-outer holds this lane's column tables and S vectors (about 11 and 18
// per column).
template <int MODEL, int CAM, int N, int L> constexpr bool row_outer() {
  constexpr int CB = (N + L - 1) / L;
  constexpr int ex = MT<MODEL>::EXTRA;
  constexpr int row_tables = ex == 0 ? 5 : (ex == 2 ? 8 : 9);
  constexpr int per_col = ex == 0 ? 11 : 18;
  return per_col * CB < row_tables * N;
}

// ---------------------------------------------------------------------------
// One Newton iteration's gradient G[P] and diagonal Hessian H[P] over this
// lane's column block (unified_kernel.jl:319-334):
//   G_k = sum_ij dudt_k cf,   H_k = sum_ij (d2udt2_k cf - dudt_k^2 df).
// Each per-pixel term factors as (column table)(row table)(cf or df), so the
// pixel loop only accumulates row-weighted cf / df sums per column
//   S1 = sum_i phiy cf    S2 = sum_i phiy^2 df   S3 = sum_i Dy cf
//   S4 = sum_i D2y cf     S5 = sum_i Dy^2 df     S6 = sum_i cf    S7 = sum_i df
//   S8 = sum_i Ay cf      S9 = sum_i AAy cf      S10 = sum_i phiy Ay df
//   S11 = sum_i Ay^2 df
// (A = derivative w.r.t. the sigma / z parameter on the row axis, AA its second
// derivative) that are then multiplied by the column's tables, which are built
// from the column's two edges on the fly.
// FAST (Poisson only): every model value exceeds 0.01 and neither 1e4 cap
// (camera_models.jl:23-28) can bind, so cf + 1 = data/model is accumulated and
// the constant row sums (sphi, sD, ...) are subtracted once per column.
// ---------------------------------------------------------------------------
template <int MODEL, int CAM, int N, int L, bool FAST>
__device__ __forceinline__ void accumulate(const float* dv, const float* vr, int c0, float nph, float bg,
                                           const AxisTab<float, N>& ry, const AxisK<float>& kc,
                                           const Widths<float>& w, const float* Ay, const float* AAy,
                                           float sphi, float sD, float sD2, float sA, float sAA, float* G,
                                           float* H) {
  using Tr = MT<MODEL>;
  constexpr int P = Tr::P;
  constexpr bool TAIL = tail_split<MODEL, CAM, N, L>();
  constexpr int CB = block_cols<MODEL, CAM, N, L>();
  constexpr bool SIG = Tr::EXTRA != 0;
  constexpr bool TWO = (Tr::EXTRA == 1 || Tr::EXTRA == 3);
#pragma unroll
  for (int k = 0; k < P; ++k) { G[k] = 0.f; H[k] = 0.f; }
  Edge<float> lo = edge_eval<float, SIG>(kc, c0);
#pragma unroll
  for (int cc = 0; cc < CB; ++cc) {
    const Edge<float> hi = edge_eval<float, SIG>(kc, c0 + cc + 1);
    const Pix<float> cx = pix_from_edges<float, SIG>(kc, lo, hi);
    lo = hi;
    const bool cvalid = (L == 1) || TAIL || (c0 + cc < N);
    const float nphx = nph * cx.phi;
    float S1 = 0.f, S2 = 0.f, S3 = 0.f, S4 = 0.f, S5 = 0.f, S6 = 0.f, S7 = 0.f;
    float S8 = 0.f, S9 = 0.f, S10 = 0.f, S11 = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float d = dv[cc * N + i];
      const float mu = fmaf(nphx, ry.phi[i], bg);
      float cf, df;
      if constexpr (CAM == IDEAL) {
        const float rc = fast_rcp(mu);
        const float q = d * rc;
        if constexpr (FAST) {
          cf = q;   // cf + 1; the constant is removed after the row sums
          df = q * rc;
        } else {
          const bool ok = mu > 0.01f;   // camera_models.jl:23-32
          cf = ok ? jmin(q - 1.f, 1.0e4f) : 0.f;
          df = ok ? jmin(q * rc, 1.0e4f) : 0.f;
        }
      } else {
        const float v = mu + vr[cc * N + i];   // camera_models.jl:36-42
        const float rc = fast_rcp(v);
        cf = (d - mu) * rc;
        df = rc;
      }
      const float py = ry.phi[i];
      S1 = fmaf(py, cf, S1);
      S2 = fmaf(py * py, df, S2);
      S3 = fmaf(ry.D[i], cf, S3);
      S4 = fmaf(ry.D2[i], cf, S4);
      S5 = fmaf(ry.D[i] * ry.D[i], df, S5);
      S6 += cf;
      S7 += df;
      if constexpr (SIG) {
        S8 = fmaf(Ay[i], cf, S8);
        S9 = fmaf(AAy[i], cf, S9);
        S11 = fmaf(Ay[i] * Ay[i], df, S11);
        if constexpr (TWO) S10 = fmaf(py * Ay[i], df, S10);
      }
    }
    if constexpr (CAM == IDEAL && FAST) {
      S1 -= sphi;
      S3 -= sD;
      S4 -= sD2;
      S6 -= (float)N;
      if constexpr (SIG) { S8 -= sA; S9 -= sAA; }
    }
    // Columns past the ROI contribute nothing (selects, not a branch).
    const float Sv[11] = {cvalid ? S1 : 0.f, cvalid ? S2 : 0.f, cvalid ? S3 : 0.f, cvalid ? S4 : 0.f,
                          cvalid ? S5 : 0.f, cvalid ? S6 : 0.f, cvalid ? S7 : 0.f, cvalid ? S8 : 0.f,
                          cvalid ? S9 : 0.f, cvalid ? S10 : 0.f, cvalid ? S11 : 0.f};
    fold_column<MODEL>(cx, w, Sv, G, H);
  }
}

// ---------------------------------------------------------------------------
// Final pass (unified_kernel.jl:353-431): at the estimate th, the LLR
// (camera_models.jl:60-85), the full Fisher matrix over pixels with model > 0
// (unified_kernel.jl:378-396) and its inversion into the CRLB, in arithmetic
// type TF.  Pixel data come from Src: VGPR arrays (RegSrc, in the iteration
// kernel) or global memory (GlobalSrc, in the Float64 final-pass kernel).
// ---------------------------------------------------------------------------
template <int N, bool SC> struct RegSrc {
  static constexpr bool kInRegisters = true;
  const float* dv;   // this lane's column block, [cc * N + i]
  const float* vr;
  const float* dt = nullptr;   // tail split: this lane's rows r + L s of the last column
  const float* vt = nullptr;
  __device__ __forceinline__ float d(int cc, int i) const { return dv[cc * N + i]; }
  __device__ __forceinline__ float var(int cc, int i) const { return SC ? vr[cc * N + i] : 0.f; }
  __device__ __forceinline__ void column(int cc, float* dcol, float* vcol) const {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      dcol[i] = dv[cc * N + i];
      if constexpr (SC) vcol[i] = vr[cc * N + i];
    }
  }
};
// n consecutive floats from p (4-byte aligned) in 16- and 8-byte loads: gfx950
// global loads need only dword alignment, and one lane's run of a column then
// takes ceil(n / 4) load instructions instead of n.
template <int n> __device__ __forceinline__ void load_run(const float* p, float* out) {
#pragma unroll
  for (int i = 0; i + 4 <= n; i += 4) {
    float4 q;
    __builtin_memcpy(&q, p + i, sizeof(q));
    out[i] = q.x;
    out[i + 1] = q.y;
    out[i + 2] = q.z;
    out[i + 3] = q.w;
  }
  constexpr int b = n - n % 4;
  if constexpr (n % 4 >= 2) {
    float2 q;
    __builtin_memcpy(&q, p + b, sizeof(q));
    out[b] = q.x;
    out[b + 1] = q.y;
  }
  if constexpr (n % 2 == 1) out[n - 1] = p[n - 1];
}

template <int N, bool SC> struct GlobalSrc {
  static constexpr bool kInRegisters = false;
  const float* data = nullptr;   // the ROI, column-major
  const float* var = nullptr;    // var_map at the ROI's first pixel
  int ny = 0;
  __device__ __forceinline__ float d(int cc, int i) const { return data[cc * N + i]; }
  __device__ __forceinline__ float var_at(int cc, int i) const { return var[(int64_t)cc * ny + i]; }
  // column cc's pixels and variances (both runs are contiguous)
  __device__ __forceinline__ void column(int cc, float* dcol, float* vcol) const {
    load_run<N>(data + cc * N, dcol);
    if constexpr (SC) load_run<N>(var + (int64_t)cc * ny, vcol);
  }
};
template <int N, bool SC> __device__ __forceinline__ float src_var(const RegSrc<N, SC>& s, int cc, int i) {
  return s.var(cc, i);
}
template <int N, bool SC> __device__ __forceinline__ float src_var(const GlobalSrc<N, SC>& s, int cc, int i) {
  return s.var_at(cc, i);
}

template <class TF, int P> struct FinalOut {
  TF crlb[P];
  TF llr;
};

// One pixel of the final pass (unified_kernel.jl:359-397): its LLR term
// (camera_models.jl:60-85) and, where the model is > 0, its Fisher
// contribution du du^T / v (v = model, or model + variance for sCMOS) to the
// upper triangle Fu.  du holds the pixel's P model derivatives.
template <int CAM, int P, class TF>
__device__ __forceinline__ void final_pixel(TF d, TF var, TF mu, const TF* du, bool cvalid, TF* Fu, TF& llr) {
  constexpr bool F32 = sizeof(TF) == 4;
  TF v = mu;
  TF lp;
  TF rv = TF(0);   // 1/mu, the ideal camera's Fisher weight
  if constexpr (CAM == IDEAL) {
    // camera_models.jl:60-71, as d log(mu/d) + d - mu with one log:
    // d log(mu/d) = -d log(d (1/mu)), and 1/mu is also the Fisher weight below.
    // Float32 takes 1/mu from v_rcp_f32 (1 ulp); the single log also avoids
    // the cancellation of d log(mu) against d log(d) in the two-log form.
    if constexpr (F32) {
      // ln x as v_log_f32 (log2) times ln 2.  x = d/mu of a pixel with d > 0
      // lies far above the denormal range; the clamp keeps even a denormal
      // product finite (its term is then ~0, as it would be).
      rv = mu > TF(0) ? fast_rcp(mu) : TF(0);
      const float x = fmaxf(d * rv, 1.17549435e-38f);
      const float lnx = __builtin_amdgcn_logf(x) * 0.693147180559945309f;
      lp = (mu > TF(0) && d > TF(0)) ? -d * lnx + (d - mu) : (mu > TF(0) && d == TF(0)) ? -mu : TF(0);
    } else {
      rv = mu > TF(0) ? TF(1) / mu : TF(0);
      lp = (mu > TF(0) && d > TF(0)) ? -d * log(d * rv) + (d - mu) : (mu > TF(0) && d == TF(0)) ? -mu : TF(0);
    }
  } else {
    // camera_models.jl:73-85.  The Fisher weight below divides again, inside
    // its own block: a 1/v shared by both keeps it live across the block, and
    // the multi-shape kernel's largest sCMOS tiles then exceed 256 VGPRs.
    v = mu + var;
    const TF res = d - mu;
    lp = v > TF(0) ? TF(-0.5) * (res * res / v) : TF(0);
  }
  if (cvalid) llr += lp;
  if (cvalid && mu > TF(0)) {
    const TF wv = CAM == IDEAL ? rv : TF(1) / v;
    int q = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const TF ak = du[k] * wv;
#pragma unroll
      for (int l = k; l < P; ++l) Fu[q++] += ak * du[l];
    }
  }
}

// The pixel's model derivatives from its column and row tables
// (psf_derivatives.jl:6-187): x, y, N, bg, and the sigma / z term.
template <int MODEL, class TF>
__device__ __forceinline__ void final_du(TF px, TF py, TF Dx, TF Dy, TF ax, TF ay, TF* du) {
  using Tr = MT<MODEL>;
  du[Tr::IX] = Dx * py;
  du[Tr::IY] = Dy * px;
  du[Tr::IN] = px * py;
  du[Tr::IBG] = TF(1);
  if constexpr (Tr::EXTRA == 2) {
    du[Tr::IE1] = ax * py;
    du[Tr::IE2] = ay * px;
  } else if constexpr (Tr::EXTRA != 0) {
    du[Tr::IE1] = ax * py + ay * px;
  }
}

// Sums across the group, the symmetric Fisher matrix and its inversion.
template <int L, int P, class TF>
__device__ __forceinline__ FinalOut<TF, P> final_reduce(TF* Fu, TF llr) {
  constexpr int NF = P * (P + 1) / 2;
  FinalOut<TF, P> out;
  out.llr = group_sum<L>(llr);
#pragma unroll
  for (int q = 0; q < NF; ++q) Fu[q] = group_sum<L>(Fu[q]);
  TF F[P * P];
  int q = 0;
#pragma unroll
  for (int k = 0; k < P; ++k)
#pragma unroll
    for (int l = k; l < P; ++l) {
      F[k + P * l] = Fu[q];
      F[l + P * k] = Fu[q];
      ++q;
    }
  invert_fisher<P, TF>(F, out.crlb);
  return out;
}

template <int MODEL, int CAM, int N, int L, class TF, class Src, bool TAIL = false>
__device__ __forceinline__ FinalOut<TF, MT<MODEL>::P> final_pass(const float* th, const float* psf, int c0,
                                                               const Src& src, int r = 0) {
  using Tr = MT<MODEL>;
  constexpr int P = Tr::P;
  constexpr int CB = TAIL ? N / L : (N + L - 1) / L;
  constexpr bool SIG = Tr::EXTRA != 0;
  TF thf[P];
#pragma unroll
  for (int k = 0; k < P; ++k) thf[k] = (TF)th[k];
  Widths<TF> wf;
  model_widths<MODEL, TF>(thf, psf, wf);
  const TF nphf = thf[Tr::IN], bgf = thf[Tr::IBG];
  AxisTab<TF, N> fy;
  axis_tables<TF, N, SIG>(thf[Tr::IY], wf.sy, nphf, 0, fy);
  const AxisK<TF> kf = axis_consts<TF>(thf[Tr::IX], wf.sx, nphf);

  constexpr int NF = P * (P + 1) / 2;
  TF Fu[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) Fu[q] = TF(0);
  TF llr = TF(0);
  Edge<TF> flo = edge_eval<TF, false>(kf, c0);
  auto column = [&](int cc) __attribute__((always_inline)) {
    const Edge<TF> fhi = edge_eval<TF, false>(kf, c0 + cc + 1);
    const Pix<TF> fx = pix_from_edges<TF, SIG>(kf, flo, fhi);
    flo = fhi;
    const TF px = fx.phi;
    const bool cvalid = (L == 1) || TAIL || (c0 + cc < N);
    TF ax = TF(0);
    if constexpr (SIG) ax = (MODEL == ASTIG) ? wf.dsx * fx.S : fx.S;
    float dcol[N], vcol[CAM == SCMOS ? N : 1];
    src.column(cc, dcol, vcol);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const TF py = fy.phi[i];
      TF ay = TF(0);
      if constexpr (SIG) ay = (MODEL == ASTIG) ? wf.dsy * fy.S[i] : fy.S[i];
      TF du[P];
      final_du<MODEL, TF>(px, py, fx.D, fy.D[i], ax, ay, du);
      const TF mu = bgf + nphf * px * py;
      final_pixel<CAM, P, TF>((TF)dcol[i], CAM == SCMOS ? (TF)vcol[i] : TF(0), mu, du, cvalid, Fu, llr);
    }
  };
  if constexpr (Src::kInRegisters) {
    // VGPR-resident pixels: the column index must be a compile-time constant.
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) column(cc);
  } else {
    // Pixels from memory: one column at a time keeps only that column's loads
    // in flight.
#pragma unroll 1
    for (int cc = 0; cc < CB; ++cc) column(cc);
  }
  if constexpr (TAIL) {
    // this lane's rows r + L s of the last column, tables from their edges
    static_assert(Src::kInRegisters, "tail pixels are register-resident");
    const Pix<TF> fx = pix_from_edges<TF, SIG>(kf, edge_eval<TF, false>(kf, N - 1), edge_eval<TF, false>(kf, N));
    const AxisK<TF> ky = axis_consts<TF>(thf[Tr::IY], wf.sy, nphf);
    TF ax = TF(0);
    if constexpr (SIG) ax = (MODEL == ASTIG) ? wf.dsx * fx.S : fx.S;
#pragma unroll
    for (int s = 0; s < row_slots<N, L>(); ++s) {
      const int ri = r + L * s;
      const Pix<TF> ry = pix_from_edges<TF, SIG>(ky, edge_eval<TF, false>(ky, ri), edge_eval<TF, false>(ky, ri + 1));
      TF ay = TF(0);
      if constexpr (SIG) ay = (MODEL == ASTIG) ? wf.dsy * ry.S : ry.S;
      TF du[P];
      final_du<MODEL, TF>(fx.phi, ry.phi, fx.D, ry.D, ax, ay, du);
      const TF mu = bgf + nphf * fx.phi * ry.phi;
      final_pixel<CAM, P, TF>((TF)src.dt[s], CAM == SCMOS ? (TF)src.vt[s] : TF(0), mu, du, ri < N, Fu, llr);
    }
  }
  return final_reduce<L, P, TF>(Fu, llr);
}

// The same final pass with rows in the outer loop, for large boxes: each
// row's tables are built from its two edges where the row is used, and a lane
// holds only its CB columns' tables, instead of every row's (which for
// N > 16 would not fit the registers).  Pixels are register-resident.
template <int MODEL, int CAM, int N, int L, class TF, class Src>
__device__ __forceinline__ FinalOut<TF, MT<MODEL>::P> final_pass_rows(const float* th, const float* psf, int c0,
                                                                    const Src& src) {
  static_assert(Src::kInRegisters, "rows outer needs register-resident pixels");
  using Tr = MT<MODEL>;
  constexpr int P = Tr::P;
  constexpr int CB = (N + L - 1) / L;
  constexpr bool SIG = Tr::EXTRA != 0;
  TF thf[P];
#pragma unroll
  for (int k = 0; k < P; ++k) thf[k] = (TF)th[k];
  Widths<TF> wf;
  model_widths<MODEL, TF>(thf, psf, wf);
  const TF nphf = thf[Tr::IN], bgf = thf[Tr::IBG];
  const AxisK<TF> kx = axis_consts<TF>(thf[Tr::IX], wf.sx, nphf);
  const AxisK<TF> ky = axis_consts<TF>(thf[Tr::IY], wf.sy, nphf);
  Pix<TF> cx[CB];
  TF ax[CB];
  {
    Edge<TF> lo = edge_eval<TF, false>(kx, c0);
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const Edge<TF> hi = edge_eval<TF, false>(kx, c0 + cc + 1);
      cx[cc] = pix_from_edges<TF, SIG>(kx, lo, hi);
      lo = hi;
      ax[cc] = SIG ? ((MODEL == ASTIG) ? wf.dsx * cx[cc].S : cx[cc].S) : TF(0);
    }
  }
  constexpr int NF = P * (P + 1) / 2;
  TF Fu[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) Fu[q] = TF(0);
  TF llr = TF(0);
  Edge<TF> rlo = edge_eval<TF, false>(ky, 0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Edge<TF> rhi = edge_eval<TF, false>(ky, i + 1);
    const Pix<TF> ry = pix_from_edges<TF, SIG>(ky, rlo, rhi);
    rlo = rhi;
    TF ay = TF(0);
    if constexpr (SIG) ay = (MODEL == ASTIG) ? wf.dsy * ry.S : ry.S;
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const bool cvalid = (L == 1) || (c0 + cc < N);
      TF du[P];
      final_du<MODEL, TF>(cx[cc].phi, ry.phi, cx[cc].D, ry.D, ax[cc], ay, du);
      const TF mu = bgf + nphf * cx[cc].phi * ry.phi;
      final_pixel<CAM, P, TF>((TF)src.d(cc, i), CAM == SCMOS ? (TF)src.var(cc, i) : TF(0), mu, du, cvalid, Fu,
                              llr);
    }
  }
  return final_reduce<L, P, TF>(Fu, llr);
}

// Out-of-range sCMOS corner: no fit, NaN outputs.
template <class TF, int P> __device__ __forceinline__ void nan_outputs(float* th, FinalOut<TF, P>& fo) {
  const float qnan = __builtin_nanf("");
#pragma unroll
  for (int k = 0; k < P; ++k) {
    th[k] = qnan;
    fo.crlb[k] = TF(qnan);
  }
  fo.llr = TF(qnan);
}

template <int P, class TF>
__device__ __forceinline__ void store_outputs(const FitArgs& a, int64_t out, int stride, const float* th,
                                              const FinalOut<TF, P>& fo) {
#pragma unroll
  for (int k = 0; k < P; ++k) {
    a.params[out * stride + k] = th[k];
    a.crlb[out * stride + k] = (float)fo.crlb[k];
  }
  a.llr[out] = (float)fo.llr;
  for (int k = P; k < stride; ++k) {   // mixed-model batches: unused parameter slots
    a.params[out * stride + k] = __builtin_nanf("");
    a.crlb[out * stride + k] = __builtin_nanf("");
  }
}

// ---------------------------------------------------------------------------
// The kernel.
// ---------------------------------------------------------------------------
// Where one ROI's pixels live and where its results go.
struct RoiRef {
  const float* data;   // the ROI's first pixel (column-major N x N)
  int64_t out;         // output slot
  int stride;          // params / crlb values per slot
  int cx, cy;          // 1-based camera corner (sCMOS)
};

// One ROI, fitted by the L lanes of its group (lane r of the group owns
// columns [r*CB, r*CB + CB)).  Every lane of the wave must call this (the
// cross-lane sums use DPP); `active` lanes store the results.
template <int MODEL, int CAM, int N, int L, int PREC>
__device__ __forceinline__ void fit_one(const FitArgs& a, const RoiRef& ref, bool active, int r) {
  using Tr = MT<MODEL>;
  constexpr int P = Tr::P;
  constexpr bool TAIL = tail_split<MODEL, CAM, N, L>();
  constexpr int CB = block_cols<MODEL, CAM, N, L>();   // full columns per lane
  constexpr int RS = row_slots<N, L>();                 // tail split: rows of the last column per lane
  constexpr bool SIG = Tr::EXTRA != 0;  // sigma / z terms present
  const int c0 = r * CB;

  // ---- load this lane's column block (and variances) once ----------------
  float dv[CB * N];
  float vr[CAM == SCMOS ? CB * N : 1];
  float dt[TAIL ? RS : 1];
  float vt[(TAIL && CAM == SCMOS) ? RS : 1];
  const float* src = ref.data;
  float dmax_l = 0.f;
  int cxc = 1, cyc = 1;
  bool inb = true;   // sCMOS ROI inside the variance map
  if constexpr (CAM == SCMOS) {
    cxc = ref.cx;
    cyc = ref.cy;
    inb = (cxc >= 1) && (cyc >= 1) && (cxc <= a.nx - N + 1) && (cyc <= a.ny - N + 1);
  }
  // Each column is a contiguous run of N floats in the ROI and in the variance
  // map: read in 16-byte loads (load_run), since a wave's lanes are ROIs apart
  // and every load instruction touches one cache line per lane.
#pragma unroll
  for (int cc = 0; cc < CB; ++cc) {
    const int col = c0 + cc;
    const bool valid = (L == 1) || TAIL || (col < N);
    if (valid) {
      load_run<N>(src + col * N, dv + cc * N);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) dv[cc * N + i] = 0.f;
    }
    if constexpr (CAM == SCMOS) {
      if (valid && inb) {
        load_run<N>(a.var_map + (int64_t)(cyc - 1) + (int64_t)a.ny * (int64_t)(cxc - 1 + col), vr + cc * N);
      } else {
#pragma unroll
        for (int i = 0; i < N; ++i) vr[cc * N + i] = 0.f;
      }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) dmax_l = fmaxf(dmax_l, dv[cc * N + i]);
  }
  if constexpr (TAIL) {
    // rows r + L s of the last column; slots past the last row hold zeros
    // (the row index is clamped, so no load leaves the ROI)
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      const int ri = r + L * s;
      const bool ok = ri < N;
      const int rc = ok ? ri : N - 1;
      const float x = src[(N - 1) * N + rc];
      dt[s] = ok ? x : 0.f;
      if constexpr (CAM == SCMOS) {
        if (inb) {
          const float v = a.var_map[(int64_t)(cyc - 1 + rc) + (int64_t)a.ny * (int64_t)(cxc - 1 + N - 1)];
          vt[s] = ok ? v : 0.f;
        } else {
          vt[s] = 0.f;
        }
      }
      dmax_l = fmaxf(dmax_l, dt[s]);
    }
  }
  const float dmax = group_max<L>(dmax_l);

  // ---- initial estimates (unified_kernel.jl:158-255) ---------------------
  float th[P];
  {
    float esum = 0.f;
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const int col = c0 + cc;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const bool edge = (i == 0) || (i == N - 1) || (col == 0) || (col == N - 1);
        if (edge) esum += dv[cc * N + i];   // invalid columns hold zeros
      }
    }
    if constexpr (TAIL) {
#pragma unroll
      for (int s = 0; s < RS; ++s) esum += dt[s];   // the last column is an edge column
    }
    esum = group_sum<L>(esum);
    const float bg0 = esum / (float)(4 * N - 4);
    float tot = 0.f, xw = 0.f, yw = 0.f;
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const bool valid = (L == 1) || TAIL || (c0 + cc < N);
      const float fj = (float)(c0 + cc + 1);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const float s = valid ? fmaxf(0.f, dv[cc * N + i] - bg0) : 0.f;
        tot += s;
        xw += s * fj;
        yw += s * (float)(i + 1);
      }
    }
    if constexpr (TAIL) {
#pragma unroll
      for (int s = 0; s < RS; ++s) {
        const int ri = r + L * s;
        const float v = ri < N ? fmaxf(0.f, dt[s] - bg0) : 0.f;
        tot += v;
        xw += v * (float)N;
        yw += v * (float)(ri + 1);
      }
    }
    tot = group_sum<L>(tot);
    xw = group_sum<L>(xw);
    yw = group_sum<L>(yw);
    const float center = (float)(N + 1) * 0.5f;
    float x0 = tot > 10.f ? xw / tot : center;
    float y0 = tot > 10.f ? yw / tot : center;
    x0 = jclamp(x0, 1.5f, (float)N - 0.5f);
    y0 = jclamp(y0, 1.5f, (float)N - 0.5f);
    const float ph = fmaxf(tot, 100.f);
    th[Tr::IX] = x0;
    th[Tr::IY] = y0;
    th[Tr::IN] = ph;
    th[Tr::IBG] = bg0;
    if constexpr (MODEL == XYNBS) th[4] = 1.3f;
    if constexpr (MODEL == XYNBSXSY) { th[4] = 1.3f; th[5] = 1.3f; }
    if constexpr (MODEL == ASTIG) {
      float mxx = 0.f, myy = 0.f;
#pragma unroll
      for (int cc = 0; cc < CB; ++cc) {
        const bool valid = (L == 1) || TAIL || (c0 + cc < N);
        const float dxp = (float)(c0 + cc + 1) - x0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float s = valid ? fmaxf(0.f, dv[cc * N + i] - bg0) : 0.f;
          const float dyp = (float)(i + 1) - y0;
          mxx += s * dxp * dxp;
          myy += s * dyp * dyp;
        }
      }
      if constexpr (TAIL) {
        const float dxp = (float)N - x0;
#pragma unroll
        for (int s = 0; s < RS; ++s) {
          const int ri = r + L * s;
          const float v = ri < N ? fmaxf(0.f, dt[s] - bg0) : 0.f;
          const float dyp = (float)(ri + 1) - y0;
          mxx += v * dxp * dxp;
          myy += v * dyp * dyp;
        }
      }
      mxx = group_sum<L>(mxx);
      myy = group_sum<L>(myy);
      float z0 = 0.f;
      if (tot > 10.f) {
        const float sxe = sqrtf(fmaxf(0.5f, mxx / tot));
        const float sye = sqrtf(fmaxf(0.5f, myy / tot));
        const float rx = sxe / a.psf[0], ry = sye / a.psf[1];
        const float dal = rx * rx - ry * ry;
        z0 = -dal * (a.psf[7] * a.psf[7]) / (4.f * a.psf[6]);
        z0 = jclamp(z0, -600.f, 600.f);
      }
      th[2] = z0;
    }
  }

  // ---- Newton-Raphson iterations (unified_kernel.jl:314-351) -------------
  for (int it = 0; it < a.iterations; ++it) {
    Widths<float> w;
    model_widths<MODEL, float>(th, a.psf, w);
    const float nph = th[Tr::IN];
    const float bg = th[Tr::IBG];

    // Poisson fast path: every pixel has model > 0.01 and no cap can bind.
    // model >= bg - N*tiny for N <= 1e5 (phi >= 0 up to erf rounding), so
    // bg >= 0.05 keeps it above 0.01, and dmax <= 5000 bg min(bg,1) keeps
    // data/model - 1 and data/model^2 below 1e4.
    bool fast = false;
    if constexpr (CAM == IDEAL) {
      bool fin = true;
#pragma unroll
      for (int k = 0; k < P; ++k) fin = fin && (fabsf(th[k]) < 3.0e38f);
      fast = fin && (bg >= 0.05f) && (nph >= 0.f) && (nph <= 1.0e5f) &&
             (dmax <= 5000.f * bg * fminf(bg, 1.f));
    }

    float G[P], H[P];
    const AxisK<float> kr = axis_consts<float>(th[Tr::IY], w.sy, nph);
    constexpr bool SHARE = share_rows<MODEL, CAM, N, L>();
    if constexpr (row_outer<MODEL, CAM, N, L>()) {
      const AxisK<float> kc = axis_consts<float>(th[Tr::IX], w.sx, nph);
      if (fast) {
        accumulate_rows<MODEL, CAM, N, L, true, SHARE>(dv, vr, dt, vt, c0, nph, bg, kr, r, kc, w, G, H);
      } else {
        accumulate_rows<MODEL, CAM, N, L, false, SHARE>(dv, vr, dt, vt, c0, nph, bg, kr, r, kc, w, G, H);
      }
    } else {
      if constexpr (L == 1) __builtin_amdgcn_s_setprio(0);   // tables and pixel sums: normal priority
      AxisTab<float, N> ry;
      const AxisK<float> kc = axis_consts<float>(th[Tr::IX], w.sx, nph);
      // tail split: the last column's rows r + L s, with this lane's own row
      // tables, folded here into Gt / Ht
      float Gt[P], Ht[P];
#pragma unroll
      for (int k = 0; k < P; ++k) { Gt[k] = 0.f; Ht[k] = 0.f; }
      if constexpr (SHARE) {
        Pix<float> own[RS];
        shared_axis_tables<N, L, SIG>(kr, r, ry, own);
        if constexpr (TAIL) {
          const Pix<float> tl = pix_from_edges<float, SIG>(kc, edge_eval<float, SIG>(kc, N - 1),
                                                           edge_eval<float, SIG>(kc, N));
          const float ncx_t = nph * tl.phi;
          float St[11], rs[6];
#pragma unroll
          for (int q = 0; q < 11; ++q) St[q] = 0.f;
#pragma unroll
          for (int q = 0; q < 6; ++q) rs[q] = 0.f;
          if (fast) {
#pragma unroll
            for (int s = 0; s < RS; ++s)
              tail_pixel<MODEL, CAM, true>(own[s], w, dt[s], CAM == SCMOS ? vt[s] : 0.f, r + L * s < N, ncx_t, bg,
                                           St, rs);
            tail_fold<MODEL, CAM, true>(St, rs, tl, w, Gt, Ht);
          } else {
#pragma unroll
            for (int s = 0; s < RS; ++s)
              tail_pixel<MODEL, CAM, false>(own[s], w, dt[s], CAM == SCMOS ? vt[s] : 0.f, r + L * s < N, ncx_t, bg,
                                            St, rs);
            tail_fold<MODEL, CAM, false>(St, rs, tl, w, Gt, Ht);
          }
        }
      } else {
        axis_tables<float, N, SIG>(th[Tr::IY], w.sy, nph, 0, ry);
      }

      // Row extras: A = d(model)/d(extra) on the row axis, AA its second derivative.
      float Ay[SIG ? N : 1], AAy[SIG ? N : 1];
      if constexpr (SIG) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          if constexpr (MODEL == ASTIG) {
            Ay[i] = w.dsy * ry.S[i];
            AAy[i] = ry.SS[i] * (w.dsy * w.dsy) + ry.S[i] * w.d2sy;
          } else {
            Ay[i] = ry.S[i];
            AAy[i] = ry.SS[i];
          }
        }
      }

      // Row sums for the q = cf + 1 form of the fast path.
      float sphi = 0.f, sD = 0.f, sD2 = 0.f, sA = 0.f, sAA = 0.f;
      if (fast) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          sphi += ry.phi[i];
          sD += ry.D[i];
          sD2 += ry.D2[i];
          if constexpr (SIG) { sA += Ay[i]; sAA += AAy[i]; }
        }
      }

      // One lane per ROI: the edge tables and pixel sums run at the normal
      // wave priority, and the short dependent chain between them (the Newton
      // step, then the next iteration's widths and axis constants) at a
      // raised one, so that the SIMD's other wave fills that chain's stalls
      // with its own pixel work.  (gfx950, XYNB 11 px, 1M ROIs: 1.403 ->
      // 1.374 ms per launch; raising the tables too gave 1.393, raising the
      // pixel sums instead 1.431.)  Setting priority 0 again here as well
      // keeps the compiler's scheduling regions as measured.
      if constexpr (L == 1) __builtin_amdgcn_s_setprio(0);
      if (fast) {
        accumulate<MODEL, CAM, N, L, true>(dv, vr, c0, nph, bg, ry, kc, w, Ay, AAy, sphi, sD, sD2, sA, sAA, G, H);
      } else {
        accumulate<MODEL, CAM, N, L, false>(dv, vr, c0, nph, bg, ry, kc, w, Ay, AAy, sphi, sD, sD2, sA, sAA, G, H);
      }
      if constexpr (L == 1) __builtin_amdgcn_s_setprio(1);
      if constexpr (TAIL) {
#pragma unroll
        for (int k = 0; k < P; ++k) { G[k] += Gt[k]; H[k] += Ht[k]; }
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      G[k] = group_sum<L>(G[k]);
      H[k] = group_sum<L>(H[k]);
    }
    // Scalar Newton step with constraints (unified_kernel.jl:337-350, constraints.jl:172-177).
    // G / H is a v_rcp_f32 and a multiply (1 ulp more than the division) where
    // both are below 1e30 in magnitude; a lane with a NaN, an infinity or a
    // value beyond that takes the IEEE division, so those cases come out as
    // the division gives them.  The division runs only in a wave that has
    // such a lane.
    float step[P];
    bool exact = true;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float g = G[k], h = H[k];
      step[k] = fabsf(h) > 1e-10f ? g * __builtin_amdgcn_rcpf(h) : 0.f;
      exact = exact && fabsf(g) < 1e30f && fabsf(h) < 1e30f;
    }
    if (__builtin_amdgcn_ballot_w64(!exact) != 0) {
#pragma unroll
      for (int k = 0; k < P; ++k)
        if (!exact) step[k] = fabsf(H[k]) > 1e-10f ? G[k] / H[k] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float ms = a.max_step[k];
      const float t = th[k] - jclamp(step[k], -ms, ms);
      th[k] = jclamp(t, a.lower[k], a.upper[k]);
    }
  }

  if constexpr (L == 1) __builtin_amdgcn_s_setprio(0);
  // ---- final pass: LLR + Fisher + CRLB (unified_kernel.jl:353-433) --------
  // Float32 here; the Float64 final pass runs as its own kernel
  // (final_pass_kernel) so that its registers do not lower this kernel's
  // occupancy.
  if constexpr (PREC == PREC_FP32) {
    const RegSrc<N, CAM == SCMOS> srcr{dv, vr, dt, vt};
    FinalOut<float, P> fo;
    if constexpr (N > kMaxSmallBox) {
      fo = final_pass_rows<MODEL, CAM, N, L, float>(th, a.psf, c0, srcr);
    } else {
      fo = final_pass<MODEL, CAM, N, L, float, RegSrc<N, CAM == SCMOS>, TAIL>(th, a.psf, c0, srcr, r);
    }
    if (!inb) nan_outputs(th, fo);
    if (active && r == 0) store_outputs<P>(a, ref.out, ref.stride, th, fo);
  } else {
    if (!inb) {
#pragma unroll
      for (int k = 0; k < P; ++k) th[k] = __builtin_nanf("");
    }
    if (active && r == 0) {
#pragma unroll
      for (int k = 0; k < P; ++k) a.params[ref.out * ref.stride + k] = th[k];
    }
  }
}

// The ROI that lane group `j` of a launch fits, and whether it stores.
// Inactive groups (past the end, or skipped batch slots) fit a valid ROI of
// the same shape and discard the result.
template <int MODEL, int CAM, int N>
__device__ __forceinline__ bool locate_roi(const FitArgs& a, int64_t j, int64_t fallback, RoiRef& ref) {
  bool active = j < a.n_rois;
  int64_t roi = active ? j : fallback;
  int64_t off = roi * (int64_t)(N * N);
  if (a.order) {
    const int32_t k = a.order[a.first + roi];
    active = active && (k >= 0);
    roi = (k >= 0) ? k : a.order[a.first + fallback];
    off = a.offsets[roi];
  }
  ref.data = a.data + off;
  ref.out = roi;
  ref.stride = a.out_stride > 0 ? a.out_stride : MT<MODEL>::P;
  ref.cx = ref.cy = 1;
  if constexpr (CAM == SCMOS) {
    ref.cx = a.x_corners[roi];
    ref.cy = a.y_corners[roi];
  }
  return active;
}

// One launch over n ROIs of one shape: 64/L ROIs per wave, in wave-sized
// tiles (a grid smaller than the tile count strides over the rest).
template <int MODEL, int CAM, int N, int L, int PREC>
__global__ __launch_bounds__(kWave) void fit_kernel(FitArgs a) {
  constexpr int RPW = kWave / L;
  const int lane = threadIdx.x;
  const int r = (L == 1) ? 0 : (lane & (L - 1));
  // A batch slot range whose first entry is empty holds no ROI of this shape.
  if (a.n_rois <= 0 || (a.order && __builtin_amdgcn_readfirstlane(a.order[a.first]) < 0)) return;
  for (int64_t base = (int64_t)blockIdx.x * RPW; base < a.n_rois; base += (int64_t)gridDim.x * RPW) {
    RoiRef ref;
    const bool active = locate_roi<MODEL, CAM, N>(a, base + lane / L, 0, ref);
    fit_one<MODEL, CAM, N, L, PREC>(a, ref, active, r);
  }
}

// Float64 final pass at the estimates the iteration kernel wrote to params,
// one lane per ROI; the ROI is re-read from HBM (once) instead of being held
// in VGPRs.  Two waves per SIMD are requested: from 12 px up the compiler then
// spills a few dozen values to scratch (astig 15 px: 284 -> 256 VGPRs, 29
// spilled), which costs less than running one wave per SIMD (gfx950: the
// astigmatic 15 px pass 18 % faster).
template <int MODEL, int CAM, int N>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void final_pass_kernel(FitArgs a) {
  constexpr int P = MT<MODEL>::P;
  if (a.n_rois <= 0 || (a.order && __builtin_amdgcn_readfirstlane(a.order[a.first]) < 0)) return;
  RoiRef ref;
  const bool active = locate_roi<MODEL, CAM, N>(a, (int64_t)blockIdx.x * kWave + threadIdx.x, 0, ref);
  if (!active) return;
  const int64_t slot = ref.out;
  float th[P];
#pragma unroll
  for (int k = 0; k < P; ++k) th[k] = a.params[slot * ref.stride + k];
  GlobalSrc<N, CAM == SCMOS> src;
  src.data = ref.data;
  bool inb = true;
  if constexpr (CAM == SCMOS) {
    const int cx = ref.cx, cy = ref.cy;
    inb = (cx >= 1) && (cy >= 1) && (cx <= a.nx - N + 1) && (cy <= a.ny - N + 1);
    src.var = a.var_map + (int64_t)(cy - 1) + (int64_t)a.ny * (int64_t)(cx - 1);
    src.ny = a.ny;
    if (!inb) src.var = nullptr;
  }
  FinalOut<double, P> fo;
  if (inb) {
    fo = final_pass<MODEL, CAM, N, 1, double>(th, a.psf, 0, src);
  } else {
    nan_outputs(th, fo);
  }
  store_outputs<P>(a, slot, ref.stride, th, fo);
  if (a.crlb64) {
#pragma unroll
    for (int k = 0; k < P; ++k) a.crlb64[slot * ref.stride + k] = fo.crlb[k];
    for (int k = P; k < ref.stride; ++k) a.crlb64[slot * ref.stride + k] = __builtin_nan("");
  }
  if (a.llr64) a.llr64[slot] = fo.llr;
}

// The Float64 final pass of large boxes (N > kMaxSmallBox): L lanes per ROI
// as in the iteration kernel, each holding its column block in VGPRs (read
// once from HBM in 16-byte runs), rows in the outer loop.
template <int MODEL, int CAM, int N, int L>
__global__ __launch_bounds__(kWave) void final_pass_lanes_kernel(FitArgs a) {
  constexpr int P = MT<MODEL>::P;
  constexpr int RPW = kWave / L;
  constexpr int CB = (N + L - 1) / L;
  const int lane = threadIdx.x;
  const int r = lane & (L - 1);
  const int c0 = r * CB;
  if (a.n_rois <= 0 || (a.order && __builtin_amdgcn_readfirstlane(a.order[a.first]) < 0)) return;
  for (int64_t base = (int64_t)blockIdx.x * RPW; base < a.n_rois; base += (int64_t)gridDim.x * RPW) {
    RoiRef ref;
    const bool active = locate_roi<MODEL, CAM, N>(a, base + lane / L, 0, ref);
    const int64_t slot = ref.out;
    float th[P];
#pragma unroll
    for (int k = 0; k < P; ++k) th[k] = a.params[slot * ref.stride + k];
    bool inb = true;
    if constexpr (CAM == SCMOS)
      inb = (ref.cx >= 1) && (ref.cy >= 1) && (ref.cx <= a.nx - N + 1) && (ref.cy <= a.ny - N + 1);
    float dv[CB * N];
    float vr[CAM == SCMOS ? CB * N : 1];
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) {
      const int col = c0 + cc;
      if (col < N) {
        load_run<N>(ref.data + col * N, dv + cc * N);
      } else {
#pragma unroll
        for (int i = 0; i < N; ++i) dv[cc * N + i] = 0.f;
      }
      if constexpr (CAM == SCMOS) {
        if (col < N && inb) {
          load_run<N>(a.var_map + (int64_t)(ref.cy - 1) + (int64_t)a.ny * (int64_t)(ref.cx - 1 + col), vr + cc * N);
        } else {
#pragma unroll
          for (int i = 0; i < N; ++i) vr[cc * N + i] = 0.f;
        }
      }
    }
    const RegSrc<N, CAM == SCMOS> src{dv, vr};
    FinalOut<double, P> fo = final_pass_rows<MODEL, CAM, N, L, double>(th, a.psf, c0, src);
    if (!inb) nan_outputs(th, fo);
    if (active && r == 0) {
      store_outputs<P>(a, slot, ref.stride, th, fo);
      if (a.crlb64) {
#pragma unroll
        for (int k = 0; k < P; ++k) a.crlb64[slot * ref.stride + k] = fo.crlb[k];
        for (int k = P; k < ref.stride; ++k) a.crlb64[slot * ref.stride + k] = __builtin_nan("");
      }
      if (a.llr64) a.llr64[slot] = fo.llr;
    }
  }
}

template <int MODEL, int CAM, int N, int L, int PREC>
bool describe_fit(KernelInfo* k) {
  k->fit = reinterpret_cast<const void*>(&fit_kernel<MODEL, CAM, N, L, PREC>);
  k->rois_per_block = kWave / L;
  k->final_pass = nullptr;
  k->final_rois_per_block = kWave;
  if constexpr (PREC == PREC_MIXED) {
    if constexpr (N > kMaxSmallBox) {
      k->final_pass = reinterpret_cast<const void*>(&final_pass_lanes_kernel<MODEL, CAM, N, L>);
      k->final_rois_per_block = kWave / L;
    } else {
      k->final_pass = reinterpret_cast<const void*>(&final_pass_kernel<MODEL, CAM, N>);
    }
  }
  return true;
}

template <int MODEL, int CAM, int N, int L, int PREC>
hipError_t launch_fit(const FitArgs& a, hipStream_t stream) {
  constexpr int RPW = kWave / L;
  const int64_t blocks = (a.n_rois + RPW - 1) / RPW;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((fit_kernel<MODEL, CAM, N, L, PREC>), dim3((unsigned)blocks), dim3(kWave), 0, stream, a);
  if constexpr (PREC == PREC_MIXED) {
    if constexpr (N > kMaxSmallBox) {
      hipLaunchKernelGGL((final_pass_lanes_kernel<MODEL, CAM, N, L>), dim3((unsigned)blocks), dim3(kWave), 0, stream,
                         a);
    } else {
      const int64_t fblocks = (a.n_rois + kWave - 1) / kWave;
      hipLaunchKernelGGL((final_pass_kernel<MODEL, CAM, N>), dim3((unsigned)fblocks), dim3(kWave), 0, stream, a);
    }
  }
  return hipGetLastError();
}

}  // namespace gmle
