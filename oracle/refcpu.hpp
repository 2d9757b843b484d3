// refcpu.hpp -- CPU restatement of GaussMLE.jl's per-ROI maximum-likelihood fit.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or as the timed CPU
// baseline -- never as a product code path.
//
// It restates, operation for operation and branch for branch, the reference's
// KernelAbstractions kernel and its callees (paths relative to the
// GaussMLE.jl v0.2.1 source tree):
//   src/unified_kernel.jl:7-60     Cholesky factor + inverse
//   src/unified_kernel.jl:64-137   no-pivot LU fallback
//   src/unified_kernel.jl:158-255  initial estimates
//   src/unified_kernel.jl:285-434  the fit (Newton loop, Fisher, LLR, CRLB)
//   src/psf_derivatives.jl:6-187   model value + first/diagonal-second derivatives
//   src/gausslib/baselibrary.jl:8-67  pixel-integrated Gaussian and derivatives
//   src/camera_models.jl:19-96     Poisson / sCMOS Newton terms and LLR terms
//   src/constraints.jl:49-54       step clamp then box clamp
//
// Templated on the arithmetic type T:
//   T = float  : the reference's own (all-Float32) CPU path.  Julia semantics are
//                mirrored: no FMA contraction (build with -ffp-contract=off), Julia
//                literal powers x^2 = x*x and x^3 = x*x*x, Float32^Int for other
//                exponents and Float32^Float32 evaluated in Float64 then rounded.
//   T = double : the FP64-everywhere oracle (same algorithm, constants, branches
//                and fixed iteration count).
// fit_roi<float, double> is the mixed variant: Float32 Newton iterations and a
// Float64 final pass (LLR, Fisher matrix, inversion, CRLB).
//
// The transcendental functions come from the C library (erff/expf/logf and the
// double versions).  The reference takes erf from SpecialFunctions.jl (OpenLibm
// erff) and exp/log from Julia's own Float32 implementations; none of those can
// run here, so bit-level agreement of erf/exp/log with the reference is not
// pinned (both sides are within about one ulp of the true value).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

namespace refcpu {

enum Model { XYNB = 0, XYNBS = 1, XYNBSXSY = 2, ASTIG = 3 };

inline int n_params(int model) {
  switch (model) {
    case XYNB: return 4;
    case XYNBS: return 5;
    case XYNBSXSY: return 6;
    case ASTIG: return 5;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Julia arithmetic helpers
// ---------------------------------------------------------------------------
template <class T> struct J;

template <> struct J<float> {
  static float erf_(float x) { return erff(x); }
  static float exp_(float x) { return expf(x); }
  static float log_(float x) { return logf(x); }
  static float sqrt_(float x) { return sqrtf(x); }
  // Float32 ^ Int (exponent not in {-2..3}): power by squaring in Float64, then rounded.
  static float powi(float x, int n) { return (float)std::pow((double)x, n); }
  // Float32 ^ Float32: evaluated in Float64, then rounded.
  static float powf_(float x, float y) { return (float)std::pow((double)x, (double)y); }
  static float pi() { return (float)M_PI; }
};

template <> struct J<double> {
  static double erf_(double x) { return std::erf(x); }
  static double exp_(double x) { return std::exp(x); }
  static double log_(double x) { return std::log(x); }
  static double sqrt_(double x) { return std::sqrt(x); }
  static double powi(double x, int n) { return std::pow(x, n); }
  static double powf_(double x, double y) { return std::pow(x, y); }
  static double pi() { return M_PI; }
};

// Julia's clamp(x, lo, hi) = ifelse(x > hi, hi, ifelse(x < lo, lo, x)); NaN passes through.
template <class T> inline T jclamp(T x, T lo, T hi) { return x > hi ? hi : (x < lo ? lo : x); }
// Julia's min/max on floats propagate NaN.
template <class T> inline T jmin(T x, T y) {
  if (std::isnan(x) || std::isnan(y)) return std::numeric_limits<T>::quiet_NaN();
  return y < x ? y : x;
}
template <class T> inline T jmax(T x, T y) {
  if (std::isnan(x) || std::isnan(y)) return std::numeric_limits<T>::quiet_NaN();
  return y > x ? y : x;
}

// ---------------------------------------------------------------------------
// PSF parameters in pixel units (after to_pixel_units, src/psf_models.jl:311-331)
// ---------------------------------------------------------------------------
template <class T> struct Psf {
  int model;
  T sigma;                    // XYNB fixed sigma (pixels)
  T sx0, sy0, Ax, Ay, Bx, By, gamma, d;  // astigmatic calibration
};

template <class T, class U> inline Psf<T> psf_cast(const Psf<U>& p) {
  Psf<T> q;
  q.model = p.model;
  q.sigma = (T)p.sigma; q.sx0 = (T)p.sx0; q.sy0 = (T)p.sy0;
  q.Ax = (T)p.Ax; q.Ay = (T)p.Ay; q.Bx = (T)p.Bx; q.By = (T)p.By;
  q.gamma = (T)p.gamma; q.d = (T)p.d;
  return q;
}

// ---------------------------------------------------------------------------
// src/gausslib/baselibrary.jl:8-14  integral_gaussian_1d
// ---------------------------------------------------------------------------
template <class T> inline T integral_gaussian_1d(int i, T position, T sigma) {
  const T half = T(1) / T(2);
  const T norm = half / (sigma * sigma);
  const T s = J<T>::sqrt_(norm);
  const T fi = (T)i;
  return half * (J<T>::erf_((fi - position + half) * s) - J<T>::erf_((fi - position - half) * s));
}

// src/gausslib/baselibrary.jl:21-33  derivative_integral_gaussian_1d
template <class T>
inline void derivative_integral_gaussian_1d(int ii, T x, T sigma, T N, T PSFy, T& dudt, T& d2udt2) {
  const T half = T(1) / T(2);
  const T two = T(1) + T(1);
  const T fi = (T)ii;
  const T ua = (fi + half - x) / sigma;
  const T ub = (fi - half - x) / sigma;
  const T factor_a = J<T>::exp_(-half * (ua * ua));
  const T factor_b = J<T>::exp_(-half * (ub * ub));
  const T constant = -N / J<T>::sqrt_(two * T(1) * J<T>::pi()) / sigma;
  dudt = constant * (factor_a - factor_b) * PSFy;
  d2udt2 = constant / (sigma * sigma) * ((fi + half - x) * factor_a - (fi - half - x) * factor_b) * PSFy;
  (void)two;
}

// src/gausslib/baselibrary.jl:40-52  derivative_integral_gaussian_1d_sigma
template <class T>
inline void derivative_integral_gaussian_1d_sigma(int i, T x, T Sx, T N, T PSFy, T& dudt, T& d2udt2) {
  const T half = T(1) / T(2);
  const T two = T(1) + T(1);
  const T pi_val = T(1) * J<T>::pi();
  const T fi = (T)i;
  const T ua = (fi + half - x) / Sx;
  const T ub = (fi - half - x) / Sx;
  const T ax = J<T>::exp_(-half * (ua * ua));
  const T bx = J<T>::exp_(-half * (ub * ub));
  const T p = fi - x + half;   // (i - x + half)
  const T m = fi - x - half;   // (i - x - half)
  const T sq2pi = J<T>::sqrt_(two * pi_val);
  dudt = -N / sq2pi / Sx / Sx * (ax * p - bx * m) * PSFy;
  d2udt2 = -two / Sx * dudt - N / sq2pi / J<T>::powi(Sx, 5) * (ax * (p * p * p) - bx * (m * m * m)) * PSFy;
}

// src/gausslib/baselibrary.jl:61-67  derivative_integral_gaussian_2d_sigma
template <class T>
inline void derivative_integral_gaussian_2d_sigma(int i, int j, T x, T y, T S, T N, T PSFx, T PSFy,
                                                  T& dudt, T& d2udt2) {
  T dSx, ddSx, dSy, ddSy;
  derivative_integral_gaussian_1d_sigma(j, x, S, N, PSFy, dSx, ddSx);
  derivative_integral_gaussian_1d_sigma(i, y, S, N, PSFx, dSy, ddSy);
  dudt = dSx + dSy;
  d2udt2 = ddSx + ddSy;
}

// ---------------------------------------------------------------------------
// src/psf_derivatives.jl:6-187  compute_pixel_derivatives (all four models)
// theta layout: XYNB [x,y,N,bg]; XYNBS [x,y,N,bg,s]; XYNBSXSY [x,y,N,bg,sx,sy];
// ASTIG [x,y,z,N,bg].  (i, j) = (row = y, column = x), 1-based.
// ---------------------------------------------------------------------------
template <class T>
inline T pixel_derivatives(int i, int j, const T* th, const Psf<T>& psf, T* dudt, T* d2) {
  switch (psf.model) {
    case XYNB: {  // psf_derivatives.jl:6-37
      const T x = th[0], y = th[1], N = th[2], bg = th[3];
      const T psf_x = integral_gaussian_1d(j, x, psf.sigma);
      const T psf_y = integral_gaussian_1d(i, y, psf.sigma);
      T dx, d2x, dy, d2y;
      derivative_integral_gaussian_1d(j, x, psf.sigma, N, psf_y, dx, d2x);
      derivative_integral_gaussian_1d(i, y, psf.sigma, N, psf_x, dy, d2y);
      const T model = bg + N * psf_x * psf_y;
      dudt[0] = dx; dudt[1] = dy; dudt[2] = psf_x * psf_y; dudt[3] = T(1);
      d2[0] = d2x; d2[1] = d2y; d2[2] = T(0); d2[3] = T(0);
      return model;
    }
    case XYNBS: {  // psf_derivatives.jl:40-73
      const T x = th[0], y = th[1], N = th[2], bg = th[3], s = th[4];
      const T psf_x = integral_gaussian_1d(j, x, s);
      const T psf_y = integral_gaussian_1d(i, y, s);
      T dx, d2x, dy, d2y, ds, d2s;
      derivative_integral_gaussian_1d(j, x, s, N, psf_y, dx, d2x);
      derivative_integral_gaussian_1d(i, y, s, N, psf_x, dy, d2y);
      derivative_integral_gaussian_2d_sigma(i, j, x, y, s, N, psf_x, psf_y, ds, d2s);
      const T model = bg + N * psf_x * psf_y;
      dudt[0] = dx; dudt[1] = dy; dudt[2] = psf_x * psf_y; dudt[3] = T(1); dudt[4] = ds;
      d2[0] = d2x; d2[1] = d2y; d2[2] = T(0); d2[3] = T(0); d2[4] = d2s;
      return model;
    }
    case XYNBSXSY: {  // psf_derivatives.jl:76-112
      const T x = th[0], y = th[1], N = th[2], bg = th[3], sx = th[4], sy = th[5];
      const T psf_x = integral_gaussian_1d(j, x, sx);
      const T psf_y = integral_gaussian_1d(i, y, sy);
      T dx, d2x, dy, d2y, dsx, d2sx, dsy, d2sy;
      derivative_integral_gaussian_1d(j, x, sx, N, psf_y, dx, d2x);
      derivative_integral_gaussian_1d(i, y, sy, N, psf_x, dy, d2y);
      derivative_integral_gaussian_1d_sigma(j, x, sx, N, psf_y, dsx, d2sx);
      derivative_integral_gaussian_1d_sigma(i, y, sy, N, psf_x, dsy, d2sy);
      const T model = bg + N * psf_x * psf_y;
      dudt[0] = dx; dudt[1] = dy; dudt[2] = psf_x * psf_y; dudt[3] = T(1); dudt[4] = dsx; dudt[5] = dsy;
      d2[0] = d2x; d2[1] = d2y; d2[2] = T(0); d2[3] = T(0); d2[4] = d2sx; d2[5] = d2sy;
      return model;
    }
    default: {  // ASTIG, psf_derivatives.jl:115-187
      const T x = th[0], y = th[1], z = th[2], N = th[3], bg = th[4];
      const T z_minus_gamma = z - psf.gamma;
      const T z_plus_gamma = z + psf.gamma;
      const T zdx = z_minus_gamma / psf.d;
      const T ax_ = T(1) + zdx * zdx + psf.Ax * (zdx * zdx * zdx) + psf.Bx * J<T>::powi(zdx, 4);
      const T zdy = z_plus_gamma / psf.d;
      const T ay_ = T(1) + zdy * zdy + psf.Ay * (zdy * zdy * zdy) + psf.By * J<T>::powi(zdy, 4);
      const T sx = psf.sx0 * J<T>::sqrt_(ax_);
      const T sy = psf.sy0 * J<T>::sqrt_(ay_);
      const T psf_x = integral_gaussian_1d(j, x, sx);
      const T psf_y = integral_gaussian_1d(i, y, sy);
      T dx, d2x, dy, d2y, dsx, d2sx, dsy, d2sy;
      derivative_integral_gaussian_1d(j, x, sx, N, psf_y, dx, d2x);
      derivative_integral_gaussian_1d(i, y, sy, N, psf_x, dy, d2y);
      derivative_integral_gaussian_1d_sigma(j, x, sx, N, psf_y, dsx, d2sx);
      derivative_integral_gaussian_1d_sigma(i, y, sy, N, psf_x, dsy, d2sy);
      const T dax = (T(2) * zdx + T(3) * psf.Ax * (zdx * zdx) + T(4) * psf.Bx * (zdx * zdx * zdx)) / psf.d;
      const T day = (T(2) * zdy + T(3) * psf.Ay * (zdy * zdy) + T(4) * psf.By * (zdy * zdy * zdy)) / psf.d;
      const T dsx_dz = (psf.sx0 / (T(2) * J<T>::sqrt_(ax_))) * dax;
      const T dsy_dz = (psf.sy0 / (T(2) * J<T>::sqrt_(ay_))) * day;
      const T dudt_z = dsx * dsx_dz + dsy * dsy_dz;
      const T d2ax = (T(2) + T(6) * psf.Ax * zdx + T(12) * psf.Bx * (zdx * zdx)) / (psf.d * psf.d);
      const T d2ay = (T(2) + T(6) * psf.Ay * zdy + T(12) * psf.By * (zdy * zdy)) / (psf.d * psf.d);
      const T d2sx_dz2 = -(psf.sx0 / (T(4) * J<T>::powf_(ax_, T(1.5)))) * (dax * dax) +
                         (psf.sx0 / (T(2) * J<T>::sqrt_(ax_))) * d2ax;
      const T d2sy_dz2 = -(psf.sy0 / (T(4) * J<T>::powf_(ay_, T(1.5)))) * (day * day) +
                         (psf.sy0 / (T(2) * J<T>::sqrt_(ay_))) * d2ay;
      const T d2z = d2sx * (dsx_dz * dsx_dz) + dsx * d2sx_dz2 + d2sy * (dsy_dz * dsy_dz) + dsy * d2sy_dz2;
      const T model = bg + N * psf_x * psf_y;
      dudt[0] = dx; dudt[1] = dy; dudt[2] = dudt_z; dudt[3] = psf_x * psf_y; dudt[4] = T(1);
      d2[0] = d2x; d2[1] = d2y; d2[2] = d2z; d2[3] = T(0); d2[4] = T(0);
      return model;
    }
  }
}

// ---------------------------------------------------------------------------
// src/camera_models.jl:19-42  Newton terms (cf, df)
// ---------------------------------------------------------------------------
template <class T> inline void likelihood_terms_poisson(T data, T model, T& cf, T& df) {
  if (model > T(0.01)) {
    cf = data / model - T(1);
    df = data / (model * model);
    cf = jmin(cf, T(1e4));
    df = jmin(df, T(1e4));
  } else {
    cf = T(0);
    df = T(0);
  }
}

template <class T> inline void likelihood_terms_scmos(T data, T model, T var, T& cf, T& df) {
  const T total_var = model + var;
  cf = (data - model) / total_var;
  df = T(1) / total_var;
}

// src/camera_models.jl:60-85  log-likelihood-ratio terms
template <class T> inline T llr_poisson(T data, T model) {
  if (model > T(0) && data > T(0)) {
    return data * J<T>::log_(model) - model - (data * J<T>::log_(data) - data);
  } else if (model > T(0) && data == T(0)) {
    return -model;
  }
  return T(0);
}

template <class T> inline T llr_scmos(T data, T model, T var) {
  const T total_var = model + var;
  if (total_var > T(0)) {
    const T r = data - model;
    return -T(0.5) * ((r * r) / total_var);
  }
  return T(0);
}

// ---------------------------------------------------------------------------
// src/unified_kernel.jl:7-60  Cholesky factor (in place) and inverse.
// Matrices are P x P, column-major like Julia: A[i + P*j] is A[i, j] (0-based).
// ---------------------------------------------------------------------------
template <class T> inline bool cholesky_decomposition(T* A, int P) {
  for (int j = 0; j < P; ++j) {
    for (int i = j; i < P; ++i) {
      T s = A[i + P * j];
      for (int k = 0; k < j; ++k) s -= A[i + P * k] * A[j + P * k];
      if (i == j) {
        if (s <= T(0)) return false;
        A[i + P * i] = J<T>::sqrt_(s);
      } else {
        A[i + P * j] = s / A[j + P * j];
      }
    }
  }
  return true;
}

template <class T> inline bool cholesky_inverse(T* Ainv, const T* L, int P) {
  T Li[36];
  for (int q = 0; q < 36; ++q) Li[q] = T(0);
  for (int j = 0; j < P; ++j) {
    Li[j + P * j] = T(1) / L[j + P * j];
    for (int i = j + 1; i < P; ++i) {
      T s = T(0);
      for (int k = j; k < i; ++k) s += L[i + P * k] * Li[k + P * j];
      Li[i + P * j] = -s / L[i + P * i];
    }
  }
  for (int i = 0; i < P; ++i) {
    for (int j = i; j < P; ++j) {
      T s = T(0);
      for (int k = j; k < P; ++k) s += Li[k + P * i] * Li[k + P * j];
      Ainv[i + P * j] = s;
      if (i != j) Ainv[j + P * i] = s;
    }
  }
  return true;
}

// src/unified_kernel.jl:64-82  LU without pivoting, relative pivot tolerance
template <class T> inline bool lu_decomposition(T* A, int P) {
  T mx = T(0);
  for (int q = 0; q < P * P; ++q) mx = jmax(mx, std::fabs(A[q]));
  const T tol = T(1e-10) * mx;
  for (int k = 0; k < P; ++k) {
    if (std::fabs(A[k + P * k]) < tol) return false;
    for (int i = k + 1; i < P; ++i) {
      A[i + P * k] /= A[k + P * k];
      for (int j = k + 1; j < P; ++j) A[i + P * j] -= A[i + P * k] * A[k + P * j];
    }
  }
  return true;
}

// src/unified_kernel.jl:85-137  back substitution and LU-based inverse
template <class T> inline bool lu_matrix_inverse(T* Ainv, const T* A, int P) {
  T LU[36];
  for (int q = 0; q < P * P; ++q) LU[q] = A[q];
  if (!lu_decomposition(LU, P)) {
    for (int i = 0; i < P; ++i)
      for (int j = 0; j < P; ++j)
        Ainv[i + P * j] = (i == j) ? std::numeric_limits<T>::infinity() : T(0);
    return false;
  }
  T x[6];
  for (int col = 0; col < P; ++col) {
    for (int i = 0; i < P; ++i) {
      x[i] = (i == col) ? T(1) : T(0);
      for (int j = 0; j < i; ++j) x[i] -= LU[i + P * j] * x[j];
    }
    for (int i = P - 1; i >= 0; --i) {
      for (int j = i + 1; j < P; ++j) x[i] -= LU[i + P * j] * x[j];
      x[i] /= LU[i + P * i];
    }
    for (int i = 0; i < P; ++i) Ainv[i + P * col] = x[i];
  }
  return true;
}

// Fisher inversion as done at src/unified_kernel.jl:399-431.  F is modified in place
// (regularised).  Writes the diagonal-based CRLB.  Returns 0 = Cholesky, 1 = LU
// fallback, 2 = singular (CRLB = Inf).
template <class T> inline int invert_fisher(T* F, int P, T* crlb) {
  T mx = T(0);
  for (int q = 0; q < P * P; ++q) mx = jmax(mx, std::fabs(F[q]));
  const T reg = T(1e-10) * mx;
  for (int k = 0; k < P; ++k) F[k + P * k] += reg;
  T L[36], Finv[36];
  for (int q = 0; q < P * P; ++q) L[q] = F[q];
  if (cholesky_decomposition(L, P) && cholesky_inverse(Finv, L, P)) {
    for (int k = 0; k < P; ++k) crlb[k] = J<T>::sqrt_(jmax(T(0), Finv[k + P * k]));
    return 0;
  }
  if (lu_matrix_inverse(Finv, F, P)) {
    for (int k = 0; k < P; ++k) crlb[k] = J<T>::sqrt_(jmax(T(0), Finv[k + P * k]));
    return 1;
  }
  for (int k = 0; k < P; ++k) crlb[k] = std::numeric_limits<T>::infinity();
  return 2;
}

// ---------------------------------------------------------------------------
// src/unified_kernel.jl:158-255  initial estimates.  roi is column-major n x n:
// roi[(i-1) + n*(j-1)] is pixel (row i, column j), 1-based.
// ---------------------------------------------------------------------------
template <class T> inline void initialize(const T* roi, int n, const Psf<T>& psf, T* th) {
  T edge_sum = T(0);
  int edge_count = 0;
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= n; ++i)
      if (i == 1 || i == n || j == 1 || j == n) {
        edge_sum += roi[(i - 1) + n * (j - 1)];
        edge_count += 1;
      }
  const T bg = edge_sum / (T)edge_count;
  T total = T(0), xw = T(0), yw = T(0);
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= n; ++i) {
      const T s = jmax(T(0), roi[(i - 1) + n * (j - 1)] - bg);
      total += s;
      xw += s * (T)j;
      yw += s * (T)i;
    }
  const T center = (T)((double)(n + 1) / 2.0);
  T x = total > T(10) ? xw / total : center;
  T y = total > T(10) ? yw / total : center;
  x = jclamp(x, T(1.5), (T)((double)n - 0.5));
  y = jclamp(y, T(1.5), (T)((double)n - 0.5));
  const T photons = jmax(total, T(100));
  switch (psf.model) {
    case XYNB: th[0] = x; th[1] = y; th[2] = photons; th[3] = bg; return;
    case XYNBS: th[0] = x; th[1] = y; th[2] = photons; th[3] = bg; th[4] = T(1.3); return;
    case XYNBSXSY:
      th[0] = x; th[1] = y; th[2] = photons; th[3] = bg; th[4] = T(1.3); th[5] = T(1.3);
      return;
    default: {
      T tot = T(0), Mxx = T(0), Myy = T(0);
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= n; ++i) {
          const T s = jmax(T(0), roi[(i - 1) + n * (j - 1)] - bg);
          tot += s;
          const T dx = (T)j - x;
          const T dy = (T)i - y;
          Mxx += s * dx * dx;
          Myy += s * dy * dy;
        }
      T z;
      if (tot > T(10)) {
        const T sxe = J<T>::sqrt_(jmax(T(0.5), Mxx / tot));
        const T sye = J<T>::sqrt_(jmax(T(0.5), Myy / tot));
        const T rx = sxe / psf.sx0, ry = sye / psf.sy0;
        const T dalpha = rx * rx - ry * ry;
        z = -dalpha * (psf.d * psf.d) / (T(4) * psf.gamma);
        z = jclamp(z, T(-600), T(600));
      } else {
        z = T(0);
      }
      th[0] = x; th[1] = y; th[2] = z; th[3] = photons; th[4] = bg;
      return;
    }
  }
}

// ---------------------------------------------------------------------------
// One ROI end-to-end: src/unified_kernel.jl:298-433.
//   TI : arithmetic type of the initialisation and Newton iterations
//   TF : arithmetic type of the final pass (LLR, Fisher, inversion, CRLB)
// The reference is fit_roi<float, float>; the FP64 oracle fit_roi<double, double>;
// the mixed-precision extension fit_roi<float, double>.
// var (sCMOS only) is the full (ny, nx) column-major variance map; (cx, cy) are
// the 1-based camera coordinates of the ROI's first pixel.  Returns the
// inversion path (0 Cholesky, 1 LU, 2 singular).
// ---------------------------------------------------------------------------
struct Bounds {
  float lower[6], upper[6], max_step[6];
};

// branches (optional, test coverage only): over all Newton iterations, the
// Poisson pixel terms that took [0] the model <= 0.01 branch, [1] the cf cap,
// [2] the df cap (camera_models.jl:23-32), and [3] the iterations in which the
// GPU kernel's precondition for skipping those branches (bg >= 0.05,
// 0 <= N <= 1e5, max(data) <= 5000 bg min(bg, 1), finite parameters) fails,
// i.e. in which the kernel evaluates them.  Not part of the reference.
template <class TI, class TF>
inline int fit_roi(const float* roi_f32, int n, const Psf<float>& psf_f32, bool scmos, const float* var,
                   int ny, int cx, int cy, const Bounds& cons, int iterations, TF* theta_out, TF* crlb_out,
                   TF* llr_out, int32_t* branches = nullptr) {
  (void)ny;
  const int P = n_params(psf_f32.model);
  const Psf<TI> psf = psf_cast<TI>(psf_f32);
  TI roi[32 * 32];
  for (int q = 0; q < n * n; ++q) roi[q] = (TI)roi_f32[q];
  auto var_at = [&](int i, int j) -> float {
    const int cam_i = cy + i - 1, cam_j = cx + j - 1;
    return var[(size_t)(cam_i - 1) + (size_t)ny * (size_t)(cam_j - 1)];
  };

  TI th[6];
  initialize(roi, n, psf, th);
  TI grad[6], hd[6], dudt[6], d2[6];
  TI dmax = TI(0);
  for (int q = 0; q < n * n; ++q) dmax = roi[q] > dmax ? roi[q] : dmax;
  const int inb = psf.model == ASTIG ? 3 : 2, ibg = psf.model == ASTIG ? 4 : 3;
  for (int it = 0; it < iterations; ++it) {
    if (branches && !scmos) {
      bool fin = true;
      for (int k = 0; k < P; ++k) fin = fin && std::fabs(th[k]) < TI(3.0e38);
      const TI bg = th[ibg], nph = th[inb];
      const bool skip = fin && bg >= TI(0.05) && nph >= TI(0) && nph <= TI(1.0e5) &&
                        dmax <= TI(5000) * bg * (bg < TI(1) ? bg : TI(1));
      if (!skip) branches[3] += 1;
    }
    for (int k = 0; k < P; ++k) { grad[k] = TI(0); hd[k] = TI(0); }
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const TI model = pixel_derivatives(i, j, th, psf, dudt, d2);
        const TI data = roi[(i - 1) + n * (j - 1)];
        TI cf, df;
        if (scmos) likelihood_terms_scmos(data, model, (TI)var_at(i, j), cf, df);
        else likelihood_terms_poisson(data, model, cf, df);
        if (branches && !scmos) {
          if (!(model > TI(0.01))) {
            branches[0] += 1;
          } else {
            if (data / model - TI(1) > TI(1e4)) branches[1] += 1;
            if (data / (model * model) > TI(1e4)) branches[2] += 1;
          }
        }
        for (int k = 0; k < P; ++k) {
          grad[k] += dudt[k] * cf;
          hd[k] += d2[k] * cf - dudt[k] * dudt[k] * df;
        }
      }
    TI step[6];
    for (int k = 0; k < P; ++k) step[k] = std::fabs(hd[k]) > TI(1e-10) ? grad[k] / hd[k] : TI(0);
    // src/constraints.jl:172-177
    for (int k = 0; k < P; ++k) {
      const TI ms = (TI)cons.max_step[k];
      const TI t = th[k] - jclamp(step[k], -ms, ms);
      th[k] = jclamp(t, (TI)cons.lower[k], (TI)cons.upper[k]);
    }
  }

  // Final pass in TF.
  const Psf<TF> psf2 = psf_cast<TF>(psf_f32);
  TF th2[6], du2[6], dd2[6];
  for (int k = 0; k < P; ++k) th2[k] = (TF)th[k];
  TF llr = TF(0);
  TF F[36];
  for (int q = 0; q < 36; ++q) F[q] = TF(0);
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= n; ++i) {
      const TF model = pixel_derivatives(i, j, th2, psf2, du2, dd2);
      const TF data = (TF)roi_f32[(i - 1) + n * (j - 1)];
      if (!scmos) llr += llr_poisson(data, model);
      else llr += llr_scmos(data, model, (TF)var_at(i, j));
      if (model > TF(0)) {
        const TF v = scmos ? model + (TF)var_at(i, j) : model;
        for (int k = 0; k < P; ++k)
          for (int l = k; l < P; ++l) {
            const TF f = du2[k] * du2[l] / v;
            F[k + P * l] += f;
            if (k != l) F[l + P * k] += f;
          }
      }
    }
  const int path = invert_fisher(F, P, crlb_out);
  for (int k = 0; k < P; ++k) theta_out[k] = th2[k];
  *llr_out = llr;
  return path;
}

// The final pass alone (src/unified_kernel.jl:353-431) at a given estimate:
// LLR, Fisher matrix, inversion and CRLB in arithmetic type TF.
template <class TF>
inline int final_pass(const float* roi_f32, int n, const Psf<float>& psf_f32, bool scmos, const float* var,
                      int ny, int cx, int cy, const double* theta, TF* crlb_out, TF* llr_out) {
  const int P = n_params(psf_f32.model);
  const Psf<TF> psf2 = psf_cast<TF>(psf_f32);
  auto var_at = [&](int i, int j) -> float {
    const int cam_i = cy + i - 1, cam_j = cx + j - 1;
    return var[(size_t)(cam_i - 1) + (size_t)ny * (size_t)(cam_j - 1)];
  };
  TF th2[6], du2[6], dd2[6];
  for (int k = 0; k < P; ++k) th2[k] = (TF)theta[k];
  TF llr = TF(0);
  TF F[36];
  for (int q = 0; q < 36; ++q) F[q] = TF(0);
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= n; ++i) {
      const TF model = pixel_derivatives(i, j, th2, psf2, du2, dd2);
      const TF data = (TF)roi_f32[(i - 1) + n * (j - 1)];
      if (!scmos) llr += llr_poisson(data, model);
      else llr += llr_scmos(data, model, (TF)var_at(i, j));
      if (model > TF(0)) {
        const TF v = scmos ? model + (TF)var_at(i, j) : model;
        for (int k = 0; k < P; ++k)
          for (int l = k; l < P; ++l) {
            const TF f = du2[k] * du2[l] / v;
            F[k + P * l] += f;
            if (k != l) F[l + P * k] += f;
          }
      }
    }
  *llr_out = llr;
  return invert_fisher(F, P, crlb_out);
}

}  // namespace refcpu
