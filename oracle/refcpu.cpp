// refcpu.cpp -- C ABI over the CPU restatement in refcpu.hpp.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg as the checker / timed CPU baseline.  The HIP
// library never links or calls it.
//
// Precision selector `prec`:
//   0 = all Float32 (the reference's CPU path, src/unified_kernel.jl:285-434)
//   1 = all Float64 (FP64-everywhere oracle)
//   2 = Float32 iterations, Float64 final pass (mixed-precision extension)
// PSF parameter array `psf` (8 floats, pixel units after to_pixel_units,
// src/psf_models.jl:311-331): XYNB -> psf[0] = sigma; ASTIG -> psf[0..7] =
// sigma_x0, sigma_y0, Ax, Ay, Bx, By, gamma, d; XYNBS / XYNBSXSY ignore it.
#include "refcpu.hpp"

#include <omp.h>

#include <cstring>

using namespace refcpu;

namespace {

Psf<float> make_psf(int model, const float* p) {
  Psf<float> q;
  std::memset(&q, 0, sizeof(q));
  q.model = model;
  if (model == XYNB) {
    q.sigma = p[0];
  } else if (model == ASTIG) {
    q.sx0 = p[0]; q.sy0 = p[1]; q.Ax = p[2]; q.Ay = p[3];
    q.Bx = p[4]; q.By = p[5]; q.gamma = p[6]; q.d = p[7];
  }
  return q;
}

template <class T>
double pixel_derivs_t(int model, const float* psf, int i, int j, const double* theta, double* dudt, double* d2) {
  const Psf<T> p = psf_cast<T>(make_psf(model, psf));
  const int P = n_params(model);
  T th[6] = {}, du[6] = {}, dd[6] = {};
  for (int k = 0; k < P; ++k) th[k] = (T)theta[k];
  const T m = pixel_derivatives(i, j, th, p, du, dd);
  for (int k = 0; k < P; ++k) { dudt[k] = du[k]; d2[k] = dd[k]; }
  return m;
}

}  // namespace

extern "C" {

int refcpu_n_params(int model) { return n_params(model); }

double refcpu_integral_gaussian_1d(int prec, int i, double pos, double sigma) {
  if (prec == 0) return integral_gaussian_1d<float>(i, (float)pos, (float)sigma);
  return integral_gaussian_1d<double>(i, pos, sigma);
}

// Model value, dudt[P] and diagonal d2[P] at pixel (i, j) (1-based row, column).
double refcpu_pixel_derivs(int prec, int model, const float* psf, int i, int j, const double* theta,
                           double* dudt, double* d2) {
  if (prec == 0) return pixel_derivs_t<float>(model, psf, i, j, theta, dudt, d2);
  return pixel_derivs_t<double>(model, psf, i, j, theta, dudt, d2);
}

// Newton terms and LLR term of one pixel.  camera 0 = Poisson, 1 = sCMOS.
void refcpu_pixel_terms(int prec, int camera, double data, double model, double var, double* cf, double* df,
                        double* llr) {
  if (prec == 0) {
    float c, d;
    if (camera) likelihood_terms_scmos<float>((float)data, (float)model, (float)var, c, d);
    else likelihood_terms_poisson<float>((float)data, (float)model, c, d);
    *cf = c; *df = d;
    *llr = camera ? llr_scmos<float>((float)data, (float)model, (float)var)
                  : llr_poisson<float>((float)data, (float)model);
  } else {
    if (camera) likelihood_terms_scmos<double>(data, model, var, *cf, *df);
    else likelihood_terms_poisson<double>(data, model, *cf, *df);
    *llr = camera ? llr_scmos<double>(data, model, var) : llr_poisson<double>(data, model);
  }
}

// Regularise + Cholesky, LU fallback, Inf (src/unified_kernel.jl:399-431).
// F is P x P column-major and is not modified.  Returns the path taken.
int refcpu_invert_fisher(int prec, int P, const double* F, double* crlb) {
  if (prec == 0) {
    float Ff[36], c[6];
    for (int q = 0; q < P * P; ++q) Ff[q] = (float)F[q];
    const int path = invert_fisher<float>(Ff, P, c);
    for (int k = 0; k < P; ++k) crlb[k] = c[k];
    return path;
  }
  double Fd[36];
  for (int q = 0; q < P * P; ++q) Fd[q] = F[q];
  return invert_fisher<double>(Fd, P, crlb);
}

// Raw inverses without regularisation, for pinning against a library inverse.
// Return 1 on success, 0 when the factorisation fails.
int refcpu_cholesky_inverse(int prec, int P, const double* A, double* Ainv) {
  if (prec == 0) {
    float L[36], Ai[36];
    for (int q = 0; q < P * P; ++q) L[q] = (float)A[q];
    if (!cholesky_decomposition<float>(L, P)) return 0;
    cholesky_inverse<float>(Ai, L, P);
    for (int q = 0; q < P * P; ++q) Ainv[q] = Ai[q];
    return 1;
  }
  double L[36];
  for (int q = 0; q < P * P; ++q) L[q] = A[q];
  if (!cholesky_decomposition<double>(L, P)) return 0;
  cholesky_inverse<double>(Ainv, L, P);
  return 1;
}

int refcpu_lu_inverse(int prec, int P, const double* A, double* Ainv) {
  if (prec == 0) {
    float Af[36], Ai[36];
    for (int q = 0; q < P * P; ++q) Af[q] = (float)A[q];
    const bool ok = lu_matrix_inverse<float>(Ai, Af, P);
    for (int q = 0; q < P * P; ++q) Ainv[q] = Ai[q];
    return ok ? 1 : 0;
  }
  return lu_matrix_inverse<double>(Ainv, A, P) ? 1 : 0;
}

void refcpu_initialize(int prec, int model, const float* psf, int box, const float* roi, double* theta) {
  const int P = n_params(model);
  if (prec == 0) {
    float th[6];
    initialize<float>(roi, box, make_psf(model, psf), th);
    for (int k = 0; k < P; ++k) theta[k] = th[k];
  } else {
    double r[32 * 32], th[6];
    for (int q = 0; q < box * box; ++q) r[q] = roi[q];
    initialize<double>(r, box, psf_cast<double>(make_psf(model, psf)), th);
    for (int k = 0; k < P; ++k) theta[k] = th[k];
  }
}

// Fit n_rois ROIs of one box size.  data is (box, box, n_rois) column-major
// Float32; outputs are (P, n_rois) / (n_rois) in Float64 (Float32 results are
// converted exactly).  var_map (ny, nx) column-major, used when camera == 1.
// x_corners / y_corners are 1-based camera coordinates.  paths (optional) gets
// the inversion path per ROI.  nthreads <= 0 uses all OpenMP threads.
// Returns 0, or -1 on a bad argument.
int refcpu_fit(int prec, int model, const float* psf, int box, int64_t n_rois, const float* data, int camera,
               const float* var_map, int ny, int nx, const int32_t* x_corners, const int32_t* y_corners,
               const float* lower, const float* upper, const float* max_step, int iterations, double* params,
               double* crlb, double* llr, int32_t* paths, int nthreads) {
  if (box < 3 || box > 32 || n_rois < 0 || model < 0 || model > 3 || prec < 0 || prec > 2) return -1;
  if (camera && (!var_map || !x_corners || !y_corners)) return -1;
  const int P = n_params(model);
  const Psf<float> p = make_psf(model, psf);
  Bounds b;
  for (int k = 0; k < P; ++k) {
    b.lower[k] = lower[k]; b.upper[k] = upper[k]; b.max_step[k] = max_step[k];
  }
  if (camera) {
    for (int64_t r = 0; r < n_rois; ++r) {
      const int cx = x_corners[r], cy = y_corners[r];
      if (cx < 1 || cy < 1 || cx + box - 1 > nx || cy + box - 1 > ny) return -1;
    }
  }
  const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (int64_t r = 0; r < n_rois; ++r) {
    const float* roi = data + (size_t)r * box * box;
    const int cx = x_corners ? x_corners[r] : 1;
    const int cy = y_corners ? y_corners[r] : 1;
    int path;
    if (prec == 0) {
      float th[6], cr[6], l;
      path = fit_roi<float, float>(roi, box, p, camera != 0, var_map, ny, cx, cy, b, iterations, th, cr, &l);
      for (int k = 0; k < P; ++k) { params[r * P + k] = th[k]; crlb[r * P + k] = cr[k]; }
      llr[r] = l;
    } else if (prec == 1) {
      path = fit_roi<double, double>(roi, box, p, camera != 0, var_map, ny, cx, cy, b, iterations,
                                     params + r * P, crlb + r * P, llr + r);
    } else {
      path = fit_roi<float, double>(roi, box, p, camera != 0, var_map, ny, cx, cy, b, iterations,
                                    params + r * P, crlb + r * P, llr + r);
    }
    if (paths) paths[r] = path;
  }
  return 0;
}

// Coverage counters of the Float32 (reference) fit of n_rois Poisson ROIs:
// branches (n_rois x 4) as fit_roi documents.  Test infrastructure only.
int refcpu_poisson_branches(int model, const float* psf, int box, int64_t n_rois, const float* data,
                            const float* lower, const float* upper, const float* max_step, int iterations,
                            int32_t* branches, int nthreads) {
  if (box < 3 || box > 32 || n_rois < 0 || model < 0 || model > 3 || !branches) return -1;
  const int P = n_params(model);
  const Psf<float> p = make_psf(model, psf);
  Bounds b;
  for (int k = 0; k < P; ++k) {
    b.lower[k] = lower[k]; b.upper[k] = upper[k]; b.max_step[k] = max_step[k];
  }
  const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (int64_t r = 0; r < n_rois; ++r) {
    float th[6], cr[6], l;
    int32_t* br = branches + 4 * r;
    br[0] = br[1] = br[2] = br[3] = 0;
    fit_roi<float, float>(data + (size_t)r * box * box, box, p, false, nullptr, 0, 1, 1, b, iterations, th, cr, &l,
                          br);
  }
  return 0;
}

// The final pass at given estimates theta (n_rois x P, Float64): LLR, Fisher,
// inversion and CRLB in Float32 (prec 0) or Float64 (prec 1 / 2).
int refcpu_final_pass(int prec, int model, const float* psf, int box, int64_t n_rois, const float* data,
                      int camera, const float* var_map, int ny, int nx, const int32_t* x_corners,
                      const int32_t* y_corners, const double* theta, double* crlb, double* llr, int nthreads) {
  if (box < 3 || box > 32 || n_rois < 0 || model < 0 || model > 3) return -1;
  if (camera && (!var_map || !x_corners || !y_corners)) return -1;
  (void)nx;
  const int P = n_params(model);
  const Psf<float> p = make_psf(model, psf);
  const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (int64_t r = 0; r < n_rois; ++r) {
    const float* roi = data + (size_t)r * box * box;
    const int cx = x_corners ? x_corners[r] : 1;
    const int cy = y_corners ? y_corners[r] : 1;
    if (prec == 0) {
      float cr[6], l;
      final_pass<float>(roi, box, p, camera != 0, var_map, ny, cx, cy, theta + r * P, cr, &l);
      for (int k = 0; k < P; ++k) crlb[r * P + k] = cr[k];
      llr[r] = l;
    } else {
      final_pass<double>(roi, box, p, camera != 0, var_map, ny, cx, cy, theta + r * P, crlb + r * P, llr + r);
    }
  }
  return 0;
}

}  // extern "C"
