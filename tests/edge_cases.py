"""Crafted ROIs that drive the reference's numerical edge branches: the
Poisson model threshold and caps (src/camera_models.jl:23-32) and every
outcome of the Fisher inversion (src/unified_kernel.jl:399-431): Cholesky,
the no-pivot LU fallback and the Inf CRLB of a singular matrix.

Each group is built so that the branch taken does not depend on rounding
(exact zeros, exactly representable entries, definite signs), so the GPU and
the oracle must take the same branch, not merely land close:
  zero            all-zero ROIs: background 0, every model value only the
                  spot (tails below the 0.01 threshold); Cholesky
  constant        flat ROIs (no signal): N falls to its lower bound; Cholesky
  point_fp        one hot pixel of 256 counts, PSF sigma 0.02 px (every pixel
                  edge lies beyond |u| = 17, where exp(-u^2) underflows in
                  Float32 and erf rounds to +-1 in both precisions, so the
                  pixel integrals are exactly 1 and 0), background free down
                  to 0: only that pixel has model > 0, its
                  derivative vector is exactly (0, 0, 1, 1) with weight 2^-8,
                  so the regularised Fisher matrix holds the exact singular
                  block [[a, a], [a, a]] (a = 2^-8) beside two ridge-only
                  pivots.  Float32: the 1e-10 ridge rounds away on that block,
                  Cholesky meets a zero pivot and LU a zero pivot -> Inf.
                  Float64: the ridge survives -> Cholesky, large finite CRLB.
  no_signal       N and bg pinned at 0 by their bounds: every model value is
                  0, the Fisher matrix is all zero, the LU tolerance is 0 and
                  0/0 gives NaN (Julia's max(0, NaN) keeps it)
  neg_variance    sCMOS with a variance map of -3 max(data) everywhere: every
                  weight 1 / (model + var) is negative, the Fisher matrix is
                  negative definite; Cholesky fails at the first pivot, LU
                  succeeds and every diagonal of the inverse is negative ->
                  CRLB 0 (sqrt(max(0, .)))
  zero iterations where a Newton step on such data would be meaningless
  (the final pass is evaluated at the initial estimate).
"""
from __future__ import annotations

import numpy as np

from gaussmle_amd import synth
from gaussmle_amd.models import GaussianXYNB, GaussianXYNBS, ParameterConstraints, default_constraints

F32 = np.float32


def _cons(psf, bounds: dict):
    """Default constraints (box 11) with parameter k bounded to bounds[k]."""
    c = default_constraints(psf, 11)
    lo, up, ms = (np.array(a, dtype=F32) for a in (c.lower, c.upper, c.max_step))
    for k, (a, b) in bounds.items():
        lo[k], up[k] = a, b
    return ParameterConstraints(lo, up, ms)


def fisher_cases(n: int = 64, seed: int = 500) -> dict:
    rng = np.random.default_rng(seed)
    out = {}
    box = 11
    psf = GaussianXYNB(1.3)
    nan_truth = lambda P: np.full((P, n), np.nan)  # noqa: E731

    zero = np.zeros((box, box, n), dtype=F32, order="F")
    out["zero_xynb_11"] = synth.Workload("zero", psf, box, zero, nan_truth(4), default_constraints(psf, 11), 20)
    out["zero_xynbs_9"] = synth.Workload("zero_s", GaussianXYNBS(), 9, np.zeros((9, 9, n), F32, order="F"),
                                         nan_truth(5), default_constraints(GaussianXYNBS(), 11), 20)

    const = np.empty((box, box, n), dtype=F32, order="F")
    const[...] = rng.integers(1, 50, n).astype(F32)[None, None, :]
    out["constant_xynb_11"] = synth.Workload("const", psf, box, const, nan_truth(4), default_constraints(psf, 11), 20)

    point = np.zeros((box, box, n), dtype=F32, order="F")
    ii = rng.integers(3, box - 3, n)
    jj = rng.integers(3, box - 3, n)
    point[ii, jj, np.arange(n)] = 256.0
    tiny = GaussianXYNB(0.02)
    out["point_fp_xynb_11"] = synth.Workload("point", tiny, box, point, nan_truth(4), _cons(tiny, {3: (0.0, 1e30)}),
                                             20)

    noise = np.empty((box, box, n), dtype=F32, order="F")
    noise[...] = rng.poisson(5.0, (box, box, n))
    out["no_signal_xynb_11"] = synth.Workload("nosig", psf, box, noise, nan_truth(4),
                                              _cons(psf, {2: (0.0, 0.0), 3: (0.0, 0.0)}), 20)

    w = synth.config0(n=n, seed=seed + 1)
    cam = 32
    var = np.full((cam, cam), -3.0 * float(w.data.max()), dtype=F32)
    xc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    yc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    out["neg_variance_xynb_11"] = synth.Workload("negvar", w.psf_pixels, box, w.data, w.truth, w.constraints, 0,
                                                 var, xc, yc)
    return out


def lowbg_cases() -> dict:
    """Poisson workloads on which the masked branch of the likelihood runs:
    low background (threshold), bright spots with interior hot pixels (both
    caps), a bright-spot workload above the fast-path background bound, and a
    width-fitting model."""
    from gaussmle_amd.models import GaussianXYNBSXSY
    return {
        "xynb_lowbg_11": lambda: synth.low_background(GaussianXYNB(1.3), 2048, 11, seed=400),
        "xynbs_lowbg_9": lambda: synth.low_background(GaussianXYNBS(), 2048, 9, seed=401),
        "xynb_hot_15": lambda: synth.low_background(GaussianXYNB(1.3), 2048, 15, seed=402, bg_range=(0.002, 0.02),
                                                    log_n_range=(2e4, 1e5), hot_fraction=0.5),
        "xynb_bright_13": lambda: synth.low_background(GaussianXYNB(1.3), 2048, 13, seed=404, bg_range=(0.05, 0.5),
                                                       log_n_range=(1e4, 1e5), bg_lower=0.01),
        "xynbsxsy_lowbg_11": lambda: synth.low_background(GaussianXYNBSXSY(), 1024, 11, seed=403),
    }


def crlb_pattern(c: np.ndarray) -> np.ndarray:
    """Per value: 'n' NaN, 'i' infinite, '0' zero, 'f' finite non-zero."""
    c = np.asarray(c, dtype=np.float64)
    return np.where(np.isnan(c), "n", np.where(np.isinf(c), "i", np.where(c == 0, "0", "f")))
