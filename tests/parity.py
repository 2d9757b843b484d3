"""Parity statistics between a GPU fit and the CPU oracle.

Tolerances (stated once, used by every GPU test):
  * x, y            : |gpu - oracle_fp32| <= 1e-3 px
  * other parameters: |gpu - oracle_fp64| <= 1e-4 * max(|oracle_fp64|, floor), floor = 1 for z and
                      bg (photons / pixel) and 0 for N and widths, i.e. relative 1e-4 except that
                      a background or z within 1 of zero is held to 1e-4 absolute
  * CRLB            : |gpu / oracle_fp64 - 1| <= 1e-4
  * LLR             : mixed precision: |gpu - oracle_fp64| <= max(1e-4 |oracle_fp64|, 1e-3);
                      Float32 mode: |gpu - oracle_fp32| <= max(1e-4 |oracle_fp32|, 1e-3, 2 s),
                      with s the ROI's rounding scale of that formula when given (llr_scale:
                      2^-24 sum_pixels (d |log d| + d), the size of the terms it cancels), and
                      either the Float32 or the Float64 oracle's LLR as reference (the Float32
                      oracle's LLR also carries the reference's erf rounding of tail pixels,
                      which the kernel's pixel integrals do not).  The
                      absolute floor: the reference's Float32 formula d log(mu) - mu - (d log(d) - d)
                      cancels terms of size d log(d) and carries ~1e-4 absolute rounding noise, and a
                      mixed-precision LLR is taken at the Float32 iterations' estimate, which on an
                      ROI not converged after the fixed iterations (3x3 boxes) moves it by ~2e-4.
                      The LLR only enters the chi^2 = -2 LLR of the p-value.

Well-posed ROIs.  A fixed-iteration Newton fit of a noisy or badly initialised
ROI may not converge; on such an ROI the reference's own Float32 path (the
Float32 oracle) already misses the Float64 oracle by more than the tolerances
above, because Float32 rounding sends the two down slightly different
trajectories.  Some ROIs are ill-conditioned even in exact arithmetic (an
astigmatic z near focus, a width on a few pixels): there the Float64 result
itself moves by more than the tolerances when every pixel value moves by one
Float32 ulp, so no Float32 implementation can be held to them.  An ROI is
well-posed when
  (1) the Float32 oracle meets every tolerance against the Float64 oracle, and
  (2) in mixed precision, when given, the oracle's own mixed mode meets the
      LLR and CRLB tolerances against the Float64 oracle (both are evaluated
      in Float64 at the Float32 iterations' estimate; the Float32 CRLB of (1)
      is not what this mode computes);
  (3) when given, the Float64 oracle on the data perturbed by a few ulp per
      pixel (perturb_ulp: 4 ulp, about the rounding of a Float32 sum over a
      hundred pixel terms, sqrt(121)/2 ulp) meets every tolerance against the
      unperturbed Float64 oracle;
  (4) in Float32 mode, when given (ref32fp), the Float32 final pass of the
      oracle at its own Float32 estimate, on the ulp-perturbed data, meets the
      CRLB tolerance against the Float64 oracle: a Float32 CRLB that moves by
      1e-4 under rounding at that level (a background pinned near zero, whose
      Fisher entry sums 1/model over tail pixels) cannot be held to it;
  (5) in Float32 mode, when given (ref32g), the oracle's Float32 final pass at
      the GPU's own estimate meets the CRLB tolerance against the Float64
      oracle: where the reference's Float32 arithmetic, evaluated at the very
      point the GPU reports, already misses it (an astigmatic z walking at its
      step clamp on a large ROI), the Float32 CRLB there is rounding-limited.
      The point only selects where the Float32 evaluation is judged; the GPU's
      CRLB is still compared with the Float64 oracle.
On well-posed ROIs the GPU must meet the tolerances (at most OUTLIER_FRACTION of
them may miss); on the others the GPU result must be finite wherever the
Float32 oracle's is, and the deviations are reported.
"""
from __future__ import annotations

import numpy as np

POS_TOL = 1e-3
REL_TOL = 1e-4
LLR_TOL = 1e-4
OUTLIER_FRACTION = 2e-3


def _rel(a, b):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.abs(a / b - 1.0)


def llr_rounding_scale(w) -> np.ndarray:
    """Per ROI: 2^-24 times the size of the terms the reference's Float32 LLR
    formula d log(mu) - mu - (d log d - d) cancels, sum_pixels (d |log d| + d)."""
    d = np.asarray(w.data, np.float64).reshape(-1, w.n, order="F")
    return 2.0 ** -24 * np.sum(np.where(d > 0, d * np.abs(np.log(np.where(d > 0, d, 1.0))), 0.0) + d, axis=0)


def _errors(model_id: int, res: dict, ref32: dict, ref64: dict, precision: str,
            llr_scale: np.ndarray | None = None) -> dict:
    gp, gc, gl = (np.asarray(res[k], np.float64) for k in ("params", "crlb", "llr"))
    n, P = gp.shape
    e = {"pos": np.abs(gp[:, :2] - ref32["params"][:, :2]).max(axis=1)}
    ibg = 4 if model_id == 3 else 3
    for k in range(2, P):
        floor = 1.0 if (k == ibg or (model_id == 3 and k == 2)) else 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            e[f"p{k}"] = np.abs(gp[:, k] - ref64["params"][:, k]) / np.maximum(floor, np.abs(ref64["params"][:, k]))
    # CRLB per parameter, relative to the Float64 oracle (both +Inf: agreed)
    ce = _rel(gc, ref64["crlb"])
    ce[np.isinf(gc) & np.isinf(ref64["crlb"])] = 0.0
    for k in range(P):
        e[f"crlb{k}"] = ce[:, k]
    if precision == "mixed":
        e["llr"] = np.abs(gl - ref64["llr"]) / np.maximum(np.abs(ref64["llr"]), 10.0)
    elif llr_scale is None:
        # expressed in units of the tolerance scale, compared against LLR_TOL
        e["llr"] = np.abs(gl - ref32["llr"]) / np.maximum(np.abs(ref32["llr"]), 10.0)
    else:
        floor = np.maximum(10.0, 2.0 * np.asarray(llr_scale) / LLR_TOL)
        e32 = np.abs(gl - ref32["llr"]) / np.maximum(np.abs(ref32["llr"]), floor)
        e64 = np.abs(gl - ref64["llr"]) / np.maximum(np.abs(ref64["llr"]), floor)
        e["llr"] = np.minimum(e32, e64)
    return e


def _bad(e: dict) -> np.ndarray:
    bad = ~(e["pos"] <= POS_TOL)
    for k, v in e.items():
        if k == "pos":
            continue
        bad |= ~(v <= (LLR_TOL if k == "llr" else REL_TOL))
    return bad


PERTURB_ULP = 4


def perturb_ulp(data: np.ndarray, ulps: int = PERTURB_ULP, seed: int = 12345) -> np.ndarray:
    """Every nonzero Float32 value moved `ulps` ulp up or down (seeded signs)."""
    d = np.asarray(data, dtype=np.float32)
    rng = np.random.default_rng(seed)
    up = rng.random(d.shape) < 0.5
    to = np.where(up, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
    p = d
    for _ in range(ulps):
        p = np.nextafter(p, to).astype(np.float32)
    return np.where(d == 0, d, p).astype(np.float32)


def perturbed_workload(w):
    """A synth.Workload with its data passed through perturb_ulp."""
    from gaussmle_amd import synth
    return synth.Workload(w.name + "_perturbed", w.psf_pixels, w.box, np.asfortranarray(perturb_ulp(w.data)),
                          w.truth, w.constraints, w.iterations, w.var_map, w.x_corners, w.y_corners)


def compare(model_id: int, gpu: dict, ref32: dict, ref64: dict, precision: str, ref64p: dict | None = None,
            refmix: dict | None = None, llr_scale: np.ndarray | None = None,
            ref32fp: dict | None = None, ref32g: dict | None = None) -> dict:
    """gpu/ref*: dict(params (n,P), crlb (n,P), llr (n,)); ref64p: the Float64
    oracle on perturb_ulp(data), or None; refmix: the oracle's mixed mode
    (Float32 iterations, Float64 final pass), or None; llr_scale: per-ROI
    rounding scale of the Float32 LLR (llr_rounding_scale), or None; ref32fp:
    the oracle's Float32 final pass at ref32's estimates on perturb_ulp(data)
    (dict with crlb), or None; ref32g: the oracle's Float32 final pass at the
    GPU's estimates (dict with crlb), or None.  Returns a report."""
    e = _errors(model_id, gpu, ref32, ref64, precision, llr_scale)
    # the reference's own Float32 path against the Float64 oracle (its LLR is
    # compared like the GPU's in the same precision mode)
    eref = _errors(model_id, ref32, ref32, ref64, "mixed" if precision == "mixed" else "fp32")
    if precision == "mixed":
        # The mixed-mode LLR is the Float64 LLR at the Float32 iterations'
        # estimate: the oracle's own mixed mode decides whether it can meet
        # the tolerance against the Float64-everywhere LLR.
        if refmix is not None:
            eref["llr"] = np.abs(np.asarray(refmix["llr"], np.float64) - ref64["llr"]) / np.maximum(
                np.abs(ref64["llr"]), 10.0)
            # likewise the CRLB: a Float64 inversion at the Float32 estimate,
            # so the oracle's mixed mode decides (a Float32 inversion of a
            # badly scaled Fisher matrix, e.g. a background near 0.01, would
            # exclude ROIs the mixed mode fits well)
            cm = _rel(np.asarray(refmix["crlb"], np.float64), ref64["crlb"])
            cm[np.isinf(refmix["crlb"]) & np.isinf(ref64["crlb"])] = 0.0
            for k in range(cm.shape[1]):
                eref[f"crlb{k}"] = cm[:, k]
        else:
            eref["llr"] = np.zeros_like(eref["llr"])
    well = ~_bad(eref)
    if ref64p is not None:
        # positions of the perturbed Float64 fit against the Float64 oracle
        well &= ~_bad(_errors(model_id, ref64p, ref64, ref64, "mixed"))
    for extra in (ref32fp, ref32g):
        if extra is not None and precision != "mixed":
            cp = _rel(np.asarray(extra["crlb"], np.float64), ref64["crlb"])
            cp[np.isinf(extra["crlb"]) & np.isinf(ref64["crlb"])] = 0.0
            well &= np.all(cp <= REL_TOL, axis=1)
    bad_gpu = _bad(e)
    fin32 = np.all(np.isfinite(ref32["params"]), axis=1)
    fin_gpu = np.all(np.isfinite(np.asarray(gpu["params"])), axis=1)
    rep = dict(e)
    rep["n"] = len(well)
    rep["well"] = well
    rep["n_ill"] = int((~well).sum())
    rep["bad"] = bad_gpu & well
    rep["n_bad"] = int(rep["bad"].sum())
    rep["finite_mismatch"] = int((fin32 & ~fin_gpu).sum())
    rep["ill_pos_max"] = float(e["pos"][~well].max()) if (~well).any() else 0.0
    return rep


def summary(rep: dict) -> str:
    parts = [f"n={rep['n']} ill_posed={rep['n_ill']} outliers={rep['n_bad']} finite_mismatch={rep['finite_mismatch']}"]
    for k, v in rep.items():
        if isinstance(v, np.ndarray) and v.dtype != bool:
            fin = v[np.isfinite(v)]
            if fin.size:
                parts.append(f"{k}:max={fin.max():.2e},p99.9={np.quantile(fin, 0.999):.2e}")
    parts.append(f"ill_pos_max={rep['ill_pos_max']:.2e}")
    return " ".join(parts)


PARAM_NAMES = {0: ["x", "y", "N", "bg"], 1: ["x", "y", "N", "bg", "sigma"],
               2: ["x", "y", "N", "bg", "sigma_x", "sigma_y"], 3: ["x", "y", "z", "N", "bg"]}


def distributions(model_id: int, rep: dict) -> dict:
    """Per output (position, each parameter, each CRLB, LLR): max, p99.9 and
    the count over tolerance, on the well-posed ROIs and on all ROIs."""
    names = PARAM_NAMES[model_id]
    out = {}
    for k, v in rep.items():
        if not (isinstance(v, np.ndarray) and v.dtype != bool):
            continue
        label = k
        if k.startswith("crlb"):
            label = "CRLB " + names[int(k[4:])]
        elif k.startswith("p") and k[1:].isdigit():
            label = names[int(k[1:])]
        elif k == "pos":
            label = "x, y (px, vs Float32)"
        tol = POS_TOL if k == "pos" else (LLR_TOL if k == "llr" else REL_TOL)
        fin = v[np.isfinite(v)]
        out[label] = {"max": float(fin.max()) if fin.size else float("nan"),
                      "p99.9": float(np.quantile(fin, 0.999)) if fin.size else float("nan"),
                      "over_tol": int((~(v <= tol)).sum()), "tol": tol}
    return out


def assert_parity(rep: dict, label: str = ""):
    allowed = int(OUTLIER_FRACTION * (rep["n"] - rep["n_ill"]))
    assert rep["n_bad"] <= allowed, f"{label}: {summary(rep)}"
    assert rep["finite_mismatch"] == 0, f"{label}: {summary(rep)}"
