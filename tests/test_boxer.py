"""ROI extraction (gmle_find_boxes / gmle_find_boxes_device, gaussmle_amd.boxer).

GaussMLE.jl takes its ROIBatch from SMLMBoxer.getboxes
(examples/full_pipeline_boxer.jl:70-80, examples/full_pipeline.jl:73-81).
SMLMBoxer is not vendored and is listed without a version
(examples/Project.toml:9), so the steps are restated with explicit conventions
(oracle/refboxer.cpp) and parity against SMLMBoxer itself is unpinned.  The
oracle is pinned here against what fixes each step independently:
  * the filter against scipy.ndimage.gaussian_filter (Float64, mode
    'nearest', truncate 3: the same weights, radius and edge rule);
  * detection against scipy.ndimage.maximum_filter on the oracle's own
    filtered stack (no ties in continuous data), and ties against a
    brute-force restatement of the rule on exact mirror-image ties;
  * boxes against numpy slicing, corners against the clamp rule, order
    against a sort.
The GPU tests then require the device path to equal the oracle bit for bit:
counts, corners, frames, the filtered value at each detection, and pixels.
"""
import numpy as np
import pytest
from scipy import ndimage

from gaussmle_amd import _lib
from gaussmle_amd import boxer as gbx
from oracle import refboxer

F32 = np.float32


def spots_stack(ny, nx, nf, n_spots, seed, bg=10.0, noise=2.0, photons=(500, 3000), sigma=1.3, edge=3):
    """Gaussian spots (integrated over pixels) on a flat background with
    Gaussian read noise, Float32, (ny, nx, nf) column-major."""
    from scipy.special import erf
    rng = np.random.default_rng(seed)
    out = np.empty((ny, nx, nf), np.float32, order="F")
    rows = np.arange(ny + 1) - 0.5
    cols = np.arange(nx + 1) - 0.5
    truth = []
    for f in range(nf):
        img = np.full((ny, nx), bg, np.float64)
        for _ in range(n_spots):
            y = rng.uniform(edge, ny - 1 - edge)
            x = rng.uniform(edge, nx - 1 - edge)
            N = rng.uniform(*photons)
            py = np.diff(0.5 * erf((rows - y) / (np.sqrt(2) * sigma)))
            px = np.diff(0.5 * erf((cols - x) / (np.sqrt(2) * sigma)))
            img += N * py[:, None] * px[None, :]
            truth.append((f, y, x))
        img += rng.normal(0, noise, img.shape)
        out[:, :, f] = img
    return out, truth


def scipy_dog(frames, ss, sl):
    a = frames.astype(np.float64)
    out = np.empty_like(a)
    for f in range(a.shape[2]):
        out[:, :, f] = (ndimage.gaussian_filter(a[:, :, f], ss, mode="nearest", truncate=3.0)
                        - ndimage.gaussian_filter(a[:, :, f], sl, mode="nearest", truncate=3.0))
    return out


# --- the oracle, pinned -----------------------------------------------------

@pytest.mark.parametrize("sigma", [0.7, 1.0, 1.1, 1.3, 2.0, 2.5, 4.0])
def test_oracle_weights_are_normalised_gaussians(sigma):
    w = refboxer.weights(sigma)
    R = int(np.floor(3 * sigma + 0.5))
    assert len(w) == R + 1
    k = np.arange(R + 1)
    s = float(np.float32(sigma))      # the ABI takes Float32 sigmas
    g = np.exp(-k ** 2 / (2 * s ** 2))
    want = g / (g[0] + 2 * g[1:].sum())
    np.testing.assert_allclose(w, want, rtol=1e-7)


@pytest.mark.parametrize("ss,sl", [(1.0, 2.0), (1.1, 1.6), (2.5, 1.0), (0.8, 4.0)])
def test_oracle_filter_matches_scipy(ss, sl):
    frames, _ = spots_stack(70, 45, 2, 8, seed=1)
    got = refboxer.dog(frames, ss, sl)
    want = scipy_dog(frames, ss, sl)
    scale = float(np.abs(frames).max())
    assert np.abs(got - want).max() <= 4e-6 * scale


def test_oracle_filter_impulse_response_is_the_kernel_difference():
    """D of a single bright pixel is w_s w_s^T - w_l w_l^T around it."""
    f = np.zeros((41, 41, 1), np.float32)
    f[20, 20, 0] = 1.0
    ws, wl = refboxer.weights(1.0), refboxer.weights(2.0)
    full = lambda w: np.concatenate([w[:0:-1], w]).astype(np.float64)  # noqa: E731
    want = np.zeros((41, 41))
    ks, kl = full(ws), full(wl)
    rs, rl = len(ws) - 1, len(wl) - 1
    want[20 - rs:21 + rs, 20 - rs:21 + rs] += np.outer(ks, ks)
    want[20 - rl:21 + rl, 20 - rl:21 + rl] -= np.outer(kl, kl)
    np.testing.assert_allclose(refboxer.dog(f, 1.0, 2.0)[:, :, 0], want, atol=2e-8)


@pytest.mark.parametrize("box,overlap,minval", [(11, 2.0, 5.0), (7, 0.0, 2.0), (9, 3.5, -np.inf)])
def test_oracle_detections_match_maximum_filter(box, overlap, minval):
    frames, _ = spots_stack(96, 80, 3, 15, seed=2)
    r = refboxer.find(frames, box, overlap, 1.0, 2.0, minval)
    D = r["dog"]
    win = gbx.suppression_radius(box, overlap)
    det = []
    for f in range(D.shape[2]):
        mx = ndimage.maximum_filter(D[:, :, f], size=2 * win + 1, mode="constant", cval=-np.inf)
        hit = (D[:, :, f] == mx) & (D[:, :, f] > minval)
        cols, rows = np.nonzero(hit.T)          # column-major order: by column, then row
        det += [(f, c, rw) for c, rw in zip(cols, rows)]
    assert r["n"] == len(det) > 0
    half = (box - 1) // 2
    ny, nx = D.shape[:2]
    for k, (f, c, rw) in enumerate(det):
        assert r["frames"][k] == f + 1
        assert r["x_corners"][k] == min(max(c - half, 0), nx - box) + 1
        assert r["y_corners"][k] == min(max(rw - half, 0), ny - box) + 1
        assert r["strength"][k] == D[rw, c, f]
        r0, c0 = r["y_corners"][k] - 1, r["x_corners"][k] - 1
        np.testing.assert_array_equal(r["rois"][:, :, k], frames[r0:r0 + box, c0:c0 + box, f])


def brute_force_detections(D, win, minval):
    """The detection rule written out: q beats p if D(q) > D(p), or equal and
    q earlier in column-major order."""
    ny, nx = D.shape
    out = []
    for c in range(nx):
        for r in range(ny):
            v = D[r, c]
            if not v > minval:
                continue
            ok = True
            for qc in range(max(0, c - win), min(nx, c + win + 1)):
                for qr in range(max(0, r - win), min(ny, r + win + 1)):
                    q = D[qr, qc]
                    if q > v or (q == v and (qc, qr) < (c, r)):
                        ok = False
            if ok:
                out.append((c, r))
    return out


def test_oracle_breaks_exact_ties_by_column_major_order():
    f = np.full((24, 30, 1), 5.0, np.float32, order="F")
    f[10:12, 8, 0] = 200.0        # two equal pixels in one column: mirror images
    f[15, 20:22, 0] = 150.0       # two equal pixels in one row
    r = refboxer.find(f, 5, 1.0, 1.0, 2.0, 1.0)
    D = r["dog"][:, :, 0]
    assert D[10, 8] == D[11, 8] and D[15, 20] == D[15, 21]
    want = brute_force_detections(D, gbx.suppression_radius(5, 1.0), 1.0)
    assert [(c, rw) for c, rw in want] == [(8, 10), (20, 15)]
    assert r["n"] == 2
    np.testing.assert_array_equal(r["x_corners"], [8 - 2 + 1, 20 - 2 + 1])
    np.testing.assert_array_equal(r["y_corners"], [10 - 2 + 1, 15 - 2 + 1])


def test_oracle_finds_isolated_spots():
    frames, truth = spots_stack(128, 128, 2, 6, seed=3, edge=12, noise=0.5, photons=(2000, 3000))
    r = refboxer.find(frames, 11, 2.0, 1.0, 2.0, 8.0)
    got = sorted(zip(r["frames"] - 1, r["y_corners"] + 4, r["x_corners"] + 4))
    # every true spot has a detection within a pixel of it (spots can merge)
    for f, y, x in truth:
        d = [max(abs(gy - y), abs(gx - x)) for gf, gy, gx in got if gf == f]
        assert min(d) <= 1.0 + 1e-9 or any(
            max(abs(y - y2), abs(x - x2)) < 9 for f2, y2, x2 in truth if f2 == f and (y2, x2) != (y, x))


# --- host-side checks of the library (no device needed) ---------------------

def test_library_weights_equal_oracle(hiplib):
    import ctypes as C
    for sigma in (0.7, 1.0, 1.3, 2.0, 4.0):
        w = np.zeros(16, np.float32)
        R = hiplib.gmle_boxer_weights(sigma, w.ctypes.data_as(C.POINTER(C.c_float)), 16)
        np.testing.assert_array_equal(w[:R + 1], refboxer.weights(sigma))
    assert hiplib.gmle_boxer_weights(4.2, w.ctypes.data_as(C.POINTER(C.c_float)), 16) == -1
    assert hiplib.gmle_boxer_weights(0.0, w.ctypes.data_as(C.POINTER(C.c_float)), 16) == -1


@pytest.mark.parametrize("kw,msg", [
    ({"boxsize": 2}, "boxsize"), ({"boxsize": 17}, "boxsize"), ({"overlap": 7.0}, "overlap"),
    ({"overlap": -1.0}, "overlap"), ({"sigma_small": 0.0}, "sigma"), ({"sigma_large": 4.5}, "4 px"),
    ({"minval": float("nan")}, "NaN"),
])
def test_find_boxes_validates_before_touching_the_device(kw, msg):
    args = dict(boxsize=7, overlap=2.0, sigma_small=1.0, sigma_large=2.0, minval=0.0)
    args.update(kw)
    with pytest.raises(_lib.ArgumentError, match=msg):
        gbx.find_boxes(np.zeros((32, 32, 1), np.float32), **args)
    with pytest.raises(_lib.ArgumentError, match="smaller than the box"):
        gbx.find_boxes(np.zeros((6, 32, 1), np.float32), boxsize=7)


def test_workspace_and_bounds():
    lib = _lib.load()
    ny, nx, nf = 100, 77, 3
    wpc = -(-ny // 64)
    ws = lib.gmle_boxer_workspace_bytes(ny, nx, nf)
    assert ws >= ny * nx * nf * 4 + 2 * nf * nx * wpc * 8
    assert lib.gmle_boxer_workspace_bytes(0, 1, 1) == 0
    assert gbx.suppression_radius(11, 2.0) == 8 and gbx.suppression_radius(7, 6.5) == 1
    assert gbx.max_detections(18, 18, 2, 11, 2.0) == 2 * 2 * 2


# --- the device path ----------------------------------------------------------

def _assert_same(got, want, label):
    assert got["n"] == want["n"], (label, got["n"], want["n"])
    for k in ("x_corners", "y_corners", "frames", "strength", "rois"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{label}: {k}")


CASES = {
    "256x256x6 box11": (lambda: spots_stack(256, 256, 6, 40, seed=10)[0], dict(boxsize=11, overlap=2.0, minval=5.0)),
    "ragged 100x77x3 box7": (lambda: spots_stack(100, 77, 3, 12, seed=11)[0], dict(boxsize=7, overlap=1.0,
                                                                                  minval=3.0)),
    "box16 no overlap": (lambda: spots_stack(130, 200, 2, 20, seed=12)[0], dict(boxsize=16, overlap=0.0,
                                                                               minval=4.0)),
    "small sigma larger radius": (lambda: spots_stack(90, 90, 2, 10, seed=13)[0],
                                  dict(boxsize=9, overlap=2.0, sigma_small=2.5, sigma_large=1.0, minval=-1.0)),
    "every maximum (minval -inf)": (lambda: spots_stack(64, 64, 2, 0, seed=14)[0],
                                    dict(boxsize=5, overlap=2.0, minval=-np.inf)),
    "exact ties": (lambda: np.asfortranarray(np.pad(np.full((24, 30, 1), 5.0, np.float32), 0)), None),
    "one frame tall": (lambda: spots_stack(300, 40, 1, 6, seed=15)[0], dict(boxsize=13, overlap=3.0, minval=5.0)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_find_boxes_matches_oracle_bit_for_bit(name):
    make, kw = CASES[name]
    frames = make()
    if kw is None:   # the tie case
        frames[10:12, 8, 0] = 200.0
        frames[15, 20:22, 0] = 150.0
        kw = dict(boxsize=5, overlap=1.0, minval=1.0)
    args = dict(sigma_small=1.0, sigma_large=2.0)
    args.update(kw)
    print(f"[boxer] {name} {frames.shape} {args}", flush=True)
    want = refboxer.find(frames, args["boxsize"], args["overlap"], args["sigma_small"], args["sigma_large"],
                         args["minval"])
    got = gbx.find_boxes(frames, **args)
    _assert_same(got, want, name)
    assert got["n"] > 0


@pytest.mark.gpu
def test_find_boxes_truncates_at_capacity():
    frames, _ = spots_stack(128, 128, 2, 25, seed=20)
    want = refboxer.find(frames, 7, 2.0, 1.0, 2.0, 4.0)
    got = gbx.find_boxes(frames, 7, 2.0, 1.0, 2.0, 4.0, max_boxes=5)
    assert got["n"] == want["n"] > 5
    for k in ("x_corners", "y_corners", "frames", "strength"):
        np.testing.assert_array_equal(got[k], want[k][:5])
    np.testing.assert_array_equal(got["rois"], want["rois"][:, :, :5])


@pytest.mark.gpu
def test_find_boxes_device_large_frames_matches_oracle():
    """2048 x 2048 frames through the device entry point, twice on the same
    buffers (the second call with fewer frames)."""
    import torch
    frames, _ = spots_stack(2048, 2048, 2, 600, seed=21)
    want = refboxer.find(frames, 11, 2.0, 1.0, 2.0, 5.0)
    d = torch.from_numpy(frames.reshape(-1, order="F").copy()).cuda()
    out = gbx.find_boxes_device(d, 2048, 2048, 2, 11, 2.0, 1.0, 2.0, 5.0, max_boxes=want["n"] + 10)
    for nf in (2, 1):
        out = gbx.find_boxes_device(d, 2048, 2048, nf, 11, 2.0, 1.0, 2.0, 5.0, max_boxes=want["n"] + 10, out=out)
        torch.cuda.synchronize()
        n = int(out["count"].item())
        m = want["n"] if nf == 2 else int(np.sum(want["frames"] == 1))
        print(f"[boxer] device 2048x2048x{nf}: {n} detections", flush=True)
        assert n == m
        np.testing.assert_array_equal(out["x_corners"][:n].cpu().numpy(), want["x_corners"][:m])
        np.testing.assert_array_equal(out["y_corners"][:n].cpu().numpy(), want["y_corners"][:m])
        np.testing.assert_array_equal(out["frames"][:n].cpu().numpy(), want["frames"][:m])
        np.testing.assert_array_equal(out["strength"][:n].cpu().numpy(), want["strength"][:m])
        rois = out["rois"][:n].cpu().numpy().reshape(n, 11, 11).transpose(2, 1, 0)
        np.testing.assert_array_equal(rois, want["rois"][:, :, :m])


@pytest.mark.gpu
def test_getboxes_then_fit_equals_fitting_the_oracle_boxes():
    """frames -> boxes -> fit on the device gives the fits of the oracle's boxes."""
    from gaussmle_amd import GaussianXYNB, GaussMLEFitter, IdealCamera, fit
    frames, _ = spots_stack(256, 256, 4, 30, seed=22)
    cam = IdealCamera.square(256, 256, 0.1)
    batch = gbx.getboxes(frames, cam, boxsize=11, overlap=2.0, sigma_small=1.0, sigma_large=2.0, minval=5.0)
    want = refboxer.find(frames, 11, 2.0, 1.0, 2.0, 5.0)
    np.testing.assert_array_equal(batch.data, want["rois"])
    fitter = GaussMLEFitter(GaussianXYNB(0.13))
    smld = fit(fitter, batch)
    assert len(smld.emitters) == want["n"] > 50
    xs = np.array([e.x for e in smld.emitters])
    assert np.all((xs > 0) & (xs < 25.6))
