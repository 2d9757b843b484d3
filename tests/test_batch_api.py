"""Mixed-shape batch entry points without a GPU: shape table, workspace size,
argument validation (every error is raised before any device work), and the
synthetic mixed-batch generator's layout."""
import numpy as np
import pytest

from gaussmle_amd import _lib, batch as gb, synth


def test_shape_index_matches_python_table(hiplib):
    for m in range(-1, 5):
        for b in range(0, 19):
            assert hiplib.gmle_shape_index(m, b) == gb.shape_index(m, b)
    assert gb.shape_index(3, 16) == _lib.NUM_SHAPES - 1
    assert all(gb.shape_of(gb.shape_index(m, b)) == (m, b) for m in range(4) for b in range(3, 17))


def test_workspace_bytes(hiplib):
    for n in (1, 1000, 4_000_000):
        ws = hiplib.gmle_batch_workspace_bytes(n)
        assert ws >= 4 * n + 2 * 4 * _lib.NUM_SHAPES
        assert ws % 256 == 0


def _args(mb, lanes=None, streams=2, psf=None, iterations=None):
    lo, up, ms = mb.constraint_tables()
    psf = mb.psf_table() if psf is None else psf
    n = mb.n
    out = [np.zeros(6 * n, np.float32), np.zeros(6 * n, np.float32), np.zeros(n, np.float32)]
    cam = mb.camera
    vflat = None if not cam else np.asfortranarray(mb.var_map).reshape(-1, order="F")
    ny, nx = mb.var_map.shape if cam else (0, 0)
    keep = [mb, lo, up, ms, psf, out, vflat, lanes]
    args = (n, _lib.i8ptr(mb.models), _lib.i8ptr(mb.boxes), _lib.fptr(mb.data), _lib.fptr(psf.reshape(-1)),
            _lib.fptr(lo.reshape(-1)), _lib.fptr(up.reshape(-1)), _lib.fptr(ms.reshape(-1)), cam, _lib.fptr(vflat),
            ny, nx, _lib.iptr(mb.x_corners if cam else None), _lib.iptr(mb.y_corners if cam else None),
            mb.iterations if iterations is None else iterations, 0, _lib.iptr(lanes), streams,
            *map(_lib.fptr, out), None, None)
    return args, keep


def test_fit_batch_rejects_bad_input(hiplib):
    mb = gb.mixed_workload(40, seed=5, scmos=True, cam=32)
    call = lambda b, **kw: hiplib.gmle_fit_batch(*_args(b, **kw)[0])  # noqa: E731
    bad = gb.MixedBatch(mb.models.copy(), mb.boxes.copy(), mb.data.copy(), mb.offsets, mb.psfs, mb.constraints,
                        mb.iterations, mb.var_map, mb.x_corners.copy(), mb.y_corners.copy())
    bad.models[5] = 9
    assert call(bad) == _lib.GMLE_ERR_ARG and "unknown PSF model" in _lib.last_error()
    bad.models[5] = mb.models[5]
    bad.boxes[7] = 2
    assert call(bad) == _lib.GMLE_ERR_ARG
    bad.boxes[7] = 17
    assert call(bad) == _lib.GMLE_ERR_UNSUPPORTED
    bad.boxes[7] = mb.boxes[7]
    bad.data[11] = np.nan
    assert call(bad) == _lib.GMLE_ERR_DATA
    bad.data[11] = mb.data[11]
    bad.x_corners[3] = 32
    assert call(bad) == _lib.GMLE_ERR_BOUNDS and "ROI 4" in _lib.last_error()
    bad.x_corners[3] = mb.x_corners[3]
    assert call(bad, streams=-1) == _lib.GMLE_ERR_ARG
    assert call(bad, streams=5) == _lib.GMLE_ERR_ARG
    assert call(bad, iterations=-1) == _lib.GMLE_ERR_ARG
    lanes = np.zeros(_lib.NUM_SHAPES, np.int32)
    lanes[gb.shape_index(int(mb.models[0]), int(mb.boxes[0]))] = 3
    assert call(bad, lanes=lanes) == _lib.GMLE_ERR_UNSUPPORTED
    psf = mb.psf_table()
    psf[0, 0] = 0.0                                  # XYNB sigma must be > 0 when XYNB ROIs are present
    assert call(bad, psf=psf) == _lib.GMLE_ERR_ARG and "sigma" in _lib.last_error()
    if _lib.device_count() == 0:
        assert call(bad) == _lib.GMLE_ERR_NODEVICE


def test_fit_batch_device_rejects_bad_pointers(hiplib):
    n = 10
    z = np.zeros(4 * 8, np.float32)
    c = np.zeros(4 * 6, np.float32)
    args = [n, None, None, None, None, _lib.fptr(z), _lib.fptr(c), _lib.fptr(c), _lib.fptr(c), 0, None, 0, 0, None,
            None, 20, 0, None, 2, None, None, None, None, None, None, None, 0, None]
    assert hiplib.gmle_fit_batch_device(*args) == _lib.GMLE_ERR_ARG
    fake = 4096   # never dereferenced: validation fails first
    args[1:5] = [fake] * 4
    args[20:23] = [fake] * 3
    args[25] = fake
    args[26] = 16                                     # workspace too small
    assert hiplib.gmle_fit_batch_device(*args) == _lib.GMLE_ERR_ARG
    assert "workspace" in _lib.last_error()
    args[0] = 0
    assert hiplib.gmle_fit_batch_device(*args) == _lib.GMLE_ERR_ARG


def test_mixed_workload_layout():
    mb = gb.mixed_workload(3000, seed=9, boxes=(7, 11), box_weights=(0.5, 0.5), models=(0, 2), scmos=True, cam=64)
    sizes = mb.boxes.astype(np.int64) ** 2
    assert mb.data.size == sizes.sum()
    # ROI segments tile the buffer without overlap (grouped by shape, not in ROI order)
    o = np.argsort(mb.offsets)
    np.testing.assert_array_equal(mb.offsets[o], np.concatenate([[0], np.cumsum(sizes[o])[:-1]]))
    counts = mb.shape_counts().reshape(4, 14)
    assert counts.sum() == 3000 and counts[1].sum() == 0 and counts[3].sum() == 0
    assert set(np.flatnonzero(counts.sum(axis=0)) + 3) == {7, 11}
    # every ROI's pixels follow its own expected image (Pearson r over a shape)
    for m, b in ((0, 7), (2, 11)):
        idx = mb.roi_indices(m, b)
        P = 4 if m == 0 else 6
        mu = synth.expected_counts(mb.psfs[m], mb.truth[:P, idx], b)
        st = mb.roi_stack(idx, b)
        assert st.shape == mu.shape
        assert np.corrcoef(st.ravel(), mu.ravel())[0, 1] > 0.95
        assert np.all(np.isnan(mb.truth[P:, idx]))
    # corners keep each ROI inside the sensor
    assert np.all(mb.x_corners + mb.boxes - 1 <= 64) and np.all(mb.x_corners >= 1)
    sub = mb.subset(np.array([5, 1, 2999]))
    np.testing.assert_array_equal(sub.offsets, np.concatenate([[0], np.cumsum(sub.boxes.astype(int) ** 2)[:-1]]))
    np.testing.assert_array_equal(sub.boxes, mb.boxes[[5, 1, 2999]])
    b1 = int(mb.boxes[1])
    np.testing.assert_array_equal(sub.data[sub.offsets[1]:sub.offsets[1] + b1 * b1],
                                  mb.data[mb.offsets[1]:mb.offsets[1] + b1 * b1])


def test_service_entry_points_without_device(hiplib):
    """Argument checks of the service entry points that need no GPU."""
    import ctypes as C
    z = np.zeros(32, np.float32)
    c = np.zeros(24, np.float32)
    lanes = np.zeros(_lib.NUM_SHAPES, np.int32)
    h = C.c_void_p()
    f = _lib.fptr
    # no shape served
    assert hiplib.gmle_service_create(0, 1 << 30, 2, 0, 20, f(z), f(c), f(c), f(c), _lib.iptr(lanes),
                                      C.byref(h)) == _lib.GMLE_ERR_ARG
    lanes[gb.shape_index(0, 11)] = 3                 # no 3-lane kernel
    assert hiplib.gmle_service_create(0, 1 << 30, 2, 0, 20, f(z), f(c), f(c), f(c), _lib.iptr(lanes),
                                      C.byref(h)) == _lib.GMLE_ERR_UNSUPPORTED
    lanes[gb.shape_index(0, 11)] = 2
    assert hiplib.gmle_service_create(0, 1 << 30, 0, 0, 20, f(z), f(c), f(c), f(c), _lib.iptr(lanes),
                                      C.byref(h)) == _lib.GMLE_ERR_ARG        # slots
    assert hiplib.gmle_service_create(0, 1 << 30, 2, 0, 20, f(z), f(c), f(c), f(c), _lib.iptr(lanes),
                                      C.byref(h)) == _lib.GMLE_ERR_ARG        # XYNB sigma 0
    z[0] = 1.3
    if _lib.device_count() == 0:
        assert hiplib.gmle_service_create(0, 1 << 30, 2, 0, 20, f(z), f(c), f(c), f(c), _lib.iptr(lanes),
                                          C.byref(h)) == _lib.GMLE_ERR_NODEVICE
    assert hiplib.gmle_service_destroy(None) == 0
    assert hiplib.gmle_host_free(None) == 0


def test_frames_generator_layout():
    from gaussmle_amd import service as sv
    maps = sv.sensor_maps(seed=3, ny=128, nx=160)
    frs = sv.make_frames(6, seed=4, maps=maps, rois=(50, 400))
    for f in frs:
        assert 50 <= f.n <= 400
        assert f.adu.dtype == np.uint16 and f.adu.size == int((f.boxes.astype(int) ** 2).sum())
        assert set(np.unique(f.boxes)) <= {9, 11, 13, 15} and set(np.unique(f.models)) <= {0, 1}
        assert np.all(f.xc >= 1) and np.all(f.xc + f.boxes - 1 <= 160)
        assert np.all(f.yc >= 1) and np.all(f.yc + f.boxes - 1 <= 128)
    # electrons follow the expected image: offset 100 ADU, gain ~0.5
    e = sv.electrons(frs[0], maps)
    assert abs(float(np.median(e))) < 100
    mb = sv.frames_batch(frs, maps, *sv.default_setup()[:2])
    assert mb.n == sum(f.n for f in frs)
    idx = mb.roi_indices(0, 11)
    from gaussmle_amd import synth
    mu = synth.expected_counts(mb.psfs[0], mb.truth[:4, idx], 11)
    assert np.corrcoef(mb.roi_stack(idx, 11).ravel(), mu.ravel())[0, 1] > 0.95


def test_one_launch_lanes(hiplib):
    for m in range(4):
        for cam in (0, 1):
            for b in range(3, 17):
                auto = hiplib.gmle_one_launch_lanes(m, cam, b, 0)
                if m == 3 and cam == 1 and b >= 14:
                    assert auto == 0          # own launch: no 4-lane tile within 256 VGPRs
                    continue
                assert auto in (1, 2, 4)
                assert hiplib.gmle_one_launch_lanes(m, cam, b, 4) == 4
                for want in (1, 2, 8, 16):
                    got = hiplib.gmle_one_launch_lanes(m, cam, b, want)
                    assert got in (want, 4)
                    if got == want:   # a smaller box never loses a variant a larger one has
                        assert all(hiplib.gmle_one_launch_lanes(m, cam, bb, want) == want for bb in range(3, b))
    assert hiplib.gmle_one_launch_lanes(0, 0, 11, 1) == 1
    assert hiplib.gmle_one_launch_lanes(1, 1, 13, 16) == 16 and hiplib.gmle_one_launch_lanes(1, 1, 13, 8) == 8
    assert hiplib.gmle_one_launch_lanes(1, 0, 16, 8) == 4     # 16 px sigma models: no 8-lane tile within 256
    assert hiplib.gmle_one_launch_lanes(9, 0, 11, 1) == 0

