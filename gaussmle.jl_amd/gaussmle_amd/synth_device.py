"""Seeded synthetic workloads drawn on the GPU, for the large benchmark sizes.

The parameter draws follow synth.py (numpy PCG64, seeded); the pixels come
from the library's own simulator (gmle_simulate_device: Philox4x32-10,
Poisson and Gaussian read noise, the reference's generate_roi_batch
semantics, src/simulator.jl:259-345), so 4,000,000 ROIs take milliseconds
instead of the tens of seconds numpy needs.  sCMOS stacks are drawn in ADU
with per-pixel offset / gain / read-noise maps and converted to electrons as
the reference's to_electrons does, (adu - offset) * gain in Float32, through
each ROI's corner (src/camera_models.jl:110-113).

Results stay in HBM; `host_subset` copies the first ROIs back for the CPU
baseline and oracle checks.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import batch as gb
from .models import GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, PSFModel, SCMOSCamera, default_constraints
from .simulator import simulate_device
from .synth import ASTIG_CONFIG_PSF, Workload, astig_constraints, scmos_maps

F32 = np.float32


@dataclass
class DeviceWorkload:
    """One shape's ROIs resident on the device (torch tensors in `d`: data,
    and for sCMOS var / xc / yc), with the host-side description."""

    name: str
    psf_pixels: PSFModel
    box: int
    n: int
    d: dict
    constraints: object
    iterations: int
    camera: str                       # "ideal" or "scmos"
    var_map: np.ndarray | None = None
    x_corners: np.ndarray | None = None
    y_corners: np.ndarray | None = None
    truth: np.ndarray | None = None   # (P, n)

    def host_subset(self, k: int) -> Workload:
        """The first k ROIs as a host synth.Workload."""
        k = min(k, self.n)
        flat = self.d["data"][: k * self.box * self.box].cpu().numpy()
        data = np.asfortranarray(flat.reshape(k, self.box, self.box).transpose(2, 1, 0))
        return Workload(self.name, self.psf_pixels, self.box, data,
                        None if self.truth is None else self.truth[:, :k], self.constraints, self.iterations,
                        self.var_map, None if self.x_corners is None else self.x_corners[:k].copy(),
                        None if self.y_corners is None else self.y_corners[:k].copy())


def _electrons(adu, box: int, xc, yc, offset, gain, chunk: int = 1 << 18):
    """(adu - offset) * gain in Float32 with the maps read at [yc + i - 1,
    xc + j - 1] (column-major (ny, nx) maps as torch tensors), in place."""
    import torch
    ny = offset.shape[0]
    dev = adu.device
    ii = torch.arange(box, device=dev).repeat(box)                  # row i of pixel p = i + box j
    jj = torch.arange(box, device=dev).repeat_interleave(box)
    off = torch.from_numpy(np.asfortranarray(offset).reshape(-1, order="F")).to(dev)
    gn = torch.from_numpy(np.asfortranarray(gain).reshape(-1, order="F")).to(dev)
    n = xc.numel()
    view = adu.view(n, box * box)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        idx = (yc[a:b, None].long() - 1 + ii[None, :]) + ny * (xc[a:b, None].long() - 1 + jj[None, :])
        view[a:b] = (view[a:b] - off[idx]) * gn[idx]
    return adu


def _draw(model: int, k: int, box: int, rng, sigma_range=(1.0, 1.6)) -> np.ndarray:
    """Parameter draws of synth.py's configurations, (P, k)."""
    c = (box + 1) / 2
    x, y = rng.uniform(c - 1.5, c + 1.5, k), rng.uniform(c - 1.5, c + 1.5, k)
    if model == 3:
        return np.stack([x, y, rng.uniform(-300, 300, k), rng.uniform(1000, 10000, k), rng.uniform(2, 30, k)])
    th = [x, y, rng.uniform(500, 5000, k), rng.uniform(1, 20, k)]
    if model == 1:
        th.append(rng.uniform(*sigma_range, k))
    elif model == 2:
        th += [rng.uniform(*sigma_range, k), rng.uniform(*sigma_range, k)]
    return np.stack(th)


def _simulate(psf, box, theta, seed, camera=None, xc=None, yc=None):
    import torch
    dev = torch.device("cuda:0")
    d_params = torch.from_numpy(np.ascontiguousarray(theta.T.astype(F32))).to(dev)
    return simulate_device(psf, box, d_params, seed, camera, xc, yc)


def config1_device(n: int = 1_000_000, seed: int = 1, box: int = 13, cam: int = 256,
                   iterations: int = 20) -> DeviceWorkload:
    """13x13 GaussianXYNBS on a 256x256 sCMOS sensor (synth.config1's draws and
    maps), pixels simulated on the device."""
    import torch
    rng = np.random.default_rng(seed)
    psf = GaussianXYNBS()
    theta = _draw(1, n, box, rng)
    offset, gain, var = scmos_maps(rng, cam, cam)
    xc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    yc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    dev = torch.device("cuda:0")
    d_xc, d_yc = torch.from_numpy(xc).to(dev), torch.from_numpy(yc).to(dev)
    camera = SCMOSCamera.square(cam, cam, 0.1, offset=offset, gain=gain, readnoise=np.sqrt(var), qe=1.0)
    adu = _simulate(psf, box, theta, seed, camera, d_xc, d_yc)
    data = _electrons(adu, box, d_xc, d_yc, offset, gain)
    d = {"data": data, "var": torch.from_numpy(np.asfortranarray(var).reshape(-1, order="F")).to(dev),
         "xc": d_xc, "yc": d_yc}
    return DeviceWorkload("xynbs_scmos_13", psf, box, n, d, default_constraints(psf, 11), iterations, "scmos", var,
                          xc, yc, theta)


def config2_device(n: int = 4_000_000, seed: int = 2, box: int = 15, iterations: int = 30) -> DeviceWorkload:
    """15x15 AstigmaticXYZNB, Poisson (synth.config2's draws), pixels simulated
    on the device."""
    rng = np.random.default_rng(seed)
    theta = _draw(3, n, box, rng)
    data = _simulate(ASTIG_CONFIG_PSF, box, theta, seed)
    return DeviceWorkload("astig_ideal_15", ASTIG_CONFIG_PSF, box, n, {"data": data}, astig_constraints(), iterations,
                          "ideal", truth=theta)


def large_box_device(n: int = 1_000_000, seed: int = 5, box: int = 21, sigma_range=(1.5, 2.5),
                     iterations: int = 20) -> DeviceWorkload:
    """Large ROIs: GaussianXYNBS of box x box px (sigma uniform in sigma_range,
    other draws as synth.py's), Poisson, with the reference's default
    constraints for the ROI's own size; pixels simulated on the device."""
    rng = np.random.default_rng(seed)
    psf = GaussianXYNBS()
    theta = _draw(1, n, box, rng, sigma_range)
    data = _simulate(psf, box, theta, seed)
    return DeviceWorkload(f"xynbs_ideal_{box}", psf, box, n, {"data": data}, default_constraints(psf, box),
                          iterations, "ideal", truth=theta)


def config3_device(n: int = 4_000_000, seed: int = 3, boxes=(7, 9, 11, 13, 15), box_weights=(0.1, 0.2, 0.4, 0.2, 0.1),
                   models=(0, 1, 2), cam: int = 256, sigma_range=(1.0, 1.8), iterations: int = 20):
    """The uneven batch of batch.config3 (box sizes and models as there, sCMOS
    256x256), pixels simulated on the device shape by shape into one buffer
    with ROIs back to back in ROI order.  Returns (MixedBatch with host
    metadata and no host pixels, dict of device tensors as batch.upload_batch
    returns)."""
    import torch
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    bsel = rng.choice(len(boxes), size=n, p=np.asarray(box_weights) / np.sum(box_weights))
    msel = rng.integers(0, len(models), n)
    box_arr = np.asarray(boxes, dtype=np.int8)[bsel]
    model_arr = np.asarray(models, dtype=np.int8)[msel]
    sizes = box_arr.astype(np.int64) ** 2
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    offset, gain, var = scmos_maps(rng, cam, cam)
    xc = np.zeros(n, np.int32)
    yc = np.zeros(n, np.int32)
    psfs, cons = {}, {}
    data = torch.empty(int(sizes.sum()), dtype=torch.float32, device=dev)
    d_off = torch.from_numpy(offsets).to(dev)
    camera = SCMOSCamera.square(cam, cam, 0.1, offset=offset, gain=gain, readnoise=np.sqrt(var), qe=1.0)
    truth = np.full((gb.STRIDE, n), np.nan)
    for m in models:
        psfs[m] = GaussianXYNB(1.3) if m == 0 else GaussianXYNBS() if m == 1 else GaussianXYNBSXSY()
        cons[m] = default_constraints(psfs[m], 11)
        for b in boxes:
            idx = np.flatnonzero((model_arr == m) & (box_arr == b))
            k = len(idx)
            if not k:
                continue
            th = _draw(m, k, b, rng, sigma_range)
            truth[:th.shape[0], idx] = th
            xs = rng.integers(1, cam - b + 2, k).astype(np.int32)
            ys = rng.integers(1, cam - b + 2, k).astype(np.int32)
            xc[idx], yc[idx] = xs, ys
            d_xs, d_ys = torch.from_numpy(xs).to(dev), torch.from_numpy(ys).to(dev)
            adu = _simulate(psfs[m], b, th, seed * 1000 + m * 100 + b, camera, d_xs, d_ys)
            e = _electrons(adu, b, d_xs, d_ys, offset, gain).view(k, b * b)
            dst = d_off[torch.from_numpy(idx).to(dev)][:, None] + torch.arange(b * b, device=dev)[None, :]
            data[dst.reshape(-1)] = e.reshape(-1)
            del adu, e, dst
    mb = gb.MixedBatch(model_arr, box_arr, np.empty(0, F32), offsets, psfs, cons, iterations, var, xc, yc, truth,
                       f"mixed_scmos_{n}")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    from . import _lib
    d = {"models": t(model_arr), "boxes": t(box_arr), "offsets": d_off, "data": data, "counts": mb.shape_counts(),
         "var": t(np.asfortranarray(var).reshape(-1, order="F")), "xc": t(xc), "yc": t(yc)}
    ws = int(_lib.load().gmle_batch_workspace_bytes(n))
    d["workspace"] = torch.empty(ws, dtype=torch.uint8, device=dev)
    return mb, d


def host_batch(mb, d: dict, k: int):
    """The first k ROIs of a device-generated mixed batch with their pixels on
    the host (for the CPU baseline and oracle checks)."""
    k = min(k, mb.n)
    end = int(mb.offsets[k - 1] + int(mb.boxes[k - 1]) ** 2)
    data = d["data"][:end].cpu().numpy()
    return gb.MixedBatch(mb.models[:k].copy(), mb.boxes[:k].copy(), data, mb.offsets[:k].copy(), mb.psfs,
                         mb.constraints, mb.iterations, mb.var_map, mb.x_corners[:k].copy(), mb.y_corners[:k].copy(),
                         None if mb.truth is None else mb.truth[:, :k].copy(), mb.name)
