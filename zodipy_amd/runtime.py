"""Device-side runtime: model block upload and the C-ABI calls.

PyTorch provides device memory and the current HIP stream; all compute happens
in the HIP library. Inputs are float64 CUDA tensors in the layouts of
include/zodi_hip.h:

* rays ``u``: (3, n_rays), rows x, y, z (ecliptic unit vectors);
* observers/Earth: (3, n_seg) per segment, or (3, n_rays) per ray;
* ``seg_offsets``: int64 (n_seg + 1,), ``seg_offsets[0] == 0``,
  ``seg_offsets[-1] == n_rays``, non-decreasing;
* ``beyond``: int32 masks from ``PackedModel.beyond_bits``.
"""

from __future__ import annotations

import numpy as np
import torch

from zodipy_amd import _lib
from zodipy_amd._lib import OBS_PER_RAY, OBS_PER_SEGMENT, OUT_COMPS, OUT_SUM, Launch
from zodipy_amd.packing import PackedModel

DEFAULT_DEVICE = "cuda"


def _require_cuda(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        msg = f"device {dev} is not a GPU device; the HIP kernels need a ROCm GPU"
        raise RuntimeError(msg)
    if not torch.cuda.is_available():
        msg = "no ROCm GPU is available to run the HIP kernels"
        raise RuntimeError(msg)
    return dev


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _f64_rows(t: torch.Tensor, rows: int, name: str, device) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t, dtype=np.float64))
    t = t.to(device=device, dtype=torch.float64)
    if t.ndim == 1 and rows == 3 and t.shape[0] == 3:
        t = t.reshape(3, 1)
    if t.ndim != 2 or t.shape[0] != rows:
        msg = f"{name} must have shape ({rows}, n), got {tuple(t.shape)}"
        raise ValueError(msg)
    return t.contiguous()


_TORCH_DTYPE = {np.float64: torch.float64, np.int64: torch.int64, np.int32: torch.int32}


class DeviceModel:
    """A packed model resident on one GPU."""

    def __init__(self, packed: PackedModel, device=DEFAULT_DEVICE) -> None:
        self.device = _require_cuda(device)
        self.lib = _lib.load()
        self.packed = packed
        self.host_block = np.ascontiguousarray(packed.block)
        self.dev_block = torch.from_numpy(self.host_block.view(np.uint8).copy()).to(self.device)

    @property
    def n_comps(self) -> int:
        return self.packed.n_comps

    @property
    def n_groups(self) -> int:
        return self.packed.n_groups

    def _ptrs(self):
        return _lib.host_ptr(self.host_block), self.dev_block.data_ptr()

    def _device_per_ray(self, obs_xyz: torch.Tensor, earth_xyz: torch.Tensor):
        """Per-ray observers already on the GPU: the cutoff bits come from the largest
        observer distance, reduced on the device (one scalar returns to the host)."""
        obs = obs_xyz.to(device=self.device, dtype=torch.float64).reshape(3, -1).contiguous()
        earth = earth_xyz.to(device=self.device, dtype=torch.float64).reshape(3, -1).contiguous()
        if earth.shape != obs.shape:
            msg = f"earth positions {tuple(earth.shape)} do not match observers {tuple(obs.shape)}"
            raise ValueError(msg)
        # r = sqrt(x^2 + y^2 + z^2) as zodipy/line_of_sight.py:71 forms it; NaN distances
        # compare false, as in (r_obs > cutoff).any(). One fused reduction kernel.
        red = torch.empty(1025, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().zodi_max_radius(obs.data_ptr(), obs.shape[1], red.data_ptr(),
                                               _stream()), "zodi_max_radius")
        r_max = float(red[0])
        beyond = self.packed.beyond_bits(np.array([[r_max], [0.0], [0.0]]), any_over_all=True)
        beyond_t = torch.from_numpy(beyond.view(np.int32).copy()).to(self.device)
        return None, 0, obs, earth, beyond_t, OBS_PER_RAY

    def prepare_observers(self, obs_xyz, earth_xyz, seg_offsets=None, per_ray: bool = False):
        """Upload observer data; returns (seg, n_seg, obs, earth, beyond, mode)."""
        return self._prepare(obs_xyz, earth_xyz, seg_offsets, per_ray)[0]

    def _prepare(self, obs_xyz, earth_xyz, seg_offsets=None, per_ray: bool = False,
                 rows: np.ndarray | None = None):
        """``prepare_observers``, plus host ``rows`` (float64, C order) uploaded in the
        same copy when all of it is small: ((seg, n_seg, obs, earth, beyond, mode),
        rows on the device or None)."""
        if per_ray and isinstance(obs_xyz, torch.Tensor) and obs_xyz.is_cuda and \
                isinstance(earth_xyz, torch.Tensor) and earth_xyz.is_cuda:
            return self._device_per_ray(obs_xyz, earth_xyz), None
        obs_np = np.asarray(obs_xyz.cpu() if isinstance(obs_xyz, torch.Tensor) else obs_xyz,
                            dtype=np.float64).reshape(3, -1)
        earth_np = np.asarray(
            earth_xyz.cpu() if isinstance(earth_xyz, torch.Tensor) else earth_xyz,
            dtype=np.float64).reshape(3, -1)
        if earth_np.shape != obs_np.shape:
            msg = f"earth positions {earth_np.shape} do not match observers {obs_np.shape}"
            raise ValueError(msg)
        if per_ray:
            beyond = self.packed.beyond_bits(obs_np, any_over_all=True)
            seg = None
            n_seg = 0
            mode = OBS_PER_RAY
        else:
            n_seg = obs_np.shape[1]
            if seg_offsets is None:
                msg = "seg_offsets required for per-segment observers"
                raise ValueError(msg)
            seg_np = np.asarray(seg_offsets.cpu() if isinstance(seg_offsets, torch.Tensor)
                                else seg_offsets, dtype=np.int64)
            if seg_np.shape != (n_seg + 1,):
                msg = f"seg_offsets must have shape ({n_seg + 1},), got {seg_np.shape}"
                raise ValueError(msg)
            if seg_np[0] != 0 or np.any(np.diff(seg_np) < 0):
                msg = "seg_offsets must start at 0 and be non-decreasing"
                raise ValueError(msg)
            beyond = self.packed.beyond_bits(obs_np, any_over_all=False)
            mode = OBS_PER_SEGMENT
        parts = [np.ascontiguousarray(obs_np, dtype=np.float64),
                 np.ascontiguousarray(earth_np, dtype=np.float64),
                 np.ascontiguousarray(beyond.view(np.int32))]
        if mode == OBS_PER_SEGMENT:
            parts.insert(0, np.ascontiguousarray(seg_np))
        small = sum(a.nbytes for a in parts) <= 1 << 20
        with_rows = rows is not None and small and sum(a.nbytes for a in parts) + rows.nbytes \
            <= 1 << 20
        if with_rows:
            parts.append(rows)
        if small:
            # One upload for all of them (a small call's fixed cost is mostly its
            # copies): the arrays at 8-byte offsets of one buffer, viewed in place.
            sizes = [a.nbytes for a in parts]
            offs = np.concatenate([[0], np.cumsum([(n + 7) // 8 * 8 for n in sizes])]).astype(int)
            buf = np.zeros(int(offs[-1]), dtype=np.uint8)
            for a, o in zip(parts, offs):
                buf[o:o + a.nbytes] = a.reshape(-1).view(np.uint8)
            dev = torch.from_numpy(buf).to(self.device)
            up = [dev[o:o + a.nbytes].view(_TORCH_DTYPE[a.dtype.type]).reshape(a.shape)
                  for a, o in zip(parts, offs)]
        else:
            # np.array copies: broadcast views (per-ray observers) are read-only
            up = [torch.from_numpy(np.array(a)).to(self.device) for a in parts]
        rows_dev = up.pop() if with_rows else None
        seg = up.pop(0) if mode == OBS_PER_SEGMENT else None
        obs, earth, beyond_t = up
        return (seg, n_seg, obs, earth, beyond_t, mode), rows_dev

    def forward_raw(self, u: torch.Tensor, seg, n_seg: int, obs: torch.Tensor,
                    earth: torch.Tensor, beyond: torch.Tensor, mode: int, out: torch.Tensor,
                    out_mode: int, launch: Launch | None = None) -> torch.Tensor:
        """Launch ``zodi_forward`` on already-resident buffers (no checks beyond the ABI's)."""
        n_rays = u.shape[1]
        if mode == OBS_PER_SEGMENT and seg is not None and seg.numel() != n_seg + 1:
            msg = "seg_offsets length does not match n_seg"
            raise ValueError(msg)
        expect = (self.n_comps, n_rays) if out_mode == OUT_COMPS else (n_rays,)
        if tuple(out.shape) != expect or out.dtype != torch.float64 or not out.is_contiguous():
            msg = f"out must be a contiguous float64 tensor of shape {expect}"
            raise ValueError(msg)
        hb, db = self._ptrs()
        st = self.lib.zodi_forward(
            hb, db, u.data_ptr(), n_rays, seg.data_ptr() if seg is not None else None, n_seg,
            obs.data_ptr(), earth.data_ptr(), mode, beyond.data_ptr(), out.data_ptr(), out_mode,
            launch, _stream(),
        )
        _lib.check(st, "zodi_forward")
        return out

    def forward(self, u, obs_xyz, earth_xyz, seg_offsets=None, per_ray: bool = False,
                out_mode: int = OUT_COMPS, launch: Launch | None = None) -> torch.Tensor:
        """Integrated emission for rays ``u`` (3, n) [MJy/sr].

        Observers: ``per_ray`` -> (3, n) per ray (one reference call, one cutoff
        test over all of them); otherwise (3, n_seg) with ``seg_offsets``;
        a single (3,) observer is one segment.
        """
        u_host = None
        if not isinstance(u, torch.Tensor):
            # host pointings: checked here, uploaded with the observer data when small
            u_host = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
            if u_host.shape == (3,):
                u_host = u_host.reshape(3, 1)
            if u_host.ndim != 2 or u_host.shape[0] != 3:
                msg = f"unit vectors must have shape (3, n), got {u_host.shape}"
                raise ValueError(msg)
            n_rays = u_host.shape[1]
        else:
            u = _f64_rows(u, 3, "unit vectors", self.device)
            n_rays = u.shape[1]
        if per_ray and isinstance(obs_xyz, torch.Tensor) and obs_xyz.is_cuda:
            if u_host is not None:
                u = _f64_rows(u_host, 3, "unit vectors", self.device)
            if obs_xyz.reshape(3, -1).shape[1] != n_rays:
                msg = f"per-ray observers must be (3, {n_rays}), got {tuple(obs_xyz.shape)}"
                raise ValueError(msg)
            seg, n_seg, obs, earth, beyond, mode = self.prepare_observers(
                obs_xyz, earth_xyz, None, True)
            shape = (self.n_comps, n_rays) if out_mode == OUT_COMPS else (n_rays,)
            out = torch.empty(shape, dtype=torch.float64, device=self.device)
            return self.forward_raw(u, seg, n_seg, obs, earth, beyond, mode, out, out_mode,
                                    launch)
        obs_np = np.asarray(obs_xyz.cpu() if isinstance(obs_xyz, torch.Tensor) else obs_xyz,
                            dtype=np.float64).reshape(3, -1)
        if not per_ray and seg_offsets is None:
            if obs_np.shape[1] != 1:
                msg = "several observers need seg_offsets or per_ray=True"
                raise ValueError(msg)
            seg_offsets = np.array([0, n_rays], dtype=np.int64)
        if per_ray and obs_np.shape[1] != n_rays:
            msg = f"per-ray observers must be (3, {n_rays}), got {obs_np.shape}"
            raise ValueError(msg)
        (seg, n_seg, obs, earth, beyond, mode), u_dev = self._prepare(
            obs_np, earth_xyz, seg_offsets, per_ray, rows=u_host)
        if u_host is not None:
            u = u_dev if u_dev is not None else _f64_rows(u_host, 3, "unit vectors", self.device)
        if mode == OBS_PER_SEGMENT and int(seg_offsets[-1]) != n_rays:
            msg = f"seg_offsets[-1] = {int(seg_offsets[-1])} but there are {n_rays} rays"
            raise ValueError(msg)
        shape = (self.n_comps, n_rays) if out_mode == OUT_COMPS else (n_rays,)
        out = torch.empty(shape, dtype=torch.float64, device=self.device)
        return self.forward_raw(u, seg, n_seg, obs, earth, beyond, mode, out, out_mode, launch)

    def line_of_sight(self, u, obs_xyz, per_ray: bool = False):
        """Per-group (start, stop) distances, each (n_groups, n_rays)."""
        u = _f64_rows(u, 3, "unit vectors", self.device)
        n_rays = u.shape[1]
        seg_offsets = None if per_ray else np.array([0, n_rays], dtype=np.int64)
        seg, n_seg, obs, _, beyond, mode = self.prepare_observers(obs_xyz, obs_xyz, seg_offsets,
                                                                  per_ray)
        start = torch.empty((self.n_groups, n_rays), dtype=torch.float64, device=self.device)
        stop = torch.empty_like(start)
        hb, db = self._ptrs()
        st = self.lib.zodi_line_of_sight(
            hb, db, u.data_ptr(), n_rays, seg.data_ptr() if seg is not None else None, n_seg,
            obs.data_ptr(), mode, beyond.data_ptr(), start.data_ptr(), stop.data_ptr(), _stream())
        _lib.check(st, "zodi_line_of_sight")
        return start, stop

    def number_density(self, X, earth_xyz) -> torch.Tensor:
        """Component densities at heliocentric points X (3, m) -> (n_comps, m)."""
        X = _f64_rows(X, 3, "points", self.device)
        earth = torch.as_tensor(np.asarray(earth_xyz, dtype=np.float64).reshape(3),
                                device=self.device)
        out = torch.empty((self.n_comps, X.shape[1]), dtype=torch.float64, device=self.device)
        hb, db = self._ptrs()
        st = self.lib.zodi_number_density(hb, db, X.data_ptr(), X.shape[1], earth.data_ptr(),
                                          out.data_ptr(), _stream())
        _lib.check(st, "zodi_number_density")
        return out


__all__ = ["DeviceModel", "OUT_COMPS", "OUT_SUM", "OBS_PER_RAY", "OBS_PER_SEGMENT", "Launch"]
