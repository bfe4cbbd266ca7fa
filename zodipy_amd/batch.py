"""Batched evaluation of many observers in one graph-captured launch.

A batch is a set of rays (3, n_rays) resident in device memory, partitioned
into contiguous segments, one per observer (``seg_offsets``). Each segment has
its own observer and Earth position and its own cutoff test, i.e. the batch is
equivalent to one reference ``Model.evaluate`` call per observer
(zodipy/model.py:119-279) but runs as a single kernel launch.

``BatchedEvaluator`` uploads the per-observer data once, allocates the output,
and captures the launch into a HIP graph, so ``run()`` is one graph replay on
the current stream with no host work and no allocation.
"""

from __future__ import annotations

import numpy as np
import torch

from zodipy_amd._lib import Launch
from zodipy_amd.runtime import OUT_COMPS, OUT_SUM, DeviceModel


class BatchedEvaluator:
    """Graph-captured forward evaluation of a fixed batch of observers.

    Args:
        device_model: the packed model on the GPU.
        unit_vectors: float64 CUDA tensor (3, n_rays), ecliptic unit vectors.
        seg_offsets: (n_obs + 1,) int64, non-decreasing, from 0 to n_rays.
        obs_xyz, earth_xyz: (3, n_obs) heliocentric ecliptic positions [AU].
        return_comps: per-component output (ncomps, n_rays) instead of the sum.
        launch: launch configuration (``None``: library default).
        use_graph: capture the launch in a HIP graph (default) or launch eagerly.
    """

    def __init__(self, device_model: DeviceModel, unit_vectors: torch.Tensor, seg_offsets,
                 obs_xyz, earth_xyz, *, return_comps: bool = False,
                 launch: Launch | None = None, use_graph: bool = True) -> None:
        self.dm = device_model
        dev = device_model.device
        if not (isinstance(unit_vectors, torch.Tensor) and unit_vectors.is_cuda):
            unit_vectors = torch.as_tensor(np.asarray(unit_vectors, dtype=np.float64), device=dev)
        if unit_vectors.dtype != torch.float64 or unit_vectors.ndim != 2 or \
                unit_vectors.shape[0] != 3:
            msg = "unit_vectors must be a float64 (3, n_rays) tensor"
            raise ValueError(msg)
        self.u = unit_vectors.contiguous()
        self.n_rays = self.u.shape[1]
        seg_np = np.asarray(seg_offsets.cpu() if isinstance(seg_offsets, torch.Tensor)
                            else seg_offsets, dtype=np.int64)
        if seg_np[-1] != self.n_rays:
            msg = f"seg_offsets[-1] = {seg_np[-1]} but there are {self.n_rays} rays"
            raise ValueError(msg)
        (self.seg, self.n_seg, self.obs, self.earth, self.beyond,
         self.mode) = device_model.prepare_observers(obs_xyz, earth_xyz, seg_np, per_ray=False)
        self.out_mode = OUT_COMPS if return_comps else OUT_SUM
        shape = (device_model.n_comps, self.n_rays) if return_comps else (self.n_rays,)
        self.out = torch.empty(shape, dtype=torch.float64, device=dev)
        self.launch = launch
        self.graph = None
        if use_graph:
            self._capture()

    def _launch(self) -> None:
        self.dm.forward_raw(self.u, self.seg, self.n_seg, self.obs, self.earth, self.beyond,
                            self.mode, self.out, self.out_mode, self.launch)

    def _capture(self) -> None:
        side = torch.cuda.Stream(device=self.dm.device)
        side.wait_stream(torch.cuda.current_stream(self.dm.device))
        with torch.cuda.stream(side):
            self._launch()  # warm-up outside capture (module load, first-launch setup)
        torch.cuda.current_stream(self.dm.device).wait_stream(side)
        torch.cuda.synchronize(self.dm.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self._launch()
        self.graph = graph

    def run(self) -> torch.Tensor:
        """Evaluate the batch on the current stream; returns the (reused) output tensor."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch()
        return self.out


class StreamedEvaluator:
    """A batch whose pointings start, and whose results end, in host memory.

    The rays are cut into ``n_chunks`` equal ranges (cuts may fall inside an
    observer's segment: each chunk carries the segment table and observer rows
    of the observers it touches, re-based to its first ray, so every ray sees
    the same observer, Earth position and cutoff bits as in the whole batch and
    gets bitwise the same result). ``run`` uploads chunk c + 1 on one copy
    stream and downloads chunk c - 1 on another while chunk c's kernel runs, so
    the PCIe transfers hide behind the compute. Two device buffers alternate for
    the pointings; the output stays device-resident at full size.

    Args:
        device_model: the packed model on the GPU.
        seg_offsets: (n_obs + 1,) int64, non-decreasing, from 0 to n_rays.
        obs_xyz, earth_xyz: (3, n_obs) heliocentric ecliptic positions [AU].
        n_chunks: number of ray ranges (at least 1).
        return_comps: per-component output (ncomps, n_rays) instead of the sum.
        launch: launch configuration (``None``: library default).
    """

    def __init__(self, device_model: DeviceModel, seg_offsets, obs_xyz, earth_xyz, *,
                 n_chunks: int = 8, return_comps: bool = False,
                 launch: Launch | None = None) -> None:
        self.dm = device_model
        dev = device_model.device
        seg = np.asarray(seg_offsets.cpu() if isinstance(seg_offsets, torch.Tensor)
                         else seg_offsets, dtype=np.int64)
        obs = np.asarray(obs_xyz, dtype=np.float64).reshape(3, -1)
        earth = np.asarray(earth_xyz, dtype=np.float64).reshape(3, -1)
        if seg.shape != (obs.shape[1] + 1,) or seg[0] != 0 or np.any(np.diff(seg) < 0):
            msg = "seg_offsets must have n_obs + 1 non-decreasing entries starting at 0"
            raise ValueError(msg)
        if n_chunks < 1:
            msg = "n_chunks must be at least 1"
            raise ValueError(msg)
        self.n_rays = int(seg[-1])
        self.return_comps = return_comps
        self.out_mode = OUT_COMPS if return_comps else OUT_SUM
        self.launch = launch
        bounds = np.linspace(0, self.n_rays, n_chunks + 1).round().astype(np.int64)
        self.chunks = []
        for r0, r1 in zip(bounds[:-1], bounds[1:]):
            if r1 <= r0:
                continue
            s0 = int(np.searchsorted(seg, r0, side="right")) - 1  # observer of ray r0
            s1 = int(np.searchsorted(seg, r1, side="left"))       # first observer from r1 on
            seg_c = np.clip(seg[s0:s1 + 1] - r0, 0, r1 - r0)
            seg_c[-1] = r1 - r0
            meta = device_model.prepare_observers(obs[:, s0:s1], earth[:, s0:s1], seg_c,
                                                  per_ray=False)
            self.chunks.append((int(r0), int(r1), meta))
        max_n = max(r1 - r0 for r0, r1, _ in self.chunks)
        self.ubuf = [torch.empty(3 * max_n, dtype=torch.float64, device=dev) for _ in range(2)]
        nc = device_model.n_comps
        self.out_shape = (nc, self.n_rays) if return_comps else (self.n_rays,)
        # summed output: device-resident at full size; per component: two chunk buffers
        self.out = None if return_comps else torch.empty(self.n_rays, dtype=torch.float64,
                                                          device=dev)
        self.obuf = ([torch.empty(nc * max_n, dtype=torch.float64, device=dev)
                      for _ in range(2)] if return_comps else None)
        self.h2d = torch.cuda.Stream(device=dev)
        self.d2h = torch.cuda.Stream(device=dev)

    def run(self, u_host: torch.Tensor, out_host: torch.Tensor) -> torch.Tensor:
        """Evaluate pointings ``u_host`` (3, n_rays) into ``out_host``; both pinned float64.

        Work is queued on the current stream and two copy streams; the current
        stream waits for both at the end, so synchronising it (or the device)
        makes ``out_host`` complete.
        """
        if tuple(u_host.shape) != (3, self.n_rays) or u_host.dtype != torch.float64:
            msg = f"u_host must be a float64 tensor of shape (3, {self.n_rays})"
            raise ValueError(msg)
        if tuple(out_host.shape) != self.out_shape or out_host.dtype != torch.float64:
            msg = f"out_host must be a float64 tensor of shape {self.out_shape}"
            raise ValueError(msg)
        cur = torch.cuda.current_stream(self.dm.device)
        self.h2d.wait_stream(cur)
        self.d2h.wait_stream(cur)
        computed_prev = [None, None]    # kernel of the last chunk that read each input buffer
        downloaded_prev = [None, None]  # download of the last chunk that used each output buffer
        for c, (r0, r1, meta) in enumerate(self.chunks):
            n = r1 - r0
            b = c % 2
            buf = self.ubuf[b][: 3 * n].view(3, n)
            if computed_prev[b] is not None:
                self.h2d.wait_event(computed_prev[b])
            with torch.cuda.stream(self.h2d):
                for k in range(3):
                    buf[k].copy_(u_host[k, r0:r1], non_blocking=True)
            uploaded = torch.cuda.Event()
            uploaded.record(self.h2d)
            cur.wait_event(uploaded)
            if self.return_comps:
                out = self.obuf[b][: self.out_shape[0] * n].view(self.out_shape[0], n)
                if downloaded_prev[b] is not None:
                    cur.wait_event(downloaded_prev[b])
            else:
                out = self.out[r0:r1]
            seg, n_seg, o, e, bey, mode = meta
            self.dm.forward_raw(buf, seg, n_seg, o, e, bey, mode, out, self.out_mode,
                                self.launch)
            computed = torch.cuda.Event()
            computed.record(cur)
            computed_prev[b] = computed
            self.d2h.wait_event(computed)
            with torch.cuda.stream(self.d2h):
                if self.return_comps:
                    for k in range(out.shape[0]):
                        out_host[k, r0:r1].copy_(out[k], non_blocking=True)
                else:
                    out_host[r0:r1].copy_(out, non_blocking=True)
            downloaded = torch.cuda.Event()
            downloaded.record(self.d2h)
            downloaded_prev[b] = downloaded
        cur.wait_stream(self.h2d)
        cur.wait_stream(self.d2h)
        return out_host


class HostMappedEvaluator:
    """A batch evaluated straight out of, and into, page-locked host memory.

    The kernel reads each ray's pointing from mapped host memory and writes its
    result there (zero-copy): the 24 + 8 bytes per ray cross PCIe once, inside
    the launch, with no copy engine involved. The kernel needs ~25 GB/s of host
    bandwidth at full speed, well inside a PCIe 5 x16 link, and each wave loads
    its pointings once before ~50 nodes of work, so the latency hides. Observer
    data stays in device memory.

    Args:
        device_model: the packed model on the GPU.
        seg_offsets: (n_obs + 1,) int64, non-decreasing, from 0 to n_rays.
        obs_xyz, earth_xyz: (3, n_obs) heliocentric ecliptic positions [AU].
        return_comps: per-component output (ncomps, n_rays) instead of the sum.
        launch: launch configuration (``None``: library default).
    """

    def __init__(self, device_model: DeviceModel, seg_offsets, obs_xyz, earth_xyz, *,
                 return_comps: bool = False, launch: Launch | None = None) -> None:
        from zodipy_amd import _lib

        self.dm = device_model
        self.lib = _lib.load()
        seg = np.asarray(seg_offsets.cpu() if isinstance(seg_offsets, torch.Tensor)
                         else seg_offsets, dtype=np.int64)
        self.n_rays = int(seg[-1])
        (self.seg, self.n_seg, self.obs, self.earth, self.beyond,
         self.mode) = device_model.prepare_observers(obs_xyz, earth_xyz, seg, per_ray=False)
        self.out_mode = OUT_COMPS if return_comps else OUT_SUM
        self.out_shape = (device_model.n_comps, self.n_rays) if return_comps else (self.n_rays,)
        self.launch = launch
        self._views: dict[int, int] = {}

    def _view(self, t: torch.Tensor, shape, name: str) -> int:
        from zodipy_amd import _lib

        if tuple(t.shape) != shape or t.dtype != torch.float64 or t.device.type != "cpu" \
                or not t.is_contiguous() or not t.is_pinned():
            msg = f"{name} must be a contiguous pinned float64 CPU tensor of shape {shape}"
            raise ValueError(msg)
        addr = t.data_ptr()
        if addr not in self._views:
            self._views[addr] = _lib.device_view(addr)
        return self._views[addr]

    def run(self, u_host: torch.Tensor, out_host: torch.Tensor) -> torch.Tensor:
        """Queue the batch on the current stream; ``out_host`` is complete after a sync."""
        from zodipy_amd import _lib
        from zodipy_amd.runtime import _stream

        du = self._view(u_host, (3, self.n_rays), "u_host")
        dout = self._view(out_host, self.out_shape, "out_host")
        hb, db = self.dm._ptrs()
        st = self.lib.zodi_forward(hb, db, du, self.n_rays, self.seg.data_ptr(), self.n_seg,
                                   self.obs.data_ptr(), self.earth.data_ptr(), self.mode,
                                   self.beyond.data_ptr(), dout, self.out_mode, self.launch,
                                   _stream())
        _lib.check(st, "zodi_forward")
        return out_host
