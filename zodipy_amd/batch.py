"""Batched evaluation of many observers in one graph-captured launch.

A batch is a set of rays (3, n_rays) resident in device memory, partitioned
into contiguous segments, one per observer (``seg_offsets``). Each segment has
its own observer and Earth position and its own cutoff test, i.e. the batch is
equivalent to one reference ``Model.evaluate`` call per observer
(zodipy/model.py:119-279) but runs as a single persistent-grid kernel launch.

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
