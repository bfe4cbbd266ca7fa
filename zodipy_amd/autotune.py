"""Launch-shape autotuner for the forward kernel.

Candidates are (threads per block, lanes per ray), each launched with one block
per tile of rays (blocks_per_cu = 0): the hardware dispatcher hands a CU a new
block whenever one retires, which is dynamic scheduling at tile granularity
already, so a persistent launch could only save the per-block prologue (staging
about 9 KB of tables into LDS, ~0.1 % of a block's work) and never shortens the
tail (one tile). Static grid-stride launches (blocks_per_cu > 0) measured 10 %
slower, because equal tile counts per block are not equal work. Every candidate
gives the same bits (lanes sharing a ray add its node terms in node order), so
the choice only sets the speed. Each candidate is timed with HIP events on a
representative slice of the real batch and the fastest is kept. Results are cached as JSON keyed by device name, model layout
(integrand, component kinds, groups), node count and a power-of-two size
bucket, in ``$ZODI_TUNE_CACHE`` or ``~/.cache/zodipy_amd/tuning.json``.
"""

from __future__ import annotations

import json
import os
import sys
from pathlib import Path

import numpy as np
import torch

from zodipy_amd import _lib
from zodipy_amd._lib import Launch
from zodipy_amd.runtime import OUT_SUM, DeviceModel

# Kernel variant: 1 = layout-specialised kernel when the model layout has one, else generic.
VARIANT_AUTO = 1

CANDIDATES = [
    (b, lanes, 0)
    for b in (64, 128, 256, 512)
    for lanes in (1, 2, 4, 8)
    if b // lanes >= 8
]


def cache_path() -> Path:
    env = os.environ.get("ZODI_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "zodipy_amd" / "tuning.json"


def layout_key(dm: DeviceModel, n_rays: int) -> str:
    b = dm.host_block[0]
    kinds = ",".join(str(int(k)) for k in b["comp_kind"][: int(b["n_comps"])])
    groups = ",".join(str(int(c)) for c in b["group_count"][: int(b["n_groups"])])
    name = torch.cuda.get_device_name(dm.device).replace(" ", "_")
    bucket = int(np.ceil(np.log2(max(n_rays, 1))))
    return (f"{name}|v{VARIANT_AUTO}|i{int(b['integrand'])}|k{kinds}|g{groups}"
            f"|n{int(b['n_nodes'])}|b{bucket}")


def _load() -> dict:
    p = cache_path()
    if p.exists():
        try:
            return json.loads(p.read_text())
        except (OSError, json.JSONDecodeError):
            return {}
    return {}


def _save(cache: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass


def time_launch(fn, reps: int = 3) -> float:
    """Median milliseconds of ``fn`` over ``reps`` event-timed runs (after one warm-up)."""
    fn()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    return float(np.median(times))


def tune(dm: DeviceModel, u: torch.Tensor, seg, n_seg: int, obs, earth, beyond, mode: int,
         *, candidates=None, max_rays: int = 4_000_000, use_cache: bool = True,
         verbose: bool = False) -> tuple[Launch, dict]:
    """Pick the fastest launch shape for this batch; returns (launch, {shape: ms})."""
    n_rays = u.shape[1]
    key = layout_key(dm, n_rays)
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        b, lanes, per_cu = cache[key]["best"]
        return Launch(b, lanes, per_cu, VARIANT_AUTO), cache[key].get("times", {})

    # Time on a prefix of the batch (whole segments) to bound tuning cost.
    if n_rays > max_rays and seg is not None:
        seg_np = seg.cpu().numpy()
        k = int(np.searchsorted(seg_np, max_rays, side="right")) - 1
        k = max(k, 1)
        n_sub = int(seg_np[k])
        u_sub = u[:, :n_sub].contiguous()
        seg_sub = seg[: k + 1].contiguous()
        nseg_sub = k
        # observer rows have length n_seg: the prefix needs its own (3, k) copies
        obs_sub = obs[:, :k].contiguous()
        earth_sub = earth[:, :k].contiguous()
        beyond_sub = beyond[:k].contiguous()
    else:
        u_sub, seg_sub, nseg_sub, obs_sub, earth_sub, beyond_sub = u, seg, n_seg, obs, earth, beyond
        n_sub = n_rays
    out = torch.empty((n_sub,), dtype=torch.float64, device=u.device)
    times = {}
    for b, lanes, per_cu in candidates or CANDIDATES:
        launch = Launch(b, lanes, per_cu, VARIANT_AUTO)

        def fn(launch=launch):
            st = dm.lib.zodi_forward(
                dm.host_block.ctypes.data, dm.dev_block.data_ptr(), u_sub.data_ptr(), n_sub,
                seg_sub.data_ptr() if seg_sub is not None else None,
                nseg_sub if seg_sub is not None else 0, obs_sub.data_ptr(), earth_sub.data_ptr(),
                mode, beyond_sub.data_ptr(), out.data_ptr(), OUT_SUM, launch,
                torch.cuda.current_stream().cuda_stream)
            _lib.check(st, "zodi_forward (tuning)")

        times[f"{b},{lanes},{per_cu}"] = time_launch(fn)
        if verbose:
            print(f"  tune {b:4d} threads, {lanes} lanes/ray, {per_cu} blocks/CU: "
                  f"{times[f'{b},{lanes},{per_cu}']:.3f} ms", file=sys.stderr, flush=True)
    best = min(times, key=times.get)
    b, lanes, per_cu = (int(v) for v in best.split(","))
    if use_cache:
        cache[key] = {"best": [b, lanes, per_cu], "times": times}
        _save(cache)
    return Launch(b, lanes, per_cu, VARIANT_AUTO), times
