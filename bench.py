"""Benchmark: batched FP64 zodiacal line-of-sight integral on one MI355X.

Workload (timed): DIRBE model at 12 um, 2000 observers on a circular 1 AU
ecliptic orbit at seeded uniform phases, each with a seeded pointing count in
[1000, 50000] (about 51 M rays) of directions uniform on the sphere, 50
Gauss-Legendre nodes, summed output. One step = one graph-replayed launch over
the whole batch, inputs resident in HBM.

Also reported (not the headline value):
* plain: the per-observer drop-in loop (host pointings -> H2D, one launch,
  D2H, sync, per observer) over the same observers;
* batched_e2e: the batch with pointings in and results out in pinned host memory:
  the best of zero-copy (the kernel reads and writes the pinned memory
  directly), streamed (chunks whose copies overlap other chunks' kernels) and
  serial (upload, launch, download), each reported;
* fwd_grad_over_fwd: chi^2 forward + gradient time over forward time, DIRBE
  25 um at nside=128; ``grad`` adds the batched gradient of a six-component
  Planck-2018-style model at 857 GHz over 1000 observers;
* ingest: the pointing-ingestion kernels (HEALPix pixel centres and lon/lat to
  ecliptic unit vectors), HBM-bound, with their GB/s;
* cpu_baseline: the NumPy oracle with the reference's loop structure and its
  fork-pool split over host cores, one pool per observer call as the reference
  creates it, on a bounded prefix of the same observers; cpu_baselines adds the
  same for the single-sky workloads at full size (DIRBE 25 um at nside 16 and
  128, the nside-256 sky with a 50-wavelength Gaussian bandpass);
* single_sky: the nside-16 one-observer evaluation (latency, host in and out)
  and the nside-256 full sky with the 50-wavelength bandpass and per-component
  output (resident kernel time and samples/s).

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

SEED = 20261202
N_OBS = 2000
COUNT_LO, COUNT_HI = 1000, 50000
WAVELENGTH_UM = 12.0
N_NODES = 50

# Algorithmic flop count per ray-node and per ray for the DIRBE model with
# albedo 0 (frozen convention: +,-,*,/,sqrt = 1, FMA = 2, exp/log/trig = 20,
# pow = log + exp + 1): per distinct range 23 + 2 transcendentals (3 ranges),
# cloud 22 + 5, each band 33 + 4, ring 20 + 1, feature 31 + 2, accumulation 4
# per component; per ray three sphere intersections of 12 + 5.
FLOP_PER_RAY_NODE_DIRBE = 3 * (23 + 2 * 20) + (22 + 5 * 20) + 3 * (33 + 4 * 20) + (20 + 20) \
    + (31 + 2 * 20) + 6 * 4  # = 785
FLOP_PER_RAY_SETUP_DIRBE = 3 * (12 + 5 * 20)  # = 336
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_observers(seed: int = SEED, n_obs: int = N_OBS):
    """Observer phases, pointing counts and segment offsets (host, deterministic)."""
    rng = np.random.default_rng(seed)
    phases = rng.uniform(0.0, 2.0 * np.pi, n_obs)
    counts = rng.integers(COUNT_LO, COUNT_HI, n_obs, endpoint=True)
    seg = np.zeros(n_obs + 1, dtype=np.int64)
    np.cumsum(counts, out=seg[1:])
    obs = np.stack([np.cos(phases), np.sin(phases), np.zeros(n_obs)])
    return seg, obs


def make_directions(n: int, device, seed: int = SEED):
    """(3, n) float64 directions uniform on the sphere, generated on the GPU (seeded)."""
    import torch

    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    z = torch.rand(n, generator=gen, dtype=torch.float64, device=device).mul_(2.0).sub_(1.0)
    phi = torch.rand(n, generator=gen, dtype=torch.float64, device=device).mul_(2.0 * np.pi)
    s = torch.sqrt(1.0 - z * z)
    return torch.stack([s * torch.cos(phi), s * torch.sin(phi), z])


_CPU_MODEL = None


def _cpu_eval_chunk(args):
    u_chunk, obs = args
    return _CPU_MODEL.evaluate_vectors(u_chunk, obs, obs).sum(axis=0)


def cpu_baseline(u, seg, obs, budget_s: float) -> dict:
    """Reference-structured CPU evaluation of a prefix of the observers.

    The oracle model is built in this process before the pools fork, as the
    reference's workers inherit its Model (zodipy/model.py:182-198).
    """
    import multiprocessing as mp
    import platform

    from oracle.zodi_oracle import OracleModel
    from zodipy_amd import units

    global _CPU_MODEL
    _CPU_MODEL = OracleModel(WAVELENGTH_UM * units.micron)

    nproc = pool_size()
    ctx = mp.get_context("fork")
    done_rays = 0
    n_done = 0
    t0 = time.perf_counter()
    for o in range(len(seg) - 1):
        a, b = int(seg[o]), int(seg[o + 1])
        chunks = np.array_split(u[:, a:b], nproc, axis=1)
        with ctx.Pool(nproc) as pool:  # one pool per evaluate call (zodipy/model.py:191)
            parts = pool.map(_cpu_eval_chunk, [(c, obs[:, o]) for c in chunks])
        np.concatenate(parts)
        done_rays += b - a
        n_done += 1
        if time.perf_counter() - t0 > budget_s and n_done >= 3:
            break
    dt = time.perf_counter() - t0
    cpu_model = cpu_model_name()
    return {
        "value": done_rays * N_NODES / dt,
        "unit": "ray-node samples/s",
        "cores": nproc,
        "affinity_cores": len(os.sched_getaffinity(0)),
        "cores_note": POOL_NOTE,
        "kind": "port",
        "sample": (f"first {n_done} of {len(seg) - 1} observers ({done_rays} rays) of the timed "
                   f"workload, NumPy oracle, fork pool of {nproc} processes per observer call; "
                   f"{dt:.1f} s wall on {cpu_model} ({platform.machine()})"),
    }


# The host's share of CPU on a one-GPU box is 16 cores (its process limits), whatever
# the affinity mask reports; the pool is the reference's nprocesses at that share.
POOL_CAP = 16
POOL_NOTE = ("pool of min(affinity cores, 16) processes: 16 is the host CPU share of one GPU "
             "on the measuring box")


def pool_size() -> int:
    return min(len(os.sched_getaffinity(0)), POOL_CAP)


def cpu_model_name() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_eval_sky(args):
    u_chunk, obs = args
    return _CPU_MODEL.evaluate_vectors(u_chunk, obs, obs)


def cpu_sky(model_factory, u, obs, label: str) -> dict:
    """One reference-structured evaluate call over all of u (3, N) with one observer:
    the oracle, np.array_split over a fork pool created for the call
    (zodipy/model.py:182-198), per-component output. Full size, one timed call after
    one warm-up call."""
    import multiprocessing as mp

    global _CPU_MODEL
    _CPU_MODEL = model_factory()
    nproc = pool_size()
    ctx = mp.get_context("fork")

    def once():
        chunks = np.array_split(u, nproc, axis=1)
        with ctx.Pool(nproc) as pool:
            parts = pool.map(_cpu_eval_sky, [(c, obs) for c in chunks])
        return np.concatenate(parts, axis=1)

    once()
    t0 = time.perf_counter()
    once()
    dt = time.perf_counter() - t0
    n = u.shape[1]
    return {"value": n * N_NODES / dt, "unit": "ray-node samples/s", "seconds": round(dt, 4),
            "cores": nproc, "affinity_cores": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": f"{label}: all {n} rays, NumPy oracle, fork pool of {nproc} processes; "
                      f"{cpu_model_name()}"}


def _event_time(fn, reps: int) -> float:
    """Average ms of ``fn`` over ``reps`` calls, HIP events on the current stream."""
    import torch

    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def single_sky(dev, reps: int = 20) -> dict:
    """The single-sky workloads on the GPU.

    nside16: DIRBE 25 um, HEALPix nside=16 (3072 pixel centres), observer = Earth =
    (1, 0, 0) AU, per-component output: latency of one Model.evaluate_vectors call
    from host pointings to host results (median of ``reps``) and of the resident
    launch alone. fullsky256: DIRBE, nside=256 (786432 pixels), Gaussian bandpass
    at 25 um with sigma 2.5 um at 50 wavelengths (normalised weights), observer on a
    circular 1 AU orbit at phase 0.3 rad, per-component output, pointings resident:
    event-timed launch and ray-node samples/s.
    """
    import torch

    from zodipy_amd import Model, healpix
    from zodipy_amd import units as zu
    from zodipy_amd.runtime import OUT_COMPS

    out = {}
    v16 = healpix.pix2vec(16)
    obs = np.array([1.0, 0.0, 0.0])
    m1 = Model(25 * zu.micron, device=dev)
    for _ in range(3):
        m1.evaluate_vectors(v16, obspos=obs, earthpos=obs, return_comps=True)
    lat = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m1.evaluate_vectors(v16, obspos=obs, earthpos=obs, return_comps=True)
        lat.append(time.perf_counter() - t0)
    dm1 = m1.device_model
    args1 = dm1.prepare_observers(obs, obs, np.array([0, v16.shape[1]], dtype=np.int64))
    u16 = torch.as_tensor(v16, device=dev).contiguous()
    o16 = torch.empty((m1.ncomps, v16.shape[1]), dtype=torch.float64, device=dev)
    t_k = _event_time(lambda: dm1.forward_raw(u16, *args1, o16, OUT_COMPS), 50)
    out["nside16_dirbe25_one_observer"] = {
        "n_rays": int(v16.shape[1]), "latency_ms_median": round(1e3 * float(np.median(lat)), 4),
        "kernel_ms": round(t_k, 4), "note": "latency: host pointings in, host results out"}
    log(f"nside 16 (3072 rays, one observer): {1e3 * np.median(lat):.3f} ms per evaluate "
        f"(host in/out), {t_k * 1e3:.1f} us resident launch")

    x = np.linspace(17.5, 32.5, 50) * zu.micron
    w = np.exp(-0.5 * ((x.value - 25.0) / 2.5) ** 2)
    m2 = Model(x, weights=w, device=dev)
    dm2 = m2.device_model
    v256 = healpix.pix2vec(256)
    obs2 = np.array([np.cos(0.3), np.sin(0.3), 0.0])
    n = int(v256.shape[1])
    u256 = torch.as_tensor(v256, device=dev).contiguous()
    args2 = dm2.prepare_observers(obs2, obs2, np.array([0, n], dtype=np.int64))
    o256 = torch.empty((m2.ncomps, n), dtype=torch.float64, device=dev)
    dm2.forward_raw(u256, *args2, o256, OUT_COMPS)
    t2 = _event_time(lambda: dm2.forward_raw(u256, *args2, o256, OUT_COMPS), reps)
    out["fullsky_nside256_bandpass50"] = {
        "n_rays": n, "kernel_ms": round(t2, 4),
        "samples_per_s": n * N_NODES / (t2 / 1e3),
        "output_GBs": round(8 * m2.ncomps * n / (t2 / 1e3) / 1e9, 1)}
    log(f"full sky nside 256, 50-wavelength bandpass, per component: {t2:.3f} ms, "
        f"{n * N_NODES / (t2 / 1e3):.4e} samples/s")
    return out


def cpu_single_sky() -> dict:
    """CPU baselines of the single-sky workloads at full size (see single_sky), and of
    the nside-128 forward evaluation the gradient workload differentiates."""
    from oracle.zodi_oracle import OracleModel
    from zodipy_amd import healpix
    from zodipy_amd import units as zu

    out = {}
    obs = np.array([1.0, 0.0, 0.0])
    out["nside16_dirbe25_one_observer"] = cpu_sky(
        lambda: OracleModel(25 * zu.micron), healpix.pix2vec(16), obs, "DIRBE 25 um nside 16")
    out["nside128_dirbe25_one_observer"] = cpu_sky(
        lambda: OracleModel(25 * zu.micron), healpix.pix2vec(128), obs, "DIRBE 25 um nside 128")
    x = np.linspace(17.5, 32.5, 50) * zu.micron
    w = np.exp(-0.5 * ((x.value - 25.0) / 2.5) ** 2)
    obs2 = np.array([np.cos(0.3), np.sin(0.3), 0.0])
    out["fullsky_nside256_bandpass50"] = cpu_sky(
        lambda: OracleModel(x, weights=w), healpix.pix2vec(256), obs2,
        "DIRBE nside 256, 50-wavelength Gaussian bandpass")
    for k, v in out.items():
        log(f"cpu baseline {k}: {v['seconds']:.2f} s ({v['value']:.3e} samples/s)")
    return out


def planck18_style_model():
    """Six-component Planck-2018-style Kelsall model (DIRBE geometry).

    The registry's planck18 has four components, so the ring and feature take
    their emissivities from planck13; T_0 and delta are DIRBE's.
    """
    from zodipy_amd import component_params as cp
    from zodipy_amd import source_params as sp
    from zodipy_amd.component import ComponentLabel as L
    from zodipy_amd.zodiacal_light_model import Kelsall

    emiss = {**sp.EMISSIVITY_PLANCK_18, L.RING: sp.EMISSIVITY_PLANCK_13[L.RING],
             L.FEATURE: sp.EMISSIVITY_PLANCK_13[L.FEATURE]}
    return Kelsall(comps=cp.DIRBE, spectrum=sp.SPECTRUM_PLANCK, emissivities=emiss,
                   T_0=sp.T_0_DIRBE, delta=sp.DELTA_DIRBE)


def grad_cost(dev, reps: int = 10) -> dict:
    """Forward + chi^2 gradient over forward.

    (a) DIRBE 25 um, HEALPix nside=128 pixel centres, observer = Earth =
        (1, 0, 0) AU: target = model at theta with every parameter perturbed by
        seeded N(0, 1%) relative noise plus N(0, 0.01 MJy/sr); loss
        chi^2 = sum (I - d)^2 / 0.01^2. The library's chi^2 call (the Jacobian
        pass, then J^T g) against the forward launch alone, both at theta0. The
        same chi^2 as a function of the spectral tables (SpectralTableFit: device
        set-up of the effective values, chi^2, chain rule) is timed beside it,
        against the plain forward and against the fit's own forward (set-up and
        forward of its block, which carries the scattering term of every
        component with an albedo table),
        and the RRM model's chi^2 on the same rays (forward of its one-component-
        per-group packing against the chi^2 call).
    (b) Planck-2018-style six-component model at 857 GHz, 30-frequency top-hat
        bandpass of +-10% (771.3-942.7 GHz, extrapolated past 857 GHz), 1000
        observers on a 1 AU orbit with seeded 5000-40000 pointings each: the
        chi^2 call over the whole batch against the batched forward; the
        registry's four-component planck18 is timed the same way. The device
        set-up of that bandpass's Planck table (30 frequencies x 100
        temperatures) and effective spectral values is timed on its own, and
        so is chi^2 with its gradient with respect to the 30 bandpass weights
        (BandpassFit: set-up, one chi^2 call giving the effective-parameter and
        Planck-table gradients, chain rule) on the six-component batch.
    """
    import torch

    from zodipy_amd import Model, healpix
    from zodipy_amd import units as zu
    from zodipy_amd.fitting import BandpassFit, LineOfSightProblem, SpectralTableFit
    from zodipy_amd.zodiacal_light_model import model_registry

    out = {}
    model = Model(25 * zu.micron, name="dirbe", device=dev)
    v = healpix.pix2vec(128)
    obs = np.array([1.0, 0.0, 0.0])
    prob = LineOfSightProblem(model, v, obs, obs)
    theta0 = prob.space.theta0
    rng = np.random.default_rng(SEED + 4)
    theta_true = theta0 * (1.0 + 0.01 * rng.normal(size=theta0.size))
    noise = torch.as_tensor(0.01 * rng.normal(size=v.shape[1]), device=dev)
    data = prob.forward(theta_true) + noise
    inv_var = 1.0 / 0.01**2
    prob.forward(theta0)
    prob.chi2_and_grad(theta0, data, inv_var)
    t_fwd = _event_time(lambda: prob.forward(theta0), reps)
    t_fg = _event_time(lambda: prob.chi2_and_grad(theta0, data, inv_var), reps)
    c2, g = prob.chi2_and_grad(theta0, data, inv_var)
    out["chi2_dirbe25_nside128"] = {
        "n_rays": int(v.shape[1]), "n_params": prob.n_params,
        "fwd_ms": round(t_fwd, 4), "fwd_grad_ms": round(t_fg, 4),
        "ratio": round(t_fg / t_fwd, 4), "chi2_at_theta0": float(c2),
        "grad_norm": float(torch.linalg.vector_norm(g)),
    }
    log(f"chi2 gradient (DIRBE 25 um, nside 128, {prob.n_params} params): fwd {t_fwd:.3f} ms, "
        f"fwd+grad {t_fg:.3f} ms, ratio {t_fg / t_fwd:.3f}")
    rmodel = Model(25 * zu.micron, name="rrm-experimental", device=dev)
    rprob = LineOfSightProblem(rmodel, v, obs, obs)
    rth = rprob.space.theta0
    rdata = rprob.forward(rth) * 1.01
    rprob.chi2_and_grad(rth, rdata)
    rt_fwd = _event_time(lambda: rprob.forward(rth), reps)
    rt_fg = _event_time(lambda: rprob.chi2_and_grad(rth, rdata), reps)
    out["chi2_rrm25_nside128"] = {
        "n_rays": int(v.shape[1]), "n_params": rprob.n_params, "fwd_ms": round(rt_fwd, 4),
        "fwd_grad_ms": round(rt_fg, 4), "ratio": round(rt_fg / rt_fwd, 4)}
    log(f"chi2 gradient (RRM 25 um, nside 128, {rprob.n_params} params): fwd {rt_fwd:.3f} ms, "
        f"fwd+grad {rt_fg:.3f} ms, ratio {rt_fg / rt_fwd:.3f}")
    del rprob
    fit = SpectralTableFit(prob)
    vals = torch.as_tensor(fit.values0, device=dev)  # a device-resident fit's table vector
    fit.chi2_and_grad(vals, data, inv_var)
    t_tab = _event_time(lambda: fit.chi2_and_grad(vals, data, inv_var), reps)
    # The fit's own forward: its block carries the scattering term of every
    # component with an albedo table (the gradient needs it where albedo is 0).
    t_tab_fwd = _event_time(lambda: fit.forward(vals), reps)
    out["spectral_tables_dirbe25_nside128"] = {
        "n_table_entries": fit.n_params, "chi2_grad_ms": round(t_tab, 4),
        "ratio_vs_fwd": round(t_tab / t_fwd, 4), "fit_fwd_ms": round(t_tab_fwd, 4),
        "ratio_vs_fit_fwd": round(t_tab / t_tab_fwd, 4)}
    log(f"chi2 gradient w.r.t. {fit.n_params} spectral table entries: {t_tab:.3f} ms "
        f"(set-up + chi2 + chain rule), ratio {t_tab / t_fwd:.3f} against the plain forward, "
        f"{t_tab / t_tab_fwd:.3f} against the fit's forward with its scattering terms "
        f"({t_tab_fwd:.3f} ms)")

    name = "planck18-style-6comp"
    try:
        model_registry.get_model(name)
    except ValueError:
        model_registry.register_model(name, planck18_style_model())
    nu = np.linspace(771.3, 942.7, 30) * zu.GHz
    rng = np.random.default_rng(SEED + 5)
    n_obs = 1000
    counts = rng.integers(5000, 40000, n_obs, endpoint=True)
    seg = np.zeros(n_obs + 1, dtype=np.int64)
    np.cumsum(counts, out=seg[1:])
    phases = rng.uniform(0.0, 2.0 * np.pi, n_obs)
    pobs = np.stack([np.cos(phases), np.sin(phases), np.zeros(n_obs)])
    u_b = make_directions(int(seg[-1]), dev, SEED + 5).contiguous()
    for key, mname in (("batched_planck18style_857GHz", name),
                       ("batched_planck18_registry_857GHz", "planck18")):
        pmodel = Model(nu, weights=np.ones(30), name=mname, extrapolate=True, device=dev)
        bprob = LineOfSightProblem(pmodel, u_b, pobs, pobs, seg_offsets=seg)
        th = bprob.space.theta0
        bdata = bprob.forward(th) * 1.01
        bprob.chi2_and_grad(th, bdata)
        bt_fwd = _event_time(lambda: bprob.forward(th), max(2, reps // 2))
        bt_fg = _event_time(lambda: bprob.chi2_and_grad(th, bdata), max(2, reps // 2))
        out[key] = {
            "n_rays": int(seg[-1]), "n_observers": n_obs, "n_comps": bprob.n_comps,
            "n_params": bprob.n_params, "fwd_ms": round(bt_fwd, 3), "fwd_grad_ms": round(bt_fg, 3),
            "ratio": round(bt_fg / bt_fwd, 4),
            "fwd_grad_ray_node_samples_per_s": int(seg[-1]) * N_NODES / (bt_fg / 1e3),
        }
        log(f"batched chi2 gradient ({mname}, 857 GHz, {int(seg[-1])} rays): fwd {bt_fwd:.2f} ms, "
            f"fwd+grad {bt_fg:.2f} ms, ratio {bt_fg / bt_fwd:.3f}")
        if mname == name:
            sfit = SpectralTableFit(bprob)
            sfit.set_tables(sfit.values0, planck=True)
            t_set = _event_time(lambda: sfit.set_tables(sfit.values0, planck=True), reps)
            out["spectral_setup_857GHz_tophat30"] = {
                "n_freq": 30, "n_temps": 100, "n_tables": len(sfit.tables),
                "ms": round(t_set, 4)}
            log(f"spectral set-up on the device (30-frequency bandpass, {len(sfit.tables)} "
                f"tables): {t_set * 1e3:.1f} us")
            del sfit
            bfit = BandpassFit(bprob)
            w = torch.as_tensor(bfit.weights0, device=dev)
            bfit.chi2_and_grad(w, bdata)
            t_bp = _event_time(lambda: bfit.chi2_and_grad(w, bdata), max(2, reps // 2))
            t_sv = _event_time(lambda: bprob.chi2_spectral_grad(th, bdata), max(2, reps // 2))
            out["bandpass_weights_857GHz_tophat30"] = {
                "n_freq": 30, "chi2_grad_ms": round(t_bp, 3), "ratio_vs_fwd": round(t_bp / bt_fwd, 4),
                "chi2_spectral_grad_ms": round(t_sv, 3)}
            log(f"chi2 gradient w.r.t. the 30 bandpass weights ({mname}): {t_bp:.2f} ms "
                f"(set-up, chi2 with the spectral and Planck-table gradients {t_sv:.2f} ms, "
                f"chain rule), ratio {t_bp / bt_fwd:.3f}")
            del bfit
        del bprob
    out["ratio"] = out["chi2_dirbe25_nside128"]["ratio"]
    return out


def ingest_cost(dev, n_rays: int, reps: int = 5) -> dict:
    """Pointing ingestion kernels: HBM-bound conversions to ecliptic unit vectors.

    pix2vec: every pixel of a nside=2048 HEALPix map (50 331 648 pixels) in the
    galactic frame; lonlat: n_rays galactic longitude/latitude pairs (degrees);
    track positions: n_rays observation times on a one-year hourly spline;
    the time-ordered pipeline: both, then the forward pass with per-ray observers.
    Algorithmic bytes: 24 written per element, plus 16 read for lonlat and 8 for
    the track times.
    """
    import torch

    from zodipy_amd import pointing

    out = {}
    nside = 2048
    npix = 12 * nside * nside
    pointing.pix2vec(nside, frame="galactic")
    t = _event_time(lambda: pointing.pix2vec(nside, frame="galactic"), reps)
    out["pix2vec_nside2048_galactic"] = {
        "n": npix, "ms": round(t, 4), "GBs": round(24 * npix / (t / 1e3) / 1e9, 1),
        "frac_of_hbm": round(24 * npix / (t / 1e3) / 1e9 / PEAK_HBM_GBS, 3)}
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED + 7)
    lon = torch.rand(n_rays, generator=gen, dtype=torch.float64, device=dev) * 360.0
    lat = torch.rand(n_rays, generator=gen, dtype=torch.float64, device=dev) * 180.0 - 90.0
    pointing.lonlat_to_vectors(lon, lat, frame="galactic", degrees=True)
    t = _event_time(lambda: pointing.lonlat_to_vectors(lon, lat, frame="galactic",
                                                       degrees=True), reps)
    out["lonlat_galactic"] = {
        "n": n_rays, "ms": round(t, 4), "GBs": round(40 * n_rays / (t / 1e3) / 1e9, 1),
        "frac_of_hbm": round(40 * n_rays / (t / 1e3) / 1e9 / PEAK_HBM_GBS, 3)}
    from zodipy_amd.ephemeris import BodyTrack

    grid = np.arange(59580.0, 59580.0 + 365.0, 1.0 / 24.0)  # one year, hourly
    ph = 2.0 * np.pi * (grid - 59580.0) / 365.25
    track = BodyTrack(grid, np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)]))
    mjd = (torch.rand(n_rays, generator=gen, dtype=torch.float64, device=dev) * 364.0
           + 59580.0)
    mjd_sorted = torch.sort(mjd).values  # time-ordered data: a scan in time
    for key, times in (("spline_positions_one_year_hourly", mjd),
                       ("spline_positions_one_year_hourly_time_ordered", mjd_sorted)):
        track.on_device(times, dev)
        t = _event_time(lambda: track.on_device(times, dev), reps)
        out[key] = {
            "n": n_rays, "knots": int(grid.size), "ms": round(t, 4),
            "times": "random" if times is mjd else "sorted",
            "GBs": round(32 * n_rays / (t / 1e3) / 1e9, 1),
            "frac_of_hbm": round(32 * n_rays / (t / 1e3) / 1e9 / PEAK_HBM_GBS, 3)}
    # The time-ordered path end to end on the device: galactic lon/lat -> ecliptic
    # vectors, sorted times -> positions on the track, then the forward with per-ray
    # observers (one reference call over the span), against that forward alone.
    from zodipy_amd import Model
    from zodipy_amd import units as zu
    from zodipy_amd.runtime import OUT_SUM

    dm = Model(12 * zu.micron, device=dev).device_model

    def pipeline():
        vec = pointing.lonlat_to_vectors(lon, lat, frame="galactic", degrees=True)
        pos = track.on_device(mjd_sorted, dev)
        return dm.forward(vec, pos, pos, per_ray=True, out_mode=OUT_SUM)

    vec0 = pointing.lonlat_to_vectors(lon, lat, frame="galactic", degrees=True)
    pos0 = track.on_device(mjd_sorted, dev)
    pipeline()
    t_pipe = _event_time(pipeline, 3)
    t_fwd = _event_time(lambda: dm.forward(vec0, pos0, pos0, per_ray=True, out_mode=OUT_SUM), 3)
    out["tod_pipeline_12um"] = {
        "n": n_rays, "ms": round(t_pipe, 3), "forward_only_ms": round(t_fwd, 3),
        "overhead_share": round((t_pipe - t_fwd) / t_pipe, 4)}
    log(f"time-ordered pipeline (lon/lat + track positions + forward, {n_rays} rays): "
        f"{t_pipe:.2f} ms, forward alone {t_fwd:.2f} ms")
    del vec0, pos0
    log(f"ingestion: pix2vec nside 2048 {out['pix2vec_nside2048_galactic']['ms']:.3f} ms "
        f"({out['pix2vec_nside2048_galactic']['GBs']} GB/s), lonlat {n_rays} "
        f"{out['lonlat_galactic']['ms']:.3f} ms ({out['lonlat_galactic']['GBs']} GB/s), "
        f"track positions {out['spline_positions_one_year_hourly']['ms']:.3f} ms "
        f"({out['spline_positions_one_year_hourly']['GBs']} GB/s) random times, "
        f"{out['spline_positions_one_year_hourly_time_ordered']['ms']:.3f} ms "
        f"({out['spline_positions_one_year_hourly_time_ordered']['GBs']} GB/s) sorted")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--observers", type=int, default=N_OBS)
    ap.add_argument("--no-plain", action="store_true", help="skip the per-observer loop")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-grad", action="store_true", help="skip the fwd+grad/fwd ratio")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--chunks", type=int, default=8,
                    help="ray chunks of the streamed end-to-end run")
    ap.add_argument("--retune", action="store_true", help="ignore the tuning cache")
    ap.add_argument("--launch", default=None,
                    help="fixed launch 'block,lanes_per_ray,blocks_per_cu' instead of tuning")
    ap.add_argument("--profile", action="store_true",
                    help="short run for profilers: only the resident batched launches "
                         "(implies --no-plain --no-cpu --no-grad)")
    ap.add_argument("--grad-only", action="store_true",
                    help="profiling aid: time only the gradient workloads and print them "
                         "(not the benchmark line)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="profiling aid: one 1 GiB device-to-device copy (the reference "
                         "kernel for the FETCH_SIZE/WRITE_SIZE scale factors), then exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for {max(args.gpus, world)}); "
              "the batched integral shards by observer with no exchange for the forward pass, "
              "see the design notes", file=sys.stderr)
        return 2
    if args.profile:
        args.no_plain = args.no_cpu = args.no_grad = True

    import torch

    from zodipy_amd import Model, autotune
    from zodipy_amd import units as zu
    from zodipy_amd.batch import BatchedEvaluator, HostMappedEvaluator, StreamedEvaluator

    if args.grad_only:
        torch.cuda.set_device(0)
        print(json.dumps(grad_cost(torch.device("cuda:0"), reps=3)), flush=True)
        return 0
    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copied 1 GiB device to device")
        return 0

    seg, obs = make_observers(SEED, args.observers)
    n_rays = int(seg[-1])
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t_gen = time.perf_counter()
    u_dev = make_directions(n_rays, dev).contiguous()
    torch.cuda.synchronize()
    log(f"workload: {args.observers} observers, {n_rays} rays "
        f"({time.perf_counter() - t_gen:.2f} s to generate on the GPU)")
    model = Model(WAVELENGTH_UM * zu.micron, name="dirbe", device=dev)
    dm = model.device_model

    seg_t, n_seg, obs_t, earth_t, beyond_t, mode = dm.prepare_observers(obs, obs, seg)
    if args.launch:
        from zodipy_amd._lib import Launch

        b, lanes, per_cu = (int(v) for v in args.launch.split(","))
        launch = Launch(b, lanes, per_cu, autotune.VARIANT_AUTO)
    else:
        launch, _ = autotune.tune(dm, u_dev, seg_t, n_seg, obs_t, earth_t, beyond_t, mode,
                                  use_cache=not args.retune, verbose=True)
    log(f"launch: block {launch.block}, lanes/ray {launch.lanes_per_ray}, "
        f"blocks/CU {launch.blocks_per_cu}")
    batch = BatchedEvaluator(dm, u_dev, seg, obs, obs, return_comps=False, launch=launch)

    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        batch.run()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    samples = n_rays * N_NODES
    value = samples / (ms_per_step / 1e3)
    log(f"batched: {ms_per_step:.3f} ms/step ({wall / args.steps * 1e3:.3f} ms wall), "
        f"{value:.4e} ray-node samples/s")

    flops = n_rays * (N_NODES * FLOP_PER_RAY_NODE_DIRBE + FLOP_PER_RAY_SETUP_DIRBE)
    achieved = flops / (ms_per_step / 1e3) / 1e12
    # HBM bytes and instruction counts per launch of this kernel on this workload come
    # from rocprofv3 --pmc runs of the same command (committed summary); the counter-
    # derived rates below divide them by this run's launch time.
    traffic = prof = None
    tfile = ROOT / "profiles" / "forward_batched_dirbe12_traffic.json"
    if tfile.exists():
        try:
            prof = json.loads(tfile.read_text())
            traffic = prof.get("hbm_bytes_per_launch")
        except (OSError, json.JSONDecodeError):
            prof = None
    roofline = {
        "bound": "valu",
        "compute": "FP64 vector ALU (no matrix instructions); peak = MI355X FP64 78.6 TFLOP/s",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
        "traffic": traffic,
        "traffic_source": "profiles/forward_batched_dirbe12_traffic.json: rocprofv3 --pmc "
                          "FETCH_SIZE and WRITE_SIZE of this kernel on this workload, scaled by "
                          "a 1 GiB copy (not measured in this run)",
        "kernel": "forward_layout (DIRBE layout)",
        "flop_per_launch": flops,
        "algorithmic_bytes_per_launch": n_rays * (24 + 8),
        "output_write_GBs": round(n_rays * 8 / (ms_per_step / 1e3) / 1e9, 3),
    }
    if prof and "counters" in prof:
        c = prof["counters"]
        hw = 64 * (2 * c["SQ_INSTS_VALU_FMA_F64"] + c["SQ_INSTS_VALU_MUL_F64"]
                   + c["SQ_INSTS_VALU_ADD_F64"] + c["SQ_INSTS_VALU_TRANS_F64"])
        issue = c["SQ_INSTS_VALU"] * 4 / (1024 * 2.4e9)  # s of VALU issue per SIMD
        roofline["counters"] = {
            "hw_fp64_tflops": round(hw / (ms_per_step / 1e3) / 1e12, 2),
            "hw_fp64_frac": round(hw / (ms_per_step / 1e3) / 1e12 / PEAK_FP64_TFLOPS, 4),
            "valu_issue_frac": round(issue / (ms_per_step / 1e3), 4),
            "valu_per_ray_node": round(c["SQ_INSTS_VALU"] * 64 / (n_rays * N_NODES), 1),
            "source": "profiles/forward_batched_dirbe12_traffic.json (SQ_INSTS_VALU*, "
                      "wave-level counts per launch) over this run's launch time",
        }

    batched_e2e = None
    u_host = None
    if not args.profile:
        # End to end: pointings start and results end in pinned host memory.
        # Streamed: chunks whose uploads and downloads overlap the kernels of other
        # chunks (StreamedEvaluator). Serial: one upload, one graph replay, one download.
        u_pinned = torch.empty((3, n_rays), dtype=torch.float64).pin_memory()
        u_pinned.copy_(u_dev)
        u_host = u_pinned.numpy()
        out_host = torch.empty((n_rays,), dtype=torch.float64).pin_memory()
        streamed = StreamedEvaluator(dm, seg, obs, obs, n_chunks=args.chunks, launch=launch)
        streamed.run(u_pinned, out_host)
        torch.cuda.synchronize()
        e2e_runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            streamed.run(u_pinned, out_host)
            torch.cuda.synchronize()
            e2e_runs.append(time.perf_counter() - t0)
        e2e = min(e2e_runs)
        streamed_out = out_host.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u_dev.copy_(u_pinned, non_blocking=True)
        out = batch.run()
        out_host.copy_(out, non_blocking=True)
        torch.cuda.synchronize()
        e2e_serial = time.perf_counter() - t0
        if not torch.equal(streamed_out, out_host):
            raise RuntimeError("streamed and resident batches disagree")
        # Zero-copy: the kernel reads the pinned pointings and writes pinned results.
        mapped = HostMappedEvaluator(dm, seg, obs, obs, launch=launch)
        mapped_out = torch.empty((n_rays,), dtype=torch.float64).pin_memory()
        mapped.run(u_pinned, mapped_out)
        torch.cuda.synchronize()
        zc_runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            mapped.run(u_pinned, mapped_out)
            torch.cuda.synchronize()
            zc_runs.append(time.perf_counter() - t0)
        e2e_zc = min(zc_runs)
        if not torch.equal(mapped_out, out_host):
            raise RuntimeError("zero-copy and resident batches disagree")
        best = min((e2e_zc, "zero-copy"), (e2e, "streamed"), (e2e_serial, "serial"))
        batched_e2e = {"seconds": round(best[0], 4), "samples_per_s": samples / best[0],
                       "method": best[1], "zero_copy_seconds": round(e2e_zc, 4),
                       "streamed_seconds": round(e2e, 4), "streamed_chunks": args.chunks,
                       "serial_seconds": round(e2e_serial, 4),
                       "note": "pinned host pointings in, pinned host results out (PCIe included)"}
        log(f"batched end-to-end (host in, host out): zero-copy {e2e_zc * 1e3:.1f} ms, "
            f"streamed {e2e * 1e3:.1f} ms ({args.chunks} chunks), serial {e2e_serial * 1e3:.1f} ms")

    plain = None
    if not args.no_plain:
        res = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for o in range(len(seg) - 1):
            a, b = int(seg[o]), int(seg[o + 1])
            res.append(model.evaluate_vectors(u_host[:, a:b], obspos=obs[:, o],
                                              earthpos=obs[:, o]).value)
        t_plain = time.perf_counter() - t0
        plain = {"seconds": round(t_plain, 4), "samples_per_s": samples / t_plain,
                 "observers": len(seg) - 1}
        got = np.concatenate(res)
        ref = out_host.numpy()
        plain["max_rel_diff_vs_batched"] = float(np.max(np.abs(got - ref) / np.abs(ref)))
        log(f"plain per-observer loop: {t_plain:.3f} s, {plain['samples_per_s']:.4e} samples/s; "
            f"batched is {value / plain['samples_per_s']:.1f}x (kernel-only) and "
            f"{batched_e2e['samples_per_s'] / plain['samples_per_s']:.1f}x (end to end)")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(u_host, seg, obs, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.3e} samples/s ({cpu['sample']})")

    ingest = None if args.profile else ingest_cost(dev, n_rays)
    sky = None if args.profile else single_sky(dev)
    cpu_sky_base = None
    if not args.no_cpu:
        cpu_sky_base = cpu_single_sky()
        if sky:
            for k in sky:
                if k in cpu_sky_base:
                    sky[k]["cpu_seconds"] = cpu_sky_base[k]["seconds"]

    grad_ratio = None
    if not args.no_grad:
        grad_ratio = grad_cost(dev)

    line = {
        "metric": "ray-node samples/sec (FP64) batched vs plain, and fwd+grad/fwd time ratio",
        "value": value,
        "unit": "ray-node samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("batched time-ordered: DIRBE 12 um, 2000 observers on a 1 AU circular "
                         "orbit, 1000-50000 uniform-sphere pointings each, one launch"),
            "n_rays": n_rays,
            "n_observers": len(seg) - 1,
            "gauss_quad_degree": N_NODES,
            "output": "summed over components",
            "launch": {"block": launch.block, "lanes_per_ray": launch.lanes_per_ray,
                       "blocks_per_cu": launch.blocks_per_cu},
        },
        "component_samples_per_s": value * dm.n_comps,
        "n_comps": dm.n_comps,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baselines": cpu_sky_base,
        "single_sky": sky,
        "plain": plain,
        "batched_e2e": batched_e2e,
        "batched_vs_plain": (value / plain["samples_per_s"]) if plain else None,
        "batched_e2e_vs_plain": (batched_e2e["samples_per_s"] / plain["samples_per_s"])
        if plain and batched_e2e else None,
        "ingest": ingest,
        "fwd_grad_over_fwd": grad_ratio["ratio"] if grad_ratio else None,
        "grad": grad_ratio,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
