"""GPU parity of the fused forward kernel against the CPU oracle.

Tolerance: 1e-10 relative per element (NaN must match NaN, exact zeros must be
exact zeros). Inputs are seeded; sizes keep the oracle to a few seconds.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle.zodi_oracle import OracleModel, evaluate_segments
from zodipy_amd import Model, healpix
from zodipy_amd import units as u
from zodipy_amd._lib import Launch
from zodipy_amd.runtime import OUT_COMPS, OUT_SUM

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def assert_parity(got, want, rtol=RTOL):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    assert np.array_equal(nan_g, nan_w), f"NaN mismatch: {nan_g.sum()} vs {nan_w.sum()}"
    g, w = got[~nan_w], want[~nan_w]
    zero = w == 0
    assert np.all(g[zero] == 0), "exact zeros differ"
    rel = np.abs(g - w) / np.maximum(np.abs(w), np.finfo(float).tiny)
    worst = rel.max(initial=0.0)
    assert worst <= rtol, f"max relative error {worst:.3e} > {rtol:.1e}"


def random_dirs(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z])


def test_dirbe_25um_nside16_one_observer(gpu):
    """DIRBE, 25 um, nside=16 pixel centres, observer = Earth = (1, 0, 0) AU."""
    v = healpix.pix2vec(16)
    obs = np.array([1.0, 0.0, 0.0])
    model = Model(25 * u.micron)
    got = model.evaluate_vectors(v, obspos=obs, earthpos=obs, return_comps=True).value
    want = OracleModel(25 * u.micron).evaluate_vectors(v, obs, obs)
    assert_parity(got, want)


@pytest.mark.parametrize(
    "name,x,kw",
    [
        ("dirbe", 1.25 * u.micron, {}),  # scattering branch (albedo != 0)
        ("dirbe", 3.5 * u.micron, {}),
        ("dirbe", 140 * u.micron, {}),
        ("planck13", 857 * u.GHz, {}),
        ("planck15", 353 * u.GHz, {}),
        ("planck18", 545 * u.GHz, {}),
        ("odegard", 100 * u.GHz, {}),
        ("rrm-experimental", 25 * u.micron, {}),
        ("rrm-experimental", 75 * u.micron, {}),
        ("dirbe", 0.9 * u.micron, {"extrapolate": True}),
    ],
)
def test_registry_models_off_ecliptic_observer(gpu, name, x, kw):
    """Every registered model; observer off the ecliptic (z0 enters c but not b)."""
    v = random_dirs(2000, seed=7)
    obs = np.array([0.62, -0.78, 0.05])
    earth = np.array([0.60, -0.80, 0.0])
    model = Model(x, name=name, **kw)
    got = model.evaluate_vectors(v, obspos=obs, earthpos=earth, return_comps=True).value
    want = OracleModel(x, name=name, **kw).evaluate_vectors(v, obs, earth)
    assert_parity(got, want)


def test_bandpass_gaussian_25um(gpu):
    """Gaussian bandpass of 50 wavelengths around 25 um (normalised weights)."""
    x = np.linspace(17.5, 32.5, 50) * u.micron
    w = np.exp(-0.5 * ((x.value - 25) / 2.5) ** 2)
    v = healpix.pix2vec(8)
    obs = np.array([np.cos(0.3), np.sin(0.3), 0.0])
    got = Model(x, weights=w).evaluate_vectors(v, obspos=obs, earthpos=obs,
                                               return_comps=True).value
    want = OracleModel(x, weights=w).evaluate_vectors(v, obs, obs)
    assert_parity(got, want)


def test_bandpass_dirbe_25um_instrument(gpu):
    """The DIRBE 25 um band's own bandpass (82 wavelengths and weights from the
    reference's test data, tests/golden/dirbe_25um_bandpass.json), on an
    off-ecliptic observer."""
    import json
    from pathlib import Path

    d = json.loads((Path(__file__).parent / "golden" / "dirbe_25um_bandpass.json").read_text())
    x = np.asarray(d["wavelength_micron"]) * u.micron
    w = np.asarray(d["weight"])
    v = random_dirs(700, seed=17)
    obs = np.array([0.7, 0.7, 0.05])
    got = Model(x, weights=w).evaluate_vectors(v, obspos=obs, earthpos=obs,
                                               return_comps=True).value
    want = OracleModel(x, weights=w).evaluate_vectors(v, obs, obs)
    assert_parity(got, want)


def test_bandpass_ghz_on_micron_model(gpu):
    """Bandpass in GHz on the micron-tabulated DIRBE model (unit-axis quirk kept)."""
    x = np.linspace(11000, 13000, 20) * u.GHz
    w = np.ones(20)
    v = random_dirs(500, seed=3)
    obs = np.array([1.0, 0.0, 0.0])
    got = Model(x, weights=w).evaluate_vectors(v, obspos=obs, earthpos=obs,
                                               return_comps=True).value
    want = OracleModel(x, weights=w).evaluate_vectors(v, obs, obs)
    assert_parity(got, want)


def test_time_ordered_per_ray_observers(gpu):
    """Per-ray observers: one cutoff test covers all observers of the call."""
    n = 1500
    v = random_dirs(n, seed=11)
    phase = np.linspace(0, 1.0, n)
    obs = np.stack([np.cos(phase), np.sin(phase), 0.01 * np.sin(3 * phase)])
    earth = np.stack([np.cos(phase), np.sin(phase), np.zeros(n)])
    model = Model(12 * u.micron)
    got = model.evaluate_vectors(v, obspos=obs, earthpos=earth, return_comps=True).value
    want = OracleModel(12 * u.micron).evaluate_vectors(v, obs, earth)
    assert_parity(got, want)
    # one observer beyond the ring/feature cutoffs zeroes those components for every ray
    obs2 = obs.copy()
    obs2[:, 17] *= 1.35
    got2 = model.evaluate_vectors(v, obspos=obs2, earthpos=earth, return_comps=True).value
    want2 = OracleModel(12 * u.micron).evaluate_vectors(v, obs2, earth)
    assert_parity(got2, want2)
    assert np.all(want2[4:] == 0)


def test_batched_segments_uneven(gpu):
    """Many observers with uneven (including empty) pointing counts in one launch."""
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 700, 23)
    counts[[3, 10]] = 0
    n = int(counts.sum())
    seg = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    v = random_dirs(n, seed=9)
    ph = rng.uniform(0, 2 * np.pi, counts.size)
    obs = np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)])
    obs[:, 4] *= 1.4  # this observer is outside the ring/feature spheres
    model = Model(12 * u.micron)
    dm = model.device_model
    got = dm.forward(v, obs, obs, seg_offsets=seg, out_mode=OUT_COMPS).cpu().numpy()
    want = evaluate_segments(OracleModel(12 * u.micron), v, seg, obs, obs)
    assert_parity(got, want)
    tot = dm.forward(v, obs, obs, seg_offsets=seg, out_mode=OUT_SUM).cpu().numpy()
    assert np.array_equal(tot, got.sum(axis=0))


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("block,lanes,per_cu", [(64, 1, 0), (128, 2, 0), (256, 4, 2),
                                                (512, 8, 1), (256, 1, 4)])
def test_launch_variants_agree(gpu, block, lanes, per_cu, variant):
    """Generic (variant 0) and layout-specialised (variant 1) kernels, every launch shape."""
    rng = np.random.default_rng(2)
    counts = rng.integers(1, 400, 9)
    seg = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    v = random_dirs(int(counts.sum()), seed=21)
    ph = rng.uniform(0, 2 * np.pi, counts.size)
    obs = np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)])
    model = Model(25 * u.micron)
    dm = model.device_model
    launch = Launch(block=block, lanes_per_ray=lanes, blocks_per_cu=per_cu, variant=variant)
    got = dm.forward(v, obs, obs, seg_offsets=seg, launch=launch).cpu().numpy()
    want = evaluate_segments(OracleModel(25 * u.micron), v, seg, obs, obs)
    assert_parity(got, want)


def test_return_comps_sum_is_exact(gpu):
    v = random_dirs(777, seed=1)
    obs = np.array([0.3, 0.95, 0.0])
    model = Model(20 * u.micron)
    comps = model.evaluate_vectors(v, obspos=obs, earthpos=obs, return_comps=True)
    total = model.evaluate_vectors(v, obspos=obs, earthpos=obs, return_comps=False)
    assert comps.shape == (6, 777) and total.shape == (777,)
    assert np.array_equal(comps.value.sum(axis=0), total.value)


@pytest.mark.parametrize("degree", [1, 7, 100])
def test_gauss_quad_degree(gpu, degree):
    v = random_dirs(300, seed=4)
    obs = np.array([1.0, 0.0, 0.0])
    got = Model(25 * u.micron, gauss_quad_degree=degree).evaluate_vectors(
        v, obspos=obs, earthpos=obs, return_comps=True).value
    want = OracleModel(25 * u.micron, gauss_quad_degree=degree).evaluate_vectors(v, obs, obs)
    assert_parity(got, want)


def test_special_directions_and_nan(gpu):
    """Poles, axis directions, and b == 0 exactly (observer on the z axis) -> NaN."""
    v = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]],
                 dtype=np.float64).T
    for obs in (np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 0.5]),
                np.array([0.0, 1.0, 0.2])):
        got = Model(25 * u.micron).evaluate_vectors(v, obspos=obs, earthpos=obs,
                                                    return_comps=True).value
        want = OracleModel(25 * u.micron).evaluate_vectors(v, obs, obs)
        assert_parity(got, want)


def test_line_of_sight_kernel(gpu):
    from oracle.zodi_oracle import get_sphere_intersection
    from zodipy_amd.cutoffs import COMPONENT_CUTOFFS

    v = random_dirs(1000, seed=8)
    obs = np.array([0.9, 0.3, 0.02])
    for name in ("dirbe", "rrm-experimental"):
        model = Model(25 * u.micron, name=name)
        dm = model.device_model
        start, stop = dm.line_of_sight(v, obs)
        start, stop = start.cpu().numpy(), stop.cpu().numpy()
        for c, label in enumerate(model.comps):
            g = dm.packed.group_of_comp[c]
            r_in, r_out = COMPONENT_CUTOFFS[label]
            want0 = np.broadcast_to(get_sphere_intersection(obs[:, None], v, r_in), (1000,))
            want1 = np.broadcast_to(get_sphere_intersection(obs[:, None], v, r_out), (1000,))
            assert_parity(start[g], want0, rtol=1e-14)
            assert_parity(stop[g], want1, rtol=1e-14)


def test_number_density_kernel(gpu):
    from zodipy_amd.number_density import density_device_model
    from zodipy_amd.model_registry import model_registry

    rng = np.random.default_rng(12)
    X = rng.uniform(-3.5, 3.5, (3, 4000))
    X[2] *= 0.3
    earth = np.array([0.7, -0.7, 0.0])
    for name in ("dirbe", "rrm-experimental"):
        zm = model_registry.get_model(name)
        got = density_device_model(zm).number_density(X, earth).cpu().numpy()
        want = OracleModel(25 * u.micron, name=name).number_density(X, earth)
        assert_parity(got, want)


def test_feature_longitude_axis_and_wrap(gpu):
    """The feature's longitude offset on its axis (x' = y' = 0, with +0 and -0,
    where the reference's atan2 of signed zeros gives 0 or +-pi) and at points
    whose offset lies at the wrap (+-pi) or crosses it, for several Earth
    longitudes including pi and -pi."""
    from zodipy_amd.model_registry import model_registry
    from zodipy_amd.number_density import density_device_model

    zm = model_registry.get_model("dirbe")
    om = OracleModel(25 * u.micron)
    z = np.linspace(-0.3, 0.3, 7)
    axis = np.concatenate([np.stack([np.zeros(7), np.zeros(7), z]),
                           np.stack([-np.zeros(7), np.zeros(7), z]),
                           np.stack([np.zeros(7), -np.zeros(7), z]),
                           np.stack([-np.zeros(7), -np.zeros(7), z])], axis=1)
    # the last three put the Earth on the feature's axis (E' = +-0): atan2 of
    # signed zeros gives its longitude (0 or pi)
    for earth in ([1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [-1.0, -0.0, 0.0], [0.3, -0.95, 0.0],
                  [0.0, 0.0, 0.0], [-0.0, 0.0, 0.1], [0.0, -0.0, -0.1]):
        earth = np.array(earth)
        lon_e = np.arctan2(earth[1], earth[0])
        theta_f = np.radians(-10.0)  # the DIRBE feature's offset
        # points 1 AU out at the longitude where the offset is pi, and 1e-9 rad to
        # either side of it
        lon = lon_e + theta_f + np.pi + np.array([-1e-9, 0.0, 1e-9, 3.0, -3.0])
        ring = np.stack([np.cos(lon), np.sin(lon), np.full(lon.size, 0.01)])
        X = np.concatenate([axis, ring], axis=1)
        got = density_device_model(zm).number_density(X, earth).cpu().numpy()
        want = om.number_density(X, earth)
        assert_parity(got, want, rtol=1e-12)


@pytest.mark.parametrize("name", ["dirbe", "rrm-experimental"])
def test_short_planck_table(gpu, name):
    """A model block with fewer than 100 tabulated temperatures (the C ABI allows
    2..100): np.interp holds B at the last one above it. With the first 40 of the
    registry table (40 - 240.9 K; an observer at 1 AU sees up to ~290 K) and
    NaN in every unused entry, the forward must equal that of the full block
    whose values beyond the 40th temperature repeat B_39 (the same piecewise-
    linear function), and read nothing past the 40th entry."""
    import dataclasses

    from zodipy_amd.runtime import DeviceModel

    model = Model(25 * u.micron, name=name)
    packed = model.device_model.packed
    n = 40
    short = packed.block.copy()
    short["n_temps"] = n
    short["tab_T"][0, n:] = np.nan
    short["tab_B"][0, n:] = np.nan
    short["tab_slope"][0, n - 1:] = np.nan
    t = short["tab_T"][0]
    short["tab_inv_step"] = (n - 1) / (t[n - 1] - t[0])
    flat = packed.block.copy()
    flat["tab_B"][0, n:] = flat["tab_B"][0, n - 1]
    flat["tab_slope"][0, n - 1:] = 0.0
    v = torch.as_tensor(healpix.pix2vec(8), device="cuda")
    obs = np.array([0.95, 0.3, 0.01])
    out = []
    for block in (short, flat):
        dm = DeviceModel(dataclasses.replace(packed, block=block))
        out.append(dm.forward(v, obs, obs).cpu().numpy())
    assert np.all(np.isfinite(out[0]))
    np.testing.assert_allclose(out[0], out[1], rtol=1e-14, atol=0)
    # some ray-node temperatures lie above the short table's last one
    assert not np.allclose(out[0], model.evaluate_vectors(
        v.cpu().numpy(), obspos=obs, earthpos=obs, return_comps=True).value, rtol=1e-6)


def test_device_tensor_output_and_stream(gpu):
    v = torch.as_tensor(random_dirs(4096, seed=30), device="cuda")
    obs = np.array([1.0, 0.0, 0.0])
    model = Model(25 * u.micron)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        out = model.evaluate_vectors(v, obspos=obs, earthpos=obs, as_tensor=True)
    s.synchronize()
    assert out.is_cuda and out.shape == (4096,)
    want = OracleModel(25 * u.micron).evaluate_vectors(v.cpu().numpy(), obs, obs).sum(axis=0)
    assert_parity(out.cpu().numpy(), want)


def _uneven_batch(seed, n_obs=17):
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, 900, n_obs)
    counts[[2, 9]] = 0
    seg = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ph = rng.uniform(0, 2 * np.pi, n_obs)
    obs = np.stack([np.cos(ph), np.sin(ph), 0.003 * np.cos(3 * ph)])
    obs[:, 5] *= 1.25  # outside the ring's stop sphere
    earth = np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)])
    v = random_dirs(int(seg[-1]), seed=seed + 1)
    return v, seg, obs, earth


@pytest.mark.parametrize("n_chunks", [1, 3, 7])
@pytest.mark.parametrize("return_comps", [False, True])
def test_streamed_evaluator_matches_resident_batch(gpu, n_chunks, return_comps):
    """Host-to-host pipelined chunks (cuts inside segments) give the resident batch's bits."""
    from zodipy_amd.batch import BatchedEvaluator, StreamedEvaluator

    v, seg, obs, earth = _uneven_batch(31)
    dm = Model(12 * u.micron).device_model
    u_dev = torch.as_tensor(v, device="cuda").contiguous()
    want = BatchedEvaluator(dm, u_dev, seg, obs, earth, return_comps=return_comps,
                            use_graph=False).run().cpu()
    st = StreamedEvaluator(dm, seg, obs, earth, n_chunks=n_chunks, return_comps=return_comps)
    u_host = torch.as_tensor(v).pin_memory()
    out_host = torch.empty(tuple(want.shape), dtype=torch.float64).pin_memory()
    for _ in range(2):  # the second run reuses the buffers
        out_host.fill_(-1.0)
        st.run(u_host, out_host)
        torch.cuda.synchronize()
        assert torch.equal(out_host, want)
    ref = evaluate_segments(OracleModel(12 * u.micron), v, seg, obs, earth)
    assert_parity(want.numpy(), ref if return_comps else ref.sum(axis=0))


def test_batched_graph_replays_match_eager(gpu):
    """Graph replays, with eager work and synchronisation between them, match eager launches."""
    from zodipy_amd.batch import BatchedEvaluator

    v, seg, obs, earth = _uneven_batch(41)
    dm = Model(25 * u.micron).device_model
    u_dev = torch.as_tensor(v, device="cuda").contiguous()
    eager = BatchedEvaluator(dm, u_dev, seg, obs, earth, use_graph=False).run().clone()
    graphed = BatchedEvaluator(dm, u_dev, seg, obs, earth, use_graph=True)
    for k in range(4):
        graphed.out.fill_(float(k))
        out = graphed.run()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
        scratch = torch.randn(1000, device="cuda").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)


@pytest.mark.parametrize("return_comps", [False, True])
def test_host_mapped_evaluator_matches_resident_batch(gpu, return_comps):
    """Zero-copy: the kernel reads pinned host pointings and writes pinned host results."""
    from zodipy_amd.batch import BatchedEvaluator, HostMappedEvaluator

    v, seg, obs, earth = _uneven_batch(51)
    dm = Model(12 * u.micron).device_model
    u_dev = torch.as_tensor(v, device="cuda").contiguous()
    want = BatchedEvaluator(dm, u_dev, seg, obs, earth, return_comps=return_comps,
                            use_graph=False).run().cpu()
    hm = HostMappedEvaluator(dm, seg, obs, earth, return_comps=return_comps)
    u_host = torch.as_tensor(v).contiguous().pin_memory()
    out_host = torch.empty(tuple(want.shape), dtype=torch.float64).pin_memory()
    for _ in range(2):
        out_host.fill_(-1.0)
        hm.run(u_host, out_host)
        torch.cuda.synchronize()
        assert torch.equal(out_host, want)


def test_host_mapped_evaluator_rejects_pageable_memory(gpu):
    from zodipy_amd.batch import HostMappedEvaluator

    v, seg, obs, earth = _uneven_batch(52)
    hm = HostMappedEvaluator(Model(12 * u.micron).device_model, seg, obs, earth)
    out = torch.empty(int(seg[-1]), dtype=torch.float64).pin_memory()
    with pytest.raises(ValueError, match="pinned"):
        hm.run(torch.as_tensor(v).contiguous(), out)  # pageable pointings


@pytest.mark.parametrize("x", [25 * u.micron, 1.25 * u.micron])
def test_custom_layout_generic_kernel(gpu, cloud_ring_model, x):
    """A registered model outside the registry layouts (cloud and ring only) runs
    the generic kernel: parity with the oracle, scattering included at 1.25 um."""
    from zodipy_amd.zodiacal_light_model import model_registry

    v = random_dirs(1500, seed=31)
    obs = np.array([0.62, -0.78, 0.05])
    earth = np.array([0.60, -0.80, 0.0])
    model = Model(x, name=cloud_ring_model)
    got = model.evaluate_vectors(v, obspos=obs, earthpos=earth, return_comps=True).value
    want = OracleModel(x, model=model_registry.get_model(cloud_ring_model)).evaluate_vectors(
        v, obs, earth)
    assert_parity(got, want)


def test_rrm_one_component_per_group(gpu):
    """The RRM model packed one component per group (as parameter fits pack it):
    the specialised kernel for that structure matches the oracle."""
    from zodipy_amd.fitting import LineOfSightProblem

    v = random_dirs(1500, seed=32)
    obs = np.array([0.62, -0.78, 0.05])
    earth = np.array([0.60, -0.80, 0.0])
    model = Model(60 * u.micron, name="rrm-experimental")
    prob = LineOfSightProblem(model, v, obs, earth)
    assert prob.space.pack(prob.space.theta0).n_groups == 8
    got = prob.forward(prob.space.theta0, return_comps=True).cpu().numpy()
    want = OracleModel(60 * u.micron, name="rrm-experimental").evaluate_vectors(v, obs, earth)
    assert_parity(got, want)


@pytest.mark.parametrize("which", ["planck18", "six-component"])
def test_857ghz_tophat_extrapolated_batched(gpu, planck18_style, which):
    """The 30-frequency top hat of +-10 % around 857 GHz (771.3-942.7 GHz, weights 1),
    which needs extrapolate=True past the last tabulated frequency
    (zodipy/zodiacal_light_model.py:62-69, zodipy/unpack_model.py:140-173), over
    uneven observer segments in one launch: registry planck18 (4 components) and the
    six-component Planck-2018-style model against the oracle at 1e-10."""
    name = "planck18" if which == "planck18" else planck18_style
    rng = np.random.default_rng(8571)
    counts = rng.integers(50, 400, 6)
    seg = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ph = rng.uniform(0, 2 * np.pi, counts.size)
    obs = np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)])
    v = random_dirs(int(seg[-1]), seed=8572)
    nu = np.linspace(771.3, 942.7, 30) * u.GHz
    w = np.ones(30)
    with pytest.raises(ValueError):
        Model(nu, weights=w, name=name)
    dm = Model(nu, weights=w, name=name, extrapolate=True).device_model
    got = dm.forward(v, obs, obs, seg_offsets=seg, out_mode=OUT_COMPS).cpu().numpy()
    want = evaluate_segments(OracleModel(nu, weights=w, name=name, extrapolate=True), v, seg,
                             obs, obs)
    assert got.shape[0] == (4 if which == "planck18" else 6)
    assert_parity(got, want)


def test_rrm_nan_direction_with_observer_beyond_cutoffs(gpu):
    """RRM with NaN pointings, b == 0 rays and an observer outside some groups' cutoff
    spheres (2 AU: beyond the fan's R_outer and both ring/feature spheres, so those
    ranges are [eps, eps]). The reference's masked densities (fan, comet, narrow and
    broad bands) are 0 at a NaN position but the Planck lookup gives NaN there, so every
    component of a NaN ray is NaN, even where the range is empty."""
    v = random_dirs(300, seed=41)
    v[:, 5] = np.nan
    v[:, 17] = [np.nan, 0.0, 0.0]
    v[:, 30] = [0.0, 0.0, 1.0]
    for obs in (np.array([2.0, 0.0, 0.0]), np.array([0.0, 0.0, 0.5]),
                np.array([0.9, -0.2, 0.01])):
        model = Model(25 * u.micron, name="rrm-experimental")
        got = model.evaluate_vectors(v, obspos=obs, earthpos=obs, return_comps=True).value
        want = OracleModel(25 * u.micron, name="rrm-experimental").evaluate_vectors(v, obs, obs)
        assert np.isnan(want[:, 5]).all()
        assert_parity(got, want)


def test_published_usage_value_through_device_pointing(gpu):
    """The reference's published output (docs/usage.md:199-203): DIRBE at 25 um, galactic
    (40, 60) deg on 2020-01-01, observer (0.87, -0.53, 0.001) AU -> 20.37750965 MJy/sr,
    through Model.evaluate: the galactic pointing is rotated to the ecliptic on the device,
    the Earth position (which only the feature uses here) comes from the analytic
    ephemeris of tests/golden/make_dirbe_kat.py."""
    import sys
    from pathlib import Path

    from zodipy_amd.coords import SkyCoord, Time

    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    from make_dirbe_kat import emb_heliocentric

    earth = emb_heliocentric(2458849.5 + 69.184 / 86400.0)
    sc = SkyCoord(40.0, 60.0, unit=u.deg, frame="galactic", obstime=Time("2020-01-01"))
    model = Model(25 * u.micron)
    got = model.evaluate(sc, obspos=np.array([0.87, -0.53, 0.001]) * u.AU,
                         earthpos=earth * u.AU).value
    assert float(np.ravel(got)[0]) == pytest.approx(20.37750965, rel=0, abs=1e-8)


def test_grid_number_density_public_api(gpu):
    """grid_number_density (zodipy/number_density.py:482-536): np.meshgrid's Cartesian
    indexing gives (ncomps, ny, nx, nz); element [c, i, j, k] is component c at
    (x[j], y[i], z[k]) with the Earth position injected into the feature. Axis sizes
    differ, so a transposed grid fails; moving the Earth changes the feature only."""
    from zodipy_amd import grid_number_density
    from zodipy_amd.coords import Time

    x = np.linspace(-2.0, 2.5, 7)
    y = np.linspace(-1.5, 1.0, 5)
    z = np.array([-0.1, 0.0, 0.2])
    earth = np.array([0.3, -0.95, 0.0])
    got = grid_number_density(x * u.AU, y * u.AU, z * u.AU, Time("2020-01-01"), model="dirbe",
                              earthpos=earth * u.AU)
    assert got.shape == (6, 5, 7, 3)
    pts = np.array([[x[j], y[i], z[k]] for i in range(5) for j in range(7) for k in range(3)]).T
    want = OracleModel(25 * u.micron).number_density(pts, earth).reshape(6, 5, 7, 3)
    assert_parity(got, want)
    earth2 = np.array([-0.95, 0.3, 0.0])
    moved = grid_number_density(x * u.AU, y * u.AU, z * u.AU, Time("2020-01-01"),
                                model="dirbe", earthpos=earth2 * u.AU)
    np.testing.assert_array_equal(moved[:5], got[:5])
    assert not np.array_equal(moved[5], got[5])
    assert_parity(moved, OracleModel(25 * u.micron).number_density(pts, earth2).reshape(
        6, 5, 7, 3))


@pytest.mark.parametrize("lanes", [1, 2])
def test_registry_variant_with_unshared_band_cutoff(gpu, lanes):
    """The DIRBE bands 1 and 3 share delta_r = 1.5, so the registry kernel with one
    lane per ray forms their radial cut-off once. A parameter update that moves band
    3's delta_r (as a fit does) must leave that sharing out: parity with the oracle
    of the updated model, and the output changes only in band 3."""
    import dataclasses

    from zodipy_amd.component import ComponentLabel
    from zodipy_amd.zodiacal_light_model import model_registry

    v = random_dirs(2000, seed=41)
    obs = np.array([0.30, 0.95, 0.02])
    model = Model(12 * u.micron)
    dm = model.device_model
    launch = Launch(block=256, lanes_per_ray=lanes, blocks_per_cu=0, variant=1)
    before = dm.forward(v, obs, obs, launch=launch).cpu().numpy()
    params = model.get_parameters()
    params["comps"]["band3"]["delta_r"] = 1.6
    model.update_parameters(params)
    got = model.device_model.forward(v, obs, obs, launch=launch).cpu().numpy()
    ipd = model_registry.get_model("dirbe")
    comps = dict(ipd.comps)
    comps[ComponentLabel.BAND3] = dataclasses.replace(comps[ComponentLabel.BAND3], delta_r=1.6)
    want = OracleModel(12 * u.micron, model=dataclasses.replace(ipd, comps=comps)
                       ).evaluate_vectors(v, obs, obs)
    assert_parity(got, want)
    band3 = list(ipd.comps).index(ComponentLabel.BAND3)
    others = [c for c in range(got.shape[0]) if c != band3]
    assert np.array_equal(got[others], before[others])
    assert not np.array_equal(got[band3], before[band3])


LANE_CASES = [("dirbe", 25.0, 50), ("dirbe", 1.25, 50), ("dirbe", 12.0, 7), ("planck18", None, 50),
              ("rrm-experimental", 25.0, 50), ("cloud_ring", 1.25, 50)]


@pytest.mark.parametrize("name,lam,degree", LANE_CASES,
                         ids=["dirbe25", "dirbe1.25_scatter", "dirbe12_7nodes", "planck18_857",
                              "rrm25", "cloud_ring_generic"])
def test_lane_splits_give_identical_bits(gpu, name, lam, degree, cloud_ring_model):
    """Every lane split and block size of one kernel variant gives the same bits: the
    L lanes of a ray evaluate L nodes at once and every lane adds their terms in node
    order (50 nodes leave a partial last step for 4 and 8 lanes, 7 nodes one shorter
    than 8 lanes). The default launch, which splits small launches, equals one lane per
    ray; so do per-ray observers. For the blocks where no component scatters (DIRBE
    25 um, planck18) the one-lane registry kernel is the one compiled without
    scattering code and the split-lane kernels keep it, so the same bits pin that
    variant too."""
    x = 857 * u.GHz if lam is None else lam * u.micron
    model = Model(x, name=cloud_ring_model if name == "cloud_ring" else name,
                  gauss_quad_degree=degree)
    dm = model.device_model
    rng = np.random.default_rng(44)
    counts = rng.integers(1, 300, 5)
    seg = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    v = random_dirs(int(seg[-1]), seed=45)
    ph = rng.uniform(0, 2 * np.pi, counts.size)
    obs = np.stack([np.cos(ph), np.sin(ph), 0.05 * np.ones_like(ph)])
    for variant in (1, 0):
        ref = dm.forward(v, obs, obs, seg_offsets=seg, launch=Launch(256, 1, 0, variant))
        for lanes in (1, 2, 4, 8):
            for block in (64, 256) if lanes > 1 else (64, 128, 512):
                got = dm.forward(v, obs, obs, seg_offsets=seg,
                                 launch=Launch(block, lanes, 0, variant))
                assert torch.equal(got, ref), (variant, lanes, block)
        if variant == 1:
            assert torch.equal(dm.forward(v, obs, obs, seg_offsets=seg), ref)
            assert_parity(ref.cpu().numpy(),
                          evaluate_segments(OracleModel(x, model=model._ipd_model,
                                                        gauss_quad_degree=degree),
                                            v, seg, obs, obs))
    per_ray = np.repeat(obs, counts, axis=1)
    one = dm.forward(v, per_ray, per_ray, per_ray=True, launch=Launch(256, 1, 0, 1))
    assert torch.equal(dm.forward(v, per_ray, per_ray, per_ray=True), one)
