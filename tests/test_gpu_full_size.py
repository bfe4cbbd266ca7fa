"""Forward parity at the full workload sizes, through properties that do not
depend on size, plus oracle parity on a seeded sample of the rays.

* Batched time-ordered workload: DIRBE 12 um, 2000 observers on a 1 AU circular
  orbit at seeded phases, seeded pointing counts uniform in [1000, 50000]
  (about 51 M rays, directions uniform on the sphere), one launch.
* Full sky: DIRBE model, HEALPix nside 256 (786 432 pixel centres), Gaussian
  bandpass centred at 25 um with sigma 2.5 um at 50 wavelengths (normalised
  weights), observer on a circular 1 AU orbit at phase 0.3 rad, per-component
  output.
* Gradient: DIRBE 25 um, nside 128 (196 608 pixel centres), observer (1, 0, 0)
  AU; target = the model with every parameter perturbed by seeded N(0, 1 %)
  relative noise plus N(0, 0.01 MJy/sr); chi^2 with sigma = 0.01 and its
  gradient with respect to all 78 parameters against central differences.
* Batched gradient: six-component Planck-2018-style model at 857 GHz with a
  30-frequency top-hat bandpass of +-10 %, 1000 observers with seeded
  5000-40000 pointings each (about 22 M rays): chi^2 and its gradient in one
  call, against central differences of chi^2 formed from the forward kernel
  for every one of its 78 parameters (1e-6 relative).

Properties: every value finite and non-negative (the cloud's positive); the
summed output equals the per-component output added in component order, bit
for bit; a launch over one observer's rays alone reproduces that observer's
slice of the batch bit for bit (batching changes nothing); sampled rays agree
with the oracle within 1e-10.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle.zodi_oracle import OracleModel
from zodipy_amd import Model, healpix
from zodipy_amd import units as u
from zodipy_amd.runtime import OUT_COMPS, OUT_SUM

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel_err(got, want):
    """Largest relative error; exact zeros of the oracle must be exact zeros."""
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape
    assert np.all(np.isfinite(want))
    zero = want == 0
    assert np.all(got[zero] == 0), "exact zeros differ"
    return float((np.abs(got[~zero] - want[~zero]) / np.abs(want[~zero])).max(initial=0.0))


def directions_on_device(n: int, seed: int) -> torch.Tensor:
    """(3, n) unit vectors uniform on the sphere, drawn on the device."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    z = torch.rand(n, generator=g, dtype=torch.float64, device="cuda") * 2 - 1
    phi = torch.rand(n, generator=g, dtype=torch.float64, device="cuda") * (2 * np.pi)
    s = torch.sqrt(torch.clamp(1 - z * z, min=0.0))
    return torch.stack([s * torch.cos(phi), s * torch.sin(phi), z]).contiguous()


def sum_in_component_order(comps: torch.Tensor) -> torch.Tensor:
    total = comps[0].clone()
    for c in range(1, comps.shape[0]):
        total += comps[c]
    return total


def test_batched_2000_observers_full_size(gpu):
    rng = np.random.default_rng(2024)
    n_obs = 2000
    counts = rng.integers(1000, 50000, n_obs, endpoint=True)
    seg = np.zeros(n_obs + 1, dtype=np.int64)
    np.cumsum(counts, out=seg[1:])
    n = int(seg[-1])
    phases = rng.uniform(0.0, 2.0 * np.pi, n_obs)
    obs = np.stack([np.cos(phases), np.sin(phases), np.zeros(n_obs)])
    print(f"batched full size: {n} rays, {n_obs} observers", flush=True)
    v = directions_on_device(n, 7)
    model = Model(12 * u.micron)
    dm = model.device_model
    comps = dm.forward(v, obs, obs, seg_offsets=seg, out_mode=OUT_COMPS)
    total = dm.forward(v, obs, obs, seg_offsets=seg, out_mode=OUT_SUM)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(total).all()) and bool((total > 0).all())
    assert torch.equal(total, sum_in_component_order(comps))
    del comps

    picks = sorted({0, n_obs - 1, int(np.argmax(counts)), int(np.argmin(counts)),
                    *rng.choice(n_obs, 8, replace=False).tolist()})
    om = OracleModel(12 * u.micron)
    worst = 0.0
    for s in picks:
        a, b = int(seg[s]), int(seg[s + 1])
        alone = dm.forward(v[:, a:b].contiguous(), obs[:, s], obs[:, s], out_mode=OUT_SUM)
        assert torch.equal(alone, total[a:b]), f"observer {s}: alone differs from the batch"
        idx = np.sort(rng.choice(b - a, 200, replace=False)) + a
        vs = v[:, torch.as_tensor(idx, device="cuda")].cpu().numpy()
        want = om.evaluate_vectors(vs, obs[:, s], obs[:, s]).sum(axis=0)
        worst = max(worst, rel_err(total[torch.as_tensor(idx, device="cuda")].cpu().numpy(),
                                   want))
    print(f"sampled oracle parity over {len(picks)} observers: max rel err {worst:.2e}")
    assert worst <= RTOL


def test_full_sky_nside256_gaussian_bandpass(gpu):
    x = np.linspace(17.5, 32.5, 50) * u.micron
    w = np.exp(-0.5 * ((x.value - 25.0) / 2.5) ** 2)
    w = w / w.sum()
    obs = np.array([np.cos(0.3), np.sin(0.3), 0.0])
    v = healpix.pix2vec(256)
    assert v.shape == (3, 786432)
    model = Model(x, weights=w)
    dm = model.device_model
    comps = dm.forward(v, obs, obs, out_mode=OUT_COMPS)
    total = dm.forward(v, obs, obs, out_mode=OUT_SUM)
    c = comps.cpu().numpy()
    assert c.shape == (6, 786432)
    assert np.all(np.isfinite(c)) and np.all(c >= 0) and np.all(c[0] > 0)
    assert torch.equal(total, sum_in_component_order(comps))
    got = model.evaluate_vectors(v, obspos=obs, earthpos=obs, return_comps=True).value
    assert np.array_equal(got, c)

    rng = np.random.default_rng(256)
    idx = np.sort(rng.choice(v.shape[1], 3000, replace=False))
    want = OracleModel(x, weights=w).evaluate_vectors(v[:, idx], obs, obs)
    err = rel_err(c[:, idx], want)
    print(f"sampled oracle parity on 3000 of 786432 pixels: max rel err {err:.2e}")
    assert err <= RTOL


def test_batched_857ghz_gradient_full_size(gpu, planck18_style):
    from zodipy_amd.fitting import LineOfSightProblem

    name = planck18_style
    rng = np.random.default_rng(857)
    n_obs = 1000
    counts = rng.integers(5000, 40000, n_obs, endpoint=True)
    seg = np.zeros(n_obs + 1, dtype=np.int64)
    np.cumsum(counts, out=seg[1:])
    phases = rng.uniform(0.0, 2.0 * np.pi, n_obs)
    obs = np.stack([np.cos(phases), np.sin(phases), np.zeros(n_obs)])
    n = int(seg[-1])
    print(f"batched gradient full size: {n} rays, {n_obs} observers", flush=True)
    v = directions_on_device(n, 857)
    nu = np.linspace(771.3, 942.7, 30) * u.GHz
    model = Model(nu, weights=np.ones(30), name=name, extrapolate=True)
    prob = LineOfSightProblem(model, v, obs, obs, seg_offsets=seg)
    theta = prob.space.theta0
    assert prob.n_comps == 6 and prob.n_params == theta.size
    data = prob.forward(theta) * 1.01
    inv_var = 1.0 / 0.01**2
    chi2, grad = prob.chi2_and_grad(theta, data, inv_var)
    g = grad.cpu().numpy()

    def chi2_fwd(th):
        return float((((prob.forward(th) - data) ** 2) * inv_var).sum())

    assert abs(float(chi2) - chi2_fwd(theta)) <= 1e-12 * float(chi2)
    names = list(prob.space.names)
    fd = np.empty_like(theta)
    for i, p in enumerate(names):
        # B(T) is linear between the Planck table's temperatures, as in the
        # reference. A step in T_0 or delta moves every node's temperature, and
        # those that cross a table temperature change slope inside the step: at
        # 1.1e9 node samples a 1e-6 step leaves FD off by ~2e-6 for delta, a
        # 1e-8 step by ~2e-8. The other parameters keep the usual 1e-6 step.
        rel_step = 1e-8 if p in ("T_0", "delta") else 1e-6
        h = rel_step * max(abs(theta[i]), 1e-3)
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (chi2_fwd(tp) - chi2_fwd(tm)) / (tp[i] - tm[i])
    worst = sensitivity_check(g, fd, theta, 1e-6, 1e-6, "batched chi2 gradient vs FD")
    print(f"{len(names)} parameters: worst relative error against FD {worst:.2e}", flush=True)


def sensitivity_check(got, want, theta, rtol, floor, what):
    """Relative error per entry where |want| |theta| is within 1e-7 of the largest;
    weaker entries are held to `floor` times that largest sensitivity."""
    got, want = np.asarray(got), np.asarray(want)
    s = np.maximum(np.abs(theta), 1e-3)
    sens = np.abs(want) * s
    big = sens.max()
    err = np.abs(got - want) * s
    strong = sens > 1e-7 * big
    rel = np.where(strong, err / np.maximum(sens, 1e-300), 0.0)
    assert rel.max() <= rtol, f"{what}: max relative error {rel.max():.3e} (entry {rel.argmax()})"
    assert (err[~strong] / big).max(initial=0.0) <= floor, f"{what}: weak-entry error"
    return float(rel.max())


def test_dirbe25_nside128_gradient_every_parameter(gpu):
    from zodipy_amd.fitting import LineOfSightProblem

    model = Model(25 * u.micron)
    v = healpix.pix2vec(128)
    obs = np.array([1.0, 0.0, 0.0])
    prob = LineOfSightProblem(model, v, obs, obs)
    theta = prob.space.theta0
    rng = np.random.default_rng(128)
    theta_true = theta * (1.0 + 0.01 * rng.normal(size=theta.size))
    data = prob.forward(theta_true) + torch.as_tensor(0.01 * rng.normal(size=v.shape[1]),
                                                      device="cuda")
    inv_var = 1.0 / 0.01**2
    chi2, grad = prob.chi2_and_grad(theta, data, inv_var)
    names = list(prob.space.names)

    def chi2_fwd(th):
        return float((((prob.forward(th) - data) ** 2) * inv_var).sum())

    assert abs(float(chi2) - chi2_fwd(theta)) <= 1e-12 * float(chi2)
    fd = np.empty_like(theta)
    for i, p in enumerate(names):
        # 1e-8 for the temperature law: B(T) is linear between table temperatures
        h = (1e-8 if p in ("T_0", "delta") else 1e-6) * max(abs(theta[i]), 1e-3)
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (chi2_fwd(tp) - chi2_fwd(tm)) / (tp[i] - tm[i])
    worst = sensitivity_check(grad.cpu().numpy(), fd, theta, 1e-6, 1e-6, "chi2 gradient vs FD")
    print(f"{len(names)} parameters: worst relative error against FD {worst:.2e}")
