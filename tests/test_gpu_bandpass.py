"""GPU tests of the Planck-table gradient (zodi_table_vjp), of the spectral
gradient (zodi_spectral_vjp) and of bandpass-weight fits.

The table gradient is checked against central differences of chi^2 with one
table entry B_k = B(T_k) of the device block moved at a time, and against the
weighted (VJP) modes; the spectral gradient's entries against zodi_adjoint's
for the same parameters, on every kernel variant (registry layouts with and
without the scattering sums, the kind switch); the bandpass fit's forward
against the model built from the same weights by the host packer, and its
gradient against central differences over each weight. HEALPix nside 8 pixel
centres, observer (1, 0, 0).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from zodipy_amd import Model, healpix
from zodipy_amd import units as u
from zodipy_amd.fitting import BandpassFit, LineOfSightProblem, chi2_bandpass

pytestmark = pytest.mark.gpu

OBS = np.array([1.0, 0.0, 0.0])


def _problem(model):
    v = healpix.pix2vec(8)
    return LineOfSightProblem(model, v, OBS, OBS)


def _data(prob, theta, seed):
    I0 = prob.forward(theta)
    rng = np.random.default_rng(seed)
    noise = torch.as_tensor(rng.normal(size=prob.n_rays), device="cuda")
    sigma = 0.01 * float(I0.abs().mean())
    return I0 * 1.01 + sigma * noise, 1.0 / sigma**2


def _chi2(prob, blocks, data, inv_var):
    I = prob.forward(None, blocks=blocks)
    return float(((I - data) ** 2).sum()) * inv_var


@pytest.mark.parametrize("name,lam", [("dirbe", 25.0), ("dirbe", 1.25), ("rrm-experimental", 25.0)],
                         ids=["dirbe25", "dirbe1.25_scatter", "rrm25"])
def test_table_grad_matches_central_fd(gpu, name, lam):
    model = Model(lam * u.micron, name=name)
    prob = _problem(model)
    theta = prob.space.theta0
    data, iv = _data(prob, theta, 3)
    c2, gB = prob.chi2_table_grad(theta, data, iv)
    host, dev, _ = prob._model(theta)
    assert float(c2) == pytest.approx(_chi2(prob, (host, dev), data, iv), rel=1e-12)
    g = gB.cpu().numpy()
    n = int(np.asarray(host["n_temps"]).reshape(-1)[0])
    assert np.all(g[n:] == 0.0)
    B = np.asarray(host["tab_B"]).reshape(-1)[:n].copy()
    sens = np.abs(g[:n] * B)
    pick = sorted(set(np.argsort(sens)[-6:].tolist()))
    for k in pick:
        # chi^2 is quadratic in each B_k (I is linear in it), so the central
        # difference is exact for any step; a large one keeps rounding out of it
        h = 1e-2 * abs(B[k])
        vals = []
        for sgn in (1.0, -1.0):
            hp = host.copy()
            hp["tab_B"][..., k] = B[k] + sgn * h
            dp = torch.from_numpy(hp.view(np.uint8).copy()).cuda()
            vals.append(_chi2(prob, (hp, dp), data, iv))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-6, abs=1e-9 * sens.max() / abs(B[k])), k


def test_table_vjp_modes_agree(gpu):
    """The VJP modes with the chi^2 weights 2 (I - d)/sigma^2 give the chi^2 mode's
    gradient; per-component weights summed over components give the summed mode's."""
    model = Model(25 * u.micron)
    prob = _problem(model)
    theta = prob.space.theta0
    data, iv = _data(prob, theta, 4)
    _, gB = prob.chi2_table_grad(theta, data, iv)
    I = prob.forward(theta)
    w = 2.0 * (I - data) * iv
    torch.testing.assert_close(prob.table_vjp(theta, w), gB, rtol=1e-11, atol=1e-11 * float(gB.abs().max()))
    wc = w.expand(prob.n_comps, -1).contiguous()
    torch.testing.assert_close(prob.table_vjp(theta, wc, return_comps=True), gB, rtol=1e-11,
                               atol=1e-11 * float(gB.abs().max()))
    # one call repeats bit for bit
    assert torch.equal(prob.table_vjp(theta, w), prob.table_vjp(theta, w))


CASES = [
    ("dirbe", np.linspace(20.0, 30.0, 8), "micron", False),
    ("dirbe", np.linspace(10.0, 15.0, 6) * 1e3, "GHz", False),
    ("planck18", np.linspace(771.3, 942.7, 12), "GHz", True),
    ("rrm-experimental", np.linspace(20.0, 30.0, 8), "micron", False),
]
IDS = ["dirbe_um", "dirbe_ghz_on_um", "planck18_857_extrapolated", "rrm_um"]


def _bandpass_model(name, xs, unit, extrap, w):
    x = xs * (u.micron if unit == "micron" else u.GHz)
    return x, Model(x, weights=w, name=name, extrapolate=extrap)


@pytest.mark.parametrize("name,xs,unit,extrap", CASES, ids=IDS)
def test_bandpass_fit_forward_matches_model(gpu, name, xs, unit, extrap):
    """Effective values and Planck table from the fit's device set-up at new
    weights equal the model the host packer builds from those weights."""
    rng = np.random.default_rng(7)
    w0 = np.exp(-0.5 * ((xs - xs.mean()) / (0.2 * np.ptp(xs))) ** 2)
    x, model = _bandpass_model(name, xs, unit, extrap, w0)
    fit = BandpassFit(_problem(model))
    w1 = w0 * (1.0 + 0.3 * rng.uniform(-1, 1, xs.size))
    got = fit.forward(w1, return_comps=True)
    _, model1 = _bandpass_model(name, xs, unit, extrap, w1)
    prob1 = _problem(model1)
    want = prob1.forward(prob1.space.theta0, return_comps=True)
    torch.testing.assert_close(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("name,xs,unit,extrap", CASES, ids=IDS)
def test_bandpass_fit_gradient_matches_central_fd(gpu, name, xs, unit, extrap):
    rng = np.random.default_rng(8)
    w0 = np.exp(-0.5 * ((xs - xs.mean()) / (0.2 * np.ptp(xs))) ** 2)
    x, model = _bandpass_model(name, xs, unit, extrap, w0)
    prob = _problem(model)
    fit = BandpassFit(prob)
    w = w0 * (1.0 + 0.2 * rng.uniform(-1, 1, xs.size))
    data, iv = _data(prob, prob.space.theta0, 9)
    c2, g = fit.chi2_and_grad(w, data, iv)
    g = g.cpu().numpy()
    assert float(c2) == pytest.approx(_chi2(prob, (fit._host, fit._dev), data, iv), rel=1e-12)
    fd = np.empty_like(w)
    for i in range(w.size):
        h = 1e-6 * w[i]
        vals = []
        for sgn in (1.0, -1.0):
            wp = w.copy()
            wp[i] += sgn * h
            fit.set_weights(wp)
            vals.append(_chi2(prob, (fit._host, fit._dev), data, iv))
        fd[i] = (vals[0] - vals[1]) / (2 * h)
    scale = np.abs(fd * w).max()
    np.testing.assert_allclose(g * w, fd * w, rtol=1e-6, atol=1e-7 * scale)
    # the normalisation removes the scale: the gradient is orthogonal to w
    assert abs(float(np.dot(g, w))) <= 1e-10 * float(np.abs(g * w).sum())
    # the autograd wrapper returns the same gradient
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    chi2_bandpass(wt, fit, data, iv).backward()
    np.testing.assert_allclose(wt.grad.cpu().numpy(), g, rtol=1e-12, atol=0)


SPECTRAL_PREFIXES = ("emissivity.", "albedo.")
SPECTRAL_SHARED = ("solar_irradiance", "C1", "C2", "C3", "calibration")


def _spectral_mask(prob):
    return np.array([n.startswith(SPECTRAL_PREFIXES) or n in SPECTRAL_SHARED
                     for n in prob.space.names])


def _generic_rrm():
    """A registered RRM model of the fan and the comet only: no RRM registry layout."""
    from zodipy_amd import component_params as cp
    from zodipy_amd import source_params as sp
    from zodipy_amd.component import ComponentLabel as L
    from zodipy_amd.zodiacal_light_model import RRM, model_registry

    name = "test-rrm-fan-comet"
    try:
        model_registry.get_model(name)
    except ValueError:
        keep = [L.FAN, L.COMET]
        model_registry.register_model(name, RRM(
            comps={k: cp.RRM[k] for k in keep}, spectrum=sp.SPECTRUM_IRAS,
            T_0={k: sp.T_0_RRM[k] for k in keep}, delta={k: sp.DELTA_RMM[k] for k in keep},
            calibration=sp.CALIBRATION_RRM))
    return name


SPECTRAL_CASES = ["dirbe25", "dirbe1.25_scatter", "planck18_style_857", "planck18_857", "rrm25",
                  "cloud_ring1.25_generic", "rrm_fan_comet_generic"]


def _spectral_model(case, planck18_style, cloud_ring_model):
    if case == "dirbe25":  # DIRBE layout; F != 0, so the scattering sums run
        return Model(25 * u.micron)
    if case == "dirbe1.25_scatter":  # scattering components, angle-dependent phase
        return Model(1.25 * u.micron)
    if case == "planck18_style_857":  # DIRBE layout without scattering
        return Model(857 * u.GHz, name=planck18_style)
    if case == "planck18_857":  # four-component layout without scattering
        return Model(857 * u.GHz, name="planck18")
    if case == "rrm25":
        return Model(25 * u.micron, name="rrm-experimental")
    if case == "cloud_ring1.25_generic":
        return Model(1.25 * u.micron, name=cloud_ring_model)
    return Model(25 * u.micron, name=_generic_rrm())


def _assert_spectral_entries(got, want, mask, what):
    got = got.cpu().numpy()
    want = want.cpu().numpy()
    assert np.all(got[~mask] == 0.0), what
    scale = float(np.abs(want[mask]).max())
    assert scale > 0.0, what
    np.testing.assert_allclose(got[mask], want[mask], rtol=1e-10, atol=1e-12 * scale,
                               err_msg=what)


@pytest.mark.parametrize("case", SPECTRAL_CASES)
def test_spectral_vjp_matches_adjoint(gpu, case, planck18_style, cloud_ring_model):
    """zodi_spectral_vjp's parameter entries equal zodi_adjoint's for the effective
    spectral parameters (the same partials, summed in another order) and are 0
    elsewhere; its table part equals zodi_table_vjp's; in the weighted modes and
    in chi^2 mode (against the Jacobian path), with the model values it writes."""
    prob = _problem(_spectral_model(case, planck18_style, cloud_ring_model))
    theta = prob.space.theta0
    mask = _spectral_mask(prob)
    rng = np.random.default_rng(11)
    w = torch.as_tensor(rng.normal(size=prob.n_rays), device="cuda")
    g_th, g_tab = prob.spectral_vjp(theta, w)
    _assert_spectral_entries(g_th, prob.vjp(theta, w), mask, f"{case}: vjp")
    gt = prob.table_vjp(theta, w)
    torch.testing.assert_close(g_tab, gt, rtol=1e-12, atol=1e-12 * float(gt.abs().max()))

    wc = torch.as_tensor(rng.normal(size=(prob.n_comps, prob.n_rays)), device="cuda")
    g_th, g_tab = prob.spectral_vjp(theta, wc, return_comps=True)
    _assert_spectral_entries(g_th, prob.vjp(theta, wc, return_comps=True), mask,
                             f"{case}: vjp per component")
    gt = prob.table_vjp(theta, wc, return_comps=True)
    torch.testing.assert_close(g_tab, gt, rtol=1e-12, atol=1e-12 * float(gt.abs().max()))

    data, iv = _data(prob, theta, 12)
    c2, g_th, g_tab, model = prob.chi2_spectral_grad(theta, data, iv, return_model=True)
    c2_j, g_j = prob.chi2_and_grad(theta, data, iv)
    assert float(c2) == pytest.approx(float(c2_j), rel=1e-12)
    _assert_spectral_entries(g_th, g_j, mask, f"{case}: chi2")
    c2_t, gt = prob.chi2_table_grad(theta, data, iv)
    assert float(c2) == float(c2_t)
    torch.testing.assert_close(g_tab, gt, rtol=1e-12, atol=1e-12 * float(gt.abs().max()))
    torch.testing.assert_close(model, prob.forward(theta), rtol=1e-14, atol=0)


def test_spectral_vjp_empty_and_launches(gpu):
    """No rays: zeros; forced lane splits agree with the default launch."""
    from zodipy_amd._lib import Launch

    model = Model(25 * u.micron)
    prob = _problem(model)
    theta = prob.space.theta0
    w = torch.as_tensor(np.random.default_rng(13).normal(size=prob.n_rays), device="cuda")
    g0, t0 = prob.spectral_vjp(theta, w)
    for split in (2, 4):
        prob.launch = Launch(0, split, 0, 0)
        try:
            g, t = prob.spectral_vjp(theta, w)
        finally:
            prob.launch = None
        torch.testing.assert_close(g, g0, rtol=1e-11, atol=1e-13 * float(g0.abs().max()))
        torch.testing.assert_close(t, t0, rtol=1e-11, atol=1e-13 * float(t0.abs().max()))
    empty = LineOfSightProblem(model, np.zeros((3, 0)), OBS, OBS)
    none = torch.zeros(0, dtype=torch.float64, device="cuda")
    c2, g, t = empty.chi2_spectral_grad(theta, none)
    assert float(c2) == 0.0 and torch.all(g == 0) and torch.all(t == 0)


def test_spectral_chi2_per_ray_inverse_variance(gpu):
    """chi^2 mode with a per-ray 1/sigma^2: the spectral entries and chi^2 equal
    zodi_adjoint's Jacobian path, the table part zodi_table_vjp's."""
    prob = _problem(Model(1.25 * u.micron))
    theta = prob.space.theta0
    data, iv = _data(prob, theta, 14)
    rng = np.random.default_rng(15)
    ivr = torch.as_tensor(iv * rng.uniform(0.5, 2.0, prob.n_rays), device="cuda")
    c2, g_th, g_tab = prob.chi2_spectral_grad(theta, data, ivr)
    c2_j, g_j = prob.chi2_and_grad(theta, data, ivr)
    assert float(c2) == pytest.approx(float(c2_j), rel=1e-12)
    _assert_spectral_entries(g_th, g_j, _spectral_mask(prob), "per-ray inverse variance")
    c2_t, gt = prob.chi2_table_grad(theta, data, ivr)
    assert float(c2) == float(c2_t)
    torch.testing.assert_close(g_tab, gt, rtol=1e-12, atol=1e-12 * float(gt.abs().max()))



def test_bandpass_fit_on_the_dirbe_25um_band(gpu):
    """A fit of the DIRBE 25 um band's own 82 weights (the reference's test data):
    the gradient against central differences on eight weights (1e-6), orthogonal
    to w."""
    import json
    from pathlib import Path

    d = json.loads((Path(__file__).parent / "golden" / "dirbe_25um_bandpass.json").read_text())
    x = np.asarray(d["wavelength_micron"]) * u.micron
    w0 = np.asarray(d["weight"], dtype=np.float64)
    prob = _problem(Model(x, weights=w0))
    fit = BandpassFit(prob)
    data, iv = _data(prob, prob.space.theta0, 16)
    w = w0 * (1.0 + 0.1 * np.random.default_rng(17).uniform(-1, 1, w0.size))
    c2, g = fit.chi2_and_grad(w, data, iv)
    g = g.cpu().numpy()
    assert float(c2) == pytest.approx(_chi2(prob, (fit._host, fit._dev), data, iv), rel=1e-12)
    assert abs(float(np.dot(g, w))) <= 1e-10 * float(np.abs(g * w).sum())
    pick = np.argsort(np.abs(g * w))[-8:]
    for i in pick:
        h = 1e-6 * w[i]
        vals = []
        for sgn in (1.0, -1.0):
            wp = w.copy()
            wp[i] += sgn * h
            fit.set_weights(wp)
            vals.append(_chi2(prob, (fit._host, fit._dev), data, iv))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6), i