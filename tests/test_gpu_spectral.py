"""Spectral set-up on the device (zodi_spectral_setup) and fits of the tables.

* The device block after set-up against the host packer's block for the same
  tables: effective spectral slots and derived slots to 1e-14, the bandpass
  Planck table to 1e-13 (its own sums and expm1), every other byte unchanged.
* The emission with perturbed tables set on the device against the host
  packer's block for a model holding the same perturbed tables (1e-12).
* d chi^2 / d(table entry) from the adjoint and the weights against central
  finite differences of chi^2 through the device set-up (1e-6, with the
  sensitivity floor of the adjoint tests); entries no weight reaches get
  exactly zero. I is linear in emissivity, albedo and solar irradiance, so
  chi^2 is quadratic in them and a central difference is exact up to rounding:
  their step is h = 1e-4 max(|v|, 1e-3), and the residuals are 0.1 % of I, which
  keeps the rounding term eps chi^2 / h far below the tolerance. chi^2 is
  strongly curved in the phase coefficients C1..C3 (the normalisation's
  denominator is small at the DIRBE values), so theirs is h = 1e-6 max(|v|, 1e-3).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from zodipy_amd import Model, healpix
from zodipy_amd import units as u
from zodipy_amd.fitting import LineOfSightProblem, SpectralTableFit, chi2_tables
from zodipy_amd.packing import (
    P_ALBEDO,
    P_EMISSIVITY,
    P_ONE_MINUS_ALBEDO,
    P_SCALE,
    pack_model,
)
from zodipy_amd.spectral import unpack_spectral

pytestmark = pytest.mark.gpu

OBS = np.array([0.95533649, 0.29552021, 0.01])
EARTH = np.array([0.95533649, 0.29552021, 0.0])

CASES = {
    "dirbe_25um": dict(x=25 * u.micron),
    "dirbe_1.25um_scatter": dict(x=1.25 * u.micron),
    "dirbe_bandpass_um": dict(x=np.linspace(17.5, 32.5, 50) * u.micron,
                              weights=np.exp(-0.5 * ((np.linspace(17.5, 32.5, 50) - 25) / 2.5)
                                             ** 2)),
    "planck18_tophat": dict(x=np.linspace(771.3, 942.7, 30) * u.GHz, weights=np.ones(30),
                            name="planck18", extrapolate=True),
    "rrm_1.25um": dict(x=1.25 * u.micron, name="rrm-experimental", extrapolate=True),
}


def _fit(case, nside=4):
    kw = dict(CASES[case])
    model = Model(kw.pop("x"), **kw)
    v = healpix.pix2vec(nside)
    prob = LineOfSightProblem(model, v, OBS, EARTH)
    return model, prob, SpectralTableFit(prob)


def _model_with_tables(model, fit, values):
    """The registry model with its spectral tables replaced by ``values`` (host)."""
    ipd = model._ipd_model
    new = {}
    k = 0
    for t in fit.tables:
        vals = list(np.asarray(values[k:k + fit.n_grid], dtype=np.float64))
        k += fit.n_grid
        key, _, lab = t.name.partition(".")
        if lab:
            field = {"emissivity": "emissivities", "albedo": "albedos"}[key]
            d = dict(new.get(field, getattr(ipd, field)))
            label = next(c for c in ipd.comps if c.value == lab)
            d[label] = vals
            new[field] = d
        else:
            new[key] = vals
    return dataclasses.replace(ipd, **new)


@pytest.mark.parametrize("case", list(CASES))
def test_setup_reproduces_host_block(gpu, case):
    model, prob, fit = _fit(case)
    fit.set_tables(fit.values0, planck=True)
    got = fit.device_block()[0]
    want = fit._host[0]
    nc = int(want["n_comps"])
    for slot in (P_EMISSIVITY, P_ALBEDO, P_ONE_MINUS_ALBEDO, P_SCALE):
        np.testing.assert_allclose(got["comp"][:nc, slot], want["comp"][:nc, slot], rtol=1e-14,
                                   atol=0)
    np.testing.assert_allclose(got["shared"], want["shared"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(got["tab_B"], want["tab_B"], rtol=1e-13, atol=0)
    scale = np.abs(want["tab_slope"]).max()
    np.testing.assert_allclose(got["tab_slope"], want["tab_slope"], rtol=1e-10,
                               atol=1e-15 * scale)
    # nothing else moved
    g, w = got.copy(), want.copy()
    for a in (g, w):
        a["comp"][:, [P_EMISSIVITY, P_ALBEDO, P_ONE_MINUS_ALBEDO, P_SCALE]] = 0
        a["shared"] = 0
        a["tab_B"] = 0
        a["tab_slope"] = 0
    assert g.tobytes() == w.tobytes()


@pytest.mark.parametrize("case", ["dirbe_25um", "dirbe_bandpass_um", "planck18_tophat",
                                  "rrm_1.25um"])
def test_forward_with_device_tables_matches_host_pack(gpu, case):
    model, prob, fit = _fit(case)
    rng = np.random.default_rng(1)
    values = fit.values0 * (1 + 0.05 * rng.normal(size=fit.n_params))
    # move some albedos off zero and set phase coefficients, so scattering is exercised
    for k, name in enumerate(fit.names):
        if name.startswith("albedo.") and values[k] == 0:
            values[k] = 0.05 + 0.1 * rng.uniform()
        if name.startswith(("C1[", "C2[", "C3[")) and values[k] == 0:
            values[k] = {"C1": -0.9, "C2": 0.5, "C3": -0.6}[name[:2]] * (1 + 0.1 * rng.normal())
        if name.startswith("solar_irradiance") and values[k] == 0:
            values[k] = 1e5
    got = fit.forward(values, return_comps=True)
    ipd = _model_with_tables(model, fit, values)
    spec = unpack_spectral(ipd, model._x, model._normalized_weights, model._extrapolate)
    packed = pack_model(ipd, spec, model._b_nu_table, model._gl_x, model._gl_w)
    host = np.ascontiguousarray(packed.block)
    dev = torch.from_numpy(host.view(np.uint8).copy()).cuda()
    want = prob.forward(None, return_comps=True, blocks=(host, dev))
    torch.testing.assert_close(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("case", ["dirbe_25um", "dirbe_1.25um_scatter", "planck18_tophat"])
def test_table_gradient_matches_fd(gpu, case):
    model, prob, fit = _fit(case, nside=2)
    rng = np.random.default_rng(2)
    values = fit.values0.copy()
    I0 = fit.forward(values)
    data = I0 * torch.as_tensor(1 + 0.001 * rng.normal(size=I0.numel()), device="cuda")
    c2, g = fit.chi2_and_grad(values, data, 2.0)
    g = g.cpu().numpy()
    reached = (fit._A.cpu().numpy() != 0).reshape(-1)
    assert np.all(g[~reached] == 0)

    def chi2_at(v):
        r = fit.forward(v) - data
        return float((r * r).sum()) * 2.0

    fd = np.zeros_like(values)
    for k in np.flatnonzero(reached):
        rel = 1e-6 if fit.names[k].startswith(("C1[", "C2[", "C3[")) else 1e-4
        h = rel * max(abs(values[k]), 1e-3)
        vp, vm = values.copy(), values.copy()
        vp[k] += h
        vm[k] -= h
        fd[k] = (chi2_at(vp) - chi2_at(vm)) / (2 * h)
    s = np.maximum(np.abs(values), 1e-3)
    sens = np.abs(fd) * s
    big = sens.max()
    assert big > 0
    strong = sens > 1e-6 * big
    err = np.abs(g - fd) * s
    assert np.max(err[strong] / sens[strong]) <= 1e-6
    assert np.max(err[~strong], initial=0.0) <= 1e-6 * big
    assert abs(float(c2) - chi2_at(values)) <= 1e-11 * float(c2)


def test_chi2_tables_autograd(gpu):
    model, prob, fit = _fit("dirbe_25um")
    values = torch.tensor(fit.values0, dtype=torch.float64, requires_grad=True)
    data = 1.02 * fit.forward(fit.values0)
    loss = 0.5 * chi2_tables(values, fit, data)
    loss.backward()
    _, g = fit.chi2_and_grad(fit.values0, data)
    torch.testing.assert_close(values.grad.cuda(), 0.5 * g, rtol=1e-15, atol=0)
