"""Spectral tables as fit parameters: the linear weights behind the device set-up.

The effective spectral parameters are linear in the tables once the bandpass is
fixed (zodipy/unpack_model.py:23-173): effective = w @ table. The weights come
from the package's own unpacking on unit tables; these tests pin them against
that unpacking on the model's and on random tables, and against closed forms
(one-hot at a grid node, 1/2-1/2 half-way between nodes, the lower node at a
nearest-neighbour tie, weights of a normalised bandpass summing to 1). The
Planck-table coefficients are pinned against the host's table.
"""

from __future__ import annotations

import numpy as np
import pytest

from zodipy_amd import Model
from zodipy_amd import units as u
from zodipy_amd.spectral import (
    CALIBRATION_UNIT_FACTOR,
    SPEC_C1,
    SPEC_CALIBRATION,
    SPEC_EMISSIVITY,
    effective_param,
    planck_coefficients,
    spectral_tables,
)

H, KB, C, MJY = 6.62607015e-34, 1.380649e-23, 299792458.0, 1e-20


def gaussian_bandpass_25um():
    x = np.linspace(17.5, 32.5, 50) * u.micron
    return x, np.exp(-0.5 * ((x.value - 25.0) / 2.5) ** 2)


def dirbe_25um_instrument_bandpass():
    """The DIRBE 25 um band's 82 wavelengths and weights (the reference's test data)."""
    import json
    from pathlib import Path

    d = json.loads((Path(__file__).parent / "golden" / "dirbe_25um_bandpass.json").read_text())
    return np.asarray(d["wavelength_micron"]) * u.micron, np.asarray(d["weight"])


CASES = {
    "dirbe_25um": dict(x=25 * u.micron),
    "dirbe_25um_instrument": dict(x=dirbe_25um_instrument_bandpass()[0],
                                  weights=dirbe_25um_instrument_bandpass()[1]),
    "dirbe_18.5um": dict(x=18.5 * u.micron),
    "dirbe_bandpass_um": dict(x=gaussian_bandpass_25um()[0], weights=gaussian_bandpass_25um()[1]),
    "dirbe_bandpass_ghz": dict(x=np.linspace(11000, 13000, 20) * u.GHz, weights=np.ones(20)),
    "planck18_857": dict(x=857 * u.GHz, name="planck18"),
    "planck18_tophat": dict(x=np.linspace(771.3, 942.7, 30) * u.GHz, weights=np.ones(30),
                            name="planck18", extrapolate=True),
    "planck13_545": dict(x=545 * u.GHz, name="planck13"),
    "rrm_1.25um": dict(x=1.25 * u.micron, name="rrm-experimental", extrapolate=True),
}


def _tables(model):
    return spectral_tables(model._ipd_model, model._x, model._normalized_weights,
                           model._extrapolate)


def _effective(model, name):
    sp = model._spectral
    if "." in name:
        key, lab = name.split(".")
        return next(v[key] for k, v in sp["comp"].items() if k.value == lab)
    return sp["shared"][name]


@pytest.mark.parametrize("case", list(CASES))
def test_weights_reproduce_the_unpacking(case):
    kw = dict(CASES[case])
    x = kw.pop("x")
    model = Model(x, **kw)
    tables = _tables(model)
    assert tables
    rng = np.random.default_rng(0)
    ipd = model._ipd_model
    xm = np.asarray(x.to_value(ipd.spectrum.unit, equivalencies=u.spectral()), dtype=np.float64)
    for t in tables:
        want = _effective(model, t.name)
        got = t.weights @ t.values
        assert abs(got - want) <= 1e-14 * max(abs(want), 1e-300), (t.name, got, want)
        # any table, not only the model's: the map is linear
        nearest = t.name in ("C1", "C2", "C3")
        scale = CALIBRATION_UNIT_FACTOR if t.target == SPEC_CALIBRATION else 1.0
        for _ in range(3):
            r = rng.normal(size=t.values.size)
            want_r = float(effective_param(xm, model._normalized_weights, ipd.spectrum, r,
                                           nearest=nearest, extrapolate=model._extrapolate))
            np.testing.assert_allclose(t.weights @ r, want_r * scale, rtol=1e-13,
                                       atol=1e-15 * np.abs(r).max() * scale)


def test_closed_forms_of_the_weights():
    grid = [1.25, 2.2, 3.5, 4.9, 12.0, 25.0, 60.0, 100.0, 140.0, 240.0]  # DIRBE bands [um]
    t25 = {t.name: t for t in _tables(Model(25 * u.micron))}
    assert t25["emissivity.cloud"].target == SPEC_EMISSIVITY
    np.testing.assert_array_equal(t25["emissivity.cloud"].weights, np.eye(10)[grid.index(25.0)])
    mid = {t.name: t for t in _tables(Model(18.5 * u.micron))}
    want = np.zeros(10)
    want[4] = want[5] = 0.5
    np.testing.assert_array_equal(mid["emissivity.ring"].weights, want)
    # nearest neighbour: a tie half-way between nodes goes to the lower one
    assert mid["C1"].target == SPEC_C1
    np.testing.assert_array_equal(mid["C1"].weights, np.eye(10)[4])
    # a normalised bandpass in the model's own unit: the weights sum to one
    x, w = gaussian_bandpass_25um()
    bp = {t.name: t for t in _tables(Model(x, weights=w))}
    assert abs(bp["emissivity.feature"].weights.sum() - 1.0) <= 1e-14
    assert np.all(bp["emissivity.feature"].weights >= 0)


@pytest.mark.parametrize("case", ["dirbe_25um", "dirbe_bandpass_um", "planck18_tophat"])
def test_planck_coefficients_reproduce_the_table(case):
    kw = dict(CASES[case])
    x = kw.pop("x")
    model = Model(x, **kw)
    nu, q = planck_coefficients(x, model._normalized_weights)
    temps, want = model._b_nu_table
    a = H * nu[:, None] / (KB * temps[None, :])
    b = (2 * H * nu[:, None] ** 3 / (C**2 * np.expm1(a))) / MJY
    np.testing.assert_allclose(q @ b, want, rtol=1e-13, atol=0)


BANDPASS_CASES = [
    ("dirbe", np.linspace(20.0, 30.0, 8), u.micron, False),
    ("dirbe", np.linspace(10.0, 15.0, 6) * 1e3, u.GHz, False),
    ("planck18", np.linspace(771.3, 942.7, 12), u.GHz, True),
    ("rrm-experimental", np.linspace(20.0, 30.0, 8), u.micron, False),
]


@pytest.mark.parametrize("name,xs,unit,extrap", BANDPASS_CASES,
                         ids=["dirbe_um", "dirbe_ghz_on_um", "planck18_857", "rrm_um"])
def test_bandpass_maps_reproduce_the_unpacking(name, xs, unit, extrap):
    """The linear maps behind bandpass-weight fits: at random weights, K @ wn equals
    the unpacking of those weights (every tabulated parameter) and
    sum_i c_i wn_i B_nu_i(T_k) the host's Planck table of them (1e-13)."""
    from zodipy_amd.spectral import (
        bandpass_maps,
        effective_value,
        normalize_weights,
        planck_table,
        unpack_spectral,
    )
    from zodipy_amd.zodiacal_light_model import model_registry

    model = model_registry.get_model(name)
    x = xs * unit
    maps = bandpass_maps(model, x, extrap)
    assert maps.K.shape == (len(maps.names), xs.size) and len(maps.names) > 0
    rng = np.random.default_rng(17)
    for _ in range(3):
        w = rng.uniform(0.2, 1.5, xs.size)
        wn = normalize_weights(x, w)
        np.testing.assert_allclose(wn, w / (maps.c @ w), rtol=1e-15)
        sp = unpack_spectral(model, x, wn, extrap)
        want = np.array([effective_value(sp, n) for n in maps.names])
        np.testing.assert_allclose(maps.K @ wn, want, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose((maps.c * wn) @ maps.bnu, planck_table(x, wn)[1], rtol=1e-13)


def test_trapezoid_coefficients():
    from zodipy_amd.spectral import trapezoid_coefficients

    rng = np.random.default_rng(5)
    x = np.sort(rng.uniform(0, 10, 17))
    y = rng.normal(size=17)
    assert trapezoid_coefficients(x) @ y == pytest.approx(np.trapezoid(y, x), rel=1e-14)
    assert trapezoid_coefficients(x[::-1]) @ y == pytest.approx(np.trapezoid(y, x[::-1]),
                                                                 rel=1e-14)
