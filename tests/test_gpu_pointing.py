"""Device pointing ingestion (zodi_pix2vec, zodi_lonlat_to_vec) against the host.

The host references are this package's NumPy HEALPix RING pixel centres
(pinned by tests/test_host.py against the closed-form ring latitudes) and the
NumPy lon/lat conversion with the same J2000 rotation matrices. The device uses
its own sin/cos, so unit-vector components agree to a few ulps (1e-15
absolute), not bit for bit; the emission computed from either agrees to 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from zodipy_amd import Model, coords, healpix, pointing
from zodipy_amd import units as u

pytestmark = pytest.mark.gpu

ATOL = 1e-15


@pytest.mark.parametrize("nside", [1, 2, 16, 64])
def test_pix2vec_all_pixels(gpu, nside):
    got = pointing.pix2vec(nside).cpu().numpy()
    want = healpix.pix2vec(nside)
    np.testing.assert_allclose(got, want, rtol=0, atol=ATOL)


@pytest.mark.parametrize("nside", [2**13, 2**20, 2**29])
def test_pix2vec_large_nside_random_pixels(gpu, nside):
    rng = np.random.default_rng(nside % 1000)
    npix = 12 * nside * nside
    ipix = np.concatenate([rng.integers(0, npix, 20000), [0, npix - 1, 2 * nside * (nside - 1),
                                                         npix - 2 * nside * (nside - 1) - 1]])
    got = pointing.pix2vec(nside, ipix).cpu().numpy()
    want = healpix.pix2vec(nside, ipix)
    np.testing.assert_allclose(got, want, rtol=0, atol=ATOL)
    np.testing.assert_allclose(np.linalg.norm(got, axis=0), 1.0, rtol=0, atol=4e-16)


@pytest.mark.parametrize("frame", ["icrs", "galactic"])
def test_pix2vec_rotated_to_ecliptic(gpu, frame):
    got = pointing.pix2vec(32, frame=frame).cpu().numpy()
    want = coords.to_ecliptic_matrix(frame) @ healpix.pix2vec(32)
    np.testing.assert_allclose(got, want, rtol=0, atol=ATOL)


def test_pix2vec_invalid_pixels_are_nan(gpu):
    got = pointing.pix2vec(4, np.array([-1, 0, 192, 191])).cpu().numpy()
    assert np.isnan(got[:, [0, 2]]).all()
    assert np.isfinite(got[:, [1, 3]]).all()


@pytest.mark.parametrize("frame", ["barycentricmeanecliptic", "icrs", "galactic"])
@pytest.mark.parametrize("degrees", [False, True])
def test_lonlat_to_vectors(gpu, frame, degrees):
    rng = np.random.default_rng(7)
    lon = rng.uniform(-360, 720, 50000)
    lat = rng.uniform(-90, 90, 50000)
    lon_in, lat_in = (lon, lat) if degrees else (np.radians(lon), np.radians(lat))
    got = pointing.lonlat_to_vectors(lon_in, lat_in, frame=frame, degrees=degrees)
    want = coords.to_ecliptic_matrix(frame) @ coords.lonlat_to_xyz(np.radians(lon),
                                                                 np.radians(lat))
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=ATOL)


def test_lonlat_large_and_non_finite_angles(gpu):
    """Angles beyond 64 radians take the library's sin/cos, non-finite ones give NaN."""
    lon = np.array([3700.0, -2.0e4, 1.0e7, 123.4, np.nan, np.inf, 30.0])
    lat = np.array([10.0, -89.0, 45.0, 4000.0, 0.0, 0.0, -np.inf])
    got = pointing.lonlat_to_vectors(lon, lat, degrees=True).cpu().numpy()
    with np.errstate(invalid="ignore"):
        want = coords.lonlat_to_xyz(np.radians(lon), np.radians(lat))
    np.testing.assert_allclose(got, want, rtol=0, atol=ATOL)
    assert np.isnan(got[:2, 4:]).all() and np.isnan(got[:, 6]).all()


def test_evaluate_skycoord_uses_device_vectors(gpu):
    """Model.evaluate on a galactic SkyCoord equals evaluate_vectors on host vectors."""
    rng = np.random.default_rng(3)
    lon, lat = rng.uniform(0, 360, 2000), rng.uniform(-90, 90, 2000)
    sc = coords.SkyCoord(lon, lat, unit=u.deg, frame="galactic",
                         obstime=coords.Time("2022-01-14"))
    model = Model(25 * u.micron)
    obs = np.array([0.3, 0.95, 0.0])
    got = model.evaluate(sc, obspos=obs * u.AU, earthpos=obs, return_comps=True).value
    want = model.evaluate_vectors(sc.ecliptic_xyz(), obspos=obs, earthpos=obs,
                                  return_comps=True).value
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


def test_empty_inputs(gpu):
    assert pointing.lonlat_to_vectors([], []).shape == (3, 0)
    assert pointing.pix2vec(8, np.array([], dtype=np.int64)).shape == (3, 0)
