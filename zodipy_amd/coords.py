"""Lightweight sky coordinates and observation times (astropy-free).

The reference takes ``astropy.coordinates.SkyCoord`` pointings and rotates them
to the ``BarycentricMeanEcliptic`` frame (zodipy/model.py:247-251). This module
provides the subset the line-of-sight integral needs when astropy is absent:

* ``Time``: observation times as Modified Julian Dates (ISO strings or MJD
  numbers); only ``size``/``mjd``/indexing are used by the integral;
* ``SkyCoord``: longitude/latitude pointings in the ``barycentricmeanecliptic``,
  ``icrs`` or ``galactic`` frame with an ``obstime``;
* fixed J2000 rotation matrices to the mean ecliptic and equinox of J2000:
  ICRS -> ecliptic is the IAU 2006 bias-precession matrix at J2000 followed by
  the rotation by the J2000 mean obliquity (ERFA ``ecm06`` at TT J2000, which is
  what astropy's ICRS -> BarycentricMeanEcliptic transform evaluates);
  Galactic -> ICRS goes through FK5 J2000, with the galactic pole
  (192.8594812065348 deg, 27.12825118085622 deg), the north celestial pole at
  galactic longitude 122.9319185680026 deg, and the FK5 -> ICRS frame bias of
  USNO Circular 179 (eta0 = -19.9 mas, xi0 = 9.1 mas, da0 = -22.9 mas).

Only directions are rotated; the rotations are identical for heliocentric and
barycentric frames, as in the reference.
"""

from __future__ import annotations

import datetime as _dt

import numpy as np

from zodipy_amd import units as u

_MJD0 = _dt.datetime(1858, 11, 17)
_ARCSEC = np.pi / (180.0 * 3600.0)


def _rot(angle_rad: float, axis: str) -> np.ndarray:
    """Frame (passive) rotation matrix about ``axis`` by ``angle_rad``."""
    c, s = np.cos(angle_rad), np.sin(angle_rad)
    if axis == "x":
        return np.array([[1, 0, 0], [0, c, s], [0, -s, c]], dtype=np.float64)
    if axis == "y":
        return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]], dtype=np.float64)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]], dtype=np.float64)


def icrs_to_ecliptic_matrix() -> np.ndarray:
    """ICRS -> mean ecliptic and equinox of J2000 (IAU 2006, at t = 0).

    Fukushima-Williams bias-precession angles at J2000: gamma_bar = -0.052928",
    phi_bar = 84381.412819", psi_bar = -0.041775", eps_A = 84381.406". The
    bias-precession matrix is R1(-eps_A) R3(-psi_bar) R1(phi_bar) R3(gamma_bar);
    rotating by the mean obliquity R1(eps_A) leaves R3(-psi_bar) R1(phi_bar) R3(gamma_bar).
    """
    gamb = -0.052928 * _ARCSEC
    phib = 84381.412819 * _ARCSEC
    psib = -0.041775 * _ARCSEC
    return _rot(-psib, "z") @ _rot(phib, "x") @ _rot(gamb, "z")


def fk5_to_icrs_matrix() -> np.ndarray:
    deg = np.pi / 180.0
    eta0 = -19.9 / 3600000.0 * deg
    xi0 = 9.1 / 3600000.0 * deg
    da0 = -22.9 / 3600000.0 * deg
    icrs_to_fk5 = _rot(-eta0, "x") @ _rot(xi0, "y") @ _rot(da0, "z")
    return icrs_to_fk5.T


def galactic_to_icrs_matrix() -> np.ndarray:
    deg = np.pi / 180.0
    fk5_to_gal = (
        _rot((180 - 122.9319185680026) * deg, "z")
        @ _rot((90 - 27.12825118085622) * deg, "y")
        @ _rot(192.8594812065348 * deg, "z")
    )
    return fk5_to_icrs_matrix() @ fk5_to_gal.T


FRAMES = ("barycentricmeanecliptic", "icrs", "galactic")


def to_ecliptic_matrix(frame: str) -> np.ndarray:
    frame = frame.lower()
    if frame == "barycentricmeanecliptic":
        return np.eye(3)
    if frame == "icrs":
        return icrs_to_ecliptic_matrix()
    if frame == "galactic":
        return icrs_to_ecliptic_matrix() @ galactic_to_icrs_matrix()
    msg = f"unsupported frame {frame!r}; supported: {', '.join(FRAMES)}"
    raise ValueError(msg)


def lonlat_to_xyz(lon_rad, lat_rad) -> np.ndarray:
    lon = np.asarray(lon_rad, dtype=np.float64)
    lat = np.asarray(lat_rad, dtype=np.float64)
    cl = np.cos(lat)
    return np.stack([cl * np.cos(lon), cl * np.sin(lon), np.sin(lat)])


class Time:
    """Observation time(s) as Modified Julian Dates."""

    def __init__(self, value, format: str | None = None) -> None:  # noqa: A002
        if isinstance(value, Time):
            self._mjd = value._mjd.copy()
            return
        if format == "mjd" or (format is None and not _is_stringish(value)):
            self._mjd = np.asarray(value, dtype=np.float64)
        else:
            vals = [value] if isinstance(value, str) else list(value)
            mjd = [(_parse_iso(v) - _MJD0).total_seconds() / 86400.0 for v in vals]
            self._mjd = np.asarray(mjd[0] if isinstance(value, str) else mjd, dtype=np.float64)

    @property
    def mjd(self):
        return self._mjd[()] if self._mjd.ndim == 0 else self._mjd

    @property
    def size(self) -> int:
        return int(self._mjd.size)

    @property
    def isscalar(self) -> bool:
        return self._mjd.ndim == 0

    def __len__(self) -> int:
        return int(self._mjd.size)

    def __getitem__(self, idx) -> "Time":
        return Time(self._mjd[idx], format="mjd")

    def __repr__(self) -> str:
        return f"Time(mjd={self._mjd!r})"


def _is_stringish(value) -> bool:
    if isinstance(value, str):
        return True
    try:
        return len(value) > 0 and isinstance(next(iter(value)), str)
    except TypeError:
        return False


def _parse_iso(text: str) -> _dt.datetime:
    text = text.strip().replace("T", " ")
    for fmt in ("%Y-%m-%d %H:%M:%S.%f", "%Y-%m-%d %H:%M:%S", "%Y-%m-%d %H:%M", "%Y-%m-%d"):
        try:
            return _dt.datetime.strptime(text, fmt)
        except ValueError:
            continue
    msg = f"cannot parse time {text!r}"
    raise ValueError(msg)


class SkyCoord:
    """Pointings given as longitude/latitude in a celestial frame, with an obstime.

    ``SkyCoord(lon, lat, unit=units.deg, frame="galactic", obstime=Time(...))``;
    ``lon``/``lat`` may be quantities (then ``unit`` is optional) or numbers.
    """

    def __init__(self, lon, lat, unit=None, frame: str = "icrs", obstime=None) -> None:
        unit = unit if unit is not None else u.deg
        self.lon = _angle(lon, unit)
        self.lat = _angle(lat, unit)
        if self.lon.shape != self.lat.shape:
            msg = "lon and lat must have the same shape"
            raise ValueError(msg)
        frame = getattr(frame, "name", frame)
        self.frame = str(frame).lower()
        to_ecliptic_matrix(self.frame)  # validate
        self.obstime = obstime if (obstime is None or isinstance(obstime, Time)) else Time(obstime)

    @property
    def size(self) -> int:
        return int(self.lon.size)

    @property
    def isscalar(self) -> bool:
        return self.lon.ndim == 0

    @property
    def shape(self):
        return self.lon.shape

    def __len__(self) -> int:
        return self.size

    def ecliptic_xyz(self) -> np.ndarray:
        """Unit vectors (3, N) in the mean ecliptic and equinox of J2000."""
        xyz = lonlat_to_xyz(self.lon.ravel(), self.lat.ravel())
        return to_ecliptic_matrix(self.frame) @ xyz


def _angle(value, unit) -> np.ndarray:
    if isinstance(value, u.Quantity):
        return np.asarray(value.to_value(u.rad), dtype=np.float64)
    return np.asarray(u.Quantity(value, unit).to_value(u.rad), dtype=np.float64)
