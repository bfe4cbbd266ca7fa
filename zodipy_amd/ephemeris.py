"""Observer and Earth positions for the front end.

The reference looks bodies up with astropy's solar-system ephemeris and, for
time-ordered data, interpolates an hourly grid with a cubic spline
(zodipy/bodies.py:16-99). That work is per observer, not per ray, and stays on
the host. When astropy is importable it is used the same way; otherwise the
caller supplies the Earth position explicitly (``earthpos=``), from which the
"earth" and "semb-l2" observers follow.
"""

from __future__ import annotations

import numpy as np
from scipy import interpolate

# Mean distance from Earth to the Sun-Earth/Moon-barycentre L2 point [AU]
# (zodipy/bodies.py:13).
MEAN_DIST_TO_L2 = 0.009896235034000056

try:  # pragma: no cover - astropy is optional and absent in the default runtime
    import astropy.coordinates as _ac
    import astropy.time as _at
    import astropy.units as _au

    HAVE_ASTROPY = True
except ImportError:  # pragma: no cover
    HAVE_ASTROPY = False


def get_semb_l2_pos(earthpos: np.ndarray) -> np.ndarray:
    """SEMB-L2 approximated at a fixed distance beyond Earth on the Sun-Earth line.

    The norm is taken over the whole array, as in zodipy/bodies.py:38-50.
    """
    earth_distance = np.linalg.norm(earthpos)
    semb_l2_distance = earth_distance + MEAN_DIST_TO_L2
    return earthpos / earth_distance * semb_l2_distance


def arrange_obstimes(t0_mjd: float, t1_mjd: float) -> np.ndarray:
    """Hourly MJD grid covering [t0, t1] (zodipy/bodies.py:16-19)."""
    dt = 1.0 / 24.0
    return np.arange(t0_mjd, t1_mjd + dt, dt)


def _require_astropy(what: str) -> None:
    if not HAVE_ASTROPY:
        msg = (
            f"{what} needs astropy's solar-system ephemeris, which is not installed; "
            "pass the heliocentric ecliptic Earth position explicitly with `earthpos=` "
            "(and `obspos=` for observers other than 'earth' and 'semb-l2')"
        )
        raise ValueError(msg)


def _astropy_time(mjd):  # pragma: no cover - needs astropy
    return _at.Time(np.asarray(mjd, dtype=np.float64), format="mjd")


def body_position(body: str, mjd, ephemeris: str) -> np.ndarray:  # pragma: no cover
    """Heliocentric mean-ecliptic position of ``body`` [AU] via astropy."""
    _require_astropy(f"the position of {body!r}")
    try:
        pos = (
            _ac.get_body(body, _astropy_time(mjd), ephemeris=ephemeris)
            .transform_to(_ac.HeliocentricMeanEcliptic)
            .cartesian.xyz.to_value(_au.AU)
        )
    except KeyError as error:
        valid = [*_ac.solar_system_ephemeris.bodies, "semb-l2"]
        msg = f"Invalid observer string: '{body}'. Valid observers are: {valid}"
        raise ValueError(msg) from error
    return np.asarray(pos, dtype=np.float64)


def interp_body_position(body: str, mjd: np.ndarray, grid_mjd: np.ndarray,
                         ephemeris: str) -> np.ndarray:  # pragma: no cover - needs astropy
    """Cubic-spline interpolation of an hourly ephemeris grid (zodipy/bodies.py:22-35)."""
    pos = body_position(body, grid_mjd, ephemeris)
    return interpolate.CubicSpline(grid_mjd, pos, axis=-1)(mjd)
