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


class BodyTrack:
    """A body's heliocentric ecliptic positions tabulated on a time grid.

    For time-ordered data the reference interpolates an hourly ephemeris grid
    with ``scipy.interpolate.CubicSpline(grid_mjd, positions, axis=-1)``
    (zodipy/bodies.py:22-35). A ``BodyTrack`` holds that spline (the grid and
    positions come from the caller, e.g. exported from an ephemeris; no network
    or astropy needed) and evaluates it either on the host (``__call__``) or per
    ray on the GPU (``on_device``, kernel ``zodi_spline_positions``), where it
    rounds bit for bit like scipy's evaluation.

    Args:
        mjd_grid: (m,) increasing Modified Julian Dates, m >= 2.
        positions: (3, m) heliocentric ecliptic positions [AU].
    """

    def __init__(self, mjd_grid, positions) -> None:
        x = np.asarray(mjd_grid, dtype=np.float64).ravel()
        pos = np.asarray(positions, dtype=np.float64)
        if x.size < 2 or pos.shape != (3, x.size):
            msg = f"need >= 2 grid times and positions of shape (3, {x.size}), got {pos.shape}"
            raise ValueError(msg)
        if np.any(np.diff(x) <= 0):
            msg = "grid times must be strictly increasing"
            raise ValueError(msg)
        self.x = x
        self.spline = interpolate.CubicSpline(x, pos, axis=-1)
        self.c = np.ascontiguousarray(self.spline.c, dtype=np.float64)  # (4, m - 1, 3)
        self._dev = {}

    @property
    def records(self) -> np.ndarray:
        """The spline as the device kernel reads it: (m - 1, 16), one row per interval.

        Row k holds x_k, the constant, linear, quadratic and cubic coefficients of
        (x, y, z) in powers of t - x_k, x_{k+1} and two zeros: 128 bytes, one
        cache line, so a ray's lookup touches a single line.
        """
        ni = self.x.size - 1
        rec = np.zeros((ni, 16), dtype=np.float64)
        rec[:, 0] = self.x[:-1]
        for j in range(4):
            rec[:, 1 + 3 * j:4 + 3 * j] = self.c[3 - j]
        rec[:, 13] = self.x[1:]
        return rec

    @classmethod
    def hourly(cls, body_position_fn, t0_mjd: float, t1_mjd: float) -> "BodyTrack":
        """Track on the reference's hourly grid over [t0, t1] from ``fn(mjd) -> (3, m)``."""
        grid = arrange_obstimes(t0_mjd, t1_mjd)
        return cls(grid, body_position_fn(grid))

    def __call__(self, mjd) -> np.ndarray:
        return self.spline(np.asarray(mjd, dtype=np.float64))

    def on_device(self, mjd, device="cuda"):
        """Positions (3, n) at times ``mjd`` (array or tensor), evaluated on the GPU."""
        import torch

        from zodipy_amd import _lib
        from zodipy_amd.runtime import _require_cuda, _stream

        dev = _require_cuda(device)
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = torch.as_tensor(self.records, device=dev)
        rec_t = self._dev[key]
        if isinstance(mjd, torch.Tensor):
            t = mjd.to(device=dev, dtype=torch.float64).reshape(-1).contiguous()
        else:
            t = torch.as_tensor(np.asarray(mjd, dtype=np.float64).ravel(), device=dev)
        out = torch.empty((3, t.numel()), dtype=torch.float64, device=dev)
        st = _lib.load().zodi_spline_positions(rec_t.data_ptr(), rec_t.shape[0], t.data_ptr(),
                                               t.numel(), out.data_ptr(), _stream())
        _lib.check(st, "zodi_spline_positions")
        return out
