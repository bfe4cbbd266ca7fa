"""HEALPix pixel centres, RING ordering (Gorski et al. 2005, ApJ 622, 759).

``pix2vec(nside, ipix)`` returns unit vectors (3, n) of pixel centres. Polar-cap
pixels use 1 - |z| = i^2 / (3 nside^2) directly for sin(theta) to avoid the
cancellation of 1 - z^2 near the poles.
"""

from __future__ import annotations

import numpy as np


def npix(nside: int) -> int:
    return 12 * nside * nside


def pix2ang_ring(nside: int, ipix) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Return (z, sin_theta, phi) of RING-scheme pixel centres."""
    p = np.asarray(ipix, dtype=np.int64)
    ns = np.int64(nside)
    n_pix = 12 * ns * ns
    if np.any((p < 0) | (p >= n_pix)):
        msg = f"pixel index outside [0, {n_pix})"
        raise ValueError(msg)
    ncap = 2 * ns * (ns - 1)
    z = np.empty(p.shape, dtype=np.float64)
    sth = np.empty(p.shape, dtype=np.float64)
    phi = np.empty(p.shape, dtype=np.float64)
    fact2 = 1.0 / (3.0 * float(ns) * float(ns))

    north = p < ncap
    if np.any(north):
        pn = p[north]
        iring = (1 + np.floor(np.sqrt(1.0 + 2.0 * pn)).astype(np.int64)) // 2
        # guard against floating error in the square root
        iring = np.where(2 * iring * (iring - 1) > pn, iring - 1, iring)
        iring = np.where(2 * (iring + 1) * iring <= pn, iring + 1, iring)
        iphi = pn + 1 - 2 * iring * (iring - 1)
        tmp = iring.astype(np.float64) ** 2 * fact2
        z[north] = 1.0 - tmp
        sth[north] = np.sqrt(tmp * (2.0 - tmp))
        phi[north] = (iphi - 0.5) * (np.pi / 2.0) / iring

    south = p >= n_pix - ncap
    if np.any(south):
        ps = n_pix - p[south]
        iring = (1 + np.floor(np.sqrt(2.0 * ps - 1.0)).astype(np.int64)) // 2
        iring = np.where(2 * iring * (iring - 1) >= ps, iring - 1, iring)
        iring = np.where(2 * (iring + 1) * iring < ps, iring + 1, iring)
        iphi = 4 * iring + 1 - (ps - 2 * iring * (iring - 1))
        tmp = iring.astype(np.float64) ** 2 * fact2
        z[south] = -(1.0 - tmp)
        sth[south] = np.sqrt(tmp * (2.0 - tmp))
        phi[south] = (iphi - 0.5) * (np.pi / 2.0) / iring

    belt = ~(north | south)
    if np.any(belt):
        pb = p[belt] - ncap
        iring = pb // (4 * ns) + ns
        iphi = pb % (4 * ns) + 1
        fodd = np.where(((iring + ns) & 1) == 1, 1.0, 0.5)
        zb = (2 * ns - iring) * (2.0 / 3.0) / float(ns)
        z[belt] = zb
        sth[belt] = np.sqrt((1.0 - zb) * (1.0 + zb))
        phi[belt] = (iphi - fodd) * (np.pi / 2.0) / float(ns)
    return z, sth, phi


def pix2vec(nside: int, ipix=None) -> np.ndarray:
    """Unit vectors (3, n) of RING pixel centres (all pixels when ``ipix`` is None)."""
    if ipix is None:
        ipix = np.arange(npix(nside), dtype=np.int64)
    z, sth, phi = pix2ang_ring(nside, ipix)
    return np.stack([sth * np.cos(phi), sth * np.sin(phi), z])
