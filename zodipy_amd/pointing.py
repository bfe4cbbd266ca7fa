"""Pointings to ecliptic unit vectors on the GPU.

The reference turns a ``SkyCoord`` into unit vectors with
``skycoord.transform_to(BarycentricMeanEcliptic).cartesian`` on the host
(zodipy/model.py:247-251), and its HEALPix examples build the pointings with
healpy's ``pix2ang`` (docs/examples/healpy_map.py:16). Here both run as HIP
kernels (``zodi_pix2vec``, ``zodi_lonlat_to_vec``) that write the SoA (3, n)
float64 array the line-of-sight kernels read, with the fixed J2000 frame
rotation of ``coords.to_ecliptic_matrix`` applied on the way.
"""

from __future__ import annotations

import numpy as np
import torch

from zodipy_amd import _lib
from zodipy_amd.coords import to_ecliptic_matrix
from zodipy_amd.runtime import _require_cuda, _stream

__all__ = ["pix2vec", "lonlat_to_vectors"]


def _device_flat(values, dtype, dev) -> torch.Tensor:
    """1-D contiguous device tensor; device tensors are used in place (no host round trip)."""
    if isinstance(values, torch.Tensor):
        return values.to(device=dev, dtype=dtype).reshape(-1).contiguous()
    np_dtype = np.int64 if dtype == torch.int64 else np.float64
    return torch.as_tensor(np.asarray(values, dtype=np_dtype).ravel(), device=dev).contiguous()


def _rotation(frame: str):
    m = to_ecliptic_matrix(frame)
    if np.array_equal(m, np.eye(3)):
        return None, None
    rot = np.ascontiguousarray(m, dtype=np.float64)
    return rot, rot.ctypes.data


def pix2vec(nside: int, ipix=None, *, frame: str = "barycentricmeanecliptic",
            device="cuda") -> torch.Tensor:
    """Ecliptic unit vectors (3, n) of HEALPix RING pixel centres given in ``frame``.

    ``ipix``: pixel indices (array or int64 tensor), or None for every pixel.
    Indices outside [0, 12 nside^2) give NaN vectors.
    """
    dev = _require_cuda(device)
    lib = _lib.load()
    rot, rot_ptr = _rotation(frame)
    if ipix is None:
        n = 12 * int(nside) * int(nside)
        pix = None
    else:
        pix = _device_flat(ipix, torch.int64, dev)
        n = int(pix.numel())
    out = torch.empty((3, n), dtype=torch.float64, device=dev)
    st = lib.zodi_pix2vec(int(nside), pix.data_ptr() if pix is not None else None, 0, n,
                          rot_ptr, out.data_ptr(), _stream())
    _lib.check(st, "zodi_pix2vec")
    del rot
    return out


def lonlat_to_vectors(lon, lat, *, frame: str = "barycentricmeanecliptic",
                      degrees: bool = False, device="cuda") -> torch.Tensor:
    """Ecliptic unit vectors (3, n) from longitude/latitude arrays given in ``frame``."""
    dev = _require_cuda(device)
    lib = _lib.load()
    lon_t = _device_flat(lon, torch.float64, dev)
    lat_t = _device_flat(lat, torch.float64, dev)
    if lon_t.shape != lat_t.shape:
        msg = "lon and lat must have the same number of elements"
        raise ValueError(msg)
    rot, rot_ptr = _rotation(frame)
    n = int(lon_t.numel())
    out = torch.empty((3, n), dtype=torch.float64, device=dev)
    st = lib.zodi_lonlat_to_vec(lon_t.data_ptr(), lat_t.data_ptr(), n, int(bool(degrees)),
                                rot_ptr, out.data_ptr(), _stream())
    _lib.check(st, "zodi_lonlat_to_vec")
    del rot
    return out
