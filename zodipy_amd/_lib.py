"""ctypes binding of the HIP library ``libzodi_hip.so`` (C ABI in include/zodi_hip.h).

The library is required: every compute call goes through it and raises
``RuntimeError`` if it cannot be loaded. There is no CPU fallback.
"""

from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

from zodipy_amd.packing import ABI_VERSION, MODEL_DTYPE

LIB_PATH = Path(__file__).resolve().parent / "libzodi_hip.so"

OBS_PER_SEGMENT, OBS_PER_RAY = 0, 1
OUT_COMPS, OUT_SUM = 0, 1
ADJ_VJP_SUM, ADJ_VJP_COMPS, ADJ_CHI2 = 0, 1, 2

EXPORTED_SYMBOLS = (
    "zodi_abi_version",
    "zodi_model_size",
    "zodi_last_error",
    "zodi_forward",
    "zodi_line_of_sight",
    "zodi_number_density",
    "zodi_pix2vec",
    "zodi_lonlat_to_vec",
    "zodi_spline_positions",
    "zodi_host_device_pointer",
    "zodi_max_radius",
    "zodi_adjoint_workspace",
    "zodi_adjoint",
    "zodi_adjoint_noalloc",
    "zodi_table_vjp_workspace",
    "zodi_table_vjp",
    "zodi_spectral_vjp_workspace",
    "zodi_spectral_vjp",
    "zodi_jacobian",
    "zodi_spectral_setup",
)


class Launch(ct.Structure):
    """``struct zodi_launch``."""

    _fields_ = [
        ("block", ct.c_int32),
        ("lanes_per_ray", ct.c_int32),
        ("blocks_per_cu", ct.c_int32),
        ("variant", ct.c_int32),
    ]


_lib = None

_p = ct.c_void_p
_i64 = ct.c_int64
_i32 = ct.c_int32


def _bind(lib: ct.CDLL) -> None:
    lib.zodi_abi_version.restype = ct.c_int32
    lib.zodi_abi_version.argtypes = []
    lib.zodi_model_size.restype = ct.c_int64
    lib.zodi_model_size.argtypes = []
    lib.zodi_last_error.restype = ct.c_char_p
    lib.zodi_last_error.argtypes = []
    lib.zodi_forward.restype = ct.c_int32
    lib.zodi_forward.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _p, _i32, _p, _p, _i32,
                                 ct.POINTER(Launch), _p]
    lib.zodi_line_of_sight.restype = ct.c_int32
    lib.zodi_line_of_sight.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _i32, _p, _p, _p, _p]
    lib.zodi_number_density.restype = ct.c_int32
    lib.zodi_number_density.argtypes = [_p, _p, _p, _i64, _p, _p, _p]
    lib.zodi_pix2vec.restype = ct.c_int32
    lib.zodi_pix2vec.argtypes = [_i64, _p, _i64, _i64, _p, _p, _p]
    lib.zodi_lonlat_to_vec.restype = ct.c_int32
    lib.zodi_lonlat_to_vec.argtypes = [_p, _p, _i64, _i32, _p, _p, _p]
    lib.zodi_spline_positions.restype = ct.c_int32
    lib.zodi_spline_positions.argtypes = [_p, _i64, _p, _i64, _p, _p]
    lib.zodi_max_radius.restype = ct.c_int32
    lib.zodi_max_radius.argtypes = [_p, _i64, _p, _p]
    lib.zodi_host_device_pointer.restype = ct.c_int32
    lib.zodi_host_device_pointer.argtypes = [_p, ct.POINTER(ct.c_void_p)]
    lib.zodi_adjoint_workspace.restype = ct.c_int64
    lib.zodi_adjoint_workspace.argtypes = [_p, _i64, _i32, ct.POINTER(Launch)]
    lib.zodi_adjoint.restype = ct.c_int32
    lib.zodi_adjoint.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _p, _i32, _p, _i32, _p, _p,
                                 _p, ct.c_double, _p, _p, _p, _i64, ct.POINTER(Launch), _p]
    lib.zodi_adjoint_noalloc.restype = ct.c_int32
    lib.zodi_adjoint_noalloc.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _p, _i32, _p, _i32, _p,
                                         _p, _p, ct.c_double, _p, _p, ct.POINTER(Launch), _p]
    lib.zodi_table_vjp_workspace.restype = _i64
    lib.zodi_table_vjp_workspace.argtypes = [_p, _i64, ct.POINTER(Launch)]
    lib.zodi_table_vjp.restype = ct.c_int32
    lib.zodi_table_vjp.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _p, _i32, _p, _i32, _p, _p,
                                   _p, ct.c_double, _p, _p, _p, _i64, ct.POINTER(Launch), _p]
    lib.zodi_spectral_vjp_workspace.restype = _i64
    lib.zodi_spectral_vjp_workspace.argtypes = [_p, _i64, ct.POINTER(Launch)]
    lib.zodi_spectral_vjp.restype = ct.c_int32
    lib.zodi_spectral_vjp.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _p, _i32, _p, _i32, _p,
                                      _p, _p, ct.c_double, _p, _p, _p, _i64, ct.POINTER(Launch),
                                      _p]
    lib.zodi_spectral_setup.restype = ct.c_int32
    lib.zodi_spectral_setup.argtypes = [_p, _p, _i32, _i32, _p, _p, _p, _i32, _p, _p, _p]
    lib.zodi_jacobian.restype = ct.c_int32
    lib.zodi_jacobian.argtypes = [_p, _p, _p, _i64, _p, _i32, _p, _p, _i32, _p, _p, _p, _i64,
                                  ct.POINTER(Launch), _p]


def load() -> ct.CDLL:
    """Load and check the HIP library (cached)."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("ZODI_HIP_LIB", LIB_PATH))
    if not path.exists():
        msg = (
            f"HIP library {path} not found; build it with `python -m zodipy_amd.csrc.build` "
            "(requires ROCm hipcc)"
        )
        raise RuntimeError(msg)
    try:
        lib = ct.CDLL(str(path))
    except OSError as error:
        msg = f"cannot load HIP library {path}: {error}"
        raise RuntimeError(msg) from error
    _bind(lib)
    if lib.zodi_abi_version() != ABI_VERSION:
        msg = f"HIP library ABI {lib.zodi_abi_version()} != expected {ABI_VERSION}"
        raise RuntimeError(msg)
    if lib.zodi_model_size() != MODEL_DTYPE.itemsize:
        msg = (
            f"struct zodi_model is {lib.zodi_model_size()} bytes in the library but "
            f"{MODEL_DTYPE.itemsize} in the packer"
        )
        raise RuntimeError(msg)
    _lib = lib
    return lib


def check(status: int, what: str) -> None:
    if status != 0:
        err = load().zodi_last_error().decode(errors="replace")
        msg = f"{what} failed (status {status}): {err}"
        raise RuntimeError(msg)


def host_ptr(arr: np.ndarray) -> int:
    return arr.ctypes.data


def device_view(host_address: int) -> int:
    """Device address of page-locked host memory (raises for anything else)."""
    out = ct.c_void_p()
    check(load().zodi_host_device_pointer(host_address, ct.byref(out)), "zodi_host_device_pointer")
    return int(out.value)
