"""Minimal physical-unit layer for the zodiacal emission front end.

The reference front end takes ``astropy.units.Quantity`` objects for the
wavelength/frequency ``x`` and for explicit observer positions, and returns a
``Quantity`` in MJy/sr (zodipy/model.py:36-45, 119-126, 202-203). astropy is
not part of this runtime, so this module provides the small subset those call
sites use: length, frequency, time, temperature, flux-density and solid-angle
units, the ``spectral()`` equivalency between wavelength and frequency, and the
``UnitConversionError`` exception type the reference raises.

A ``Quantity`` is an ``ndarray`` subclass that carries a ``Unit``. Arithmetic
between two quantities is not unit-aware; the front end always converts to
plain values with ``to_value`` before doing numerics.
"""

from __future__ import annotations

from typing import Iterable

import numpy as np

__all__ = [
    "Unit",
    "Quantity",
    "UnitsError",
    "UnitConversionError",
    "spectral",
    "m",
    "cm",
    "mm",
    "um",
    "micron",
    "nm",
    "km",
    "AU",
    "Hz",
    "kHz",
    "MHz",
    "GHz",
    "THz",
    "s",
    "K",
    "Jy",
    "MJy",
    "sr",
    "deg",
    "rad",
    "dimensionless",
]

# Speed of light [m/s], exact in SI.
C_SI = 299792458.0
# Astronomical unit [m], exact (IAU 2012 Resolution B2).
AU_SI = 149597870700.0

# Base dimensions: length, time, mass, temperature, angle, solid angle.
_BASES = ("m", "s", "kg", "K", "rad", "sr")


class UnitsError(Exception):
    """Base class for unit errors."""


class UnitConversionError(UnitsError, ValueError):
    """Raised when a quantity cannot be converted to the requested unit."""


class Unit:
    """A scaled product of base dimensions.

    ``scale`` converts a value in this unit to the SI-like base unit of the
    same dimension (e.g. ``micron.scale == 1e-6``).
    """

    __slots__ = ("name", "scale", "dims")
    # Make ``ndarray * unit`` defer to ``Unit.__rmul__`` instead of broadcasting.
    __array_ufunc__ = None

    def __init__(self, name: str, scale: float, dims: Iterable[int]) -> None:
        self.name = name
        self.scale = float(scale)
        self.dims = tuple(int(d) for d in dims)

    # -- algebra ---------------------------------------------------------
    def __mul__(self, other):
        if isinstance(other, Unit):
            return Unit(
                f"{self.name} {other.name}",
                self.scale * other.scale,
                (a + b for a, b in zip(self.dims, other.dims)),
            )
        return Quantity(other, self)

    def __rmul__(self, other):
        if isinstance(other, Unit):  # pragma: no cover - handled by __mul__
            return other.__mul__(self)
        return Quantity(other, self)

    def __truediv__(self, other):
        if isinstance(other, Unit):
            return Unit(
                f"{self.name} / {other.name}",
                self.scale / other.scale,
                (a - b for a, b in zip(self.dims, other.dims)),
            )
        return NotImplemented

    def __pow__(self, power: int):
        return Unit(
            f"{self.name}^{power}", self.scale**power, (a * power for a in self.dims)
        )

    def __eq__(self, other) -> bool:
        if not isinstance(other, Unit):
            return False
        return self.dims == other.dims and np.isclose(self.scale, other.scale, rtol=1e-15, atol=0)

    def __hash__(self) -> int:
        return hash((self.dims, round(np.log10(self.scale), 12)))

    def __repr__(self) -> str:
        return f"Unit({self.name!r})"

    def __str__(self) -> str:
        return self.name

    @property
    def physical_type(self) -> str:
        return _PHYSICAL_TYPES.get(self.dims, "unknown")

    def is_equivalent(self, other: "Unit", equivalencies=None) -> bool:
        try:
            _conversion(self, other, equivalencies)
        except UnitConversionError:
            return False
        return True


_LENGTH = (1, 0, 0, 0, 0, 0)
_FREQ = (0, -1, 0, 0, 0, 0)
_TIME = (0, 1, 0, 0, 0, 0)
_PHYSICAL_TYPES = {
    _LENGTH: "length",
    _FREQ: "frequency",
    _TIME: "time",
    (0, 0, 0, 1, 0, 0): "temperature",
    (0, 0, 0, 0, 1, 0): "angle",
    (0, 0, 0, 0, 0, 1): "solid angle",
    (0, 0, 0, 0, 0, 0): "dimensionless",
}

dimensionless = Unit("", 1.0, (0,) * 6)
m = Unit("m", 1.0, _LENGTH)
cm = Unit("cm", 1e-2, _LENGTH)
mm = Unit("mm", 1e-3, _LENGTH)
um = Unit("micron", 1e-6, _LENGTH)
micron = um
nm = Unit("nm", 1e-9, _LENGTH)
km = Unit("km", 1e3, _LENGTH)
AU = Unit("AU", AU_SI, _LENGTH)
Hz = Unit("Hz", 1.0, _FREQ)
kHz = Unit("kHz", 1e3, _FREQ)
MHz = Unit("MHz", 1e6, _FREQ)
GHz = Unit("GHz", 1e9, _FREQ)
THz = Unit("THz", 1e12, _FREQ)
s = Unit("s", 1.0, _TIME)
K = Unit("K", 1.0, (0, 0, 0, 1, 0, 0))
# Jansky: 1e-26 W m^-2 Hz^-1 = 1e-26 kg s^-2.
Jy = Unit("Jy", 1e-26, (0, -2, 1, 0, 0, 0))
MJy = Unit("MJy", 1e-20, (0, -2, 1, 0, 0, 0))
rad = Unit("rad", 1.0, (0, 0, 0, 0, 1, 0))
deg = Unit("deg", np.pi / 180.0, (0, 0, 0, 0, 1, 0))
sr = Unit("sr", 1.0, (0, 0, 0, 0, 0, 1))


def spectral():
    """Wavelength <-> frequency equivalency (``nu = c / lambda``)."""
    return [(_LENGTH, _FREQ, lambda lam: C_SI / lam, lambda nu: C_SI / nu)]


def _conversion(src: Unit, dst: Unit, equivalencies):
    """Return a function mapping values in ``src`` to values in ``dst``."""
    if src.dims == dst.dims:
        if src.scale == dst.scale:
            return lambda v: v
        factor = src.scale / dst.scale
        return lambda v: v * factor
    for a, b, fwd, back in equivalencies or ():
        if src.dims == a and dst.dims == b:
            return lambda v: fwd(v * src.scale) / dst.scale
        if src.dims == b and dst.dims == a:
            return lambda v: back(v * src.scale) / dst.scale
    msg = f"'{src}' ({src.physical_type}) and '{dst}' ({dst.physical_type}) are not convertible"
    raise UnitConversionError(msg)


class Quantity(np.ndarray):
    """An array of values with a unit attached."""

    def __new__(cls, value, unit: Unit | None = None, dtype=np.float64):
        if isinstance(value, Quantity) and unit is None:
            unit = value.unit
        if isinstance(value, Quantity):
            if unit is not None and not (value.unit == unit):
                value = value.to_value(unit)
            else:
                value = value.view(np.ndarray)
        if isinstance(value, (list, tuple)) and any(isinstance(v, Quantity) for v in value):
            first = next(v for v in value if isinstance(v, Quantity))
            unit = unit or first.unit
            value = [v.to_value(unit) if isinstance(v, Quantity) else v for v in value]
        arr = np.asarray(value, dtype=dtype).view(cls)
        arr.unit = unit if unit is not None else dimensionless
        return arr

    def __array_finalize__(self, obj) -> None:
        self.unit = getattr(obj, "unit", dimensionless)

    def __reduce__(self):
        state = super().__reduce__()
        return (state[0], state[1], (*state[2], self.unit))

    def __setstate__(self, state) -> None:
        self.unit = state[-1]
        super().__setstate__(state[:-1])

    @property
    def value(self):
        v = self.view(np.ndarray)
        return v[()] if v.ndim == 0 else v

    @property
    def isscalar(self) -> bool:
        return self.ndim == 0

    def to_value(self, unit: Unit | None = None, equivalencies=None):
        if unit is None:
            return self.value
        if not isinstance(unit, Unit):
            msg = f"target unit must be a Unit, got {type(unit).__name__}"
            raise TypeError(msg)
        out = _conversion(self.unit, unit, equivalencies)(self.view(np.ndarray).astype(np.float64))
        return out[()] if np.ndim(out) == 0 else out

    def to(self, unit: Unit, equivalencies=None) -> "Quantity":
        return Quantity(self.to_value(unit, equivalencies), unit)

    def __repr__(self) -> str:
        return f"<Quantity {self.view(np.ndarray)!r} {self.unit}>"

    def __str__(self) -> str:
        return f"{self.view(np.ndarray)} {self.unit}"

    def __getitem__(self, key):
        out = super().__getitem__(key)
        if isinstance(out, Quantity):
            return out
        return Quantity(out, self.unit)

    def __iter__(self):
        for v in self.view(np.ndarray):
            yield Quantity(v, self.unit)

    def __lshift__(self, unit: Unit) -> "Quantity":
        return Quantity(self.view(np.ndarray), unit)
