"""Zodiacal component parameter schema.

Each component type is a dataclass whose *init* fields are the model
parameters; their names and order define the parameter dictionaries returned by
``Model.get_parameters`` and accepted by ``Model.update_parameters``, and hence
the order of the gradient vector. The schema follows zodipy/component.py:14-223
(field names, field order and the ``ComponentLabel`` values must be identical
for the parameter dictionaries to be interchangeable with the reference's).

Derived quantities (sines/cosines of the inclination ``i`` and ascending node
``Omega``, angles in radians, the offset column vector ``X_0``) are read-only
properties computed with the same NumPy expressions as the reference
(zodipy/component.py:39-44, 87-89, 133-136), so the packed device parameters
are bit-identical to the values the reference binds into its density partials.
"""

from __future__ import annotations

from dataclasses import dataclass, fields
from enum import Enum

import numpy as np


class ComponentLabel(Enum):
    """Component labels (values are the dictionary keys used in parameter dicts)."""

    CLOUD = "cloud"
    BAND1 = "band1"
    BAND2 = "band2"
    BAND3 = "band3"
    RING = "ring"
    FEATURE = "feature"
    FAN = "fan"
    COMET = "comet"
    INTERSTELLAR = "interstellar"
    INNER_NARROW_BAND = "inner_narrow_band"
    OUTER_NARROW_BAND = "outer_narrow_band"
    BROAD_BAND = "broad_band"
    RING_RRM = "ring_rrm"
    FEATURE_RRM = "feature_rrm"


@dataclass
class ZodiacalComponent:
    """Common geometry: offset of the symmetry centre and orientation of the symmetry plane.

    Args:
        x_0, y_0, z_0: offset from the Sun in heliocentric ecliptic coordinates [AU].
        i: inclination with respect to the ecliptic plane [deg].
        Omega: longitude of the ascending node [deg].
    """

    x_0: float
    y_0: float
    z_0: float
    i: float
    Omega: float

    @property
    def X_0(self) -> np.ndarray:
        return np.expand_dims([self.x_0, self.y_0, self.z_0], axis=-1)

    @property
    def sin_i_rad(self) -> float:
        return np.sin(np.radians(self.i))

    @property
    def cos_i_rad(self) -> float:
        return np.cos(np.radians(self.i))

    @property
    def sin_Omega_rad(self) -> float:
        return np.sin(np.radians(self.Omega))

    @property
    def cos_Omega_rad(self) -> float:
        return np.cos(np.radians(self.Omega))

    def init_params(self) -> dict:
        """Return the init fields (the model parameters) in declaration order."""
        return {f.name: getattr(self, f.name) for f in fields(self) if f.init}


@dataclass
class Cloud(ZodiacalComponent):
    """Kelsall diffuse cloud: ``n_0`` at 1 AU, radial index ``alpha``, vertical shape ``beta, gamma, mu``."""

    n_0: float
    alpha: float
    beta: float
    gamma: float
    mu: float


@dataclass
class Band(ZodiacalComponent):
    """Kelsall asteroidal dust band: ``n_0``, ``delta_zeta`` [deg], shape ``v, p``, inner cutoff ``delta_r``."""

    n_0: float
    delta_zeta: float
    v: float
    p: float
    delta_r: float

    @property
    def delta_zeta_rad(self) -> float:
        return np.radians(self.delta_zeta)


@dataclass
class Ring(ZodiacalComponent):
    """Kelsall circumsolar ring: ``n_0``, peak radius ``R``, dispersions ``sigma_r, sigma_z``."""

    n_0: float
    R: float
    sigma_r: float
    sigma_z: float


@dataclass
class Feature(ZodiacalComponent):
    """Kelsall Earth-trailing feature: ring parameters plus longitude ``theta`` and ``sigma_theta`` [deg]."""

    n_0: float
    R: float
    sigma_r: float
    sigma_z: float
    theta: float
    sigma_theta: float

    @property
    def theta_rad(self) -> float:
        return np.radians(self.theta)

    @property
    def sigma_theta_rad(self) -> float:
        return np.radians(self.sigma_theta)


@dataclass
class Fan(ZodiacalComponent):
    """Rowan-Robinson & May fan."""

    gamma: float
    Z_0: float
    Q: float
    P: float
    R_outer: float


@dataclass
class Comet(ZodiacalComponent):
    """Rowan-Robinson & May cometary component."""

    gamma: float
    Z_0: float
    P: float
    amp: float
    R_inner: float
    R_outer: float


@dataclass
class Interstellar(ZodiacalComponent):
    """Rowan-Robinson & May interstellar dust (constant density ``amp``)."""

    amp: float


@dataclass
class NarrowBand(ZodiacalComponent):
    """Rowan-Robinson & May narrow asteroidal band (latitudes in degrees)."""

    gamma: float
    A: float
    G: float
    R_inner: float
    R_outer: float
    beta_nb: float


@dataclass
class BroadBand(ZodiacalComponent):
    """Rowan-Robinson & May broad asteroidal band (latitudes in degrees)."""

    gamma: float
    A: float
    R_inner: float
    R_outer: float
    beta_bb: float
    sigma_bb: float


@dataclass
class RingRRM(Ring):
    """Rowan-Robinson & May ring: the Kelsall ring scaled by ``A``."""

    A: float


@dataclass
class FeatureRRM(Feature):
    """Rowan-Robinson & May feature: the Kelsall feature scaled by ``A``."""

    A: float
