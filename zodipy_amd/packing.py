"""Packing of a model into the device model block (``struct zodi_model``, include/zodi_hip.h).

The block is built once per ``Model`` construction or ``update_parameters``
call and holds everything the kernels read besides the rays themselves:

* component kinds and parameters, with the derived quantities the reference
  binds into its density partials (sin/cos of ``i`` and ``Omega``, radians,
  ``3 * n_0`` for the bands; zodipy/component.py:39-44, 87-89, 133-136,
  zodipy/number_density.py:101) computed with the same NumPy expressions;
* the effective spectral parameters (emissivity, albedo, ``1 - albedo`` per
  component; C1-C3, solar irradiance, phase normalisation or the RRM
  calibration);
* the line-of-sight cutoffs of each component (zodipy/line_of_sight.py:19-52),
  grouped into runs of consecutive components sharing cutoffs and temperature
  law, at most four per group;
* the Gauss-Legendre nodes/weights and the Planck table with its slopes.

``beyond_bits`` computes the per-observer "outside the cutoff sphere" masks
(zodipy/line_of_sight.py:71-73) on the host with the reference's expression
for the observer distance.
"""

from __future__ import annotations

import math

from dataclasses import dataclass

import numpy as np

from zodipy_amd.component import (
    Band,
    BroadBand,
    Cloud,
    Comet,
    ComponentLabel,
    Fan,
    Feature,
    FeatureRRM,
    Interstellar,
    NarrowBand,
    Ring,
    RingRRM,
    ZodiacalComponent,
)
from zodipy_amd.cutoffs import COMPONENT_CUTOFFS
from zodipy_amd.spectral import phase_normalization
from zodipy_amd.zodiacal_light_model import Integrand, ZodiacalLightModel

ABI_VERSION = 2
MAX_COMPS = 16
MAX_GROUPS = 16
MAX_NODES = 512
N_TEMPS = 100
COMP_SLOTS = 32
SHARED_SLOTS = 8
MAX_PARAMS = 256
MAX_SLOTS_PER_GROUP = 4

# Component kinds (enum in include/zodi_hip.h).
KIND_CLOUD, KIND_BAND, KIND_RING, KIND_FEATURE = 0, 1, 2, 3
KIND_FAN, KIND_COMET, KIND_INTERSTELLAR = 4, 5, 6
KIND_NARROW_BAND, KIND_BROAD_BAND, KIND_RING_RRM, KIND_FEATURE_RRM = 7, 8, 9, 10

CF_CENTERED, CF_SCATTER, CF_P_IS_4 = 1, 2, 4

P_X0, P_Y0, P_Z0, P_NX, P_NY, P_NZ, P_K0, P_D0 = 0, 1, 2, 3, 4, 5, 6, 13
P_EMISSIVITY, P_ALBEDO, P_ONE_MINUS_ALBEDO, P_SCALE = 20, 21, 22, 23
P_SIN_I, P_COS_I, P_SIN_OMEGA, P_COS_OMEGA = 24, 25, 26, 27
S_C1, S_C2, S_C3, S_SOLAR, S_PHASE_NORM, S_CALIBRATION, S_DD_DC3 = 0, 1, 2, 3, 4, 5, 6

MODEL_DTYPE = np.dtype(
    [
        ("abi_version", "<i4"),
        ("integrand", "<i4"),
        ("n_comps", "<i4"),
        ("n_groups", "<i4"),
        ("n_nodes", "<i4"),
        ("n_temps", "<i4"),
        ("flags", "<i4"),
        ("reserved", "<i4"),
        ("comp_kind", "<i4", (MAX_COMPS,)),
        ("comp_flags", "<u4", (MAX_COMPS,)),
        ("group_first", "<i4", (MAX_GROUPS,)),
        ("group_count", "<i4", (MAX_GROUPS,)),
        ("grad_comp", "<i4", (MAX_COMPS,)),
        ("grad_spec_a", "<i4", (MAX_COMPS,)),
        ("grad_spec_b", "<i4", (MAX_COMPS,)),
        ("grad_shared", "<i4", (8,)),
        ("n_params", "<i4"),
        ("reserved2", "<i4"),
        ("group_r_in", "<f8", (MAX_GROUPS,)),
        ("group_r_out", "<f8", (MAX_GROUPS,)),
        ("group_T0", "<f8", (MAX_GROUPS,)),
        ("group_delta", "<f8", (MAX_GROUPS,)),
        ("shared", "<f8", (SHARED_SLOTS,)),
        ("comp", "<f8", (MAX_COMPS, COMP_SLOTS)),
        ("gl_x", "<f8", (MAX_NODES,)),
        ("gl_w", "<f8", (MAX_NODES,)),
        ("tab_T", "<f8", (N_TEMPS,)),
        ("tab_B", "<f8", (N_TEMPS,)),
        ("tab_slope", "<f8", (N_TEMPS,)),
        ("tab_inv_step", "<f8"),
        ("pad_", "<f8", (7,)),
    ]
)


def _kind_of(comp: ZodiacalComponent) -> int:
    # Subclasses before their parents (RingRRM < Ring, FeatureRRM < Feature).
    for cls, kind in (
        (FeatureRRM, KIND_FEATURE_RRM), (RingRRM, KIND_RING_RRM), (Cloud, KIND_CLOUD),
        (Band, KIND_BAND), (Feature, KIND_FEATURE), (Ring, KIND_RING), (Fan, KIND_FAN),
        (Comet, KIND_COMET), (Interstellar, KIND_INTERSTELLAR),
        (NarrowBand, KIND_NARROW_BAND), (BroadBand, KIND_BROAD_BAND),
    ):
        if isinstance(comp, cls):
            return kind
    msg = f"component type {type(comp).__name__} has no device implementation"
    raise TypeError(msg)


def _derived_slots(comp: ZodiacalComponent, kind: int) -> list[float]:
    """Reciprocals and constants the kernels multiply by instead of dividing."""
    c = comp
    if kind == KIND_CLOUD:
        return [1.0 / (2 * c.mu), c.mu / 2]
    if kind == KIND_BAND:
        # the plane normal over delta_zeta (zeta/delta_zeta = |Z'|/R) and 1/delta_r^2
        inv = 1.0 / c.delta_zeta_rad
        return [inv, 1.0 / c.v, 1.0 / c.delta_r, c.sin_Omega_rad * c.sin_i_rad * inv,
                c.cos_Omega_rad * c.sin_i_rad * inv, c.cos_i_rad * inv, 1.0 / c.delta_r**2]
    if kind in (KIND_RING, KIND_RING_RRM):
        return [1.0 / c.sigma_r**2, 1.0 / c.sigma_z, 1.0 / c.sigma_r]
    if kind in (KIND_FEATURE, KIND_FEATURE_RRM):
        return [1.0 / c.sigma_r**2, 1.0 / c.sigma_z, 1.0 / c.sigma_theta_rad**2,
                1.0 / c.sigma_r, 1.0 / c.sigma_theta_rad, float(np.cos(c.theta_rad)),
                float(np.sin(c.theta_rad))]
    if kind in (KIND_FAN, KIND_COMET):
        return [1.0 / c.Z_0]
    if kind == KIND_NARROW_BAND:
        return [1.0 / c.R_outer]
    if kind == KIND_BROAD_BAND:
        return [1.0 / c.sigma_bb, 1.0 / c.R_outer]
    return []


def _kind_slots(comp: ZodiacalComponent, kind: int) -> list[float]:
    c = comp
    if kind == KIND_CLOUD:
        return [c.n_0, c.alpha, c.beta, c.gamma, c.mu]
    if kind == KIND_BAND:
        return [3 * c.n_0, c.delta_zeta_rad, c.p, c.v, c.delta_r]
    if kind == KIND_RING:
        return [c.n_0, c.R, c.sigma_r, c.sigma_z]
    if kind == KIND_FEATURE:
        return [c.n_0, c.R, c.sigma_r, c.sigma_z, c.theta_rad, c.sigma_theta_rad]
    if kind == KIND_FAN:
        return [c.Q, c.P, c.gamma, c.Z_0, c.R_outer]
    if kind == KIND_COMET:
        return [c.gamma, c.Z_0, c.P, c.amp, c.R_inner, c.R_outer]
    if kind == KIND_INTERSTELLAR:
        return [c.amp]
    if kind == KIND_NARROW_BAND:
        return [c.beta_nb, c.G, c.gamma, c.A, c.R_inner, c.R_outer]
    if kind == KIND_BROAD_BAND:
        return [c.beta_bb, c.sigma_bb, c.gamma, c.A, c.R_inner, c.R_outer]
    if kind == KIND_RING_RRM:
        return [c.n_0, c.R, c.sigma_r, c.sigma_z, c.A]
    if kind == KIND_FEATURE_RRM:
        return [c.n_0, c.R, c.sigma_r, c.sigma_z, c.theta_rad, c.sigma_theta_rad, c.A]
    raise AssertionError(kind)


@dataclass
class ParamLayout:
    """Positions of the model parameters in the gradient/parameter vector.

    Order: each component's init fields in declaration order (the order of the
    reference's parameter dictionary, zodipy/zodiacal_light_model.py:38-55), then
    Kelsall: T_0, delta, the effective emissivity and albedo of every component,
    solar irradiance, C1, C2, C3; RRM: T_0 and delta of every component, calibration.
    """

    names: list[str]
    comp: list[int]
    spec_a: list[int]
    spec_b: list[int]
    shared: list[int]

    @property
    def n_params(self) -> int:
        return len(self.names)


def param_layout(model: ZodiacalLightModel) -> ParamLayout:
    names: list[str] = []
    comp_off = []
    for label, comp in model.comps.items():
        comp_off.append(len(names))
        names += [f"comps.{label.value}.{k}" for k in comp.init_params()]
    labels = list(model.comps)
    if model.integrand == Integrand.KELSALL:
        shared = [len(names), len(names) + 1]
        names += ["T_0", "delta"]
        spec_a = [len(names) + c for c in range(len(labels))]
        names += [f"emissivity.{lab.value}" for lab in labels]
        spec_b = [len(names) + c for c in range(len(labels))]
        names += [f"albedo.{lab.value}" for lab in labels]
        shared += [len(names) + k for k in range(4)]
        names += ["solar_irradiance", "C1", "C2", "C3"]
    else:
        spec_a = [len(names) + c for c in range(len(labels))]
        names += [f"T_0.{lab.value}" for lab in labels]
        spec_b = [len(names) + c for c in range(len(labels))]
        names += [f"delta.{lab.value}" for lab in labels]
        shared = [len(names)]
        names += ["calibration"]
    return ParamLayout(names, comp_off, spec_a, spec_b, shared)


def param_vector(model: ZodiacalLightModel, spectral: dict) -> np.ndarray:
    """Current values of the parameter vector (see ``param_layout``)."""
    vals: list[float] = []
    for comp in model.comps.values():
        vals += [float(v) for v in comp.init_params().values()]
    labels = list(model.comps)
    sh = spectral["shared"]
    if model.integrand == Integrand.KELSALL:
        vals += [sh["T_0"], sh["delta"]]
        vals += [spectral["comp"][lab]["emissivity"] for lab in labels]
        vals += [spectral["comp"][lab]["albedo"] for lab in labels]
        vals += [sh["solar_irradiance"], sh["C1"], sh["C2"], sh["C3"]]
    else:
        vals += [spectral["comp"][lab]["T_0"] for lab in labels]
        vals += [spectral["comp"][lab]["delta"] for lab in labels]
        vals += [sh["calibration"]]
    return np.asarray(vals, dtype=np.float64)


@dataclass
class PackedModel:
    """Host copy of the device model block plus the metadata the front end needs."""

    block: np.ndarray  # shape (1,), dtype MODEL_DTYPE
    labels: list[ComponentLabel]
    group_of_comp: list[int]
    group_r_in: np.ndarray
    group_r_out: np.ndarray

    @property
    def n_comps(self) -> int:
        return len(self.labels)

    @property
    def n_groups(self) -> int:
        return len(self.group_r_in)

    def tobytes(self) -> bytes:
        return self.block.tobytes()

    def beyond_bits(self, obs_xyz: np.ndarray, any_over_all: bool) -> np.ndarray:
        """Cutoff masks for observers ``obs_xyz`` (3, n).

        Bit 2g: observer outside group g's start sphere; bit 2g+1: outside its
        stop sphere. With ``any_over_all`` a single mask is returned for the
        whole set (one reference ``evaluate`` call covers every observer in it).
        """
        obs = np.asarray(obs_xyz, dtype=np.float64).reshape(3, -1)
        if obs.shape[1] <= _FEW_OBSERVERS:
            # the same operations per observer in plain floats (x * x is x**2 exactly,
            # and both square roots are correctly rounded)
            r_obs = np.array([math.sqrt(x * x + y * y + z * z) for x, y, z in obs.T.tolist()])
        else:
            x_0, y_0, z_0 = obs
            r_obs = np.sqrt(x_0**2 + y_0**2 + z_0**2)
        return beyond_bits_for(r_obs, self.group_r_in, self.group_r_out, any_over_all)


# Up to this many observers the cutoff masks are formed in plain Python: a single
# call's array passes cost more than the arithmetic.
_FEW_OBSERVERS = 8


def beyond_bits_for(r_obs: np.ndarray, group_r_in, group_r_out, any_over_all: bool) -> np.ndarray:
    """Cutoff masks from observer distances to the Sun and a block's group radii.

    Bit 2g: r_obs > group_r_in[g]; bit 2g+1: r_obs > group_r_out[g] (NaN sets
    neither); with ``any_over_all`` the OR over all observers.
    """
    r_obs = np.asarray(r_obs, dtype=np.float64)
    if r_obs.size <= _FEW_OBSERVERS:
        radii = list(zip(np.asarray(group_r_in, dtype=np.float64).tolist(),
                         np.asarray(group_r_out, dtype=np.float64).tolist()))
        vals = []
        for r in r_obs.reshape(-1).tolist():
            b = 0
            shift = 0
            for a, c in radii:
                b |= ((r > a) << shift) | ((r > c) << (shift + 1))
                shift += 2
            vals.append(b)
        if any_over_all:
            acc = 0
            for b in vals:
                acc |= b
            return np.array([acc], dtype=np.uint32)
        return np.array(vals, dtype=np.uint32).reshape(r_obs.shape)
    bits = np.zeros(r_obs.shape, dtype=np.uint32)
    for g in range(len(group_r_in)):
        bits |= (r_obs > group_r_in[g]).astype(np.uint32) << np.uint32(2 * g)
        bits |= (r_obs > group_r_out[g]).astype(np.uint32) << np.uint32(2 * g + 1)
    if any_over_all:
        return np.bitwise_or.reduce(bits, keepdims=True).astype(np.uint32)
    return bits


def pack_model(model: ZodiacalLightModel, spectral: dict, b_nu_table: np.ndarray,
               gl_x: np.ndarray, gl_w: np.ndarray, force_scatter=(),
               merge_groups: bool = True) -> PackedModel:
    """Build the device block for ``model`` with effective ``spectral`` parameters.

    ``force_scatter``: labels of components whose scattering branch is kept even at
    albedo 0 (their albedo may change on the device; a zero albedo then adds an
    exact zero). ``merge_groups=False`` gives every component a group of its own,
    a structure that no parameter value can change (merged groups follow equal
    temperature laws, which RRM parameter vectors can separate).
    """
    labels = list(model.comps.keys())
    n_comps = len(labels)
    if not 1 <= n_comps <= MAX_COMPS:
        msg = f"models with {n_comps} components are not supported (1..{MAX_COMPS})"
        raise ValueError(msg)
    if not 1 <= len(gl_x) <= MAX_NODES:
        msg = f"gauss_quad_degree {len(gl_x)} outside [1, {MAX_NODES}]"
        raise ValueError(msg)
    integrand = model.integrand
    blk = np.zeros(1, dtype=MODEL_DTYPE)
    b = blk[0]
    b["abi_version"] = ABI_VERSION
    b["integrand"] = int(integrand)
    b["n_comps"] = n_comps
    b["n_nodes"] = len(gl_x)
    b["n_temps"] = N_TEMPS
    b["gl_x"][: len(gl_x)] = gl_x
    b["gl_w"][: len(gl_w)] = gl_w

    temps = np.asarray(b_nu_table[0], dtype=np.float64)
    bnu = np.asarray(b_nu_table[1], dtype=np.float64)
    if temps.shape != (N_TEMPS,) or bnu.shape != (N_TEMPS,):
        msg = f"Planck table must be (2, {N_TEMPS})"
        raise ValueError(msg)
    b["tab_T"] = temps
    b["tab_B"] = bnu
    slope = np.zeros(N_TEMPS)
    slope[:-1] = (bnu[1:] - bnu[:-1]) / (temps[1:] - temps[:-1])
    b["tab_slope"] = slope
    b["tab_inv_step"] = (N_TEMPS - 1) / (temps[-1] - temps[0])

    shared = spectral["shared"]
    if integrand == Integrand.KELSALL:
        C1, C2, C3 = shared["C1"], shared["C2"], shared["C3"]
        b["shared"][S_C1] = C1
        b["shared"][S_C2] = C2
        b["shared"][S_C3] = C3
        b["shared"][S_SOLAR] = shared["solar_irradiance"]
        # N is used by the forward pass only for components with albedo != 0, and by
        # the adjoint for d/d(albedo) of every component.
        b["shared"][S_PHASE_NORM] = phase_normalization(C1, C2, C3)
        e3 = np.exp(np.pi * C3)
        b["shared"][S_DD_DC3] = (np.pi * e3 * (C3**2 + 1) - (e3 + 1) * 2 * C3) / (C3**2 + 1) ** 2
    else:
        b["shared"][S_CALIBRATION] = shared["calibration"]

    temps_law = []
    for c, label in enumerate(labels):
        comp = model.comps[label]
        kind = _kind_of(comp)
        b["comp_kind"][c] = kind
        flags = 0
        if comp.x_0 == 0 and comp.y_0 == 0 and comp.z_0 == 0:
            flags |= CF_CENTERED
        slots = b["comp"][c]
        slots[P_X0], slots[P_Y0], slots[P_Z0] = comp.x_0, comp.y_0, comp.z_0
        slots[P_NX] = comp.sin_Omega_rad * comp.sin_i_rad
        slots[P_NY] = comp.cos_Omega_rad * comp.sin_i_rad
        slots[P_NZ] = comp.cos_i_rad
        slots[P_SIN_I], slots[P_COS_I] = comp.sin_i_rad, comp.cos_i_rad
        slots[P_SIN_OMEGA], slots[P_COS_OMEGA] = comp.sin_Omega_rad, comp.cos_Omega_rad
        ks = _kind_slots(comp, kind)
        slots[P_K0:P_K0 + len(ks)] = ks
        ds = _derived_slots(comp, kind)
        slots[P_D0:P_D0 + len(ds)] = ds
        if kind == KIND_BAND and comp.p == 4.0:
            flags |= CF_P_IS_4
        if integrand == Integrand.KELSALL:
            sp = spectral["comp"][label]
            slots[P_EMISSIVITY] = sp["emissivity"]
            slots[P_ALBEDO] = sp["albedo"]
            slots[P_ONE_MINUS_ALBEDO] = 1 - sp["albedo"]
            slots[P_SCALE] = (1 - sp["albedo"]) * sp["emissivity"]
            if sp["albedo"] != 0 or label in force_scatter:
                flags |= CF_SCATTER
            temps_law.append((float(shared["T_0"]), float(shared["delta"])))
        else:
            sp = spectral["comp"][label]
            temps_law.append((float(sp["T_0"]), float(sp["delta"])))
        b["comp_flags"][c] = flags

    # Group consecutive components with identical cutoffs and temperature law.
    groups: list[list[int]] = []
    keys = []
    for c, label in enumerate(labels):
        r_in, r_out = COMPONENT_CUTOFFS[label]
        key = (float(r_in), float(r_out), *temps_law[c])
        if (merge_groups and groups and keys[-1] == key
                and len(groups[-1]) < MAX_SLOTS_PER_GROUP):
            groups[-1].append(c)
        else:
            groups.append([c])
            keys.append(key)
    if len(groups) > MAX_GROUPS:
        msg = f"model needs {len(groups)} component groups (max {MAX_GROUPS})"
        raise ValueError(msg)
    b["n_groups"] = len(groups)
    group_of_comp = [0] * n_comps
    for g, (members, key) in enumerate(zip(groups, keys)):
        b["group_first"][g] = members[0]
        b["group_count"][g] = len(members)
        b["group_r_in"][g], b["group_r_out"][g] = key[0], key[1]
        b["group_T0"][g], b["group_delta"][g] = key[2], key[3]
        for c in members:
            group_of_comp[c] = g
    if any(f & CF_SCATTER for f in b["comp_flags"][:n_comps]):
        b["flags"] = 1
    lay = param_layout(model)
    if lay.n_params > MAX_PARAMS:
        msg = f"model has {lay.n_params} parameters (max {MAX_PARAMS})"
        raise ValueError(msg)
    b["grad_comp"][:n_comps] = lay.comp
    b["grad_spec_a"][:n_comps] = lay.spec_a
    b["grad_spec_b"][:n_comps] = lay.spec_b
    b["grad_shared"][: len(lay.shared)] = lay.shared
    b["n_params"] = lay.n_params
    return PackedModel(
        block=blk,
        labels=labels,
        group_of_comp=group_of_comp,
        group_r_in=np.array([k[0] for k in keys]),
        group_r_out=np.array([k[1] for k in keys]),
    )
