"""``Model``: the drop-in front end of the zodiacal emission integral.

Same constructor and ``evaluate`` signature, validation, exception types,
registry names, parameter dictionaries and MJy/sr output as the reference's
``zodipy.Model`` (zodipy/model.py:33-333). The line-of-sight integral itself
(zodipy/model.py:253-277) runs in the fused HIP kernel; there is no CPU path.

Differences forced by the runtime (astropy is not installed):

* ``x`` and explicit ``obspos`` are ``zodipy_amd.units.Quantity`` values
  (``25 * units.micron``, ``[0.9, 0.1, 0.0] * units.AU``);
* pointings are ``zodipy_amd.coords.SkyCoord`` objects (ecliptic, ICRS or
  galactic lon/lat with an obstime);
* without astropy the Earth position must be given with ``earthpos=``
  ((3,) for one instant, (3, N) for time-ordered pointings);
* ``nprocesses`` is accepted and ignored: the GPU grid replaces the process pool.

``evaluate_vectors`` exposes the kernel-level interface directly (ecliptic
unit vectors, observer and Earth positions) and can return a device tensor.
"""

from __future__ import annotations

import numpy as np
import torch

from zodipy_amd import coords, ephemeris, pointing, units
from zodipy_amd.component import ComponentLabel
from zodipy_amd.model_registry import model_registry
from zodipy_amd.packing import PackedModel, pack_model
from zodipy_amd.runtime import OUT_COMPS, OUT_SUM, DeviceModel, Launch
from zodipy_amd.spectral import normalize_weights, planck_table, unpack_spectral

MJY_SR = units.MJy / units.sr


class Model:
    """Zodiacal light model evaluated on an MI355X GPU."""

    def __init__(
        self,
        x,
        *,
        weights=None,
        name: str = "dirbe",
        gauss_quad_degree: int = 50,
        extrapolate: bool = False,
        ephemeris: str = "builtin",
        device="cuda",
    ) -> None:
        try:
            if not x.isscalar and weights is None:
                msg = "Bandpass weights must be provided for non-scalar `x`."
                raise ValueError(msg)
        except AttributeError as error:
            msg = "The input 'x' must be a Quantity (e.g. 25 * zodipy_amd.units.micron)."
            raise TypeError(msg) from error
        if x.isscalar and weights is not None:
            msg = "Bandpass weights should not be provided for scalar `x`."
            raise ValueError(msg)

        self._ipd_model = model_registry.get_model(name)
        if not extrapolate and not self._ipd_model.is_valid_at(x):
            msg = (
                "The requested frequencies are outside the valid range of the model. "
                "If this was intended, set the extrapolate argument to True."
            )
            raise ValueError(msg)

        if weights is not None:
            weights = np.asarray(weights)
            if x.size != weights.size:
                msg = "Number of wavelengths and weights must be the same in the bandpass."
                raise ValueError(msg)

        self._x = x
        self._extrapolate = extrapolate
        self._normalized_weights = normalize_weights(x, weights)
        self._b_nu_table = planck_table(x, self._normalized_weights)
        self._gl_x, self._gl_w = np.polynomial.legendre.leggauss(gauss_quad_degree)
        self._ephemeris = ephemeris
        self._device = device
        self._device_model: DeviceModel | None = None
        self._init_ipd_model_partials()

    # ------------------------------------------------------------------ setup
    def _init_ipd_model_partials(self) -> None:
        """Effective spectral parameters and the packed device block."""
        self._spectral = unpack_spectral(self._ipd_model, self._x, self._normalized_weights,
                                         self._extrapolate)
        self._packed: PackedModel = pack_model(self._ipd_model, self._spectral,
                                               self._b_nu_table, self._gl_x, self._gl_w)
        self._device_model = None

    @property
    def device_model(self) -> DeviceModel:
        if self._device_model is None:
            self._device_model = DeviceModel(self._packed, self._device)
        return self._device_model

    @property
    def ncomps(self) -> int:
        return self._ipd_model.ncomps

    @property
    def comps(self) -> list[ComponentLabel]:
        return list(self._ipd_model.comps)

    # ------------------------------------------------------------------ evaluate
    def evaluate(
        self,
        skycoord,
        *,
        obspos="earth",
        return_comps: bool = False,
        nprocesses: int = 1,
        earthpos=None,
    ) -> units.Quantity:
        """Return the zodiacal emission [MJy/sr] towards every coordinate of ``skycoord``.

        Follows zodipy/model.py:119-203 (validation, instantaneous vs
        time-ordered observers, component sum in model order). ``earthpos`` (and
        ``obspos``) may also be an ``ephemeris.BodyTrack``: for time-ordered data
        its spline is then evaluated per ray on the GPU.
        """
        try:
            if skycoord.obstime is None:
                msg = "The `obstime` attribute of the `SkyCoord` object is not set."
                raise ValueError(msg)
        except AttributeError as error:
            msg = "The input coordinates must be a SkyCoord object."
            raise TypeError(msg) from error

        obspos_isstr = isinstance(obspos, str)
        if not obspos_isstr and not isinstance(obspos, ephemeris.BodyTrack):
            if not isinstance(obspos, units.Quantity):
                msg = "The observer position is not a string or a Quantity."
                raise TypeError(msg)
            try:  # the unit before any position is looked up, as the reference fails
                obspos.to_value(units.AU)
            except units.UnitConversionError as error:
                msg = "The observer position must be in length units."
                raise units.UnitConversionError(msg) from error
            if obspos.ndim > 2 or (obspos.ndim == 2 and obspos.shape[0] != 3) or (
                obspos.ndim == 1 and obspos.shape[0] != 3
            ):
                msg = "The observer position must have shape (3,) or (3, ncoords)."
                raise ValueError(msg)
            if (obspos.ndim > 1 and skycoord.obstime.size != obspos.shape[-1]) or (
                obspos.ndim == 1 and skycoord.obstime.size != 1
            ):
                msg = "The number of obstime (ncoords) and obspos (3, ncoords) does not match."
                raise ValueError(msg)

        if skycoord.obstime.size > skycoord.size:
            msg = "The size of obstime must be either 1 or ncoords."
            raise ValueError(msg)
        instantaneous = skycoord.obstime.size == 1
        if not instantaneous and skycoord.obstime.size != skycoord.size:
            msg = "The size of obstime must be either 1 or ncoords."
            raise ValueError(msg)

        earth_xyz = self._earth_position(skycoord, earthpos, instantaneous)
        obs_xyz = self._observer_position(skycoord, obspos, earth_xyz, instantaneous)

        if hasattr(skycoord, "ecliptic_xyz"):
            # the frame rotation and cartesian conversion run on the GPU
            unit_vectors = pointing.lonlat_to_vectors(skycoord.lon, skycoord.lat,
                                                      frame=skycoord.frame, device=self._device)
        else:  # pragma: no cover - astropy SkyCoord
            unit_vectors = _astropy_ecliptic_xyz(skycoord)

        out = self.evaluate_vectors(unit_vectors, obspos=obs_xyz, earthpos=earth_xyz,
                                    return_comps=return_comps, per_ray=not instantaneous)
        return out

    def _earth_position(self, skycoord, earthpos, instantaneous: bool) -> np.ndarray:
        if isinstance(earthpos, ephemeris.BodyTrack):
            mjd = np.asarray(_mjd(skycoord.obstime), dtype=np.float64)
            if instantaneous:
                return earthpos(mjd.reshape(())).reshape(3)
            return earthpos.on_device(mjd, self._device)
        if earthpos is not None:
            e = earthpos.to_value(units.AU) if isinstance(earthpos, units.Quantity) else earthpos
            e = np.asarray(e, dtype=np.float64)
            expect = (3,) if instantaneous else (3, skycoord.size)
            if e.shape != expect and not (instantaneous and e.shape == (3, 1)):
                msg = f"earthpos must have shape {expect}, got {e.shape}"
                raise ValueError(msg)
            return e.reshape(3) if instantaneous else e
        mjd = np.asarray(_mjd(skycoord.obstime), dtype=np.float64)
        if instantaneous:
            return ephemeris.body_position("earth", mjd.reshape(()), self._ephemeris).reshape(3)
        grid = ephemeris.arrange_obstimes(mjd[0], mjd[-1])
        return ephemeris.interp_body_position("earth", mjd, grid, self._ephemeris)

    def _observer_position(self, skycoord, obspos, earth_xyz, instantaneous: bool) -> np.ndarray:
        if isinstance(obspos, ephemeris.BodyTrack):
            mjd = np.asarray(_mjd(skycoord.obstime), dtype=np.float64)
            if instantaneous:
                return obspos(mjd.reshape(())).reshape(3)
            return obspos.on_device(mjd, self._device)
        if isinstance(obspos, str):
            if obspos == "semb-l2":
                if isinstance(earth_xyz, torch.Tensor):
                    # the norm over the whole array, as zodipy/bodies.py:47 takes it
                    dist = torch.linalg.vector_norm(earth_xyz)
                    return earth_xyz / dist * (dist + ephemeris.MEAN_DIST_TO_L2)
                return ephemeris.get_semb_l2_pos(earth_xyz)
            if obspos == "earth":
                return earth_xyz
            mjd = np.asarray(_mjd(skycoord.obstime), dtype=np.float64)
            if instantaneous:
                return ephemeris.body_position(obspos, mjd.reshape(()),
                                               self._ephemeris).reshape(3)
            grid = ephemeris.arrange_obstimes(mjd[0], mjd[-1])
            return ephemeris.interp_body_position(obspos, mjd, grid, self._ephemeris)
        try:
            obs = np.asarray(obspos.to_value(units.AU), dtype=np.float64)
        except units.UnitConversionError as error:
            msg = "The observer position must be in length units."
            raise units.UnitConversionError(msg) from error
        return obs.reshape(3) if instantaneous else obs

    def evaluate_vectors(
        self,
        unit_vectors,
        *,
        obspos,
        earthpos,
        return_comps: bool = False,
        per_ray: bool | None = None,
        as_tensor: bool = False,
        launch: Launch | None = None,
    ):
        """Kernel-level evaluation.

        ``unit_vectors``: (3, N) ecliptic unit vectors (array or float64 tensor);
        ``obspos``/``earthpos``: heliocentric ecliptic positions in AU, (3,) for
        one instant or (3, N) per ray (then one cutoff test covers all of them,
        as in one reference ``evaluate`` call). Returns (ncomps, N) with
        ``return_comps`` else (N,), as a ``Quantity`` in MJy/sr, or as a device
        tensor with ``as_tensor=True``.
        """
        on_device = isinstance(obspos, torch.Tensor) and obspos.is_cuda
        obs = obspos if on_device else np.asarray(obspos, dtype=np.float64)
        earth = earthpos if isinstance(earthpos, torch.Tensor) and earthpos.is_cuda else \
            np.asarray(earthpos.cpu() if isinstance(earthpos, torch.Tensor) else earthpos,
                       dtype=np.float64)
        if per_ray is None:
            per_ray = obs.ndim == 2 and obs.shape[1] > 1
        if isinstance(unit_vectors, torch.Tensor):
            u = unit_vectors
        else:
            u = np.asarray(unit_vectors, dtype=np.float64)
            if u.ndim == 1:
                u = u.reshape(3, 1)
        n = u.shape[-1]
        if per_ray and on_device:
            if not (isinstance(earth, torch.Tensor) and earth.is_cuda):
                earth = torch.as_tensor(np.broadcast_to(np.asarray(earth).reshape(3, -1),
                                                        (3, n)).copy(), device=obs.device)
        elif per_ray:
            obs = np.broadcast_to(obs.reshape(3, -1), (3, n))
            earth = np.broadcast_to(earth.reshape(3, -1), (3, n))
        out_mode = OUT_COMPS if return_comps else OUT_SUM
        res = self.device_model.forward(u, obs, earth, per_ray=per_ray, out_mode=out_mode,
                                        launch=launch)
        if as_tensor:
            return res
        return units.Quantity(res.cpu().numpy(), MJY_SR)

    # ------------------------------------------------------------------ parameters
    def get_parameters(self) -> dict:
        """Return the model parameter dictionary (zodipy/model.py:303-309)."""
        return self._ipd_model.to_dict()

    def update_parameters(self, parameters: dict) -> None:
        """Rebuild the model from a parameter dictionary (zodipy/model.py:311-333)."""
        new = parameters.copy()
        new["comps"] = {}
        for key, value in parameters.items():
            if key == "comps":
                for comp_key, comp_value in value.items():
                    label = ComponentLabel(comp_key)
                    new["comps"][label] = type(self._ipd_model.comps[label])(**comp_value)
            elif isinstance(value, dict):
                new[key] = {ComponentLabel(k): v for k, v in value.items()}
        self._ipd_model = self._ipd_model.__class__(**new)
        self._init_ipd_model_partials()


def _mjd(obstime):
    if isinstance(obstime, coords.Time):
        return obstime.mjd
    return obstime.mjd  # astropy Time


def _astropy_ecliptic_xyz(skycoord) -> np.ndarray:  # pragma: no cover - needs astropy
    import astropy.coordinates as ac

    sc = skycoord.transform_to(ac.BarycentricMeanEcliptic)
    xyz = sc.cartesian.xyz.value
    return xyz[:, np.newaxis] if sc.isscalar else xyz.reshape(3, -1)
