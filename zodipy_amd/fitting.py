"""Gradients of the line-of-sight integral for zodiacal-model fitting.

The reference evaluates the model only; it has no derivatives. Its parameter
dictionary (zodipy/model.py:303-333, zodipy/zodiacal_light_model.py:38-55)
fixes the order of the component parameters here. The parameter vector theta is

* every component's init fields in declaration order (x_0, y_0, z_0, i, Omega,
  then the kind's own fields; angles in degrees as in the dictionary),
* Kelsall models: T_0, delta, the effective emissivity and albedo of every
  component at the model's bandpass, solar irradiance, C1, C2, C3;
  RRM models: T_0 and delta of every component, calibration.

``ParameterSpace.names`` lists the entries. Gradients come from the
hand-written adjoint kernel (``zodi_adjoint`` in the HIP library), which
recomputes the integrand instead of storing node values:

* ``emission(theta, problem)`` -- a ``torch.autograd.Function``: forward is the
  line-of-sight integral, backward is one adjoint launch with the incoming
  gradient as weights;
* ``chi2(theta, problem, data, inv_var)`` -- one library call: one pass writes
  the model values and the per-ray Jacobian into the workspace and a second
  forms chi^2 and J^T g (or, with ``store_jacobian=False`` or a Jacobian over
  72 GiB, the forward kernel writes the model values and one adjoint launch
  forms chi^2 and d(chi^2)/d(theta)); backward scales the stored gradient;
* ``LineOfSightProblem.jacobian(theta)`` -- the model values and the per-ray
  Jacobian dI_r/dtheta_p themselves, for Gauss-Newton or Fisher-matrix work.

Supported models: every component layout. The registry layouts -- the Kelsall
layouts (cloud and three bands, optionally followed by the ring and the
feature) and the RRM model (packed one component per group, since a parameter
vector may give its components different temperature laws) -- run kernels
specialised for them; any other combination of component kinds runs one
generic kernel that takes a group's components one per pass over its nodes.
``SpectralTableFit`` fits the spectral tables themselves on the device, and
``BandpassFit`` the bandpass weights (with ``LineOfSightProblem.chi2_spectral_grad``:
the gradient with respect to the effective spectral parameters and the Planck
table from one pass over the nodes, ``zodi_spectral_vjp``; ``table_vjp`` gives
the Planck-table part alone).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from zodipy_amd import _lib
from zodipy_amd._lib import (
    ADJ_CHI2,
    ADJ_VJP_COMPS,
    ADJ_VJP_SUM,
    OBS_PER_RAY,
    OUT_COMPS,
    OUT_SUM,
    Launch,
)
from zodipy_amd.packing import (
    N_TEMPS,
    PackedModel,
    beyond_bits_for,
    pack_model,
    param_layout,
    param_vector,
)
from zodipy_amd.runtime import DeviceModel, _f64_rows, _stream
from zodipy_amd.zodiacal_light_model import Integrand

__all__ = ["ParameterSpace", "LineOfSightProblem", "SpectralTableFit", "BandpassFit", "emission",
           "chi2", "chi2_tables", "chi2_bandpass"]


class ParameterSpace:
    """Map between a parameter vector and the packed model block of one ``Model``.

    The bandpass, quadrature and Planck table of the model stay fixed; only the
    entries of theta change.
    """

    def __init__(self, model) -> None:
        self._ipd = model._ipd_model
        self._spectral = model._spectral
        self._b_nu_table = model._b_nu_table
        self._gl_x, self._gl_w = model._gl_x, model._gl_w
        self._x = model._x
        self._normalized_weights = model._normalized_weights
        self._extrapolate = model._extrapolate
        self.layout = param_layout(self._ipd)
        self.names: tuple[str, ...] = tuple(self.layout.names)
        self.theta0 = param_vector(self._ipd, self._spectral)

    @property
    def n_params(self) -> int:
        return len(self.names)

    def index(self, name: str) -> int:
        return self.names.index(name)

    def _check(self, theta) -> np.ndarray:
        if isinstance(theta, torch.Tensor):
            theta = theta.detach().cpu().numpy()
        theta = np.asarray(theta, dtype=np.float64)
        if theta.shape != (self.n_params,):
            msg = f"theta must have shape ({self.n_params},), got {theta.shape}"
            raise ValueError(msg)
        return theta

    def pack(self, theta, force_scatter=()) -> PackedModel:
        """Packed device block of the model at ``theta`` (``force_scatter``: see pack_model)."""
        theta = self._check(theta)
        lay = self.layout
        comps = {}
        for (label, comp), off in zip(self._ipd.comps.items(), lay.comp):
            keys = list(comp.init_params())
            comps[label] = type(comp)(**{k: float(theta[off + i]) for i, k in enumerate(keys)})
        ipd = dataclasses.replace(self._ipd, comps=comps)
        spec = {
            "comp": {lab: dict(v) for lab, v in self._spectral["comp"].items()},
            "shared": dict(self._spectral["shared"]),
        }
        labels = list(self._ipd.comps)
        sh = spec["shared"]
        if self._ipd.integrand == Integrand.KELSALL:
            sh["T_0"], sh["delta"] = float(theta[lay.shared[0]]), float(theta[lay.shared[1]])
            for c, lab in enumerate(labels):
                spec["comp"][lab]["emissivity"] = float(theta[lay.spec_a[c]])
                spec["comp"][lab]["albedo"] = float(theta[lay.spec_b[c]])
            for k, key in enumerate(("solar_irradiance", "C1", "C2", "C3")):
                sh[key] = float(theta[lay.shared[2 + k]])
        else:
            for c, lab in enumerate(labels):
                spec["comp"][lab]["T_0"] = float(theta[lay.spec_a[c]])
                spec["comp"][lab]["delta"] = float(theta[lay.spec_b[c]])
            sh["calibration"] = float(theta[lay.shared[0]])
        # Kelsall groups follow the fixed cutoffs and one shared temperature law;
        # RRM components carry their own, so each keeps a group of its own.
        return pack_model(ipd, spec, self._b_nu_table, self._gl_x, self._gl_w,
                          force_scatter=force_scatter,
                          merge_groups=self._ipd.integrand == Integrand.KELSALL)


class LineOfSightProblem:
    """Rays and observers resident on one GPU, evaluated for any parameter vector.

    ``unit_vectors``: (3, n) ecliptic unit vectors; ``obspos``/``earthpos``:
    heliocentric ecliptic positions [AU], (3,) for one observer, (3, n_seg) with
    ``seg_offsets`` (n_seg + 1,) for observers owning consecutive ray ranges, or
    (3, n) with ``per_ray=True`` (one reference call over time-ordered data).
    ``store_jacobian``: let chi^2 keep the per-ray Jacobian (n_params x n doubles)
    in its workspace, which saves the separate forward pass.
    """

    def __init__(self, model, unit_vectors, obspos, earthpos, seg_offsets=None,
                 per_ray: bool = False, store_jacobian: bool = True) -> None:
        self.space = ParameterSpace(model)
        base: DeviceModel = model.device_model
        self.device = base.device
        self.lib = _lib.load()
        self.u = _f64_rows(unit_vectors, 3, "unit vectors", self.device)
        self.n_rays = int(self.u.shape[1])
        obs = np.asarray(obspos.cpu() if isinstance(obspos, torch.Tensor) else obspos,
                         dtype=np.float64).reshape(3, -1)
        if not per_ray and seg_offsets is None:
            if obs.shape[1] != 1:
                msg = "several observers need seg_offsets or per_ray=True"
                raise ValueError(msg)
            seg_offsets = np.array([0, self.n_rays], dtype=np.int64)
        (self.seg, self.n_seg, self.obs, self.earth, beyond,
         self.mode) = base.prepare_observers(obs, earthpos, seg_offsets, per_ray)
        if self.seg is not None and int(self.seg[-1]) != self.n_rays:
            msg = f"seg_offsets[-1] = {int(self.seg[-1])} but there are {self.n_rays} rays"
            raise ValueError(msg)
        # Observer distances to the Sun: the cutoff bits depend on the group radii
        # of the block evaluated, and a parameter vector can regroup components
        # (a temperature law that no longer matches its neighbour's).
        o = self.obs
        r_obs = torch.sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])
        if self.mode == OBS_PER_RAY:  # one call over all rays: only the farthest matters
            r_obs = r_obs[~torch.isnan(r_obs)].max().reshape(1) if r_obs.numel() else r_obs
        self._r_obs = r_obs.cpu().numpy()
        self._beyond_cache = {self._group_key(base.packed.block): beyond}
        # Adjoint launch (None = library defaults); ``variant`` = components per pass
        # (1, 2, or 4: the Jacobian pass's default, up to two of them summing in LDS).
        self.launch: Launch | None = None
        self.store_jacobian = bool(store_jacobian)
        self._theta_key: bytes | None = None
        self._block: tuple[np.ndarray, torch.Tensor, PackedModel] | None = None
        self._workspace: torch.Tensor | None = None

    @property
    def n_params(self) -> int:
        return self.space.n_params

    @property
    def n_comps(self) -> int:
        return len(self.space._ipd.comps)

    def _model(self, theta):
        """(host block, device block) of the model at theta (cached for the last theta)."""
        th = self.space._check(theta)
        key = th.tobytes()
        if key != self._theta_key:
            packed = self.space.pack(th)
            host = np.ascontiguousarray(packed.block)
            dev = torch.from_numpy(host.view(np.uint8).copy()).to(self.device)
            self._block = (host, dev, packed)
            self._theta_key = key
        return self._block

    @staticmethod
    def _group_key(block: np.ndarray):
        b = block.reshape(-1)[0]
        ng = int(b["n_groups"])
        return ng, b["group_r_in"][:ng].tobytes(), b["group_r_out"][:ng].tobytes()

    def _beyond(self, host: np.ndarray) -> torch.Tensor:
        """Cutoff bits of the observers for the groups of block ``host`` (cached)."""
        key = self._group_key(host)
        if key not in self._beyond_cache:
            b = host.reshape(-1)[0]
            ng = int(b["n_groups"])
            bits = beyond_bits_for(self._r_obs, b["group_r_in"][:ng], b["group_r_out"][:ng],
                                   any_over_all=self.mode == OBS_PER_RAY)
            self._beyond_cache[key] = torch.from_numpy(bits.view(np.int32).copy()).to(
                self.device)
        return self._beyond_cache[key]

    def _ptr_args(self, host: np.ndarray):
        return (self.u.data_ptr(), self.n_rays,
                self.seg.data_ptr() if self.seg is not None else None, self.n_seg,
                self.obs.data_ptr(), self.earth.data_ptr(), self.mode,
                self._beyond(host).data_ptr())

    def forward(self, theta, return_comps: bool = False, blocks=None) -> torch.Tensor:
        """Integrated emission [MJy/sr]: (n_comps, n) or (n,).

        ``blocks``: (host block, device block) to evaluate instead of theta's.
        """
        host, dev = blocks if blocks is not None else self._model(theta)[:2]
        out_mode = OUT_COMPS if return_comps else OUT_SUM
        shape = (self.n_comps, self.n_rays) if return_comps else (self.n_rays,)
        out = torch.empty(shape, dtype=torch.float64, device=self.device)
        st = self.lib.zodi_forward(_lib.host_ptr(host), dev.data_ptr(), *self._ptr_args(host),
                                   out.data_ptr(), out_mode, None, _stream())
        _lib.check(st, "zodi_forward")
        return out

    def _ensure_workspace(self, host: np.ndarray, mode: int) -> tuple[torch.Tensor, int]:
        """Workspace for ``mode`` and the byte count to pass (which selects the chi^2 path)."""
        if mode == ADJ_CHI2 and not self.store_jacobian:
            mode = ADJ_VJP_SUM  # the recompute path's size
        need = int(self.lib.zodi_adjoint_workspace(_lib.host_ptr(host), self.n_rays, mode,
                                                   self.launch))
        if need < 0:
            err = self.lib.zodi_last_error().decode(errors="replace")
            msg = f"zodi_adjoint_workspace rejected the model or launch: {err}"
            raise RuntimeError(msg)
        n = (need + 7) // 8
        if self._workspace is None or self._workspace.numel() < n:
            self._workspace = None
            self._workspace = torch.empty(max(n, 1), dtype=torch.float64, device=self.device)
        return self._workspace, need

    def _adjoint(self, theta, mode: int, weights=None, target=None, inv_var=None,
                 inv_var_scalar: float = 1.0, model_out=None, blocks=None) -> torch.Tensor:
        host, dev = blocks if blocks is not None else self._model(theta)[:2]
        ws, need = self._ensure_workspace(host, mode)
        grad = torch.empty(self.n_params + 1, dtype=torch.float64, device=self.device)

        def ptr(t):
            return t.data_ptr() if t is not None else None

        st = self.lib.zodi_adjoint(
            _lib.host_ptr(host), dev.data_ptr(), *self._ptr_args(host), mode, ptr(weights),
            ptr(target), ptr(inv_var), float(inv_var_scalar), ptr(model_out), grad.data_ptr(),
            ws.data_ptr(), need, self.launch, _stream())
        _lib.check(st, "zodi_adjoint")
        return grad

    def jacobian(self, theta) -> tuple[torch.Tensor, torch.Tensor]:
        """(I (n,), J (n_params, n)) with J[p, r] = dI_r/dtheta_p, on the device."""
        host, dev, _ = self._model(theta)
        model = torch.empty(self.n_rays, dtype=torch.float64, device=self.device)
        jac = torch.empty((self.n_params, self.n_rays), dtype=torch.float64, device=self.device)
        st = self.lib.zodi_jacobian(_lib.host_ptr(host), dev.data_ptr(), *self._ptr_args(host),
                                    model.data_ptr(), jac.data_ptr(), self.n_rays, self.launch,
                                    _stream())
        _lib.check(st, "zodi_jacobian")
        return model, jac

    def normal_equations(self, theta, data, inv_var=None, chunk_rays: int = 1 << 22):
        """(chi^2, J^T W r, J^T W J) at theta for a least-squares fit.

        r = I - data, W = diag(inv_var) (None: 1; scalar or (n,)); J from
        ``jacobian``. The two products are FP64 GEMMs (the library GEMM on the
        GPU's matrix cores), accumulated over ray chunks so that only one
        scaled copy of a chunk of J exists at a time. d chi^2/d theta = 2 J^T W r.
        """
        I, J = self.jacobian(theta)
        d = self._rows(data, (self.n_rays,), "data")
        if inv_var is None:
            w = torch.ones_like(d)
        elif isinstance(inv_var, torch.Tensor) and inv_var.ndim == 1:
            w = self._rows(inv_var, (self.n_rays,), "inv_var")
        else:
            w = torch.full_like(d, float(inv_var))
        r = I - d
        wr = w * r
        chi2 = torch.dot(r, wr)
        g = torch.mv(J, wr)
        H = torch.zeros((self.n_params, self.n_params), dtype=torch.float64, device=self.device)
        for a in range(0, self.n_rays, chunk_rays):
            Jc = J[:, a:a + chunk_rays]
            H.addmm_(Jc * w[a:a + chunk_rays], Jc.t())
        return chi2, g, H

    def gauss_newton_step(self, theta, data, inv_var=None, damping: float = 0.0, free=None):
        """One Levenberg-Marquardt step: theta + delta with
        (H + damping diag(H)) delta = -J^T W r over the ``free`` entries
        (boolean mask or indices; default: every entry with diag(H) > 0).

        Returns (new theta as a NumPy array, chi^2 at theta).
        """
        th = self.space._check(theta)
        chi2, g, H = self.normal_equations(th, data, inv_var)
        diag = torch.diagonal(H)
        if free is None:
            mask = diag > 0
        else:
            mask = torch.zeros(self.n_params, dtype=torch.bool, device=self.device)
            mask[torch.as_tensor(free, device=self.device)] = True
            mask &= diag > 0
        idx = torch.nonzero(mask).reshape(-1)
        A = H[idx][:, idx] + damping * torch.diag(diag[idx])
        delta = torch.linalg.solve(A, -g[idx])
        out = th.copy()
        out[idx.cpu().numpy()] += delta.cpu().numpy()
        return out, float(chi2)

    def _rows(self, t, shape, name) -> torch.Tensor:
        t = torch.as_tensor(t, dtype=torch.float64, device=self.device)
        if tuple(t.shape) != shape:
            msg = f"{name} must have shape {shape}, got {tuple(t.shape)}"
            raise ValueError(msg)
        return t.contiguous()

    def _table_adjoint(self, theta, mode: int, weights=None, target=None, inv_var=None,
                       inv_var_scalar: float = 1.0, model_out=None, blocks=None,
                       eff: bool = False) -> torch.Tensor:
        """zodi_table_vjp (eff False) or zodi_spectral_vjp (eff True): the gradient row."""
        host, dev = blocks if blocks is not None else self._model(theta)[:2]
        name = "zodi_spectral_vjp" if eff else "zodi_table_vjp"
        need = int(getattr(self.lib, name + "_workspace")(_lib.host_ptr(host), self.n_rays,
                                                          self.launch))
        if need < 0:
            err = self.lib.zodi_last_error().decode(errors="replace")
            msg = f"{name}_workspace rejected the model or launch: {err}"
            raise RuntimeError(msg)
        ws = torch.empty(max((need + 7) // 8, 1), dtype=torch.float64, device=self.device)
        n_row = N_TEMPS + 1 + (self.n_params if eff else 0)
        grad = torch.empty(n_row, dtype=torch.float64, device=self.device)

        def ptr(t):
            return t.data_ptr() if t is not None else None

        st = getattr(self.lib, name)(
            _lib.host_ptr(host), dev.data_ptr(), *self._ptr_args(host), mode, ptr(weights),
            ptr(target), ptr(inv_var), float(inv_var_scalar), ptr(model_out), grad.data_ptr(),
            ws.data_ptr(), need, self.launch, _stream())
        _lib.check(st, name)
        return grad

    def table_vjp(self, theta, weights, return_comps: bool = False, blocks=None) -> torch.Tensor:
        """sum_r (sum_c) weights * dI/dB_k over the Planck table entries B_k (100,).

        B_k = B(T_k) is the bandpass-weighted blackbody emission the model block
        tabulates (zodipy/blackbody.py:33-49); entries past the block's table are 0.
        """
        if return_comps:
            w = self._rows(weights, (self.n_comps, self.n_rays), "weights")
            return self._table_adjoint(theta, ADJ_VJP_COMPS, weights=w, blocks=blocks)[:-1]
        w = self._rows(weights, (self.n_rays,), "weights")
        return self._table_adjoint(theta, ADJ_VJP_SUM, weights=w, blocks=blocks)[:-1]

    def chi2_table_grad(self, theta, data, inv_var=None, blocks=None):
        """(chi^2, d chi^2 / dB_k (100,)) for the Planck table entries, on the device."""
        d = self._rows(data, (self.n_rays,), "data")
        iv_t, iv_s = self._inv_var(inv_var)
        g = self._table_adjoint(theta, ADJ_CHI2, target=d, inv_var=iv_t, inv_var_scalar=iv_s,
                                blocks=blocks)
        return g[-1], g[:-1]

    def spectral_vjp(self, theta, weights, return_comps: bool = False, blocks=None):
        """(g_theta (n_params,), g_table (100,)): ``table_vjp`` together with the
        entries of ``vjp`` for the effective spectral parameters (emissivities,
        albedos, solar irradiance, C1..C3; RRM calibration), every other entry of
        g_theta 0, from one pass over the nodes (``zodi_spectral_vjp``)."""
        if return_comps:
            w = self._rows(weights, (self.n_comps, self.n_rays), "weights")
            g = self._table_adjoint(theta, ADJ_VJP_COMPS, weights=w, blocks=blocks, eff=True)
        else:
            w = self._rows(weights, (self.n_rays,), "weights")
            g = self._table_adjoint(theta, ADJ_VJP_SUM, weights=w, blocks=blocks, eff=True)
        n = self.n_params
        return g[:n], g[n:-1]

    def chi2_spectral_grad(self, theta, data, inv_var=None, return_model: bool = False,
                           blocks=None):
        """(chi^2, g_theta (n_params,), g_table (100,)[, model (n,)]): chi^2 and its
        gradient with respect to the effective spectral parameters (other entries
        of g_theta 0) and the Planck table entries, in one library call (the
        forward pass, then one pass of ``zodi_spectral_vjp``)."""
        d = self._rows(data, (self.n_rays,), "data")
        iv_t, iv_s = self._inv_var(inv_var)
        model = (torch.empty(self.n_rays, dtype=torch.float64, device=self.device)
                 if return_model else None)
        g = self._table_adjoint(theta, ADJ_CHI2, target=d, inv_var=iv_t, inv_var_scalar=iv_s,
                                model_out=model, blocks=blocks, eff=True)
        n = self.n_params
        if return_model:
            return g[-1], g[:n], g[n:-1], model
        return g[-1], g[:n], g[n:-1]

    def _inv_var(self, inv_var):
        """(per-ray tensor or None, scalar) from None, a scalar or an (n,) array."""
        if inv_var is None:
            return None, 1.0
        if np.ndim(inv_var) == 0 and not isinstance(inv_var, torch.Tensor):
            return None, float(inv_var)
        if isinstance(inv_var, torch.Tensor) and inv_var.ndim == 0:
            return None, float(inv_var)
        return self._rows(inv_var, (self.n_rays,), "inv_var"), 1.0

    def vjp(self, theta, weights, return_comps: bool = False) -> torch.Tensor:
        """sum_r (sum_c) weights * dI/dtheta for weights of the forward output's shape."""
        if return_comps:
            w = self._rows(weights, (self.n_comps, self.n_rays), "weights")
            return self._adjoint(theta, ADJ_VJP_COMPS, weights=w)[:-1]
        w = self._rows(weights, (self.n_rays,), "weights")
        return self._adjoint(theta, ADJ_VJP_SUM, weights=w)[:-1]

    def chi2_and_grad(self, theta, data, inv_var=None, return_model: bool = False,
                      blocks=None):
        """chi^2 = sum_r (I_r - data_r)^2 inv_var_r and its gradient, in one library call.

        ``inv_var``: None (1), a scalar, or (n,) per ray. Returns (chi2 (0-d
        tensor), grad (n_params,)[, model (n,)]) on the device. ``blocks``: as
        in ``forward``.
        """
        d = self._rows(data, (self.n_rays,), "data")
        iv_t, iv_s = self._inv_var(inv_var)
        model = (torch.empty(self.n_rays, dtype=torch.float64, device=self.device)
                 if return_model else None)
        g = self._adjoint(theta, ADJ_CHI2, target=d, inv_var=iv_t, inv_var_scalar=iv_s,
                          model_out=model, blocks=blocks)
        if return_model:
            return g[-1], g[:-1], model
        return g[-1], g[:-1]


class SpectralTableFit:
    """chi^2 fits of a model's spectral tables, with every step on the device.

    The effective spectral parameters are linear in the tables for a fixed
    bandpass (zodipy/unpack_model.py:23-173): effective = A @ table, A holding
    the interpolation and trapezoid weights. The fit's parameter vector is every
    entry of every tabulated spectral parameter (``names``: ``emissivity.cloud[0]``
    .. in ``spectral.spectral_tables`` order); everything else stays at
    ``theta`` (ParameterSpace's effective vector; default theta0). Per call,
    ``zodi_spectral_setup`` writes A @ table and its derived slots into the fit's
    own device block, the chi^2 mode of ``zodi_adjoint`` gives d chi^2/d(effective
    values), and the chain rule d chi^2/d table[t, j] = A[t, j] d chi^2/d eff_t
    runs on the device. (``chi2_spectral_grad`` gives the same entries from a
    forward pass and the spectral sums alone; with the Jacobian pass's registry
    kernels both cost about the same, so the fit keeps the full gradient.) Components with albedo tables are packed with their
    scattering branch, so an albedo may leave 0.
    """

    def __init__(self, problem: LineOfSightProblem, theta=None) -> None:
        from zodipy_amd.spectral import planck_coefficients, spectral_tables

        self.problem = problem
        space = problem.space
        self.theta = space.theta0.copy() if theta is None else space._check(theta)
        self.tables = spectral_tables(space._ipd, space._x, space._normalized_weights,
                                      space._extrapolate)
        if not self.tables:
            msg = "the model has no spectral tables"
            raise ValueError(msg)
        self.n_grid = int(self.tables[0].values.size)
        self.names = tuple(f"{t.name}[{j}]" for t in self.tables for j in range(self.n_grid))
        dev = problem.device
        self._A = torch.as_tensor(np.stack([t.weights for t in self.tables]), device=dev)
        self.values0 = np.concatenate([t.values for t in self.tables])
        self._targets = np.array([t.target for t in self.tables], dtype=np.int32)
        self._eff_index = torch.as_tensor([space.index(t.name) for t in self.tables],
                                          device=dev)
        scatter = tuple(lab for lab in space._ipd.comps
                        if any(t.name == f"albedo.{lab.value}" for t in self.tables))
        packed = space.pack(self.theta, force_scatter=scatter)
        self._host = np.ascontiguousarray(packed.block)
        self._dev = torch.from_numpy(self._host.view(np.uint8).copy()).to(dev)
        nu, q = planck_coefficients(space._x, space._normalized_weights)
        self._nu = torch.as_tensor(nu, device=dev)
        self._q = torch.as_tensor(q, device=dev)

    @property
    def n_params(self) -> int:
        return len(self.names)

    def set_tables(self, values, planck: bool = False) -> None:
        """Write the effective values of ``values`` (n_params,) into the device block.

        ``planck``: also rebuild the block's Planck table for the model's bandpass
        on the device (it does not depend on the tables).
        """
        v = torch.as_tensor(values, dtype=torch.float64, device=self.problem.device)
        if v.shape != (self.n_params,):
            msg = f"table values must have shape ({self.n_params},), got {tuple(v.shape)}"
            raise ValueError(msg)
        v = v.contiguous()
        n_freq = int(self._nu.numel()) if planck else 0
        st = self.problem.lib.zodi_spectral_setup(
            _lib.host_ptr(self._host), self._dev.data_ptr(), len(self.tables), self.n_grid,
            self._A.data_ptr(), v.data_ptr(), self._targets.ctypes.data, n_freq,
            self._nu.data_ptr() if planck else None, self._q.data_ptr() if planck else None,
            _stream())
        _lib.check(st, "zodi_spectral_setup")

    def device_block(self) -> np.ndarray:
        """The fit's device model block, copied to the host."""
        return self._dev.cpu().numpy().view(self._host.dtype).copy()

    def forward(self, values, return_comps: bool = False) -> torch.Tensor:
        self.set_tables(values)
        return self.problem.forward(None, return_comps, blocks=(self._host, self._dev))

    def chi2_and_grad(self, values, data, inv_var=None):
        """(chi^2, d chi^2 / d table entries (n_params,)) on the device."""
        self.set_tables(values)
        c2, g_eff = self.problem.chi2_and_grad(None, data, inv_var,
                                               blocks=(self._host, self._dev))
        g = self._A * g_eff[self._eff_index][:, None]
        return c2, g.reshape(-1)


class BandpassFit:
    """chi^2 fits of a model's bandpass weights at its fixed frequencies, on the device.

    The model normalises the weights w by their trapezoid over x
    (zodipy/model.py:82-93): wn = w / (c . w), c the trapezoid coefficients. The
    effective spectral parameters (zodipy/unpack_model.py:140-173) and the Planck
    table (zodipy/blackbody.py:33-49) are then linear in wn: eff = K @ wn, K's
    column i the unpacking of a unit weight at x_i (every interpolation,
    extrapolation and unit rule of the unpacking itself), and
    B_k = sum_i c_i wn_i B_nu_i(T_k). Per call ``zodi_spectral_setup`` writes both
    into the fit's own device block; chi^2, d chi^2/d eff and d chi^2/dB_k come
    from one ``zodi_spectral_vjp`` call (a forward pass, then one pass over the
    nodes), and the chain rule, through the normalisation, runs on the device:

        d chi^2/d wn_i = sum_t K[t, i] d chi^2/d eff_t + c_i sum_k B_nu_i(T_k) d chi^2/dB_k,
        d chi^2/d w_j = (d chi^2/d wn_j - c_j sum_i wn_i d chi^2/d wn_i) / (c . w).

    Everything else stays at ``theta`` (default theta0). The gradient is
    orthogonal to w (the normalisation removes the scale).
    """

    def __init__(self, problem: LineOfSightProblem, theta=None) -> None:
        from zodipy_amd.spectral import bandpass_maps

        self.problem = problem
        space = problem.space
        if space._normalized_weights is None or np.size(space._x.value) < 2:
            msg = "the model has no bandpass (weights over two or more frequencies)"
            raise ValueError(msg)
        self.theta = space.theta0.copy() if theta is None else space._check(theta)
        maps = bandpass_maps(space._ipd, space._x, space._extrapolate)
        self.n_freq = int(maps.c.size)
        if self.n_freq > 256:  # zodi_spectral_setup's largest grid
            msg = f"bandpass fits take up to 256 frequencies, the model has {self.n_freq}"
            raise ValueError(msg)
        self.weights0 = np.asarray(space._normalized_weights, dtype=np.float64).copy()
        self.names = maps.names
        dev = problem.device
        self._c = torch.as_tensor(maps.c, device=dev)
        self._K = torch.as_tensor(maps.K, device=dev)
        self._bnu = torch.as_tensor(maps.bnu, device=dev)  # (n_freq, 100)
        self._nu = torch.as_tensor(maps.nu, device=dev)
        self._targets = maps.targets
        self._eff_index = torch.as_tensor([space.index(n) for n in self.names], device=dev,
                                          dtype=torch.int64)
        scatter = tuple(lab for lab in space._ipd.comps
                        if f"albedo.{lab.value}" in self.names)
        packed = space.pack(self.theta, force_scatter=scatter)
        self._host = np.ascontiguousarray(packed.block)
        self._dev = torch.from_numpy(self._host.view(np.uint8).copy()).to(dev)

    def _normalised(self, weights) -> tuple[torch.Tensor, torch.Tensor]:
        w = torch.as_tensor(weights, dtype=torch.float64, device=self.problem.device)
        if w.shape != (self.n_freq,):
            msg = f"weights must have shape ({self.n_freq},), got {tuple(w.shape)}"
            raise ValueError(msg)
        S = torch.dot(self._c, w)
        return w, w / S

    def set_weights(self, weights) -> None:
        """Write the effective values and the Planck table of ``weights`` into the device block."""
        _, wn = self._normalised(weights)
        n_tab = len(self.names)
        tabs = wn.expand(max(n_tab, 1), -1).contiguous()
        q = (self._c * wn).contiguous()
        st = self.problem.lib.zodi_spectral_setup(
            _lib.host_ptr(self._host), self._dev.data_ptr(), n_tab, self.n_freq,
            self._K.data_ptr() if n_tab else None, tabs.data_ptr() if n_tab else None,
            self._targets.ctypes.data if n_tab else None, self.n_freq, self._nu.data_ptr(),
            q.data_ptr(), _stream())
        _lib.check(st, "zodi_spectral_setup")

    def forward(self, weights, return_comps: bool = False) -> torch.Tensor:
        self.set_weights(weights)
        return self.problem.forward(None, return_comps, blocks=(self._host, self._dev))

    def chi2_and_grad(self, weights, data, inv_var=None):
        """(chi^2, d chi^2 / d weights (n_freq,)) on the device."""
        w, wn = self._normalised(weights)
        self.set_weights(w)
        c2, g, g_table = self.problem.chi2_spectral_grad(None, data, inv_var,
                                                         blocks=(self._host, self._dev))
        d_wn = self._c * (self._bnu @ g_table)
        if len(self.names):
            d_wn = d_wn + self._K.t() @ g[self._eff_index]
        S = torch.dot(self._c, w)
        d_w = (d_wn - self._c * torch.dot(wn, d_wn)) / S
        return c2, d_w


class _Emission(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, problem, return_comps):
        ctx.problem = problem
        ctx.return_comps = return_comps
        ctx.save_for_backward(theta)
        return problem.forward(theta.detach(), return_comps)

    @staticmethod
    def backward(ctx, grad_out):
        (theta,) = ctx.saved_tensors
        g = ctx.problem.vjp(theta.detach(), grad_out.to(torch.float64).contiguous(),
                            ctx.return_comps)
        return g.to(device=theta.device, dtype=theta.dtype), None, None


class _Chi2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, problem, data, inv_var):
        value, grad = problem.chi2_and_grad(theta.detach(), data, inv_var)
        ctx.save_for_backward(grad)
        ctx.theta_meta = (theta.device, theta.dtype)
        return value

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        device, dtype = ctx.theta_meta
        g = grad * grad_out.to(device=grad.device, dtype=torch.float64)
        return g.to(device=device, dtype=dtype), None, None, None


def emission(theta: torch.Tensor, problem: LineOfSightProblem,
             return_comps: bool = False) -> torch.Tensor:
    """Differentiable line-of-sight emission [MJy/sr] at parameter vector theta."""
    return _Emission.apply(theta, problem, return_comps)


def chi2(theta: torch.Tensor, problem: LineOfSightProblem, data, inv_var=None) -> torch.Tensor:
    """Differentiable chi^2 = sum (I - data)^2 inv_var, forward and gradient fused."""
    return _Chi2.apply(theta, problem, data, inv_var)


class _Chi2Tables(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, fit, data, inv_var):
        value, grad = fit.chi2_and_grad(values.detach(), data, inv_var)
        ctx.save_for_backward(grad)
        ctx.meta = (values.device, values.dtype)
        return value

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        device, dtype = ctx.meta
        g = grad * grad_out.to(device=grad.device, dtype=torch.float64)
        return g.to(device=device, dtype=dtype), None, None, None


def chi2_tables(values: torch.Tensor, fit: SpectralTableFit, data, inv_var=None) -> torch.Tensor:
    """Differentiable chi^2 as a function of the spectral table entries."""
    return _Chi2Tables.apply(values, fit, data, inv_var)


def chi2_bandpass(weights: torch.Tensor, fit: BandpassFit, data, inv_var=None) -> torch.Tensor:
    """Differentiable chi^2 as a function of the bandpass weights."""
    return _Chi2Tables.apply(weights, fit, data, inv_var)
