"""FP64 autograd twin of the oracle's Kelsall and RRM line-of-sight integrals.

TEST INFRASTRUCTURE ONLY: the reference gradient the adjoint kernel is checked
against. Only tests may import it; the product path never calls it.

It restates the oracle functions of zodi_oracle.py (which follow
zodipy/brightness.py:21-83, zodipy/number_density.py:47-404,
zodipy/scattering.py:11-59, zodipy/blackbody.py:16-30 and
zodipy/line_of_sight.py:55-105) with torch float64 operations on the CPU, so
that torch.autograd differentiates the same arithmetic with respect to a
parameter vector theta. Two deliberate differences from the value path, both
invisible in the values:

* the scattering term a F Phi / R^2 is always formed (the reference adds it only
  for albedo != 0, where it is 0 anyway), so d/d(albedo) includes it at a = 0;
* 1 - exp(-x) of the bands uses -expm1(-x) (exact for the value, smooth).

theta follows the parameter order of the reference's parameter dictionary
(zodipy/zodiacal_light_model.py:38-55: component init fields, angles in degrees)
followed by T_0, delta, emissivity.<comp>, albedo.<comp>, solar_irradiance,
C1, C2, C3 (Kelsall) or T_0.<comp>, delta.<comp>, calibration (RRM) -- entries
are addressed by these names, so the twin does not depend on the product's
packing code. The RRM densities' masks (R_inner, R_outer) enter through
torch.where, so their derivative is 0, as the kernels take it.
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import zodi_oracle as O

_DEG = np.pi / 180.0


def _interp(T, xs, ys):
    """np.interp(T, xs, ys) with clamping, differentiable in T (zodipy/brightness.py:48)."""
    n = xs.shape[0]
    idx = torch.searchsorted(xs, T.detach().contiguous(), right=True) - 1
    idx = idx.clamp(0, n - 2)
    x0 = xs[idx]
    slope = (ys[idx + 1] - ys[idx]) / (xs[idx + 1] - xs[idx])
    Tc = torch.clamp(T, xs[0], xs[-1])
    return ys[idx] + slope * (Tc - x0)


def _tilted(X, p):
    X0 = torch.stack([p["x_0"], p["y_0"], p["z_0"]]).reshape(3, 1, 1)
    i = p["i"] * _DEG
    Om = p["Omega"] * _DEG
    Xc = X - X0
    R = torch.sqrt(Xc[0] ** 2 + Xc[1] ** 2 + Xc[2] ** 2)
    Z = (Xc[0] * torch.sin(Om) * torch.sin(i) - Xc[1] * torch.cos(Om) * torch.sin(i)
         + Xc[2] * torch.cos(i))
    return Xc, R, Z


def _density(kind, X, p, earth):
    """Component densities, zodipy/number_density.py:47-181."""
    Xc, R, Z = _tilted(X, p)
    if kind == "cloud":
        zeta = torch.abs(Z / R)
        g = torch.where(zeta < p["mu"], zeta**2 / (2 * p["mu"]), zeta - p["mu"] / 2)
        return p["n_0"] * R ** (-p["alpha"]) * torch.exp(-p["beta"] * g ** p["gamma"])
    if kind == "band":
        zeta = torch.abs(Z / R)
        zz = zeta / (p["delta_zeta"] * _DEG)
        t1 = 3 * p["n_0"] / R
        t2 = torch.exp(-(zz**6))
        t3 = 1 + zz ** p["p"] / p["v"]
        t4 = -torch.expm1(-((R / p["delta_r"]) ** 20))
        return t1 * t2 * t3 * t4
    if kind == "ring":
        return p["n_0"] * torch.exp(-((R - p["R"]) ** 2) / p["sigma_r"] ** 2
                                    - torch.abs(Z) / p["sigma_z"])
    if kind == "feature":
        E = earth.reshape(3, 1, -1) - torch.stack([p["x_0"], p["y_0"], p["z_0"]]).reshape(3, 1, 1)
        theta_comp = torch.atan2(Xc[1], Xc[0]) - torch.atan2(E[1], E[0])
        dth = theta_comp - p["theta"] * _DEG
        dth = torch.remainder(dth + np.pi, 2 * np.pi) - np.pi
        e = (R - p["R"]) ** 2 / p["sigma_r"] ** 2
        e = e + torch.abs(Z) / p["sigma_z"]
        e = e + dth**2 / (p["sigma_theta"] * _DEG) ** 2
        return p["n_0"] * torch.exp(-e)
    if kind in ("ring_rrm", "feature_rrm"):
        return p["A"] * _density(kind[:-4], X, p, earth)
    if kind == "interstellar":
        return p["amp"] * torch.ones_like(R)
    sin_beta = Z / R
    beta = torch.arcsin(sin_beta)
    if kind in ("fan", "comet"):
        # zodipy/number_density.py:184-256
        za = torch.abs(Z)
        eps = torch.where(za < p["Z_0"], 2 - za / p["Z_0"], torch.ones_like(za))
        f = torch.exp(-p["P"] * torch.sin(torch.abs(beta) ** eps))
        if kind == "fan":
            inside = R <= p["R_outer"]
            val = R ** (-p["gamma"]) * torch.cos(beta) ** p["Q"] * f
        else:
            inside = (R >= p["R_inner"]) & (R <= p["R_outer"])
            val = p["amp"] * R ** (-p["gamma"]) * f
        return torch.where(inside, val, torch.zeros_like(val))
    inside = (R >= p["R_inner"]) & (R <= p["R_outer"])
    beta_deg = beta * (180.0 / np.pi)
    if kind == "narrow_band":
        # zodipy/number_density.py:267-304
        babs = torch.abs(beta_deg)
        f = torch.where(babs < p["beta_nb"], torch.exp(p["G"] * (babs - p["beta_nb"])),
                        torch.zeros_like(babs))
    elif kind == "broad_band":
        # zodipy/number_density.py:307-342
        f = (torch.exp(-0.5 * ((beta_deg - p["beta_bb"]) / p["sigma_bb"]) ** 2)
             + torch.exp(-0.5 * ((beta_deg + p["beta_bb"]) / p["sigma_bb"]) ** 2))
    else:
        msg = f"no twin for component kind {kind}"
        raise ValueError(msg)
    val = p["A"] * (R / p["R_outer"]) ** (-p["gamma"]) * f
    return torch.where(inside, val, torch.zeros_like(val))


def _kind(label: str) -> str:
    if label == "cloud":
        return "cloud"
    if label.startswith("band"):
        return "band"
    if label.endswith("narrow_band"):
        return "narrow_band"
    return label


class KelsallTwin:
    """Twin of ``OracleModel`` for a Kelsall model, parameterised by theta.

    ``oracle_model`` supplies the fixed inputs (Planck table, quadrature and the
    registry model for the cutoffs); ``names`` labels the entries of theta.
    """

    def __init__(self, oracle_model: O.OracleModel, names) -> None:
        self.om = oracle_model
        self.names = list(names)
        self.labels = [lab.value for lab in oracle_model.ipd_model.comps]
        tab = oracle_model.b_nu_table
        self.tab_T = torch.as_tensor(tab[0], dtype=torch.float64)
        self.tab_B = torch.as_tensor(tab[1], dtype=torch.float64)

    def _params(self, theta):
        idx = {n: k for k, n in enumerate(self.names)}
        comps = {}
        for lab in self.labels:
            pre = f"comps.{lab}."
            comps[lab] = {n[len(pre):]: theta[k] for n, k in idx.items() if n.startswith(pre)}
        sh = {k: theta[idx[k]] for k in ("T_0", "delta", "solar_irradiance", "C1", "C2", "C3")}
        spec = {lab: (theta[idx[f"emissivity.{lab}"]], theta[idx[f"albedo.{lab}"]])
                for lab in self.labels}
        return comps, sh, spec

    def emission(self, theta, unit_vectors, obs_xyz, earth_xyz):
        """(ncomps, N) emission for one observer (3,) or per-ray observers (3, N)."""
        uvec = np.asarray(unit_vectors, dtype=np.float64)
        obs = np.asarray(obs_xyz, dtype=np.float64).reshape(3, -1)
        earth = np.asarray(earth_xyz, dtype=np.float64).reshape(3, -1)
        start, stop = O.get_line_of_sight_range(self.om.ipd_model.comps.keys(), uvec, obs)
        comps, sh, spec = self._params(theta)
        x = torch.as_tensor(self.om.points, dtype=torch.float64).reshape(-1, 1)
        w = torch.as_tensor(self.om.weights, dtype=torch.float64).reshape(-1, 1)
        U = torch.as_tensor(uvec).reshape(3, 1, -1)
        Xo = torch.as_tensor(obs).reshape(3, 1, -1)
        Ea = torch.as_tensor(earth)
        C1, C2, C3 = sh["C1"], sh["C2"], sh["C3"]
        N = 1 / (2 * np.pi * (2 * C1 + np.pi * C2 + (torch.exp(C3 * np.pi) + 1) / (C3**2 + 1)))
        out = []
        for label, ipd_label in zip(self.labels, self.om.ipd_model.comps):
            a = torch.as_tensor(start[ipd_label]).reshape(1, -1)
            b = torch.as_tensor(stop[ipd_label]).reshape(1, -1)
            R_los = 0.5 * (b - a) * x + 0.5 * (b + a)          # (K, N)
            X_los = R_los.unsqueeze(0) * U                       # (3, K, N)
            X = X_los + Xo
            R = torch.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
            T = sh["T_0"] * R ** (-sh["delta"])
            B = _interp(T.reshape(-1), self.tab_T, self.tab_B).reshape(T.shape)
            eps, alb = spec[label]
            em = (1 - alb) * (eps * B)
            cos_t = (X_los * X).sum(0) / (R_los * R)
            theta_s = torch.arccos(-torch.clamp(cos_t, -1, 1))
            phi = N * (C1 + C2 * theta_s + torch.exp(C3 * theta_s))
            em = em + alb * (sh["solar_irradiance"] / R**2) * phi
            n = _density(_kind(label), X, comps[label], Ea)
            f = em * n * 0.5 * (b - a)
            out.append((f * w).sum(0))
        return torch.stack(out)


class RRMTwin:
    """Twin of ``OracleModel`` for the RRM model, parameterised by theta.

    f = B(T_c) n_c calibration (stop - start)/2 with T_c = T_0,c R^-delta_c per
    component (zodipy/brightness.py:59-83).
    """

    def __init__(self, oracle_model: O.OracleModel, names) -> None:
        self.om = oracle_model
        self.names = list(names)
        self.labels = [lab.value for lab in oracle_model.ipd_model.comps]
        tab = oracle_model.b_nu_table
        self.tab_T = torch.as_tensor(tab[0], dtype=torch.float64)
        self.tab_B = torch.as_tensor(tab[1], dtype=torch.float64)

    def emission(self, theta, unit_vectors, obs_xyz, earth_xyz):
        """(ncomps, N) emission for one observer (3,) or per-ray observers (3, N)."""
        uvec = np.asarray(unit_vectors, dtype=np.float64)
        obs = np.asarray(obs_xyz, dtype=np.float64).reshape(3, -1)
        earth = np.asarray(earth_xyz, dtype=np.float64).reshape(3, -1)
        start, stop = O.get_line_of_sight_range(self.om.ipd_model.comps.keys(), uvec, obs)
        idx = {n: k for k, n in enumerate(self.names)}
        x = torch.as_tensor(self.om.points, dtype=torch.float64).reshape(-1, 1)
        w = torch.as_tensor(self.om.weights, dtype=torch.float64).reshape(-1, 1)
        U = torch.as_tensor(uvec).reshape(3, 1, -1)
        Xo = torch.as_tensor(obs).reshape(3, 1, -1)
        Ea = torch.as_tensor(earth)
        cal = theta[idx["calibration"]]
        out = []
        for label, ipd_label in zip(self.labels, self.om.ipd_model.comps):
            pre = f"comps.{label}."
            p = {n[len(pre):]: theta[k] for n, k in idx.items() if n.startswith(pre)}
            a = torch.as_tensor(start[ipd_label]).reshape(1, -1)
            b = torch.as_tensor(stop[ipd_label]).reshape(1, -1)
            R_los = 0.5 * (b - a) * x + 0.5 * (b + a)
            X = R_los.unsqueeze(0) * U + Xo
            R = torch.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
            T = theta[idx[f"T_0.{label}"]] * R ** (-theta[idx[f"delta.{label}"]])
            B = _interp(T.reshape(-1), self.tab_T, self.tab_B).reshape(T.shape)
            n = _density(_kind(label), X, p, Ea)
            f = B * n * cal * 0.5 * (b - a)
            out.append((f * w).sum(0))
        return torch.stack(out)
