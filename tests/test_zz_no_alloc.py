"""The allocation-free adjoint (zodi_adjoint_noalloc) against zodi_adjoint, through
the C ABI: same kernels and launch plan, so the gradients (and chi^2, and the model
values) agree bit for bit on zodi_adjoint's recompute path. Two problems in a row,
the second smaller, so leftovers of the first call in the shared per-GPU array
would show."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from zodipy_amd import Model, _lib
from zodipy_amd import units as u
from zodipy_amd._lib import ADJ_CHI2, ADJ_VJP_COMPS, ADJ_VJP_SUM, Launch

pytestmark = pytest.mark.gpu

ERR_TOO_LARGE = 4


def _dirs(n, seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    z = torch.rand(n, generator=g, dtype=torch.float64, device=device) * 2.0 - 1.0
    phi = torch.rand(n, generator=g, dtype=torch.float64, device=device) * (2.0 * np.pi)
    s = torch.sqrt(1.0 - z * z)
    return torch.stack([s * torch.cos(phi), s * torch.sin(phi), z]).contiguous()


def _problem(name, n, seed):
    from zodipy_amd.fitting import LineOfSightProblem

    model = Model(25 * u.micron, name=name)
    v = _dirs(n, seed, torch.device("cuda"))
    ph = 0.37 * seed
    obs = np.array([np.cos(ph), np.sin(ph), 0.01])
    return LineOfSightProblem(model, v, obs, obs)


def _both(prob, mode, weights=None, target=None, launch=None):
    lib = _lib.load()
    host, dev, _ = prob._model(prob.space.theta0)
    hp = _lib.host_ptr(host)
    # zodi_adjoint on its recompute path: a workspace of that path's size (the chi^2
    # mode takes the Jacobian path only with room for J)
    need = int(lib.zodi_adjoint_workspace(hp, prob.n_rays, ADJ_VJP_SUM, launch))
    assert need > 0
    ws = torch.empty((need + 7) // 8, dtype=torch.float64, device="cuda")
    np1 = prob.n_params + 1
    g_ref = torch.full((np1,), np.nan, dtype=torch.float64, device="cuda")
    g_na = torch.full((np1,), np.nan, dtype=torch.float64, device="cuda")
    i_ref = torch.full((prob.n_rays,), np.nan, dtype=torch.float64, device="cuda")
    i_na = torch.full((prob.n_rays,), np.nan, dtype=torch.float64, device="cuda")

    def ptr(t):
        return t.data_ptr() if t is not None else None

    stream = torch.cuda.current_stream().cuda_stream
    args = (hp, dev.data_ptr(), *prob._ptr_args(host), mode, ptr(weights), ptr(target), None,
            1.0e4)
    _lib.check(lib.zodi_adjoint(*args, i_ref.data_ptr(), g_ref.data_ptr(), ws.data_ptr(), need,
                                launch, stream), "zodi_adjoint")
    _lib.check(lib.zodi_adjoint_noalloc(*args, i_na.data_ptr(), g_na.data_ptr(), launch,
                                        stream), "zodi_adjoint_noalloc")
    torch.cuda.synchronize()
    return g_ref, g_na, i_ref, i_na


@pytest.mark.parametrize("name", ["dirbe", "rrm-experimental"])
def test_noalloc_adjoint_matches_adjoint_twice(gpu, name):
    for n, seed in ((40000, 3), (5000, 4)):  # the second call smaller
        prob = _problem(name, n, seed)
        print(f"zodi_adjoint_noalloc: {name}, {n} rays", flush=True)
        w = torch.as_tensor(np.random.default_rng(seed).normal(size=n), device="cuda")
        g_ref, g_na, _, _ = _both(prob, ADJ_VJP_SUM, weights=w)
        assert torch.isfinite(g_ref[:-1]).all() and torch.equal(g_ref, g_na)
        wc = torch.as_tensor(np.random.default_rng(seed + 1).normal(size=(prob.n_comps, n)),
                             device="cuda")
        g_ref, g_na, _, _ = _both(prob, ADJ_VJP_COMPS, weights=wc)
        assert torch.equal(g_ref, g_na)
        target = prob.forward(prob.space.theta0) * 1.01
        g_ref, g_na, i_ref, i_na = _both(prob, ADJ_CHI2, target=target)
        assert float(g_ref[-1]) > 0.0  # chi^2
        assert torch.equal(g_ref, g_na) and torch.equal(i_ref, i_na)


def test_noalloc_adjoint_rejects_launches_beyond_its_array(gpu):
    lib = _lib.load()
    prob = _problem("dirbe", 2_000_000, 5)
    host, dev, _ = prob._model(prob.space.theta0)
    grad = torch.empty(prob.n_params + 1, dtype=torch.float64, device="cuda")
    w = torch.ones(prob.n_rays, dtype=torch.float64, device="cuda")
    launch = Launch(0, 1, 64, 0)  # 64 blocks per CU: far more partial rows than 4 MiB hold
    st = lib.zodi_adjoint_noalloc(_lib.host_ptr(host), dev.data_ptr(), *prob._ptr_args(host),
                                  ADJ_VJP_SUM, w.data_ptr(), None, None, 1.0, None,
                                  grad.data_ptr(), launch, torch.cuda.current_stream().cuda_stream)
    assert st == ERR_TOO_LARGE
    assert "allocation-free" in lib.zodi_last_error().decode()
    # chi^2 mode without model_out is an argument error
    st = lib.zodi_adjoint_noalloc(_lib.host_ptr(host), dev.data_ptr(), *prob._ptr_args(host),
                                  ADJ_CHI2, None, w.data_ptr(), None, 1.0, None,
                                  grad.data_ptr(), None, torch.cuda.current_stream().cuda_stream)
    assert st == 1
