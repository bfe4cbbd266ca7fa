"""Host-side logic: the C ABI library, the model block packer, the parameter
space and the reference's front-end behaviour (tests/test_model.py of zodipy).

CPU only. Calls into the library are restricted to entry points and argument
errors that return before any device work; those tests are skipped where a GPU
is present, since their device pointers are placeholders.
"""

from __future__ import annotations

import ctypes as ct
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from zodipy_amd import Model, _lib, healpix
from zodipy_amd import units as u
from zodipy_amd.component import ComponentLabel
from zodipy_amd.fitting import ParameterSpace
from zodipy_amd.packing import (
    ABI_VERSION,
    CF_CENTERED,
    CF_P_IS_4,
    CF_SCATTER,
    MODEL_DTYPE,
    P_EMISSIVITY,
    P_SCALE,
    pack_model,
    param_layout,
)
from zodipy_amd.zodiacal_light_model import ZodiacalLightModel, model_registry

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "zodi_hip.h"
NO_GPU = not torch.cuda.is_available()
cpu_only = pytest.mark.skipif(not NO_GPU, reason="placeholder device pointers: CPU-only check")

KELSALL_MODELS = ["dirbe", "planck13", "planck15", "planck18", "odegard"]


def header_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(zodi_\w+)\s*\(", text))


# --------------------------------------------------------------------- C ABI
def test_library_exports_every_declared_function():
    lib = _lib.load()
    declared = header_functions()
    assert declared == set(_lib.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_abi_version_and_model_size():
    lib = _lib.load()
    assert lib.zodi_abi_version() == ABI_VERSION == 2
    assert lib.zodi_model_size() == MODEL_DTYPE.itemsize == 15848


def _packed(name="dirbe", x=25 * u.micron, **kw):
    return Model(x, name=name, **kw)._packed


def _call_adjoint(lib, block, mode=0, launch=None, ws_bytes=1 << 20, n_rays=4):
    dummy = ct.c_void_p(0x1000)  # never dereferenced: every case fails host validation
    grad = ct.c_void_p(0x2000)
    return lib.zodi_adjoint(block.ctypes.data, dummy, dummy, n_rays, dummy, 1, dummy, dummy, 0,
                            dummy, mode, dummy, dummy, None, 1.0, None, grad, dummy, ws_bytes,
                            launch, None)


@cpu_only
def test_host_validation_errors():
    lib = _lib.load()
    blk = np.ascontiguousarray(_packed().block)
    dummy = ct.c_void_p(0x1000)
    # zodi_forward: null model, wrong ABI version, negative ray count, bad modes
    assert lib.zodi_forward(None, dummy, dummy, 1, dummy, 1, dummy, dummy, 0, dummy, dummy, 0,
                            None, None) == 1
    bad = blk.copy()
    bad["abi_version"] = 7
    assert lib.zodi_forward(bad.ctypes.data, dummy, dummy, 1, dummy, 1, dummy, dummy, 0, dummy,
                            dummy, 0, None, None) == 2
    assert b"abi_version" in lib.zodi_last_error()
    assert lib.zodi_forward(blk.ctypes.data, dummy, dummy, -1, dummy, 1, dummy, dummy, 0, dummy,
                            dummy, 0, None, None) == 1
    assert lib.zodi_forward(blk.ctypes.data, dummy, dummy, 1, dummy, 1, dummy, dummy, 5, dummy,
                            dummy, 0, None, None) == 1
    assert lib.zodi_forward(blk.ctypes.data, dummy, dummy, 1, dummy, 1, dummy, dummy, 0, dummy,
                            dummy, 9, None, None) == 1
    dst = ct.c_void_p(0x3000)
    assert lib.zodi_pix2vec(0, None, 0, 10, None, dst, None) == 1            # nside < 1
    assert lib.zodi_pix2vec(4, None, 150, 60, None, dst, None) == 1          # past 192 pixels
    assert lib.zodi_pix2vec(4, None, 0, 10, None, None, None) == 1           # no output
    assert lib.zodi_lonlat_to_vec(None, None, 5, 0, None, dst, None) == 1
    assert lib.zodi_lonlat_to_vec(None, None, 0, 0, None, None, None) == 0   # nothing to do
    assert lib.zodi_spline_positions(dst, 0, dst, 5, dst, None) == 1         # no interval
    assert lib.zodi_spline_positions(dst, 4, None, 5, dst, None) == 1
    assert lib.zodi_spline_positions(ct.c_void_p(0x3008), 4, dst, 5, dst, None) == 1  # align
    assert lib.zodi_max_radius(dst, -1, dst, None) == 1
    assert lib.zodi_max_radius(dst, 5, None, None) == 1
    assert lib.zodi_max_radius(None, 5, dst, None) == 1
    out = ct.c_void_p()
    pageable = np.zeros(16)
    assert lib.zodi_host_device_pointer(pageable.ctypes.data, ct.byref(out)) == 1
    assert lib.zodi_host_device_pointer(None, ct.byref(out)) == 1
    groups = blk.copy()
    groups["group_count"][0, 0] = 3  # groups no longer cover the components
    assert lib.zodi_forward(groups.ctypes.data, dummy, dummy, 1, dummy, 1, dummy, dummy, 0,
                            dummy, dummy, 0, None, None) == 2


@cpu_only
def test_adjoint_host_validation_errors(cloud_ring_model):
    lib = _lib.load()
    blk = np.ascontiguousarray(_packed().block)
    assert _call_adjoint(lib, blk, mode=7) == 1
    assert b"mode" in lib.zodi_last_error()
    assert _call_adjoint(lib, blk, launch=ct.byref(_lib.Launch(128, 0, 0, 0))) == 1
    assert _call_adjoint(lib, blk, launch=ct.byref(_lib.Launch(0, 3, 0, 0))) == 1
    assert _call_adjoint(lib, blk, launch=ct.byref(_lib.Launch(0, 0, 0, 5))) == 1
    assert _call_adjoint(lib, blk, ws_bytes=8) == 1
    assert b"workspace" in lib.zodi_last_error()
    corrupt = blk.copy()
    corrupt["grad_comp"][0, 1] = 75  # band1's ten slots would pass the end of 78 parameters
    assert _call_adjoint(lib, corrupt) == 2
    assert b"grad_comp" in lib.zodi_last_error()
    corrupt = blk.copy()
    corrupt["grad_shared"][0, 5] = -1
    assert _call_adjoint(lib, corrupt) == 2
    # a custom layout (the generic adjoint): the same validation before any launch
    odd = np.ascontiguousarray(_packed(cloud_ring_model).block)
    assert _call_adjoint(lib, odd, launch=ct.byref(_lib.Launch(0, 3, 0, 0))) == 1
    corrupt = odd.copy()
    corrupt["grad_spec_b"][0, 1] = 400
    assert _call_adjoint(lib, corrupt) == 2
    assert b"grad_spec" in lib.zodi_last_error()


@cpu_only
def test_adjoint_workspace_size():
    lib = _lib.load()
    blk = np.ascontiguousarray(_packed().block)
    # without a device the library assumes 256 CUs: grid = min(tiles / 4, 2 * 256) blocks
    n = 1_000_000
    ws = lib.zodi_adjoint_workspace
    for mode in (_lib.ADJ_VJP_SUM, _lib.ADJ_VJP_COMPS):
        assert ws(blk.ctypes.data, n, mode, None) == (512 * 79 + n) * 8
    assert ws(blk.ctypes.data, 64, 0, None) == (1 * 79 + 64) * 8
    # chi^2 stores the Jacobian [78][n], the J^T g partials of its ray chunks (about
    # n / 1024 rays each, a multiple of 256 in [256, 8192]), then I
    chunks = -(-n // 1024)  # 977 rays -> 1024 per chunk
    assert ws(blk.ctypes.data, 196608, _lib.ADJ_CHI2, None) == (78 * 196608 + 768 * 79
                                                                + 196608) * 8
    assert ws(blk.ctypes.data, n, _lib.ADJ_CHI2, None) == (78 * n + chunks * 79 + n) * 8
    # ... unless the launch asks for lanes sharing a ray, or J would pass 72 GiB
    two = ct.byref(_lib.Launch(0, 2, 0, 0))
    assert ws(blk.ctypes.data, n, _lib.ADJ_CHI2, two) == ws(blk.ctypes.data, n, 0, two)
    big = 130_000_000  # 78 x 1.3e8 x 8 B = 81 GB
    assert ws(blk.ctypes.data, big, _lib.ADJ_CHI2, None) == (512 * 79 + big) * 8
    assert ws(blk.ctypes.data, -1, 0, None) == -1
    assert ws(blk.ctypes.data, n, 0, ct.byref(_lib.Launch(0, 3, 0, 0))) == -1
    assert ws(blk.ctypes.data, n, 9, None) == -1


def _call_table(lib, block, spectral, mode=0, grad=0x2000, ws_bytes=1 << 20, n_rays=4):
    dummy = ct.c_void_p(0x1000)  # never dereferenced: every case fails host validation
    fn = lib.zodi_spectral_vjp if spectral else lib.zodi_table_vjp
    return fn(block.ctypes.data, dummy, dummy, n_rays, dummy, 1, dummy, dummy, 0, dummy, mode,
              dummy, dummy, None, 1.0, None, ct.c_void_p(grad) if grad else None, dummy,
              ws_bytes, None, None)


@cpu_only
def test_table_and_spectral_vjp_workspace_and_validation(cloud_ring_model, planck18_style):
    """Workspace: block partials (rows of the table's 100 entries and chi^2, plus the
    parameter vector for the spectral gradient), then I; the launch keeps as many
    blocks per CU as the kernel's waves (without a device the library assumes 256
    CUs): four for the registry layouts, three for the Kelsall layouts' spectral
    sums with scattering and for the kind switch."""
    lib = _lib.load()
    n = 1_000_000
    cases = [  # (block, table-only blocks per CU, spectral blocks per CU)
        (_packed(), 4, 3),  # DIRBE 25 um: F != 0, so the spectral pass sums scattering
        (_packed(planck18_style, 857 * u.GHz), 4, 4),  # DIRBE layout, no scattering
        (_packed("planck18", 857 * u.GHz), 4, 4),  # one group, no scattering
        (_packed("rrm-experimental"), 4, 4),
        (_packed(cloud_ring_model), 3, 3),
    ]
    for pk, per_cu_table, per_cu_spec in cases:
        blk = np.ascontiguousarray(pk.block)
        np_ = int(blk["n_params"][0])
        assert lib.zodi_table_vjp_workspace(blk.ctypes.data, n, None) == (
            per_cu_table * 256 * 101 + n) * 8
        assert lib.zodi_spectral_vjp_workspace(blk.ctypes.data, n, None) == (
            per_cu_spec * 256 * (101 + np_) + n) * 8
        # a small launch: one block of four waves
        assert lib.zodi_spectral_vjp_workspace(blk.ctypes.data, 64, None) == (
            (101 + np_) + 64) * 8
    blk = np.ascontiguousarray(_packed().block)
    for spectral in (False, True):
        assert _call_table(lib, blk, spectral, mode=7) == 1
        assert b"mode" in lib.zodi_last_error()
        assert _call_table(lib, blk, spectral, grad=0) == 1
        assert _call_table(lib, blk, spectral, ws_bytes=8) == 1
        assert b"workspace" in lib.zodi_last_error()
    # the spectral gradient writes the parameter rows: their layout must be valid
    corrupt = blk.copy()
    corrupt["grad_spec_a"][0, 2] = 300
    assert lib.zodi_spectral_vjp_workspace(corrupt.ctypes.data, n, None) == -1
    assert _call_table(lib, corrupt, True) == 2
    assert b"grad_spec" in lib.zodi_last_error()
    assert lib.zodi_table_vjp_workspace(corrupt.ctypes.data, n, None) > 0


@cpu_only
def test_spectral_setup_host_validation_errors():
    from zodipy_amd.spectral import SPEC_ALBEDO, SPEC_CALIBRATION, SPEC_EMISSIVITY, SPEC_SOLAR

    lib = _lib.load()
    blk = np.ascontiguousarray(_packed().block)  # DIRBE 25 um: albedo 0, no scattering branch
    d = ct.c_void_p(0x1000)  # never dereferenced: every case fails host validation

    def call(block, targets, n_grid=10, n_freq=0):
        t = np.asarray(targets, dtype=np.int32)
        return lib.zodi_spectral_setup(block.ctypes.data if block is not None else None, d,
                                       t.size, n_grid, d, d, t.ctypes.data, n_freq,
                                       d if n_freq else None, d if n_freq else None, None)

    assert call(None, [SPEC_EMISSIVITY]) == 1
    assert call(blk, [SPEC_EMISSIVITY | 6]) == 1  # DIRBE has components 0..5
    assert b"no such Kelsall component" in lib.zodi_last_error()
    assert call(blk, [SPEC_ALBEDO | 2]) == 2
    assert b"scattering branch" in lib.zodi_last_error()
    assert call(blk, [SPEC_CALIBRATION]) == 1
    assert call(blk, [0x999]) == 1
    assert call(blk, [SPEC_SOLAR] * 65) == 1
    assert call(blk, [SPEC_SOLAR], n_grid=0) == 1
    assert call(blk, [], n_freq=-1) == 1
    assert call(blk, [], n_freq=0) == 0  # nothing to write
    rrm = np.ascontiguousarray(_packed("rrm-experimental", 1.25 * u.micron,
                                       extrapolate=True).block)
    assert call(rrm, [SPEC_EMISSIVITY]) == 1
    m = Model(25 * u.micron)
    forced = pack_model(m._ipd_model, m._spectral, m._b_nu_table, m._gl_x, m._gl_w,
                        force_scatter=[ComponentLabel.BAND2]).block
    assert forced["comp_flags"][0, 2] & CF_SCATTER and forced["flags"][0] == 1
    assert not forced["comp_flags"][0, 1] & CF_SCATTER


@cpu_only
def test_jacobian_host_validation_errors(cloud_ring_model):
    lib = _lib.load()
    blk = np.ascontiguousarray(_packed().block)
    d = ct.c_void_p(0x1000)  # never dereferenced: every case fails host validation

    def call(block, n=4, out=d, jac=d, ld=4, launch=None):
        return lib.zodi_jacobian(block.ctypes.data if block is not None else None, d, d, n, d, 1,
                                 d, d, 0, d, out, jac, ld, launch, None)

    assert call(None) == 1
    assert call(blk, ld=3) == 1
    assert b"ld_jac" in lib.zodi_last_error()
    assert call(blk, jac=None) == 1
    assert call(blk, launch=ct.byref(_lib.Launch(0, 2, 0, 0))) == 1
    assert b"one lane per ray" in lib.zodi_last_error()
    odd = np.ascontiguousarray(_packed(cloud_ring_model).block)  # the generic adjoint
    assert call(odd, launch=ct.byref(_lib.Launch(0, 2, 0, 0))) == 1
    assert call(odd, n=0, out=None, jac=None, ld=0) == 0
    assert call(blk, n=0, out=None, jac=None, ld=0) == 0  # nothing to do


# --------------------------------------------------------------------- packer
def test_dirbe_block_groups_and_flags():
    pk = _packed()
    b = pk.block[0]
    assert b["n_comps"] == 6 and b["n_groups"] == 3
    assert list(b["group_count"][:3]) == [4, 1, 1]
    assert list(b["group_first"][:3]) == [0, 4, 5]
    flags = b["comp_flags"][:6]
    assert all(f & CF_CENTERED for f in flags[1:])  # bands, ring, feature sit at the Sun
    assert not flags[0] & CF_CENTERED  # the cloud is offset
    assert all(f & CF_P_IS_4 for f in flags[1:4])
    assert not any(f & CF_SCATTER for f in flags)  # albedo 0 at 25 um
    assert b["n_nodes"] == 50 and b["n_params"] == 78
    scat = _packed(x=1.25 * u.micron).block[0]
    assert all(f & CF_SCATTER for f in scat["comp_flags"][:6])


def test_beyond_bits_follow_cutoffs():
    pk = _packed()
    obs = np.array([[1.0, 1.25, 0.5], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    bits = pk.beyond_bits(obs, any_over_all=False)
    # groups: 0 = cloud+bands (eps, 5.2), 1 = ring (0.8, 1.2), 2 = feature (0.7, 1.3);
    # bit 2g: beyond group g's start sphere, bit 2g+1: beyond its stop sphere
    assert bits[0] == 1 | (1 << 2) | (1 << 4)             # beyond every start sphere
    assert bits[1] == 1 | (1 << 2) | (1 << 3) | (1 << 4)  # also beyond the ring's stop
    assert bits[2] == 1                                   # inside all but the eps sphere
    allb = pk.beyond_bits(obs, any_over_all=True)
    assert np.all(allb == (bits[0] | bits[1] | bits[2]))



def test_beyond_bits_few_observers_match_array_path():
    """A few observers take a plain-Python path; it gives the array path's masks,
    at the cutoff radii themselves, for NaN, infinite and zero distances, with and
    without the OR over observers, and through PackedModel.beyond_bits."""
    from zodipy_amd.packing import beyond_bits_for

    pk = _packed("rrm-experimental")
    radii = np.concatenate([pk.group_r_in, pk.group_r_out, [np.nan, np.inf, 0.0, 1.0]])
    rng = np.random.default_rng(5)
    for _ in range(300):
        n = int(rng.integers(1, 9))
        r = np.where(rng.uniform(size=n) < 0.5, rng.choice(radii, size=n),
                     rng.uniform(0.0, 6.0, size=n))
        many = np.concatenate([r, np.full(20, 3.0)])  # past the few-observer size
        want = beyond_bits_for(many, pk.group_r_in, pk.group_r_out, False)[:n]
        got = beyond_bits_for(r, pk.group_r_in, pk.group_r_out, False)
        assert got.dtype == np.uint32 and np.array_equal(got, want)
        got_any = beyond_bits_for(r, pk.group_r_in, pk.group_r_out, True)
        assert got_any.shape == (1,) and got_any[0] == np.bitwise_or.reduce(want)
    obs = rng.normal(size=(3, 5))
    many = np.concatenate([obs, rng.normal(size=(3, 20))], axis=1)
    assert np.array_equal(pk.beyond_bits(obs, False), pk.beyond_bits(many, False)[:5])

def test_param_layout_dirbe():
    lay = param_layout(model_registry.get_model("dirbe"))
    assert lay.n_params == 78
    assert lay.names[:5] == ["comps.cloud.x_0", "comps.cloud.y_0", "comps.cloud.z_0",
                             "comps.cloud.i", "comps.cloud.Omega"]
    assert lay.names[lay.comp[1] + 5] == "comps.band1.n_0"
    assert lay.names[lay.shared[0]] == "T_0" and lay.names[lay.shared[5]] == "C3"
    assert lay.names[lay.spec_a[4]] == "emissivity.ring"
    assert lay.names[lay.spec_b[5]] == "albedo.feature"
    assert len(set(lay.comp + lay.spec_a + lay.spec_b + lay.shared)) == 6 * 3 + 6


@pytest.mark.parametrize("name", KELSALL_MODELS)
def test_parameter_space_round_trip(name):
    x = 857 * u.GHz if name.startswith("planck") else 25 * u.micron
    model = Model(x, name=name, extrapolate=True)
    sp = ParameterSpace(model)
    assert sp.pack(sp.theta0).block.tobytes() == model._packed.block.tobytes()


def test_parameter_space_moves_only_its_slots():
    model = Model(25 * u.micron)
    sp = ParameterSpace(model)
    th = sp.theta0.copy()
    k = sp.index("emissivity.band2")
    th[k] *= 1.5
    a, b = model._packed.block[0], sp.pack(th).block[0]
    diff = np.argwhere(a["comp"] != b["comp"])
    assert sorted(map(tuple, diff)) == [(2, P_EMISSIVITY), (2, P_SCALE)]
    for f in MODEL_DTYPE.names:
        if f != "comp":
            assert np.array_equal(a[f], b[f]), f


def test_model_block_matches_header_layout():
    """Field offsets of the packer's dtype follow the struct in include/zodi_hip.h."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = text[text.index("typedef struct zodi_model {"):text.index("} zodi_model;")]
    fields = re.findall(r"\b(?:int32_t|uint32_t|double)\s+(\w+)", body)
    assert fields == list(MODEL_DTYPE.names)


# --------------------------------------------------------------------- front end
def test_x_input():
    with pytest.raises(u.UnitConversionError):
        Model(x=20 * u.s)
    with pytest.raises(TypeError):
        Model(x=20)
    with pytest.raises(ValueError):
        Model(np.linspace(20, 25, 10) * u.micron)



def test_invalid_sky_coord():
    """zodipy's tests/test_evaluate.py:98-116: the same exception types, raised
    before any position is looked up or any GPU work."""
    from zodipy_amd.coords import SkyCoord, Time

    m = Model(x=20 * u.micron)
    with pytest.raises(TypeError):
        m.evaluate(20)
    with pytest.raises(ValueError):
        m.evaluate(SkyCoord(20, 30, unit=u.deg))
    with pytest.raises(ValueError):
        m.evaluate(SkyCoord(20, 30, unit=u.deg,
                            obstime=Time(["2021-01-01T00:00:00", "2021-01-01T00:02:00"])))


def test_evaluate_invalid_obspos():
    """zodipy's tests/test_evaluate.py:119-132."""
    from zodipy_amd.coords import SkyCoord, Time

    m = Model(x=20 * u.micron)
    sc = SkyCoord(20, 30, unit=u.deg, obstime=Time("2021-01-01T00:00:00"))
    earth = np.array([1.0, 0.0, 0.0]) * u.AU
    with pytest.raises(ValueError):
        m.evaluate(sc, obspos="invalid", earthpos=earth)
    with pytest.raises(TypeError):
        m.evaluate(sc, obspos=[30, 40, 50])
    with pytest.raises(u.UnitConversionError):
        m.evaluate(sc, obspos=[30, 40, 50] * u.s)
    with pytest.raises(ValueError):
        m.evaluate(sc, obspos=[[30, 40, 50], [30, 40, 50]] * u.AU)


def test_evaluate_input_shape_errors():
    """The mismatched obstime/obspos/ncoords cases of zodipy's tests/test_evaluate.py:151-195."""
    from zodipy_amd.coords import SkyCoord, Time

    m = Model(x=20 * u.micron)
    t0 = Time("2021-01-01T00:00:00")
    obstimes = Time([float(t0.mjd) + i for i in range(4)], format="mjd")
    obsposes = np.array([[1.0, 0.1, 0.0], [0.9, 0.3, 0.0], [0.8, 0.5, 0.0],
                         [0.7, 0.7, 0.0]]).T * u.AU  # (3, 4)
    four = SkyCoord([20, 30, 40, 50], [30, 40, 30, 20], unit=u.deg, obstime=obstimes)
    one = SkyCoord(20, 30, unit=u.deg, obstime=t0)
    earth1, earth4 = np.array([1.0, 0.0, 0.0]) * u.AU, obsposes
    with pytest.raises(ValueError, match="obstime"):  # obstime > sky_coord
        m.evaluate(SkyCoord([20, 30], [30, 40], unit=u.deg, obstime=obstimes), earthpos=earth4)
    with pytest.raises(ValueError, match="obstime"):  # obspos > sky_coord / obstime
        m.evaluate(one, obspos=obsposes, earthpos=earth1)
    with pytest.raises(ValueError, match="obstime"):  # obstimes > obspos
        m.evaluate(four, obspos=[1, 2, 3] * u.AU, earthpos=earth4)
    with pytest.raises(ValueError, match="shape"):
        m.evaluate(four, obspos=[[1, -0.4, 0.2]] * u.AU, earthpos=earth4)
    with pytest.raises(ValueError, match="shape"):
        m.evaluate(four, obspos=obsposes[:2], earthpos=earth4)  # (2, 4)

def test_weights_input():
    bandpass = np.random.default_rng(0).random(10)
    with pytest.raises(ValueError):
        Model(20 * u.micron, weights=bandpass)
    with pytest.raises(ValueError):
        Model(np.linspace(20, 25, 15) * u.micron, weights=bandpass, name="dirbe")


def test_extrapolate_input():
    with pytest.raises(ValueError):
        Model(0.5 * u.micron, name="dirbe", extrapolate=False)
    with pytest.raises(ValueError):
        Model(400 * u.micron, name="dirbe", extrapolate=False)
    Model(20 * u.micron, extrapolate=True)
    x = np.linspace(40, 125, 10) * u.GHz
    bandpass = np.random.default_rng(1).random(10)
    with pytest.raises(ValueError):
        Model(x, weights=bandpass, name="planck18", extrapolate=False)
    Model(x, weights=bandpass, name="planck18", extrapolate=True)


def test_parameters_round_trip():
    model = Model(20 * u.micron, name="dirbe")
    params = model.get_parameters()
    assert isinstance(params, dict)
    before = model._packed.block.tobytes()
    model.update_parameters(params)
    assert model._packed.block.tobytes() == before
    params["comps"]["ring"]["sigma_r"] = 0.5
    model.update_parameters(params)
    assert model._packed.block.tobytes() != before


def test_registry():
    with pytest.raises(ValueError):
        model_registry.get_model("invalid")
    with pytest.raises(ValueError):
        model_registry.register_model("DIRBE", model_registry.get_model("DIRBE"))
    with pytest.raises(TypeError):
        model_registry.register_model("not a model", 2)
    for name in ["dirbe", "planck13", "planck15", "planck18", "odegard", "rrm-experimental"]:
        assert isinstance(model_registry.get_model(name), ZodiacalLightModel)


def test_pack_rejects_unknown_component_kind():
    import dataclasses

    dirbe = model_registry.get_model("dirbe")

    class Odd:
        pass

    odd = dataclasses.replace(dirbe, comps={ComponentLabel.CLOUD: Odd()})
    model = Model(25 * u.micron)
    with pytest.raises((TypeError, AttributeError)):
        pack_model(odd, model._spectral, model._b_nu_table, model._gl_x, model._gl_w)


def test_no_gpu_means_loud_failure():
    if not NO_GPU:
        pytest.skip("a GPU is present")
    model = Model(25 * u.micron)
    with pytest.raises(RuntimeError, match="GPU"):
        model.evaluate_vectors(np.array([1.0, 0.0, 0.0]), obspos=[1.0, 0, 0], earthpos=[1.0, 0, 0])


# --------------------------------------------------------------------- HEALPix
@pytest.mark.parametrize("nside", [1, 4, 16])
def test_pix2vec_ring_properties(nside):
    v = np.asarray(healpix.pix2vec(nside))
    npix = 12 * nside * nside
    assert v.shape == (3, npix)
    np.testing.assert_allclose(np.linalg.norm(v, axis=0), 1.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(v.mean(axis=1), 0.0, atol=1e-12)
    # RING scheme (Gorski et al. 2005, ApJ 622, 759, eqs. 4-5): 4 nside - 1 rings, z
    # non-increasing with the pixel index, ring i at z = 1 - i^2 / (3 nside^2) in the
    # polar caps (i < nside) and z = 4/3 - 2 i / (3 nside) in the equatorial belt
    assert np.all(np.diff(v[2]) <= 1e-15)
    i = np.arange(1, 4 * nside)
    z_ring = np.where(i < nside, 1 - i**2 / (3 * nside**2), 4 / 3 - 2 * i / (3 * nside))
    z_ring = np.where(i > 3 * nside, -(1 - (4 * nside - i) ** 2 / (3 * nside**2)), z_ring)
    zs = np.unique(np.round(v[2], 12))[::-1]
    np.testing.assert_allclose(zs, z_ring, rtol=0, atol=1e-12)


# Published RING pixel centres (Gorski et al. 2005, ApJ 622, 759, figs. 3-4), ring by
# ring from the north: (z, pixels in the ring, longitude of the ring's first pixel in
# units of pi/4, spacing in units of pi/4). nside 1 holds the 12 base pixels: the first
# ring's pixels sit at longitudes pi/4 + k pi/2 and the equatorial ring starts on the
# x axis. nside 2 alternates rings whose first pixel sits half a pixel east of
# longitude 0 with rings that start at longitude 0.
_HEALPIX_RINGS = {
    1: [(2 / 3, 4, 1, 2), (0.0, 4, 0, 2), (-2 / 3, 4, 1, 2)],
    2: [(11 / 12, 4, 1, 2), (2 / 3, 8, 0.5, 1), (1 / 3, 8, 0, 1), (0.0, 8, 0.5, 1),
        (-1 / 3, 8, 0, 1), (-2 / 3, 8, 0.5, 1), (-11 / 12, 4, 1, 2)],
}


@pytest.mark.parametrize("nside", [1, 2])
def test_pix2vec_published_pixel_centres(nside):
    """Every pixel centre of nside 1 and 2 against the published positions, so a half-pixel
    longitude phase error or a swapped ring fails (the zero-mean check above cannot see
    either)."""
    want = []
    for z, n, first, step in _HEALPIX_RINGS[nside]:
        phi = (first + step * np.arange(n)) * (np.pi / 4)
        s = np.sqrt(1.0 - z * z)
        want.append(np.stack([s * np.cos(phi), s * np.sin(phi), np.full(n, z)]))
    want = np.concatenate(want, axis=1)
    got = np.asarray(healpix.pix2vec(nside))
    assert got.shape == want.shape == (3, 12 * nside * nside)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


# --------------------------------------------------------------------- bench contract
def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_body_track_host_and_validation():
    from zodipy_amd.ephemeris import BodyTrack, arrange_obstimes

    grid = arrange_obstimes(59000.0, 59002.0)
    assert np.allclose(np.diff(grid), 1 / 24)
    pos = np.stack([np.cos(grid / 58.1), np.sin(grid / 58.1), 1e-4 * np.cos(grid)])
    tr = BodyTrack(grid, pos)
    # the spline interpolates its knots (the last one through the last interval)
    np.testing.assert_array_equal(tr(grid[:-1]), pos[:, :-1])
    np.testing.assert_allclose(tr(grid[-1]), pos[:, -1], rtol=1e-14, atol=1e-17)
    assert tr.c.shape == (4, grid.size - 1, 3)
    rec = tr.records
    assert rec.shape == (grid.size - 1, 16)
    np.testing.assert_array_equal(rec[:, 0], grid[:-1])
    np.testing.assert_array_equal(rec[:, 13], grid[1:])
    np.testing.assert_array_equal(rec[:, 1:4], tr.c[3])   # constant term
    np.testing.assert_array_equal(rec[:, 10:13], tr.c[0])  # cubic term
    assert not rec[:, 14:].any()
    with pytest.raises(ValueError):
        BodyTrack(grid[:1], pos[:, :1])
    with pytest.raises(ValueError):
        BodyTrack(grid[::-1], pos)


@pytest.mark.parametrize("n_obs", [2, 50000])
def test_prepare_observers_uploads(n_obs):
    """Observer data go to the device in one packed upload (small calls) or one
    upload per array (large ones); either way the views hold the arrays' values
    and dtypes. Checked with host memory standing in for the device."""
    import torch

    from zodipy_amd.runtime import DeviceModel

    model = Model(25 * u.micron)
    dm = DeviceModel.__new__(DeviceModel)
    dm.packed = model._packed
    dm.device = torch.device("cpu")
    rng = np.random.default_rng(3)
    obs = rng.uniform(-2, 2, (3, n_obs))
    earth = obs * 1.01
    seg = np.concatenate([[0], np.cumsum(rng.integers(0, 5, n_obs))])
    s, n, o, e, b, mode = dm.prepare_observers(obs, earth, seg, False)
    assert n == n_obs and mode == 0
    assert s.dtype == torch.int64 and np.array_equal(s.numpy(), seg)
    assert o.dtype == torch.float64 and np.array_equal(o.numpy(), obs)
    assert np.array_equal(e.numpy(), earth)
    want = model._packed.beyond_bits(obs, any_over_all=False).view(np.int32)
    assert b.dtype == torch.int32 and np.array_equal(b.numpy(), want)
    # per-ray observers from read-only broadcast views
    ro = np.broadcast_to(obs[:, :1], (3, 7))
    s, n, o, e, b, mode = dm.prepare_observers(ro, np.broadcast_to(earth[:, :1], (3, 7)), None,
                                               True)
    assert s is None and n == 0 and mode == 1
    assert np.array_equal(o.numpy(), np.asarray(ro)) and tuple(e.shape) == (3, 7)
    assert np.array_equal(b.numpy(),
                          model._packed.beyond_bits(np.asarray(ro), any_over_all=True)
                          .view(np.int32))
