"""Accuracy of the kernels' atan2/acos (zodipy_amd/csrc/zodi_trig.hpp), host build.

The header compiles for the host with g++ (same arithmetic; the host forms the
reciprocal with a division where the device refines a hardware estimate).
Errors are measured in ulps against mpmath (exact to the precision used) and
against the C library on many points; IEEE special cases follow NumPy.
"""

from __future__ import annotations

import ctypes as ct
import shutil
import subprocess
from pathlib import Path

import mpmath
import numpy as np
import pytest

CSRC = Path(__file__).resolve().parents[1] / "zodipy_amd" / "csrc"

HARNESS = r"""
#include "zodi_trig.hpp"
extern "C" void eval_atan2(const double* y, const double* x, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = zodi_trig::atan2_fast(y[i], x[i]);
}
extern "C" void eval_acos(const double* x, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = zodi_trig::acos_fast(x[i]);
}
extern "C" void eval_sincos(const double* x, double* s, double* c, long n) {
  for (long i = 0; i < n; ++i) zodi_trig::sincos_moderate(x[i], s + i, c + i);
}
"""

P = ct.POINTER(ct.c_double)


@pytest.fixture(scope="module")
def trig(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("trig")
    src = d / "harness.cpp"
    src.write_text(HARNESS)
    lib = d / "libtrig.so"
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC",
                    f"-I{CSRC}", str(src), "-o", str(lib)], check=True)
    h = ct.CDLL(str(lib))

    def at2(y, x):
        y = np.ascontiguousarray(y, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(y)
        h.eval_atan2(y.ctypes.data_as(P), x.ctypes.data_as(P), out.ctypes.data_as(P), y.size)
        return out

    def ac(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        h.eval_acos(x.ctypes.data_as(P), out.ctypes.data_as(P), x.size)
        return out

    return at2, ac


def ulps(got, exact):
    return np.abs(got - exact) / np.spacing(np.abs(exact))


def test_atan2_against_mpmath(trig):
    at2, _ = trig
    rng = np.random.default_rng(0)
    n = 2000
    y = rng.normal(size=n) * 10.0 ** rng.uniform(-8, 3, n)
    x = rng.normal(size=n) * 10.0 ** rng.uniform(-8, 3, n)
    got = at2(y, x)
    mpmath.mp.dps = 40
    exact = np.array([float(mpmath.atan2(mpmath.mpf(a), mpmath.mpf(b))) for a, b in zip(y, x)])
    assert ulps(got, exact).max() <= 2.0


def test_atan2_against_libm_many_points(trig):
    at2, _ = trig
    rng = np.random.default_rng(1)
    n = 400_000
    y = rng.normal(size=n) * 10.0 ** rng.uniform(-6, 2, n)
    x = rng.normal(size=n) * 10.0 ** rng.uniform(-6, 2, n)
    err = ulps(at2(y, x), np.arctan2(y, x))
    assert err.max() <= 3.0  # libm itself is within an ulp
    assert err.mean() < 0.3


def test_atan2_special_values(trig):
    at2, _ = trig
    cases = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (0.0, 0.0), (-0.0, -0.0),
             (0.0, -0.0), (1.0, 0.0), (-1.0, 0.0), (1.0, 1.0), (-2.0, -2.0), (3.0, -1e-300)]
    y, x = np.array(cases).T
    np.testing.assert_array_equal(np.signbit(at2(y, x)), np.signbit(np.arctan2(y, x)))
    np.testing.assert_allclose(at2(y, x), np.arctan2(y, x), rtol=3e-16, atol=0)
    assert np.isnan(at2([np.nan, 1.0], [1.0, np.nan])).all()


def test_acos_against_mpmath(trig):
    _, ac = trig
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1, 1, 1500), 1 - 10.0 ** rng.uniform(-16, 0, 250),
                        -1 + 10.0 ** rng.uniform(-16, 0, 250), [0.5, -0.5, 0.0, -0.0]])
    got = ac(x)
    mpmath.mp.dps = 40
    exact = np.array([float(mpmath.acos(mpmath.mpf(v))) for v in x])
    assert ulps(got, exact).max() <= 1.0


def test_acos_special_values(trig):
    _, ac = trig
    assert ac([1.0])[0] == 0.0
    assert ac([-1.0])[0] == np.pi
    assert np.isnan(ac([np.nan, 1.5, -1.5])).all()


@pytest.fixture(scope="module")
def sincos(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("sincos")
    src = d / "harness.cpp"
    src.write_text(HARNESS)
    lib = d / "libsincos.so"
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC",
                    f"-I{CSRC}", str(src), "-o", str(lib)], check=True)
    h = ct.CDLL(str(lib))

    def sc(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        s = np.empty_like(x)
        c = np.empty_like(x)
        h.eval_sincos(x.ctypes.data_as(P), s.ctypes.data_as(P), c.ctypes.data_as(P), x.size)
        return s, c

    return sc


def test_sincos_moderate_against_mpmath(sincos):
    """The HEALPix kernel's sin/cos on [0, 2 pi): within 1.5 ulp of the exact values
    (and within 2.3e-16 absolute), including the quadrant boundaries and the
    doubles next to multiples of pi/2, where the reduced argument is tiny."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0.0, 2 * np.pi, 3000), rng.uniform(-64.0, 64.0, 500),
                        np.arange(9) * (np.pi / 4), np.nextafter(np.arange(1, 5) * (np.pi / 2), 0),
                        np.nextafter(np.arange(1, 5) * (np.pi / 2), 10), [0.0, 1e-300, 2 * np.pi]])
    s, c = sincos(x)
    mpmath.mp.dps = 50
    es = np.array([float(mpmath.sin(mpmath.mpf(v))) for v in x])
    ec = np.array([float(mpmath.cos(mpmath.mpf(v))) for v in x])
    assert ulps(s, es).max() <= 1.5
    assert ulps(c, ec).max() <= 1.5
    assert np.abs(s - es).max() <= 2.3e-16 and np.abs(c - ec).max() <= 2.3e-16
    assert s[-3] == 0.0 and c[-3] == 1.0 and s[-2] == 1e-300


def test_sincos_moderate_against_libm_many_points(sincos):
    x = np.random.default_rng(4).uniform(0.0, 2 * np.pi, 1_000_000)
    s, c = sincos(x)
    assert np.abs(s - np.sin(x)).max() <= 2.3e-16
    assert np.abs(c - np.cos(x)).max() <= 2.3e-16
