"""Accuracy of the kernels' table-driven exp and log (zodipy_amd/csrc/zodi_math.hpp).

The header compiles for the host with g++: the same operation sequences and the
same tables (FMA is exact-rounded on both sides; -ffp-contract=off keeps the host
from fusing more than the device source asks for). Errors are measured against
mpmath and against the C library:

* exp_fast_bounded and exp_fast within 1 ulp of exp over the range the kernels
  use and correctly rounded below ln2/512 in magnitude, exact 0 below -746
  (exp_fast's clamp, including -inf), NaN passing;
* log_pos within 3e-16 absolute plus half an ulp (its error enters the kernels
  only through exp(y log x), where the absolute error counts), log_nonneg(0) =
  -inf.
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
#include "zodi_math.hpp"
extern "C" void eval(int which, const double* x, double* out, long n) {
  for (long i = 0; i < n; ++i) {
    switch (which) {
      case 0: out[i] = zodi::exp_fast_bounded(x[i]); break;
      case 1: out[i] = zodi::exp_fast(x[i]); break;
      case 2: out[i] = zodi::log_pos(x[i]); break;
      default: out[i] = zodi::log_nonneg(x[i]); break;
    }
  }
}
"""

P = ct.POINTER(ct.c_double)


@pytest.fixture(scope="module")
def mathlib(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("math")
    src = d / "harness.cpp"
    src.write_text(HARNESS)
    lib = d / "libmath.so"
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC",
                    f"-I{CSRC}", str(src), "-o", str(lib)], check=True)
    h = ct.CDLL(str(lib))

    def f(which, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        h.eval(which, x.ctypes.data_as(P), out.ctypes.data_as(P), x.size)
        return out

    return f


def ulps(got, exact):
    return np.abs(got - exact) / np.spacing(np.abs(exact))


@pytest.mark.parametrize("which", [0, 1])
def test_exp_against_mpmath(mathlib, which):
    rng = np.random.default_rng(10 + which)
    x = np.concatenate([rng.uniform(-745.0, 50.0, 3000), rng.uniform(-1e-3, 1e-3, 500),
                        rng.uniform(-3.0, 3.0, 1500)])
    mpmath.mp.dps = 40
    exact = np.array([float(mpmath.exp(mpmath.mpf(v))) for v in x])
    got = mathlib(which, x)
    assert ulps(got, exact).max() <= 1.0


@pytest.mark.parametrize("which", [0, 1])
def test_exp_against_libm_many_points(mathlib, which):
    rng = np.random.default_rng(20 + which)
    x = rng.uniform(-700.0, 40.0, 1_000_000)
    err = ulps(mathlib(which, x), np.exp(x))
    assert err.max() <= 2.0  # libm itself is within an ulp
    assert err.mean() < 0.3


def test_exp_special_values(mathlib):
    x = np.array([0.0, -0.0, -746.0, -746.5, -1e4, -5e6, -np.inf, np.nan, 1e-20, -1e-20])
    b = mathlib(0, x[[0, 1, 2, 3, 4, 5, 7, 8, 9]])
    np.testing.assert_array_equal(b[:6], [1.0, 1.0, 0.0, 0.0, 0.0, 0.0])
    assert np.isnan(b[6]) and b[7] == 1.0 and b[8] == 1.0
    c = mathlib(1, x)
    np.testing.assert_array_equal(c[:7], [1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    assert np.isnan(c[7]) and c[8] == 1.0 and c[9] == 1.0
    # below ln2/512 in magnitude exp is correctly rounded, so 1 - exp(-x) (the band's
    # radial cut-off) is as accurate as a correctly rounded exp allows
    s = np.concatenate([10.0 ** np.linspace(-12, -3.2, 50), [1e-300, 5e-324]])
    mpmath.mp.dps = 40
    for sign in (-1.0, 1.0):
        exact = np.array([float(mpmath.exp(mpmath.mpf(sign * v))) for v in s])
        np.testing.assert_array_equal(mathlib(0, sign * s), exact)
        np.testing.assert_array_equal(mathlib(1, sign * s), exact)


def test_log_against_mpmath(mathlib):
    rng = np.random.default_rng(30)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 2000), rng.uniform(0.5, 2.0, 2000),
                        1.0 + rng.uniform(-1e-6, 1e-6, 500), 10.0 ** rng.uniform(-3, 2, 1500)])
    mpmath.mp.dps = 40
    got = mathlib(2, x)
    # error against the exact value (not its rounding): 3e-16 plus the final rounding
    err = np.array([float(abs(mpmath.mpf(g) - mpmath.log(mpmath.mpf(v)))) for g, v in zip(got, x)])
    bound = 3e-16 + 0.5 * np.spacing(np.abs(got))
    assert (err <= bound).all(), (err - bound).max()
    np.testing.assert_array_equal(mathlib(3, x), got)


def test_log_against_libm_many_points(mathlib):
    rng = np.random.default_rng(31)
    x = 10.0 ** rng.uniform(-30, 30, 1_000_000)
    want = np.log(x)  # libm: within an ulp of the exact value
    err = np.abs(mathlib(2, x) - want)
    assert (err <= 3e-16 + 1.5 * np.spacing(np.abs(want))).all()


def test_log_special_values(mathlib):
    assert mathlib(3, np.array([0.0]))[0] == -np.inf
    assert np.isnan(mathlib(2, np.array([np.nan]))[0])
    np.testing.assert_array_equal(mathlib(2, np.array([1.0, 2.0, 0.5])),
                                  np.log(np.array([1.0, 2.0, 0.5])))
