"""Accuracy of the library's own FP64 sin/cos and sin(x)/x
(glass_amd/csrc/fastmath.h), compiled for the host and compared with
glibc on the argument ranges the slow series produces (phases up to ~2e7
rad, transfer-function arguments below 1.75)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def fm(tmp_path_factory):
    d = tmp_path_factory.mktemp("fm")
    src = d / "fm.cpp"
    src.write_text('#include "fastmath.h"\n'
                   'extern "C" void fm_sincos(const double* x, double* s, double* c, int n)'
                   '{ for (int i = 0; i < n; i++) glass::sincos_red(x[i], s + i, c + i); }\n'
                   'extern "C" void fm_sinc(const double* x, double* y, int n)'
                   '{ for (int i = 0; i < n; i++) y[i] = glass::sinc_ratio(x[i]); }\n')
    so = d / "fm.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "glass_amd", "csrc"), str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    dp = ctypes.POINTER(ctypes.c_double)
    L.fm_sincos.argtypes = [dp, dp, dp, ctypes.c_int]
    L.fm_sinc.argtypes = [dp, dp, ctypes.c_int]
    return L, dp


def ulps(a, ref, scale):
    return np.max(np.abs(a - ref) / np.spacing(np.maximum(np.abs(ref), scale)))


@pytest.mark.parametrize("lo,hi", [(-4.0, 4.0), (1e3, 1e6), (1e6, 3e7), (-3e7, -1e6), (2e8, 5.3e8)])
def test_sincos_vs_glibc(fm, lo, hi):
    L, dp = fm
    rng = np.random.default_rng(int(abs(lo)) + 1)
    x = rng.uniform(lo, hi, 200000)
    s = np.empty_like(x)
    c = np.empty_like(x)
    L.fm_sincos(x.ctypes.data_as(dp), s.ctypes.data_as(dp), c.ctypes.data_as(dp), len(x))
    # error in ulps of max(|result|, 2^-3): absolute accuracy near zeros of
    # sin/cos is what an amplitude-weighted sum sees
    assert ulps(s, np.sin(x), 0.125) <= 4
    assert ulps(c, np.cos(x), 0.125) <= 4


def test_sincos_special(fm):
    L, dp = fm
    x = np.array([0.0, -0.0, np.pi / 2, np.pi, 1e-300, 5.0e8, 1e10, np.inf, np.nan])
    s = np.empty_like(x)
    c = np.empty_like(x)
    L.fm_sincos(x.ctypes.data_as(dp), s.ctypes.data_as(dp), c.ctypes.data_as(dp), len(x))
    fin = np.abs(x) < 2.0 ** 29  # the domain; outside it the result is NaN
    assert np.allclose(s[fin], np.sin(x[fin]), rtol=0, atol=1e-15)
    assert np.allclose(c[fin], np.cos(x[fin]), rtol=0, atol=1e-15)
    assert np.all(np.isnan(s[~fin])) and np.all(np.isnan(c[~fin]))


def test_sinc_ratio(fm):
    L, dp = fm
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-1.75, 1.75, 100000), rng.uniform(-1e-8, 1e-8, 1000), [1e-300, 1.9, 3.0]])
    y = np.empty_like(x)
    L.fm_sinc(x.ctypes.data_as(dp), y.ctypes.data_as(dp), len(x))
    ref = np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))
    assert ulps(y, ref, 0.0) <= 3
