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
                   '{ for (int i = 0; i < n; i++) y[i] = glass::sinc_ratio(x[i]); }\n'
                   'extern "C" void fm_sincos_cr(const double* x, double* s, double* c, int n)'
                   '{ for (int i = 0; i < n; i++) glass::sincos_cr(x[i], s + i, c + i); }\n'
                   'extern "C" void fm_atan2_cr(const double* y, const double* x, double* r, int n)'
                   '{ for (int i = 0; i < n; i++) r[i] = glass::atan2_cr(y[i], x[i]); }\n'
                   'extern "C" void fm_exp_cr(const double* x, double* r, int n)'
                   '{ for (int i = 0; i < n; i++) r[i] = glass::exp_cr(x[i]); }\n')
    so = d / "fm.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "glass_amd", "csrc"), str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    dp = ctypes.POINTER(ctypes.c_double)
    L.fm_sincos.argtypes = [dp, dp, dp, ctypes.c_int]
    L.fm_sinc.argtypes = [dp, dp, ctypes.c_int]
    L.fm_sincos_cr.argtypes = [dp, dp, dp, ctypes.c_int]
    L.fm_atan2_cr.argtypes = [dp, dp, dp, ctypes.c_int]
    L.fm_exp_cr.argtypes = [dp, dp, ctypes.c_int]
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


# The correctly rounded variants used on the wavelet template's phase chain:
# compared with values rounded from 150-bit mpmath evaluations.

def _mp():
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 150
    return mp


@pytest.mark.parametrize("lo,hi", [(-7.0, 7.0), (1e5, 2e6), (-5e8, 5e8)])
def test_sincos_cr_correctly_rounded(fm, lo, hi):
    mp = _mp()
    L, dp = fm
    x = np.random.default_rng(int(abs(hi))).uniform(lo, hi, 1500)
    s = np.empty_like(x)
    c = np.empty_like(x)
    L.fm_sincos_cr(x.ctypes.data_as(dp), s.ctypes.data_as(dp), c.ctypes.data_as(dp), len(x))
    assert np.array_equal(s, [float(mp.sin(mp.mpf(v))) for v in x])
    assert np.array_equal(c, [float(mp.cos(mp.mpf(v))) for v in x])


@pytest.mark.parametrize("scale", [1.0, 1e-23, 1e5])
def test_atan2_cr_correctly_rounded(fm, scale):
    mp = _mp()
    L, dp = fm
    rng = np.random.default_rng(int(scale * 7) % 1000 + 1)
    y = rng.standard_normal(1500) * scale
    x = rng.standard_normal(1500) * scale
    r = np.empty_like(x)
    L.fm_atan2_cr(y.ctypes.data_as(dp), x.ctypes.data_as(dp), r.ctypes.data_as(dp), len(x))
    assert np.array_equal(r, [float(mp.atan2(mp.mpf(a), mp.mpf(b))) for a, b in zip(y, x)])
    z = np.array([0.0, -0.0, 1.0, -1.0])
    for a in z:
        for b in z:
            L.fm_atan2_cr(np.array([a]).ctypes.data_as(dp), np.array([b]).ctypes.data_as(dp),
                          r.ctypes.data_as(dp), 1)
            assert r[0] == np.arctan2(a, b) and np.signbit(r[0]) == np.signbit(np.arctan2(a, b))


@pytest.mark.parametrize("lo,hi", [(-55.0, -46.0), (-1.0, 1.0), (-700.0, 700.0)])
def test_exp_cr_correctly_rounded(fm, lo, hi):
    mp = _mp()
    L, dp = fm
    x = np.random.default_rng(int(abs(lo))).uniform(lo, hi, 1500)
    r = np.empty_like(x)
    L.fm_exp_cr(x.ctypes.data_as(dp), r.ctypes.data_as(dp), len(x))
    assert np.array_equal(r, [float(mp.exp(mp.mpf(v))) for v in x])
