"""Pins the oracle FFT (kissfft algorithm restated in oracle/orc_fft.c).

Checks: bitwise equality with kissfft compiled from the reference's own
sources (oracle/_ref, only where the reference tree was present at build
time), a brute-force O(N^2) DFT, and the impulse known answers of
apps/src/fft_wdm_unit_tests.c:454-545."""
import ctypes

import numpy as np
import pytest


SIZES = [4, 8, 16, 32, 64, 128, 256, 1024, 2048, 12, 30, 1536, 77, 7 * 11 * 13]


def brute_dft(x, sign=-1):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", SIZES)
def test_fft_vs_bruteforce(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = orc.fft(x)
    assert np.max(np.abs(y - brute_dft(x))) <= 1e-12 * np.max(np.abs(y))
    yi = orc.fft(x, inverse=True)
    assert np.max(np.abs(yi - brute_dft(x, +1))) <= 1e-12 * np.max(np.abs(yi))


@pytest.mark.parametrize("n", SIZES)
def test_fft_bitwise_reference_kissfft(orc, n):
    K = orc.ref_kissfft()
    if K is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(1000 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for inv in (0, 1):
        cfg = K.kiss_fft_alloc(n, inv, None, None)
        ref = np.zeros(n, complex)
        K.kiss_fft(cfg, x.view(np.float64).ctypes.data_as(orc._dp),
                   ref.view(np.float64).ctypes.data_as(orc._dp))
        ctypes.CDLL(None).free(ctypes.c_void_p(cfg))
        assert np.array_equal(orc.fft(x, inverse=bool(inv)), ref)


@pytest.mark.parametrize("n", [8, 64, 512, 2048, 1536, 60])
def test_fftr_bitwise_reference_kissfft(orc, n):
    K = orc.ref_kissfft()
    if K is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(2000 + n)
    x = rng.standard_normal(n)
    cfg = K.kiss_fftr_alloc(n, 0, None, None)
    ref = np.zeros(n // 2 + 1, complex)
    K.kiss_fftr(cfg, x.ctypes.data_as(orc._dp), ref.view(np.float64).ctypes.data_as(orc._dp))
    ctypes.CDLL(None).free(ctypes.c_void_p(cfg))
    F = orc.fftr(x)
    assert np.array_equal(F, ref)
    cfgi = K.kiss_fftr_alloc(n, 1, None, None)
    back = np.zeros(n)
    K.kiss_fftri(cfgi, ref.view(np.float64).ctypes.data_as(orc._dp), back.ctypes.data_as(orc._dp))
    ctypes.CDLL(None).free(ctypes.c_void_p(cfgi))
    assert np.array_equal(orc.fftri(F, n), back)


@pytest.mark.parametrize("n", [64, 1536, 2048])
def test_fftr_vs_numpy(orc, n):
    rng = np.random.default_rng(3000 + n)
    x = rng.standard_normal(n)
    F = orc.fftr(x)
    assert np.max(np.abs(F - np.fft.rfft(x))) <= 1e-12 * np.max(np.abs(F))
    assert np.max(np.abs(orc.fftri(F, n) / n - x)) <= 1e-13


def test_impulse_known_answers(orc):
    """fft_wdm_unit_tests.c:454-545: FFT(impulse) == 1 in every bin, and the
    inverse returns the impulse (atol 1e-10, rtol 1e-5 there)."""
    n = 1536 * 30
    x = np.zeros(n, complex)
    x[0] = 1.0
    y = orc.fft(x)
    assert np.allclose(y, 1.0, atol=1e-10, rtol=1e-5)
    back = orc.fft(y, inverse=True) / n
    assert np.allclose(back, x, atol=1e-10, rtol=1e-5)
    xr = np.zeros(n)
    xr[0] = 1.0
    F = orc.fftr(xr)
    assert np.allclose(F, 1.0, atol=1e-10, rtol=1e-5)
    assert np.allclose(orc.fftri(F, n) / n, xr, atol=1e-10, rtol=1e-5)


def test_plant_mistake_detected(orc):
    """A dropped twiddle conjugation (inverse instead of forward) must fail
    the brute-force check: the check is not vacuous."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    wrong = orc.fft(x, inverse=True)
    assert np.max(np.abs(wrong - brute_dft(x))) > 1e-3
