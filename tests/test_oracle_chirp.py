"""Pins the oracle's chirp terms and its bandwidth rule by physics, not by
retyping the reference's expressions.

* ucb_waveform slow series (oracle/orc_fd.c, following
  ucb/src/glass_ucb_waveform.c:804-937): on a point-like constellation
  (arm 1 m, so k.x/c ~ 1e-9 s and the transfer function is flat) the slow
  series of a chirping source divided by the same source with fdot =
  fddot = 0 must be the chirp alone: phase = 2 pi * integral of the
  frequency drift (pi fdot t^2 + pi/3 fddot t^3) and amplitude = the first
  order of the quadrupole law h ~ f^(2/3), 1 + (2/3) fdot t / f0.  Opposite
  signs and wrong factors fail by orders of magnitude.
* The fast/slow Fourier-domain waveform against the wavelet model's
  independent barycentre-phase series (ucb_phase / ucb_amplitude,
  glass_ucb_waveform.c:332-353, through the interpolated TDI series of
  :1064-1178) at Tobs = 4 yr with fdot of both signs and a drift
  fdot T^2 / 2 of 5-8 bins: match > 0.99 and norms within 1 %, while the
  same comparison against -fdot fails (match < 0.9).
* ucb_barycenter_waveform itself: phase and amplitude of a chirp.
* ucb_bandwidth (glass_ucb_waveform.c:698-725): the returned power of two
  switches exactly where the spread "1 + twice the largest frequency
  excursion in bins" passes 2^k + 1, for the chirp drift fdot T^2, the
  orbital Doppler excursion f v/c sin(theta) T (v = 2 pi AU / yr, the
  orbit's speed) and two bins per year of annual sidebands on each side;
  the Doppler coefficient is checked against the excursion measured on the
  oracle's own slow series; the SNR spread switches where the SNR passes
  2^k + 1, an SNR linear in A and in sqrt(T) that stays within a factor 2
  of the true sky-averaged SNR of the oracle's templates; the clamps
  [16, 2^floor(log2(N/2))]; monotone in |fdot|, A and |sin theta|.  The
  product's host ucb_bandwidth (glass_hip_ucb_bandwidth) is compared with
  the oracle at the same switching points and on random draws.
"""
import ctypes
import math

import numpy as np
import pytest

from glass_amd import synth

YEAR_T = 31457280.0
T2 = 2 * YEAR_T
T4 = 4 * YEAR_T
CLIGHT = 299792458.0
AU = 1.49597870700e11
YEAR = 3.15581497632e7
V_ORBIT = 2 * math.pi * AU / YEAR   # speed of a circular 1 AU orbit


def point_orbit(orc):
    """A constellation of 1 m arms: k.x/c and f/f* are ~1e-9, so the slow
    series carries the source's own phase and amplitude only."""
    L = 1.0
    return orc.Orbit(L, CLIGHT / (2 * math.pi * L), L / (2 * math.sqrt(3.0)))


def slow_time_series(orc, p, T, BW, orbit):
    """The six links' slow series as time samples t_n = T n / BW: the oracle
    returns them FFT'd (kissfft, unnormalised), half-swapped and scaled by
    1 / (2 BW); numpy undoes the three steps."""
    d = orc.slow_series(p, T, BW, 0.0, orbit)
    return np.fft.ifft(np.fft.ifftshift(d, axes=1) * (2 * BW), axis=1) * BW


@pytest.mark.parametrize("fdot", [2e-13, -2e-13])
def test_slow_series_chirp_closed_form(orc, fdot):
    f0, BW = 2e-3, 1024
    o = point_orbit(orc)
    p = np.array([f0 * T4, 1.0, 0.3, np.log(1e-21), 0.4, 0.7, 1.1, fdot * T4 * T4])
    p0 = p.copy()
    p0[7] = 0.0
    r = slow_time_series(orc, p, T4, BW, o) / slow_time_series(orc, p0, T4, BW, o)
    t = T4 * np.arange(BW) / BW
    amp = 1 + 2.0 / 3.0 * fdot * t / f0       # first order of (f(t)/f0)^(2/3)
    chirp = amp * np.exp(1j * math.pi * fdot * t * t)
    # |pi fdot T^2| ~ 1e4 rad: its rounding leaves ~1e-12
    assert np.max(np.abs(r / chirp - 1)) < 1e-10
    # the amplitude drift is 0.8 % at the end, the phase 1e4 rad: a wrong
    # sign or factor of either term misses by orders of magnitude
    assert np.max(np.abs(np.abs(r) - 1)) > 5e-3
    wrong_sign = amp * np.exp(-1j * math.pi * fdot * t * t)
    assert np.max(np.abs(r / wrong_sign - 1)) > 1.0
    wrong_amp = (1 + 4.0 / 3.0 * fdot * t / f0) * np.exp(1j * math.pi * fdot * t * t)
    assert np.max(np.abs(r / wrong_amp - 1)) > 3e-3
    # the full f^(2/3) law differs from its first order only at second order
    full = (1 + fdot * t / f0) ** (2.0 / 3.0)
    assert np.max(np.abs(full - amp)) < (fdot * T4 / f0) ** 2


def test_slow_series_fddot_closed_form(orc):
    f0, fdot, fddot, BW = 2e-3, 1e-14, 5e-23, 1024
    o = point_orbit(orc)
    p = np.array([f0 * T4, -1.0, 0.3, np.log(1e-21), 0.4, 0.7, 1.1, fdot * T4 ** 2, fddot * T4 ** 3])
    p0 = p.copy()
    p0[7] = p0[8] = 0.0
    r = slow_time_series(orc, p, T4, BW, o) / slow_time_series(orc, p0, T4, BW, o)
    t = T4 * np.arange(BW) / BW
    # phase = 2 pi * integral of (fdot t + fddot t^2 / 2)
    phase = math.pi * fdot * t ** 2 + math.pi / 3 * fddot * t ** 3
    assert math.pi / 3 * fddot * T4 ** 3 > 50.0      # the cubic term is tested
    expect = (1 + 2.0 / 3.0 * fdot * t / f0) * np.exp(1j * phase)
    assert np.max(np.abs(r / expect - 1)) < 1e-10
    nocubic = (1 + 2.0 / 3.0 * fdot * t / f0) * np.exp(1j * math.pi * fdot * t ** 2)
    assert np.max(np.abs(r / nocubic - 1)) > 1.0


def test_barycenter_waveform_closed_form(orc):
    f0, fdot = 3e-3, -4e-14
    p = np.array([f0 * T4, 0.1, 0.2, np.log(2e-22), 0.3, 0.4, 0.9, fdot * T4 ** 2])
    t = np.linspace(0, T4, 97)
    ph, am = orc.ucb_barycenter_waveform(p, t, T4)
    # -phi0 (the LDC phase convention) + 2 pi * integral of f(t)
    want = -0.9 + 2 * math.pi * (f0 * t + 0.5 * fdot * t ** 2)
    assert np.max(np.abs(ph - want)) < 1e-16 * np.max(np.abs(want)) * 8
    assert np.max(np.abs(am / (2e-22 * (1 + 2.0 / 3.0 * fdot * t / f0)) - 1)) < 1e-14
    assert np.max(np.abs(am / 2e-22 - 1)) > 1e-3


@pytest.fixture(scope="module")
def orbit4(orc):
    return orc.iorbit(T4)


@pytest.mark.parametrize("fdot,seed", [(1e-15, 11), (-1e-15, 12), (6e-16, 13)])
def test_fd_waveform_vs_barycentre_series_with_fdot(orc, orbit4, fdot, seed):
    """Two independent models of the reference (fast/slow FD, wavelet
    series built from the SSB phase) agree on a chirping source."""
    ND = 2 ** 22                      # 30 s cadence: Nyquist 16.7 mHz
    p = synth.draw_params(np.random.default_rng(seed), 1, T4)[0]
    p[0] = np.floor(4e-3 * T4) + 0.3
    p[7] = fdot * T4 * T4
    assert abs(p[7]) / 2 >= 4.5       # a drift of >= 4.5 bins over Tobs
    ser = orc.ucb_wavelet_series(orbit4, T4, p, ND, T4 / ND, 0.0)
    F = np.fft.rfft(ser, axis=1) * np.sqrt(T4) / (2 * ND)
    BW = 512
    q = int(p[0])
    fd = orc.ucb_waveform(p, T4, BW)
    pm = p.copy()
    pm[7] = -p[7]
    fdm = orc.ucb_waveform(pm, T4, BW)
    for c in range(3):
        a, b = fd[c], F[c, q - BW // 2:q + BW // 2]
        match = np.abs(np.vdot(a, b)) / np.linalg.norm(a) / np.linalg.norm(b)
        assert match > 0.99, (c, match)
        assert abs(np.linalg.norm(a) / np.linalg.norm(b) - 1) < 0.01
        wrong = np.abs(np.vdot(fdm[c], b)) / np.linalg.norm(fdm[c]) / np.linalg.norm(b)
        assert wrong < 0.9, (c, wrong)


# ------------------------------------------------------------ bandwidth

def _switch(g, lo, hi, geometric=False):
    """The argument where the integer g changes value between lo and hi
    (g(lo) != g(hi), g monotone), to the last bit."""
    a = g(lo)
    assert g(hi) != a
    for _ in range(2000):
        mid = math.sqrt(lo * hi) if geometric else 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        if g(mid) == a:
            lo = mid
        else:
            hi = mid
    return lo, hi


def _bw(orc, f, fdot, costh, A, T, N=1 << 22):
    o = orc.analytic_orbit()
    return orc.lib().orc_ucb_bandwidth(o.L, o.fstar, f, fdot, costh, A, T, N)


def _spread(f, fdot, sinth, T):
    """Twice the largest excursion from the carrier, in bins of 1/T: the
    chirp drift |fdot| T^2, the orbital Doppler f v sin(theta) / c T and two
    annual sidebands (2 / yr) of the constellation's motion."""
    return 2 * T * (2 * 2.0 / YEAR + f * V_ORBIT * sinth / CLIGHT + abs(fdot) * T)


@pytest.mark.parametrize("T", [T2, T4])
def test_bandwidth_switches_on_chirp_drift(orc, T):
    # cos(theta) = 1: no orbital Doppler; A tiny: the SNR spread is 16
    f, A = 3e-3, 1e-30
    for k in range(5, 11):
        g = lambda fd: _bw(orc, f, fd, 1.0, A, T) > 2 ** k  # noqa: E731
        lo, hi = _switch(g, 0.0, 1e-12)
        assert _bw(orc, f, lo, 1.0, A, T) == 2 ** k
        assert _bw(orc, f, hi, 1.0, A, T) == 2 ** (k + 1)
        # the switch sits where spread - 1 passes 2^k (not at spread = 2^k)
        assert abs(_spread(f, hi, 0.0, T) - (2 ** k + 1)) < 1e-9 * 2 ** k
        # |fdot|: the same switch for a source whose frequency falls
        assert _bw(orc, f, -lo, 1.0, A, T) == 2 ** k
        assert _bw(orc, f, -hi, 1.0, A, T) == 2 ** (k + 1)


def test_bandwidth_switches_on_orbital_doppler(orc):
    # cos(theta) = 0: the source in the ecliptic, full Doppler; fdot = 0
    A = 1e-30
    for T, ks in ((T2, (5, 6, 7)), (T4, (6, 7, 8))):
        for k in ks:
            g = lambda f: _bw(orc, f, 0.0, 0.0, A, T) > 2 ** k  # noqa: E731
            lo, hi = _switch(g, 1e-4, 0.05)
            assert abs(_spread(hi, 0.0, 1.0, T) - (2 ** k + 1)) < 1e-9 * 2 ** k


def test_doppler_coefficient_is_the_orbits_doppler(orc):
    """The measured Doppler excursion of the oracle's own slow series, for a
    source in the ecliptic, equals f v / c T (v = 2 pi AU / yr) to 2 %: the
    ratio of the slow series at two carrier frequencies (whole bins, so the
    heterodyne cancels) carries the phase -2 pi (f_a - f_b) k.x / c."""
    T = T2
    BW = 2048
    qa, qb = 630000, 315000            # 10.0 and 5.0 mHz
    fa, fb = qa / T, qb / T
    p = np.array([qa, 0.0, 1.3, np.log(1e-21), 0.5, 0.3, 0.7, 0.0])
    pb = p.copy()
    pb[0] = qb
    o = orc.analytic_orbit()
    r = slow_time_series(orc, p, T, BW, o)[0] / slow_time_series(orc, pb, T, BW, o)[0]
    ph = np.unwrap(np.angle(r))
    finst = np.diff(ph) / (2 * math.pi) * BW / T        # Hz
    exc = 0.5 * (finst.max() - finst.min())
    assert abs(exc / ((fa - fb) * V_ORBIT / CLIGHT) - 1) < 0.02


def test_bandwidth_snr_switches(orc):
    """The SNR spread switches where SNR - 1 passes 2^k, with the SNR linear
    in A (ratios of consecutive switching amplitudes (2^(k+1)+1)/(2^k+1),
    not 2) and in sqrt(T)."""
    f = 5e-3
    amps = {}
    # (the Doppler-free spread is 16 bins at 1 yr and 32 at 4 yr)
    for T in (YEAR_T, T4):
        for k in range(5, 11):
            g = lambda A: _bw(orc, f, 0.0, 1.0, A, T) > 2 ** k  # noqa: E731
            lo, hi = _switch(g, 1e-26, 1e-17, geometric=True)
            amps[T, k] = hi
        for k in range(5, 10):
            ratio = amps[T, k + 1] / amps[T, k]
            assert abs(ratio - (2 ** (k + 1) + 1) / (2 ** k + 1)) < 1e-12
    for k in range(5, 11):
        assert abs(amps[YEAR_T, k] / amps[T4, k] - 2.0) < 1e-12


def test_bandwidth_snr_matches_template_snr(orc):
    """The analytic SNR behind the SNR spread stays within a factor 2 of the
    rms SNR of the oracle's templates over random sky positions and
    orientations (X, Y, Z with the full noise covariance), from 0.7 to
    12 mHz, and scales with sqrt(T) exactly as they do."""
    ratios = {}
    for T in (YEAR_T, T4):
        for f in (0.7e-3, 2e-3, 5e-3, 12e-3):
            g = lambda A: _bw(orc, f, 0.0, 1.0, A, T) > 64  # noqa: E731
            _, A64 = _switch(g, 1e-26, 1e-17, geometric=True)
            snrm = 65.0 / A64 * 1e-21          # analytic SNR at A = 1e-21
            rng = np.random.default_rng(3)
            s2 = []
            for _ in range(24):
                p = synth.draw_params(rng, 1, T)[0]
                p[0], p[3], p[7] = f * T, np.log(1e-21), 0.0
                BW = 256
                h = orc.ucb_waveform(p, T, BW)
                qmin = int(p[0]) - BW // 2
                _, invC, _ = orc.xyz_noise(T, qmin / T, BW)
                s2.append(orc.snr(h[0], h[1], h[2], invC, BW) ** 2)
            ratios[T, f] = math.sqrt(np.mean(s2)) / snrm
            assert 1.0 < ratios[T, f] < 2.0, (T, f, ratios[T, f])
    for f in (0.7e-3, 2e-3, 5e-3, 12e-3):
        assert abs(ratios[YEAR_T, f] / ratios[T4, f] - 1) < 1e-4


def test_bandwidth_clamps(orc):
    for N in (32, 64, 100, 1000, 1023, 1024, 4096, 5000, 1 << 20):
        nmax = 1
        while 2 * nmax <= N // 2:
            nmax *= 2
        assert _bw(orc, 5e-3, 0.0, 0.3, 1e-15, T2, N) == max(nmax, 16)
        assert _bw(orc, 5e-3, 0.0, 1.0, 1e-30, T2, N) == 16


def test_bandwidth_monotone(orc):
    rng = np.random.default_rng(17)
    for _ in range(300):
        f = rng.uniform(0.5e-3, 20e-3)
        fd = rng.uniform(-1e-15, 1e-15)
        c = rng.uniform(-1, 1)
        A = 10 ** rng.uniform(-24, -20)
        T = [YEAR_T, T2, T4][rng.integers(3)]
        b = _bw(orc, f, fd, c, A, T)
        assert b & (b - 1) == 0 and b >= 16
        assert _bw(orc, f, 1.5 * fd, c, A, T) >= b
        assert _bw(orc, f, fd, c, 1.5 * A, T) >= b
        assert _bw(orc, f, fd, 0.5 * c, A, T) >= b


def test_library_bandwidth_equals_oracle(orc, gl_host):
    """The product's host ucb_bandwidth (the one that sizes templates) at
    the oracle's switching points and on random draws."""
    lib = gl_host
    o = orc.analytic_orbit()

    def both(f, fd, c, A, T, N=1 << 22):
        a = lib.glass_hip_ucb_bandwidth(o.L, o.fstar, f, fd, c, A, T, N)
        b = orc.lib().orc_ucb_bandwidth(o.L, o.fstar, f, fd, c, A, T, N)
        assert a == b, (f, fd, c, A, T, N, a, b)

    for k in range(5, 11):
        lo, hi = _switch(lambda fd: _bw(orc, 3e-3, fd, 1.0, 1e-30, T2) > 2 ** k, 0.0, 1e-12)
        both(3e-3, lo, 1.0, 1e-30, T2)
        both(3e-3, hi, 1.0, 1e-30, T2)
    for k in range(5, 11):
        lo, hi = _switch(lambda A: _bw(orc, 5e-3, 0.0, 1.0, A, T4) > 2 ** k, 1e-26, 1e-17,
                         geometric=True)
        both(5e-3, 0.0, 1.0, lo, T4)
        both(5e-3, 0.0, 1.0, hi, T4)
    rng = np.random.default_rng(23)
    for _ in range(2000):
        both(rng.uniform(0.1e-3, 25e-3), rng.uniform(-1e-15, 1e-15), rng.uniform(-1, 1),
             10 ** rng.uniform(-25, -19), [YEAR_T, T2, T4][rng.integers(3)],
             int(rng.choice([64, 256, 1000, 4096, 8192])))


@pytest.fixture(scope="module")
def gl_host():
    """libglass_hip.so's host-only entry point (no GPU call)."""
    import glass_amd
    L = glass_amd.lib()
    L.glass_hip_ucb_bandwidth.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int]
    L.glass_hip_ucb_bandwidth.restype = ctypes.c_int
    return L
