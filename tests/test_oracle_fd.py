"""Pins the oracle's Fourier-domain path (oracle/orc_fd.c) with checks that
do not retype its formulas:

* ucb_waveform: linearity in A, the phi0 global phase e^{-i dphi0}, the
  psi -> psi + pi/2 sign flip, the three TDI formulas of the reference
  against each other (LISA_tdi_FF X = -i LISA_tdi_Sangria X, phase X =
  Sangria X / (2 i f/f*), utils/src/glass_lisa.c:458-710), and the carrier
  placement of the signal energy at f0*T + fdot*T^2/2.
* fourier_nwip / snr / gaussian_log_likelihood: closed-form numpy sums.
* invert_noise_covariance_matrix: C @ invC = I and log det C.
* ucb_bandwidth: hand-evaluated closed form at chosen inputs.
* ucb_fisher (finite-difference, reference-compatible): symmetry,
  positive semi-definiteness, diagonal == snr(dh_i)^2.
"""
import math

import numpy as np
import pytest

from glass_amd import synth
from conftest import normwise

T = synth.YEAR_T


def P(f0=5e-3, costh=0.3, phi=1.0, A=1e-22, cosi=0.2, psi=0.5, phi0=1.2, fdot=1e-17, T=T):
    return np.array([f0 * T, costh, phi, np.log(A), cosi, psi, phi0, fdot * T * T])


def test_linear_in_amplitude(orc):
    p = P()
    h1 = orc.ucb_waveform(p, T, 128)
    q = p.copy()
    q[3] += 0.7
    h2 = orc.ucb_waveform(q, T, 128)
    assert normwise(h2, np.exp(0.7) * h1) < 1e-13


def test_phi0_is_global_phase(orc):
    p = P()
    h1 = orc.ucb_waveform(p, T, 128)
    q = p.copy()
    q[6] += 0.3
    h2 = orc.ucb_waveform(q, T, 128)
    # arg2 = 2 pi f0 xi - phi0 - ...: phi0 enters as exp(-i phi0); its
    # rounding at |arg2| ~ 1e6 rad allows ~1e-10 movement
    assert normwise(h2, np.exp(-0.3j) * h1) < 1e-8
    assert normwise(h2, np.exp(+0.3j) * h1) > 1e-2


def test_psi_quarter_turn_flips_sign(orc):
    p = P()
    h1 = orc.ucb_waveform(p, T, 128)
    q = p.copy()
    q[5] += np.pi / 2
    assert normwise(orc.ucb_waveform(q, T, 128), -h1) < 1e-10


@pytest.mark.parametrize("f0", [1.3e-3, 4e-3, 9.7e-3])
def test_tdi_formats_agree(orc, f0):
    p = P(f0=f0)
    BW = 128
    hs = orc.ucb_waveform(p, T, BW, fmt="sangria")
    hf = orc.ucb_waveform(p, T, BW, fmt="frequency")
    hp = orc.ucb_waveform(p, T, BW, fmt="phase")
    q = int(p[0] / T * T)
    f = (q + np.arange(BW) - BW // 2) / T
    x = f / synth.FSTAR
    for c in range(3):
        # (a e^{-2ix} + b e^{-ix})(e^{-ix} - e^{ix}) = -2i sin x (...)
        assert normwise(hf[c], -1j * hs[c]) < 1e-12
        assert normwise(hp[c] * (2j * x), hs[c]) < 1e-12
    # A/E: LDC combinations; "phase" conjugates A,E (glass_lisa.c:544-545)
    assert normwise(hs[3], (hs[2] - hs[0]) / np.sqrt(2)) < 1e-14
    assert normwise(hs[4], (hs[0] - 2 * hs[1] + hs[2]) / np.sqrt(6)) < 1e-14
    assert normwise(hp[3], np.conj((hp[2] - hp[0]) / np.sqrt(2))) < 1e-14


@pytest.mark.parametrize("f0,fdot", [(2e-3, 0.0), (6e-3, 5e-17), (9e-3, -1e-17)])
def test_carrier_placement(orc, f0, fdot):
    p = P(f0=f0, fdot=fdot)
    BW = 256
    h = orc.ucb_waveform(p, T, BW)
    q = int(p[0] / T * T)
    k = q - BW // 2 + np.arange(BW)
    w = np.sum(np.abs(h[:3]) ** 2, axis=0)
    centroid = np.sum(k * w) / np.sum(w)
    expect = f0 * T + 0.5 * fdot * T * T
    # annual Doppler excursion f0 * v_orbit / c, in bins
    doppler = f0 * (2 * np.pi * 1.49597870700e11 / 3.15581497632e7) / synth.CLIGHT * T
    assert abs(centroid - expect) < 0.3 * doppler + 2.0
    # little energy outside the Doppler band (the slow series is not
    # periodic over T, so its FFT leaks a ~1/k tail)
    far = np.abs(k - expect) > 3 * doppler + 8
    assert w[far].sum() < 1e-2 * w.sum()


def test_bandwidth_closed_form(orc):
    o = orc.analytic_orbit()
    AU, C, YEAR = 1.49597870700e11, 299792458.0, 3.15581497632e7
    f, fdot, costh = 5e-3, 2e-17, 0.0
    bw = 2 * T * ((4.0 + 2 * math.pi * f * (AU / C) * 1.0) / YEAR + abs(fdot) * T)
    DS = 2 ** (int(math.log2(bw - 1)) + 1)
    # tiny amplitude: SNR spreading is clamped to the 16-bin floor
    got = orc.lib().orc_ucb_bandwidth(o.L, o.fstar, f, fdot, costh, 1e-30, T, 4096)
    assert got == max(DS, 16) == 64
    # huge amplitude: SNR spreading hits Nmax = 2^floor(log2(NFFT/2))
    assert orc.lib().orc_ucb_bandwidth(o.L, o.fstar, f, fdot, costh, 1e-15, T, 4096) == 2048
    assert orc.lib().orc_ucb_bandwidth(o.L, o.fstar, f, fdot, costh, 1e-15, T, 1000) == 256


def test_nwip_closed_forms(orc):
    rng = np.random.default_rng(3)
    N = 37
    a = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    w = rng.uniform(0.5, 2, N)
    ref = 2 * np.sum((a * np.conj(b)).real * w)
    assert abs(orc.fourier_nwip(a, b, w, N) - ref) < 1e-13 * abs(ref) * 10
    lst = rng.choice(N, 11, replace=False)
    ar, br = a.real.copy(), b.real.copy()
    refw = np.sum(ar[lst] * br[lst] * w[lst])
    assert abs(orc.wavelet_nwip(ar, br, w, lst) - refw) < 1e-13


def test_noise_inverse(orc):
    C, invC, ld = orc.xyz_noise(T, 4e-3, 64)
    Cm, Im = np.moveaxis(C, -1, 0), np.moveaxis(invC, -1, 0)
    assert np.max(np.abs(Cm @ Im - np.eye(3))) < 1e-10
    assert np.max(np.abs(ld - np.log(np.linalg.det(Cm)))) < 1e-10
    # the product's synthetic-input generator restates the same PSD
    assert np.max(np.abs(synth.xyz_noise_cov(T, 4e-3, 64) - C) / np.abs(C)) < 1e-13


def test_loglikelihood_and_snr_closed_form(orc):
    rng = np.random.default_rng(5)
    W = 96
    C, invC, _ = orc.xyz_noise(T, 3e-3, W)
    d = np.zeros((5, W), complex)
    m = np.zeros((5, W), complex)
    d[:3] = synth.noise_draw(rng, C)
    m[:3] = 0.3 * synth.noise_draw(rng, C)
    r = d[:3] - m[:3]
    chi2 = 0.0
    for a in range(3):
        for b in range(3):
            chi2 += 2 * np.sum((r[a] * np.conj(r[b])).real * invC[a, b])
    got = orc.gaussian_log_likelihood(d, m, invC)
    assert abs(got - (-0.5 * chi2)) < 1e-12 * abs(chi2)
    s2 = 0.0
    for a in range(3):
        for b in range(3):
            s2 += 2 * np.sum((m[a] * np.conj(m[b])).real * invC[a, b])
    assert abs(orc.snr(m[0], m[1], m[2], invC, W) - np.sqrt(s2)) < 1e-12 * np.sqrt(s2)


def test_fisher_fd_properties(orc):
    NFFT = 256
    p = P(A=3e-22)
    f0 = p[0] / T
    qmin = synth.window_for(f0, T, NFFT)
    C, invC, _ = orc.xyz_noise(T, qmin / T, NFFT)
    F = orc.ucb_fisher_fd(p, T, NFFT, qmin, invC)
    assert np.array_equal(F, F.T)
    ev = np.linalg.eigvalsh(F)
    assert ev.min() > -1e-9 * ev.max()
    # diagonal == snr(dh_i)^2 with the same (bin-0 aligned) invC
    BW, _, _ = orc.ucb_alignment(p, T, NFFT, qmin)
    h0 = orc.ucb_waveform(p, T, BW)
    for i in [0, 3, 6]:
        q = p.copy()
        q[i] += 1e-6
        BWp, _, _ = orc.ucb_alignment(q, T, NFFT, qmin)
        assert BWp == BW
        dh = (orc.ucb_waveform(q, T, BW) - h0) * 1e6
        s = orc.snr(dh[0], dh[1], dh[2], invC, BW)
        assert abs(F[i, i] - s * s) < 1e-9 * F[i, i]
    # amplitude direction: d h / d lnA = h, so F[3,3] = snr(h)^2 up to the
    # forward-difference truncation (1e-6)
    s0 = orc.snr(h0[0], h0[1], h0[2], invC, BW)
    assert abs(F[3, 3] / s0 ** 2 - 1) < 1e-5


def test_numeric_orbit_reproduces_analytic(orc):
    """interpolate_orbits' splines over knots sampled from the analytic orbit
    (400 per year) give the analytic template to the spline's interpolation
    error (< 1e-6 normwise), and the average arm length of those samples is
    the analytic 2.5e9 m; a 3000 km wobble of the knots moves the template
    by ~1e-3 (the numeric path really reads the knots)."""
    import ctypes
    o = orc.analytic_orbit()
    N = 400
    buf = T / (N - 1)
    t = -buf + np.arange(N) * (T + 2 * buf) / (N - 1)
    pos = np.zeros((9, N))
    for n in range(N):
        x, y, z = np.zeros(4), np.zeros(4), np.zeros(4)
        orc.lib().orc_analytic_orbits(ctypes.byref(o), t[n], orc._d(x), orc._d(y), orc._d(z))
        for i in range(3):
            pos[3 * i:3 * i + 3, n] = x[i + 1], y[i + 1], z[i + 1]
    p = P(f0=4e-3)
    ana = orc.ucb_waveform(p, T, 256)
    with orc.numeric_orbit(t, pos) as on:
        assert abs(on.orbit.L / 2.5e9 - 1) < 1e-9
        h = orc.ucb_waveform(p, T, 256, orbit=on.orbit)
    assert normwise(h, ana) < 1e-6
    wob = pos + 3e6 * np.sin(2 * np.pi * t / 1.2e7)[None, :] * np.arange(1, 10)[:, None] / 9
    with orc.numeric_orbit(t, wob) as on:
        hw = orc.ucb_waveform(p, T, 256, orbit=on.orbit)
    assert normwise(hw, ana) > 1e-4
    # outside the block the analytic orbit is back
    assert np.array_equal(orc.ucb_waveform(p, T, 256), ana)
