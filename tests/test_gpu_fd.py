"""GPU parity of the Fourier-domain drop-ins against the oracle.

Gates (the reference's FP64 CPU output as restated by the oracle):
waveform bins normwise max|dh|/max|h| <= 1e-12; logL, snr, nwip relative
error <= 1e-10."""
import ctypes

import numpy as np
import pytest

from glass_amd import synth
from conftest import normwise

pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-12
LOGL_TOL = 1e-10


@pytest.fixture(scope="module")
def cfg0(orc):
    return synth.config0(seed=20240, waveform=lambda p, T, BW: orc.ucb_waveform(p, T, BW))


def test_waveform_config0(gl, orc, cfg0):
    """64 vectors, BW=128, T=1 yr, 'sangria' (the plain method case)."""
    o = gl.analytic_orbit()
    worst = 0.0
    for i in range(len(cfg0["params"])):
        p = cfg0["params"][i]
        print(f"vector {i}", flush=True)
        h_gpu = gl.ucb_waveform(o, "sangria", cfg0["T"], 0.0, p, 8, 128)
        h_ref = orc.ucb_waveform(p, cfg0["T"], 128)
        worst = max(worst, normwise(h_gpu, h_ref))
    print("worst normwise", worst)
    assert worst <= WAVE_TOL


@pytest.mark.parametrize("fmt", ["sangria", "phase", "frequency"])
@pytest.mark.parametrize("BW", [16, 32, 64, 256, 512, 1024, 2048])
def test_waveform_formats_sizes(gl, orc, fmt, BW):
    rng = np.random.default_rng(BW)
    T = synth.T_2YR
    o = gl.analytic_orbit()
    P = synth.draw_params(rng, 4, T, 5e-4, 2e-2)
    for p in P:
        h_gpu = gl.ucb_waveform(o, fmt, T, 0.0, p, 8, BW)
        h_ref = orc.ucb_waveform(p, T, BW, fmt=fmt)
        assert normwise(h_gpu, h_ref) <= WAVE_TOL, (fmt, BW, p)


@pytest.mark.parametrize("NP", [7, 9])
def test_waveform_nparams(gl, orc, NP):
    rng = np.random.default_rng(NP)
    T = synth.YEAR_T
    P = synth.draw_params(rng, 3, T)
    o = gl.analytic_orbit()
    for p in P:
        q = np.concatenate([p, [1e-26 * T ** 3]])[:NP]
        h_gpu = gl.ucb_waveform(o, "sangria", T, 0.0, q, NP, 128)
        h_ref = orc.ucb_waveform(q, T, 128)
        assert normwise(h_gpu, h_ref) <= WAVE_TOL


def test_waveform_t0_and_ni(gl, orc):
    T = synth.YEAR_T
    p = synth.draw_params(np.random.default_rng(9), 1, T)[0]
    o = gl.analytic_orbit()
    h_gpu = gl.ucb_waveform(o, "phase", T, 1.5e6, p, 8, 64, NI=1)
    h_ref = orc.ucb_waveform(p, T, 64, t0=1.5e6, fmt="phase", NI=1)
    assert normwise(h_gpu[0], h_ref[0]) <= WAVE_TOL
    assert np.all(h_gpu[1:] == 0)  # NI=1: only X written


def test_unsupported_format_raises(gl):
    o = gl.analytic_orbit()
    with pytest.raises(gl.GlassHipError):
        gl.ucb_waveform(o, "tdi2", synth.YEAR_T, 0.0, np.zeros(8) + 0.1, 8, 64)


def test_loglikelihood_config0(gl, orc, cfg0):
    """Full-window logL of the data against the injected template."""
    T, W, BW = cfg0["T"], cfg0["W"], cfg0["BW"]
    worst = 0.0
    for i in range(len(cfg0["params"])):
        p = cfg0["params"][i]
        _, invC, _ = orc.xyz_noise(T, cfg0["qmin"][i] / T, W)
        data = np.zeros((5, W), complex)
        data[:3] = cfg0["data"][i]
        model = np.zeros((5, W), complex)
        # a slightly detuned template, as an MCMC proposal would be
        q = p.copy()
        q[0] += 0.01
        h = orc.ucb_waveform(q, T, BW)
        bw, qm, imin = orc.ucb_alignment(q, T, W, int(cfg0["qmin"][i]))
        synth.scatter(model[:3], h[:3], imin)
        ref = orc.gaussian_log_likelihood(data, model, invC)
        got = gl.gaussian_log_likelihood(data, model, invC)
        worst = max(worst, abs(got - ref) / abs(ref))
        s_ref = orc.snr(h[0], h[1], h[2], invC, BW)
        s_got = gl.snr(h[0], h[1], h[2], invC, BW)
        assert abs(s_got - s_ref) <= LOGL_TOL * s_ref
    print("worst logL rel err", worst)
    assert worst <= LOGL_TOL


def test_nwip_and_edges(gl, orc):
    rng = np.random.default_rng(11)
    for N in [1, 7, 256, 4096, 100003]:
        a = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        w = rng.uniform(0.1, 3, N)
        ref = orc.fourier_nwip(a, b, w, N)
        assert abs(gl.fourier_nwip(a, b, w, N) - ref) <= LOGL_TOL * abs(ref) + 1e-300
    assert gl.fourier_nwip(np.zeros(2), np.zeros(2), np.zeros(1), 0) == 0.0
    n = 5000
    a, b, w = rng.standard_normal(n), rng.standard_normal(n), rng.uniform(0.1, 3, n)
    lst = rng.choice(n, 777, replace=False).astype(np.int32)
    ref = orc.wavelet_nwip(a, b, w, lst)
    assert abs(gl.wavelet_nwip(a, b, w, lst) - ref) <= LOGL_TOL * abs(ref)
    with pytest.raises(gl.GlassHipError):
        gl.wavelet_nwip(a, b, w, np.array([n + 5], dtype=np.int32))


def test_logl_two_and_one_channel(gl, orc):
    rng = np.random.default_rng(12)
    W = 300
    C, invC, _ = orc.xyz_noise(synth.YEAR_T, 2e-3, W)
    d = np.zeros((5, W), complex)
    m = np.zeros((5, W), complex)
    d[:] = rng.standard_normal((5, W)) + 1j * rng.standard_normal((5, W))
    m[:] = 0.1 * (rng.standard_normal((5, W)) + 1j * rng.standard_normal((5, W)))
    for nc in (1, 2):
        ref = orc.gaussian_log_likelihood(d, m, invC, nchannel=nc)
        got = gl.gaussian_log_likelihood(d, m, invC, Nchannel=nc)
        assert abs(got - ref) <= LOGL_TOL * abs(ref)


def test_invert_noise_covariance(gl, orc):
    C, invC_ref, ld_ref = orc.xyz_noise(synth.YEAR_T, 1e-3, 2048)
    invC, ld = gl.invert_noise_covariance_matrix(C)
    assert np.max(np.abs(invC - invC_ref) / np.abs(invC_ref)) < 1e-13
    assert np.max(np.abs(ld - ld_ref)) < 1e-12


def _orbit_file(orc, T, t0, N=400, seed=3):
    """An orbit 'file': knots over [t0 - T/(N-1), t0 + T + T/(N-1)] with the
    analytic positions plus a smooth few-thousand-km wobble per spacecraft
    (so the constellation is not the rigid analytic one)."""
    o = orc.analytic_orbit()
    buf = T / (N - 1)
    t = t0 - buf + np.arange(N) * (T + 2 * buf) / (N - 1)
    rng = np.random.default_rng(seed)
    pos = np.zeros((9, N))
    for n in range(N):
        x, y, z = np.zeros(4), np.zeros(4), np.zeros(4)
        orc.lib().orc_analytic_orbits(ctypes.byref(o), t[n], orc._d(x), orc._d(y), orc._d(z))
        for i in range(3):
            pos[3 * i:3 * i + 3, n] = x[i + 1], y[i + 1], z[i + 1]
    ph = rng.uniform(0, 2 * np.pi, 9)
    pos += 3e6 * np.sin(2 * np.pi * t[None, :] / (0.37 * 3.15581497632e7) + ph[:, None])
    return t, pos


@pytest.mark.parametrize("fmt", ["sangria", "phase", "frequency"])
def test_numeric_orbit_waveform_vs_oracle(gl, orc, fmt):
    """ucb_waveform with a numeric orbit (interpolate_orbits' splines over an
    orbit file's knots) against the oracle on the same knots: the average
    arm length bitwise, templates 1e-12 normwise for BW 32..2048; the orbit
    differs from the analytic one, so the analytic template differs."""
    T = synth.YEAR_T
    t, pos = _orbit_file(orc, T, 0.0)
    no = gl.numeric_orbit(t, pos[0::3], pos[1::3], pos[2::3])
    with orc.numeric_orbit(t, pos) as onum:
        assert no.L == onum.orbit.L and no.fstar == onum.orbit.fstar
        P = synth.draw_params(np.random.default_rng(9), 6, T)
        worst = 0.0
        for BW, p in zip((32, 64, 256, 512, 1024, 2048), P):
            got = gl.ucb_waveform_numeric(no, fmt, T, 0.0, p, 8, BW)
            ref = orc.ucb_waveform(p, T, BW, fmt=fmt, orbit=onum.orbit)
            n = 5 if fmt == "sangria" else 1
            worst = max(worst, normwise(got[:n], ref[:n]))
    ana = orc.ucb_waveform(P[-1], T, 2048, fmt=fmt)
    print(fmt, "numeric-orbit waveform worst normwise", worst)
    assert worst < 1e-12
    assert normwise(got[:1], ana[:1]) > 1e-6
    with pytest.raises(gl.GlassHipError):
        gl.ucb_waveform_numeric(no, fmt, 2 * T, 0.0, P[0], 8, 64)   # samples beyond the knots
