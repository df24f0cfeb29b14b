"""Pins of the F-statistic oracle (oracle/orc_fstat.c) that do not retype it:

* the F-statistic identity (Cornish & Crowder 2005): on noise-free data made
  of one binary with the cell's own intrinsic parameters (f0, fdot, sky) the
  four filters span the signal, so F = (d|d)/2 for every amplitude,
  inclination, polarisation and phase -- for X and for A+E;
* F is quadratic in the data;
* the fit recovers the injected cos(iota) and the amplitude up to one fixed
  factor;
* the 4x4 inverse against numpy.
"""
import numpy as np
import pytest

from oracle import orc

T2 = 62914560.0


@pytest.fixture(scope="module")
def orbit():
    return orc.analytic_orbit()


def segment(orbit, seed, NFFT=512, fq=3.1e-3):
    rng = np.random.default_rng(seed)
    qmin0 = int(fq * T2)
    C, _, _ = orc.xyz_noise(T2, qmin0 / T2, NFFT, orbit=orbit)
    return rng, qmin0, NFFT, C[0, 0].copy()


def inject(orbit, qmin0, NFFT, p, M):
    """Noise-free data of params p on the filters' M bins."""
    h = orc.ucb_waveform(p, T2, M, NI=2, orbit=orbit)
    X, A, E = (np.zeros(NFFT, dtype=np.complex128) for _ in range(3))
    q = int(p[0])
    for i in range(M):
        k = q - qmin0 + i - M // 2
        if 0 < k < NFFT:
            X[k], A[k], E[k] = h[0][i], h[3][i], h[4][i]
    return X, A, E


def inner(a, b, C00):
    return 4.0 * float(np.sum((a.real * b.real + a.imag * b.imag) / C00))


def test_fstat_identity(orbit):
    rng, qmin0, NFFT, C00 = segment(orbit, 1)
    for trial in range(6):
        f0 = (qmin0 + rng.uniform(150, 350)) / T2
        theta, phi = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        fdot = rng.uniform(-1e-17, 1e-16) if trial % 2 else 0.0
        M = orc.ucb_bandwidth(f0, fdot, np.cos(theta), 1e-22, T2, NFFT, orbit=orbit)
        p = np.array([f0 * T2, np.cos(theta), phi, np.log(10 ** rng.uniform(-22.5, -21.5)),
                      rng.uniform(-1, 1), rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi), fdot * T2 * T2])
        X, A, E = inject(orbit, qmin0, NFFT, p, M)
        fd = orc.FstatData(T2, qmin0, X, A, E, C00, orbit=orbit)
        lx, lae, fp, Mf = fd.logl(f0, fdot, theta, phi)
        assert Mf == M
        dd_ae = inner(A, A, C00) + inner(E, E, C00)
        dd_x = inner(X, X, C00)
        print(trial, M, lae / (0.5 * dd_ae), lx / (0.5 * dd_x))
        assert abs(lae - 0.5 * dd_ae) < 1e-9 * dd_ae
        assert abs(lx - 0.5 * dd_x) < 1e-9 * dd_x
        # quadratic in the data
        fd3 = orc.FstatData(T2, qmin0, 3 * X, 3 * A, 3 * E, C00, orbit=orbit)
        assert abs(fd3.logl(f0, fdot, theta, phi)[1] - 9 * lae) < 1e-12 * 9 * lae


def test_fstat_fit_recovers_inclination_and_amplitude(orbit):
    rng, qmin0, NFFT, C00 = segment(orbit, 2)
    ratios = []
    for _ in range(6):
        f0 = (qmin0 + rng.uniform(150, 350)) / T2
        theta, phi = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        M = orc.ucb_bandwidth(f0, 0.0, np.cos(theta), 1e-22, T2, NFFT, orbit=orbit)
        A_inj = 10 ** rng.uniform(-22.5, -21.5)
        cosi = rng.uniform(-0.9, 0.9)
        p = np.array([f0 * T2, np.cos(theta), phi, np.log(A_inj), cosi, rng.uniform(0, np.pi),
                      rng.uniform(0, 2 * np.pi), 0.0])
        X, A, E = inject(orbit, qmin0, NFFT, p, M)
        _, _, fp, _ = orc.FstatData(T2, qmin0, X, A, E, C00, orbit=orbit).logl(f0, 0.0, theta, phi)
        assert abs(abs(np.cos(fp[1])) - abs(cosi)) < 1e-6, (np.cos(fp[1]), cosi)
        ratios.append(fp[0] / A_inj)
    assert np.ptp(ratios) < 1e-6 * np.mean(ratios)


def test_invert4_vs_numpy():
    rng = np.random.default_rng(3)
    for _ in range(100):
        B = rng.standard_normal((4, 4))
        m = B @ B.T + 0.1 * np.eye(4)
        assert np.allclose(orc.invert4(m), np.linalg.inv(m), rtol=1e-10, atol=1e-12 * np.abs(np.linalg.inv(m)).max())


def test_ucb_fdot_chirp_formula():
    # 96/5 pi^(8/3) (G Mc/c^3)^(5/3) f^(11/3), G Msun/c^3 = 4.92569043916e-6 s
    for Mc, f in ((0.3, 2e-3), (1.0, 1e-2)):
        M = Mc * 4.92569043916e-6
        ref = 96.0 / 5.0 * np.pi ** (8 / 3) * M ** (5 / 3) * f ** (11 / 3)
        assert abs(orc.ucb_fdot(Mc, f) - ref) < 1e-13 * ref
