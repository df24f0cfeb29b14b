"""The reference's UCB file formats (glass_amd.files over include/glass_hip_io.h),
host code only, so these run without a GPU.

Pins that do not restate the C++:
* every line is compared with Python's own %-formatting of the same values
  (an independent printf implementation, correctly rounded like glibc's);
* the verification-binary fixtures (data files shipped with the reference,
  ucb/etc/sources/verification/*.dat, copied under tests/golden/) are parsed
  by Python's float(); the amplitude is checked against the textbook GW
  strain amplitude A = 2 (G Mc / c^3)^(5/3) (pi f)^(2/3) c / D, and the drawn
  phase and polarisation against glibc's own rand_r through ctypes;
* files written and read back keep every value to its printed precision,
  and the injection count follows the reference's read loop (a file without
  a final newline loses its last source).
"""
import ctypes
import glob
import math
import os

import numpy as np
import pytest

from glass_amd import files

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "verification")
# template.dat is the header-only template of the format (an error case)
VB = sorted(p for p in glob.glob(os.path.join(GOLD, "*.dat")) if not p.endswith("template.dat"))
T = 62914560.0


def named(p, T):
    return dict(f0=p[0] / T, dfdt=p[7] / T**2, amp=math.exp(p[3]), phi=p[2], costheta=p[1], cosi=p[4], psi=p[5],
                phiref=p[6])


def py_source_line(p, T):
    s = named(p, T)
    return ("%.16g " % s["f0"]) + "".join("%.12g " % s[k] for k in
                                          ("dfdt", "amp", "phi", "costheta", "cosi", "psi", "phiref"))


def random_params(rng, n):
    p = np.empty((n, 8))
    p[:, 0] = rng.uniform(1e-4, 2e-2, n) * T
    p[:, 1] = rng.uniform(-1, 1, n)
    p[:, 2] = rng.uniform(0, 2 * np.pi, n)
    p[:, 3] = np.log(10 ** rng.uniform(-24, -21, n))
    p[:, 4] = rng.uniform(-1, 1, n)
    p[:, 5] = rng.uniform(0, np.pi, n)
    p[:, 6] = rng.uniform(0, 2 * np.pi, n)
    p[:, 7] = rng.uniform(-1e-17, 1e-15, n) * T * T
    return p


def test_source_line_matches_printf():
    rng = np.random.default_rng(1)
    for p in random_params(rng, 200):
        assert files.print_source_params(p, T) == py_source_line(p, T)


def test_state_lines_match_printf():
    rng = np.random.default_rng(2)
    for _ in range(50):
        step, nlive = int(rng.integers(-1000, 100000)), int(rng.integers(0, 30))
        logL, logLnorm, t0 = rng.normal(0, 1e4), rng.normal(0, 1e6), rng.uniform(0, 1e8)
        assert files.print_chain_state(step, nlive, logL, logLnorm, t0) == \
            "%i %i %g %g %.12g \n" % (step, nlive, logL, logLnorm, t0)
        eta = rng.uniform(0.5, 2, 3)
        assert files.print_psd_state(step, logL, logLnorm, eta) == \
            "%i %g %g " % (step, logL, logLnorm) + "".join("%g " % e for e in eta) + "\n"
        v = rng.normal(0, 1e3, 16)
        assert files.value_line(step, v) == "%i " % step + "".join("%g " % x for x in v) + "\n"


def test_scan_round_trip():
    rng = np.random.default_rng(3)
    for p in random_params(rng, 200):
        line = files.print_source_params(p, T)
        q, n = files.scan_source_params(line, T)
        assert n == len(line.rstrip())
        s, r = named(p, T), named(q, T)
        # each value survives to its printed precision (f0 and fdot through
        # the T scalings, A through params[3] = ln A: ~|ln A| ulp)
        assert float("%.16g" % s["f0"]) == pytest.approx(r["f0"], rel=3e-16, abs=0)
        assert float("%.12g" % s["amp"]) == pytest.approx(r["amp"], rel=2e-14, abs=0)
        for k in ("dfdt", "phi", "costheta", "cosi", "psi", "phiref"):
            assert float("%.12g" % s[k]) == pytest.approx(r[k], rel=3e-16, abs=0)


@pytest.mark.parametrize("path", VB, ids=[os.path.basename(p) for p in VB])
def test_verification_binary(path):
    NFFT, iseed = 1024, 1234
    p, qmin, fmin, fmax = files.read_verification_binary(path, T, NFFT, iseed)
    with open(path) as fh:
        fh.readline()
        f0, fdot, costheta, phi, m1, m2, cosi, d = (float(x) for x in fh.readline().split()[:8])
    s = named(p, T)
    assert p[0] == f0 * T and p[7] == fdot * T * T
    assert (s["costheta"], s["phi"], s["cosi"]) == (costheta, phi, cosi)
    # textbook strain amplitude with G Msun / c^3 = 4.92569043916e-6 s and
    # 1 pc = 3.0856775815e16 m
    Mc = (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2
    A = 2 * (Mc * 4.92569043916e-6) ** (5 / 3) * (math.pi * f0) ** (2 / 3) / (d * 1e3 * 3.0856775815e16 / 299792458.)
    assert s["amp"] == pytest.approx(A, rel=1e-13)
    # phase and polarisation: glibc rand_r seeded with iseed
    libc = ctypes.CDLL(None)
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    seed = ctypes.c_uint(iseed)
    u1 = libc.rand_r(ctypes.byref(seed)) / 2147483647.0
    u2 = libc.rand_r(ctypes.byref(seed)) / 2147483647.0
    assert p[6] == u1 * math.pi * 2.0 and p[5] == u2 * math.pi / 4.0
    assert qmin == int((f0 - (NFFT // 2) / T) * T) and fmin == qmin / T and fmax == (qmin + NFFT) / T


def test_injection_file(tmp_path):
    rng = np.random.default_rng(4)
    rows = []
    for k in range(7):
        psi = rng.uniform(0, 2 * np.pi) if k % 2 else rng.uniform(0, np.pi)
        phi0 = rng.uniform(0, 4 * np.pi)
        rows.append([rng.uniform(1e-3, 1e-2), rng.uniform(0, 1e-16), rng.uniform(-np.pi / 2, np.pi / 2),
                     rng.uniform(0, 2 * np.pi), 10 ** rng.uniform(-23, -21), rng.uniform(0, np.pi), psi, phi0])
    text = "".join(" ".join("%.17g" % v for v in r) + "\n" for r in rows)
    f = tmp_path / "inj.dat"
    f.write_text(text)
    p = files.read_injection_file(f, T)
    assert len(p) == len(rows)
    for r, q in zip(rows, p):
        f0, dfdt, theta, phi, amp, iota, psi, phi0 = r
        if psi > math.pi:
            psi -= math.pi
        if phi0 > 2 * math.pi:
            phi0 -= 2 * math.pi
        s = named(q, T)
        assert q[0] == f0 * T and q[7] == dfdt * T * T and s["phi"] == phi
        assert s["costheta"] == math.cos(math.pi / 2 - theta) and s["cosi"] == math.cos(iota)
        assert s["psi"] == psi and s["phiref"] == phi0
        assert s["amp"] == pytest.approx(amp, rel=1e-15)
    # the reference's count drops a last line that has no newline
    f.write_text(text.rstrip("\n"))
    assert len(files.read_injection_file(f, T)) == len(rows) - 1


def test_chain_state_file_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    p = random_params(rng, 5)
    eta = np.array([1.01, 0.98, 1.2])
    path = tmp_path / "chain_state_3.dat"
    files.save_chain_state(path, 1234.5678, 42, -987.654321, 1.5e5, 12345.0, eta, p, T)
    expect = "%.12g\n" % 1234.5678 + "%i %i %g %g %.12g \n" % (42, 5, -987.654321, 1.5e5, 12345.0)
    expect += "%i %g %g " % (42, -987.654321, 1.5e5) + "".join("%g " % e for e in eta) + "\n"
    expect += "".join(py_source_line(q, T) + "\n" for q in p)
    assert path.read_text() == expect
    r = files.restore_chain_state(path, T, 3)
    assert r["step"] == 42 and r["nlive"] == 5 and r["logLmax"] == float("%.12g" % 1234.5678)
    assert r["logL"] == float("%g" % -987.654321) and r["t0"] == 12345.0
    assert np.array_equal(r["eta"], [float("%g" % e) for e in eta])
    for q, s in zip(r["params"], p):
        assert files.print_source_params(q, T) == files.print_source_params(s, T)


def test_bad_files(tmp_path):
    from glass_amd import GlassHipError
    with pytest.raises(GlassHipError):
        files.read_verification_binary(os.path.join(GOLD, "template.dat"), T, 512, 1)
    with pytest.raises(GlassHipError):
        files.read_verification_binary(tmp_path / "missing.dat", T, 512, 1)
    with pytest.raises(GlassHipError):
        files.scan_source_params("1 2 3", T)
    path = tmp_path / "s.dat"
    files.save_chain_state(path, 0.0, 1, 0.0, 0.0, 0.0, [1.0], random_params(np.random.default_rng(6), 3), T)
    with pytest.raises(GlassHipError):
        files.restore_chain_state(path, T, 1, maxsrc=2)
    # a checkpoint cut inside its last source line
    text = path.read_text()
    path.write_text(text[:text.rstrip().rfind(" ")])
    with pytest.raises(GlassHipError):
        files.restore_chain_state(path, T, 1)
