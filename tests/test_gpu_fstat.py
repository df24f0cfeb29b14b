"""GPU parity of the F-statistic (glass_hip_fstat_*, glass_hip_get_Fstat_logL)
against the oracle (oracle/orc_fstat.c, itself pinned by the F-statistic
identity).  Tolerances: F-statistics 1e-10 relative (the device template
against the oracle's, ~1e-14, amplified by the Gram inverse); the fitted
extrinsic parameters 1e-7 (the arccos/arctan2 of near-degenerate fits);
the proposal density 1e-9 relative to its maximum."""
import numpy as np
import pytest

from test_oracle_fstat import T2, inject, segment

pytestmark = pytest.mark.gpu


def noisy_segment(orc, seed, NFFT=512):
    o = orc.analytic_orbit()
    rng, qmin0, NFFT, C00 = segment(o, seed, NFFT)
    # two binaries plus white noise at the C00 level
    X, A, E = (np.zeros(NFFT, dtype=np.complex128) for _ in range(3))
    for _ in range(2):
        f0 = (qmin0 + rng.uniform(120, 380)) / T2
        theta, phi = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        M = orc.ucb_bandwidth(f0, 0.0, np.cos(theta), 1e-22, T2, NFFT, orbit=o)
        p = np.array([f0 * T2, np.cos(theta), phi, np.log(3e-22), rng.uniform(-1, 1), rng.uniform(0, np.pi),
                      rng.uniform(0, 2 * np.pi), 0.0])
        x, a, e = inject(o, qmin0, NFFT, p, M)
        X, A, E = X + x, A + a, E + e
    s = np.sqrt(C00 / 2)
    for ch in (X, A, E):
        ch += s * (rng.standard_normal(NFFT) + 1j * rng.standard_normal(NFFT))
    return o, qmin0, NFFT, X, A, E, C00, rng


def test_fstat_cells_match_oracle(gl, orc):
    import torch
    from glass_amd.fstat import Fstat, get_Fstat_logL
    o, qmin0, NFFT, X, A, E, C00, rng = noisy_segment(orc, 11)
    fd = orc.FstatData(T2, qmin0, X, A, E, C00, orbit=o)
    n = 300
    cells = np.stack([(qmin0 + rng.uniform(40, 470, n)) / T2, rng.uniform(-1e-17, 1e-16, n) * (rng.random(n) < 0.5),
                      np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * np.pi, n)], axis=1)
    F = Fstat(gl.analytic_orbit(), "sangria", T2, 0.0, qmin0, X, A, E, C00, max_cells=n)
    lx, la, fp = (t.cpu().numpy() for t in F.eval(torch.tensor(cells, device="cuda")))
    wl = wp = 0.0
    for c in range(n):
        rx, ra, rp, _ = fd.logl(*cells[c])
        wl = max(wl, abs(lx[c] - rx) / abs(rx), abs(la[c] - ra) / abs(ra))
        wp = max(wp, float(np.max(np.abs(fp[c] - rp))))
    print("max rel F", wl, "max abs Fparams", wp)
    assert wl < 1e-10 and wp < 1e-7
    # the per-call drop-in
    gx, ga, gp = get_Fstat_logL(gl.analytic_orbit(), "sangria", T2, 0.0, qmin0, X, A, E, C00, *cells[5])
    rx, ra, rp, _ = fd.logl(*cells[5])
    assert abs(gx - rx) < 1e-10 * abs(rx) and abs(ga - ra) < 1e-10 * abs(ra)
    F.close()


def test_fstat_proposal_matches_oracle(gl, orc):
    from glass_amd.fstat import Fstat
    o, qmin0, NFFT, X, A, E, C00, rng = noisy_segment(orc, 12)
    fd = orc.FstatData(T2, qmin0, X, A, E, C00, orbit=o)
    fmin, fmax = (qmin0 + 150) / T2, (qmin0 + 156) / T2
    F = Fstat(gl.analytic_orbit(), "sangria", T2, 0.0, qmin0, X, A, E, C00, max_cells=4096)
    ten, norm, maxp, nf, df = F.proposal(fmin, fmax, 20, 20)
    rten, rnorm, rmaxp, rnf, rdf = orc.fstat_proposal(fd, fmin, fmax, 20, 20)
    assert nf == rnf and df == rdf
    print("proposal: max", maxp, "norm", norm, "rel", np.max(np.abs(ten - rten)) / rmaxp)
    assert np.max(np.abs(ten - rten)) < 1e-9 * rmaxp
    assert abs(norm - rnorm) < 1e-10 * rnorm and abs(maxp - rmaxp) < 1e-9 * rmaxp
    F.close()
