"""Pins of the UCB sampler oracle (oracle/orc_sampler.c) by checks that do
not retype it:

* the Jacobi eigen-decomposition against numpy.linalg.eigh (LAPACK) on
  random matrices and against a 50-digit mpmath eigensolve on graded
  (Fisher-like) ones, A v = lambda v, orthonormality, the order and sign
  conventions;
* the prior: its value for in-range parameters is minus the sum of the log
  prior volumes, out-of-range parameters give -inf, angles wrap into range;
* the proposals: Fisher jumps have covariance F^-1 / 8 (one of the 8
  eigen-directions at a time), fm-shift jumps are whole multiples of
  T/YEAR in {-3..3}\\{0} with the sky angle moved by U(-0.2, 0.2),
  psi-phi jumps keep psi - phi/2 fixed modulo pi, uniform prior draws fill
  the box, proposal types appear in proportion to their weights;
* the accept rule accepts with probability min(1, exp(dlogL / T)); the PT
  swap with min(1, exp(dlogL H)); the ladder update follows its formula in
  a hand-worked case.
"""
import math

import numpy as np
import pytest

from oracle import orc

YEAR = 3.15581497632e7
T2 = 62914560.0


def box(qmin=1000.0, W=512.0):
    return np.array([[qmin, qmin + W], [-1, 1], [0, 2 * np.pi], [-54, -46], [-1, 1], [0, np.pi], [0, 2 * np.pi],
                     [-1, 1]], dtype=np.float64)


def inside(rng, pr):
    return pr[:, 0] + rng.uniform(0.1, 0.9, 8) * (pr[:, 1] - pr[:, 0])


def graded_spd(rng, decades=12):
    Q, _ = np.linalg.qr(rng.standard_normal((8, 8)))
    lam = 10.0 ** rng.uniform(0, decades, 8)
    A = (Q * lam) @ Q.T
    return 0.5 * (A + A.T)


def test_jacobi_vs_numpy_random():
    rng = np.random.default_rng(1)
    for _ in range(200):
        A = rng.standard_normal((8, 8))
        A = A + A.T
        ev, la, sw = orc.jacobi_eigen8(A)
        ref = np.linalg.eigvalsh(A)
        nrm = np.linalg.norm(A)
        assert np.max(np.abs(la - ref)) < 1e-14 * nrm
        assert np.allclose(ev.T @ ev, np.eye(8), atol=1e-14)
        assert np.max(np.abs(A @ ev - ev * la)) < 1e-13 * nrm
        assert np.all(np.diff(la) >= 0)
        for j in range(8):
            assert ev[np.argmax(np.abs(ev[:, j])), j] > 0
        assert sw < 15


def test_jacobi_rotated_wide_spectrum():
    """A rotated matrix over 12 decades: every eigenvalue to eps * ||A||, as
    LAPACK (the rounding of the matrix itself allows no better)."""
    rng = np.random.default_rng(2)
    for _ in range(100):
        A = graded_spd(rng)
        _, la, _ = orc.jacobi_eigen8(A)
        ref = np.linalg.eigvalsh(A)
        assert np.max(np.abs(la - ref)) < 1e-14 * np.linalg.norm(A)


def test_jacobi_graded_relative_accuracy():
    """Fisher matrices in parameter units are graded, F = D H D with H well
    conditioned and D the parameter scales; Jacobi keeps every eigenvalue,
    the smallest included, to high relative accuracy (checked against a
    50-digit eigensolve)."""
    import mpmath
    mpmath.mp.dps = 50
    rng = np.random.default_rng(12)
    for _ in range(10):
        B = rng.standard_normal((8, 8))
        H = B @ B.T + 8 * np.eye(8)
        D = np.diag(10.0 ** rng.uniform(-3, 3, 8))
        F = D @ H @ D
        F = 0.5 * (F + F.T)
        _, la, _ = orc.jacobi_eigen8(F)
        ref = np.array(sorted(float(x) for x in mpmath.eigsy(mpmath.matrix(F.tolist()))[0]))
        # eps * cond(H) * small constant; an eps * norm(F) method would give ~1e-5 on the smallest
        assert np.max(np.abs(la - ref) / ref) < 1e-10, np.abs(la - ref) / ref


def test_jacobi_known_answer():
    # block diagonal: a 2x2 [[2,1],[1,2]] (eigenvalues 1, 3) and a diagonal
    A = np.diag([5.0, 7.0, 2.0, 2.0, 11.0, 13.0, 17.0, 19.0])
    A[2, 3] = A[3, 2] = 1.0
    ev, la, _ = orc.jacobi_eigen8(A)
    assert np.allclose(la, [1, 3, 5, 7, 11, 13, 17, 19], rtol=0, atol=1e-15)
    s2 = 1 / np.sqrt(2)
    assert np.allclose(ev[2:4, 0], [s2, -s2], atol=1e-15) or np.allclose(ev[2:4, 0], [-s2, s2], atol=1e-15)
    assert np.allclose(np.abs(ev[2:4, 1]), [s2, s2], atol=1e-15)


def test_prior_values_and_wraps():
    rng = np.random.default_rng(3)
    pr = box()
    lpv = np.log(pr[:, 1] - pr[:, 0])
    for _ in range(50):
        p = inside(rng, pr)
        lp, q = orc.ucb_prior(p, pr)
        assert abs(lp + lpv.sum()) < 1e-12 and np.array_equal(p, q)
    p = inside(rng, pr)
    p2 = p.copy()
    p2[5] += 3 * np.pi   # psi, period pi
    p2[6] -= 4 * np.pi   # phi0, period 2 pi
    p2[2] += 2 * np.pi   # sky longitude
    lp, q = orc.ucb_prior(p2, pr)
    assert abs(lp + lpv.sum()) < 1e-12
    assert np.allclose(q, p, atol=1e-12)
    for k, v in ((0, pr[0, 1] + 1), (1, 1.5), (3, -40.0), (4, -1.2), (7, 2.0), (0, np.nan)):
        p3 = p.copy()
        p3[k] = v
        assert orc.ucb_prior(p3, pr)[0] == -np.inf, k


def run_props(w, nrep, eig=None, seed0=1):
    rng = np.random.default_rng(4)
    pr = box()
    x = inside(rng, pr)
    if eig is None:
        ev, la = np.eye(8).reshape(1, 64), np.ones((1, 8))
    else:
        ev, la = eig
    P = orc.UcbProposer(w, T2)
    out = []
    seed = seed0
    for _ in range(nrep):
        sp = x.reshape(1, 8).copy()
        seed, n, npr, y, lpx, lpy, lq = P.propose(seed, sp, ev, la, pr)
        out.append((npr, y.copy(), lpx, lpy))
    return x, pr, out


def test_fisher_jump_covariance():
    rng = np.random.default_rng(5)
    F = graded_spd(rng, decades=4) * 1e2
    ev, la, _ = orc.jacobi_eigen8(F)
    w = np.zeros(9)
    w[4] = 1.0
    x, pr, out = run_props(w, 40000, (ev.reshape(1, 64), la.reshape(1, 8)))
    J = np.array([y - x for _, y, _, _ in out])
    C = J.T @ J / len(J)
    ref = np.linalg.inv(F) / 8
    # relative error of a sample covariance of 40000 draws
    assert np.max(np.abs(C - ref)) < 0.05 * np.max(np.abs(ref))
    assert all(npr == 4 for npr, _, _, _ in out)


def test_fm_shift_and_psi_phi_laws():
    w = np.zeros(9)
    w[5] = w[6] = 0.5
    x, pr, out = run_props(w, 8000)
    fm = T2 / YEAR
    nfm = npp = 0
    for npr, y, lpx, lpy in out:
        if npr == 5:
            nfm += 1
            k = (y[0] - x[0]) / fm
            assert abs(k - round(k)) < 1e-6 and round(k) in (-3, -2, -1, 1, 2, 3)
            d2 = y[2] - x[2]
            assert -0.2 <= d2 <= 0.2 + 1e-12 or (lpy > -np.inf and abs(d2 % (2 * np.pi)) < 0.21)
        else:
            assert npr == 6
            npp += 1
            # psi - phi0/2 is unchanged modulo pi/2 (both are wrapped into range by the prior)
            dpsi, dphi = y[5] - x[5], y[6] - x[6]
            r = (dpsi - 0.5 * dphi) / (np.pi / 2)
            assert abs(r - round(r)) < 1e-9 or abs((dpsi + 0.5 * dphi) / (np.pi / 2) -
                                                     round((dpsi + 0.5 * dphi) / (np.pi / 2))) < 1e-9
    assert abs(nfm / len(out) - 0.5) < 0.03


def test_uniform_draws_fill_box_and_weights():
    w = np.array([0.1, 0, 0, 0, 0.7, 0.1, 0.1, 0, 0])
    x, pr, out = run_props(w, 20000)
    cnt = np.bincount([o[0] for o in out], minlength=9) / len(out)
    assert np.all(cnt[[1, 2, 3, 7, 8]] == 0)
    assert np.allclose(cnt[[0, 4, 5, 6]], w[[0, 4, 5, 6]], atol=0.015)
    ys = np.array([o[1] for o in out if o[0] == 0])
    assert np.all(ys >= pr[:, 0]) and np.all(ys <= pr[:, 1])
    frac = (ys - pr[:, 0]) / (pr[:, 1] - pr[:, 0])
    assert np.all(np.abs(frac.mean(axis=0) - 0.5) < 0.05)


def test_accept_probability():
    P = orc.UcbProposer(np.ones(9), T2)
    rng = np.random.default_rng(6)
    for dlogl, T in ((-1.0, 1.0), (-3.0, 2.5), (0.5, 1.0)):
        n, acc = 20000, 0
        for s in rng.integers(1, 2 ** 31, n):
            _, a = P.accept(int(s), dlogl, 0.0, 10.0, T, -1.0, -1.0)
            acc += a
        p = min(1.0, math.exp(dlogl / T))
        assert abs(acc / n - p) < 4 * math.sqrt(p * (1 - p) / n) + 1e-3
    # the snr cut rejects, prior -inf rejects
    assert not P.accept(5, 100.0, 0.0, 4.9, 1.0, 0.0, 0.0)[1]
    assert not P.accept(5, 100.0, 0.0, 10.0, 1.0, 0.0, -np.inf)[1]


def test_ptmcmc_rate_with_lognorm():
    T = np.array([1.0, 2.0])
    rng = np.random.default_rng(7)
    logL = np.array([0.0, -1.0])
    norm = np.array([0.5, 0.2])  # the swap sees logL + logLnorm
    n, acc = 20000, 0
    for s in rng.integers(1, 2 ** 31, n):
        _, _, a = orc.ucb_ptmcmc([0, 1], T, logL, norm, [int(s), 1])
        acc += int(a[0])
    H = (T[1] - T[0]) / (T[1] * T[0])
    p = min(1.0, math.exp(((logL[1] + norm[1]) - (logL[0] + norm[0])) * H))
    assert abs(acc / n - p) < 4 * math.sqrt(p * (1 - p) / n) + 1e-3


def test_adapt_ladder_hand_case():
    T = np.array([1.0, 1.5, 2.5, 4.0])
    acc = np.array([1, 0, 1, 0], dtype=np.int32)
    mcmc = 30000
    out = orc.adapt_temperature_ladder(T, acc, mcmc)
    k = (10000.0 / (mcmc + 10000.0)) / 10.0
    # ic = 1: S = log(0.5) + (acc0 - acc1) k; ic = 2: S = log(1.0) + (acc1 - acc2) k
    t1 = 1.0 + 0.5 * math.exp(k)
    t2 = t1 + 1.0 * math.exp(-k)
    assert abs(out[1] - t1) < 1e-15 and abs(out[2] - t2) < 1e-15 and out[3] == 4.0 and out[0] == 1.0
    # the minimum ratio 1.1
    out = orc.adapt_temperature_ladder(np.array([1.0, 1.01, 3.0]), np.array([0, 1, 0], dtype=np.int32), 0)
    assert abs(out[1] - 1.1) < 1e-15


def test_fstat_draws_follow_the_density():
    """draw_from_fstatistic / jump_from_fstatistic land in a cell with
    probability proportional to its density; the frequency/sky of the draw
    lie in that cell; the proposal densities in the Hastings ratio are the
    log densities of the two cells."""
    rng = np.random.default_rng(9)
    nf, nt, nph = 4, 5, 6
    ten = rng.uniform(0.1, 3.0, (nf, nt, nph))
    ten *= ten.size / ten.sum()
    qmin = 1000.0
    pr = box(qmin=qmin, W=nf)  # the frequency prior spans the density's band
    d_f = 1.0 / T2
    fs, keep = orc.fstat_prop(ten, d_f, ten.max(), qmin)
    for npw in (1, 2):
        w = np.zeros(9)
        w[npw] = 1.0
        P = orc.UcbProposer(w, T2)
        x = inside(rng, pr)
        counts = np.zeros(ten.shape)
        seed = 77
        n = 30000
        for _ in range(n):
            sp = x.reshape(1, 8).copy()
            seed, _, npr, y, lpx, lpy, lq = P.propose(seed, sp, np.eye(8).reshape(1, 64), np.ones((1, 8)), pr,
                                                      fs)
            assert npr == npw
            i = int(np.floor((y[0] - qmin) / (d_f * T2)))
            j = int(np.floor((y[1] + 1) / (2.0 / nt)))
            k = int(np.floor(y[2] / (2 * np.pi / nph)))
            counts[i, j, k] += 1
            ix = int(np.floor((x[0] - qmin) / (d_f * T2)))
            jx = int(np.floor((x[1] + 1) / (2.0 / nt)))
            kx = int(np.floor(sp[0, 2] / (2 * np.pi / nph)))
            assert abs(lq - (np.log(ten[ix, jx, kx]) - np.log(ten[i, j, k]))) < 1e-12
            if npw == 2:  # the jump keeps the other parameters
                assert np.array_equal(y[3:], x[3:])
        expect = n * ten / ten.sum()
        chi2 = float(np.sum((counts - expect) ** 2 / expect))
        assert chi2 < ten.size + 6 * np.sqrt(2 * ten.size), chi2
