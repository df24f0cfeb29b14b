"""Pins of the noise oracle (oracle/orc_noise.c) by checks that do not retype
its formulas:

* the instrument model at equal link levels equals the reference's
  independent closed forms (get_noise_levels("sangria") with XYZnoise_FF /
  XYZcross_FF, utils/src/glass_lisa.c:1082-1135, restated separately in the
  oracle), up to the 12 significant digits of the unit constants the
  reference hard-codes in generate_instrument_noise_model;
* relabelling the spacecraft cyclically (1->2->3->1) permutes X, Y, Z, and
  the model is linear in every link parameter -- these fix which link enters
  which covariance entry, which equal levels cannot;
* the likelihoods equal numpy's per-bin inverse and log-determinant;
* the random numbers are the C library's rand_r, and the polar normals have
  the moments of N(0, 1);
* the PT swap accepts with the Metropolis probability min(1, exp(dlogL H));
* the MCMC cold chain, started away from the injected levels, moves to them.
"""
import ctypes

import numpy as np
import pytest

from oracle import orc

T3M = 7864320.0  # the reference noise_mcmc example's --duration
PERM = [2, 3, 5, 4, 1, 0]  # links (12,21,23,32,13,31) -> (23,32,31,13,21,12)


@pytest.fixture(scope="module")
def fstar():
    return orc.analytic_orbit().fstar


def band(fmin, n, T=T3M):
    return fmin + np.arange(n) / T


def rand_links(rng):
    return rng.uniform(0.3, 3.0, 6) * 5.76e-30, rng.uniform(0.3, 3.0, 6) * 1.28e-22


def test_equal_links_match_sangria_closed_form(fstar):
    f = band(1e-4, 3000)
    C = orc.instrument_noise_model(fstar, f, np.full(6, 5.76e-30), np.full(6, 1.28e-22))
    C2, _, _ = orc.xyz_noise(T3M, 1e-4, f.size)
    assert np.max(np.abs(C - C2) / np.abs(C2)) < 5e-12


def test_equal_links_above_acc_cutoff(fstar):
    """Above 10 mHz the acceleration term is dropped: only the OMS part of
    the closed form remains."""
    from glass_amd.synth import noise_levels_sangria
    f = np.linspace(0.0101, 0.02, 50)
    C = orc.instrument_noise_model(fstar, f, np.full(6, 5.76e-30), np.full(6, 1.28e-22))
    _, Sop = noise_levels_sangria(f)
    x = f / fstar
    dg = 16.0 * np.sin(x) ** 2 * Sop / 4.0
    cr = -8.0 * np.sin(x) ** 2 * np.cos(x) * Sop / 4.0
    assert np.max(np.abs(C[0, 0] - dg) / np.abs(dg)) < 5e-12
    assert np.max(np.abs(C[0, 1] - cr) / np.abs(cr)) < 5e-12


def test_cyclic_relabelling_permutes_channels(fstar):
    rng = np.random.default_rng(3)
    f = band(2e-4, 500)
    sa, so = rand_links(rng)
    C = orc.instrument_noise_model(fstar, f, sa, so)
    Cp = orc.instrument_noise_model(fstar, f, sa[PERM], so[PERM])
    # channel a of the relabelled constellation is channel a+1 of the original
    for a in range(3):
        for b in range(3):
            ref = C[(a + 1) % 3, (b + 1) % 3]
            assert np.allclose(Cp[a, b], ref, rtol=1e-14, atol=0), (a, b)
    # a slip in the link assignment (two links exchanged) breaks it
    for i, j in ((0, 1), (0, 2), (2, 5)):
        bad = list(PERM)
        bad[i], bad[j] = bad[j], bad[i]
        Cb = orc.instrument_noise_model(fstar, f, sa[bad], so[bad])
        worst = max(np.max(np.abs(Cb[a, b] - C[(a + 1) % 3, (b + 1) % 3]) / np.abs(C[(a + 1) % 3, (b + 1) % 3]))
                    for a in range(3) for b in range(3))
        assert worst > 1e-6, (i, j, worst)


def test_linear_in_link_parameters(fstar):
    rng = np.random.default_rng(4)
    f = band(1e-4, 400)
    a1, o1 = rand_links(rng)
    a2, o2 = rand_links(rng)
    C1 = orc.instrument_noise_model(fstar, f, a1, o1)
    C2 = orc.instrument_noise_model(fstar, f, a2, o2)
    C12 = orc.instrument_noise_model(fstar, f, a1 + a2, o1 + o2)
    assert np.allclose(C12, C1 + C2, rtol=1e-13, atol=0)


def test_one_and_two_channel_cases(fstar):
    f = band(1e-4, 300)
    sa, so = np.full(6, 5.76e-30), np.full(6, 1.28e-22)
    C3 = orc.instrument_noise_model(fstar, f, sa, so, nch=3)
    C1 = orc.instrument_noise_model(fstar, f, sa, so, nch=1)
    assert np.max(np.abs(C1[0, 0] - C3[0, 0]) / C3[0, 0]) < 1e-13
    # A = (Z - X)/sqrt2, E = (X - 2Y + Z)/sqrt6: var(A) = C_XX - C_XZ for
    # equal links (the two-channel model is the A/E PSD)
    C2 = orc.instrument_noise_model(fstar, f, sa, so, nch=2, C=np.full((2, 2, f.size), 8.0))
    varA = C3[0, 0] - C3[0, 2]
    assert np.max(np.abs(C2[0, 0] - varA) / varA) < 1e-12
    assert np.all(C2[0, 1] == 0) and np.all(C2[1, 0] == 2.0)  # untouched entry only divided by 4


def test_wavelet_model_is_layer_centres(fstar):
    rng = np.random.default_rng(5)
    sa, so = rand_links(rng)
    df, ND, N = 1.0 / 122880.0, 512 * 4096, 64
    f = (37 + np.arange(N)) * df
    C, na = orc.instrument_noise_model_wavelet(fstar, f, df, ND, sa, so)
    assert na == N
    Cf = orc.instrument_noise_model(fstar, f, sa, so)
    assert np.allclose(C, Cf / ND, rtol=1e-15, atol=0)


def cov_and_data(fstar, N, seed, Cadd=None):
    rng = np.random.default_rng(seed)
    f = band(1e-4, N)
    sa, so = rand_links(rng)
    C = orc.instrument_noise_model(fstar, f, sa, so)
    if Cadd is not None:
        C = C + Cadd
    L = np.linalg.cholesky(np.moveaxis(C, -1, 0))
    w = (rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))) / np.sqrt(2)
    d = np.einsum("nab,nb->na", L, w)
    return f, C, d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(), np.concatenate([sa, so])


def numpy_logl(C, X, Y, Z):
    Cn = np.moveaxis(C, -1, 0)
    inv = np.linalg.inv(Cn)
    sign, ld = np.linalg.slogdet(Cn)
    r = np.stack([X, Y, Z], axis=1)
    q = np.einsum("na,nab,nb->n", r.conj(), inv, r).real
    return float(-np.sum(q) - np.sum(ld)), float(np.sum(ld))


def test_likelihoods_vs_numpy(fstar):
    f, C, X, Y, Z, _ = cov_and_data(fstar, 300, 6)
    invC, ld = orc.invert_noise(C)
    ref, _ = numpy_logl(C, X, Y, Z)
    N = f.size
    my = orc.my_noise_log_likelihood(X, Y, Z, invC, ld)
    plain = orc.noise_log_likelihood(X, Y, Z, invC, ld)
    assert abs(my - (ref - 3 * N * np.log(2 * np.pi))) < 1e-10 * abs(my)
    assert abs(plain - ref) < 1e-10 * abs(plain)


def test_wavelet_likelihood_vs_numpy(fstar):
    rng = np.random.default_rng(7)
    NT, Nl = 16, 5
    n = NT * Nl
    sa, so = rand_links(rng)
    f = (20 + np.arange(n)) * (1.0 / 122880.0)
    C = orc.instrument_noise_model(fstar, f, sa, so)
    invC, ld = orc.invert_noise(C)
    X, Y, Z = rng.standard_normal((3, n)) * np.sqrt(C[0, 0])
    got = orc.my_noise_log_likelihood_wavelet(X, Y, Z, invC, ld, NT, Nl, 12345)
    Cn = np.moveaxis(C, -1, 0)
    r = np.stack([X, Y, Z], axis=1)
    q = np.einsum("na,nab,nb->n", r, np.linalg.inv(Cn), r)
    ref = -0.5 * np.sum(q) - 0.5 * np.sum(np.linalg.slogdet(Cn)[1]) - 0.5 * 3 * 12345 * np.log(2 * np.pi)
    assert abs(got - ref) < 1e-10 * abs(ref)


def test_rand_is_libc_rand_r():
    libc = ctypes.CDLL("libc.so.6")
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    libc.rand_r.restype = ctypes.c_int
    s = ctypes.c_uint(12345)
    want = np.array([libc.rand_r(ctypes.byref(s)) for _ in range(1000)]) / 2147483647.0
    u, _ = orc.rand_stream(12345, 1000)
    assert np.array_equal(u, want)


def test_polar_normals_moments():
    _, g = orc.rand_stream(777, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1) < 0.01
    assert abs(np.mean(g ** 4) - 3) < 0.05


def test_ptswap_acceptance_probability():
    """Two slots, fixed states: the swap rate is min(1, exp(dlogL H))."""
    T = np.array([1.0, 1.5])
    rng = np.random.default_rng(8)
    for dlogl in (-2.0, -0.3, 0.7):
        logL = np.array([0.0, dlogl])
        n, acc = 20000, 0
        seeds = rng.integers(1, 2 ** 31, n)
        for s in seeds:
            _, _, a = orc.noise_ptmcmc([0, 1], T, logL, [int(s), 1])
            acc += int(a[0])
        H = (T[1] - T[0]) / (T[1] * T[0])
        p = min(1.0, np.exp(dlogl * H))
        assert abs(acc / n - p) < 4 * np.sqrt(p * (1 - p) / n) + 1e-3, (dlogl, acc / n, p)


def test_mcmc_moves_to_injection(fstar):
    """The cold chain, started with every level 40% high, climbs to the
    likelihood of the injected levels (a flipped Hastings ratio, or a
    temperature applied the wrong way, walks away from them)."""
    rng = np.random.default_rng(9)
    f = band(1e-4, 1500)
    sa, so = rand_links(rng)
    so = np.repeat(so[::2], 2)  # one OMS level per link pair, as the sampler keeps them
    truth = np.concatenate([sa, so])
    C = orc.instrument_noise_model(fstar, f, sa, so)
    L = np.linalg.cholesky(np.moveaxis(C, -1, 0))
    w = (rng.standard_normal((f.size, 3)) + 1j * rng.standard_normal((f.size, 3))) / np.sqrt(2)
    d = np.einsum("nab,nb->na", L, w)
    nd = orc.NoiseData(fstar, f, d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy())
    l_true = nd.logl(truth)
    x = truth * 1.4
    lg = l0 = nd.logl(x)
    seed, nacc = 4242, 0
    for _ in range(150):
        x, lg, seed, na = nd.mcmc_call(x, lg, seed, 1.0)
        nacc += na
    print("logL start", l0 - l_true, "end", lg - l_true, "accepted", nacc)
    assert nacc > 30
    assert abs(lg - nd.logl(x)) < 1e-9 * abs(lg)
    assert l0 < l_true - 100
    assert lg > l_true - 30
