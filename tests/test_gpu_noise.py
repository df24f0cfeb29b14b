"""GPU parity of the instrument-noise model, the noise likelihoods and the
batched instrument-noise sampler against the oracle (oracle/orc_noise.c).

Tolerances: covariance entries 1e-13 relative (the device's sin, cos and
pow against the C library's, ulp level); likelihoods 1e-12 relative (the
fused kernel's quotient inverse differs from the reference's log-domain
division by ~|log det| x 1e-16 per entry); random numbers bit for bit
(uniforms) and 4e-15 (normals, the device's log); sampler runs: the same
accept and swap decisions, parameters and logL within 1e-12.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T3M = 7864320.0


@pytest.fixture(scope="module")
def nz(gl):
    from glass_amd import noise
    return noise


@pytest.fixture(scope="module")
def fstar(orc):
    return orc.analytic_orbit().fstar


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def links(rng, tie=False):
    sa = rng.uniform(0.3, 3.0, 6) * 5.76e-30
    so = rng.uniform(0.3, 3.0, 6) * 1.28e-22
    if tie:
        so = np.repeat(so[::2], 2)
    return sa, so


def segment(orc, fstar, N, seed, fmin=1e-4, cadd=False):
    rng = np.random.default_rng(seed)
    f = fmin + np.arange(N) / T3M
    sa, so = links(rng, tie=True)
    C = orc.instrument_noise_model(fstar, f, sa, so)
    Cadd = None
    if cadd:
        # a smooth extra (foreground-like) covariance, positive definite
        g = 0.3 * C[0, 0] * np.exp(-((f - f.mean()) / (np.ptp(f) + 1e-30)) ** 2)
        Cadd = np.zeros_like(C)
        for a in range(3):
            for b in range(3):
                Cadd[a, b] = g * (1.0 if a == b else -0.5)
        C = C + Cadd
    L = np.linalg.cholesky(np.moveaxis(C, -1, 0))
    w = (rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))) / np.sqrt(2)
    d = np.einsum("nab,nb->na", L, w)
    return f, d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(), np.concatenate([sa, so]), Cadd


@pytest.mark.parametrize("nch", [1, 2, 3])
def test_instrument_model_dropin(nz, orc, fstar, nch):
    rng = np.random.default_rng(10 + nch)
    f = np.concatenate([[0.0], 1e-4 + np.arange(3000) / T3M, np.linspace(0.0099, 0.0102, 40)])
    sa, so = links(rng)
    init = rng.standard_normal((nch, nch, f.size))
    got = nz.instrument_noise_model(fstar, f, sa, so, nch=nch, C=init.copy())
    ref = orc.instrument_noise_model(fstar, f, sa, so, nch=nch, C=init.copy())
    print("nch", nch, "rel", rel(got[..., 1:], ref[..., 1:]))
    assert rel(got[..., 1:], ref[..., 1:]) < 1e-13
    assert np.array_equal(got[..., 0], ref[..., 0])  # the f = 0 bin: zeros and untouched entries


def test_instrument_model_wavelet_dropin(nz, orc, fstar):
    rng = np.random.default_rng(20)
    sa, so = links(rng)
    df, ND = 1.0 / 122880.0, 512 * 4096
    f = (5 + np.arange(300)) * df
    got, na = nz.instrument_noise_model_wavelet(fstar, f, df, ND, sa, so)
    ref, nr = orc.instrument_noise_model_wavelet(fstar, f, df, ND, sa, so)
    assert na == nr == 300
    assert rel(got, ref) < 1e-13


def test_likelihood_dropins(nz, orc, fstar):
    f, X, Y, Z, p, _ = segment(orc, fstar, 5000, 21)
    C = orc.instrument_noise_model(fstar, f, p[:6], p[6:])
    invC, ld = orc.invert_noise(C)
    # per-bin terms, their exactly rounded sum, and the rounding scale of the
    # reference's summation: noise_log_likelihood sums every channel pair
    # over all bins separately (fourier_nwip) and then the log-determinants,
    # so its rounding grows with sum_n sum_ab |r_a||r_b||invC_ab| + sum|log det|,
    # which towards 0.1 mHz (nearly singular XYZ covariance) is far above the
    # likelihood itself
    r = np.stack([X, Y, Z])
    re = np.einsum("an,bn->abn", r.real, r.real) + np.einsum("an,bn->abn", r.imag, r.imag)
    terms = -np.einsum("abn,abn->n", re, invC) - ld
    exact = math.fsum(terms)
    scale = float(np.sum(np.abs(r[:, None]) * np.abs(r[None, :]) * np.abs(invC)) + np.sum(np.abs(ld)))
    for name, cst in (("my_noise_log_likelihood", -3 * f.size * np.log(2 * np.pi)), ("noise_log_likelihood", 0.0)):
        got = getattr(nz, name)(X, Y, Z, invC, ld)
        ref = getattr(orc, name)(X, Y, Z, invC, ld)
        print(name, got, ref, "rel", abs(got - ref) / abs(ref), "vs exact", abs(got - exact - cst) / abs(ref),
              "scale/|logL|", scale / abs(ref))
        assert abs(got - (exact + cst)) < 1e-13 * abs(ref)
        assert abs(got - ref) < 1e-13 * scale
    NT, Nl = 512, 3
    rng = np.random.default_rng(22)
    fw = (40 + np.arange(NT * Nl)) / 122880.0
    Cw = orc.instrument_noise_model(fstar, fw, p[:6], p[6:])
    iw, lw = orc.invert_noise(Cw)
    Xw, Yw, Zw = rng.standard_normal((3, NT * Nl)) * np.sqrt(Cw[0, 0])
    got = nz.my_noise_log_likelihood_wavelet(Xw, Yw, Zw, iw, lw, NT, Nl, 2 ** 21)
    ref = orc.my_noise_log_likelihood_wavelet(Xw, Yw, Zw, iw, lw, NT, Nl, 2 ** 21)
    assert abs(got - ref) < 1e-12 * abs(ref)


def test_rand_r_stream(nz, orc):
    libc = ctypes.CDLL("libc.so.6")
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    libc.rand_r.restype = ctypes.c_int
    for seed in (1, 12345, 2 ** 31 + 7, 4294967295):
        u, g = nz.rand_stream(seed, 4000)
        s = ctypes.c_uint(seed)
        want = np.array([libc.rand_r(ctypes.byref(s)) for _ in range(4000)]) / 2147483647.0
        assert np.array_equal(u, want), seed
        ou, og = orc.rand_stream(seed, 4000)
        assert np.array_equal(u, ou)
        assert rel(g, og) < 4e-15, seed


@pytest.mark.parametrize("cadd", [False, True])
def test_batched_logl(nz, orc, fstar, cadd):
    import torch
    N = 20000
    f, X, Y, Z, p, Cadd = segment(orc, fstar, N, 23, cadd=cadd)
    plan = nz.NoisePlan(fstar, f, X, Y, Z, nladder=5, ntemp=4, Cadd=Cadd)
    rng = np.random.default_rng(24)
    P = np.stack([p * rng.uniform(0.9, 1.1, 12) for _ in range(plan.nslot)])
    P[0] = p
    got = plan.logl(torch.tensor(P, device="cuda")).cpu().numpy()
    od = orc.NoiseData(fstar, f, X, Y, Z, Cadd)
    ref = np.array([od.logl(q) for q in P])
    print("batched logL rel", rel(got, ref), plan.info())
    assert rel(got, ref) < 1e-12
    # fewer parameter sets than the plan holds
    got3 = plan.logl(torch.tensor(P[:3], device="cuda")).cpu().numpy()
    assert rel(got3, ref[:3]) < 1e-12
    plan.close()


def oracle_sweeps(orc, od, st, temps, nladder, ntemp, ncalls, nsweep):
    """The reference's loop: per slot ncalls noise_instrument_model_mcmc
    calls on its state, then noise_ptmcmc per ladder."""
    x, lg, seeds, idx = st["params"].copy(), st["logL"].copy(), st["seeds"].astype(np.int64), st["index"].copy()
    nacc = np.zeros(len(seeds), dtype=np.int64)
    swaps = np.zeros(len(seeds), dtype=np.int64)
    for _ in range(nsweep):
        for s in range(len(seeds)):
            stt = idx[s]
            for _ in range(ncalls):
                x[stt], lg[stt], seeds[s], na = od.mcmc_call(x[stt], lg[stt], seeds[s], temps[s])
                nacc[s] += na
        for l in range(nladder):
            o = slice(l * ntemp, (l + 1) * ntemp)
            i2, s2, a2 = orc.noise_ptmcmc(idx[o], temps[o], lg, seeds[o])
            idx[o], seeds[o] = i2, s2
            swaps[o] += a2
    return x, lg, seeds, idx, nacc, swaps


@pytest.mark.parametrize("graph", [False, True])
def test_mcmc_sweeps_match_oracle(nz, orc, fstar, graph):
    # a well-conditioned band (the oracle's term-by-term running sum loses
    # digits where the XYZ covariance is nearly singular, below ~0.5 mHz)
    N, nladder, ntemp, ncalls = 800, 2, 4, 2
    f, X, Y, Z, p, Cadd = segment(orc, fstar, N, 25, fmin=1e-3, cadd=True)
    plan = nz.NoisePlan(fstar, f, X, Y, Z, nladder=nladder, ntemp=ntemp, Cadd=Cadd)
    rng = np.random.default_rng(26)
    S = plan.nslot
    P = np.stack([p * rng.uniform(0.95, 1.05, 12) for _ in range(S)])
    P[:, 6:] = np.repeat(P[:, 6::2], 2, axis=1)
    seeds = rng.integers(1, 2 ** 32, S, dtype=np.uint64).astype(np.uint32)
    temps = np.tile(1.3 ** np.arange(ntemp), nladder)
    plan.set_state(P, seeds, temps)
    st0 = plan.state()
    od = orc.NoiseData(fstar, f, X, Y, Z, Cadd)
    ref0 = np.array([od.logl(q) for q in P])
    assert rel(st0["logL"], ref0) < 1e-12
    nsweep = 2
    for _ in range(nsweep):
        plan.mcmc(ncalls, swap=True, graph=graph)
    st = plan.state()
    x, lg, sd, idx, nacc, swaps = oracle_sweeps(orc, od, st0, temps, nladder, ntemp, ncalls, nsweep)
    print("accepted", st["nacc"], "swaps", st["swaps"])
    assert np.array_equal(st["nacc"], nacc)
    assert np.array_equal(st["swaps"], swaps)
    assert np.array_equal(st["index"], idx)
    assert np.array_equal(st["seeds"], sd.astype(np.uint32))
    assert rel(st["params"], x) < 1e-12
    assert rel(st["logL"], lg) < 1e-12
    assert st["nacc"].sum() > 0
    plan.close()


def test_bad_plan_arguments(nz, gl, fstar):
    f = np.array([0.0, 1e-4, 2e-4])
    X = np.zeros(3, dtype=np.complex128)
    with pytest.raises(gl.GlassHipError):
        nz.NoisePlan(fstar, f, X, X, X, nladder=1, ntemp=2)
    with pytest.raises(gl.GlassHipError):
        nz.instrument_noise_model(fstar, f[1:], np.ones(6), np.ones(6), nch=4, C=np.zeros((4, 4, 2)))
