"""Pins the oracle's wavelet-domain restatement (oracle/orc_wdm.c).

What fixes each function independently of the restatement:
  * incomplete_beta          -> scipy.special.betainc
  * WDM forward / inverse    -> the reference's own known answers
    (apps/src/fft_wdm_unit_tests.c:547-590: the WDM of an impulse maps back
    to the all-ones spectrum and to the impulse), at its NF=1536, NT=30 grid
    and at power-of-two grids; round trips of random series
  * by-layers transform      -> for a bin-centred tone, the heterodyned
    by-layers pixels times sqrt(L+1) equal the full transform's pixels in
    those layers, to rounding (odd and even first layer)
  * splines                  -> scipy CubicSpline (natural), lookups by
    np.searchsorted
  * amplitude/phase split    -> M = A cos(phi + phiR), Mf = -A sin(phi + phiR)
  * the wavelet waveform     -> (a) the full-cadence series through the full
    transform (the reference's ucb_waveform_wavelet_slow route) agrees with
    the heterodyned template to ~1e-6; (b) its series agrees with the
    reference's Fourier-domain fast/slow waveform (an independent model) to
    a match > 0.99; (c) linearity in A and the psi -> psi + pi/2 and
    phi0 -> phi0 + pi sign flips
  * wavelet logL / snr       -> closed forms evaluated with numpy
"""
import numpy as np
import pytest

from glass_amd import synth

T1 = synth.YEAR_T


def test_incomplete_beta_vs_scipy(orc):
    from scipy.special import betainc
    x = np.linspace(0.0, 1.0, 101)
    got = np.array([orc.incomplete_beta(6.0, 6.0, v) for v in x])
    # the series stops at terms below 1e-8
    assert np.max(np.abs(got - betainc(6.0, 6.0, x))) < 1e-7


@pytest.mark.parametrize("NF,NT", [(1536, 30), (64, 32), (96, 64)])
def test_wdm_impulse_known_answers(orc, NF, NT):
    """fft_wdm_unit_tests.c:547-590: WDM(impulse) -> inverse -> FFT of the
    impulse (all ones), and back to the impulse (atol 1e-10, rtol 1e-5)."""
    w = orc.Wdm(NF, NT, 7680.0)
    ND = NF * NT
    x = np.zeros(ND)
    x[0] = 1.0
    pix = orc.wavelet_transform_timefreq(w, x)
    F = orc.wavelet_transform_inverse_freq(w, pix)
    ref = np.zeros(ND + 2)
    ref[0::2] = 1.0
    assert np.all(np.abs(F - ref) <= 1e-10 + 1e-5 * np.abs(ref))
    back = orc.fftri(F.view(np.complex128), ND) / ND
    assert np.all(np.abs(back - x) <= 1e-10 + 1e-5 * np.abs(x))


@pytest.mark.parametrize("NF,NT", [(64, 32), (96, 64), (256, 128)])
def test_wdm_round_trip_random(orc, NF, NT):
    w = orc.Wdm(NF, NT, 100.0)
    ND = NF * NT
    x = np.random.default_rng(NF + NT).standard_normal(ND)
    pix = orc.wavelet_transform_timefreq(w, x)
    F = orc.wavelet_transform_inverse_freq(w, pix)
    back = orc.fftri(F.view(np.complex128), ND) / ND
    assert np.max(np.abs(back - x)) < 1e-11 * np.max(np.abs(x))
    # the forward transform from the spectrum equals the one from the series
    pix2 = orc.wavelet_transform_freq(w, orc.fftr(x).view(np.float64))
    assert np.array_equal(pix, pix2)


@pytest.mark.parametrize("jmin,L,kb", [(11, 2, 3), (12, 2, 20), (20, 3, 5), (21, 1, 9), (30, 3, 30), (31, 3, 41),
                                       (2, 7, 5), (3, 15, 11), (2, 31, 40)])
def test_by_layers_heterodyne_identity(orc, jmin, L, kb):
    NF, NT, dt = 64, 32, 100.0
    w = orc.Wdm(NF, NT, dt)
    ND, T = NF * NT, NT * dt
    k = (jmin - 1) * NT // 2 + kb
    f = k / T
    ts = np.arange(ND) * dt / NF
    full = orc.wavelet_transform_timefreq(w, 1.3 * np.cos(2 * np.pi * f * ts + 0.4))
    tds = np.arange(NT * (L + 1)) * dt / (L + 1)
    xh = 1.3 * np.cos(2 * np.pi * (f - (jmin - 1) * w.df) * tds + 0.4)
    bl = orc.wavelet_transform_by_layers(w, jmin, L, orc.wdm_filter(L + 1, NT), xh)
    a = np.concatenate([full[NT * (jmin + o):NT * (jmin + o + 1)] for o in range(L)])
    b = np.concatenate([bl[np.arange(NT) * L + o] for o in range(L)])
    assert np.max(np.abs(a - np.sqrt(L + 1) * b)) < 1e-10 * np.max(np.abs(a))


def test_spline_vs_scipy(orc):
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.uniform(0.5, 1.5, 40))
    y = np.sin(x) * 3 + x
    c = orc.spline_coefficients(x, y)
    cs = CubicSpline(x, y, bc_type="natural")
    xv = np.concatenate([rng.uniform(x[0], x[-1], 500), x])
    got = orc.spline_eval(x, c, xv)
    assert np.max(np.abs(got - cs(xv))) < 1e-12 * np.max(np.abs(y))
    # knots: the value itself, through the last knot (index N-1)
    assert np.array_equal(orc.spline_eval(x, c, x), y)
    for v in list(x) + list(xv[:50]):
        assert orc.binary_search(x, v) == np.searchsorted(x, v, side="right") - 1


def test_extract_amplitude_and_phase_reconstructs(orc):
    rng = np.random.default_rng(5)
    n = 300
    t = np.linspace(0, 1, n)
    amp = np.cos(2 * np.pi * 1.3 * t) + 0.05            # changes sign: flips
    ph = 3.0 * t + 0.2 * np.sin(7 * t)
    phiR = 1e6 + 2e3 * t + rng.uniform(0, 1e-3, n)
    M = amp * np.cos(ph + phiR)
    Mf = -amp * np.sin(ph + phiR)
    As, D = orc.extract_amplitude_and_phase(M, Mf, phiR)
    assert np.allclose(As * np.cos(D + phiR), M, atol=1e-9)
    assert np.allclose(-As * np.sin(D + phiR), Mf, atol=1e-9)
    U = orc.unwrap_phase(D)
    assert np.all(np.abs(np.diff(U)) <= np.pi + 1e-12)
    assert np.allclose(np.cos(U - D), 1.0)


def test_iorbit_knots(orc):
    t, pos, coef = orc.iorbit(T1)
    N = len(t)
    assert N == int(200 * T1 / 3.15581497632e7) == 199
    buf = T1 / (N - 1)
    assert np.allclose(t, np.arange(N) * (T1 + 2 * buf) / (N - 1) - buf, rtol=0, atol=1e-6)
    o = orc.analytic_orbit()
    x, y, z = (np.zeros(4) for _ in range(3))
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    for n in (0, 57, N - 1):
        orc.lib().orc_analytic_orbits(ctypes.byref(o), t[n], x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                                      z.ctypes.data_as(dp))
        for i in range(3):
            assert pos[3 * i, n] == x[i + 1] and pos[3 * i + 1, n] == y[i + 1] and pos[3 * i + 2, n] == z[i + 1]
    # between knots the spline stays on the rigid orbit (AU scale, 1e-6 rel)
    tm = 0.5 * (t[10] + t[11])
    orc.lib().orc_analytic_orbits(ctypes.byref(o), tm, x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                                  z.ctypes.data_as(dp))
    v = orc.spline_eval(t, coef[0], [tm])[0]
    assert abs(v - x[1]) < 1e-6 * 1.5e11


def test_wavelet_layers_closed_form(orc):
    w = orc.Wdm(4096, 512, T1 / 512)
    df = w.df
    VE = 0.00010103671
    for jf, fdotT2, want in [(100.3, 0.0, (100, 2)), (100.7, 0.0, (100, 2)), (100.995, 0.0, (100, 3)),
                             (200.2, 0.0, (200, 2))]:
        fmin = jf * df
        f0 = fmin / (1 - VE)
        p = np.array([f0 * T1, 0, 0, 0, 0, 0, 0, fdotT2])
        j0, L = orc.ucb_wavelet_layers(w, T1, p)
        j = int(np.rint(fmin / df))
        assert j0 == (j - 1 if fmin - j * df < 0 else j)
        fmax = f0 * (1 + VE)
        assert j0 + L - 1 == (j + 1 if fmax - j * df > 0 else j)
        assert (j0, L) == want


@pytest.fixture(scope="module")
def grid2(orc):
    w = orc.Wdm(4096, 512, T1 / 512)
    return w, orc.iorbit(T1)


def _params(seed, n):
    return synth.draw_params(np.random.default_rng(seed), n, T1)


def test_wavelet_waveform_vs_full_cadence(orc, grid2):
    w, orb = grid2
    NF, NT = w.NF, w.NT
    ND = NF * NT
    for p in _params(21, 2):
        lst, v = orc.ucb_waveform_wavelet(w, orb, T1, p)
        j0, L = orc.ucb_wavelet_layers(w, T1, p)
        ser = orc.ucb_wavelet_series(orb, T1, p, ND, T1 / ND, 0.0)
        full = np.stack([orc.wavelet_transform_timefreq(w, ser[c]) for c in range(3)])
        k = lst + w.kmin
        a = full[:, k]
        i = k % NT
        inner = (i >= w.imin) & (i <= w.imax)
        err = np.abs(a - np.sqrt(L + 1) * v)[:, inner].max() / np.abs(a).max()
        assert err < 1e-5, err


@pytest.mark.parametrize("layers", [7, 15])
def test_wavelet_waveform_forced_layers_vs_full_cadence(orc, grid2, layers):
    """A forced layer count (the timing harness's override of Nlayers) keeps
    the template an exact by-layers transform: sqrt(layers + 1) times its
    pixels equal the full transform of the full-cadence series on every
    listed pixel, also in the layers the reference rule would not touch
    (where both are small); the list covers layers jmin .. jmin + layers - 1."""
    w, orb = grid2
    NF, NT = w.NF, w.NT
    ND = NF * NT
    for p in _params(23, 2):
        lst, v = orc.ucb_waveform_wavelet(w, orb, T1, p, layers=layers)
        j0, L = orc.ucb_wavelet_layers(w, T1, p)
        k = lst + w.kmin
        assert set(np.unique(k // NT)) == set(range(j0, j0 + layers))
        ser = orc.ucb_wavelet_series(orb, T1, p, ND, T1 / ND, 0.0)
        full = np.stack([orc.wavelet_transform_timefreq(w, ser[c]) for c in range(3)])
        a = full[:, k]
        i = k % NT
        inner = (i >= w.imin) & (i <= w.imax)
        err = np.abs(a - np.sqrt(layers + 1) * v)[:, inner].max() / np.abs(a).max()
        assert err < 1e-5, err
        # the rule's own template, rescaled, is the same signal on its pixels
        lr, vr = orc.ucb_waveform_wavelet(w, orb, T1, p)
        pos = {int(x): n for n, x in enumerate(lst)}
        sel = np.array([pos[int(x)] for x in lr])
        d = np.abs(np.sqrt(layers + 1) * v[:, sel] - np.sqrt(L + 1) * vr)[:, inner[sel]]
        assert d.max() < 2e-5 * np.abs(a).max()


def test_wavelet_series_vs_fd_waveform(orc, grid2):
    _, orb = grid2
    ND = 4096 * 512
    rng = np.random.default_rng(3)
    for trial in range(2):
        p = _params(30 + trial, 1)[0]
        p[0] = np.floor(p[0]) + 0.3
        p[7] = 0.0
        ser = orc.ucb_wavelet_series(orb, T1, p, ND, T1 / ND, 0.0)
        F = np.fft.rfft(ser, axis=1) * np.sqrt(T1) / (2 * ND)
        BW = 256
        fd = orc.ucb_waveform(p, T1, BW)
        q = int(p[0])
        for c in range(3):
            a, b = fd[c], F[c, q - BW // 2:q + BW // 2]
            match = np.abs(np.vdot(a, b)) / np.linalg.norm(a) / np.linalg.norm(b)
            assert match > 0.99, (trial, c, match)
            assert abs(np.linalg.norm(a) / np.linalg.norm(b) - 1) < 0.01
    del rng


def test_wavelet_waveform_invariants(orc, grid2):
    w, orb = grid2
    for p in _params(40, 3):
        lst, v = orc.ucb_waveform_wavelet(w, orb, T1, p)
        q = p.copy()
        q[3] += np.log(2.0)
        l2, v2 = orc.ucb_waveform_wavelet(w, orb, T1, q)
        assert np.array_equal(lst, l2)
        assert np.max(np.abs(v2 - 2 * v)) < 1e-13 * np.max(np.abs(v))
        for k, d in ((5, np.pi / 2), (6, np.pi)):
            q = p.copy()
            q[k] += d
            l3, v3 = orc.ucb_waveform_wavelet(w, orb, T1, q)
            assert np.array_equal(lst, l3)
            assert np.max(np.abs(v3 + v)) < 1e-8 * np.max(np.abs(v))


def test_wavelet_logl_and_snr_closed_forms(orc):
    rng = np.random.default_rng(9)
    NF, NT = 16, 256
    w = orc.Wdm(NF, NT, 1000.0)
    N = w.npix
    d = rng.standard_normal((3, N))
    A = rng.standard_normal((N, 3, 3))
    C = np.einsum("nij,nkj->nik", A, A) + 3 * np.eye(3)
    invC9 = np.ascontiguousarray(np.linalg.inv(C).reshape(N, 9).T)
    iC = invC9.T.reshape(N, 3, 3)
    q = lambda r: np.einsum("an,nab,bn->", r, iC, r)  # noqa: E731
    # empty model
    got = orc.gaussian_log_likelihood_wavelet(w, d, np.zeros_like(d), np.zeros(0, np.int32), invC9)
    assert abs(got + 0.5 * q(d)) < 1e-12 * q(d)
    # a model on a list spanning interior and edge time pixels
    lst = np.arange(0, N, 7).astype(np.int32)
    m = np.zeros_like(d)
    m[:, lst] = rng.standard_normal((3, len(lst)))
    r = d.copy()
    r[:, lst] -= m[:, lst]
    i = (lst + w.kmin) % NT
    edge = lst[(i < w.imin) | (i > w.imax)]
    r[:, edge] = 0.0
    got = orc.gaussian_log_likelihood_wavelet(w, d, m, lst, invC9)
    assert abs(got + 0.5 * q(r)) < 1e-12 * q(r)
    # snr_wavelet over the list
    s = orc.snr_wavelet(m[0], m[1], m[2], invC9, lst)
    mm = np.zeros_like(m)
    mm[:, lst] = m[:, lst]
    assert abs(s - np.sqrt(q(mm))) < 1e-12 * s


def test_batch_dlogl_is_logl_difference(orc, grid2):
    w, orb = grid2
    d, invC9 = synth.wdm_random_noise(w.NF, w.NT, w.kmin, w.kmax, 4)
    P = _params(50, 3)
    P[:, 3] += np.log(1e3)      # loud enough that the template matters
    dl, sn, nw, _ = orc.wdm_eval_batch(w, orb, T1, P, d, invC9)
    L0 = orc.gaussian_log_likelihood_wavelet(w, d, np.zeros((3, 1)), np.zeros(0, np.int32), invC9)
    for e in range(3):
        lst, v = orc.ucb_waveform_wavelet(w, orb, T1, P[e])
        model = np.zeros_like(d)
        model[:, lst] = v
        L1 = orc.gaussian_log_likelihood_wavelet(w, d, model, lst, invC9)
        assert abs((L1 - L0) - dl[e]) < 1e-12 * abs(L0), (e, L1 - L0, dl[e])
        assert nw[e] == len(lst)
        assert abs(sn[e] - orc.snr_wavelet(model[0], model[1], model[2], invC9, lst)) < 1e-12 * sn[e]


def test_wdm_whitening_layer_var(orc):
    """Like the reference's variance test (rtol 1e-2 on whitened noise):
    WDM pixels of white noise have the layer variance of the PSD rule."""
    NF, NT = 256, 128
    w = orc.Wdm(NF, NT, 100.0)
    ND = NF * NT
    rng = np.random.default_rng(11)
    psd = np.full(ND // 2 + 1, float(ND))          # E|FFT|^2 of unit white noise
    lv = orc.wdm_layer_var(w, psd)
    acc = np.zeros(NF)
    reps = 8
    for _ in range(reps):
        pix = orc.wavelet_transform_timefreq(w, rng.standard_normal(ND)).reshape(NF, NT)
        acc += np.mean(pix ** 2, axis=1)
    emp = acc[1:NF] / reps
    assert abs(np.mean(emp / lv[1:NF]) - 1) < 1e-2


# ----------------------------------------------- ucb_fisher_wavelet pins --

def _edge_inner(w, lst):
    i = (lst + w.kmin) % w.NT
    return (i > w.imin) & (i < w.imax)


def _quad9(u, v, ic9, k):
    """sum over pixels k of u^T C v with the nine planes (u, v: [3, len(k)])."""
    return sum(np.sum(u[a] * v[b] * ic9[3 * a + b, k]) for a in range(3) for b in range(3))


def test_fisher_wavelet_lnA_is_edge_snr2(orc, grid2):
    """The template is linear in A = exp(params[3]), so the central
    difference of params[3] is (exp(dp) - exp(dm))/eps h, with dp, dm the
    steps actually taken (p3 +- eps/2 round to the ulp of p3 ~ 50, 7e-15,
    i.e. ~1e-8 of eps; p+ - p3 is exact): F[3][3] = that factor^2 times
    the sum of h^T C h over the template's pixels strictly inside (imin,
    imax).  The two templates differ in every rounding of their A-scaled
    intermediates, ~1e-15/eps per pixel and correlated along the splines:
    residuals of 1e-10..1e-7 relative; the bound 1e-6 still fails a wrong
    step, edge rule or channel pair by orders of magnitude."""
    w, orb = grid2
    cfg = synth.config2(n=8, seed=12)
    eps = 1e-6
    for p in cfg["params"]:
        F = orc.ucb_fisher_wavelet(w, orb, T1, p, cfg["invC9"])
        lst, v = orc.ucb_waveform_wavelet(w, orb, T1, p)
        m = _edge_inner(w, lst)
        dp = (p[3] + eps / 2.0) - p[3]
        dm = (p[3] - eps / 2.0) - p[3]
        want = _quad9(v[:, m], v[:, m], cfg["invC9"], lst[m]) * ((np.exp(dp) - np.exp(dm)) / eps) ** 2
        assert abs(F[3, 3] - want) < 1e-6 * want, (F[3, 3], want)
        # the pixels outside the edges do contribute to snr: the rule matters
        assert _quad9(v, v, cfg["invC9"], lst) > 1.01 * want
        assert np.array_equal(F, F.T)


def test_fisher_wavelet_vs_five_point(orc, grid2):
    """Against a Fisher built independently from five-point central
    differences with larger steps on the same pixels: agreement set by the
    oracle's 1e-6 central-difference noise (~1e-5 on the f0 terms)."""
    w, orb = grid2
    cfg = synth.config2(n=2, seed=13)
    steps = np.array([1e-3, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4, 1e-3])
    for p in cfg["params"]:
        p = p.copy()
        p[1] = np.clip(p[1], -0.9, 0.9)
        p[4] = np.clip(p[4], -0.9, 0.9)
        F = orc.ucb_fisher_wavelet(w, orb, T1, p, cfg["invC9"])
        lst0, _ = orc.ucb_waveform_wavelet(w, orb, T1, p)
        m0 = _edge_inner(w, lst0)
        keep = lst0[m0]
        pos = {int(k): n for n, k in enumerate(keep)}
        D = np.zeros((8, 3, len(keep)))
        for i in range(8):
            acc = np.zeros((3, len(keep)))
            for c, s in ((8, 1), (-8, -1), (-1, 2), (1, -2)):
                q = p.copy()
                q[i] += s * steps[i]
                l, v = orc.ucb_waveform_wavelet(w, orb, T1, q)
                assert np.array_equal(l, lst0)
                acc += c * v[:, m0]
            D[i] = acc / (12 * steps[i])
        G = np.array([[_quad9(D[i], D[j], cfg["invC9"], keep) for j in range(8)] for i in range(8)])
        norm = np.sqrt(np.outer(np.diag(G), np.diag(G)))
        err = np.max(np.abs(F - G) / norm)
        print("five-point vs oracle FD Fisher, normalised", err)
        assert err < 1e-3, err
        del pos


def test_wdm_layer_var_white_closed_form(orc):
    """The Meyer window's normalisation makes the WDM basis orthonormal:
    a white DFT variance of ND (unit-variance samples) gives interior layer
    variance 1 exactly, and each edge layer the sum of its one-sided
    window weights (l = 0 quartered) over NT^2; linear in the PSD."""
    for NF, NT in ((64, 32), (256, 128), (1536, 64)):
        w = orc.Wdm(NF, NT, 100.0)
        ND = NF * NT
        lv = orc.wdm_layer_var(w, np.full(ND // 2 + 1, float(ND)))
        assert np.max(np.abs(lv[1:NF] - 1.0)) < 1e-12
        phif = orc.wdm_filter(NF, NT)
        edge = (0.25 * phif[0] ** 2 + np.sum(phif[1:NT // 2] ** 2)) * ND / NT ** 2
        assert abs(lv[0] - edge) < 1e-12 * edge and abs(lv[NF] - edge) < 1e-12 * edge
        rng = np.random.default_rng(NF)
        a, b = rng.uniform(1, 2, ND // 2 + 1), rng.uniform(1, 2, ND // 2 + 1)
        lab = orc.wdm_layer_var(w, 2 * a + 3 * b)
        assert np.max(np.abs(lab - (2 * orc.wdm_layer_var(w, a) + 3 * orc.wdm_layer_var(w, b))) / lab) < 1e-13
