"""GPU parity of the wavelet-domain path against the oracle.

* transforms (forward, from spectrum, inverse; device API) at grids from
  64 x 32 up to the 4096 x 512 grid of a year at 15 s, random series and the
  reference's impulse known answer: normwise 1e-12;
* interpolated orbit knots: bitwise;
* ucb_waveform_wavelet drop-in: identical pixel lists; values normwise 1e-12
  per template against the oracle with correctly rounded sin/cos.  Against
  the glibc oracle the bound per template is 1e-12 plus that template's own
  glibc-vs-correctly-rounded difference: glibc misrounds ~0.1% of sin/cos
  arguments and the TDI sums cancel by (f/f*)^2, so the reference's output
  itself moves by up to ~3e-11 between libms (1.5% of templates > 1e-12);
* batched templates: lists/values as above, snr and dlogL relative 1e-10;
* snr_wavelet: relative 1e-12; gaussian_log_likelhood_wavelet: the
  reference's formula cancels (XX and 2XY of ~1e14 for the XYZ spectrum near
  DC), so its own rounding is ~1e-16 * sum|terms|; the bound is 1e-12 of
  that scale (and 1e-10 relative on a well-conditioned covariance).
"""
import numpy as np
import pytest

from glass_amd import synth

pytestmark = pytest.mark.gpu

T1 = synth.YEAR_T


def normwise(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def W():
    from glass_amd import wdm
    return wdm


def grids(orc, W, NF, NT, dt=100.0):
    return orc.Wdm(NF, NT, dt), W.Wdm.grid(NF, NT, dt)


@pytest.mark.parametrize("NF,NT", [(64, 32), (96, 64), (1536, 64), (256, 512), (4096, 512)])
def test_transform_forward_vs_oracle(orc, W, NF, NT):
    ow, gw = grids(orc, W, NF, NT)
    x = np.random.default_rng(NF * 7 + NT).standard_normal(NF * NT)
    ref = orc.wavelet_transform_timefreq(ow, x)
    got = W.wavelet_transform_timefreq(gw, x)
    err = normwise(got, ref)
    print(NF, NT, "forward normwise", err)
    assert err < 1e-12


@pytest.mark.parametrize("NF,NT", [(64, 32), (1536, 64), (4096, 512)])
def test_transform_freq_and_inverse_vs_oracle(orc, W, NF, NT):
    ow, gw = grids(orc, W, NF, NT)
    x = np.random.default_rng(NF + NT).standard_normal(NF * NT)
    F = orc.fftr(x).view(np.float64)
    assert normwise(W.wavelet_transform_freq(gw, F), orc.wavelet_transform_freq(ow, F)) < 1e-12
    pix = orc.wavelet_transform_timefreq(ow, x)
    got = W.wavelet_transform_inverse_freq(gw, pix)
    ref = orc.wavelet_transform_inverse_freq(ow, pix)
    assert normwise(got, ref) < 1e-12


@pytest.mark.parametrize("NF,NT", [(1536, 32), (64, 64)])
def test_transform_impulse_known_answer(W, NF, NT):
    """fft_wdm_unit_tests.c:547-590 through the GPU transforms."""
    gw = W.Wdm.grid(NF, NT, 7680.0)
    x = np.zeros(NF * NT)
    x[0] = 1.0
    F = W.wavelet_transform_inverse_freq(gw, W.wavelet_transform_timefreq(gw, x))
    ref = np.zeros(NF * NT + 2)
    ref[0::2] = 1.0
    assert np.all(np.abs(F - ref) <= 1e-10 + 1e-5 * np.abs(ref))


@pytest.mark.parametrize("NT,L,jmin", [(512, 1, 7), (512, 2, 300), (512, 3, 4000), (256, 2, 0), (256, 3, 4094),
                                       (256, 1, 4096)])
def test_by_layers_vs_oracle(orc, W, NT, L, jmin):
    """wavelet_transform_timefreq_by_layers drop-in, including the m = 0 and
    m = NF edge layers (normwise 1e-12)."""
    NF = 4096
    ow, gw = grids(orc, W, NF, NT, T1 / NT)
    win = orc.wdm_filter(L + 1, NT, 0.0, True)
    x = np.random.default_rng(NT + 10 * L + jmin).standard_normal(NT * (L + 1))
    ref = orc.wavelet_transform_by_layers(ow, jmin, L, win, x)
    got = W.wavelet_transform_timefreq_by_layers(gw, jmin, L, win, x)
    err = normwise(got, ref)
    print(NT, L, jmin, "by-layers normwise", err)
    assert err < 1e-12
    assert np.all(got[NT * L:] == 0.0)


def test_wdm_transform_device_api(W):
    import torch
    gw = W.Wdm.grid(4096, 512, T1 / 512)
    x = np.random.default_rng(3).standard_normal(4096 * 512)
    d = torch.tensor(x, device="cuda")
    out = W.wdm_transform(gw, d)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), W.wavelet_transform_timefreq(gw, x))


def test_transform_bad_grid(W):
    from glass_amd import GlassHipError
    with pytest.raises(GlassHipError):
        W.wavelet_transform_timefreq(W.Wdm.grid(64, 48, 100.0), np.zeros(64 * 48))   # NT not 2^a
    with pytest.raises(GlassHipError):
        W.wavelet_transform_timefreq(W.Wdm.grid(80, 32, 100.0), np.zeros(80 * 32))   # NF 5*2^4


def test_interpolated_orbit_bitwise(orc, gl, W):
    t, x, y, z = W.interpolated_analytic_orbit(gl.analytic_orbit(), T1)
    ot, pos, _ = orc.iorbit(T1)
    assert np.array_equal(t, ot)
    for i in range(3):
        assert np.array_equal(x[i], pos[3 * i]) and np.array_equal(y[i], pos[3 * i + 1])
        assert np.array_equal(z[i], pos[3 * i + 2])


@pytest.fixture(scope="module")
def g2(orc, gl, W):
    NF, NT = 4096, 512
    ow, gw = grids(orc, W, NF, NT, T1 / NT)
    return ow, gw, orc.iorbit(T1), W.interpolated_analytic_orbit(gl.analytic_orbit(), T1)


def oracle_templates(orc, ow, oorb, p):
    """(list, values with glibc sin/cos, values with correctly rounded sin/cos)"""
    lst, v = orc.ucb_waveform_wavelet(ow, oorb, T1, p)
    with orc.cr_trig():
        l2, vcr = orc.ucb_waveform_wavelet(ow, oorb, T1, p)
    assert np.array_equal(lst, l2)
    return lst, v, vcr


def test_ucb_waveform_wavelet_dropin(orc, W, g2):
    ow, gw, oorb, gorb = g2
    P = synth.draw_params(np.random.default_rng(17), 48, T1)
    X = np.zeros(gw.npix)
    Y, Z = np.zeros_like(X), np.zeros_like(X)
    worst_cr, worst_env = 0.0, 0.0
    for e, p in enumerate(P):
        lst, v, vcr = oracle_templates(orc, ow, oorb, p)
        glst, X, Y, Z = W.ucb_waveform_wavelet(gw, gorb, T1, p, X=X, Y=Y, Z=Z)
        assert np.array_equal(glst, lst), e
        got = np.stack([X[glst], Y[glst], Z[glst]])
        e_cr = normwise(got, vcr)
        env = normwise(v, vcr)
        e_gl = normwise(got, v)
        assert e_gl <= 1e-12 + env, (e, e_gl, env)
        worst_cr = max(worst_cr, e_cr)
        worst_env = max(worst_env, env)
        X[glst] = Y[glst] = Z[glst] = 0.0
    print("wavelet waveform worst normwise vs CR oracle", worst_cr, "; oracle glibc-vs-CR envelope", worst_env)
    assert worst_cr < 1e-12


def test_wavelet_batch_vs_oracle(orc, W, g2):
    import torch
    ow, gw, oorb, gorb = g2
    cfg = synth.config2(n=256, seed=5)
    P = cfg["params"]
    B = W.WdmBatch(gw, gorb, T1, cfg["data"], cfg["invC6"], max_templates=256)
    dl, sn = B.eval(torch.tensor(P, device="cuda"))
    dl, sn = dl.cpu().numpy(), sn.cpu().numpy()
    with orc.cr_trig():
        ref_dl, ref_sn, ref_nw, _ = orc.wdm_eval_batch(ow, oorb, T1, P, cfg["data"], cfg["invC9"])
    nw, lst, val = B.templates(256)
    assert np.array_equal(nw, ref_nw)
    worst = 0.0
    for e in range(0, 256, 16):
        olst, ov, ovcr = oracle_templates(orc, ow, oorb, P[e])
        assert np.array_equal(lst[e, :nw[e]], olst)
        worst = max(worst, normwise(val[e, :, :nw[e]], ovcr))
        assert normwise(val[e, :, :nw[e]], ov) <= 1e-12 + normwise(ov, ovcr)
    print("batch template worst normwise (CR oracle)", worst)
    assert worst < 1e-12
    rel_dl = np.max(np.abs(dl - ref_dl) / np.maximum(np.abs(ref_dl), 1.0))
    rel_sn = np.max(np.abs(sn - ref_sn) / ref_sn)
    print("batch dlogL rel", rel_dl, "snr rel", rel_sn)
    assert rel_dl < 1e-10 and rel_sn < 1e-10
    B.close()


@pytest.mark.parametrize("layers", [7, 15, 31, 63])
def test_wavelet_batch_forced_layers_vs_oracle(orc, W, g2, layers):
    """A plan that forces every template to `layers` frequency layers (the
    timing knob for templates touching 8-64 layers; long-series layer kernel
    at NT 512): pixel lists, values, dlogL and snr against the oracle with
    the same override; the wavelet Fisher refuses such a plan."""
    import torch
    ow, gw, oorb, gorb = g2
    n = 16
    cfg = synth.config2(n=n, seed=8)
    P = cfg["params"]
    B = W.WdmBatch(gw, gorb, T1, cfg["data"], cfg["invC6"], max_templates=n, layers=layers)
    assert B.vcap == gw.NT * layers
    dl, sn = B.eval(torch.tensor(P, device="cuda"))
    dl, sn = dl.cpu().numpy(), sn.cpu().numpy()
    nw, lst, val = B.templates(n)
    with orc.cr_trig():
        ref_dl, ref_sn, ref_nw, _ = orc.wdm_eval_batch(ow, oorb, T1, P, cfg["data"], cfg["invC9"], layers=layers)
    assert np.array_equal(nw, ref_nw)
    worst = 0.0
    for e in range(0, n, 5):
        with orc.cr_trig():
            olst, ov = orc.ucb_waveform_wavelet(ow, oorb, T1, P[e], layers=layers)
        assert np.array_equal(lst[e, :nw[e]], olst)
        worst = max(worst, normwise(val[e, :, :nw[e]], ov))
    rel_dl = np.max(np.abs(dl - ref_dl) / np.maximum(np.abs(ref_dl), 1.0))
    rel_sn = np.max(np.abs(sn - ref_sn) / ref_sn)
    print("forced", layers, "layers: template worst normwise (CR oracle)", worst, "dlogL rel", rel_dl, "snr rel", rel_sn)
    assert worst < 1e-12
    assert rel_dl < 1e-10 and rel_sn < 1e-10
    with pytest.raises(Exception, match="forces"):
        B.fisher(torch.tensor(P[:1], device="cuda"))
    B.close()


def test_wavelet_batch_forced_layers_refused(W, g2):
    _, gw, _, gorb = g2
    cfg = synth.config2(n=2, seed=9)
    for bad in (5, 8, 127):
        with pytest.raises(Exception, match="forced layer count"):
            W.WdmBatch(gw, gorb, T1, cfg["data"], cfg["invC6"], max_templates=2, layers=bad)


def test_snr_and_logl_wavelet_dropins(orc, W, g2):
    ow, gw, oorb, gorb = g2
    cfg = synth.config2(n=4, seed=6)
    d, ic9 = cfg["data"], cfg["invC9"]
    lst, v = orc.ucb_waveform_wavelet(ow, oorb, T1, cfg["params"][0])
    m = np.zeros_like(d)
    m[:, lst] = v
    s_ref = orc.snr_wavelet(m[0], m[1], m[2], ic9, lst)
    s = W.snr_wavelet(m[0], m[1], m[2], ic9, lst)
    assert abs(s - s_ref) < 1e-12 * s_ref
    L_ref = orc.gaussian_log_likelihood_wavelet(ow, d, m, lst, ic9)
    L = W.gaussian_log_likelihood_wavelet(gw, d, m, lst, ic9)
    # rounding scale of the reference's XX + 2 XY over all pixels
    scale = sum(np.sum(np.abs(d[a] * d[b] * ic9[3 * a + b])) for a in range(3) for b in range(3))
    print("logL wavelet", L, L_ref, "diff / scale", abs(L - L_ref) / scale)
    assert abs(L - L_ref) < 1e-12 * scale


def test_logl_wavelet_well_conditioned(orc, W, g2):
    ow, gw, oorb, gorb = g2
    d, ic9 = synth.wdm_random_noise(ow.NF, ow.NT, ow.kmin, ow.kmax, 8)
    p = synth.draw_params(np.random.default_rng(8), 1, T1)[0]
    lst, v = orc.ucb_waveform_wavelet(ow, oorb, T1, p)
    m = np.zeros_like(d)
    m[:, lst] = 1e3 * v
    L_ref = orc.gaussian_log_likelihood_wavelet(ow, d, m, lst, ic9)
    L = W.gaussian_log_likelihood_wavelet(gw, d, m, lst, ic9)
    assert abs(L - L_ref) < 1e-10 * abs(L_ref)


@pytest.fixture(scope="module")
def gref(orc, gl, W):
    """The reference's default wavelet grid at Tobs = 1 yr (initialize_wavelet,
    utils/src/glass_wavelet.c:210-216): NT = ceil(T / 7680 s) = 4096 pixels
    of 7680 s, NF = 7680 s / 5 s = 1536 layers of 1/15360 Hz."""
    NF, NT = 1536, 4096
    ow, gw = grids(orc, W, NF, NT, 7680.0)
    assert abs(gw.df - 6.51041666666667e-5) < 1e-15 and NT * 7680.0 == T1
    return ow, gw, orc.iorbit(T1), W.interpolated_analytic_orbit(gl.analytic_orbit(), T1)


def test_ucb_waveform_wavelet_reference_grid(orc, W, gref):
    """Templates on the reference's own grid (NT 4096: the long-series layer
    kernel): identical pixel lists, values against the correctly rounded
    oracle to 1e-12 normwise and against glibc within its envelope."""
    ow, gw, oorb, gorb = gref
    P = synth.draw_params(np.random.default_rng(23), 24, T1)
    # a layer boundary inside [f0 (1 - v/c), f0 (1 + v/c)]: three layers
    # (an NT x 4 = 16384-sample series); the others take two (12288)
    for i in range(4):
        P[i, 0] = int(P[i, 0] / T1 / gw.df) * gw.df * T1
        P[i, 7] = 0.0
    X = np.zeros(gw.npix)
    Y, Z = np.zeros_like(X), np.zeros_like(X)
    worst, layers = 0.0, []
    for e, p in enumerate(P):
        lst, v, vcr = oracle_templates(orc, ow, oorb, p)
        layers.append(orc.ucb_wavelet_layers(ow, T1, p)[1])
        glst, X, Y, Z = W.ucb_waveform_wavelet(gw, gorb, T1, p, X=X, Y=Y, Z=Z)
        assert np.array_equal(glst, lst), e
        got = np.stack([X[glst], Y[glst], Z[glst]])
        assert normwise(got, v) <= 1e-12 + normwise(v, vcr), e
        worst = max(worst, normwise(got, vcr))
        X[glst] = Y[glst] = Z[glst] = 0.0
    print("reference grid: layers per template", sorted(set(layers)), "worst normwise vs CR oracle", worst)
    assert set(layers) == {2, 3}
    assert worst < 1e-12


@pytest.mark.parametrize("NF,NT", [(2048, 1024), (1024, 2048)])
def test_ucb_waveform_wavelet_long_series_grids(orc, gl, W, NF, NT):
    """The long-series layer kernel at NT = 1024 and 2048 (two and three
    layers): lists identical, values 1e-12 normwise against the CR oracle."""
    ow, gw = grids(orc, W, NF, NT, T1 / NT)
    oorb, gorb = orc.iorbit(T1), W.interpolated_analytic_orbit(gl.analytic_orbit(), T1)
    P = synth.draw_params(np.random.default_rng(NT + 1), 8, T1)
    for i in range(3):
        P[i, 0] = int(P[i, 0] / T1 / gw.df) * gw.df * T1
        P[i, 7] = 0.0
    worst, layers = 0.0, set()
    for e, p in enumerate(P):
        lst, v, vcr = oracle_templates(orc, ow, oorb, p)
        layers.add(orc.ucb_wavelet_layers(ow, T1, p)[1])
        glst, X, Y, Z = W.ucb_waveform_wavelet(gw, gorb, T1, p)
        assert np.array_equal(glst, lst), e
        worst = max(worst, normwise(np.stack([X[glst], Y[glst], Z[glst]]), vcr))
    print(NF, NT, "layers", sorted(layers), "worst normwise vs CR oracle", worst)
    assert worst < 1e-12


def test_wavelet_batch_reference_grid(orc, W, gref):
    """The batched templates with fused wavelet_nwip on the reference's grid
    (6.3 M pixels, well-conditioned random pixel covariance): lists, values,
    dlogL and snr against the oracle."""
    import torch
    ow, gw, oorb, gorb = gref
    d, ic9 = synth.wdm_random_noise(ow.NF, ow.NT, ow.kmin, ow.kmax, 29)
    ic6 = ic9[[0, 4, 8, 1, 2, 5]]
    P = synth.draw_params(np.random.default_rng(31), 40, T1)
    B = W.WdmBatch(gw, gorb, T1, d, ic6, max_templates=40)
    dl, sn = B.eval(torch.tensor(P, device="cuda"))
    dl, sn = dl.cpu().numpy(), sn.cpu().numpy()
    with orc.cr_trig():
        ref_dl, ref_sn, ref_nw, _ = orc.wdm_eval_batch(ow, oorb, T1, P, d, ic9)
    nw, lst, val = B.templates(40)
    assert np.array_equal(nw, ref_nw)
    for e in range(0, 40, 8):
        olst, ov, ovcr = oracle_templates(orc, ow, oorb, P[e])
        assert np.array_equal(lst[e, :nw[e]], olst)
        assert normwise(val[e, :, :nw[e]], ovcr) < 1e-12
    rel_dl = np.max(np.abs(dl - ref_dl) / np.maximum(np.abs(ref_dl), 1.0))
    rel_sn = np.max(np.abs(sn - ref_sn) / ref_sn)
    print("reference grid batch: dlogL rel", rel_dl, "snr rel", rel_sn)
    assert rel_dl < 1e-10 and rel_sn < 1e-10
    B.close()


def test_wavelet_bad_inputs(gl, W, g2):
    from glass_amd import GlassHipError
    ow, gw, oorb, gorb = g2
    for NT in (8192, 768):
        bad = W.Wdm.grid(256, NT, T1 / NT)
        with pytest.raises(GlassHipError):
            W.ucb_waveform_wavelet(bad, gorb, T1, synth.draw_params(np.random.default_rng(1), 1, T1)[0])
    with pytest.raises(GlassHipError):
        X = np.zeros(8)
        W.snr_wavelet(X, X, X, np.zeros((9, 8)), [0, 9])


# ------------------------------------------------------ wavelet Fisher --

def boundary_params(p, df, j0):
    """params with f0 (1 - v_earth) = j0 df exactly and fdot = 0: the plus
    and minus f0 steps of ucb_fisher_wavelet then take different layer
    sets (jmin j0 vs j0 - 1), so only the plus template's pixels carry
    derivatives (glass_ucb_waveform.c:592-603)."""
    VE = 0.00010103671    # VEARTH of ucb_wavelet_layers (ucb/src/glass_ucb_waveform.c:1003-1037)
    q = p.copy()
    q[0] = j0 * df / (1.0 - VE) * T1
    q[7] = 0.0
    return q


def fisher_err(F, R):
    """max_ij |F - R| / sqrt(R_ii R_jj)."""
    d = np.sqrt(np.abs(np.diag(R)))
    return float(np.max(np.abs(F - R) / np.outer(d, d)))


def test_fisher_wavelet_dropin_vs_oracle(orc, W, g2):
    """Drop-in vs the oracle with correctly rounded sin/cos: the central
    differences divide the ~1e-14 template differences by eps = 1e-6, so
    the matrices agree to ~1e-8 of their scale; bound 1e-6, as for the
    Fourier-domain FD Fisher."""
    ow, gw, oorb, gorb = g2
    cfg = synth.config2(n=20, seed=14)
    P = list(cfg["params"])
    P += [boundary_params(cfg["params"][0], ow.df, 611), boundary_params(cfg["params"][1], ow.df, 300)]
    for q in P[-2:]:
        a, b = q.copy(), q.copy()
        a[0] += 5e-7
        b[0] -= 5e-7
        assert orc.ucb_wavelet_layers(ow, T1, a) != orc.ucb_wavelet_layers(ow, T1, b)
    worst = 0.0
    for e, p in enumerate(P):
        F = W.ucb_fisher_wavelet(gw, gorb, T1, p, cfg["invC9"])
        with orc.cr_trig():
            R = orc.ucb_fisher_wavelet(ow, oorb, T1, p, cfg["invC9"])
        err = fisher_err(F, R)
        worst = max(worst, err)
        assert np.array_equal(F, F.T)
        assert err < 1e-6, (e, err)
    print("wavelet Fisher drop-in worst normalised error vs oracle", worst)


def test_fisher_wavelet_boundary_quirk(orc, g2):
    """At a layer boundary the plus-list rule is what the result follows:
    summing derivatives over the union of both lists would change F."""
    ow, gw, oorb, gorb = g2
    cfg = synth.config2(n=2, seed=15)
    p = boundary_params(cfg["params"][0], ow.df, 611)
    e = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[0] += e / 2
    pm[0] -= e / 2
    lp, _ = orc.ucb_waveform_wavelet(ow, oorb, T1, pp)
    lm, _ = orc.ucb_waveform_wavelet(ow, oorb, T1, pm)
    assert len(np.setdiff1d(lm, lp)) > 0      # the minus template has pixels the plus one lacks


def test_wdm_batch_fisher_matches_dropin(orc, W, g2):
    import torch
    ow, gw, oorb, gorb = g2
    cfg = synth.config2(n=24, seed=16)
    P = cfg["params"].copy()
    P[5] = boundary_params(P[5], ow.df, 450)
    B = W.WdmBatch(gw, gorb, T1, cfg["data"], cfg["invC6"], max_templates=16 * 24)
    F = B.fisher(torch.tensor(P, device="cuda")).cpu().numpy()
    # same templates and kernel; the batch reads 6 symmetric invC planes
    # where the drop-in reads all 9 (numpy's inverse is symmetric only to
    # an ulp)
    for e in range(0, 24, 3):
        R = W.ucb_fisher_wavelet(gw, gorb, T1, P[e], cfg["invC9"])
        assert fisher_err(F[e], R) < 1e-13, (e, fisher_err(F[e], R))
    with orc.cr_trig():
        R = orc.ucb_fisher_wavelet(ow, oorb, T1, P[5], cfg["invC9"])
    assert fisher_err(F[5], R) < 1e-6
    from glass_amd import GlassHipError
    with pytest.raises(GlassHipError):
        B.fisher(torch.tensor(np.tile(P, (2, 1)), device="cuda"))   # 48 sources > 384 / 16
    B.close()


@pytest.mark.parametrize("NF,NT", [(4096, 512), (1536, 4096), (64, 32)])
def test_stationary_psd_to_wdm_vs_oracle(orc, W, NF, NT):
    """stationary_dft_psd_to_wdm_psd on the device: layer variances against
    the oracle's running sum (all terms >= 0: relative 1e-13), every pixel
    of the NF x NT map from its layer (row 0: even n layer 0, odd n NF)."""
    ow, gw = grids(orc, W, NF, NT)
    ND = NF * NT
    rng = np.random.default_rng(NF + NT)
    k = np.arange(ND // 2 + 1)
    psd = (1.0 + (k + 1.0) ** -1.5 * 1e3) * rng.uniform(0.5, 1.5, ND // 2 + 1)
    ref = orc.wdm_layer_var(ow, psd)
    lv = W.stationary_dft_psd_to_wdm_layer_var(gw, psd)
    assert np.max(np.abs(lv - ref) / ref) < 1e-13
    pix = W.stationary_dft_psd_to_wdm_psd(gw, psd).reshape(NF, NT)
    assert np.array_equal(pix[1:], np.repeat(lv[1:NF, None], NT, axis=1))
    assert np.all(pix[0, 0::2] == lv[0]) and np.all(pix[0, 1::2] == lv[NF])
