"""GPU Fisher / gradient checks.

* glass_hip_ucb_fisher (reference-compatible forward differences) against
  the oracle's restatement of ucb_fisher.  The reference differences two
  templates 1e-6 apart, which turns their ~1e-15 rounding into ~1e-9
  relative noise on the matrix; parity is asserted at 1e-6.
* glass_hip_fd_source_derivs (exact forward-mode derivatives) against
  central finite differences of the oracle's L(theta) = (R|h) - (h|h)/2 and
  of its templates, with q and BW held fixed, and (mode 1) against the
  reference's own finite-difference Fisher.
* the torch.autograd binding.
"""
import numpy as np
import pytest

from glass_amd import synth

pytestmark = pytest.mark.gpu

T4 = synth.T_4YR


def bform(u, v, invC):
    """2 sum_ab Re(u_a conj v_b) invC_ab summed over bins; u,v [3,W]."""
    s = 0.0
    for a in range(3):
        for b in range(3):
            s += 2.0 * np.sum((u[a] * np.conj(v[b])).real * invC[a, b])
    return s


def make_sources(orc, n, seed, BW=256, W=512, T=T4):
    rng = np.random.default_rng(seed)
    P = synth.draw_params(rng, n, T)
    # keep f0*T away from an integer so the carrier bin q survives the steps
    P[:, 0] = np.floor(P[:, 0]) + rng.uniform(0.25, 0.75, n)
    P[:, 3] = np.log(10.0 ** rng.uniform(-22.0, -21.0, n))  # loud enough to matter
    srcs = []
    for p in P:
        q = int(p[0] / T * T)
        qmin = q - W // 2
        C = synth.xyz_noise_cov(T, qmin / T, W)
        invC = orc.invert_noise(C)[0]
        R = synth.noise_draw(rng, C)
        imin = (q - BW // 2) - qmin
        srcs.append(dict(p=p, qmin=qmin, invC=invC, R=R, imin=imin))
    return srcs


def oracle_h(orc, p, T, BW, W, imin):
    h = orc.ucb_waveform(p, T, BW)[:3]
    win = np.zeros((3, W), complex)
    synth.scatter(win, h, imin)
    return win


def oracle_L(orc, p, s, T, BW, W):
    h = oracle_h(orc, p, T, BW, W, s["imin"])
    return bform(s["R"], h, s["invC"]) - 0.5 * bform(h, h, s["invC"])


def run_derivs(gl, srcs, T, BW, W, mode=0, with_resid=True):
    import torch
    from glass_amd.batch import pack_invC
    from glass_amd.derivs import source_derivs
    dev = torch.device("cuda")
    params = torch.tensor(np.stack([s["p"] for s in srcs]), device=dev)
    imin = torch.tensor([s["imin"] for s in srcs], dtype=torch.int32, device=dev)
    invC = torch.tensor(np.stack([pack_invC(s["invC"]) for s in srcs]), device=dev)
    R = torch.tensor(np.stack([s["R"] for s in srcs]), device=dev) if with_resid else None
    L, g, F = source_derivs(gl.analytic_orbit(), T, params, BW, imin, invC, R, mode=mode)
    return L.cpu().numpy(), g.cpu().numpy(), F.cpu().numpy()


STEPS = np.array([1e-3, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5, 1e-4])


def stencil(f, x, k, h=None):
    """Five-point central derivative along parameter k (error O(h^4))."""
    h = STEPS[k] if h is None else h
    e = np.zeros(8)
    e[k] = h
    return (-f(x + 2 * e) + 8 * f(x + e) - 8 * f(x - e) + f(x - 2 * e)) / (12 * h)


def test_value_and_gradient_vs_central_fd(gl, orc):
    BW, W = 256, 512
    srcs = make_sources(orc, 4, 11, BW, W)
    L, g, F = run_derivs(gl, srcs, T4, BW, W)
    for e, s in enumerate(srcs):
        Lref = oracle_L(orc, s["p"], s, T4, BW, W)
        assert abs(L[e] - Lref) <= 1e-10 * max(abs(Lref), 1.0), (e, L[e], Lref)
        # Finite differences of the waveform are limited by its own FP64
        # phase rounding: arg2 ~ 1e7 rad carries ~1e-10 rad of noise that is
        # not smooth in theta, so a step moving the phase by ~1e-3 rad leaves
        # ~1e-6 of noise in the Fisher metric.  Steps STEPS (five-point
        # stencil, truncation O(h^4)) sit near that floor.
        sig = np.sqrt(np.diag(F[e]))
        gfd = np.array([stencil(lambda x: oracle_L(orc, x, s, T4, BW, W), s["p"], k) for k in range(8)])
        # compare in the Fisher metric: dimensionless error per direction
        err = np.abs(g[e] - gfd) / sig
        print(e, "grad err / sigma^-1:", err)
        assert err.max() < 2e-5, (e, g[e], gfd)


def test_fisher_vs_central_fd_templates(gl, orc):
    BW, W = 256, 512
    srcs = make_sources(orc, 3, 12, BW, W)
    _, _, F = run_derivs(gl, srcs, T4, BW, W, with_resid=False)
    for e, s in enumerate(srcs):
        dh = [stencil(lambda x: oracle_h(orc, x, T4, BW, W, s["imin"]), s["p"], k) for k in range(8)]
        Ffd = np.array([[bform(dh[a], dh[b], s["invC"]) for b in range(8)] for a in range(8)])
        d = np.sqrt(np.outer(np.diag(Ffd), np.diag(Ffd)))
        err = np.max(np.abs(F[e] - Ffd) / d)
        print(e, "fisher err (normalised):", err)
        # same finite-difference floor as the gradient test
        assert err < 2e-5
        assert np.allclose(F[e], F[e].T, rtol=0, atol=1e-12 * np.abs(F[e]).max())
        assert np.linalg.eigvalsh(F[e]).min() > -1e-9 * np.abs(F[e]).max()


def test_mode1_matches_reference_fd_fisher(gl, orc):
    """Mode 1 uses the reference's invC alignment; its exact Fisher must
    agree with the reference's forward-difference one up to the forward
    difference's truncation (~1e-5 relative)."""
    W = 512
    T = synth.YEAR_T
    rng = np.random.default_rng(13)
    P = synth.draw_params(rng, 3, T)
    P[:, 0] = np.floor(P[:, 0]) + 0.5
    for p in P:
        q = int(p[0] / T * T)
        qmin = q - W // 2
        C = synth.xyz_noise_cov(T, qmin / T, W)
        invC = orc.invert_noise(C)[0]
        BW, _, imin = orc.ucb_alignment(p, T, W, qmin)
        Fref = orc.ucb_fisher_fd(p, T, W, qmin, invC)
        s = dict(p=p, qmin=qmin, invC=invC, R=np.zeros((3, W), complex), imin=imin)
        _, _, F = run_derivs(gl, [s], T, BW, W, mode=1, with_resid=False)
        d = np.sqrt(np.outer(np.diag(Fref), np.diag(Fref)))
        err = np.max(np.abs(F[0] - Fref) / d)
        print("mode-1 vs reference FD fisher:", err)
        assert err < 1e-4


def test_fd_fisher_dropin_matches_oracle(gl, orc):
    from glass_amd.derivs import ucb_fisher
    W = 256
    cfg = synth.config0(seed=5, n=6, W=W)
    o = gl.analytic_orbit()
    for i in range(6):
        p = cfg["params"][i]
        qmin = int(cfg["qmin"][i])
        invC = orc.invert_noise(cfg["C"][i])[0]
        Fref = orc.ucb_fisher_fd(p, cfg["T"], W, qmin, invC)
        F, w, v = ucb_fisher(o, "sangria", cfg["T"], 0.0, W, qmin, p, invC)
        rel = np.linalg.norm(F - Fref) / np.linalg.norm(Fref)
        print(i, "FD fisher rel (Frobenius):", rel)
        assert rel < 1e-6
        assert np.allclose(v @ np.diag(w) @ v.T, F, rtol=1e-10, atol=1e-10 * np.abs(F).max())


def test_autograd_binding(gl, orc):
    import torch
    from glass_amd.batch import pack_invC
    from glass_amd.derivs import UcbSourceLogL, source_derivs
    BW, W = 128, 256
    srcs = make_sources(orc, 5, 14, BW, W, T=synth.YEAR_T)
    dev = torch.device("cuda")
    params = torch.tensor(np.stack([s["p"] for s in srcs]), device=dev, requires_grad=True)
    imin = torch.tensor([s["imin"] for s in srcs], dtype=torch.int32, device=dev)
    invC = torch.tensor(np.stack([pack_invC(s["invC"]) for s in srcs]), device=dev)
    R = torch.tensor(np.stack([s["R"] for s in srcs]), device=dev)
    L = UcbSourceLogL().apply(params, gl.analytic_orbit(), synth.YEAR_T, BW, imin, invC, R)
    wts = torch.arange(1, 6, dtype=torch.float64, device=dev)
    (L * wts).sum().backward()
    L2, g, F = source_derivs(gl.analytic_orbit(), synth.YEAR_T, params.detach(), BW, imin, invC, R)
    # the backward is the adjoint kernel's gradient: the forward-mode one to
    # rounding, in the Fisher metric
    sig = torch.sqrt(torch.diagonal(F, dim1=1, dim2=2))
    assert torch.max(torch.abs(params.grad - wts[:, None] * g) / (wts[:, None] * sig)).item() < 1e-10
    assert torch.max(torch.abs(L.detach() - L2) / torch.abs(L2)).item() < 1e-12
    # without a gradient to compute, ucb_source_logl runs the likelihood only
    from glass_amd.derivs import ucb_source_logl
    with torch.no_grad():
        L3 = ucb_source_logl(params, gl.analytic_orbit(), synth.YEAR_T, BW, imin, invC, R)
    assert torch.equal(L3, L.detach())
    L4 = ucb_source_logl(params, gl.analytic_orbit(), synth.YEAR_T, BW, imin, invC, R)
    assert L4.requires_grad and torch.equal(L4.detach(), L.detach())


@pytest.mark.parametrize("BW,W,fmt,mode", [(32, 64, "sangria", 0), (64, 128, "phase", 0),
                                            (128, 256, "frequency", 1), (256, 512, "sangria", 0),
                                            (512, 1024, "sangria", 1), (1024, 2048, "phase", 0)])
def test_adjoint_gradient_vs_forward_mode(gl, orc, BW, W, fmt, mode):
    """The reverse-mode gradient (glass_hip_fd_source_grad) against the
    forward-mode tangents (glass_hip_fd_source_derivs) on the same sources:
    the same derivative by another route, so rounding only: 1e-10 in the
    Fisher metric; L to 1e-12; every TDI format, both weight modes, with
    and without a residual, sources with |cos theta| = 1 included."""
    import torch
    from glass_amd.batch import pack_invC
    from glass_amd.derivs import source_derivs, source_grad
    srcs = make_sources(orc, 6, BW + W, BW, W)
    srcs[0]["p"][1] = 1.0
    srcs[1]["p"][1] = -1.0
    dev = torch.device("cuda")
    params = torch.tensor(np.stack([s["p"] for s in srcs]), device=dev)
    imin = torch.tensor([s["imin"] for s in srcs], dtype=torch.int32, device=dev)
    invC = torch.tensor(np.stack([pack_invC(s["invC"]) for s in srcs]), device=dev)
    for R in (torch.tensor(np.stack([s["R"] for s in srcs]), device=dev), None):
        o = gl.analytic_orbit()
        Lf, gf, F = source_derivs(o, T4, params, BW, imin, invC, R, fmt=fmt, mode=mode)
        La, ga = source_grad(o, T4, params, BW, imin, invC, R, fmt=fmt, mode=mode)
        L0, none = source_grad(o, T4, params, BW, imin, invC, R, fmt=fmt, mode=mode, want_grad=False)
        assert none is None and torch.equal(L0, La)
        Lf, gf, F, La, ga = (x.cpu().numpy() for x in (Lf, gf, F, La, ga))
        sig = np.sqrt(np.abs(np.einsum("nii->ni", F)))
        nr = np.sqrt(6.0 * BW + np.abs(F[:, 3, 3]))
        err = np.max(np.abs(ga - gf) / (sig * nr[:, None]))
        eL = np.max(np.abs(La - Lf) / (np.abs(F[:, 3, 3]) + np.sqrt(6.0 * BW * np.abs(F[:, 3, 3]))))
        print(BW, fmt, mode, "resid" if R is not None else "no resid", "adjoint vs forward-mode:", err, eL)
        assert np.all(np.isfinite(ga))
        assert err < 1e-10 and eL < 1e-12


def test_adjoint_gradient_vs_central_fd(gl, orc):
    """The adjoint gradient against five-point central differences of the
    oracle's L(theta) (the forward-mode test's bound and steps)."""
    import torch
    from glass_amd.batch import pack_invC
    from glass_amd.derivs import source_grad
    BW, W = 256, 512
    srcs = make_sources(orc, 4, 19, BW, W)
    dev = torch.device("cuda")
    L, g = source_grad(gl.analytic_orbit(), T4, torch.tensor(np.stack([s["p"] for s in srcs]), device=dev), BW,
                       torch.tensor([s["imin"] for s in srcs], dtype=torch.int32, device=dev),
                       torch.tensor(np.stack([pack_invC(s["invC"]) for s in srcs]), device=dev),
                       torch.tensor(np.stack([s["R"] for s in srcs]), device=dev))
    _, _, F = run_derivs(gl, srcs, T4, BW, W)
    L, g = L.cpu().numpy(), g.cpu().numpy()
    for e, s in enumerate(srcs):
        Lref = oracle_L(orc, s["p"], s, T4, BW, W)
        assert abs(L[e] - Lref) <= 1e-10 * max(abs(Lref), 1.0)
        gfd = np.array([stencil(lambda x: oracle_L(orc, x, s, T4, BW, W), s["p"], k) for k in range(8)])
        err = np.abs(g[e] - gfd) / np.sqrt(np.diag(F[e]))
        print(e, "adjoint grad err / sigma^-1:", err)
        assert err.max() < 2e-5
