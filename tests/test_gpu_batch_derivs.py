"""Exact derivatives from the batched plan's state (glass_hip_fd_batch_derivs)
against glass_hip_fd_source_derivs on the same source with the residual
without the source built explicitly on the host (data minus every other
source's drop-in template), and after accepts.  Bound: 1e-10 of the
matrix scale (the two paths differ only in how the residual is rounded).
"""
import numpy as np
import pytest

from glass_amd import synth
from test_gpu_batch import make_cfg, new_plan

pytestmark = pytest.mark.gpu


def explicit(gl, cfg, sp, c, s):
    """(params, BW, imin, W, residual without source s [3, W], invC [3,3,W])."""
    from glass_amd.batch import ucb_alignment
    o = gl.analytic_orbit()
    T = cfg["T"]
    g = int(cfg["chain_group"][c])
    W, qmin = int(cfg["group_W"][g]), int(cfg["group_qmin"][g])
    r = np.array(cfg["data"][g], dtype=np.complex128)
    for s2 in range(int(cfg["chain_nsrc"][c])):
        if s2 == s:
            continue
        BW, _, imin = ucb_alignment(o, T, W, qmin, sp[c, s2])
        h = gl.ucb_waveform(o, "sangria", T, 0.0, sp[c, s2], 8, BW)[:3]
        lo, hi = max(0, imin), min(W, imin + BW)
        r[:, lo:hi] -= h[:, lo - imin:hi - imin]
    BW, _, imin = ucb_alignment(o, T, W, qmin, sp[c, s])
    return sp[c, s], BW, imin, W, r, cfg["invC"][g]


def compare(gl, cfg, plan, sp, items):
    import torch
    from glass_amd import derivs
    from glass_amd.batch import pack_invC
    chain = [c for c, s in items]
    src = [s for c, s in items]
    L, g, F = plan.derivs(chain, src)
    L, g, F = L.cpu().numpy(), g.cpu().numpy(), F.cpu().numpy()
    worst = 0.0
    for e, (c, s) in enumerate(items):
        p, BW, imin, W, r, ic = explicit(gl, cfg, sp, c, s)
        assert np.isfinite(L[e]) and np.all(np.isfinite(F[e])), (c, s, BW)
        dev = torch.device("cuda")
        Lr, gr, Fr = derivs.source_derivs(gl.analytic_orbit(), cfg["T"],
                                          torch.tensor(np.asarray(p)[None], device=dev), BW,
                                          torch.tensor([imin], dtype=torch.int32, device=dev),
                                          torch.tensor(pack_invC(np.asarray(ic)), device=dev),
                                          resid=torch.tensor(r, device=dev))
        Lr, gr, Fr = (x.cpu().numpy()[0] for x in (Lr, gr, Fr))
        # scales: Fisher sqrt(F_kk F_ll); gradient (R - h | dh_k) <= |R - h| sqrt(F_kk)
        # with |R - h|^2 ~ 6 BW (noise) + F_33 (the template's own norm);
        # L = (R|h) - (h|h)/2 ~ F_33 + sqrt(6 BW F_33)
        d = np.sqrt(np.abs(np.diag(Fr)))
        nr = np.sqrt(6.0 * BW + abs(Fr[3, 3]))
        eF = np.max(np.abs(F[e] - Fr) / np.outer(d, d))
        eg = np.max(np.abs(g[e] - gr) / (d * nr))
        eL = abs(L[e] - Lr) / (abs(Fr[3, 3]) + np.sqrt(6.0 * BW * abs(Fr[3, 3])))
        worst = max(worst, eF, eg, eL)
    return worst


def test_batch_derivs_vs_source_derivs(gl, orc):
    cfg = make_cfg(orc, seed=21, ngroup=5, ntemp=2, nprop=4)
    E = len(cfg["p_chain"])
    plan = new_plan(gl, cfg, E)
    _, sp = plan.state()
    items = [(c, s) for c in range(len(cfg["chain_group"])) for s in range(int(cfg["chain_nsrc"][c]))]
    worst = compare(gl, cfg, plan, sp, items)
    print("plan derivs vs explicit-residual source_derivs, worst", worst, "over", len(items))
    assert worst < 1e-10
    # after accepting one proposal per chain the plan's residuals and
    # templates follow the new parameters
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    plan.eval(E)
    nchain = len(cfg["chain_group"])
    acc = [int(np.nonzero(cfg["p_chain"] == c)[0][0]) for c in range(nchain)]
    plan.accept(acc)
    _, sp2 = plan.state()
    worst = compare(gl, cfg, plan, sp2, items)
    print("after accepts, worst", worst)
    assert worst < 1e-10
    plan.close()


def test_batch_derivs_bad_items(gl, orc):
    cfg = make_cfg(orc, seed=22, ngroup=2, ntemp=1, nprop=2)
    plan = new_plan(gl, cfg, 8)
    L, g, F = plan.derivs([0, 99, 0], [0, 0, 7])
    L = L.cpu().numpy()
    assert np.isfinite(L[0]) and np.isnan(L[1]) and np.isnan(L[2])
    plan.close()


@pytest.mark.parametrize("BWt", [2048, 4096])
def test_batch_derivs_big_bw_vs_fd(gl, orc, BWt):
    """BW 2048 and 4096 (the decimated derivative kernels) against the oracle's
    L(theta) = (R|h) - (h|h)/2 (1e-10) and five-point central differences of
    L and of the templates (q and BW held fixed), as the BW <= 1024 checks
    do.  These binaries sit at 12-24 mHz over 4 yr: slow-series phases of
    ~1.6e7 rad carry ~4e-9 rad of rounding that is not smooth in theta, so
    the differences have a noise floor of ~1e-5 in the Fisher metric
    (bound 2e-4; a misplaced bin or tangent would be O(1)).  The stencil is
    repeated at three times the steps: noise falls as 1/h.  BW 4096 needs
    loud sources (SNR > 1025 sets the spread) in 4096-bin windows; the
    per-source API is checked against the plan on the same sources."""
    import torch
    from glass_amd import derivs
    from glass_amd.batch import FdBatch, pack_invC
    from test_gpu_derivs import STEPS, bform, oracle_h, oracle_L, stencil
    T = synth.T_4YR
    W = 4096
    rng = np.random.default_rng(31 if BWt == 2048 else 37)
    o = orc.analytic_orbit()
    lA = (-22.0, -21.0) if BWt == 2048 else (-21.0, -20.0)
    srcs = []
    while len(srcs) < 3:
        p = synth.draw_params(rng, 1, T, f_lo=12e-3, f_hi=24e-3)[0]
        p[0] = np.floor(p[0]) + rng.uniform(0.25, 0.75)
        p[3] = np.log(10.0 ** rng.uniform(*lA))
        q = int(p[0] / T * T)
        qmin = q - W // 2
        BW, _, imin = orc.ucb_alignment(p, T, W, qmin, orbit=o)
        if BW != BWt:
            continue
        C = synth.xyz_noise_cov(T, qmin / T, W)
        invC = orc.invert_noise(C)[0]
        R = synth.noise_draw(rng, C)
        srcs.append(dict(p=p, qmin=qmin, invC=invC, R=R, imin=imin))
    n = len(srcs)
    plan = FdBatch(gl.analytic_orbit(), "sangria", T, 0.0, [s["qmin"] for s in srcs], [s["R"] for s in srcs],
                   [s["invC"] for s in srcs], np.arange(n), np.ones(n, dtype=np.int32),
                   np.stack([np.concatenate([s["p"][None], np.zeros((7, 8))]) for s in srcs]), 4)
    L, g, F = (x.cpu().numpy() for x in plan.derivs(np.arange(n), np.zeros(n)))
    plan.close()
    # the per-source API on the same sources (residual = data, window offset imin)
    dev = torch.device("cuda")
    Ls, gs, Fs = (x.cpu().numpy() for x in derivs.source_derivs(
        gl.analytic_orbit(), T, torch.tensor(np.stack([s["p"] for s in srcs]), device=dev), BWt,
        torch.tensor([s["imin"] for s in srcs], dtype=torch.int32, device=dev),
        torch.tensor(np.stack([pack_invC(s["invC"]) for s in srcs]), device=dev),
        resid=torch.tensor(np.stack([s["R"] for s in srcs]), device=dev)))
    for e in range(n):
        d = np.sqrt(np.abs(np.diag(F[e])))
        assert np.max(np.abs(Fs[e] - F[e]) / np.outer(d, d)) < 1e-10
        assert abs(Ls[e] - L[e]) <= 1e-10 * max(abs(L[e]), 1.0)
    for e, s in enumerate(srcs):
        # the plan's residual is data - h; the derivatives add h back: R = data
        Lref = oracle_L(orc, s["p"], s, T, BWt, W)
        assert abs(L[e] - Lref) <= 1e-10 * max(abs(Lref), 1.0), (e, L[e], Lref)
        sig = np.sqrt(np.diag(F[e]))
        errs = []
        for scale in (1.0, 3.0):
            steps = STEPS * scale
            gfd = np.array([stencil(lambda x: oracle_L(orc, x, s, T, BWt, W), s["p"], k, steps[k])
                            for k in range(8)])
            eg = np.max(np.abs(g[e] - gfd) / sig)
            dh = [stencil(lambda x: oracle_h(orc, x, T, BWt, W, s["imin"]), s["p"], k, steps[k]) for k in range(8)]
            Ffd = np.array([[bform(dh[a], dh[b], s["invC"]) for b in range(8)] for a in range(8)])
            d = np.sqrt(np.outer(np.diag(Ffd), np.diag(Ffd)))
            eF = np.max(np.abs(F[e] - Ffd) / d)
            errs.append((eg, eF))
        print(e, f"BW {BWt}: (grad err / sigma^-1, fisher err) at steps x1, x3:", errs)
        assert max(max(x) for x in errs) < 2e-4
