"""Exact derivatives from the batched plan's state (glass_hip_fd_batch_derivs)
against glass_hip_fd_source_derivs on the same source with the residual
without the source built explicitly on the host (data minus every other
source's drop-in template), and after accepts.  Bound: 1e-10 of the
matrix scale (the two paths differ only in how the residual is rounded).
"""
import numpy as np
import pytest

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
        if BW > 1024:
            assert np.isnan(L[e]) and np.all(np.isnan(F[e]))
            continue
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
