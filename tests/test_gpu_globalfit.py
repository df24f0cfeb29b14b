"""The global-fit layout at reduced chain count: density-weighted segments
holding more than eight binaries per chain (glass_hip_fd_batch_create_ex),
proposals checked against the oracle, the one-graph sweep (eval + exact
Fisher items) reproducing the eager calls bit for bit, and the plan's
template memory following the per-window slot rule.

Two bands: 0.1-1.2 mHz (60 segments, many binaries per chain) and the full
0.1-25 mHz band of the global fit (1,530 segments, capped at 2048 bins,
templates of 128-2048 bins, 2 x 2 chains per segment).  Towards 0.1 mHz
the XYZ noise covariance is nearly singular (X+Y+Z carries almost no
noise: condition ~1e5), so chi^2 cancels between the diagonal and cross
terms and any two summation orders differ by ~1e-16 of
sum |d_a| |d_b| |invC_ab|.  The bound is 1e-12 of that scale for every
proposal, and relative 1e-10 where the window starts above 0.5 mHz."""
import numpy as np
import pytest

from glass_amd import synth
from test_gpu_batch import TOL, oracle_chains

pytestmark = pytest.mark.gpu


BANDS = {"0.1-1.2mHz": 1.2e-3, "0.1-25mHz": 2.5e-2}


@pytest.fixture(scope="module", params=sorted(BANDS))
def gf(request, gl, orc):
    o = orc.analytic_orbit()
    cfg = synth.config4(seed=9, nsrc=600, nseg=60, ntemp=2, nchain=2, f_lo=1e-4, f_hi=BANDS[request.param],
                        waveform=lambda p, T, BW: orc.ucb_waveform(p, T, BW),
                        alignment=lambda p, T, W, q: orc.ucb_alignment(p, T, W, q, orbit=o))
    from glass_amd.batch import FdBatch
    E = len(cfg["p_chain"])
    plan = FdBatch(gl.analytic_orbit(), "sangria", cfg["T"], 0.0, (cfg["group_qmin"], cfg["group_W"]),
                   cfg["data_flat"], cfg["invC_flat"], cfg["chain_group"], cfg["chain_nsrc"], cfg["src_params"], E,
                   flat=True)
    yield cfg, plan, E
    plan.close()


def test_layout(gf):
    cfg, plan, E = gf
    assert cfg["max_src"] > 8 and plan.max_src == cfg["max_src"]
    W = cfg["group_W"]
    cap = np.array([min(4096, 1 << int(np.floor(np.log2(w)))) for w in W])
    per = len(cfg["chain_group"]) // len(W)
    want = per * int((cfg["nsrc_per_group"] * 3 * cap).sum()) * 16
    tb, total = plan.memory()
    assert tb == want and total > tb


def test_eval_vs_oracle(gf, orc):
    cfg, plan, E = gf
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    logL, snr = plan.eval(E)
    logL, snr = logL.cpu().numpy(), snr.cpu().numpy()
    _, sp = plan.state()
    import ctypes
    idx = np.random.default_rng(1).choice(E, 60, replace=False)
    cs, _ = oracle_chains(orc, cfg, sp, chains=np.unique(cfg["p_chain"][idx]))
    T = cfg["T"]
    worst_scaled, worst_rel_hi = 0.0, 0.0
    for e in idx:
        c, s, p = int(cfg["p_chain"][e]), int(cfg["p_src"][e]), cfg["p_params"][e]
        ref = orc.lib().orc_eval_logl(ctypes.byref(orc.analytic_orbit()), T, 0.0, ctypes.byref(cs.arr[c]), s,
                                      p.ctypes.data_as(orc._dp))
        g = int(cfg["chain_group"][c])
        d, ic = np.abs(cfg["data"][g]), np.abs(cfg["invC"][g])
        scale = float(sum(np.sum(d[a] * d[b] * ic[a, b]) for a in range(3) for b in range(3)))
        worst_scaled = max(worst_scaled, abs(logL[e] - ref) / scale)
        if cfg["group_qmin"][g] / T > 5e-4:
            worst_rel_hi = max(worst_rel_hi, abs(logL[e] - ref) / abs(ref))
    print("global-fit layout: worst |dlogL| / scale", worst_scaled, "; rel above 0.5 mHz", worst_rel_hi)
    assert worst_scaled <= 1e-12
    assert worst_rel_hi <= TOL
    # snr: the same two gates on snr^2 = sum h_a conj(h_b) invC_ab, whose
    # terms cancel the same way where the covariance is nearly singular
    worst_snr_scaled, worst_snr_hi = 0.0, 0.0
    for e in idx[:16]:
        c, p = int(cfg["p_chain"][e]), cfg["p_params"][e]
        g = int(cfg["chain_group"][c])
        W, qmin = int(cfg["group_W"][g]), int(cfg["group_qmin"][g])
        b, _, _ = orc.ucb_alignment(p, T, W, qmin)
        h = orc.ucb_waveform(p, T, b)[:3]
        sref = orc.snr(h[0], h[1], h[2], cfg["invC"][g], b)
        ic = np.abs(cfg["invC"][g])[:, :, :b]
        scale = 2.0 * float(sum(np.sum(np.abs(h[a]) * np.abs(h[bb]) * ic[a, bb]) for a in range(3) for bb in range(3)))
        worst_snr_scaled = max(worst_snr_scaled, abs(snr[e] ** 2 - sref ** 2) / scale)
        if qmin / T > 5e-4:
            worst_snr_hi = max(worst_snr_hi, abs(snr[e] - sref) / sref)
    print("global-fit layout: snr^2 worst |d| / scale", worst_snr_scaled, "; snr rel above 0.5 mHz", worst_snr_hi)
    assert worst_snr_scaled <= 1e-12
    assert worst_snr_hi <= TOL


def test_sweep_graph_matches_eager(gf):
    import torch
    cfg, plan, E = gf
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    l0, s0 = plan.eval(E)
    l0, s0 = l0.cpu().numpy().copy(), s0.cpu().numpy().copy()
    dch = torch.tensor(cfg["f_chain"], device="cuda")
    dsr = torch.tensor(cfg["f_src"], device="cuda")
    L0, g0, F0 = (x.cpu().numpy() for x in plan.derivs(dch, dsr))
    n = dch.numel()
    dL = torch.empty(n, dtype=torch.float64, device="cuda")
    dg = torch.empty((n, 8), dtype=torch.float64, device="cuda")
    dF = torch.empty((n, 8, 8), dtype=torch.float64, device="cuda")
    for r in range(3):
        plan.logL.zero_()
        dF.zero_()
        plan.sweep_graph(E, dch, dsr, dL, dg, dF)
        torch.cuda.synchronize()
        torch.randn(1000, device="cuda").sum().item()
        assert np.array_equal(plan.logL[:E].cpu().numpy(), l0), r
        assert np.array_equal(dF.cpu().numpy(), F0, equal_nan=True), r
        assert np.array_equal(dg.cpu().numpy(), g0, equal_nan=True)
        assert np.array_equal(dL.cpu().numpy(), L0, equal_nan=True)
    assert np.isfinite(F0).all()
