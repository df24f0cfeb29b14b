"""GPU parity of the UCB sampler glue (glass_hip_ucb_sampler_*, the batched
eigensolver) against the oracle (oracle/orc_sampler.c + the oracle's
waveform, likelihood and snr).

The device runs every slot's steps in lockstep; the oracle replays each
slot on its own with the same rand_r seed and the eigen data the device
stored, regenerating a source's template on accept.  Tolerances: the same
proposal types, accepts, swaps and rand_r states; parameters 1e-12 and
chain logL 1e-10 relative (the likelihood gate); eigen-decompositions
1e-13 (the same Jacobi arithmetic with contraction off on both sides).
"""
import ctypes

import numpy as np
import pytest

from test_gpu_batch import make_cfg, new_plan, oracle_chains, oracle_snr

pytestmark = pytest.mark.gpu

W_DEFAULT = np.array([0.1, 0, 0, 0, 0.7, 0.1, 0.1, 0, 0])


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_fisher_eigen_vs_oracle(gl, orc):
    import torch
    from glass_amd import sampler
    rng = np.random.default_rng(1)
    F = []
    for k in range(600):
        B = rng.standard_normal((8, 8))
        H = B @ B.T + 8 * np.eye(8)
        D = np.diag(10.0 ** rng.uniform(-3, 3, 8)) if k % 2 else np.eye(8)
        A = D @ H @ D
        F.append(0.5 * (A + A.T))
    F = np.array(F)
    ev, la, sw = sampler.fisher_eigen(torch.tensor(F, device="cuda"))
    ev, la, sw = ev.cpu().numpy(), la.cpu().numpy(), sw.cpu().numpy()
    worst_v = worst_l = 0.0
    for k in range(len(F)):
        e2, l2, s2 = orc.jacobi_eigen8(F[k])
        worst_l = max(worst_l, rel(la[k], l2))
        worst_v = max(worst_v, float(np.max(np.abs(ev[k] - e2))))
        assert sw[k] == s2
    print("eigen: max rel eval", worst_l, "max abs evec", worst_v)
    assert worst_l < 1e-13 and worst_v < 1e-13
    e1, l1 = sampler.matrix_eigenstuff(F[3])
    e2, l2, _ = orc.jacobi_eigen8(F[3])
    assert rel(l1, l2) < 1e-13 and np.max(np.abs(e1 - e2)) < 1e-13


def priors(cfg):
    G = len(cfg["group_W"])
    pr = np.zeros((G, 8, 2))
    for g in range(G):
        q, W = float(cfg["group_qmin"][g]), float(cfg["group_W"][g])
        pr[g] = [[q, q + W], [-1, 1], [0, 2 * np.pi], [-60, -40], [-1, 1], [0, np.pi], [0, 2 * np.pi], [-1, 1]]
    return pr


def setup_sampler(gl, orc, seed=21, ngroup=3, ntemp=4, snr_min=1.0):
    from glass_amd.sampler import UcbSampler
    cfg = make_cfg(orc, seed=seed, ngroup=ngroup, ntemp=ntemp, nprop=1)
    C = len(cfg["chain_group"])
    assert np.array_equal(cfg["chain_group"], np.repeat(np.arange(ngroup), ntemp))
    plan = new_plan(gl, cfg, C)
    pr = priors(cfg)
    smp = UcbSampler(plan, ngroup, ntemp, pr, weights=W_DEFAULT, snr_min=snr_min)
    rng = np.random.default_rng(seed + 1)
    seeds = rng.integers(1, 2 ** 32, C, dtype=np.uint64).astype(np.uint32)
    temps = np.tile(1.25 ** np.arange(ntemp), ngroup)
    smp.set_state(seeds, temps)
    return cfg, plan, smp, pr, seeds, temps


def oracle_steps(orc, cfg, pr, seeds, temps, evec, eval_, nsteps, snr_min, logl0, w=W_DEFAULT, fsp=None):
    """Replays nsteps ucb_mcmc steps of every slot (slot s on chain s)."""
    T = cfg["T"]
    P = orc.UcbProposer(w, T, snr_min=snr_min)
    cs, chains = oracle_chains(orc, cfg, cfg["src_params"])
    C = len(chains)
    out_seeds = np.zeros(C, dtype=np.int64)
    trials = np.zeros((C, 9), dtype=np.int64)
    accepts = np.zeros((C, 9), dtype=np.int64)
    logl = np.array(logl0, dtype=np.float64)
    o = orc.analytic_orbit()
    for c in range(C):
        bufs = cs.keep[c]
        nlive = int(chains[c]["nsrc"])
        sp = bufs["src_params"]
        tm = bufs["tmpl"].reshape(nlive, 3, -1)
        g = int(cfg["chain_group"][c])
        W, qmin = int(cfg["group_W"][g]), int(cfg["group_qmin"][g])
        ev = np.ascontiguousarray(evec[c, :nlive].reshape(nlive, 64))
        la = np.ascontiguousarray(eval_[c, :nlive])
        seed = int(seeds[c])
        for _ in range(nsteps):
            seed, n, npr, y, lpx, lpy, lq = P.propose(seed, sp, ev, la, pr[g], None if fsp is None else fsp[g][0])
            trials[c, npr] += 1
            if lpy == -np.inf:
                continue
            ly = orc.lib().orc_eval_logl(ctypes.byref(o), T, 0.0, ctypes.byref(cs.arr[c]), n,
                                         y.ctypes.data_as(orc._dp))
            sn = oracle_snr(orc, cfg, c, y)
            seed, acc = P.accept(seed, ly, logl[c], sn, temps[c], lpx, lpy, lq)
            if acc:
                accepts[c, npr] += 1
                b, _, i = orc.ucb_alignment(y, T, W, qmin)
                h = orc.ucb_waveform(y, T, b)
                tm[n] = 0.0
                tm[n, :, :2 * b] = h[:3].view(np.float64).reshape(3, -1)
                bufs["src_bw"][n], bufs["src_imin"][n] = b, i
                sp[n] = y
                logl[c] = ly
        out_seeds[c] = seed
    params = np.zeros_like(cfg["src_params"])
    for c in range(C):
        n = int(chains[c]["nsrc"])
        params[c, :n] = cs.keep[c]["src_params"][:n]
    return params, logl, out_seeds, trials, accepts


@pytest.fixture(scope="module")
def stepped(gl, orc):
    cfg, plan, smp, pr, seeds, temps = setup_sampler(gl, orc)
    smp.fisher()
    ev, la = smp.eigen()
    logl0, _ = plan.state()
    # the plan's Fisher matrices of every live source at the initial state
    C = len(cfg["chain_group"])
    ch = np.concatenate([np.full(int(cfg["chain_nsrc"][c]), c) for c in range(C)]).astype(np.int32)
    sr = np.concatenate([np.arange(int(cfg["chain_nsrc"][c])) for c in range(C)]).astype(np.int32)
    F0 = plan.derivs(ch, sr)[2].cpu().numpy()
    nsteps = 24
    smp.steps(nsteps)
    st = smp.state()
    logl, sp = plan.state()
    ref = oracle_steps(orc, cfg, pr, seeds, temps, ev, la, nsteps, 1.0, logl0)
    yield dict(cfg=cfg, plan=plan, smp=smp, st=st, logl=logl, sp=sp, ref=ref, ev=ev, la=la, temps=temps,
               F0=(ch, sr, F0))
    smp.close()
    plan.close()


def test_stored_eigen_is_eigen_of_plan_fisher(stepped, orc):
    """The Fisher update stores, per (chain, source), the eigen-decomposition
    of the plan's exact Fisher matrix."""
    ev, la = stepped["ev"], stepped["la"]
    ch, sr, F0 = stepped["F0"]
    worst_l = worst_v = 0.0
    for k in range(len(ch)):
        e2, l2, _ = orc.jacobi_eigen8(F0[k])
        worst_l = max(worst_l, rel(la[ch[k], sr[k]], l2))
        worst_v = max(worst_v, float(np.max(np.abs(ev[ch[k], sr[k]] - e2))))
    print("checked", len(ch), "sources, max rel eigenvalue", worst_l, "max abs eigenvector", worst_v)
    assert len(ch) > 10 and worst_l < 1e-13 and worst_v < 1e-13


def test_steps_match_oracle(stepped):
    st, sp, logl = stepped["st"], stepped["sp"], stepped["logl"]
    params, lref, sref, trials, accepts = stepped["ref"]
    print("device accepts", st["accepts"].sum(axis=0), "trials", st["trials"].sum(axis=0))
    assert np.array_equal(st["trials"], trials)
    assert np.array_equal(st["accepts"], accepts)
    assert np.array_equal(st["seeds"], sref.astype(np.uint32))
    assert rel(sp, params) < 1e-12
    assert rel(logl, lref) < 1e-10
    assert st["accepts"].sum() > 0


def test_ptswap_adapt_match_oracle(stepped, orc):
    smp, plan, temps = stepped["smp"], stepped["plan"], stepped["temps"]
    cfg = stepped["cfg"]
    G, K = len(cfg["group_W"]), smp.ntemp
    st0 = smp.state()
    logl, _ = plan.state()
    smp.ptswap(adapt=True)
    st = smp.state()
    for l in range(G):
        o = slice(l * K, (l + 1) * K)
        idx, sd, acc = orc.ucb_ptmcmc(st0["index"][o], st0["temperature"][o], logl, np.zeros_like(logl),
                                      st0["seeds"][o])
        t2 = orc.adapt_temperature_ladder(st0["temperature"][o], acc, 0)
        assert np.array_equal(st["index"][o], idx)
        assert np.array_equal(st["seeds"][o], sd)
        assert rel(st["temperature"][o], t2) < 1e-15
        assert np.array_equal(st["swaps"][o] - st0["swaps"][o], acc)


def test_sweep_graph_equals_eager(gl, orc):
    """One reference iteration (steps, Fisher update, swap + adapt) as a
    graph, twice, equals the same run eagerly on an identical plan."""
    runs = []
    for graph in (False, True):
        cfg, plan, smp, pr, seeds, temps = setup_sampler(gl, orc, seed=31)
        smp.fisher()
        for _ in range(2):
            smp.sweep(nsteps=6, adapt=True, graph=graph)
        st = smp.state()
        logl, sp = plan.state()
        ev, la = smp.eigen()
        runs.append((st, logl, sp, la))
        smp.close()
        plan.close()
    (a, la_, sa, ea), (b, lb, sb, eb) = runs
    for k in ("seeds", "index", "trials", "accepts", "swaps"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["temperature"], b["temperature"])
    assert np.array_equal(la_, lb) and np.array_equal(sa, sb) and np.array_equal(ea, eb)


def test_bad_sampler_arguments(gl, orc):
    from glass_amd.sampler import UcbSampler
    cfg = make_cfg(orc, seed=5, ngroup=2, ntemp=2, nprop=1)
    plan = new_plan(gl, cfg, 4)
    pr = priors(cfg)
    w = W_DEFAULT.copy()
    w[3] = 0.2  # extrinsic-prior draw: not built
    with pytest.raises(gl.GlassHipError):
        UcbSampler(plan, 2, 2, pr, weights=w)
    with pytest.raises(gl.GlassHipError):
        UcbSampler(plan, 3, 2, pr)
    plan.close()


def test_steps_with_fstat_proposals_match_oracle(gl, orc):
    """The reference's default proposal table, F-statistic draw and jump
    included, with a proposal density per window."""
    from glass_amd.sampler import FSTAT_WEIGHTS, UcbSampler
    w = np.array(FSTAT_WEIGHTS)
    cfg = make_cfg(orc, seed=41, ngroup=3, ntemp=4, nprop=1)
    C = len(cfg["chain_group"])
    plan = new_plan(gl, cfg, C)
    pr = priors(cfg)
    smp = UcbSampler(plan, 3, 4, pr, weights=w, snr_min=1.0)
    rng = np.random.default_rng(42)
    seeds = rng.integers(1, 2 ** 32, C, dtype=np.uint64).astype(np.uint32)
    temps = np.tile(1.25 ** np.arange(4), 3)
    smp.set_state(seeds, temps)
    with pytest.raises(gl.GlassHipError):
        smp.steps(1)  # F-statistic weights without densities
    T = cfg["T"]
    fsp = []
    for g in range(3):
        W, q = int(cfg["group_W"][g]), int(cfg["group_qmin"][g])
        ten = rng.uniform(0.2, 2.0, (W, 20, 20))
        ten *= ten.size / ten.sum()
        smp.set_fstat(g, ten, 1.0 / T, ten.max(), q)
        fsp.append(orc.fstat_prop(ten, 1.0 / T, ten.max(), q))
    smp.fisher()
    ev, la = smp.eigen()
    logl0, _ = plan.state()
    nsteps = 20
    smp.steps(nsteps)
    st = smp.state()
    logl, sp = plan.state()
    params, lref, sref, trials, accepts = oracle_steps(orc, cfg, pr, seeds, temps, ev, la, nsteps, 1.0, logl0, w=w,
                                                       fsp=fsp)
    print("accepts", st["accepts"].sum(axis=0), "trials", st["trials"].sum(axis=0))
    assert st["trials"][:, [1, 2]].sum() > 0
    assert np.array_equal(st["trials"], trials)
    assert np.array_equal(st["accepts"], accepts)
    assert np.array_equal(st["seeds"], sref.astype(np.uint32))
    assert rel(sp, params) < 1e-12
    assert rel(logl, lref) < 1e-10
    smp.close()
    plan.close()


def test_steps_with_bw4096_sources_match_oracle(gl, orc):
    """Loud high-frequency sources in 4096-bin windows (BW 4096 templates):
    their exact Fisher matrices are finite and positive definite, so the
    Fisher proposal draws along real eigen-directions (no 0.01 x fallback
    of a failed decomposition, ucb_sampler.hip's guard), and the device
    steps match the oracle step for step."""
    from glass_amd.batch import ucb_alignment
    from glass_amd.sampler import UcbSampler
    cfg = make_cfg(orc, seed=71, ngroup=2, ntemp=3, nprop=1, windows=(4096,), f_lo=15e-3, f_hi=18e-3)
    T = cfg["T"]
    o = gl.analytic_orbit()
    loud = np.log(6e-21)
    cfg["src_params"][:, 0, 3] = loud
    C = len(cfg["chain_group"])
    bws = []
    for c in range(C):
        g = int(cfg["chain_group"][c])
        bws.append(ucb_alignment(o, T, int(cfg["group_W"][g]), int(cfg["group_qmin"][g]),
                                 cfg["src_params"][c, 0])[0])
    print("BW of the loud sources", bws)
    assert sum(b == 4096 for b in bws) >= 2
    plan = new_plan(gl, cfg, C)
    pr = priors(cfg)
    smp = UcbSampler(plan, 2, 3, pr, weights=W_DEFAULT, snr_min=1.0)
    rng = np.random.default_rng(72)
    seeds = rng.integers(1, 2 ** 32, C, dtype=np.uint64).astype(np.uint32)
    temps = np.tile(1.25 ** np.arange(3), 2)
    smp.set_state(seeds, temps)
    smp.fisher()
    ev, la = smp.eigen()
    for c in range(C):
        n = int(cfg["chain_nsrc"][c])
        assert np.all(np.isfinite(la[c, :n])) and np.all(la[c, :n] > 0), (c, la[c, :n])
        assert np.all(np.isfinite(ev[c, :n]))
    logl0, _ = plan.state()
    nsteps = 12
    smp.steps(nsteps)
    st = smp.state()
    logl, sp = plan.state()
    params, lref, sref, trials, accepts = oracle_steps(orc, cfg, pr, seeds, temps, ev, la, nsteps, 1.0, logl0)
    print("accepts", st["accepts"].sum(axis=0), "trials", st["trials"].sum(axis=0))
    assert st["trials"][:, 4].sum() > 0
    assert np.array_equal(st["trials"], trials)
    assert np.array_equal(st["accepts"], accepts)
    assert np.array_equal(st["seeds"], sref.astype(np.uint32))
    assert rel(sp, params) < 1e-12
    assert rel(logl, lref) < 1e-10
    smp.close()
    plan.close()
