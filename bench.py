#!/usr/bin/env python3
"""bench.py -- batched UCB waveform + log-likelihood throughput on one MI355X.

One step = one parallel-tempering sweep of likelihood evaluations of the
batched PTMCMC case: 256 data windows (512-4096 complex128 bins, Tobs = 2 yr)
x 16 temperatures = 4096 chains, each with 1-8 galactic binaries
(f0 ~ U[0.5, 20] mHz), 80 proposals per chain = 327,680 evaluations.  One
evaluation = the reference's per-proposal likelihood work: regenerate one
source's template (ucb_alignment + ucb_waveform) and return the chain's
logL and the template's snr (ucb/src/glass_ucb_sampler.c:139-155).  All
inputs are resident in HBM before the timed region; the proposals are
evaluated by the library's per-BW-class kernels.

Printed: one JSON line with the throughput, the plain per-call drop-in rate
(the "plain" method) for comparison, a parity check of a sample of the
timed evaluations against the FP64 oracle, the roofline of the dominant
kernel and the oracle's CPU throughput on the host cores.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "UCB waveform+logL evals/s, batched vs plain, at logL rel err ≤1e-10 vs FP64"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


# HBM bytes per launch of the batched per-class kernels, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this workload and scaled by the same counters on a 256 MiB device copy (fetch x2.00,
# write x1.00): profiles/eval_kernel_batched_fd_sweep.md.  Keyed by template bins N.
MEASURED_TRAFFIC = {512: 2.3859e9, 1024: 1.8438e9, 2048: 1.5761e9, 256: 8.5133e8, 4096: 1.6860e8,
                    128: 2.4454e8, 64: 7.8533e7, 32: 1.8974e6}

# HBM bytes per launch of logl_chunk_kernel on the inner-product leg (rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, same calibration): profiles/logl_chunk_kernel_batched_inner_product.md.
INNER_TRAFFIC = 1.1350e9


def wave_flops(N):
    """Algorithmic FP64 work of one template of N bins plus its delta-logL:
    940 N + 30 N log2 N for the slow series, six FFTs and TDI (each sin/cos
    counted as 20 operations), + 36 N for the 3-channel weighted inner
    products."""
    return 940.0 * N + 30.0 * N * math.log2(N) + 36.0 * N


def eval_bytes(N):
    """Algorithmic HBM bytes of one fused evaluation over N bins: chain
    residual 48 N + invC planes 48 N + the source's stored template 48 N +
    parameters 64 + logL/snr out 16."""
    return 144.0 * N + 80.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_inputs(gl, synth, ngroup, ntemp, nprop, seed):
    from glass_amd.batch import ucb_alignment
    o = gl.analytic_orbit()
    cfg = synth.config1(seed=seed, ngroup=ngroup, ntemp=ntemp, nprop_per_chain=nprop,
                        waveform=lambda p, T, BW: gl.ucb_waveform(o, "sangria", T, 0.0, p, 8, BW),
                        alignment=lambda p, T, W, q: ucb_alignment(o, T, W, q, p))
    cfg["invC"] = [gl.invert_noise_covariance_matrix(C)[0] for C in cfg["C"]]
    return o, cfg


def plain_rate(gl, synth, seconds=2.0):
    """The plain method: per-call drop-ins ucb_waveform + gaussian_log_likelihood
    on host memory (64 vectors, BW 128, W 256, Tobs 1 yr)."""
    cfg = synth.config0(seed=7, waveform=None)
    o = gl.analytic_orbit()
    T, W, BW = cfg["T"], cfg["W"], cfg["BW"]
    invCs = [gl.invert_noise_covariance_matrix(C)[0] for C in cfg["C"]]
    datas = []
    for i in range(len(cfg["params"])):
        d = np.zeros((5, W), complex)
        d[:3] = cfg["data"][i]
        datas.append(d)
    n = 0
    t0 = time.perf_counter()
    while True:
        i = n % len(cfg["params"])
        h = gl.ucb_waveform(o, "sangria", T, 0.0, cfg["params"][i], 8, BW)
        m = np.zeros((5, W), complex)
        synth.scatter(m[:3], h[:3], int(cfg["imin"][i]))
        gl.gaussian_log_likelihood(datas[i], m, invCs[i])
        n += 1
        if n % 64 == 0 and time.perf_counter() - t0 > seconds:
            break
    return n / (time.perf_counter() - t0)


def plain_leg(gl, cfg, logL_batched, seconds=2.0, nthreads=16, seed=5):
    """The plain method on the same workload: a sample of this sweep's
    proposals (all proposals of 32 random chains), each evaluated as the
    reference's sampler does per proposal (glass_ucb_sampler.c:139-155) but
    through the per-call drop-ins on host memory: ucb_alignment, the
    ucb_waveform drop-in, the window model (the chain's other sources'
    stored templates plus the new one), the gaussian_log_likelihood drop-in
    and the snr drop-in.  Timed from one host thread and from nthreads host
    threads (the reference calls these from its OpenMP threads over chains,
    apps/src/ucb_mcmc.c:230; each thread has its own HIP stream).  Also
    checks that the drop-in logL equals the batched one."""
    from concurrent.futures import ThreadPoolExecutor
    from glass_amd.batch import ucb_alignment
    o = gl.analytic_orbit()
    T = cfg["T"]
    rng = np.random.default_rng(seed)
    C = len(cfg["chain_group"])
    chains = rng.choice(C, size=min(32, C), replace=False)
    base = {}
    for c in chains:
        g = int(cfg["chain_group"][c])
        W, qmin = int(cfg["group_W"][g]), int(cfg["group_qmin"][g])
        n = int(cfg["chain_nsrc"][c])
        tm = []
        for s_ in range(n):
            b, _, i = ucb_alignment(o, T, W, qmin, cfg["src_params"][c, s_])
            h = gl.ucb_waveform(o, "sangria", T, 0.0, cfg["src_params"][c, s_], 8, b)
            m = np.zeros((5, W), complex)
            lo, hi = max(0, i), min(W, i + b)
            m[:, lo:hi] = h[:, lo - i:hi - i]
            tm.append(m)
        tot = np.sum(tm, axis=0)
        d = np.zeros((5, W), complex)
        d[:3] = cfg["data"][g]
        base[int(c)] = dict(W=W, qmin=qmin, data=d, invC=np.ascontiguousarray(cfg["invC"][g]),
                            minus=[tot - t for t in tm])
    sel = np.nonzero(np.isin(cfg["p_chain"], chains))[0]
    rng.shuffle(sel)

    def one(q):
        c, s_ = int(cfg["p_chain"][q]), int(cfg["p_src"][q])
        p = cfg["p_params"][q]
        B = base[c]
        W = B["W"]
        b, _, i = ucb_alignment(o, T, W, B["qmin"], p)
        h = gl.ucb_waveform(o, "sangria", T, 0.0, p, 8, b)
        m = B["minus"][s_].copy()
        lo, hi = max(0, i), min(W, i + b)
        m[:, lo:hi] += h[:, lo - i:hi - i]
        ll = gl.gaussian_log_likelihood(B["data"], m, B["invC"])
        gl.snr(h[0], h[1], h[2], B["invC"], b)
        return ll

    ll = np.array([one(q) for q in sel[:64]])
    rel = float(np.max(np.abs(ll - logL_batched[sel[:64]]) / np.abs(logL_batched[sel[:64]])))

    def timed(nt):
        k, t0 = 0, time.perf_counter()
        if nt == 1:
            while time.perf_counter() - t0 < seconds:
                one(sel[k % len(sel)])
                k += 1
        else:
            with ThreadPoolExecutor(nt) as ex:
                while time.perf_counter() - t0 < seconds:
                    list(ex.map(one, sel[:4 * nt] if len(sel) >= 4 * nt else sel))
                    k += min(len(sel), 4 * nt)
        return k / (time.perf_counter() - t0)

    return dict(rate1=timed(1), rateN=timed(nthreads), nthreads=nthreads, n=len(sel), nchains=len(chains),
                max_rel_vs_batched=rel)


def fisher_leg(gl, synth, n=8192, BW=256, W=512, steps=5, seed=3):
    """Exact-gradient Fisher evaluations (adjoint gradients case): n sources,
    BW 256, Tobs 4 yr, each with its own 512-bin window (noise residual and
    invC); one launch returns L, dL/dtheta and the 8x8 Fisher matrix per
    source.  Compared with the plain per-call forward-difference ucb_fisher
    drop-in (9 waveforms + Gram per call)."""
    import torch
    from glass_amd.batch import pack_invC
    from glass_amd.derivs import source_derivs, ucb_fisher
    T = synth.T_4YR
    rng = np.random.default_rng(seed)
    P = synth.draw_params(rng, n, T)
    q = P[:, 0].astype(np.int64)
    qmin = q - W // 2
    imin = (q - BW // 2) - qmin
    # windows share the PSD shape up to the band: one covariance per source,
    # vectorised
    f = qmin[:, None] / T + np.arange(W)[None, :] / T
    Spm, Sop = synth.noise_levels_sangria(f)
    x = f / synth.FSTAR
    tr = np.sin(x) ** 2
    diag = 16.0 * tr * (2.0 * (1.0 + np.cos(x) ** 2) * Spm + Sop) / 4.0
    cross = -8.0 * tr * np.cos(x) * (4.0 * Spm + Sop) / 4.0
    dev = torch.device("cuda")
    # inverse of [[d,c,c],[c,d,c],[c,c,d]]: diag (d+c)/((d-c)(d+2c)), off -c/(...)
    den = (diag - cross) * (diag + 2 * cross)
    idg, ioff = (diag + cross) / den, -cross / den
    invC = torch.tensor(np.stack([idg, idg, idg, ioff, ioff, ioff], axis=1), device=dev)
    sd = np.sqrt(np.maximum(diag - cross, 0) / 2)
    R = torch.tensor((rng.standard_normal((n, 3, W)) + 1j * rng.standard_normal((n, 3, W))) * sd[:, None, :],
                     device=dev)
    params = torch.tensor(P, device=dev)
    im = torch.tensor(imin, dtype=torch.int32, device=dev)
    o = gl.analytic_orbit()
    for _ in range(2):
        source_derivs(o, T, params, BW, im, invC, R)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        source_derivs(o, T, params, BW, im, invC, R)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # reverse mode: L + dL/dtheta from the adjoint kernel on the same
    # sources (one backward pass instead of the forward-mode tangents)
    from glass_amd.derivs import source_grad
    Lf, gf, _ = source_derivs(o, T, params, BW, im, invC, R)
    La, ga = source_grad(o, T, params, BW, im, invC, R)
    torch.cuda.synchronize()
    scale = torch.clamp(gf.abs().amax(dim=1, keepdim=True), min=1e-300)
    grad_err = float(((ga - gf).abs() / scale).max())
    L_err = float(((La - Lf).abs() / Lf.abs().clamp(min=1e-300)).max())
    e0.record()
    for _ in range(steps):
        source_grad(o, T, params, BW, im, invC, R)
    e1.record()
    torch.cuda.synchronize()
    ms_grad = e0.elapsed_time(e1) / steps
    e0.record()
    for _ in range(steps):
        source_grad(o, T, params, BW, im, invC, R, want_grad=False)
    e1.record()
    torch.cuda.synchronize()
    ms_L = e0.elapsed_time(e1) / steps
    # plain: the per-call forward-difference drop-in on host arrays
    invC33 = np.zeros((3, 3, W))
    nplain, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        i = nplain % 64
        for a in range(3):
            for b in range(3):
                invC33[a, b] = idg[i] if a == b else ioff[i]
        ucb_fisher(o, "sangria", T, 0.0, W, int(qmin[i]), P[i], invC33, eigen=False)
        nplain += 1
    plain = nplain / (time.perf_counter() - t0)
    return dict(value=n / (ms * 1e-3), ms=ms, plain=plain, n=n, ms_grad=ms_grad, ms_L=ms_L, grad_err=grad_err,
                L_err=L_err)


def wdm_leg(gl, synth, n=4096, steps=5, seed=7, cpu_seconds=6.0, nthreads=0, parity_n=64, plain=True):
    """Wavelet-domain case: n templates (config-0 distributions, Tobs 1 yr)
    on the NT 512 x NF 4096 grid (15 s cadence), stationary XYZ pixel noise;
    one launch = every template's ucb_waveform_wavelet pixels plus its
    wavelet_nwip terms against the data (snr and the logL change).  Also one
    full forward WDM transform of a 2^21-sample noise series, and the
    per-call drop-in for comparison."""
    import torch
    from glass_amd import wdm as W
    T = synth.YEAR_T
    cfg = synth.config2(n=n, seed=seed)
    w = W.Wdm.grid(cfg["NF"], cfg["NT"], cfg["dt"])
    orb = W.interpolated_analytic_orbit(gl.analytic_orbit(), T)
    B = W.WdmBatch(w, orb, T, cfg["data"], cfg["invC6"], max_templates=n)
    P = torch.tensor(cfg["params"], device="cuda")
    for _ in range(2):
        B.eval(P)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        dl, sn = B.eval(P)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    dl, sn = dl.cpu().numpy(), sn.cpu().numpy()
    nw, _, _ = B.templates(n)
    layers = {int(L): int(np.sum(nw == L * cfg["NT"])) for L in (1, 2, 3)}
    # full transform of a noise series on the device
    ND = cfg["NF"] * cfg["NT"]
    x = torch.randn(ND, dtype=torch.float64, device="cuda")
    out = torch.empty_like(x)
    W.wdm_transform(w, x, out)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        W.wdm_transform(w, x, out)
    e1.record()
    torch.cuda.synchronize()
    tms_call = e0.elapsed_time(e1) / steps
    # the device rate: the three launches of one transform replayed from a
    # graph (per call, the host's issue time exceeds the ~60 us of kernels);
    # every replay must give the eager pixels bit for bit
    ref_pix = out.clone()
    gs = torch.cuda.Stream()
    gs.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=gs):
        W.wdm_transform(w, x, out)
    torch.cuda.synchronize()
    graph_equal = True
    for _ in range(3):
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        graph_equal = graph_equal and bool(torch.equal(out, ref_pix))
    # median of 5 groups of 20 replays: a 60 us sequence is easily stretched
    # by anything else on the device or host for one group
    groups = []
    for _ in range(5):
        e0.record()
        for _ in range(20):
            graph.replay()
        e1.record()
        torch.cuda.synchronize()
        groups.append(e0.elapsed_time(e1) / 20)
    tms = float(np.median(groups))
    del graph
    # plain: the per-call drop-in (host arrays)
    rate = None
    if plain:
        X = np.zeros(w.npix)
        Y, Z = np.zeros_like(X), np.zeros_like(X)
        nplain, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 2.0:
            lst, X, Y, Z = W.ucb_waveform_wavelet(w, orb, T, cfg["params"][nplain % n], X=X, Y=Y, Z=Z)
            X[lst] = Y[lst] = Z[lst] = 0.0
            nplain += 1
        rate = nplain / (time.perf_counter() - t0)
    # wavelet Fisher (ucb_fisher_wavelet before the eigensolve): n / 16
    # sources per launch (16 templates each) vs the per-call drop-in
    nf = n // 16
    Pf = P[:nf].contiguous()
    B.fisher(Pf)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        B.fisher(Pf)
    e1.record()
    torch.cuda.synchronize()
    fms = e0.elapsed_time(e1) / steps
    frate = None
    if plain:
        nfp, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 1.0:
            W.ucb_fisher_wavelet(w, orb, T, cfg["params"][nfp % n], cfg["invC9"])
            nfp += 1
        frate = nfp / (time.perf_counter() - t0)
    res = dict(value=n / (ms * 1e-3), ms=ms, n=n, layers=layers, transform_ms=tms, transform_ms_call=tms_call,
               transform_graph_equal=graph_equal, ND=ND, plain=rate,
               fisher=dict(n=nf, ms=fms, value=nf / (fms * 1e-3), plain=frate))
    B.close()
    # the same templates with a forced layer count (templates touching 8-64
    # frequency layers: NT (L + 1) heterodyned samples, L inverse transforms
    # and NT L pixels each), first n_forced of them
    n_forced = min(n, 1024)
    forced = {}
    for L in (7, 15, 31, 63):
        BF = W.WdmBatch(w, orb, T, cfg["data"], cfg["invC6"], max_templates=n_forced, layers=L)
        PF = P[:n_forced].contiguous()
        BF.eval(PF)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(2):
            BF.eval(PF)
        e1.record()
        torch.cuda.synchronize()
        fl_ms = e0.elapsed_time(e1) / 2
        forced[str(L)] = dict(ms_per_launch=round(fl_ms, 3), templates=n_forced,
                              value=round(n_forced / (fl_ms * 1e-3), 1),
                              pixels_per_template=L * cfg["NT"])
        BF.close()
    res["forced_layers"] = forced
    # parity and CPU baseline with the oracle
    try:
        from oracle import orc
    except Exception as exc:  # noqa: BLE001
        log(f"wdm oracle unavailable ({exc}); no parity / cpu baseline")
        return res
    ow = orc.Wdm(cfg["NF"], cfg["NT"], cfg["dt"])
    oorb = orc.iorbit(T)
    idx = np.random.default_rng(seed + 1).choice(n, parity_n, replace=False)
    with orc.cr_trig():
        rdl, rsn, _, _ = orc.wdm_eval_batch(ow, oorb, T, cfg["params"][idx], cfg["data"], cfg["invC9"],
                                            nthreads=nthreads)
    res["parity"] = {"checked": int(parity_n),
                     "max_dlogL_rel_err": float(np.max(np.abs(dl[idx] - rdl) / np.maximum(np.abs(rdl), 1.0))),
                     "max_snr_rel_err": float(np.max(np.abs(sn[idx] - rsn) / rsn)), "gate": 1e-10}
    xs = x.cpu().numpy()
    ref = orc.wavelet_transform_timefreq(ow, xs)
    res["transform_normwise_err"] = float(np.max(np.abs(out.cpu().numpy() - ref)) / np.max(np.abs(ref)))
    if cpu_seconds > 0:
        reps, tot, m = 0, 0.0, 0
        sample = cfg["params"][:256]
        while tot < cpu_seconds:
            _, _, _, secs = orc.wdm_eval_batch(ow, oorb, T, sample, cfg["data"], cfg["invC9"], nthreads=nthreads)
            tot += secs
            reps += 1
            m += len(sample)
        res["cpu"] = dict(rate=m / tot, n=len(sample), reps=reps, secs=tot)
    return res


def inner_leg(gl, synth, n=4096, steps=20, warmup=3, seed=11, parity_n=64):
    """Batched full-window likelihood (glass_hip_logl_batch_eval): n
    independent gaussian_log_likelihood items with the batched case's
    windows (512-4096 bins, Tobs 2 yr), each with its own data, model and
    invC, resident in HBM.  144 algorithmic bytes per bin (3 data + 3
    model channels complex128, 6 invC planes).  Returns the launch time
    from HIP events on the launch stream and oracle parity on a sample."""
    import torch
    from glass_amd.batch import LoglBatch
    from oracle import orc

    r = synth.logl_batch_inputs(n, seed=seed, device="cuda")
    W, off = r["W"], r["off"]
    bins = int(W.astype(np.int64).sum())
    B = LoglBatch(W, off, off, off)
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        B.eval(r["data"], r["model"], r["invC"], out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        B.eval(r["data"], r["model"], r["invC"], out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    got = out.cpu().numpy()
    rng = np.random.default_rng(seed + 1)
    worst = 0.0
    for e in rng.choice(n, min(parity_n, n), replace=False):
        w, o = int(W[e]), int(off[e])
        D = np.zeros((5, w), dtype=np.complex128)
        M = np.zeros((5, w), dtype=np.complex128)
        D[:3] = r["data"][3 * o:3 * (o + w)].cpu().numpy().reshape(3, w)
        M[:3] = r["model"][3 * o:3 * (o + w)].cpu().numpy().reshape(3, w)
        ic = r["invC"][6 * o:6 * (o + w)].cpu().numpy().reshape(6, w)
        I9 = np.stack([ic[0], ic[3], ic[4], ic[3], ic[1], ic[5], ic[4], ic[5], ic[2]]).reshape(3, 3, w)
        ref = orc.gaussian_log_likelihood(D, M, I9, 3)
        worst = max(worst, abs(got[e] - ref) / abs(ref))
    B.close()
    del r
    torch.cuda.empty_cache()
    return dict(n=n, bins=bins, ms=ms, bytes=144.0 * bins, max_rel=worst, checked=min(parity_n, n))


def noise_data(orc_or_none, fstar, N, T, fmin, seed):
    """Synthetic data for the noise leg: the instrument model at the
    reference's default link levels plus a smooth foreground-like covariance
    (kept fixed, as the foreground model is while the instrument model is
    sampled), drawn per bin from the 3x3 covariance."""
    from glass_amd import noise
    rng = np.random.default_rng(seed)
    f = fmin + np.arange(N) / T
    sa, so = noise.default_links()
    C = noise.instrument_noise_model(fstar, f, sa, so)
    g = 0.5 * C[0, 0] * (1e-3 / f) ** (7.0 / 3.0) * np.exp(-f / 2e-3)
    Cadd = np.zeros_like(C)
    for a in range(3):
        for b in range(3):
            Cadd[a, b] = g * (1.0 if a == b else -0.5)
    L = np.linalg.cholesky(np.moveaxis(C + Cadd, -1, 0))
    w = (rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))) / np.sqrt(2)
    d = np.einsum("nab,nb->na", L, w)
    return f, d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(), np.concatenate([sa, so]), Cadd


def noise_leg(gl, steps=3, seed=31, pcalls=2, ladders=64):
    """Instrument-noise sampler on the device (glass_hip_noise_*): the band
    of the reference's noise_mcmc example (--duration 7864320 --fmin 1e-4
    --fmax 8e-3: 62,128 bins) with a fixed foreground covariance.  One sweep
    is the reference's outer iteration: per chain 10 calls of
    noise_instrument_model_mcmc (10 moves, one likelihood of every bin
    each), then noise_ptmcmc -- here one hipGraph.  Runs the reference's 12
    chains, and `ladders` ladders of 12 for throughput; times the fused
    likelihood kernel alone; checks sampled likelihoods and one sweep of 12
    chains (pcalls calls each) against the oracle, which is also the
    single-core CPU baseline."""
    import torch
    from glass_amd import noise
    from oracle import orc
    T, fmin, fmax, ntemp = 7864320.0, 1e-4, 8e-3, 12
    N = int(round((fmax - fmin) * T))
    fstar = gl.analytic_orbit().fstar
    f, X, Y, Z, p0, Cadd = noise_data(None, fstar, N, T, fmin, seed)
    rng = np.random.default_rng(seed + 1)
    temps1 = 1.2 ** np.arange(ntemp)
    res = dict(N=N, ntemp=ntemp)
    od = orc.NoiseData(fstar, f, X, Y, Z, Cadd)
    for nl in (1, ladders):
        S = nl * ntemp
        plan = noise.NoisePlan(fstar, f, X, Y, Z, nladder=nl, ntemp=ntemp, Cadd=Cadd)
        P = np.stack([p0 * rng.uniform(0.9, 1.1, 12) for _ in range(S)])
        P[:, 6:] = np.repeat(P[:, 6::2], 2, axis=1)
        seeds = rng.integers(1, 2 ** 32, S, dtype=np.uint64).astype(np.uint32)
        temps = np.tile(temps1, nl)
        plan.set_state(P, seeds, temps)
        if nl == 1:
            # parity: one sweep of the reference's loop against the oracle
            st0 = plan.state()
            plan.mcmc(pcalls, swap=True, graph=False)
            st = plan.state()
            x, lg = st0["params"].copy(), st0["logL"].copy()
            sd, idx = st0["seeds"].astype(np.int64), st0["index"].copy()
            nacc = np.zeros(S, dtype=np.int64)
            t0 = time.perf_counter()
            for s_ in range(S):
                for _ in range(pcalls):
                    x[idx[s_]], lg[idx[s_]], sd[s_], na = od.mcmc_call(x[idx[s_]], lg[idx[s_]], sd[s_], temps[s_])
                    nacc[s_] += na
            cpu_s = time.perf_counter() - t0
            i2, s2, _ = orc.noise_ptmcmc(idx, temps, lg, sd)
            same = (np.array_equal(st["nacc"], nacc) and np.array_equal(st["index"], i2) and
                    np.array_equal(st["seeds"], s2.astype(np.uint32)))
            prel = float(np.max(np.abs(st["params"] - x) / np.abs(x)))
            lrel = float(np.max(np.abs(st["logL"] - lg) / np.abs(lg)))
            res["parity"] = {"chains": S, "moves_per_chain": 10 * pcalls, "same_decisions": bool(same),
                             "max_param_rel_err": prel, "max_logL_rel_err": lrel,
                             "gate": "same accept/swap decisions; params 1e-12, logL 1e-10 relative",
                             "pass": bool(same and prel <= 1e-12 and lrel <= 1e-10)}
            res["cpu"] = {"rate": S * 10 * pcalls / cpu_s, "secs": cpu_s, "evals": S * 10 * pcalls}
        plan.mcmc(10, swap=True, graph=True)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            plan.mcmc(10, swap=True, graph=True)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # the fused likelihood kernel alone, all slots at once
        dP = torch.tensor(P, device="cuda")
        out = plan.logl(dP)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(20):
            plan.logl(dP, out=out)
        e1.record()
        torch.cuda.synchronize()
        kms = e0.elapsed_time(e1) / 20
        st = plan.state()
        moves = 100 * (steps + 1) + (10 * pcalls if nl == 1 else 0)
        res[nl] = dict(slots=S, ms_sweep=ms, logl_ms=kms, info=plan.info(),
                       acc_rate=float(st["nacc"].sum() / (S * moves)))
        if nl == ladders:
            smp = rng.choice(S, 8, replace=False)
            got = out.cpu().numpy()[smp]
            ref = np.array([od.logl(P[i]) for i in smp])
            res["logl_parity"] = float(np.max(np.abs(got - ref) / np.abs(ref)))
        plan.close()
        del plan
    return res


def formats_leg(plan, cfg):
    """The reference's checkpoint (save_chain_state / restore_chain_state,
    ucb/src/glass_ucb_io.c:418-490) of every chain of the batched plan: the
    device state copied and written as chain_state_<c>.dat, then every file
    read back; the read-back parameters must print as the plan's do."""
    import shutil
    import tempfile
    import torch
    from glass_amd import files
    d = tempfile.mkdtemp(prefix="glass_ckpt_")
    try:
        logl, sp = plan.state()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        files.save_chain_states(plan, d, step=1, logLmax=float(logl.max()), t0=0.0, eta=[1.0, 1.0, 1.0])
        t_write = time.perf_counter() - t0
        nchain = len(cfg["chain_group"])
        t0 = time.perf_counter()
        back = [files.restore_chain_state(os.path.join(d, f"chain_state_{c}.dat"), cfg["T"], 3,
                                          maxsrc=sp.shape[1]) for c in range(nchain)]
        t_read = time.perf_counter() - t0
        nsrc = int(sum(b["nlive"] for b in back))
        # values kept to the file's precision: f0 to 16 digits (then scaled by
        # T), the rest to 12 (A through ln A)
        T = cfg["T"]
        worst = 0.0
        for c, b in enumerate(back):
            n = b["nlive"]
            if n == 0:
                continue
            a, r = sp[c, :n], b["params"]
            worst = max(worst, float(np.max(np.abs(r[:, 0] - a[:, 0]) / np.abs(a[:, 0])) / 1e-15))
            for k in (1, 2, 4, 5, 6, 7):
                den = np.maximum(np.abs(a[:, k]), 1e-300)
                worst = max(worst, float(np.max(np.abs(r[:, k] - a[:, k]) / den)) / 1e-11)
            worst = max(worst, float(np.max(np.abs(np.exp(r[:, 3]) / np.exp(a[:, 3]) - 1.0))) / 1e-11)
        if not worst <= 1.0:
            raise SystemExit("formats: checkpoint read-back differs from the plan's state beyond the printed "
                             "precision (worst %.3g of the bound)" % worst)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"chains": nchain, "sources": nsrc, "write_s": t_write, "read_s": t_read,
            "sources_per_s": 2 * nsrc / (t_write + t_read),
            "parity": {"worst_over_bound": round(worst, 4), "bound": "f0 1e-15, A and the other parameters 1e-11 "
                       "relative (16 and 12 printed digits)"},
            "config": "save_chain_state / restore_chain_state of the batched case's chains, one file per chain, "
                      "host formatting after one device-to-host copy"}


def sampler_leg(plan, cfg, nladder, ntemp, nsweep=3, nsteps=80, seed=41):
    """The UCB sampler on the device over the batched plan's chains
    (glass_hip_ucb_sampler_*): one iteration of the reference's outer loop
    (apps/src/ucb_mcmc.c:230-265) = per chain 10 * NP = 80 ucb_mcmc steps
    (one proposal, waveform, likelihood and accept each), the Fisher update
    of every live source (exact derivatives + Jacobi eigen-decomposition),
    ptmcmc and adapt_temperature_ladder -- one hipGraph.  Default proposal
    weights without the F-statistic / GMM / covariance proposals, window
    priors, snr cut 5."""
    import torch
    from glass_amd.sampler import UcbSampler
    G = len(cfg["group_W"])
    pr = np.zeros((G, 8, 2))
    for g in range(G):
        q, W = float(cfg["group_qmin"][g]), float(cfg["group_W"][g])
        pr[g] = [[q, q + W], [-1, 1], [0, 2 * np.pi], [-60, -40], [-1, 1], [0, np.pi], [0, 2 * np.pi], [-1, 1]]
    smp = UcbSampler(plan, nladder, ntemp, pr)
    rng = np.random.default_rng(seed)
    C = nladder * ntemp
    smp.set_state(rng.integers(1, 2 ** 32, C, dtype=np.uint64).astype(np.uint32), np.tile(1.2 ** np.arange(ntemp),
                                                                                           nladder))
    smp.fisher()
    t0 = time.perf_counter()
    smp.sweep(nsteps, adapt=True, graph=True)  # capture + first replay
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(nsweep):
        smp.sweep(nsteps, adapt=True, graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / nsweep
    # the parts: steps alone, Fisher update alone
    e0.record()
    smp.steps(nsteps)
    e1.record()
    torch.cuda.synchronize()
    ms_steps = e0.elapsed_time(e1)
    e0.record()
    smp.fisher()
    e1.record()
    torch.cuda.synchronize()
    ms_fisher = e0.elapsed_time(e1)
    st = smp.state()
    smp.close()
    tr, ac = st["trials"].sum(axis=0), st["accepts"].sum(axis=0)
    return dict(chains=C, nsteps=nsteps, ms=ms, ms_steps_eager=ms_steps, ms_fisher=ms_fisher, first_s=t_first,
                live=int(np.asarray(cfg["chain_nsrc"]).sum()),
                acceptance={k: round(float(ac[i] / max(tr[i], 1)), 4) for i, k in
                            ((0, "uniform_prior"), (4, "fisher"), (5, "fm_shift"), (6, "psi_phi"))},
                swap_rate=round(float(st["swaps"].sum() / max(1, (nsweep + 2) * nladder * (ntemp - 1))), 4))


def fstat_leg(gl, seed=51, NFFT=1024, fq=3e-3, cpu_cells=200):
    """The F-statistic proposal density of one data segment
    (glass_hip_fstat_proposal, build_fstatistic_proposal of
    ucb/src/glass_ucb_proposal.c:1086-1180): NFFT = 1024 bins at 3 mHz,
    Tobs 2 yr, three binaries in noise; n_f = 1024 frequency cells x 20 x 20
    sky cells, each an F-statistic of four filter templates, plus the
    chirp-mass scan of fdot (10 more) for the cells with logL > 1.  The
    oracle's get_Fstat_logL on a sample of the cells is the CPU baseline."""
    import torch
    from glass_amd.fstat import Fstat
    from oracle import orc
    T = 62914560.0
    o = orc.analytic_orbit()
    rng = np.random.default_rng(seed)
    qmin0 = int(fq * T)
    C, _, _ = orc.xyz_noise(T, qmin0 / T, NFFT, orbit=o)
    C00 = C[0, 0].copy()
    X, A, E = (np.zeros(NFFT, dtype=np.complex128) for _ in range(3))
    for _ in range(3):
        f0 = (qmin0 + rng.uniform(100, NFFT - 100)) / T
        theta, phi = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        M = orc.ucb_bandwidth(f0, 0.0, np.cos(theta), 1e-22, T, NFFT, orbit=o)
        p = np.array([f0 * T, np.cos(theta), phi, np.log(2e-22), rng.uniform(-1, 1), rng.uniform(0, np.pi),
                      rng.uniform(0, 2 * np.pi), 0.0])
        h = orc.ucb_waveform(p, T, M, NI=2, orbit=o)
        q = int(p[0])
        for i in range(M):
            k = q - qmin0 + i - M // 2
            if 0 < k < NFFT:
                X[k] += h[0][i]
                A[k] += h[3][i]
                E[k] += h[4][i]
    s = np.sqrt(C00 / 2)
    for ch in (X, A, E):
        ch += s * (rng.standard_normal(NFFT) + 1j * rng.standard_normal(NFFT))
    fmin, fmax = (qmin0 + 1) / T, (qmin0 + NFFT - 1) / T
    F = Fstat(gl.analytic_orbit(), "sangria", T, 0.0, qmin0, X, A, E, C00, max_cells=1 << 19)
    F.proposal(fmin, fmin + 8 / T)  # warm up (kernels, tables)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ten, norm, maxp, nf, df = F.proposal(fmin, fmax)
    secs = time.perf_counter() - t0
    ncell = ten.size
    # the fdot = 0 pass alone, as device time
    cells = np.zeros((ncell, 4))
    d_theta, d_phi = 2.0 / 20, 6.283185307179586 / 20
    ii, jj, kk = np.meshgrid(np.arange(nf), np.arange(20), np.arange(20), indexing="ij")
    cells[:, 0] = (fmin + ii.reshape(-1) * df) + df / 2
    cells[:, 2] = np.arccos(-1.0 + jj.reshape(-1) * d_theta + d_theta / 2)
    cells[:, 3] = kk.reshape(-1) * d_phi + d_phi / 2
    dc = torch.tensor(cells[:1 << 19], device="cuda")
    F.eval(dc)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    lx, la, fp = F.eval(dc)
    e1.record()
    torch.cuda.synchronize()
    ms_eval = e0.elapsed_time(e1)
    n_eval = dc.shape[0]
    # parity on a sample of cells, and the CPU baseline
    fd = orc.FstatData(T, qmin0, X, A, E, C00, orbit=o)
    smp = rng.choice(n_eval, cpu_cells, replace=False)
    la_h = la.cpu().numpy()
    t1 = time.perf_counter()
    ref = [fd.logl(*cells[c])[1] for c in smp]
    cpu_s = time.perf_counter() - t1
    worst = float(np.max(np.abs(la_h[smp] - np.array(ref)) / np.abs(ref)))
    F.close()
    return dict(ncell=ncell, nf=nf, secs=secs, maxp=maxp, ms_eval=ms_eval, n_eval=n_eval, max_rel=worst,
                cpu_rate=cpu_cells / cpu_s, cpu_s=cpu_s, cpu_cells=cpu_cells)


def globalfit_leg(gl, synth, steps=3, seed=4, cpu_seconds=2.0, nthreads=0, serial=False):
    """Global-fit UCB block at full scale on one GPU: 20,000 binaries (f0
    log-uniform 0.1-25 mHz, A log-uniform 1e-24..1e-21), Tobs 4 yr, the
    band in 4,000 density-weighted segments of 64-2048 bins, 8 temperatures
    x 32 chains per segment (1,024,000 chains).  One sweep = one proposal
    per chain with sources (waveform + delta-logL) and the exact gradient
    and Fisher matrix of one source on every 8th of those chains, replayed
    as one hipGraph, tiles from the autotuner."""
    import torch
    from glass_amd.batch import FdBatch, ucb_alignment
    o = gl.analytic_orbit()
    t0 = time.perf_counter()
    cfg = synth.config4(seed=seed, waveform=lambda p, T, BW: gl.ucb_waveform(o, "sangria", T, 0.0, p, 8, BW),
                        alignment=lambda p, T, W, q: ucb_alignment(o, T, W, q, p))
    t_gen = time.perf_counter() - t0
    E = len(cfg["p_chain"])
    free0 = torch.cuda.mem_get_info()[0]
    t0 = time.perf_counter()
    plan = FdBatch(o, "sangria", cfg["T"], 0.0, (cfg["group_qmin"], cfg["group_W"]), cfg["data_flat"],
                   cfg["invC_flat"], cfg["chain_group"], cfg["chain_nsrc"], cfg["src_params"], E, flat=True)
    if serial:
        plan.set_concurrent(False)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    tb, total = plan.memory()
    log(f"globalfit: inputs {t_gen:.1f} s, plan {t_plan:.1f} s, {total / 1e9:.1f} GB on the device "
        f"({tb / 1e9:.1f} GB templates), {len(cfg['chain_group'])} chains, {E} proposals per sweep")
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    t0 = time.perf_counter()
    tiles, tms = plan.autotune(E, reps=2)
    t_tune = time.perf_counter() - t0
    dch = torch.tensor(cfg["f_chain"], device="cuda")
    dsr = torch.tensor(cfg["f_src"], device="cuda")
    nF = dch.numel()
    dL = torch.empty(nF, dtype=torch.float64, device="cuda")
    dg = torch.empty((nF, 8), dtype=torch.float64, device="cuda")
    dF = torch.empty((nF, 8, 8), dtype=torch.float64, device="cuda")
    plan.sweep_graph(E, dch, dsr, dL, dg, dF)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan.sweep_graph(E, dch, dsr, dL, dg, dF)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    logL = plan.logL[:E].cpu().numpy()
    nan_f = int(torch.isnan(dF.reshape(nF, -1)[:, 0]).sum().item())
    # breakdown
    plan.eval(E, graph=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        plan.eval(E, graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms_eval = e0.elapsed_time(e1) / steps
    e0.record()
    for _ in range(steps):
        plan.derivs(dch, dsr)
    e1.record()
    torch.cuda.synchronize()
    ms_der = e0.elapsed_time(e1) / steps
    # per-class kernel times of the eval part (HIP events around each class
    # kernel on its launch stream, the classes in turn): the roofline of the
    # dominant class kernel, and the algorithmic FP64 rate of all classes
    plan.profile(True)
    for _ in range(steps):
        plan.eval(E)
    kms, klaunch, kcount = plan.profile_read()
    plan.profile(False)
    torch.cuda.synchronize()
    # dominant class: the largest share of the algorithmic FP64 work, as in
    # the sweep's roofline; the longest class beside it
    work = np.array([float(kcount[k]) * wave_flops(32 << k) if klaunch[k] else 0.0 for k in range(8)])
    kd = int(np.argmax(work))
    kt = int(np.argmax(kms))
    Nd = 32 << kd
    avg_ms = kms[kd] / klaunch[kd]
    ach_tf = kcount[kd] * wave_flops(Nd) / (avg_ms * 1e-3) / 1e12
    ach_gbs = kcount[kd] * eval_bytes(Nd) / (avg_ms * 1e-3) / 1e9
    t = tiles[kd]
    sweep_flops = sum(float(kcount[k]) * wave_flops(32 << k) for k in range(8) if klaunch[k])
    roofline = {"bound": "fp64-valu", "kernel": f"eval_kernel<{Nd},{t[0]},{'true' if t[1] else 'false'},{t[2]}>",
                "achieved": round(ach_tf, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach_tf / FP64_PEAK_TFLOPS, 5), "traffic": None,
                "avg_launch_ms": round(avg_ms, 4), "evals_per_launch": int(kcount[kd]),
                "hbm_algorithmic_GBs": round(ach_gbs, 2), "hbm_frac": round(ach_gbs / HBM_PEAK_GBS, 5),
                "classes": {str(32 << k): {"evals": int(kcount[k]), "ms_per_launch": round(kms[k] / klaunch[k], 4)}
                            for k in range(8) if klaunch[k]},
                "longest_class": {"N": 32 << kt, "ms_per_launch": round(kms[kt] / klaunch[kt], 4),
                                  "frac": round(kcount[kt] * wave_flops(32 << kt) / (kms[kt] / klaunch[kt] * 1e-3)
                                                / 1e12 / FP64_PEAK_TFLOPS, 5)},
                "eval_sweep_achieved": round(sweep_flops / (ms_eval * 1e-3) / 1e12, 4),
                "eval_sweep_frac": round(sweep_flops / (ms_eval * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 5),
                "note": "dominant = the class with the largest algorithmic FP64 work; "
                        "achieved = evals_per_launch x wave_flops(N) / the class kernel's launch duration; "
                        "eval_sweep_* = the algorithmic FP64 of every class / the eval-only graph sweep time"}
    log(f"globalfit: dominant eval class (work) BW {Nd} {avg_ms:.3f} ms/launch, {ach_tf:.2f} TFLOP/s "
        f"({ach_tf / FP64_PEAK_TFLOPS:.3f} of FP64 peak); all classes {roofline['eval_sweep_achieved']:.2f} TFLOP/s")
    res = dict(ms=ms, ms_eval=ms_eval, ms_derivs=ms_der, E=E, nF=nF, nan_fisher=nan_f, chains=len(cfg["chain_group"]),
               roofline=roofline,
               segments=len(cfg["group_W"]), bins=int(cfg["group_W"].sum()), max_src=cfg["max_src"],
               device_GB=total / 1e9, template_GB=tb / 1e9, free_GB_before=free0 / 1e9, t_gen=t_gen,
               t_plan=t_plan, t_tune=t_tune,
               tiles={str(32 << k): list(t) for k, t in enumerate(tiles) if tms[k] > 0})
    plan.close()
    del plan
    torch.cuda.empty_cache()
    try:
        leg = oracle_leg(cfg, logL, cpu_seconds, nthreads, seed + 5)
        ok = leg["max_scaled"] <= 1e-12 and (leg["max_rel_above_05mHz"] or 0.0) <= 1e-10
        res["parity"] = {"checked": leg["n"], "max_logL_rel_err": leg["max_rel"],
                         "max_logL_rel_err_above_0.5mHz": leg["max_rel_above_05mHz"],
                         "max_err_over_chi2_scale": leg["max_scaled"],
                         "gate": "1e-10 relative above 0.5 mHz; 1e-12 of sum|d||d||invC| everywhere "
                                 "(the XYZ covariance is nearly singular towards 0.1 mHz)",
                         "pass": bool(ok)}
        res["cpu"] = leg
    except Exception as exc:  # noqa: BLE001
        log(f"globalfit oracle check unavailable: {exc}")
    return res


def oracle_leg(cfg, logL_gpu, seconds, nthreads, seed, variants=False):
    """CPU baseline (the oracle restatement of the reference, OpenMP over
    proposals) on a bounded sample of the same workload, and the parity of
    the GPU results for those proposals."""
    import ctypes
    from oracle import orc
    orc.lib()
    rng = np.random.default_rng(seed)
    C = len(cfg["chain_group"])
    chains = rng.choice(C, size=min(48, C), replace=False)
    T = cfg["T"]
    bwmax = 4096
    structs, local = [], {}
    for j, c in enumerate(chains):
        g = cfg["chain_group"][c]
        W, qmin = int(cfg["group_W"][g]), int(cfg["group_qmin"][g])
        n = int(cfg["chain_nsrc"][c])
        tm = np.zeros((n, 3, 2 * bwmax))
        bw = np.zeros(n, np.int32)
        im = np.zeros(n, np.int32)
        for s in range(n):
            b, _, i = orc.ucb_alignment(cfg["src_params"][c, s], T, W, qmin)
            tm[s, :, :2 * b] = orc.ucb_waveform(cfg["src_params"][c, s], T, b)[:3].view(np.float64).reshape(3, -1)
            bw[s], im[s] = b, i
        structs.append(dict(nsrc=n, W=W, qmin=qmin, bwmax=bwmax,
                            data=np.ascontiguousarray(cfg["data"][g]).view(np.float64).reshape(-1).copy(),
                            invC=np.ascontiguousarray(cfg["invC"][g]).reshape(-1).copy(),
                            src_params=cfg["src_params"][c, :n].copy(), tmpl=tm.reshape(-1).copy(),
                            src_bw=bw, src_imin=im))
        local[int(c)] = j
    cs = orc.ChainSet(structs)
    sel = np.nonzero(np.isin(cfg["p_chain"], chains))[0]
    rng.shuffle(sel)
    ch = np.array([local[int(c)] for c in cfg["p_chain"][sel]], np.int32)
    # one pass over the sample (also the parity check), then repeat it to
    # about `seconds` of CPU work in total
    n = len(sel)
    out, secs = orc.eval_batch(cs, T, ch, cfg["p_src"][sel], cfg["p_params"][sel], nthreads=nthreads)
    rel = np.abs(logL_gpu[sel] - out) / np.abs(out)
    # rounding scale of chi^2 in each window: sum |d_a| |d_b| |invC_ab| (the XYZ
    # covariance is nearly singular towards 0.1 mHz, where chi^2 cancels)
    scale = {}
    for c in chains:
        g = int(cfg["chain_group"][c])
        if g not in scale:
            d, ic = np.abs(np.asarray(cfg["data"][g])), np.abs(np.asarray(cfg["invC"][g]))
            scale[g] = float(sum(np.sum(d[a] * d[b] * ic[a, b]) for a in range(3) for b in range(3)))
    grp = cfg["chain_group"][cfg["p_chain"][sel]]
    scaled = np.abs(logL_gpu[sel] - out) / np.array([scale[int(g)] for g in grp])
    hi = cfg["group_qmin"][grp] / T > 5e-4
    reps = max(0, int(round(seconds / max(secs, 1e-3))) - 1)
    for _ in range(reps):
        _, s = orc.eval_batch(cs, T, ch, cfg["p_src"][sel], cfg["p_params"][sel], nthreads=nthreads)
        secs += s
    res = dict(rate=n * (reps + 1) / secs, n=n, reps=reps + 1, secs=secs, nchains=len(chains),
               max_rel=float(rel.max()), max_scaled=float(scaled.max()),
               max_rel_above_05mHz=float(rel[hi].max()) if hi.any() else None)
    if variants:
        # the same sample at the reference's release flags, and both builds
        # on one core, each for about a quarter of `seconds`
        res["variants"] = {"parity_flags": {str(nthreads): round(res["rate"], 2)}, "release_flags": {}}
        for flags, release in (("release_flags", True), ("parity_flags", False)):
            for nt in ((nthreads, 1) if release else (1,)):
                k, tot = 0, 0.0
                while tot < seconds / 4 or k == 0:
                    _, s = orc.eval_batch(cs, T, ch, cfg["p_src"][sel], cfg["p_params"][sel], nthreads=nt,
                                          release=release)
                    tot += s
                    k += 1
                res["variants"][flags][str(nt)] = round(n * k / tot, 2)
    return res


def cpu_host():
    """nproc, the cores this process may use and the CPU model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--groups", type=int, default=256, help="data windows (256 = the full case)")
    ap.add_argument("--temps", type=int, default=16)
    ap.add_argument("--props", type=int, default=80, help="proposals per chain per sweep")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fisher", action="store_true")
    ap.add_argument("--no-wdm", action="store_true")
    ap.add_argument("--no-inner", action="store_true")
    ap.add_argument("--no-noise", action="store_true")
    ap.add_argument("--no-sampler", action="store_true")
    ap.add_argument("--no-fstat", action="store_true")
    ap.add_argument("--no-formats", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tiles (skip the autotuner)")
    ap.add_argument("--serial-classes", action="store_true",
                    help="run the BW-class kernels of each sweep in turn on one stream (profiling runs)")
    ap.add_argument("--globalfit", action="store_true",
                    help="run the global-fit block (the default; kept for older command lines)")
    ap.add_argument("--no-globalfit", action="store_true",
                    help="skip the global-fit block at full scale (1,024,000 chains, ~170 GB of HBM)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run a 256 MiB device copy, the reference point that scales the "
                         "FETCH_SIZE/WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU "
              "path; run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import glass_amd as gl
    from glass_amd import synth
    from glass_amd.batch import FdBatch

    gl.lib()
    t0 = time.perf_counter()
    o, cfg = build_inputs(gl, synth, args.groups, args.temps, args.props, args.seed)
    E = len(cfg["p_chain"])
    plan = FdBatch(o, "sangria", cfg["T"], 0.0, cfg["group_qmin"], cfg["data"], cfg["invC"],
                   cfg["chain_group"], cfg["chain_nsrc"], cfg["src_params"], E)
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    if args.serial_classes:
        plan.set_concurrent(False)
    torch.cuda.synchronize()
    log(f"inputs + plan: {time.perf_counter() - t0:.1f} s, {E} evaluations per step")

    tune = None
    if not args.no_tune:
        t1 = time.perf_counter()
        tiles, tms = plan.autotune(E, reps=3)
        tune = {"tiles": {str(32 << k): {"threads": t[0], "grid_mult": t[1], "waves": t[2],
                                         "ms": round(float(tms[k]), 4)}
                          for k, t in enumerate(tiles) if tms[k] > 0},
                "seconds": round(time.perf_counter() - t1, 2)}
        log(f"autotune {tune['seconds']} s: " + ", ".join(f"BW {32 << k}: {t[0]}x{t[1]}w{t[2]}"
                                                          for k, t in enumerate(tiles) if tms[k] > 0))
    tiles_now = plan.tiles()

    for _ in range(args.warmup):
        plan.eval(E, graph=True)
    torch.cuda.synchronize()
    # graph replay (the production path), for reference
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g0.record()
    for _ in range(max(1, min(args.steps, 5))):
        plan.eval(E, graph=True)
    g1.record()
    torch.cuda.synchronize()
    graph_ms = g0.elapsed_time(g1) / max(1, min(args.steps, 5))

    # timed region: K eager sweeps as a caller issues them (BW classes on
    # concurrent streams unless --serial-classes)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        plan.eval(E)
    e1.record()
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    ms_step = total_ms / args.steps
    value = E * args.steps / (total_ms * 1e-3)
    # per-class kernel times (the roofline's launch duration): the same K
    # sweeps again with HIP events around each class kernel on the launch
    # stream, the classes in turn
    plan.profile(True)
    for _ in range(args.steps):
        plan.eval(E)
    kms, klaunch, kcount = plan.profile_read()
    plan.profile(False)
    logL = plan.logL[:E].cpu().numpy()
    log(f"sweep {ms_step:.3f} ms ({value:.4e} evals/s), graph replay {graph_ms:.3f} ms")
    for k in range(8):
        if klaunch[k]:
            log(f"  BW {32 << k:5d}: {kcount[k]:7d} evals, {kms[k] / klaunch[k]:8.3f} ms/launch")

    # the dominant class: the largest share of the sweep's algorithmic FP64
    # work (stable run to run; the longest class is reported beside it)
    work = np.array([float(kcount[k]) * wave_flops(32 << k) if klaunch[k] else 0.0 for k in range(8)])
    kd = int(np.argmax(work))
    kt = int(np.argmax(kms))
    Nd = 32 << kd
    avg_ms = kms[kd] / klaunch[kd]
    flops = kcount[kd] * wave_flops(Nd)
    byts = kcount[kd] * eval_bytes(Nd)
    ach_tf = flops / (avg_ms * 1e-3) / 1e12
    ach_gbs = byts / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "fp64-valu", "pipe": "FP64 vector ALU (no MFMA: nothing on the path is a dense contraction)",
                "kernel": f"eval_kernel<{Nd},{tiles_now[kd][0]},{'true' if tiles_now[kd][1] else 'false'},"
                          f"{tiles_now[kd][2]}>",
                "achieved": round(ach_tf, 4),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach_tf / FP64_PEAK_TFLOPS, 5),
                "traffic": MEASURED_TRAFFIC.get(Nd),
                "traffic_source": "constant, not measured in this run: profiles/eval_kernel_batched_fd_sweep.md "
                                  "(rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of this workload, calibrated "
                                  "on a 256 MiB copy); bytes per launch (L2/MALL reuse keeps it below the "
                                  "algorithmic bytes)",
                "avg_launch_ms": round(avg_ms, 4), "evals_per_launch": int(kcount[kd]),
                "hbm_algorithmic_GBs": round(ach_gbs, 2), "hbm_frac": round(ach_gbs / HBM_PEAK_GBS, 5),
                "dominant": "the class kernel with the largest share of the sweep's algorithmic FP64 work "
                            "(evaluations x wave_flops(N)); longest_class: the one with the longest time in "
                            "the per-class timing pass",
                "longest_class": {"kernel": f"eval_kernel<{32 << kt},{tiles_now[kt][0]},"
                                            f"{'true' if tiles_now[kt][1] else 'false'},{tiles_now[kt][2]}>",
                                  "ms_per_launch": round(kms[kt] / klaunch[kt], 4),
                                  "frac": round(kcount[kt] * wave_flops(32 << kt) / (kms[kt] / klaunch[kt] * 1e-3)
                                                / 1e12 / FP64_PEAK_TFLOPS, 5)},
                "classes": {str(32 << k): {
                    "kernel": f"eval_kernel<{32 << k},{tiles_now[k][0]},{'true' if tiles_now[k][1] else 'false'},"
                              f"{tiles_now[k][2]}>",
                    "evals": int(kcount[k]), "ms_per_launch": round(kms[k] / klaunch[k], 4),
                    "achieved": round(kcount[k] * wave_flops(32 << k) / (kms[k] / klaunch[k] * 1e-3) / 1e12, 4),
                    "frac": round(kcount[k] * wave_flops(32 << k) / (kms[k] / klaunch[k] * 1e-3) / 1e12
                                  / FP64_PEAK_TFLOPS, 5)}
                    for k in range(8) if klaunch[k] and kcount[k]}}
    log(f"roofline: {roofline['kernel']} {avg_ms:.3f} ms/launch, {ach_tf:.2f} TFLOP/s = {roofline['frac']:.3f} of "
        f"FP64 peak (largest algorithmic work); longest class {roofline['longest_class']['kernel']} "
        f"{roofline['longest_class']['ms_per_launch']:.3f} ms, {roofline['longest_class']['frac']:.3f}")

    if args.copy_calibrate:
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 256 MiB read + 256 MiB written")
        del a, b

    plain = plain0 = None
    if not args.no_plain:
        nthr_p = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        plain = plain_leg(gl, cfg, logL, nthreads=nthr_p)
        plain0 = plain_rate(gl, synth)
        log(f"plain per-call drop-ins on this sweep's proposals: {plain['rate1']:.1f} evals/s from one host "
            f"thread, {plain['rateN']:.1f} from {nthr_p}; config-0 case {plain0:.1f} evals/s")

    fisher = None
    if not args.no_fisher:
        fl = fisher_leg(gl, synth)
        fisher = {"value": round(fl["value"], 1), "unit": "exact-gradient Fisher evals/s",
                  "ms_per_launch": round(fl["ms"], 4),
                  "plain_fd_fisher": round(fl["plain"], 1),
                  "vs_plain": round(fl["value"] / fl["plain"], 1),
                  "config": "8192 sources, BW 256, Tobs 4 yr, own 512-bin window each; one launch -> logL, "
                            "dlogL/dtheta[8], Fisher[8x8] per source; plain = per-call forward-difference "
                            "ucb_fisher drop-in"}
        fisher["adjoint_gradient"] = {
            "value": round(fl["n"] / (fl["ms_grad"] * 1e-3), 1), "unit": "logL + dlogL/dtheta[8] evals/s",
            "ms_per_launch": round(fl["ms_grad"], 4), "ms_logL_only": round(fl["ms_L"], 4),
            "vs_forward_mode": round(fl["ms"] / fl["ms_grad"], 2),
            "parity_vs_forward_mode": {"max_grad_err_over_row_max": fl["grad_err"], "max_logL_rel_err": fl["L_err"],
                                       "gate": {"grad": 1e-9, "logL": 1e-10},
                                       "pass": bool(fl["grad_err"] <= 1e-9 and fl["L_err"] <= 1e-10)},
            "method": "reverse mode (glass_hip_fd_source_grad): template forward, TDI and FFT transposed on the "
                      "residual-weighted adjoint, one contraction per parameter; same 8192 sources"}
        log(f"exact Fisher: {fl['value']:.4e} evals/s ({fl['ms']:.3f} ms / {fl['n']}), "
            f"plain FD ucb_fisher {fl['plain']:.1f}/s; adjoint gradient {fl['ms_grad']:.3f} ms "
            f"(logL only {fl['ms_L']:.3f} ms), grad err {fl['grad_err']:.2e}")

    wdm = None
    if not args.no_wdm:
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        wl = wdm_leg(gl, synth, cpu_seconds=0.0 if args.no_cpu_baseline else args.cpu_seconds / 2,
                     nthreads=nthr, plain=not args.no_plain)
        byts = 16.0 * wl["ND"]
        wdm = {"value": round(wl["value"], 1), "unit": "wavelet templates + wavelet_nwip evals/s",
               "ms_per_launch": round(wl["ms"], 4), "templates": wl["n"], "layers_histogram": wl["layers"],
               "plain_dropin": None if wl["plain"] is None else round(wl["plain"], 1),
               "vs_plain": None if wl["plain"] is None else round(wl["value"] / wl["plain"], 1),
               "transform": {"ND": wl["ND"], "ms": round(wl["transform_ms"], 4),
                             "timing": "graph replay of the three launches (device rate, median of 5 groups "
                                       "of 20 replays); ms_per_call: the same through the per-call API from Python",
                             "ms_per_call": round(wl["transform_ms_call"], 4),
                             "graph_replay_bitwise_equal": wl["transform_graph_equal"],
                             "algorithmic_GBs": round(byts / (wl["transform_ms"] * 1e-3) / 1e9, 1),
                             "hbm_frac": round(byts / (wl["transform_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "normwise_err_vs_oracle": wl.get("transform_normwise_err")},
               "fisher": {"value": round(wl["fisher"]["value"], 1), "unit": "wavelet FD Fisher evals/s",
                          "sources_per_launch": wl["fisher"]["n"], "ms_per_launch": round(wl["fisher"]["ms"], 4),
                          "plain_dropin": None if wl["fisher"]["plain"] is None else round(wl["fisher"]["plain"], 1),
                          "method": "ucb_fisher_wavelet central differences: 16 templates + 36-element "
                                    "nine-term Gram per source"},
               "parity": wl.get("parity"),
               "forced_layers": {"unit": "wavelet templates + wavelet_nwip evals/s", "by_layers": wl.get("forced_layers"),
                                 "method": "the first 1024 templates with every template forced to L frequency "
                                           "layers (glass_hip_wdm_batch_create_layers; the reference's "
                                           "ucb_wavelet_layers gives 1-3): NT (L + 1) heterodyned samples, "
                                           "L inverse transforms, NT L pixels per template"},
               "config": "Tobs 1 yr, NT 512 x NF 4096 (15 s cadence), 4096 templates (config-0 parameter "
                         "distributions), stationary XYZ pixel noise; plain = per-call ucb_waveform_wavelet "
                         "drop-in on host arrays"}
        if "cpu" in wl:
            wdm["cpu_baseline"] = {"value": round(wl["cpu"]["rate"], 1), "unit": "templates/s", "cores": nthr,
                                   "kind": "port",
                                   "sample": f"{wl['cpu']['n']} templates x {wl['cpu']['reps']} by the oracle "
                                             f"restatement, OpenMP over templates, {wl['cpu']['secs']:.1f} s"}
        log(f"wdm: {wl['value']:.4e} templates/s ({wl['ms']:.3f} ms / {wl['n']}), transform "
            f"{wl['transform_ms']:.3f} ms (graph; per call {wl['transform_ms_call']:.3f} ms, replays equal "
            f"{wl['transform_graph_equal']}), plain {wl['plain']}/s, parity {wl.get('parity')}; forced layers "
            + ", ".join(f"L={k}: {v['value']:.3e}/s" for k, v in (wl.get('forced_layers') or {}).items()))

    inner = None
    if not args.no_inner:
        il = inner_leg(gl, synth)
        gbs = il["bytes"] / (il["ms"] * 1e-3) / 1e9
        inner = {"value": round(il["n"] / (il["ms"] * 1e-3), 1), "unit": "full-window logL evals/s",
                 "ms_per_launch": round(il["ms"], 4), "items": il["n"], "bins": il["bins"],
                 "roofline": {"bound": "hbm", "kernel": "logl_chunk_kernel + logl_finish_kernel",
                              "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": INNER_TRAFFIC,
                              "algorithmic_bytes": il["bytes"]},
                 "parity": {"checked": il["checked"], "max_logL_rel_err": il["max_rel"], "gate": 1e-10,
                            "pass": bool(il["max_rel"] <= 1e-10)},
                 "config": "batched inner product: %d independent gaussian_log_likelihood items (3 channels), "
                           "windows 512-4096 bins at Tobs 2 yr, own data/model/invC each, 144 B/bin" % il["n"]}
        log(f"inner: {il['ms']:.4f} ms per launch of {il['n']} items / {il['bins']} bins, {gbs:.1f} GB/s, "
            f"parity {il['max_rel']:.3e}")

    nzl = None
    if not args.no_noise:
        nl = noise_leg(gl)
        big = nl[max(k for k in nl if isinstance(k, int))]
        one = nl[1]
        N = nl["N"]
        nzl = {"value": round(big["slots"] * 100 / (big["ms_sweep"] * 1e-3), 1),
               "unit": "noise-model likelihoods/s (each over all bins, inside the MCMC sweep)",
               "bins": N, "chains": big["slots"], "ms_per_sweep": round(big["ms_sweep"], 3),
               "reference_config": {"chains": one["slots"], "ms_per_sweep": round(one["ms_sweep"], 3),
                                    "likelihoods_per_s": round(one["slots"] * 100 / (one["ms_sweep"] * 1e-3), 1)},
               "fused_logl_kernel": {"ms_per_launch": round(big["logl_ms"], 4), "chains": big["slots"],
                                     "chain_bins_per_s": round(big["slots"] * N / (big["logl_ms"] * 1e-3), 1),
                                     "hbm_bytes": 128.0 * N * ((big["slots"] + 7) // 8),
                                     "plan": big["info"]},
               "acceptance": round(big["acc_rate"], 4),
               "parity": dict(nl["parity"], logl_sample_max_rel_err=nl.get("logl_parity")),
               "config": "instrument-noise MCMC, band 0.1-8 mHz at Tobs 7,864,320 s (62,128 bins) with a "
                         "fixed foreground covariance; sweep = per chain 10 x noise_instrument_model_mcmc "
                         "(10 moves each) + noise_ptmcmc, one hipGraph; 12 temperatures per ladder"}
        if "cpu" in nl:
            nzl["cpu_baseline"] = {"value": round(nl["cpu"]["rate"], 2), "unit": "likelihoods/s", "cores": 1,
                                   "kind": "port",
                                   "sample": f"the 12 reference chains, {nl['cpu']['evals']} moves, by "
                                             f"the oracle restatement, single-threaded, "
                                             f"{nl['cpu']['secs']:.1f} s"}
        log(f"noise: {nzl['value']:.4e} likelihoods/s at {big['slots']} chains ({big['ms_sweep']:.3f} ms/sweep); "
            f"12 chains {one['ms_sweep']:.3f} ms/sweep; fused kernel {big['logl_ms']:.4f} ms; "
            f"parity {nzl['parity']}")

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        leg = oracle_leg(cfg, logL, args.cpu_seconds / 2, nthr, args.seed + 17, variants=True)
        rel = leg["variants"]["release_flags"][str(nthr)]
        cpu = {"value": rel, "unit": "evals/s", "cores": nthr, "kind": "port",
               "flags": "release: -O3 -march=x86-64-v3, FP contraction on (the reference's -O3 -march=native, "
                        "CMakeLists.txt:37-39)",
               "sample": f"{leg['n']} proposals (all of {leg['nchains']} randomly chosen chains) of the "
                         f"same sweep, each evaluated by the oracle restatement of generate_ucb_model + "
                         f"gaussian_log_likelihood, OpenMP over proposals (the reference's parallelism over "
                         f"chains); about {args.cpu_seconds / 8:.1f} s per variant",
               "variants_evals_per_s": {"release_flags": leg["variants"]["release_flags"],
                                        "parity_flags (-O2 -ffp-contract=off)": leg["variants"]["parity_flags"]},
               "host": cpu_host()}
        parity = {"checked": leg["n"], "max_logL_rel_err": leg["max_rel"], "gate": 1e-10,
                  "pass": bool(leg["max_rel"] <= 1e-10)}
        log(f"cpu baseline {rel:.1f} evals/s on {nthr} threads (release flags), variants {leg['variants']}; "
            f"parity max rel {leg['max_rel']:.3e}")

    fst = None
    if not args.no_fstat:
        fl = fstat_leg(gl)
        fst = {"value": round(fl["n_eval"] / (fl["ms_eval"] * 1e-3), 1),
               "unit": "F-statistic cells/s (4 filter templates + 28 sums + two 4x4 inverses per cell)",
               "cells_per_launch": fl["n_eval"], "ms_per_launch": round(fl["ms_eval"], 3),
               "proposal_build_s": round(fl["secs"], 3), "proposal_cells": fl["ncell"], "frequency_cells": fl["nf"],
               "parity": {"checked": fl["cpu_cells"], "max_logL_AE_rel_err": fl["max_rel"], "gate": 1e-10,
                          "pass": bool(fl["max_rel"] <= 1e-10)},
               "cpu_baseline": {"value": round(fl["cpu_rate"], 1), "unit": "cells/s", "cores": 1, "kind": "port",
                                "sample": f"{fl['cpu_cells']} cells by the oracle's get_Fstat_logL, "
                                          f"single-threaded, {fl['cpu_s']:.2f} s"},
               "config": "F-statistic proposal of a 1024-bin segment at 3 mHz, Tobs 2 yr, three binaries in noise: "
                         "1022 x 20 x 20 cells at fdot = 0 + the chirp-mass fdot scan of cells with logL > 1 "
                         "(build_fstatistic_proposal)"}
        log(f"fstat: {fst['value']:.4e} cells/s ({fl['ms_eval']:.2f} ms / {fl['n_eval']} cells), proposal of "
            f"{fl['ncell']} cells in {fl['secs']:.3f} s, parity {fl['max_rel']:.2e}, oracle {fl['cpu_rate']:.0f} "
            f"cells/s")

    smpl = None
    if not args.no_sampler:
        sl = sampler_leg(plan, cfg, args.groups, args.temps)
        smpl = {"value": round(sl["chains"] * sl["nsteps"] / (sl["ms"] * 1e-3), 1),
                "unit": "ucb_mcmc steps/s (proposal + waveform + likelihood + accept per chain step)",
                "ms_per_iteration": round(sl["ms"], 3), "chains": sl["chains"], "steps_per_iteration": sl["nsteps"],
                "ms_steps_eager": round(sl["ms_steps_eager"], 3), "ms_fisher_update": round(sl["ms_fisher"], 3),
                "live_sources": sl["live"], "graph_capture_and_first_run_s": round(sl["first_s"], 2),
                "acceptance": sl["acceptance"], "pt_swap_rate": sl["swap_rate"],
                "config": "the batched case's %d windows x %d temperatures; per iteration 80 ucb_mcmc steps per "
                          "chain, exact Fisher + Jacobi eigen of every live source, ptmcmc + "
                          "adapt_temperature_ladder, one hipGraph; proposals: uniform prior 0.1, Fisher 0.7, "
                          "fm shift 0.1, psi-phi 0.1; snr cut 5" % (args.groups, args.temps)}
        log(f"sampler: {smpl['value']:.4e} steps/s, {sl['ms']:.2f} ms per iteration (steps eager "
            f"{sl['ms_steps_eager']:.2f} ms, Fisher update {sl['ms_fisher']:.2f} ms), acceptance {sl['acceptance']}")

    fmts = None
    if not args.no_formats:
        fmts = formats_leg(plan, cfg)
        log(f"formats: checkpoint of {fmts['chains']} chains / {fmts['sources']} sources written in "
            f"{fmts['write_s']:.3f} s, read back in {fmts['read_s']:.3f} s ({fmts['sources_per_s']:.3e} sources/s)")

    gfit = None
    if not args.no_globalfit:
        plan.close()
        torch.cuda.empty_cache()
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        gl_ = globalfit_leg(gl, synth, nthreads=nthr, serial=args.serial_classes)
        gfit = {"value": round(gl_["E"] / (gl_["ms"] * 1e-3), 1), "unit": "evals/s (sweep incl. Fisher items)",
                "ms_per_sweep": round(gl_["ms"], 3), "evals_per_sweep": gl_["E"], "fisher_items": gl_["nF"],
                "ms_eval_only": round(gl_["ms_eval"], 3), "ms_derivs_only": round(gl_["ms_derivs"], 3),
                "eval_rate": round(gl_["E"] / (gl_["ms_eval"] * 1e-3), 1),
                "fisher_rate": round(gl_["nF"] / (gl_["ms_derivs"] * 1e-3), 1),
                "fisher_items_nan": gl_["nan_fisher"],
                "chains": gl_["chains"], "segments": gl_["segments"], "bins": gl_["bins"],
                "max_sources_per_chain": gl_["max_src"], "device_GB": round(gl_["device_GB"], 1),
                "template_GB": round(gl_["template_GB"], 1), "tiles": gl_["tiles"],
                "setup_s": {"inputs": round(gl_["t_gen"], 1), "plan": round(gl_["t_plan"], 1),
                            "autotune": round(gl_["t_tune"], 1)},
                "parity": gl_.get("parity"), "roofline": gl_["roofline"],
                "config": "20,000 binaries (f0 log-uniform 0.1-25 mHz, A log-uniform 1e-24..1e-21), Tobs 4 yr, "
                          "4,000 density-weighted segments of 64-2048 bins, 8 temperatures x 32 chains per "
                          "segment; one proposal per chain + exact Fisher of one source on every 8th chain, "
                          "one hipGraph per sweep"}
        if "cpu" in gl_:
            gfit["cpu_baseline"] = {"value": round(gl_["cpu"]["rate"], 1), "unit": "evals/s", "cores": nthr,
                                    "kind": "port"}
        log(f"globalfit: sweep {gl_['ms']:.2f} ms (eval {gl_['ms_eval']:.2f}, derivs {gl_['ms_derivs']:.2f}), "
            f"{gfit['value']:.4e} evals/s, parity {gl_.get('parity')}")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched PTMCMC likelihood sweep: %d windows x %d temperatures, 1-8 sources "
                               "per chain, %d proposals per chain, Tobs 2 yr, windows 512-4096 bins" %
                               (args.groups, args.temps, args.props),
                   "evals_per_step": E, "chains": len(cfg["chain_group"]), "format": "sangria",
                   "autotune": tune, "concurrent_classes": not args.serial_classes,
                   "graph_replay_ms_per_step": round(graph_ms, 4)},
        "plain": None if plain is None else {
            "value": round(plain["rate1"], 1), "unit": "evals/s",
            "threads": {"1": round(plain["rate1"], 1), str(plain["nthreads"]): round(plain["rateN"], 1)},
            "method": "per-call drop-ins on host memory, per proposal: ucb_alignment, ucb_waveform, window model, "
                      "gaussian_log_likelihood, snr; the same sweep's proposals (all proposals of %d chains, %d "
                      "evaluations); from 1 host thread and from %d (the reference's OpenMP callers)"
                      % (plain["nchains"], plain["n"], plain["nthreads"]),
            "max_logL_rel_vs_batched": plain["max_rel_vs_batched"],
            "config0_case": {"value": round(plain0, 1),
                             "method": "the same drop-ins on the plain method's own case: 64 vectors, BW 128, "
                                       "W 256, Tobs 1 yr, one host thread"}},
        "batched_vs_plain": None if plain is None else round(value / plain["rate1"], 1),
        "batched_vs_plain_threads": None if plain is None else round(value / plain["rateN"], 1),
        "fisher": fisher,
        "wdm": wdm,
        "inner_product": inner,
        "globalfit": gfit,
        "noise": nzl,
        "sampler": smpl,
        "formats": fmts,
        "fstat": fst,
        "parity": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if args.no_globalfit:
        plan.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
