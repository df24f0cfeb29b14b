"""Synthetic inputs for the UCB likelihood path (parity tests and bench).

Parameter arrays follow the reference layout (map_ucb_params_to_array,
ucb/src/glass_ucb_model.c:40-53):
    [f0*T, cos(theta), phi, ln A, cos(iota), psi, phi0, fdot*T^2]
Noise follows GetNoiseModel's 3-channel "sangria" branch without confusion
noise (utils/src/glass_data.c:1156-1242) and AddNoise's per-bin Cholesky
draw with variance C/2 on the real and imaginary parts
(utils/src/glass_data.c:1323-1403); the random numbers come from numpy's
PCG64, not the reference's rand_r, so realisations differ while the
statistics agree.
"""
import numpy as np

YEAR_T = 31457280.0          # 2^21 * 15 s, the reference default Tobs
T_2YR = 62914560.0
T_4YR = 125829120.0
CLIGHT = 299792458.0
PI2 = 6.283185307179586
LARM = 2.5e9
FSTAR = CLIGHT / (2.0 * np.pi * LARM)


def draw_params(rng, n, T, f_lo=1e-3, f_hi=1e-2, fdot=(-1e-17, 1e-16), log10A=(-23.0, -21.0),
                log_f=False):
    """UCB parameter vectors with the configs' distributions."""
    if log_f:
        f0 = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), n))
    else:
        f0 = rng.uniform(f_lo, f_hi, n)
    p = np.empty((n, 8))
    p[:, 0] = f0 * T
    p[:, 1] = rng.uniform(-1, 1, n)
    p[:, 2] = rng.uniform(0, 2 * np.pi, n)
    p[:, 3] = np.log(10.0 ** rng.uniform(log10A[0], log10A[1], n))
    p[:, 4] = rng.uniform(-1, 1, n)
    p[:, 5] = rng.uniform(0, np.pi, n)
    p[:, 6] = rng.uniform(0, 2 * np.pi, n)
    p[:, 7] = rng.uniform(fdot[0], fdot[1], n) * T * T
    return p


def noise_levels_sangria(f):
    """get_noise_levels("sangria") (utils/src/glass_lisa.c:1107-1112)."""
    Spm = 5.76e-30 / (PI2 * f * CLIGHT) / (PI2 * f * CLIGHT) * (1.0 + (0.4e-3 / f) ** 2) * \
        (1.0 + (f / 8.0e-3) ** 4)
    Sop = 1.28e-22 * (PI2 * f / CLIGHT) * (PI2 * f / CLIGHT) * (1.0 + (2.0e-3 / f) ** 4)
    return Spm, Sop


def xyz_noise_cov(T, fmin, N, fstar=FSTAR):
    """C[3,3,N] for bins f = fmin + n/T: XYZnoise_FF / XYZcross_FF over 4
    (utils/src/glass_lisa.c:1121-1135, utils/src/glass_data.c:1206-1242)."""
    f = fmin + np.arange(N) / T
    Spm, Sop = noise_levels_sangria(f)
    x = f / fstar
    cx = np.cos(x)
    tr = np.sin(x) ** 2
    diag = 16.0 * tr * (2.0 * (1.0 + cx * cx) * Spm + Sop) / 4.0
    cross = -8.0 * tr * cx * (4.0 * Spm + Sop) / 4.0
    C = np.empty((3, 3, N))
    for a in range(3):
        for b in range(3):
            C[a, b] = diag if a == b else cross
    return C


def noise_draw(rng, C):
    """Complex Gaussian noise [3, N] with E|n|^2 = C (re/im variance C/2)."""
    N = C.shape[-1]
    Cm = np.moveaxis(C, -1, 0)                   # [N,3,3]
    Lc = np.linalg.cholesky(Cm)                  # lower
    u = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    n = np.einsum("nij,nj->ni", Lc, u) / np.sqrt(2.0)
    return n.T.copy()


def window_for(f0, T, nfft):
    """Data window centred on f0 (ucb/src/glass_ucb_data.c:317-326)."""
    fmin = f0 - (nfft // 2) / T
    qmin = int(fmin * T)
    return qmin


def scatter(window, tmpl, imin):
    """add_ucb_model (ucb/src/glass_ucb_model.c:192-232) for one template:
    window [C, W] += tmpl [C, BW] at offset imin, clipped to the window."""
    W = window.shape[-1]
    BW = tmpl.shape[-1]
    lo, hi = max(0, imin), min(W, imin + BW)
    if hi > lo:
        window[..., lo:hi] += tmpl[..., lo - imin:hi - imin]


def config0(seed=20240, n=64, T=YEAR_T, BW=128, W=256, waveform=None):
    """Plain-method case: n vectors, BW=128, one injected source in a window
    of W bins plus noise.  waveform(params, T, BW) -> complex [5, BW]."""
    rng = np.random.default_rng(seed)
    P = draw_params(rng, n, T)
    qmin = np.array([window_for(P[i, 0] / T, T, W) for i in range(n)], dtype=np.int64)
    C = np.stack([xyz_noise_cov(T, qmin[i] / T, W) for i in range(n)])
    data = np.zeros((n, 3, W), dtype=np.complex128)
    imin = np.zeros(n, dtype=np.int64)
    for i in range(n):
        q = int(P[i, 0] / T * T)
        imin[i] = (q - BW // 2) - qmin[i]
        if waveform is not None:
            h = waveform(P[i], T, BW)[:3]
            scatter(data[i], h, imin[i])
        data[i] += noise_draw(rng, C[i])
    return dict(params=P, T=T, BW=BW, W=W, qmin=qmin, imin=imin, C=C, data=data, seed=seed)


# proposal jump scales per parameter (f0*T, cos th, phi, ln A, cos i, psi,
# phi0, fdot*T^2): small moves of the kind a Fisher proposal makes
JUMP = np.array([0.05, 0.01, 0.01, 0.02, 0.01, 0.01, 0.02, 0.01])


def _clip_cos(p):
    p[..., 1] = np.clip(p[..., 1], -1.0, 1.0)
    p[..., 4] = np.clip(p[..., 4], -1.0, 1.0)
    return p


def config1(seed=1, ngroup=256, ntemp=16, nprop_per_chain=80, T=T_2YR, f_lo=0.5e-3, f_hi=20e-3,
            windows=(512, 1024, 2048, 4096), max_src=8, waveform=None, alignment=None):
    """Batched PT-MCMC case: ngroup data windows x ntemp chains, 1..max_src
    sources per chain, nprop_per_chain proposals per chain.

    waveform(params, T, BW) -> complex [5, BW] and
    alignment(params, T, NFFT, qmin) -> (BW, qmin, imin) build the injected
    data (any FP64 implementation; the data are only an input)."""
    rng = np.random.default_rng(seed)
    G = ngroup
    Wg = rng.choice(np.asarray(windows), G).astype(np.int32)
    fc = rng.uniform(f_lo, f_hi, G)
    gqmin = (fc * T).astype(np.int64) - Wg // 2
    nsrc = rng.integers(1, max_src + 1, G)
    data, invC33, inj = [], [], []
    for g in range(G):
        W = int(Wg[g])
        P = draw_params(rng, int(nsrc[g]), T)
        bins = rng.uniform(gqmin[g] + W / 8, gqmin[g] + 7 * W / 8, int(nsrc[g]))
        P[:, 0] = bins
        inj.append(P)
        C = xyz_noise_cov(T, gqmin[g] / T, W)
        d = noise_draw(rng, C)
        if waveform is not None:
            for p in P:
                BW, _, imin = alignment(p, T, W, int(gqmin[g]))
                scatter(d, waveform(p, T, BW)[:3], imin)
        data.append(d)
        invC33.append(C)  # inverted by the caller
    C = G * ntemp
    chain_group = np.repeat(np.arange(G), ntemp).astype(np.int32)
    chain_nsrc = nsrc[chain_group].astype(np.int32)
    src_params = np.zeros((C, 8, 8))
    for c in range(C):
        g = chain_group[c]
        n = chain_nsrc[c]
        src_params[c, :n] = _clip_cos(inj[g] + JUMP * rng.standard_normal((n, 8)))
    # proposals: nprop_per_chain per chain, cycling over its sources
    E = C * nprop_per_chain
    p_chain = np.repeat(np.arange(C), nprop_per_chain).astype(np.int32)
    k = np.tile(np.arange(nprop_per_chain), C)
    p_src = (k % chain_nsrc[p_chain]).astype(np.int32)
    p_params = _clip_cos(src_params[p_chain, p_src] + 0.5 * JUMP * rng.standard_normal((E, 8)))
    return dict(T=T, group_qmin=gqmin.astype(np.int32), group_W=Wg, data=data, C=invC33,
                chain_group=chain_group, chain_nsrc=chain_nsrc, src_params=src_params,
                p_chain=p_chain, p_src=p_src, p_params=p_params, injected=inj, seed=seed)


def wdm_pixel_cov(T, NF, NT):
    """3x3 pixel covariance per WDM layer m = 0..NF, [NF+1, 3, 3], of
    stationary noise with the analytic XYZ spectrum: the layer-centre
    approximation of stationary_dft_psd_to_wdm_psd (utils/src/glass_wavelet.c:
    902-921: DFT variance / ND at bin m NT/2), with the DFT variance
    4 ND^2 C(f)/T of this module's Fourier convention (C = E|d_k|^2)."""
    ND = NF * NT
    f = np.arange(NF + 1) * (NT / 2) / T
    f[0] = f[1]                          # the spectrum diverges at DC
    Cl = np.empty((NF + 1, 3, 3))
    for m in range(NF + 1):
        Cm = xyz_noise_cov(T, f[m], 1)[:, :, 0]
        Cl[m] = 4.0 * ND * Cm / T
    return Cl


def config2(seed=7, n=4096, NF=4096, NT=512, T=YEAR_T, inject=8):
    """Wavelet-domain case: an NT x NF grid over Tobs (15 s cadence at 1 yr:
    NT 512, NF 4096, dt = T/NT), n templates with config-0 distributions,
    stationary Gaussian pixel noise with per-layer 3x3 covariance, and the
    first `inject` templates' parameters (for the caller to inject).
    Returns the pixel data of the data range [kmin, kmax) = layers 1..NF-2
    as [3, npix], invC as 9 row-major planes [9, npix] and 6 planes
    (XX, YY, ZZ, XY, XZ, YZ) [6, npix]."""
    rng = np.random.default_rng(seed)
    dt = T / NT
    P = draw_params(rng, n, T)
    Cl = wdm_pixel_cov(T, NF, NT)
    kmin, kmax = NT, NT * (NF - 1)
    layers = np.arange(kmin, kmax) // NT
    Lc = np.linalg.cholesky(Cl)                       # [NF+1, 3, 3]
    z = rng.standard_normal((kmax - kmin, 3))
    data = np.einsum("kij,kj->ik", Lc[layers], z)     # [3, npix]
    iC = np.linalg.inv(Cl)                            # [NF+1, 3, 3]
    invC9 = np.ascontiguousarray(iC[layers].reshape(-1, 9).T)
    invC6 = np.ascontiguousarray(invC9[[0, 4, 8, 1, 2, 5]])
    return dict(params=P, T=T, NF=NF, NT=NT, dt=dt, kmin=kmin, kmax=kmax, data=np.ascontiguousarray(data),
                invC9=invC9, invC6=invC6, inject=P[:inject].copy(), seed=seed)


def wdm_random_noise(NF, NT, kmin, kmax, seed):
    """Per-layer random SPD pixel covariance (condition ~10) on the pixel
    range [kmin, kmax): (data [3, npix], invC [9, npix]).  With it the
    full-grid wavelet logL is accurate to rounding, unlike the XYZ spectrum
    near DC whose covariance is nearly singular."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((NF + 1, 3, 3))
    Cl = np.einsum("mij,mkj->mik", A, A) + 3 * np.eye(3)
    layers = np.arange(kmin, kmax) // NT
    d = np.einsum("kij,kj->ik", np.linalg.cholesky(Cl)[layers], rng.standard_normal((len(layers), 3)))
    invC9 = np.ascontiguousarray(np.linalg.inv(Cl)[layers].reshape(-1, 9).T)
    return np.ascontiguousarray(d), invC9


def logl_batch_inputs(n, seed=11, T=T_2YR, windows=(512, 1024, 2048, 4096), f_lo=0.5e-3, f_hi=20e-3,
                      device="cpu"):
    """n independent full-window likelihood items of the batched tempered
    sampler's shape: window W_e drawn from `windows`, centred at f ~
    U[f_lo, f_hi]; per bin the analytic 3x3 XYZ covariance C = a I + b J
    (xyz_noise_cov: equal diagonals d and cross terms b, a = d - b), its
    closed-form inverse (1/a)(I - b/(a + 3b) J), data = a noise draw and
    model = data - (a second noise draw), so the residual has the noise
    statistics of a model that fits.  Noise: n = sqrt(a/2)(u - mean(u))
    + sqrt((a + 3b)/2) mean(u) for complex unit normals u (covariance
    C/2 on the real and imaginary parts).

    Generated with torch on `device`.  Returns dict W [n] int32, off [n]
    int64 (bin offsets, identical for data, model and invC), data and
    model complex128 [sum 3 W] (item e: X, Y, Z of W_e bins at
    3 off[e]), invC float64 [sum 6 W] (6 planes XX YY ZZ XY XZ YZ at
    6 off[e])."""
    import torch
    rng = np.random.default_rng(seed)
    W = rng.choice(np.asarray(windows), n).astype(np.int32)
    qmin = (rng.uniform(f_lo, f_hi, n) * T).astype(np.int64) - W // 2
    off = np.zeros(n, dtype=np.int64)
    off[1:] = np.cumsum(W[:-1].astype(np.int64))
    nb = int(W.astype(np.int64).sum())
    dev = torch.device(device)
    Wt = torch.as_tensor(W, device=dev, dtype=torch.int64)
    ot = torch.as_tensor(off, device=dev)
    item = torch.repeat_interleave(torch.arange(n, device=dev), Wt)
    j = torch.arange(nb, device=dev) - ot[item]
    f = (torch.as_tensor(qmin, device=dev, dtype=torch.float64)[item] + j.to(torch.float64)) / T
    Spm = 5.76e-30 / (PI2 * f * CLIGHT) / (PI2 * f * CLIGHT) * (1.0 + (0.4e-3 / f) ** 2) * (1.0 + (f / 8.0e-3) ** 4)
    Sop = 1.28e-22 * (PI2 * f / CLIGHT) * (PI2 * f / CLIGHT) * (1.0 + (2.0e-3 / f) ** 4)
    x = f / FSTAR
    cx = torch.cos(x)
    tr = torch.sin(x) ** 2
    d = 16.0 * tr * (2.0 * (1.0 + cx * cx) * Spm + Sop) / 4.0
    b = -8.0 * tr * cx * (4.0 * Spm + Sop) / 4.0
    a = d - b
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))

    def draw():
        u = torch.complex(torch.randn((3, nb), generator=g, device=dev, dtype=torch.float64),
                          torch.randn((3, nb), generator=g, device=dev, dtype=torch.float64))
        m = u.mean(dim=0, keepdim=True)
        return torch.sqrt(a / 2.0) * (u - m) + torch.sqrt((a + 3.0 * b) / 2.0) * m

    # layout: item e, channel c, bin j -> c*W_e + j past 3*off_e (6*off_e for invC)
    idx3 = 3 * ot[item] + j
    data = torch.empty(3 * nb, dtype=torch.complex128, device=dev)
    model = torch.empty_like(data)
    nd, nr = draw(), draw()
    for c in range(3):
        data[idx3 + c * Wt[item]] = nd[c]
        model[idx3 + c * Wt[item]] = nd[c] - nr[c]
    k = b / (a + 3.0 * b)
    idiag = (1.0 - k) / a
    ioff = -k / a
    invC = torch.empty(6 * nb, dtype=torch.float64, device=dev)
    idx6 = 6 * ot[item] + j
    for c in range(6):
        invC[idx6 + c * Wt[item]] = idiag if c < 3 else ioff
    return dict(W=W, off=off, data=data, model=model, invC=invC, qmin=qmin, T=T, seed=seed)


class _Windows:
    """Per-window views [3, W_g] (complex data) or [3, 3, W_g] (invC from
    the 6 stored planes) over the flat arrays of a segmented band."""

    def __init__(self, flat, off, W, kind):
        self.flat, self.off, self.W, self.kind = flat, off, W, kind

    def __len__(self):
        return len(self.W)

    def __getitem__(self, g):
        o, w = int(self.off[g]), int(self.W[g])
        if self.kind == "data":
            return self.flat[3 * o:3 * (o + w)].reshape(3, w)
        p = self.flat[6 * o:6 * (o + w)].reshape(6, w)
        return np.stack([p[0], p[3], p[4], p[3], p[1], p[5], p[4], p[5], p[2]]).reshape(3, 3, w)


def segment_band(T, nseg, f_lo, f_hi, wmin=64, wmax=2048):
    """Contiguous segments [q_g, q_g + W_g) tiling the bins of [f_lo, f_hi]
    with widths W = clip(c q, wmin, wmax) (density-weighted for a
    log-uniform population: equal source counts per segment where the clip
    allows), c chosen by bisection so that there are nseg of them."""
    qlo, qhi = int(f_lo * T), int(f_hi * T)

    def build(c):
        out, q = [], qlo
        while q < qhi:
            w = int(min(max(c * q, wmin), wmax))
            out.append((q, w))
            q += w
        return out

    lo, hi = 1e-9, 1.0
    for _ in range(80):
        c = np.sqrt(lo * hi)
        if len(build(c)) > nseg:
            lo = c
        else:
            hi = c
    segs = build(hi)
    q0 = np.array([s[0] for s in segs], dtype=np.int64)
    W = np.array([s[1] for s in segs], dtype=np.int32)
    return q0, W


def config4(seed=4, nsrc=20000, nseg=4000, ntemp=8, nchain=32, T=T_4YR, f_lo=1e-4, f_hi=2.5e-2,
            log10A=(-24.0, -21.0), fisher_every=8, waveform=None, alignment=None):
    """Global-fit UCB block: nsrc binaries with f0 log-uniform in [f_lo,
    f_hi] and A log-uniform, the band cut into nseg density-weighted
    segments of 64-2048 bins (segment_band), each segment a data window of
    stationary XYZ noise (closed-form draw and inverse of C = a I + b J, as
    logl_batch_inputs) plus its injected sources, and ntemp x nchain chains
    per segment holding the segment's sources (jittered).  One proposal per
    chain with sources (a random source, half a jump), and one Fisher item
    (chain, random source) on every fisher_every-th such chain.

    waveform(params, T, BW) -> complex [5, BW] and alignment(params, T, W,
    qmin) -> (BW, qmin, imin) build the injections.  Flat layouts: data
    complex [sum 3 W_g], invC float [sum 6 W_g] (XX YY ZZ XY XZ YZ);
    data_w / invC_w give per-window views."""
    rng = np.random.default_rng(seed)
    q0, W = segment_band(T, nseg, f_lo, f_hi)
    G = len(W)
    off = np.zeros(G, dtype=np.int64)
    off[1:] = np.cumsum(W[:-1].astype(np.int64))
    nb = int(W.astype(np.int64).sum())
    # sources
    P = draw_params(rng, nsrc, T, log10A=log10A)
    P[:, 0] = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), nsrc)) * T
    qs = np.floor(P[:, 0]).astype(np.int64)
    grp = np.clip(np.searchsorted(q0, qs, side="right") - 1, 0, G - 1)
    order = np.argsort(grp, kind="stable")
    P, grp = P[order], grp[order]
    ng = np.bincount(grp, minlength=G)
    first = np.zeros(G, dtype=np.int64)
    first[1:] = np.cumsum(ng[:-1])
    # noise: per bin C = a I + b J (xyz_noise_cov), closed-form draw and inverse
    seg_of_bin = np.repeat(np.arange(G), W)
    f = (q0[seg_of_bin] + (np.arange(nb) - off[seg_of_bin])) / T
    Spm, Sop = noise_levels_sangria(f)
    x = f / FSTAR
    cx, tr = np.cos(x), np.sin(x) ** 2
    d = 16.0 * tr * (2.0 * (1.0 + cx * cx) * Spm + Sop) / 4.0
    b = -8.0 * tr * cx * (4.0 * Spm + Sop) / 4.0
    a = d - b
    u = rng.standard_normal((3, nb)) + 1j * rng.standard_normal((3, nb))
    m = u.mean(axis=0, keepdims=True)
    noise = np.sqrt(a / 2.0) * (u - m) + np.sqrt((a + 3.0 * b) / 2.0) * m
    data = np.empty(3 * nb, dtype=np.complex128)
    invC = np.empty(6 * nb, dtype=np.float64)
    k = b / (a + 3.0 * b)
    idiag, ioff = (1.0 - k) / a, -k / a
    for g in range(G):
        o, w = int(off[g]), int(W[g])
        sl = slice(o, o + w)
        dg = noise[:, sl].copy()
        if waveform is not None:
            for p in P[first[g]:first[g] + ng[g]]:
                BW, _, imin = alignment(p, T, w, int(q0[g]))
                scatter(dg, waveform(p, T, BW)[:3], imin)
        data[3 * o:3 * (o + w)] = dg.reshape(-1)
        invC[6 * o:6 * (o + w)] = np.concatenate([idiag[sl]] * 3 + [ioff[sl]] * 3)
    # chains
    per = ntemp * nchain
    C = G * per
    chain_group = np.repeat(np.arange(G), per).astype(np.int32)
    chain_nsrc = ng[chain_group].astype(np.int32)
    max_src = max(8, int(ng.max()))
    src_params = np.zeros((C, max_src, 8))
    slots = np.arange(max_src)
    live = slots[None, :] < chain_nsrc[:, None]
    srcidx = np.minimum(first[chain_group][:, None] + slots[None, :], nsrc - 1)
    src_params[live] = P[srcidx[live]]
    src_params[live] += JUMP * rng.standard_normal((int(live.sum()), 8))
    _clip_cos(src_params)
    # proposals: one per chain with sources
    pc = np.nonzero(chain_nsrc > 0)[0].astype(np.int32)
    ps = (rng.random(len(pc)) * chain_nsrc[pc]).astype(np.int32)
    pp = _clip_cos(src_params[pc, ps] + 0.5 * JUMP * rng.standard_normal((len(pc), 8)))
    fc = pc[::fisher_every]
    fs = (rng.random(len(fc)) * chain_nsrc[fc]).astype(np.int32)
    return dict(T=T, group_qmin=q0.astype(np.int32), group_W=W, group_off=off, data_flat=data, invC_flat=invC,
                data=_Windows(data, off, W, "data"), invC=_Windows(invC, off, W, "invC"),
                chain_group=chain_group, chain_nsrc=chain_nsrc, max_src=max_src, src_params=src_params,
                p_chain=pc, p_src=ps, p_params=pp, f_chain=fc.astype(np.int32), f_src=fs,
                sources=P, nsrc_per_group=ng, seed=seed)
