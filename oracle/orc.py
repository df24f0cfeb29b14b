"""ctypes binding of the oracle library (test infrastructure only).

The oracle is the CPU restatement of the reference's UCB likelihood path in
``oracle/*.c``.  Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` load it, as the checker; the product
(``glass_amd``) never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LIB_REL_PATH = os.path.join(HERE, "liboracle_rel.so")
REF_KISSFFT = os.path.join(HERE, "_ref", "libkissfft_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Orbit(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double), ("fstar", ctypes.c_double), ("R", ctypes.c_double)]


class Chain(ctypes.Structure):
    _fields_ = [
        ("nsrc", ctypes.c_int), ("W", ctypes.c_int), ("qmin", ctypes.c_int),
        ("data", _dp), ("invC", _dp), ("src_params", _dp), ("tmpl", _dp),
        ("src_bw", _ip), ("src_imin", _ip), ("bwmax", ctypes.c_int),
    ]


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def build():
    """Compile the oracle (and, where the reference tree exists, oracle/_ref)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def _declare(L):
    if True:
        L.orc_fft.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_fftr.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_fftri.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_init_analytic_orbit.argtypes = [ctypes.POINTER(Orbit)]
        L.orc_analytic_orbits.argtypes = [ctypes.POINTER(Orbit), ctypes.c_double, _dp, _dp, _dp]
        L.orc_ucb_waveform.argtypes = [ctypes.POINTER(Orbit), ctypes.c_char_p, ctypes.c_double,
                                       ctypes.c_double, _dp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                       _dp, ctypes.c_int, ctypes.c_int]
        L.orc_ucb_slow_series.argtypes = [ctypes.POINTER(Orbit), ctypes.c_double, ctypes.c_double,
                                          _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_ucb_bandwidth.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int]
        L.orc_ucb_bandwidth.restype = ctypes.c_int
        L.orc_ucb_alignment.argtypes = [ctypes.POINTER(Orbit), ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, _dp, _ip, _ip, _ip]
        L.orc_AEnoise.argtypes = [ctypes.c_double] * 3
        L.orc_AEnoise.restype = ctypes.c_double
        L.orc_xyz_noise_covariance.argtypes = [ctypes.POINTER(Orbit), ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, _dp]
        L.orc_invert_noise_covariance.argtypes = [_dp, _dp, _dp, ctypes.c_int]
        L.orc_fourier_nwip.argtypes = [_dp, _dp, _dp, ctypes.c_int]
        L.orc_fourier_nwip.restype = ctypes.c_double
        L.orc_wavelet_nwip.argtypes = [_dp, _dp, _dp, _ip, ctypes.c_int]
        L.orc_wavelet_nwip.restype = ctypes.c_double
        L.orc_snr.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_snr.restype = ctypes.c_double
        L.orc_gaussian_log_likelihood.argtypes = [_dp] * 11 + [ctypes.c_int, ctypes.c_int]
        L.orc_gaussian_log_likelihood.restype = ctypes.c_double
        L.orc_add_ucb_model.argtypes = [_dp] * 6 + [ctypes.c_int] * 3
        L.orc_ucb_fisher_fd.argtypes = [ctypes.POINTER(Orbit), ctypes.c_char_p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_eval_logl.argtypes = [ctypes.POINTER(Orbit), ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(Chain), ctypes.c_int, _dp]
        L.orc_eval_logl.restype = ctypes.c_double
        L.orc_eval_batch.argtypes = [ctypes.POINTER(Orbit), ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(Chain), ctypes.c_int, _ip, _ip, _dp, _dp,
                                     ctypes.c_int]
        L.orc_eval_batch.restype = ctypes.c_double
        _wdm_signatures(L)
        _noise_signatures(L)
        _sampler_signatures(L)
        _fstat_signatures(L)
    return L


def lib():
    """The oracle at the parity flags (-O2 -ffp-contract=off): the checker."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


_lib_rel = None


def lib_release():
    """The same restatement built with the reference's release flags
    (-O3 with FP contraction, CMakeLists.txt:37-39; -march=x86-64-v3 stands
    in for -march=native since the library is built on one host and timed
    on another): timing only, never the checker."""
    global _lib_rel
    if _lib_rel is None:
        if not os.path.exists(LIB_REL_PATH):
            build()
        _lib_rel = _declare(ctypes.CDLL(LIB_REL_PATH))
    return _lib_rel


def analytic_orbit():
    o = Orbit()
    lib().orc_init_analytic_orbit(ctypes.byref(o))
    return o


def fft(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_fft(x.view(np.float64).ctypes.data_as(_dp), out.view(np.float64).ctypes.data_as(_dp),
                  len(x), int(inverse))
    return out


def fftr(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x) // 2 + 1, dtype=np.complex128)
    lib().orc_fftr(_d(x), out.view(np.float64).ctypes.data_as(_dp), len(x))
    return out


def fftri(F, n):
    F = np.ascontiguousarray(F, dtype=np.complex128)
    out = np.zeros(n, dtype=np.float64)
    lib().orc_fftri(F.view(np.float64).ctypes.data_as(_dp), _d(out), n)
    return out


def ucb_waveform(params, T, BW, t0=0.0, fmt="sangria", NI=3, orbit=None):
    """Returns X, Y, Z, A, E as complex arrays of BW bins."""
    orbit = orbit or analytic_orbit()
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros((5, 2 * BW))
    rc = lib().orc_ucb_waveform(ctypes.byref(orbit), fmt.encode(), T, t0, _d(p), len(p),
                                *[_d(out[c]) for c in range(5)], BW, NI)
    if rc != 0:
        raise ValueError("unsupported format " + fmt)
    return out.view(np.complex128)


def slow_series(params, T, BW, t0=0.0, orbit=None):
    orbit = orbit or analytic_orbit()
    p = np.ascontiguousarray(params, dtype=np.float64)
    d = np.zeros((6, 2 * BW))
    lib().orc_ucb_slow_series(ctypes.byref(orbit), T, t0, _d(p), len(p), BW, _d(d))
    return d.view(np.complex128)


def ucb_alignment(params, T, NFFT, data_qmin, orbit=None):
    orbit = orbit or analytic_orbit()
    p = np.ascontiguousarray(params, dtype=np.float64)
    bw, qm, im = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().orc_ucb_alignment(ctypes.byref(orbit), T, NFFT, data_qmin, _d(p), ctypes.byref(bw),
                            ctypes.byref(qm), ctypes.byref(im))
    return bw.value, qm.value, im.value


def xyz_noise(T, fmin, N, orbit=None):
    """Returns (C[3,3,N], invC[3,3,N], logdetC[N])."""
    orbit = orbit or analytic_orbit()
    C = np.zeros((3, 3, N))
    lib().orc_xyz_noise_covariance(ctypes.byref(orbit), T, fmin, N, _d(C))
    invC = np.zeros_like(C)
    ld = np.zeros(N)
    lib().orc_invert_noise_covariance(_d(C), _d(invC), _d(ld), N)
    return C, invC, ld


def invert_noise(C):
    C = np.ascontiguousarray(C, dtype=np.float64)
    N = C.shape[-1]
    invC = np.zeros_like(C)
    ld = np.zeros(N)
    lib().orc_invert_noise_covariance(_d(C), _d(invC), _d(ld), N)
    return invC, ld


def fourier_nwip(a, b, invC, N):
    a = np.ascontiguousarray(a).view(np.float64) if np.iscomplexobj(a) else np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b).view(np.float64) if np.iscomplexobj(b) else np.ascontiguousarray(b, dtype=np.float64)
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    return lib().orc_fourier_nwip(_d(a), _d(b), _d(invC), N)


def wavelet_nwip(a, b, invC, lst):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    return lib().orc_wavelet_nwip(_d(a), _d(b), _d(invC), _i(lst), len(lst))


def snr(X, Y, Z, invC, N):
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.complex128).view(np.float64) for v in (X, Y, Z))
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    return lib().orc_snr(_d(X), _d(Y), _d(Z), _d(invC), invC.shape[-1], N)


def gaussian_log_likelihood(data, model, invC, nchannel=3):
    """data, model: complex arrays [5, NFFT] (X,Y,Z,A,E); invC [3,3,NFFT]."""
    data = np.ascontiguousarray(data, dtype=np.complex128)
    model = np.ascontiguousarray(model, dtype=np.complex128)
    NFFT = data.shape[1]
    dv = data.view(np.float64)
    mv = model.view(np.float64)
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    return lib().orc_gaussian_log_likelihood(*[_d(dv[c]) for c in range(5)],
                                             *[_d(mv[c]) for c in range(5)], _d(invC), NFFT,
                                             nchannel)


def ucb_fisher_fd(params, T, NFFT, data_qmin, invC, t0=0.0, fmt="sangria", orbit=None):
    orbit = orbit or analytic_orbit()
    p = np.ascontiguousarray(params, dtype=np.float64)
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    F = np.zeros((8, 8))
    lib().orc_ucb_fisher_fd(ctypes.byref(orbit), fmt.encode(), T, t0, NFFT, data_qmin, _d(p),
                            _d(invC), _d(F))
    return F


class ChainSet:
    """Holds numpy buffers for a list of orc_chain structs (keeps them alive)."""

    def __init__(self, chains):
        self.keep = []
        self.arr = (Chain * len(chains))()
        for c, ch in enumerate(chains):
            bufs = {k: np.ascontiguousarray(ch[k]) for k in
                    ("data", "invC", "src_params", "tmpl", "src_bw", "src_imin")}
            self.keep.append(bufs)
            s = self.arr[c]
            s.nsrc = int(ch["nsrc"])
            s.W = int(ch["W"])
            s.qmin = int(ch["qmin"])
            s.bwmax = int(ch["bwmax"])
            s.data = _d(bufs["data"])
            s.invC = _d(bufs["invC"])
            s.src_params = _d(bufs["src_params"])
            s.tmpl = _d(bufs["tmpl"])
            s.src_bw = _i(bufs["src_bw"])
            s.src_imin = _i(bufs["src_imin"])


def eval_batch(chainset, T, chain_of, sid_of, params, t0=0.0, nthreads=0, orbit=None, release=False):
    """logL of proposals (chain, source, params) and the wall seconds
    (omp_get_wtime) of the OpenMP loop; release=True runs the release-flags
    build (timing only)."""
    orbit = orbit or analytic_orbit()
    chain_of = np.ascontiguousarray(chain_of, dtype=np.int32)
    sid_of = np.ascontiguousarray(sid_of, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(len(chain_of))
    secs = (lib_release() if release else lib()).orc_eval_batch(ctypes.byref(orbit), T, t0, chainset.arr, len(chain_of),
                                _i(chain_of), _i(sid_of), _d(params), _d(out), nthreads)
    return out, secs


_ref = None


def ref_kissfft():
    """The reference's own kissfft, built from its sources in oracle/_ref (or None)."""
    global _ref
    if _ref is None and os.path.exists(REF_KISSFFT):
        L = ctypes.CDLL(REF_KISSFFT)
        L.kiss_fft_alloc.restype = ctypes.c_void_p
        L.kiss_fft_alloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.kiss_fft.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.kiss_fftr_alloc.restype = ctypes.c_void_p
        L.kiss_fftr_alloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.kiss_fftr.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.kiss_fftri.argtypes = [ctypes.c_void_p, _dp, _dp]
        _ref = L
    return _ref


# ----------------------------------------------------------------- WDM -----

class WdmStruct(ctypes.Structure):
    _fields_ = [("NF", ctypes.c_int), ("NT", ctypes.c_int), ("dt", ctypes.c_double),
                ("df", ctypes.c_double), ("A", ctypes.c_double), ("kmin", ctypes.c_int),
                ("kmax", ctypes.c_int), ("imin", ctypes.c_int), ("imax", ctypes.c_int),
                ("window", _dp)]


def _wdm_signatures(L):
    c_int, c_d = ctypes.c_int, ctypes.c_double
    W = ctypes.POINTER(WdmStruct)
    L.orc_wdm_init.argtypes = [W, c_int, c_int, c_d, _dp]
    L.orc_incomplete_beta.argtypes = [c_d, c_d, c_d]
    L.orc_incomplete_beta.restype = c_d
    L.orc_wdm_filter.argtypes = [_dp, c_int, c_int, c_d, c_int]
    L.orc_wavelet_transform_freq.argtypes = [W, _dp, _dp]
    L.orc_wavelet_transform_timefreq.argtypes = [W, _dp]
    L.orc_wavelet_transform_inverse_freq.argtypes = [W, _dp, _dp]
    L.orc_wavelet_transform_timefreq_by_layers.argtypes = [W, c_int, c_int, _dp, _dp]
    L.orc_wdm_layer_var.argtypes = [W, _dp, _dp]
    L.orc_spline_coefficients.argtypes = [c_int, _dp, _dp, _dp]
    L.orc_spline_eval.argtypes = [c_int, _dp, _dp, c_int, c_d]
    L.orc_spline_eval.restype = c_d
    L.orc_binary_search.argtypes = [_dp, c_int, c_int, c_d]
    L.orc_binary_search.restype = c_int
    L.orc_even_search.argtypes = [_dp, c_int, c_int, c_d]
    L.orc_even_search.restype = c_int
    L.orc_iorbit_size.argtypes = [c_d]
    L.orc_iorbit_size.restype = c_int
    L.orc_iorbit_fill.argtypes = [ctypes.POINTER(Orbit), c_d, c_d, c_int, _dp, _dp, _dp]
    L.orc_extract_amplitude_and_phase.argtypes = [c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_unwrap_phase.argtypes = [c_int, _dp]
    L.orc_ucb_wavelet_layers.argtypes = [W, c_d, _dp, _ip, _ip]
    L.orc_ucb_waveform_wavelet_compact.argtypes = [W, c_int, _dp, _dp, c_d, c_d, _dp, _ip, _ip,
                                                   _dp, _dp, _dp]
    L.orc_ucb_waveform_wavelet_compact.restype = c_int
    L.orc_ucb_wavelet_series.argtypes = [c_int, _dp, _dp, c_d, c_d, _dp, c_int, c_d, c_d, _dp]
    L.orc_ucb_barycenter_waveform.argtypes = [_dp, c_int, _dp, _dp, _dp, c_d]
    L.orc_set_numeric_orbit.argtypes = [c_int, _dp, _dp]
    L.orc_numeric_orbit_armlength.argtypes = [c_int, _dp]
    L.orc_numeric_orbit_armlength.restype = c_d
    L.orc_set_cr_trig.argtypes = [c_int]
    L.orc_set_wavelet_layers.argtypes = [c_int]
    L.orc_ucb_wavelet_response.argtypes = [c_int, _dp, _dp, c_d, c_d, _dp, _dp, _dp, _dp, _dp]
    L.orc_snr_wavelet.argtypes = [_dp, _dp, _dp, _dp, c_int, _ip, c_int]
    L.orc_snr_wavelet.restype = c_d
    L.orc_gaussian_log_likelihood_wavelet.argtypes = [W] + [_dp] * 6 + [_ip, c_int, _dp, c_int]
    L.orc_gaussian_log_likelihood_wavelet.restype = c_d
    L.orc_wdm_eval_batch.argtypes = [W, c_int, _dp, _dp, c_d, c_d, c_int, _dp, _dp, _dp, c_int,
                                     _dp, _dp, _ip, c_int]
    L.orc_wdm_eval_batch.restype = c_d
    L.orc_ucb_fisher_wavelet.argtypes = [W, c_int, _dp, _dp, c_d, c_d, _dp, _dp, c_int, _dp]
    L.orc_ucb_fisher_wavelet.restype = c_int


class Wdm:
    """Hand-filled WDM geometry (NF layers x NT time pixels of duration dt)
    with its forward-normalised Meyer window."""

    def __init__(self, NF, NT, dt):
        self.window = np.zeros(NT // 2 + 1)
        self.s = WdmStruct()
        lib().orc_wdm_init(ctypes.byref(self.s), NF, NT, float(dt), _d(self.window))
        for f, _ in WdmStruct._fields_:
            if f != "window":
                setattr(self, f, getattr(self.s, f))

    @property
    def npix(self):
        return self.kmax - self.kmin

    @property
    def ref(self):
        return ctypes.byref(self.s)


def incomplete_beta(a, b, x):
    return lib().orc_incomplete_beta(a, b, x)


def wdm_filter(Nf, Nt, A=0.0, forward=True):
    phif = np.zeros(Nt // 2 + 1)
    lib().orc_wdm_filter(_d(phif), Nf, Nt, A, int(forward))
    return phif


def wavelet_transform_timefreq(w, x):
    out = np.array(x, dtype=np.float64, copy=True)
    lib().orc_wavelet_transform_timefreq(w.ref, _d(out))
    return out


def wavelet_transform_freq(w, freq):
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    out = np.zeros(w.NF * w.NT)
    lib().orc_wavelet_transform_freq(w.ref, _d(freq), _d(out))
    return out


def wavelet_transform_inverse_freq(w, pix):
    pix = np.ascontiguousarray(pix, dtype=np.float64)
    out = np.zeros(w.NF * w.NT + 2)
    lib().orc_wavelet_transform_inverse_freq(w.ref, _d(pix), _d(out))
    return out


def wavelet_transform_by_layers(w, jmin, nlayers, phif, x):
    out = np.array(x, dtype=np.float64, copy=True)
    phif = np.ascontiguousarray(phif, dtype=np.float64)
    lib().orc_wavelet_transform_timefreq_by_layers(w.ref, jmin, nlayers, _d(phif), _d(out))
    return out


def wdm_layer_var(w, dft_psd):
    psd = np.ascontiguousarray(dft_psd, dtype=np.float64)
    out = np.zeros(w.NF + 1)
    lib().orc_wdm_layer_var(w.ref, _d(psd), _d(out))
    return out


def spline_coefficients(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    c = np.zeros(4 * len(x))
    lib().orc_spline_coefficients(len(x), _d(x), _d(y), _d(c))
    return c


def spline_eval(x, c, xv, even=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    return np.array([lib().orc_spline_eval(len(x), _d(x), _d(c), int(even), float(v))
                     for v in np.atleast_1d(xv)])


def binary_search(a, x):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_binary_search(_d(a), 0, len(a), float(x))


def iorbit(Tobs, t0=0.0, orbit=None):
    """Knots, positions [9,N] and spline coefficients [9,4N] of the
    interpolated analytic orbit."""
    orbit = orbit or analytic_orbit()
    N = lib().orc_iorbit_size(float(Tobs))
    t = np.zeros(N)
    pos = np.zeros((9, N))
    coef = np.zeros((9, 4 * N))
    lib().orc_iorbit_fill(ctypes.byref(orbit), float(Tobs), float(t0), N, _d(t), _d(pos), _d(coef))
    return t, pos, coef


def extract_amplitude_and_phase(M, Mf, phiR):
    M, Mf, phiR = (np.ascontiguousarray(v, dtype=np.float64) for v in (M, Mf, phiR))
    As = np.zeros(len(M))
    Dphi = np.zeros(len(M))
    lib().orc_extract_amplitude_and_phase(len(M), _d(As), _d(Dphi), _d(M), _d(Mf), _d(phiR))
    return As, Dphi


def unwrap_phase(ph):
    out = np.array(ph, dtype=np.float64, copy=True)
    lib().orc_unwrap_phase(len(out), _d(out))
    return out


def ucb_wavelet_layers(w, Tobs, params):
    p = np.ascontiguousarray(params, dtype=np.float64)
    j0, nl = ctypes.c_int(), ctypes.c_int()
    lib().orc_ucb_wavelet_layers(w.ref, float(Tobs), _d(p), ctypes.byref(j0), ctypes.byref(nl))
    return j0.value, nl.value


class cr_trig:
    """Context manager: the wavelet template path of the oracle uses
    correctly rounded sin/cos inside the block (glibc's otherwise)."""

    def __enter__(self):
        lib().orc_set_cr_trig(1)
        return self

    def __exit__(self, *a):
        lib().orc_set_cr_trig(0)


def ucb_waveform_wavelet(w, orb, Tobs, params, t0=0.0, layers=0):
    """Compact ucb_waveform_wavelet: (list [Nw] int32, values [3, Nw]).
    layers > 0 forces that layer count (orc_set_wavelet_layers)."""
    t, _, coef = orb
    p = np.ascontiguousarray(params, dtype=np.float64)
    cap = max(4, layers + 1) * w.NT
    lst = np.zeros(cap, dtype=np.int32)
    vals = np.zeros((3, cap))
    nw = ctypes.c_int()
    lib().orc_set_wavelet_layers(int(layers))
    try:
        lib().orc_ucb_waveform_wavelet_compact(w.ref, len(t), _d(t), _d(coef), float(Tobs), float(t0),
                                               _d(p), _i(lst), ctypes.byref(nw), _d(vals[0]),
                                               _d(vals[1]), _d(vals[2]))
    finally:
        lib().orc_set_wavelet_layers(0)
    n = nw.value
    return lst[:n].copy(), vals[:, :n].copy()


def ucb_wavelet_series(orb, Tobs, params, N, dt_s, fhet, t0=0.0):
    """Interpolated TDI series [3, N] at t0 + i dt_s, heterodyned by fhet."""
    t, _, coef = orb
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros((3, N))
    lib().orc_ucb_wavelet_series(len(t), _d(t), _d(coef), float(Tobs), float(t0), _d(p), int(N),
                                 float(dt_s), float(fhet), _d(out))
    return out


class numeric_orbit:
    """Context manager: inside the block the oracle's ucb_waveform takes its
    spacecraft positions from splines over the knots t [N] and positions
    pos [9, N] (rows x1 y1 z1 x2 ... as an orbit file's columns), as
    interpolate_orbits does for a numeric orbit.  .orbit is the Orbit with
    that file's average arm length and transfer frequency."""

    def __init__(self, t, pos):
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        N = len(self.t)
        self.coef = np.zeros((9, 4 * N))
        for s in range(9):
            self.coef[s] = spline_coefficients(self.t, pos[s]).reshape(-1)
        L = lib().orc_numeric_orbit_armlength(N, _d(pos))
        self.orbit = Orbit(L, 299792458.0 / (2.0 * np.pi * L), L / (2.0 * 1.73205080757 * 1.49597870700e11)
                           * 1.49597870700e11)

    def __enter__(self):
        lib().orc_set_numeric_orbit(len(self.t), _d(self.t), _d(self.coef))
        return self

    def __exit__(self, *a):
        lib().orc_set_numeric_orbit(0, None, None)


def ucb_barycenter_waveform(params, times, T):
    """(phase, amp) of ucb_barycenter_waveform at the given SSB times."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    tt = np.ascontiguousarray(times, dtype=np.float64)
    ph, am = np.zeros(len(tt)), np.zeros(len(tt))
    lib().orc_ucb_barycenter_waveform(_d(p), len(tt), _d(tt), _d(ph), _d(am), float(T))
    return ph, am


def ucb_wavelet_response(orb, Tobs, params, t0=0.0):
    """(phase_ssb [N], phase_sc [N], amp [3,N], phase [3,N]) of the response stage."""
    t, _, coef = orb
    N = len(t)
    p = np.ascontiguousarray(params, dtype=np.float64)
    ps, pc, a3, p3 = np.zeros(N), np.zeros(N), np.zeros((3, N)), np.zeros((3, N))
    lib().orc_ucb_wavelet_response(N, _d(t), _d(coef), float(Tobs), float(t0), _d(p), _d(ps), _d(pc), _d(a3),
                                   _d(p3))
    return ps, pc, a3, p3


def ucb_fisher_wavelet(w, orb, Tobs, params, invC9, t0=0.0):
    """ucb_fisher_wavelet before the eigensolve -> Fisher [8, 8]."""
    t, _, coef = orb
    p = np.ascontiguousarray(params, dtype=np.float64)
    ic = np.ascontiguousarray(invC9, dtype=np.float64)
    F = np.zeros((8, 8))
    rc = lib().orc_ucb_fisher_wavelet(w.ref, len(t), _d(t), _d(coef), float(Tobs), float(t0), _d(p), _d(ic),
                                      ic.shape[-1], _d(F))
    if rc:
        raise MemoryError("orc_ucb_fisher_wavelet")
    return F


def snr_wavelet(X, Y, Z, invC9, lst):
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, Y, Z))
    ic = np.ascontiguousarray(invC9, dtype=np.float64)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    return lib().orc_snr_wavelet(_d(X), _d(Y), _d(Z), _d(ic), len(X), _i(lst), len(lst))


def gaussian_log_likelihood_wavelet(w, data, model, mlist, invC9):
    """data, model: [3, Npix]; invC9: [9, Npix] (row-major 3x3 planes)."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    m = np.ascontiguousarray(model, dtype=np.float64)
    ic = np.ascontiguousarray(invC9, dtype=np.float64)
    ml = np.ascontiguousarray(mlist, dtype=np.int32)
    npix = d.shape[1]
    return lib().orc_gaussian_log_likelihood_wavelet(w.ref, _d(d[0]), _d(d[1]), _d(d[2]), _d(m[0]),
                                                     _d(m[1]), _d(m[2]), _i(ml), len(ml), _d(ic),
                                                     npix)


def wdm_eval_batch(w, orb, Tobs, params, data, invC9, t0=0.0, nthreads=0, layers=0):
    """(dlogL, snr, nwave, seconds); layers > 0 forces that layer count."""
    t, _, coef = orb
    p = np.ascontiguousarray(params, dtype=np.float64)
    d = np.ascontiguousarray(data, dtype=np.float64)
    ic = np.ascontiguousarray(invC9, dtype=np.float64)
    n = p.shape[0]
    dl = np.zeros(n)
    sn = np.zeros(n)
    nw = np.zeros(n, dtype=np.int32)
    lib().orc_set_wavelet_layers(int(layers))
    try:
        secs = lib().orc_wdm_eval_batch(w.ref, len(t), _d(t), _d(coef), float(Tobs), float(t0), n, _d(p),
                                        _d(d), _d(ic), d.shape[1], _d(dl), _d(sn), _i(nw), nthreads)
    finally:
        lib().orc_set_wavelet_layers(0)
    return dl, sn, nw, secs


# ---------------------------------------------------------- noise model ---

def _noise_signatures(L):
    _up = ctypes.POINTER(ctypes.c_uint)
    d, i = ctypes.c_double, ctypes.c_int
    L.orc_instrument_noise_model.argtypes = [d, _dp, i, i, _dp, _dp, _dp]
    L.orc_instrument_noise_model_wavelet.argtypes = [d, _dp, i, d, i, _dp, _dp, _dp]
    L.orc_my_noise_log_likelihood.argtypes = [_dp] * 5 + [i]
    L.orc_my_noise_log_likelihood.restype = d
    L.orc_noise_log_likelihood.argtypes = [_dp] * 5 + [i]
    L.orc_noise_log_likelihood.restype = d
    L.orc_my_noise_log_likelihood_wavelet.argtypes = [_dp] * 5 + [i] * 4
    L.orc_my_noise_log_likelihood_wavelet.restype = d
    L.orc_rand_U_0_1.argtypes = [_up]
    L.orc_rand_U_0_1.restype = d
    L.orc_rand_N_0_1.argtypes = [_up]
    L.orc_rand_N_0_1.restype = d
    L.orc_noise_logl_params.argtypes = [d, _dp, i, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_noise_logl_params.restype = d
    L.orc_noise_mcmc_call.argtypes = [d, _dp, i, _dp, _dp, _dp, _dp, d, _dp, _dp, _up, _dp]
    L.orc_noise_mcmc_call.restype = i
    L.orc_noise_ptmcmc.argtypes = [i, _ip, _dp, _dp, _up, _ip]


def instrument_noise_model(fstar, f, sacc, soms, nch=3, C=None):
    """generate_instrument_noise_model -> C[nch, nch, N] (in place if C given)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    if C is None:
        C = np.zeros((nch, nch, f.size))
    sa = np.ascontiguousarray(sacc, dtype=np.float64)
    so = np.ascontiguousarray(soms, dtype=np.float64)
    lib().orc_instrument_noise_model(fstar, _d(f), f.size, nch, _d(sa), _d(so), _d(C))
    return C


def instrument_noise_model_wavelet(fstar, f, df, ND, sacc, soms, C=None):
    f = np.ascontiguousarray(f, dtype=np.float64)
    if C is None:
        C = np.zeros((3, 3, f.size))
    sa = np.ascontiguousarray(sacc, dtype=np.float64)
    so = np.ascontiguousarray(soms, dtype=np.float64)
    na = lib().orc_instrument_noise_model_wavelet(fstar, _d(f), f.size, df, ND, _d(sa), _d(so), _d(C))
    return C, na


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


def my_noise_log_likelihood(X, Y, Z, invC, logdetC):
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    ld = np.ascontiguousarray(logdetC, dtype=np.float64)
    return lib().orc_my_noise_log_likelihood(_d(_cplx(X)), _d(_cplx(Y)), _d(_cplx(Z)), _d(invC), _d(ld), ld.size)


def noise_log_likelihood(X, Y, Z, invC, logdetC):
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    ld = np.ascontiguousarray(logdetC, dtype=np.float64)
    return lib().orc_noise_log_likelihood(_d(_cplx(X)), _d(_cplx(Y)), _d(_cplx(Z)), _d(invC), _d(ld), ld.size)


def my_noise_log_likelihood_wavelet(X, Y, Z, invC, logdetC, NT, Nlayer, Ntot):
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, Y, Z))
    invC = np.ascontiguousarray(invC, dtype=np.float64)
    ld = np.ascontiguousarray(logdetC, dtype=np.float64)
    return lib().orc_my_noise_log_likelihood_wavelet(_d(X), _d(Y), _d(Z), _d(invC), _d(ld), NT, Nlayer, Ntot,
                                                     ld.size)


def rand_stream(seed, n):
    """n rand_r_U_0_1 draws then n rand_r_N_0_1 draws from one rand_r seed."""
    s = ctypes.c_uint(seed & 0xFFFFFFFF)
    u = np.array([lib().orc_rand_U_0_1(ctypes.byref(s)) for _ in range(n)])
    g = np.array([lib().orc_rand_N_0_1(ctypes.byref(s)) for _ in range(n)])
    return u, g


class NoiseData:
    """One data segment for the noise oracle: f [N], X, Y, Z complex [N],
    Cadd [3, 3, N] or None."""

    def __init__(self, fstar, f, X, Y, Z, Cadd=None):
        self.fstar = float(fstar)
        self.f = np.ascontiguousarray(f, dtype=np.float64)
        self.N = self.f.size
        self.X, self.Y, self.Z = _cplx(X), _cplx(Y), _cplx(Z)
        self.Cadd = None if Cadd is None else np.ascontiguousarray(Cadd, dtype=np.float64)
        self.work = np.zeros(19 * self.N)

    def _ca(self):
        return None if self.Cadd is None else _d(self.Cadd)

    def logl(self, p12):
        p = np.ascontiguousarray(p12, dtype=np.float64)
        return lib().orc_noise_logl_params(self.fstar, _d(self.f), self.N, _d(self.X), _d(self.Y), _d(self.Z),
                                           self._ca(), _d(p), _d(self.work))

    def mcmc_call(self, x, logLx, seed, temperature):
        """One noise_instrument_model_mcmc call; returns (x, logLx, seed, nacc)."""
        xx = np.array(x, dtype=np.float64)
        lg = np.array([logLx], dtype=np.float64)
        s = ctypes.c_uint(int(seed) & 0xFFFFFFFF)
        na = lib().orc_noise_mcmc_call(self.fstar, _d(self.f), self.N, _d(self.X), _d(self.Y), _d(self.Z),
                                       self._ca(), float(temperature), _d(xx), _d(lg), ctypes.byref(s),
                                       _d(self.work))
        return xx, float(lg[0]), s.value, na


def noise_ptmcmc(index, temperature, logL, seeds):
    """noise_ptmcmc over one ladder; returns (index, seeds, acceptance)."""
    idx = np.ascontiguousarray(index, dtype=np.int32).copy()
    t = np.ascontiguousarray(temperature, dtype=np.float64)
    lg = np.ascontiguousarray(logL, dtype=np.float64)
    sd = np.ascontiguousarray(seeds, dtype=np.uint32).copy()
    acc = np.zeros(idx.size, dtype=np.int32)
    lib().orc_noise_ptmcmc(idx.size, _i(idx), _d(t), _d(lg), sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                           _i(acc))
    return idx, sd, acc


# ---------------------------------------------------------- UCB sampler ---

def _sampler_signatures(L):
    _up = ctypes.POINTER(ctypes.c_uint)
    d, i = ctypes.c_double, ctypes.c_int
    L.orc_jacobi_eigen8.argtypes = [_dp, _dp, _dp]
    L.orc_jacobi_eigen8.restype = i
    L.orc_ucb_prior.argtypes = [_dp, _dp, _dp]
    L.orc_ucb_prior.restype = d
    L.orc_ucb_log_prior_volume.argtypes = [_dp, _dp]
    L.orc_ucb_propose.argtypes = [_up, i, _dp, _dp, _dp, _dp, _dp, _dp, d, i, i, i, ctypes.POINTER(FstatProp), _ip,
                                  _ip, _dp, _dp, _dp, _dp]
    L.orc_ucb_accept.argtypes = [_up, d, d, d, d, d, d, d, d]
    L.orc_ucb_accept.restype = i
    L.orc_ucb_ptmcmc.argtypes = [i, _ip, _dp, _dp, _dp, _up, _ip]
    L.orc_adapt_temperature_ladder.argtypes = [i, _dp, _ip, i]


def jacobi_eigen8(F):
    """(evec[8,8], eval[8], sweeps) of the symmetric matrix F."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    ev, la = np.zeros((8, 8)), np.zeros(8)
    sw = lib().orc_jacobi_eigen8(_d(F), _d(ev), _d(la))
    return ev, la, sw


def log_prior_volume(pr):
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    lpv = np.zeros(8)
    lib().orc_ucb_log_prior_volume(_d(pr), _d(lpv))
    return lpv


def ucb_prior(p, pr, lpv=None):
    """(logP, p wrapped) of evaluate_ucb_prior."""
    p = np.array(p, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    lpv = log_prior_volume(pr) if lpv is None else np.ascontiguousarray(lpv, dtype=np.float64)
    return lib().orc_ucb_prior(_d(p), _d(pr), _d(lpv)), p


class FstatProp(ctypes.Structure):
    _fields_ = [("ten", _dp), ("nf", ctypes.c_int), ("nt", ctypes.c_int), ("nph", ctypes.c_int),
                ("d_f", ctypes.c_double), ("maxp", ctypes.c_double), ("qmin", ctypes.c_double)]


def fstat_prop(tensor, d_f, maxp, qmin):
    """(FstatProp, keep-alive array) of a proposal density [nf, nt, nph]."""
    t = np.ascontiguousarray(tensor, dtype=np.float64)
    return FstatProp(ten=_d(t), nf=t.shape[0], nt=t.shape[1], nph=t.shape[2], d_f=float(d_f), maxp=float(maxp),
                     qmin=float(qmin)), t


class UcbProposer:
    """Per-window prior box and the proposal weights/flags of the sampler."""

    def __init__(self, w, T, fix_sky=0, fix_freq=0, fix_fdot=0, snr_min=5.0):
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.T = float(T)
        self.flags = (int(fix_sky), int(fix_freq), int(fix_fdot))
        self.snr_min = float(snr_min)

    def propose(self, seed, sparams, evec, eval_, pr, fs=None):
        """sparams [nlive, 8] (wrapped in place), evec [nlive, 64], eval
        [nlive, 8], pr [8, 2], fs an FstatProp (or None) -> (seed, n, np, y,
        lpx, lpy, lq)."""
        s = ctypes.c_uint(int(seed) & 0xFFFFFFFF)
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        lpv = log_prior_volume(pr)
        n, npr = ctypes.c_int(), ctypes.c_int()
        y = np.zeros(8)
        lpx, lpy, lq = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ev = np.ascontiguousarray(evec, dtype=np.float64)
        la = np.ascontiguousarray(eval_, dtype=np.float64)
        lib().orc_ucb_propose(ctypes.byref(s), sparams.shape[0], _d(sparams), _d(ev), _d(la), _d(pr), _d(lpv),
                              _d(self.w), self.T, *self.flags, None if fs is None else ctypes.byref(fs),
                              ctypes.byref(n), ctypes.byref(npr), _d(y), ctypes.byref(lpx), ctypes.byref(lpy),
                              ctypes.byref(lq))
        return s.value, n.value, npr.value, y, lpx.value, lpy.value, lq.value

    def accept(self, seed, logLy, logLx, snr, temperature, lpx, lpy, lq=0.0):
        s = ctypes.c_uint(int(seed) & 0xFFFFFFFF)
        a = lib().orc_ucb_accept(ctypes.byref(s), logLy, logLx, snr, self.snr_min, temperature, lpx, lpy, lq)
        return s.value, bool(a)


def ucb_ptmcmc(index, temperature, logL, lognorm, seeds):
    idx = np.ascontiguousarray(index, dtype=np.int32).copy()
    t = np.ascontiguousarray(temperature, dtype=np.float64)
    lg = np.ascontiguousarray(logL, dtype=np.float64)
    ln = np.ascontiguousarray(lognorm, dtype=np.float64)
    sd = np.ascontiguousarray(seeds, dtype=np.uint32).copy()
    acc = np.zeros(idx.size, dtype=np.int32)
    lib().orc_ucb_ptmcmc(idx.size, _i(idx), _d(t), _d(lg), _d(ln), sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                         _i(acc))
    return idx, sd, acc


def adapt_temperature_ladder(temperature, acceptance, mcmc):
    t = np.array(temperature, dtype=np.float64)
    a = np.ascontiguousarray(acceptance, dtype=np.int32)
    lib().orc_adapt_temperature_ladder(t.size, _d(t), _i(a), int(mcmc))
    return t


# ---------------------------------------------------------- F-statistic ---

def _fstat_signatures(L):
    d, i = ctypes.c_double, ctypes.c_int
    L.orc_get_Fstat_logL.argtypes = [ctypes.POINTER(Orbit), ctypes.c_char_p, d, d, i, i, _dp, _dp, _dp, _dp, d, d, d,
                                     d, _dp, _dp, _dp]
    L.orc_get_Fstat_logL.restype = i
    L.orc_invert4.argtypes = [_dp, _dp]
    L.orc_ucb_fdot.argtypes = [d, d]
    L.orc_ucb_fdot.restype = d


class FstatData:
    """A data segment for the F-statistic: X, A, E complex [NFFT] from bin
    qmin0, C00 [NFFT]."""

    def __init__(self, T, qmin0, X, A, E, C00, t0=0.0, fmt="sangria", orbit=None):
        self.orbit = orbit or analytic_orbit()
        self.T, self.t0, self.qmin0, self.fmt = float(T), float(t0), int(qmin0), fmt
        self.X = np.ascontiguousarray(X, dtype=np.complex128).view(np.float64)
        self.A = np.ascontiguousarray(A, dtype=np.complex128).view(np.float64)
        self.E = np.ascontiguousarray(E, dtype=np.complex128).view(np.float64)
        self.C00 = np.ascontiguousarray(C00, dtype=np.float64)
        self.NFFT = self.C00.size

    def logl(self, f0, fdot, theta, phi):
        """(logL_X, logL_AE, Fparams[4], filter length) of get_Fstat_logL."""
        lx, la = ctypes.c_double(), ctypes.c_double()
        fp = np.zeros(4)
        M = lib().orc_get_Fstat_logL(ctypes.byref(self.orbit), self.fmt.encode(), self.T, self.t0, self.NFFT,
                                     self.qmin0, _d(self.X), _d(self.A), _d(self.E), _d(self.C00), float(f0),
                                     float(fdot), float(theta), float(phi), ctypes.byref(lx), ctypes.byref(la),
                                     _d(fp))
        return lx.value, la.value, fp, M


def invert4(m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros((4, 4))
    lib().orc_invert4(_d(m), _d(out))
    return out


def ucb_fdot(Mc, f0):
    return lib().orc_ucb_fdot(float(Mc), float(f0))


def fstat_proposal(fd, fmin, fmax, n_theta=20, n_phi=20):
    """build_fstatistic_proposal on the oracle: (tensor [n_f, n_theta,
    n_phi], norm, maxp, n_f, d_f)."""
    T = fd.T
    n_f = int((fmax - fmin) * T)
    d_f = int((fmax - fmin) * T) / float(n_f) / T
    d_theta, d_phi = 2.0 / n_theta, 6.283185307179586 / n_phi
    ten = np.zeros((n_f, n_theta, n_phi))
    d_fdot = (1.1 - 0.1) / 10
    for i in range(n_f):
        f = fmin + i * d_f
        f += d_f / 2
        for j in range(n_theta):
            for k in range(n_phi):
                theta = np.arccos(-1.0 + j * d_theta + d_theta / 2)
                phi = k * d_phi + d_phi / 2
                la = fd.logl(f, 0.0, theta, phi)[1]
                lmax = la
                if la > 1:
                    for n in range(10):
                        la = fd.logl(f, ucb_fdot(0.1 + n * d_fdot, f), theta, phi)[1]
                        if la > lmax:
                            lmax = la
                        if la < 1:
                            break
                ten[i, j, k] = lmax * lmax
    norm = 0.0
    for v in ten.reshape(-1):
        norm += v
    norm = (n_f * n_theta * n_phi) / norm
    ten *= norm
    return ten, norm, float(ten.max()), n_f, d_f


def ucb_bandwidth(f, fdot, costheta, A, T, N, orbit=None):
    """ucb_bandwidth (ucb/src/glass_ucb_waveform.c:698-725)."""
    orbit = orbit or analytic_orbit()
    return lib().orc_ucb_bandwidth(orbit.L, orbit.fstar, float(f), float(fdot), float(costheta), float(A),
                                   float(T), int(N))
