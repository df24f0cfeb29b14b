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


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
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
        _lib = L
    return _lib


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


def eval_batch(chainset, T, chain_of, sid_of, params, t0=0.0, nthreads=0, orbit=None):
    orbit = orbit or analytic_orbit()
    chain_of = np.ascontiguousarray(chain_of, dtype=np.int32)
    sid_of = np.ascontiguousarray(sid_of, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(len(chain_of))
    secs = lib().orc_eval_batch(ctypes.byref(orbit), T, t0, chainset.arr, len(chain_of),
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
