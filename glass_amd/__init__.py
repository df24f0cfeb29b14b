"""glass_amd -- MI355X (gfx950) implementation of GLASS's galactic-binary
(UCB) likelihood path.

The compute lives in ``libglass_hip.so`` (hand-written HIP kernels behind a C
ABI, ``include/glass_hip.h``).  This package is the host-side mirror of the
reference's C interface for that path: the same function names, the same
argument meaning, and an exception where the reference would print and exit.
There is no CPU fallback: if the library is missing or no GPU is present,
every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglass_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class GlassHipError(RuntimeError):
    pass


class Orbit(ctypes.Structure):
    """glass_hip_orbit: analytic LISA orbit constants (L, fstar, R)."""
    _fields_ = [("L", ctypes.c_double), ("fstar", ctypes.c_double), ("R", ctypes.c_double)]


class NumericOrbit(ctypes.Structure):
    """glass_hip_numeric_orbit: an orbit read from a file (knots t [Norb],
    positions x, y, z [3, Norb] each, average arm length L, fstar)."""
    _fields_ = [("Norb", ctypes.c_int), ("t", ctypes.POINTER(ctypes.c_double)),
                ("x", ctypes.POINTER(ctypes.c_double)), ("y", ctypes.POINTER(ctypes.c_double)),
                ("z", ctypes.POINTER(ctypes.c_double)), ("L", ctypes.c_double), ("fstar", ctypes.c_double)]


_lib = None


def _declare(L):
    d, i = ctypes.c_double, ctypes.c_int
    L.glass_hip_analytic_orbit.argtypes = [ctypes.POINTER(Orbit)]
    L.glass_hip_analytic_orbit.restype = None
    L.glass_hip_set_device.argtypes = [i]
    L.glass_hip_last_error.restype = ctypes.c_char_p
    L.glass_hip_ucb_waveform.argtypes = [ctypes.POINTER(Orbit), ctypes.c_char_p, d, d, _dp, i,
                                         _dp, _dp, _dp, _dp, _dp, i, i]
    L.glass_hip_ucb_waveform_numeric.argtypes = [ctypes.POINTER(NumericOrbit), ctypes.c_char_p, d, d, _dp, i,
                                                 _dp, _dp, _dp, _dp, _dp, i, i]
    L.glass_hip_numeric_orbit_armlength.argtypes = [i, _dp, _dp, _dp, _dp, _dp]
    L.glass_hip_fourier_nwip.argtypes = [_dp, _dp, _dp, i]
    L.glass_hip_fourier_nwip.restype = d
    L.glass_hip_wavelet_nwip.argtypes = [_dp, _dp, _dp, _ip, i, i]
    L.glass_hip_wavelet_nwip.restype = d
    L.glass_hip_snr.argtypes = [_dp, _dp, _dp, _dp, i, i]
    L.glass_hip_snr.restype = d
    L.glass_hip_gaussian_log_likelihood.argtypes = [_dp] * 11 + [i, i]
    L.glass_hip_gaussian_log_likelihood.restype = d
    L.glass_hip_invert_noise_covariance.argtypes = [_dp, _dp, _dp, i]
    for name, args, res in _EXTRA:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


# (name, argtypes, restype) of the device-pointer entry points, filled by
# the submodules that wrap them
_EXTRA = []


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GlassHipError(
                f"{LIB_PATH} is missing: build it with `python -m glass_amd.build` "
                "(there is no CPU fallback)")
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Loading it first makes this library bind to
        # that one runtime instead of pulling in a second copy from
        # /opt/rocm, which would leave two HIP runtimes in one process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def last_error():
    return lib().glass_hip_last_error().decode()


def check(rc, what):
    if rc != 0:
        raise GlassHipError(f"{what}: {last_error()} (status {rc})")


def checkd(v, what):
    if v != v:  # NaN is the drop-ins' error value
        msg = last_error()
        if msg:
            raise GlassHipError(f"{what}: {msg}")
    return v


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    a = np.ascontiguousarray(a)
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)
    return np.ascontiguousarray(a, dtype=np.float64)


def analytic_orbit():
    """initialize_analytic_orbit (utils/src/glass_lisa.c:125-136)."""
    o = Orbit()
    lib().glass_hip_analytic_orbit(ctypes.byref(o))
    return o


def set_device(device):
    check(lib().glass_hip_set_device(int(device)), "set_device")


def ucb_waveform(orbit, format, T, t0, params, NParams, BW, NI=3):
    """ucb_waveform (ucb/src/glass_ucb_waveform.c:738-1001).

    Returns X, Y, Z, A, E as complex128 arrays of BW template-local bins
    (bin m <-> absolute bin q - BW/2 + m).  Raises GlassHipError for an
    unsupported format (the reference exits)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros((5, 2 * BW))
    rc = lib().glass_hip_ucb_waveform(ctypes.byref(orbit), format.encode(), float(T), float(t0),
                                      _d(p), int(NParams), *[_d(out[c]) for c in range(5)],
                                      int(BW), int(NI))
    check(rc, "ucb_waveform")
    return out.view(np.complex128)


def numeric_orbit(t, x, y, z):
    """A NumericOrbit over host arrays (kept alive on the returned object):
    knots t [Norb], positions x, y, z [3, Norb] (row i = spacecraft i); L
    and fstar as initialize_numeric_orbit computes them."""
    t, x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (t, x, y, z))
    N = len(t)
    if x.shape != (3, N) or y.shape != (3, N) or z.shape != (3, N):
        raise ValueError("x, y, z must be [3, Norb]")
    Lv, fs = ctypes.c_double(), ctypes.c_double()
    check(lib().glass_hip_numeric_orbit_armlength(N, _d(x), _d(y), _d(z), ctypes.byref(Lv), ctypes.byref(fs)),
          "numeric_orbit_armlength")
    o = NumericOrbit(N, _d(t), _d(x), _d(y), _d(z), Lv.value, fs.value)
    o._keep = (t, x, y, z)
    return o


def ucb_waveform_numeric(orbit, format, T, t0, params, NParams, BW, NI=3):
    """ucb_waveform with a numeric orbit (interpolate_orbits); as ucb_waveform."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros((5, 2 * BW))
    rc = lib().glass_hip_ucb_waveform_numeric(ctypes.byref(orbit), format.encode(), float(T), float(t0),
                                              _d(p), int(NParams), *[_d(out[c]) for c in range(5)],
                                              int(BW), int(NI))
    check(rc, "ucb_waveform_numeric")
    return out.view(np.complex128)


def fourier_nwip(a, b, invC, N):
    """fourier_nwip (utils/src/glass_math.c:362-380)."""
    a, b, invC = _f64(a), _f64(b), _f64(invC)
    if a.size < 2 * N or b.size < 2 * N or invC.size < N:
        raise GlassHipError("fourier_nwip: arrays shorter than N")
    return checkd(lib().glass_hip_fourier_nwip(_d(a), _d(b), _d(invC), int(N)), "fourier_nwip")


def wavelet_nwip(a, b, invC, lst, N=None):
    """wavelet_nwip (utils/src/glass_math.c:382-391)."""
    a, b, invC = _f64(a), _f64(b), _f64(invC)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    N = len(lst) if N is None else int(N)
    nmax = min(a.size, b.size, invC.size)
    return checkd(lib().glass_hip_wavelet_nwip(_d(a), _d(b), _d(invC),
                                               lst.ctypes.data_as(_ip), N, nmax), "wavelet_nwip")


def snr(X, Y, Z, invC, N):
    """snr (utils/src/glass_math.c:403-426), 3 channels; invC [3,3,Ninv]."""
    X, Y, Z, ic = _f64(X), _f64(Y), _f64(Z), _f64(invC)
    Ninv = ic.size // 9
    return checkd(lib().glass_hip_snr(_d(X), _d(Y), _d(Z), _d(ic), Ninv, int(N)), "snr")


def gaussian_log_likelihood(data, model, invC, Nchannel=3):
    """gaussian_log_likelihood (utils/src/glass_model.c:726-774).

    data, model: complex arrays [5, NFFT] (X, Y, Z, A, E) or dicts of
    channel arrays; invC [3, 3, NFFT]."""
    def chans(v):
        if isinstance(v, dict):
            return [None if v.get(k) is None else _f64(v[k]) for k in "XYZAE"]
        v = np.ascontiguousarray(v, dtype=np.complex128)
        return [v[c].view(np.float64) for c in range(v.shape[0])] + [None] * (5 - v.shape[0])

    dd, mm = chans(data), chans(model)
    ic = _f64(invC)
    NFFT = ic.size // 9
    ptr = lambda a: None if a is None else _d(a)
    return checkd(lib().glass_hip_gaussian_log_likelihood(*[ptr(a) for a in dd], *[ptr(a) for a in mm],
                                                          _d(ic), NFFT, int(Nchannel)),
                  "gaussian_log_likelihood")


def invert_noise_covariance_matrix(C):
    """invert_noise_covariance_matrix (utils/src/glass_math.c:240-289), 3
    channels: C [3,3,N] -> (invC [3,3,N], logdetC [N])."""
    C = _f64(C)
    N = C.size // 9
    invC = np.zeros((3, 3, N))
    ld = np.zeros(N)
    check(lib().glass_hip_invert_noise_covariance(_d(C), _d(invC), _d(ld), N),
          "invert_noise_covariance_matrix")
    return invC, ld
