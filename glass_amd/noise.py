"""Instrument-noise model, noise likelihoods and the batched instrument-noise
sampler on the GPU, over ``include/glass_hip_noise.h``.

Mirrors the reference's noise package (noise/src/glass_noise_model.c,
noise/src/glass_noise_sampler.c):

* instrument_noise_model, instrument_noise_model_wavelet: the covariance
  C[nch, nch, N] of generate_instrument_noise_model(_wavelet) for link
  parameters sacc[6], soms[6] (link order 12, 21, 23, 32, 13, 31);
* my_noise_log_likelihood, noise_log_likelihood,
  my_noise_log_likelihood_wavelet: drop-ins on host arrays;
* NoisePlan: many chains on one data segment -- the fused model, inverse and
  likelihood of many parameter sets per launch, and the whole
  noise_instrument_model_mcmc / noise_ptmcmc sweep on the device.

Parameter vectors are 12 doubles: sacc[6] then soms[6].
"""
import ctypes

import numpy as np

from . import GlassHipError, _f64, check, checkd, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_vp = ctypes.c_void_p
_i, _d = ctypes.c_int, ctypes.c_double

# initialize_instrument_model's levels (noise/src/glass_noise_model.c:1392-1393)
SACC_DEFAULT = 5.76e-30
SOMS_DEFAULT = 1.28e-22

_SIGS = [
    ("glass_hip_instrument_noise_model", [_d, _dp, _i, _i, _dp, _dp, _dp], _i),
    ("glass_hip_instrument_noise_model_wavelet", [_d, _dp, _i, _d, _i, _dp, _dp, _dp, _ip], _i),
    ("glass_hip_my_noise_log_likelihood", [_dp, _dp, _dp, _dp, _dp, _i], _d),
    ("glass_hip_noise_log_likelihood", [_dp, _dp, _dp, _dp, _dp, _i], _d),
    ("glass_hip_my_noise_log_likelihood_wavelet", [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _i], _d),
    ("glass_hip_noise_plan_create", [_vp, _i, _i, _i, _d, _dp, _dp, _dp, _dp, _dp, _vp], _i),
    ("glass_hip_noise_plan_destroy", [_vp], _i),
    ("glass_hip_noise_logl", [_vp, _i, _vp, _vp, _vp], _i),
    ("glass_hip_noise_set_state", [_vp, _dp, _up, _dp, _vp], _i),
    ("glass_hip_noise_get_state", [_vp, _dp, _dp, _up, _ip, _ip, _ip, _vp], _i),
    ("glass_hip_noise_mcmc", [_vp, _i, _i, _i, _vp], _i),
    ("glass_hip_noise_ptswap", [_vp, _vp], _i),
    ("glass_hip_noise_plan_info", [_vp, _ip, _ip, ctypes.POINTER(ctypes.c_size_t)], _i),
    ("glass_hip_rand_stream", [ctypes.c_uint, _i, _dp, _dp], _i),
]
_declared = False


def _L():
    global _declared
    L = lib()
    if not _declared:
        for name, args, res in _SIGS:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _declared = True
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def default_links():
    """sacc[6], soms[6] at the reference's default levels."""
    return np.full(6, SACC_DEFAULT), np.full(6, SOMS_DEFAULT)


def instrument_noise_model(fstar, f, sacc, soms, nch=3, C=None):
    """C[nch, nch, N] of generate_instrument_noise_model.  C (if given) is
    updated in place and keeps the entries the reference does not write."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    N = f.size
    if C is None:
        C = np.zeros((nch, nch, N))
    if C.shape != (nch, nch, N) or C.dtype != np.float64 or not C.flags.c_contiguous:
        raise ValueError("C must be a C-contiguous float64 array of shape (nch, nch, N)")
    sa = np.ascontiguousarray(sacc, dtype=np.float64)
    so = np.ascontiguousarray(soms, dtype=np.float64)
    if sa.size != 6 or so.size != 6:
        raise ValueError("sacc and soms hold 6 link parameters each")
    check(_L().glass_hip_instrument_noise_model(float(fstar), _p(f), N, int(nch), _p(sa), _p(so), _p(C)),
          "instrument_noise_model")
    return C


def instrument_noise_model_wavelet(fstar, f, df, ND, sacc, soms, C=None):
    """(C[3, 3, N], number of active layers) of
    generate_instrument_noise_model_wavelet."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    N = f.size
    if C is None:
        C = np.zeros((3, 3, N))
    sa = np.ascontiguousarray(sacc, dtype=np.float64)
    so = np.ascontiguousarray(soms, dtype=np.float64)
    na = ctypes.c_int()
    check(_L().glass_hip_instrument_noise_model_wavelet(float(fstar), _p(f), N, float(df), int(ND), _p(sa),
                                                        _p(so), _p(C), ctypes.byref(na)),
          "instrument_noise_model_wavelet")
    return C, na.value


def _given(fn, X, Y, Z, invC, logdetC, N, what):
    X, Y, Z = _f64(X), _f64(Y), _f64(Z)
    iC = np.ascontiguousarray(invC, dtype=np.float64).reshape(-1)
    ld = np.ascontiguousarray(logdetC, dtype=np.float64)
    if X.size < 2 * N or Y.size < 2 * N or Z.size < 2 * N or iC.size < 9 * N or ld.size < N:
        raise GlassHipError(f"{what}: arrays shorter than N")
    return checkd(fn(_p(X), _p(Y), _p(Z), _p(iC), _p(ld), int(N)), what)


def my_noise_log_likelihood(X, Y, Z, invC, logdetC):
    """my_noise_log_likelihood; X, Y, Z complex [N], invC [3, 3, N]."""
    N = np.asarray(logdetC).size
    return _given(_L().glass_hip_my_noise_log_likelihood, X, Y, Z, invC, logdetC, N, "my_noise_log_likelihood")


def noise_log_likelihood(X, Y, Z, invC, logdetC):
    """noise_log_likelihood (three channels)."""
    N = np.asarray(logdetC).size
    return _given(_L().glass_hip_noise_log_likelihood, X, Y, Z, invC, logdetC, N, "noise_log_likelihood")


def my_noise_log_likelihood_wavelet(X, Y, Z, invC, logdetC, NT, Nlayer, Ntot):
    """my_noise_log_likelihood_wavelet over NT * Nlayer real pixels; invC
    [3, 3, ldc], logdetC [ldc]."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    iC = np.ascontiguousarray(invC, dtype=np.float64)
    ld = np.ascontiguousarray(logdetC, dtype=np.float64)
    ldc = ld.size
    n = NT * Nlayer
    if X.size < n or Y.size < n or Z.size < n or iC.size != 9 * ldc:
        raise GlassHipError("my_noise_log_likelihood_wavelet: array sizes")
    return checkd(_L().glass_hip_my_noise_log_likelihood_wavelet(_p(X), _p(Y), _p(Z), _p(iC.reshape(-1)), _p(ld),
                                                                 int(NT), int(Nlayer), int(Ntot), int(ldc)),
                  "my_noise_log_likelihood_wavelet")


def rand_stream(seed, n):
    """(uniforms, normals) of the device's rand_r restatement from one seed:
    n rand_r_U_0_1 draws, then n rand_r_N_0_1 draws continuing the stream."""
    u, g = np.zeros(n), np.zeros(n)
    check(_L().glass_hip_rand_stream(ctypes.c_uint(seed & 0xFFFFFFFF), int(n), _p(u), _p(g)), "rand_stream")
    return u, g


class NoisePlan:
    """Chains of the instrument-noise model on one data segment.

    f [N] bin frequencies (all > 0), X, Y, Z complex [N] data, Cadd the
    fixed extra covariance [3, 3, N] (or None), nladder ladders of ntemp
    temperature slots.
    """

    def __init__(self, fstar, f, X, Y, Z, nladder, ntemp, Cadd=None, stream=None):
        import torch
        self.torch = torch
        self.N = int(np.asarray(f).size)
        self.nladder, self.ntemp = int(nladder), int(ntemp)
        self.nslot = self.nladder * self.ntemp
        self._stream = stream
        f = np.ascontiguousarray(f, dtype=np.float64)
        Xa, Ya, Za = _f64(X), _f64(Y), _f64(Z)
        ca = None if Cadd is None else np.ascontiguousarray(Cadd, dtype=np.float64).reshape(-1)
        self.h = ctypes.c_void_p()
        check(_L().glass_hip_noise_plan_create(ctypes.byref(self.h), self.N, self.nladder, self.ntemp, float(fstar),
                                               _p(f), _p(Xa), _p(Ya), _p(Za), None if ca is None else _p(ca),
                                               self._sp()), "noise_plan_create")

    def _sp(self):
        s = self._stream if self._stream is not None else self.torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def close(self):
        if self.h:
            _L().glass_hip_noise_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def info(self):
        nc, ch, by = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        check(_L().glass_hip_noise_plan_info(self.h, ctypes.byref(nc), ctypes.byref(ch), ctypes.byref(by)),
              "noise_plan_info")
        return {"nchunk": nc.value, "chunk": ch.value, "bytes": by.value}

    def logl(self, params, out=None):
        """my_noise_log_likelihood of the model (+ Cadd) of each parameter
        set; params a device float64 tensor [n, 12]; returns a device
        tensor [n] (out, if given)."""
        torch = self.torch
        if params.dtype != torch.float64 or params.dim() != 2 or params.shape[1] != 12 or not params.is_cuda:
            raise ValueError("params must be a CUDA float64 tensor of shape (n, 12)")
        params = params.contiguous()
        n = params.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device=params.device)
        elif out.dtype != torch.float64 or out.numel() < n or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 tensor of at least n elements")
        check(_L().glass_hip_noise_logl(self.h, int(n), ctypes.c_void_p(params.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr()), self._sp()), "noise_logl")
        return out

    def set_state(self, params, seeds, temperature):
        """params [nslot, 12] of the states, rand_r seeds and temperatures
        [nslot] of the slots (slot s = ladder * ntemp + t)."""
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(self.nslot, 12)
        s = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(self.nslot)
        t = np.ascontiguousarray(temperature, dtype=np.float64).reshape(self.nslot)
        check(_L().glass_hip_noise_set_state(self.h, _p(p), s.ctypes.data_as(_up), _p(t), self._sp()),
              "noise_set_state")

    def state(self):
        S = self.nslot
        p, lg = np.zeros((S, 12)), np.zeros(S)
        sd = np.zeros(S, dtype=np.uint32)
        idx, na, sw = (np.zeros(S, dtype=np.int32) for _ in range(3))
        check(_L().glass_hip_noise_get_state(self.h, _p(p), _p(lg), sd.ctypes.data_as(_up), idx.ctypes.data_as(_ip),
                                             na.ctypes.data_as(_ip), sw.ctypes.data_as(_ip), self._sp()),
              "noise_get_state")
        return {"params": p, "logL": lg, "seeds": sd, "index": idx, "nacc": na, "swaps": sw}

    def mcmc(self, ncalls, swap=True, graph=False):
        """ncalls calls of noise_instrument_model_mcmc per slot, then one
        noise_ptmcmc pass per ladder if swap (asynchronous)."""
        check(_L().glass_hip_noise_mcmc(self.h, int(ncalls), 1 if swap else 0, 1 if graph else 0, self._sp()),
              "noise_mcmc")

    def ptswap(self):
        check(_L().glass_hip_noise_ptswap(self.h, self._sp()), "noise_ptswap")
