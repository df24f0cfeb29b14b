"""The F-statistic of the UCB F-statistic proposal on the GPU, over
``include/glass_hip_fstat.h`` (get_Fstat_logL, ucb/src/glass_ucb_fstatistic.c:
646-703; build_fstatistic_proposal, ucb/src/glass_ucb_proposal.c:1086-1180).

* get_Fstat_logL: the drop-in on host arrays;
* Fstat: one data segment resident on the device; eval() of many grid
  cells (f0, fdot, theta, phi) per launch, proposal() builds the proposal
  density over a band.
"""
import ctypes

import numpy as np

from . import _f64, check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i, _d = ctypes.c_int, ctypes.c_double

_SIGS = [
    ("glass_hip_get_Fstat_logL", [_vp, ctypes.c_char_p, _d, _d, _i, _i, _dp, _dp, _dp, _dp, _d, _d, _d, _d, _dp, _dp,
                                  _dp], _i),
    ("glass_hip_fstat_create", [_vp, _vp, ctypes.c_char_p, _d, _d, _i, _i, _dp, _dp, _dp, _dp, _i, _vp], _i),
    ("glass_hip_fstat_destroy", [_vp], _i),
    ("glass_hip_fstat_eval", [_vp, _i, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fstat_proposal", [_vp, _d, _d, _i, _i, _dp, _ip, _dp, _dp, _dp, _vp], _i),
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


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def get_Fstat_logL(orbit, fmt, T, t0, qmin0, X, A, E, C00, f0, fdot, theta, phi):
    """(logL_X, logL_AE, Fparams[4]) of one cell; X, A, E complex [NFFT]."""
    Xa, Aa, Ea = _f64(X), _f64(A), _f64(E)
    C = np.ascontiguousarray(C00, dtype=np.float64)
    lx, la = ctypes.c_double(), ctypes.c_double()
    fp = np.zeros(4)
    check(_L().glass_hip_get_Fstat_logL(ctypes.cast(ctypes.byref(orbit), _vp), fmt.encode(), float(T), float(t0),
                                        C.size, int(qmin0), _p(Xa), _p(Aa), _p(Ea), _p(C), float(f0), float(fdot),
                                        float(theta), float(phi), ctypes.byref(lx), ctypes.byref(la), _p(fp)),
          "get_Fstat_logL")
    return lx.value, la.value, fp


class Fstat:
    """One data segment (X, A, E complex [NFFT] from bin qmin0; C00 [NFFT])
    resident on the device, for up to max_cells cells per eval."""

    def __init__(self, orbit, fmt, T, t0, qmin0, X, A, E, C00, max_cells=1 << 20):
        import torch
        self.torch = torch
        Xa, Aa, Ea = _f64(X), _f64(A), _f64(E)
        C = np.ascontiguousarray(C00, dtype=np.float64)
        self.NFFT, self.T, self.max_cells = C.size, float(T), int(max_cells)
        self.h = _vp()
        check(_L().glass_hip_fstat_create(ctypes.byref(self.h), ctypes.cast(ctypes.byref(orbit), _vp), fmt.encode(),
                                          float(T), float(t0), C.size, int(qmin0), _p(Xa), _p(Aa), _p(Ea), _p(C),
                                          self.max_cells, _stream()), "fstat_create")

    def close(self):
        if getattr(self, "h", None):
            _L().glass_hip_fstat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def eval(self, cells):
        """cells: CUDA float64 [n, 4] (f0, fdot, theta, phi) -> device
        (logL_X [n], logL_AE [n], Fparams [n, 4])."""
        t = self.torch
        cells = cells.contiguous()
        n = cells.shape[0]
        lx = t.empty(n, dtype=t.float64, device=cells.device)
        la = t.empty(n, dtype=t.float64, device=cells.device)
        fp = t.empty((n, 4), dtype=t.float64, device=cells.device)
        check(_L().glass_hip_fstat_eval(self.h, int(n), _vp(cells.data_ptr()), _vp(lx.data_ptr()),
                                        _vp(la.data_ptr()), _vp(fp.data_ptr()), _stream()), "fstat_eval")
        return lx, la, fp

    def proposal(self, fmin, fmax, n_theta=20, n_phi=20):
        """build_fstatistic_proposal -> (tensor [n_f, n_theta, n_phi], norm,
        maxp, n_f, d_f)."""
        nf, df = ctypes.c_int(), ctypes.c_double()
        check(_L().glass_hip_fstat_proposal(self.h, float(fmin), float(fmax), int(n_theta), int(n_phi), None,
                                            ctypes.byref(nf), ctypes.byref(df), None, None, _stream()),
              "fstat_proposal")
        ten = np.zeros((nf.value, n_theta, n_phi))
        norm, maxp = ctypes.c_double(), ctypes.c_double()
        check(_L().glass_hip_fstat_proposal(self.h, float(fmin), float(fmax), int(n_theta), int(n_phi), _p(ten),
                                            ctypes.byref(nf), ctypes.byref(df), ctypes.byref(norm),
                                            ctypes.byref(maxp), _stream()), "fstat_proposal")
        return ten, norm.value, maxp.value, nf.value, df.value
