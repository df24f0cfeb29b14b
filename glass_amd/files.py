"""The reference's UCB file formats, over ``include/glass_hip_io.h``.

Writers and readers with the reference's printf / scanf formats (files are
byte-identical to the reference's for the same values):

* print_source_params / scan_source_params (ucb/src/glass_ucb_io.c:598-636),
  print_chain_state (utils/src/glass_sampler.c:153-161), print_psd_state
  (ucb/src/glass_ucb_io.c:580-589), the likelihood / temperature lines of
  print_chain_files (ucb/src/glass_ucb_io.c:494-509);
* save_chain_state / restore_chain_state (ucb/src/glass_ucb_io.c:418-490), per
  chain and for every chain of a batched plan (its device state);
* the verification-binary file (GetVerificationBinary,
  ucb/src/glass_ucb_data.c:627-673) and the injection source file
  (UCBInjectSimulatedSource, ucb/src/glass_ucb_data.c:270-384).

Parameters are the reference's params[8] layout (f0 T, cos theta, phi, ln A,
cos i, psi, phi0, fdot T^2).
"""
import ctypes

import numpy as np

from . import check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p
_i, _d = ctypes.c_int, ctypes.c_double

_SIGS = [
    ("glass_hip_format_source_params", [_dp, _d, _cp, _i], _i),
    ("glass_hip_format_chain_state", [_i, _i, _d, _d, _d, _cp, _i], _i),
    ("glass_hip_format_psd_state", [_i, _d, _d, _dp, _i, _cp, _i], _i),
    ("glass_hip_format_value_line", [_i, _dp, _i, _cp, _i], _i),
    ("glass_hip_scan_source_params", [_cp, _d, _dp, _ip], _i),
    ("glass_hip_read_verification_binary", [_cp, _d, _i, ctypes.c_uint, _dp, _ip, _dp, _dp], _i),
    ("glass_hip_read_injection_file", [_cp, _d, _i, _dp, _ip], _i),
    ("glass_hip_write_chain_state_file", [_cp, _d, _i, _i, _d, _d, _d, _dp, _i, _dp, _d], _i),
    ("glass_hip_read_chain_state_file", [_cp, _i, _i, _d, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _dp], _i),
    ("glass_hip_fd_batch_save_chain_states", [_vp, _cp, _i, _d, _dp, _d, _dp, _i], _i),
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


def _text(fn, *args, cap=4096):
    buf = ctypes.create_string_buffer(cap)
    n = fn(*args, buf, cap)
    if n < 0:
        check(-1, fn.__name__)
    return buf.value.decode()


def print_source_params(params, T):
    """The source's line without its newline."""
    p = np.ascontiguousarray(params, dtype=np.float64).reshape(8)
    return _text(_L().glass_hip_format_source_params, _p(p), float(T))


def print_chain_state(step, nlive, logL, logLnorm, t0):
    return _text(_L().glass_hip_format_chain_state, int(step), int(nlive), float(logL), float(logLnorm), float(t0))


def print_psd_state(step, logL, logLnorm, eta):
    e = np.ascontiguousarray(eta, dtype=np.float64).reshape(-1)
    return _text(_L().glass_hip_format_psd_state, int(step), float(logL), float(logLnorm), _p(e), len(e))


def value_line(step, values):
    """A line of the likelihood (logL + logLnorm per chain) or temperature
    (1/T per chain) file."""
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    return _text(_L().glass_hip_format_value_line, int(step), _p(v), len(v), cap=32 + 32 * len(v))


def scan_source_params(text, T):
    """-> (params[8], characters consumed)."""
    p = np.zeros(8)
    n = ctypes.c_int()
    check(_L().glass_hip_scan_source_params(text.encode(), float(T), _p(p), ctypes.byref(n)), "scan_source_params")
    return p, n.value


def read_verification_binary(path, T, NFFT, iseed):
    """-> (params[8], qmin, fmin, fmax)."""
    p = np.zeros(8)
    q = ctypes.c_int()
    lo, hi = ctypes.c_double(), ctypes.c_double()
    check(_L().glass_hip_read_verification_binary(str(path).encode(), float(T), int(NFFT), int(iseed), _p(p),
                                                  ctypes.byref(q), ctypes.byref(lo), ctypes.byref(hi)),
          "read_verification_binary")
    return p, q.value, lo.value, hi.value


def read_injection_file(path, T, max_sources=4096):
    """-> params[n, 8] (n as the reference counts the file's sources)."""
    p = np.zeros((max_sources, 8))
    n = ctypes.c_int()
    check(_L().glass_hip_read_injection_file(str(path).encode(), float(T), int(max_sources), _p(p),
                                             ctypes.byref(n)), "read_injection_file")
    return p[:min(n.value, max_sources)].copy()


def save_chain_state(path, logLmax, step, logL, logLnorm, t0, eta, params, T):
    """One chain's checkpoint file; params[nlive, 8]."""
    e = np.ascontiguousarray(eta, dtype=np.float64).reshape(-1)
    p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 8)
    check(_L().glass_hip_write_chain_state_file(str(path).encode(), float(logLmax), int(step), len(p), float(logL),
                                                float(logLnorm), float(t0), _p(e), len(e), _p(p), float(T)),
          "write_chain_state_file")


def restore_chain_state(path, T, neta, maxsrc=64):
    """-> dict(logLmax, step, nlive, logL, logLnorm, t0, eta, params[nlive, 8])."""
    eta = np.zeros(max(neta, 1))
    p = np.zeros((maxsrc, 8))
    lm, lg, ln, t0 = (ctypes.c_double() for _ in range(4))
    st, nl = ctypes.c_int(), ctypes.c_int()
    check(_L().glass_hip_read_chain_state_file(str(path).encode(), int(maxsrc), int(neta), float(T),
                                               ctypes.byref(lm), ctypes.byref(st), ctypes.byref(nl),
                                               ctypes.byref(lg), ctypes.byref(ln), ctypes.byref(t0), _p(eta),
                                               _p(p)), "read_chain_state_file")
    return {"logLmax": lm.value, "step": st.value, "nlive": nl.value, "logL": lg.value, "logLnorm": ln.value,
            "t0": t0.value, "eta": eta[:neta].copy(), "params": p[:nl.value].copy()}


def save_chain_states(plan, directory, step, logLmax, logLnorm=None, t0=0.0, eta=None):
    """chain_state_<c>.dat for every chain of a batched plan (FdBatch), from
    its device state."""
    ln = None if logLnorm is None else np.ascontiguousarray(logLnorm, dtype=np.float64)
    e = np.zeros(0) if eta is None else np.ascontiguousarray(eta, dtype=np.float64).reshape(-1)
    check(_L().glass_hip_fd_batch_save_chain_states(plan.h, str(directory).encode(), int(step), float(logLmax),
                                                    None if ln is None else _p(ln), float(t0),
                                                    _p(e) if len(e) else None, len(e)),
          "fd_batch_save_chain_states")
