"""LDC HDF5 readers (``libglass_ldc.so``, C ABI ``include/glass_ldc.h``).

Host code: the two readers the reference's ReadHDF5
(utils/src/glass_data.c:654-660) calls, returning the time-domain TDI
channels and the sample spacing.  Errors raise ``LdcError`` where the
reference would read garbage or crash.
"""
import ctypes
import os
from collections import namedtuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglass_ldc.so")

_dp = ctypes.POINTER(ctypes.c_double)

# struct TDI's time-domain fields as the readers fill them
# (utils/src/glass_lisa.h:116-144): X, Y, Z, A, E, T, delta, N
Tdi = namedtuple("Tdi", "X Y Z A E T delta N")


class LdcError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LdcError(f"{LIB_PATH} is not built (python -m glass_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        c, i, ip = ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)
        L.glass_ldc_last_error.restype = c
        L.glass_ldc_tdi_samples.argtypes = [c, c, ip]
        L.glass_ldc_read_tdi.argtypes = [c, c, i] + [_dp] * 7
        L.glass_ldc_radler_samples.argtypes = [c, ip]
        L.glass_ldc_read_radler_tdi.argtypes = [c, i] + [_dp] * 7
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise LdcError(f"{what}: {lib().glass_ldc_last_error().decode()} (code {rc})")


def _fill(call, N):
    out = [np.empty(N) for _ in range(6)]
    delta = ctypes.c_double()
    ptrs = [a.ctypes.data_as(_dp) for a in out]
    call(N, ptrs, ctypes.byref(delta))
    return Tdi(*out, delta.value, N)


def read_tdi(fileName, dataName="/obs/tdi"):
    """LISA_Read_HDF5_LDC_TDI (utils/src/glass_lisa.c:1394-1454): the
    compound {t, X, Y, Z} dataset dataName of fileName ("/obs/tdi" for
    Sangria data, "/sky/<component>/tdi" for its sky parts)."""
    f, d = os.fsencode(fileName), dataName.encode()
    n = ctypes.c_int()
    _check(lib().glass_ldc_tdi_samples(f, d, ctypes.byref(n)), "glass_ldc_tdi_samples")

    def call(N, p, delta):
        _check(lib().glass_ldc_read_tdi(f, d, N, *p, delta), "glass_ldc_read_tdi")
    return _fill(call, n.value)


def read_radler_tdi(fileName):
    """LISA_Read_HDF5_LDC_RADLER_TDI (utils/src/glass_lisa.c:1456-1506):
    columns 1..3 of "/H5LISA/PreProcess/TDIdata", the first Nrow/2 rows."""
    f = os.fsencode(fileName)
    n = ctypes.c_int()
    _check(lib().glass_ldc_radler_samples(f, ctypes.byref(n)), "glass_ldc_radler_samples")

    def call(N, p, delta):
        _check(lib().glass_ldc_read_radler_tdi(f, N, *p, delta), "glass_ldc_read_radler_tdi")
    return _fill(call, n.value)
