"""The allocation-free batched likelihood (glass_hip_logl_batch_eval_noalloc)
against the plan-based glass_hip_logl_batch_eval, through the C ABI: two
calls in a row on the shared static scratch, the second with fewer and
shorter windows, each bitwise equal to the plan's result for the same
items; windows above max_W give NaN; oversize requests are refused."""
import ctypes

import numpy as np
import pytest
import torch

from glass_amd import synth

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _fn(gl):
    L = gl.lib()
    f = L.glass_hip_logl_batch_eval_noalloc
    f.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
    f.restype = ctypes.c_int
    return f


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _noalloc(gl, r, n, max_W, W=None):
    dev = {k: r[k].cuda() for k in ("data", "model", "invC")}
    Wd = torch.tensor(np.asarray(r["W"][:n] if W is None else W, dtype=np.int32), device="cuda")
    off = torch.tensor(np.asarray(r["off"][:n], dtype=np.int64), device="cuda")
    out = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _fn(gl)(n, _ptr(Wd), _ptr(off), _ptr(off), _ptr(off), max_W, _ptr(dev["data"]), _ptr(dev["model"]),
                 _ptr(dev["invC"]), _ptr(out), s)
    assert rc == 0, gl.last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _plan(r, n):
    from glass_amd.batch import LoglBatch
    B = LoglBatch(r["W"][:n], r["off"][:n], r["off"][:n], r["off"][:n])
    got = B.eval(r["data"].cuda(), r["model"].cuda(), r["invC"].cuda()).cpu().numpy()
    B.close()
    return got


def test_noalloc_equals_plan_eval_twice(gl):
    windows = (1, 17, 512, 2047, 2048, 2049, 4096, 6000)
    r = synth.logl_batch_inputs(48, seed=31, windows=windows)
    print("noalloc call 1: 48 items, max_W 6000", flush=True)
    a = _noalloc(gl, r, 48, 6000)
    assert np.array_equal(a, _plan(r, 48))
    # second call on the same scratch: fewer items, smaller max_W, so the
    # partials of the first call sit beyond what this one reads and writes
    r2 = synth.logl_batch_inputs(20, seed=32, windows=(1, 300, 2047, 2048, 2049))
    print("noalloc call 2: 20 items, max_W 2049", flush=True)
    b = _noalloc(gl, r2, 20, 2049)
    assert np.array_equal(b, _plan(r2, 20))
    assert np.all(np.isfinite(b))


def test_noalloc_window_above_max_w_is_nan(gl):
    r = synth.logl_batch_inputs(6, seed=33, windows=(100, 3000))
    W = np.asarray(r["W"][:6], dtype=np.int32)
    got = _noalloc(gl, r, 6, 2048, W)
    ref = _plan(r, 6)
    for e in range(6):
        if W[e] > 2048:
            assert np.isnan(got[e])
        else:
            assert got[e] == ref[e]


def test_noalloc_refuses_oversize(gl):
    f = _fn(gl)
    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    i = torch.zeros(16, dtype=torch.int64, device="cuda")
    w = torch.zeros(16, dtype=torch.int32, device="cuda")
    # 2^20 partials: 1025 items of 2^21 bins need 1025 * 1024 chunks
    rc = f(1025, _ptr(w), _ptr(i), _ptr(i), _ptr(i), 1 << 21, _ptr(x), _ptr(x), _ptr(x), _ptr(x), None)
    assert rc == -1 and "static scratch" in gl.last_error()
    rc = f(0, _ptr(w), _ptr(i), _ptr(i), _ptr(i), 16, _ptr(x), _ptr(x), _ptr(x), _ptr(x), None)
    assert rc == -1
