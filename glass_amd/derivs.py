"""Fisher matrices and exact gradients of the UCB likelihood.

* ucb_fisher: drop-in for the reference's forward-difference ucb_fisher
  (ucb/src/glass_ucb_waveform.c:364-522) on host arrays; the eigensystem the
  reference then takes with LAPACKE_dgeev (utils/src/glass_math.c:475-499) is
  returned by numpy.
* source_derivs: exact L, dL/dtheta and Fisher for many sources on device
  tensors (include/glass_hip_batch.h).
* source_grad: L and dL/dtheta only, from the reverse-mode (adjoint) kernel.
* UcbSourceLogL: torch.autograd.Function whose forward is L(theta) from the
  adjoint kernel and whose backward returns the adjoint gradient;
  ucb_source_logl runs the likelihood-only kernel when no gradient can flow.
"""
import ctypes

import numpy as np

from . import _EXTRA, _f64, check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i, _d, _ll = ctypes.c_int, ctypes.c_double, ctypes.c_longlong

_EXTRA += [
    ("glass_hip_ucb_fisher", [_vp, ctypes.c_char_p, _d, _d, _i, _i, _dp, _dp, _dp], _i),
    ("glass_hip_fd_source_derivs", [_vp, ctypes.c_char_p, _d, _d, _i, _vp, _i, _vp, _i, _vp, _ll, _vp, _ll, _i,
                                    _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_source_grad", [_vp, ctypes.c_char_p, _d, _d, _i, _vp, _i, _vp, _i, _vp, _ll, _vp, _ll, _i,
                                  _vp, _vp, _vp], _i),
]


def _L():
    L = lib()
    for name, args, res in _EXTRA:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def ucb_fisher(orbit, fmt, T, t0, NFFT, data_qmin, params, invC, eigen=True):
    """Forward-difference Fisher matrix of ucb_fisher.  invC: [3,3,NFFT].
    Returns F (8x8) or (F, evalues, evectors) with eigen=True."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    ic = _f64(invC)
    F = np.zeros((8, 8))
    check(_L().glass_hip_ucb_fisher(ctypes.cast(ctypes.byref(orbit), _vp), fmt.encode(), float(T), float(t0),
                                    int(NFFT), int(data_qmin), p.ctypes.data_as(_dp), ic.ctypes.data_as(_dp),
                                    F.ctypes.data_as(_dp)), "ucb_fisher")
    if not eigen:
        return F
    w, v = np.linalg.eigh(F)
    return F, w, v


def source_derivs(orbit, T, params, BW, imin, invC, resid=None, t0=0.0, fmt="sangria", mode=0, stream=None):
    """Exact L, dL/dtheta, Fisher for n sources (torch tensors on the GPU).

    params [n,8] float64; imin [n] int32; invC [6,W] (shared) or [n,6,W];
    resid None or complex128 [3,W] (shared) or [n,3,W] -- the residual
    without the source.  Returns (L [n], grad [n,8], fisher [n,8,8])."""
    import torch
    n = params.shape[0]
    W = invC.shape[-1]
    dev = params.device
    params = params.contiguous().to(torch.float64)
    imin = imin.contiguous().to(torch.int32)
    invC = invC.contiguous().to(torch.float64)
    istride = 0 if invC.dim() == 2 else 6 * W
    if resid is not None:
        resid = torch.view_as_real(resid.contiguous().to(torch.complex128)).contiguous()
        rstride = 0 if resid.dim() == 3 else 6 * W
        rptr = _vp(resid.data_ptr())
    else:
        rstride, rptr = 0, None
    out_L = torch.empty(n, dtype=torch.float64, device=dev)
    grad = torch.empty((n, 8), dtype=torch.float64, device=dev)
    fisher = torch.empty((n, 8, 8), dtype=torch.float64, device=dev)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(_L().glass_hip_fd_source_derivs(ctypes.cast(ctypes.byref(orbit), _vp), fmt.encode(), float(T),
                                          float(t0), int(n), _vp(params.data_ptr()), int(BW),
                                          _vp(imin.data_ptr()), int(W), rptr, rstride, _vp(invC.data_ptr()),
                                          istride, int(mode), _vp(out_L.data_ptr()), _vp(grad.data_ptr()),
                                          _vp(fisher.data_ptr()), _vp(s)), "fd_source_derivs")
    return out_L, grad, fisher


def source_grad(orbit, T, params, BW, imin, invC, resid=None, t0=0.0, fmt="sangria", mode=0, stream=None,
                want_grad=True):
    """L and dL/dtheta for n sources from the adjoint kernel (torch tensors
    on the GPU; arguments as source_derivs, BW 32..1024).  Returns (L [n],
    grad [n,8]) or (L, None) with want_grad=False."""
    import torch
    n = params.shape[0]
    W = invC.shape[-1]
    dev = params.device
    params = params.contiguous().to(torch.float64)
    imin = imin.contiguous().to(torch.int32)
    invC = invC.contiguous().to(torch.float64)
    istride = 0 if invC.dim() == 2 else 6 * W
    if resid is not None:
        resid = torch.view_as_real(resid.contiguous().to(torch.complex128)).contiguous()
        rstride = 0 if resid.dim() == 3 else 6 * W
        rptr = _vp(resid.data_ptr())
    else:
        rstride, rptr = 0, None
    out_L = torch.empty(n, dtype=torch.float64, device=dev)
    grad = torch.empty((n, 8), dtype=torch.float64, device=dev) if want_grad else None
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(_L().glass_hip_fd_source_grad(ctypes.cast(ctypes.byref(orbit), _vp), fmt.encode(), float(T), float(t0),
                                        int(n), _vp(params.data_ptr()), int(BW), _vp(imin.data_ptr()), int(W), rptr,
                                        rstride, _vp(invC.data_ptr()), istride, int(mode), _vp(out_L.data_ptr()),
                                        None if grad is None else _vp(grad.data_ptr()), _vp(s)), "fd_source_grad")
    return out_L, grad


def _autograd_class():
    import torch

    class UcbSourceLogL(torch.autograd.Function):
        """L(theta) = (R|h) - (h|h)/2 per source; backward = exact gradient."""

        @staticmethod
        def forward(ctx, params, orbit, T, BW, imin, invC, resid):
            # one adjoint launch: L and the gradient the backward returns
            # (never the Fisher kernel)
            L, grad = source_grad(orbit, T, params.detach(), BW, imin, invC, resid)
            ctx.save_for_backward(grad)
            return L

        @staticmethod
        def backward(ctx, gL):
            (grad,) = ctx.saved_tensors
            return gL[:, None] * grad, None, None, None, None, None, None

    return UcbSourceLogL


_cls = None


def UcbSourceLogL():
    """The torch.autograd.Function class (built on first use)."""
    global _cls
    if _cls is None:
        _cls = _autograd_class()
    return _cls


def ucb_source_logl(params, orbit, T, BW, imin, invC, resid=None):
    """L(theta) per source as a differentiable torch function: through
    UcbSourceLogL (the adjoint kernel's L and gradient) when a gradient can
    flow to params, else the likelihood-only kernel (no gradient work)."""
    import torch
    if torch.is_grad_enabled() and params.requires_grad:
        return UcbSourceLogL().apply(params, orbit, T, BW, imin, invC, resid)
    return source_grad(orbit, T, params.detach(), BW, imin, invC, resid, want_grad=False)[0]
