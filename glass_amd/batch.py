"""Python binding of the batched device API (include/glass_hip_batch.h).

Device memory and streams come from torch (ROCm build); the compute is the
HIP library's.  FdBatch mirrors one PT-MCMC sweep's likelihood work: many
chains x temperatures x sources, one proposal per work item.
"""
import ctypes

import numpy as np

from . import _EXTRA, _f64, check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i, _d = ctypes.c_int, ctypes.c_double



class Tile(ctypes.Structure):
    """glass_hip_tile: threads per workgroup, grid multiplier (0 = one
    workgroup per proposal slot) and waves per SIMD of one BW class's eval
    kernel."""
    _fields_ = [("nthr", ctypes.c_int), ("gmul", ctypes.c_int), ("waves", ctypes.c_int)]


_tp = ctypes.POINTER(Tile)

_EXTRA += [
    ("glass_hip_fd_batch_tile_candidates", [_i, _tp, _i, _ip], _i),
    ("glass_hip_fd_batch_get_tiles", [_vp, _tp], _i),
    ("glass_hip_fd_batch_set_tiles", [_vp, _tp], _i),
    ("glass_hip_fd_batch_autotune", [_vp, _i, _vp, _vp, _vp, _i, _tp, _dp, _vp], _i),
    ("glass_hip_fd_batch_derivs", [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_batch_create_ex", [ctypes.POINTER(_vp), _vp, ctypes.c_char_p, _d, _d, _i, _ip, _ip, _dp, _dp,
                                      _i, _ip, _ip, _i, _dp, _i], _i),
    ("glass_hip_fd_batch_sweep_graph", [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_batch_memory", [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)], _i),
    ("glass_hip_fd_batch_create", [ctypes.POINTER(_vp), _vp, ctypes.c_char_p, _d, _d, _i, _ip, _ip, _dp, _dp,
                                   _i, _ip, _ip, _dp, _i], _i),
    ("glass_hip_fd_batch_destroy", [_vp], _i),
    ("glass_hip_fd_batch_buffers", [_vp, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_batch_eval", [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_batch_eval_graph", [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_batch_accept", [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_fd_batch_refresh", [_vp, _vp], _i),
    ("glass_hip_fd_batch_state", [_vp, _dp, _dp], _i),
    ("glass_hip_fd_batch_profile", [_vp, _i], _i),
    ("glass_hip_fd_batch_set_concurrent", [_vp, _i], _i),
    ("glass_hip_fd_batch_profile_read", [_vp, _dp, _ip, _ip], _i),
    ("glass_hip_ucb_bandwidth", [_d] * 7 + [_i], _i),
    ("glass_hip_ucb_alignment", [_vp, _d, _i, _i, _dp, _ip, _ip, _ip], _i),
    ("glass_hip_logl_batch_create", [ctypes.POINTER(_vp), _i, _ip, _vp, _vp, _vp], _i),
    ("glass_hip_logl_batch_destroy", [_vp], _i),
    ("glass_hip_logl_batch_eval", [_vp, _vp, _vp, _vp, _vp, _vp], _i),
]


def _L():
    L = lib()
    # _declare ran before this module extended _EXTRA if lib() was loaded
    # earlier: (re)declare the entry points here
    for name, args, res in _EXTRA:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def ucb_alignment(orbit, T, NFFT, data_qmin, params):
    """ucb_alignment (ucb/src/glass_ucb_waveform.c:727-736) -> (BW, qmin, imin)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    bw, qm, im = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(_L().glass_hip_ucb_alignment(ctypes.cast(ctypes.byref(orbit), _vp), float(T), int(NFFT),
                                       int(data_qmin), p.ctypes.data_as(_dp), ctypes.byref(bw),
                                       ctypes.byref(qm), ctypes.byref(im)), "ucb_alignment")
    return bw.value, qm.value, im.value


NCLASS = 8   # BW classes 32 << k, k = 0..7


def tile_candidates(k):
    """Compiled (nthr, gmul, waves) launch shapes of BW class k (BW = 32 <<
    k), the default first."""
    out = (Tile * 128)()
    n = ctypes.c_int()
    check(_L().glass_hip_fd_batch_tile_candidates(int(k), out, 128, ctypes.byref(n)), "tile_candidates")
    return [(out[i].nthr, out[i].gmul, out[i].waves) for i in range(min(n.value, 128))]


def pack_invC(invC33):
    """[3,3,W] -> [6,W] planes XX, YY, ZZ, XY, XZ, YZ."""
    return np.stack([invC33[0, 0], invC33[1, 1], invC33[2, 2], invC33[0, 1], invC33[0, 2], invC33[1, 2]])


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class FdBatch:
    """A device-resident set of data windows and chains (see the header)."""

    def __init__(self, orbit, fmt, T, t0, group_qmin, data, invC, chain_group, chain_nsrc, src_params,
                 max_prop, flat=False):
        """data: list of complex [3, W_g]; invC: list of [3,3,W_g];
        src_params: [nchain, max_src, 8] (max_src 8 uses
        glass_hip_fd_batch_create, others glass_hip_fd_batch_create_ex).
        With flat=True, data is already the flat complex array of all
        windows and invC the flat [6 W_g] planes per window, and group_W is
        passed as data's window sizes via the `W` attribute of group_qmin's
        tuple (group_qmin, group_W)."""
        import torch
        self.torch = torch
        L = _L()
        if flat:
            group_qmin, W = group_qmin
            W = np.ascontiguousarray(W, dtype=np.int32)
            dflat = np.ascontiguousarray(data, dtype=np.complex128)
            iflat = np.ascontiguousarray(invC, dtype=np.float64)
        else:
            W = np.array([d.shape[-1] for d in data], dtype=np.int32)
            dflat = np.concatenate([np.ascontiguousarray(d, dtype=np.complex128).reshape(-1) for d in data])
            iflat = np.concatenate([pack_invC(np.asarray(c)).reshape(-1) for c in invC])
        gq = np.ascontiguousarray(group_qmin, dtype=np.int32)
        cg = np.ascontiguousarray(chain_group, dtype=np.int32)
        cn = np.ascontiguousarray(chain_nsrc, dtype=np.int32)
        sp = np.ascontiguousarray(src_params, dtype=np.float64)
        assert sp.ndim == 3 and sp.shape[0] == len(cg) and sp.shape[2] == 8
        self.nchain = len(cg)
        self.max_src = sp.shape[1]
        self.max_prop = int(max_prop)
        h = _vp()
        dv = _f64(dflat)
        iv = _f64(iflat)
        args = [ctypes.byref(h), ctypes.cast(ctypes.byref(orbit), _vp), fmt.encode(), float(T), float(t0), len(W),
                gq.ctypes.data_as(_ip), W.ctypes.data_as(_ip), dv.ctypes.data_as(_dp), iv.ctypes.data_as(_dp),
                len(cg), cg.ctypes.data_as(_ip), cn.ctypes.data_as(_ip)]
        if self.max_src == 8:
            check(L.glass_hip_fd_batch_create(*args, sp.ctypes.data_as(_dp), self.max_prop), "fd_batch_create")
        else:
            check(L.glass_hip_fd_batch_create_ex(*args, self.max_src, sp.ctypes.data_as(_dp), self.max_prop),
                  "fd_batch_create_ex")
        self.h = h
        dev = torch.device("cuda")
        self.chain = torch.zeros(self.max_prop, dtype=torch.int32, device=dev)
        self.src = torch.zeros(self.max_prop, dtype=torch.int32, device=dev)
        self.params = torch.zeros((self.max_prop, 8), dtype=torch.float64, device=dev)
        self.logL = torch.zeros(self.max_prop, dtype=torch.float64, device=dev)
        self.snr = torch.zeros(self.max_prop, dtype=torch.float64, device=dev)

    def close(self):
        if getattr(self, "h", None):
            _L().glass_hip_fd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_proposals(self, chain, src, params):
        n = len(chain)
        t = self.torch
        self.chain[:n].copy_(t.as_tensor(np.asarray(chain, dtype=np.int32)))
        self.src[:n].copy_(t.as_tensor(np.asarray(src, dtype=np.int32)))
        self.params[:n].copy_(t.as_tensor(np.asarray(params, dtype=np.float64)))
        return n

    def eval(self, n, graph=False):
        """Evaluates the first n proposals in the plan's torch buffers on the
        current torch stream; returns (logL, snr) views (device)."""
        s = ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        fn = _L().glass_hip_fd_batch_eval_graph if graph else _L().glass_hip_fd_batch_eval
        check(fn(self.h, int(n), _ptr(self.chain), _ptr(self.src), _ptr(self.params), _ptr(self.logL),
                 _ptr(self.snr), s), "fd_batch_eval")
        return self.logL[:n], self.snr[:n]

    def accept(self, idx):
        """Accept proposals idx (indices into the last evaluated buffers)."""
        t = self.torch
        d_idx = t.as_tensor(np.asarray(idx, dtype=np.int32)).to(self.chain.device)
        s = ctypes.c_void_p(t.cuda.current_stream().cuda_stream)
        check(_L().glass_hip_fd_batch_accept(self.h, len(idx), _ptr(d_idx), _ptr(self.chain), _ptr(self.src),
                                             _ptr(self.params), _ptr(self.logL), s), "fd_batch_accept")
        t.cuda.current_stream().synchronize()

    def refresh(self):
        s = ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        check(_L().glass_hip_fd_batch_refresh(self.h, s), "fd_batch_refresh")

    def profile(self, on=True):
        """Per-BW-class kernel timing of eager evals (makes them synchronous)."""
        check(_L().glass_hip_fd_batch_profile(self.h, int(bool(on))), "fd_batch_profile")

    def set_concurrent(self, on=True):
        """BW-class kernels on forked side streams (default) or in turn."""
        check(_L().glass_hip_fd_batch_set_concurrent(self.h, int(bool(on))), "fd_batch_set_concurrent")

    def profile_read(self):
        """-> (ms[8], launches[8], count[8]) per BW class (BW = 32 << k)."""
        ms = np.zeros(8)
        la = np.zeros(8, dtype=np.int32)
        cn = np.zeros(8, dtype=np.int32)
        check(_L().glass_hip_fd_batch_profile_read(self.h, ms.ctypes.data_as(_dp), la.ctypes.data_as(_ip),
                                                   cn.ctypes.data_as(_ip)), "fd_batch_profile_read")
        return ms, la, cn

    def tiles(self):
        """Current (nthr, gmul, waves) per BW class."""
        t = (Tile * NCLASS)()
        check(_L().glass_hip_fd_batch_get_tiles(self.h, t), "fd_batch_get_tiles")
        return [(x.nthr, x.gmul, x.waves) for x in t]

    def set_tiles(self, tiles):
        t = (Tile * NCLASS)(*[Tile(int(a), int(b), int(c)) for a, b, c in tiles])
        check(_L().glass_hip_fd_batch_set_tiles(self.h, t), "fd_batch_set_tiles")

    def autotune(self, n, reps=3):
        """Times every compiled tile of each BW class over the first n
        proposals in the plan's buffers and keeps the fastest; returns
        (tiles, ms per launch of each class's chosen tile)."""
        t = (Tile * NCLASS)()
        ms = np.zeros(NCLASS)
        s = ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        check(_L().glass_hip_fd_batch_autotune(self.h, int(n), _ptr(self.chain), _ptr(self.src), _ptr(self.params),
                                               int(reps), t, ms.ctypes.data_as(_dp), s), "fd_batch_autotune")
        return [(x.nthr, x.gmul, x.waves) for x in t], ms

    def derivs(self, chain, src):
        """Exact (L, grad [n, 8], Fisher [n, 8, 8]) of sources (chain[e],
        src[e]) at their current parameters (device tensors, current
        stream)."""
        t = self.torch
        dev = self.chain.device
        c = t.as_tensor(np.asarray(chain, dtype=np.int32)).to(dev) if not t.is_tensor(chain) else chain
        s_ = t.as_tensor(np.asarray(src, dtype=np.int32)).to(dev) if not t.is_tensor(src) else src
        n = int(c.numel())
        L = t.empty(n, dtype=t.float64, device=dev)
        g = t.empty((n, 8), dtype=t.float64, device=dev)
        F = t.empty((n, 8, 8), dtype=t.float64, device=dev)
        st = ctypes.c_void_p(t.cuda.current_stream().cuda_stream)
        check(_L().glass_hip_fd_batch_derivs(self.h, n, _ptr(c), _ptr(s_), _ptr(L), _ptr(g), _ptr(F), st),
              "fd_batch_derivs")
        return L, g, F

    def sweep_graph(self, n, dchain, dsrc, dL, dgrad, dfisher):
        """Eval of the first n proposals in the plan's buffers plus the exact
        derivatives of the (dchain, dsrc) device tensors, replayed as one
        hipGraph on the current stream."""
        st = ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        check(_L().glass_hip_fd_batch_sweep_graph(self.h, int(n), _ptr(self.chain), _ptr(self.src),
                                                  _ptr(self.params), _ptr(self.logL), _ptr(self.snr),
                                                  int(dchain.numel()), _ptr(dchain), _ptr(dsrc), _ptr(dL),
                                                  _ptr(dgrad), _ptr(dfisher), st), "fd_batch_sweep_graph")

    def memory(self):
        """(template bytes, total device bytes) the plan holds."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        check(_L().glass_hip_fd_batch_memory(self.h, ctypes.byref(a), ctypes.byref(b)), "fd_batch_memory")
        return a.value, b.value

    def state(self):
        self.torch.cuda.current_stream().synchronize()
        lg = np.zeros(self.nchain)
        sp = np.zeros((self.nchain, self.max_src, 8))
        check(_L().glass_hip_fd_batch_state(self.h, lg.ctypes.data_as(_dp), sp.ctypes.data_as(_dp)),
              "fd_batch_state")
        return lg, sp


class LoglBatch:
    """n independent gaussian_log_likelihood calls (Nchannel 3) in one launch.

    Item e reads a window of W[e] bins: data and model as complex [3, W[e]]
    at bin offsets data_off[e] / model_off[e] of flat complex128 device
    buffers (3 W per item), invC as [6, W[e]] planes (XX YY ZZ XY XZ YZ) at
    bin offset invC_off[e] of a flat float64 device buffer (6 W per item).
    """

    def __init__(self, W, data_off, model_off, invC_off):
        W = np.ascontiguousarray(W, dtype=np.int32)
        offs = [np.ascontiguousarray(o, dtype=np.int64) for o in (data_off, model_off, invC_off)]
        assert all(len(o) == len(W) for o in offs)
        self.n = len(W)
        h = _vp()
        check(_L().glass_hip_logl_batch_create(ctypes.byref(h), self.n, W.ctypes.data_as(_ip),
                                               *(o.ctypes.data_as(_vp) for o in offs)), "logl_batch_create")
        self.h = h

    def eval(self, data, model, invC, out=None):
        """data, model: complex128 (or float64 re/im) device tensors; invC:
        float64 device tensor.  Returns logL[n] (device), on the current
        torch stream."""
        import torch
        if out is None:
            out = torch.empty(self.n, dtype=torch.float64, device=data.device)
        for t in (data, model, invC, out):
            assert t.is_cuda and t.is_contiguous()
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(_L().glass_hip_logl_batch_eval(self.h, _ptr(data), _ptr(model), _ptr(invC), _ptr(out), s),
              "logl_batch_eval")
        return out

    def close(self):
        if getattr(self, "h", None):
            _L().glass_hip_logl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
