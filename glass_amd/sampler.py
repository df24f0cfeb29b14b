"""The UCB sampler's per-step work on the device, over ``include/glass_hip_sampler.h``.

Mirrors ucb_mcmc (ucb/src/glass_ucb_sampler.c:75-173) with the reference's
proposal table (ucb/src/glass_ucb_proposal.c:615-760), priors
(ucb/src/glass_ucb_prior.c:396-646), ptmcmc and adapt_temperature_ladder
(utils/src/glass_sampler.c:19-94), for every chain of a batched plan at
once:

* UcbSampler: ladders of temperature slots over the chains of an FdBatch;
  steps / Fisher update / PT swap, or a whole iteration of the reference's
  outer loop as one hipGraph;
* fisher_eigen, matrix_eigenstuff: the eigen-decomposition that replaces
  matrix_eigenstuff's LAPACKE_dgeev (utils/src/glass_math.c:475-499).
"""
import ctypes
import os

import numpy as np

from . import check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_vp = ctypes.c_void_p
_i = ctypes.c_int

# The reference's proposal weights without F-statistic, GMM or covariance
# proposals (initialize_ucb_proposal: uniform prior 0.1, fm shift 0.1,
# psi-phi jump 0.1, Fisher the rest).
DEFAULT_WEIGHTS = (0.1, 0.0, 0.0, 0.0, 0.7, 0.1, 0.1, 0.0, 0.0)
# ... and with the F-statistic draw and jump (0.2 each), as the reference
# runs by default (ucb/src/glass_ucb_proposal.c:636-747)
FSTAT_WEIGHTS = (0.1, 0.2, 0.2, 0.0, 0.3, 0.1, 0.1, 0.0, 0.0)


class Opts(ctypes.Structure):
    _fields_ = [("weight", ctypes.c_double * 9), ("fix_sky", _i), ("fix_freq", _i), ("fix_fdot", _i),
                ("snr_min", ctypes.c_double)]


_SIGS = [
    ("glass_hip_fisher_eigen", [_i, _vp, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_matrix_eigenstuff", [_dp, _dp, _dp, _i], _i),
    ("glass_hip_ucb_sampler_create", [_vp, _vp, _i, _i, _dp, ctypes.POINTER(Opts), _vp], _i),
    ("glass_hip_ucb_sampler_destroy", [_vp], _i),
    ("glass_hip_ucb_sampler_set_state", [_vp, _up, _dp, _dp, _i, _vp], _i),
    ("glass_hip_ucb_sampler_fisher", [_vp, _vp], _i),
    ("glass_hip_ucb_sampler_set_eigen", [_vp, _dp, _dp, _vp], _i),
    ("glass_hip_ucb_sampler_get_eigen", [_vp, _dp, _dp, _vp], _i),
    ("glass_hip_ucb_sampler_steps", [_vp, _i, _vp], _i),
    ("glass_hip_ucb_sampler_ptswap", [_vp, _i, _vp], _i),
    ("glass_hip_ucb_sampler_sweep", [_vp, _i, _i, _i, _vp], _i),
    ("glass_hip_ucb_sampler_get_state", [_vp, _up, _ip, _dp, _ip, _ip, _ip, _vp], _i),
    ("glass_hip_ucb_sampler_set_fstat", [_vp, _i, _dp, _i, ctypes.c_double, _i, _i, ctypes.c_double, _i, _vp], _i),
    ("glass_hip_ucb_sampler_save_chain_states", [_vp, ctypes.POINTER(ctypes.c_char_p), _i, _dp, _dp,
                                                 ctypes.c_double, _dp, _i, _vp], _i),
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


def fisher_eigen(F):
    """(evec, eval, sweeps) of a CUDA float64 tensor F [n, 8, 8]: evec[k][i][j]
    = component i of eigenvector j, eigenvalues ascending."""
    import torch
    F = F.contiguous()
    n = F.shape[0]
    ev = torch.empty_like(F)
    la = torch.empty((n, 8), dtype=torch.float64, device=F.device)
    sw = torch.empty(n, dtype=torch.int32, device=F.device)
    check(_L().glass_hip_fisher_eigen(int(n), ctypes.c_void_p(F.data_ptr()), ctypes.c_void_p(ev.data_ptr()),
                                      ctypes.c_void_p(la.data_ptr()), ctypes.c_void_p(sw.data_ptr()), _stream()),
          "fisher_eigen")
    return ev, la, sw


def matrix_eigenstuff(F):
    """matrix_eigenstuff on a host 8x8 matrix -> (evectors, evalues)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    ev, la = np.zeros((8, 8)), np.zeros(8)
    check(_L().glass_hip_matrix_eigenstuff(_p(F), _p(ev), _p(la), 8), "matrix_eigenstuff")
    return ev, la


class UcbSampler:
    """nladder ladders of ntemp temperature slots over the chains of an
    FdBatch (nladder * ntemp == its chains; slot s starts on chain s).
    prior: [G, 8, 2] per data window (model->prior)."""

    def __init__(self, plan, nladder, ntemp, prior, weights=DEFAULT_WEIGHTS, fix_sky=False, fix_freq=False,
                 fix_fdot=False, snr_min=5.0):
        self.plan = plan
        self.nladder, self.ntemp = int(nladder), int(ntemp)
        self.nslot = self.nladder * self.ntemp
        self.max_src = plan.max_src
        pr = np.ascontiguousarray(prior, dtype=np.float64)
        o = Opts()
        for k in range(9):
            o.weight[k] = float(weights[k])
        o.fix_sky, o.fix_freq, o.fix_fdot = int(fix_sky), int(fix_freq), int(fix_fdot)
        o.snr_min = float(snr_min)
        self.h = _vp()
        check(_L().glass_hip_ucb_sampler_create(ctypes.byref(self.h), plan.h, self.nladder, self.ntemp, _p(pr),
                                                ctypes.byref(o), _stream()), "ucb_sampler_create")

    def close(self):
        if getattr(self, "h", None):
            _L().glass_hip_ucb_sampler_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def set_state(self, seeds, temperature, lognorm=None, adapt_iter0=0):
        sd = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(self.nslot)
        t = np.ascontiguousarray(temperature, dtype=np.float64).reshape(self.nslot)
        ln = None if lognorm is None else np.ascontiguousarray(lognorm, dtype=np.float64).reshape(self.nslot)
        check(_L().glass_hip_ucb_sampler_set_state(self.h, sd.ctypes.data_as(_up), _p(t),
                                                   None if ln is None else _p(ln), int(adapt_iter0), _stream()),
              "ucb_sampler_set_state")

    def fisher(self):
        check(_L().glass_hip_ucb_sampler_fisher(self.h, _stream()), "ucb_sampler_fisher")

    def set_fstat(self, g, tensor, d_f, maxp, data_qmin):
        """The F-statistic proposal density [n_f, n_theta, n_phi] of window g."""
        t = np.ascontiguousarray(tensor, dtype=np.float64)
        check(_L().glass_hip_ucb_sampler_set_fstat(self.h, int(g), _p(t), t.shape[0], float(d_f), t.shape[1],
                                                   t.shape[2], float(maxp), int(data_qmin), _stream()),
              "ucb_sampler_set_fstat")

    def eigen(self):
        ev = np.zeros((self.nslot, self.max_src, 8, 8))
        la = np.zeros((self.nslot, self.max_src, 8))
        check(_L().glass_hip_ucb_sampler_get_eigen(self.h, _p(ev), _p(la), _stream()), "ucb_sampler_get_eigen")
        return ev, la

    def set_eigen(self, evec, eval_):
        ev = np.ascontiguousarray(evec, dtype=np.float64)
        la = np.ascontiguousarray(eval_, dtype=np.float64)
        assert ev.size == self.nslot * self.max_src * 64 and la.size == self.nslot * self.max_src * 8
        check(_L().glass_hip_ucb_sampler_set_eigen(self.h, _p(ev), _p(la), _stream()), "ucb_sampler_set_eigen")

    def steps(self, n):
        check(_L().glass_hip_ucb_sampler_steps(self.h, int(n), _stream()), "ucb_sampler_steps")

    def ptswap(self, adapt=False):
        check(_L().glass_hip_ucb_sampler_ptswap(self.h, 1 if adapt else 0, _stream()), "ucb_sampler_ptswap")

    def sweep(self, nsteps=80, adapt=False, graph=False):
        check(_L().glass_hip_ucb_sampler_sweep(self.h, int(nsteps), 1 if adapt else 0, 1 if graph else 0,
                                               _stream()), "ucb_sampler_sweep")

    def state(self):
        S = self.nslot
        sd = np.zeros(S, dtype=np.uint32)
        idx = np.zeros(S, dtype=np.int32)
        t = np.zeros(S)
        tr = np.zeros((S, 9), dtype=np.int32)
        ac = np.zeros((S, 9), dtype=np.int32)
        sw = np.zeros(S, dtype=np.int32)
        check(_L().glass_hip_ucb_sampler_get_state(self.h, sd.ctypes.data_as(_up), idx.ctypes.data_as(_ip), _p(t),
                                                   tr.ctypes.data_as(_ip), ac.ctypes.data_as(_ip),
                                                   sw.ctypes.data_as(_ip), _stream()), "ucb_sampler_get_state")
        return {"seeds": sd, "index": idx, "temperature": t, "trials": tr, "accepts": ac, "swaps": sw}

    def save_chain_states(self, dirs, step, logLmax=None, lognorm=None, t0=0.0, eta=None):
        """save_chain_state of every ladder: dirs[l]/chain_state_<t>.dat holds
        the chain now at temperature t of ladder l."""
        assert len(dirs) == self.nladder
        arr = (ctypes.c_char_p * self.nladder)(*[os.fsencode(d) for d in dirs])
        lm = None if logLmax is None else np.ascontiguousarray(logLmax, dtype=np.float64).reshape(self.nladder)
        ln = None if lognorm is None else np.ascontiguousarray(lognorm, dtype=np.float64).reshape(self.nslot)
        et = None if eta is None else np.ascontiguousarray(eta, dtype=np.float64).ravel()
        check(_L().glass_hip_ucb_sampler_save_chain_states(self.h, arr, int(step), None if lm is None else _p(lm),
                                                           None if ln is None else _p(ln), float(t0),
                                                           None if et is None else _p(et),
                                                           0 if et is None else len(et), _stream()),
              "ucb_sampler_save_chain_states")
