"""Wavelet-domain (WDM) part of the UCB path on the GPU.

Mirrors the reference's C interface (utils/src/glass_wavelet.c,
ucb/src/glass_ucb_waveform.c:1061-1240, utils/src/glass_math.c:428-438,
utils/src/glass_model.c:805-855) over ``include/glass_hip_wdm.h``:

* Wdm: the struct Wavelets fields the path reads (NF, NT, dt, df, A, kmin,
  kmax, imin, imax), filled by hand for an NF x NT grid as the reference's
  own WDM tests do;
* interpolated_analytic_orbit: the orbit knots and positions of
  initialize_interpolated_analytic_orbits;
* wavelet_transform_timefreq / _freq / _inverse_freq, ucb_waveform_wavelet,
  snr_wavelet, gaussian_log_likelihood_wavelet: drop-ins on host arrays;
* wdm_transform: forward transform of a device tensor;
* WdmBatch: many templates per launch with the fused inner products.
"""
import ctypes

import numpy as np

from . import _f64, check, checkd, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i, _d = ctypes.c_int, ctypes.c_double


class Wdm(ctypes.Structure):
    """glass_hip_wdm.  Wdm.grid(NF, NT, dt) fills it the way the reference's
    setup_wdm_basis / initialize_wavelet do (A = 0, df = 1/(2 dt), edge
    buffer 100, data pixels from layer 1 up to layer NF-1)."""
    _fields_ = [("NF", _i), ("NT", _i), ("dt", _d), ("df", _d), ("A", _d), ("kmin", _i), ("kmax", _i),
                ("imin", _i), ("imax", _i)]

    @classmethod
    def grid(cls, NF, NT, dt):
        return cls(NF=NF, NT=NT, dt=float(dt), df=1.0 / (2.0 * dt), A=0.0, kmin=NT, kmax=NT * (NF - 1),
                   imin=100, imax=NT - 100)

    @property
    def npix(self):
        return self.kmax - self.kmin


_W = ctypes.POINTER(Wdm)
_SIGS = [
    ("glass_hip_interpolated_analytic_orbit", [_vp, _d, _d, _ip, _dp, _dp, _dp, _dp], _i),
    ("glass_hip_wavelet_transform_timefreq", [_W, _dp], _i),
    ("glass_hip_wavelet_transform_freq", [_W, _dp, _dp], _i),
    ("glass_hip_wavelet_transform_inverse_freq", [_W, _dp, _dp], _i),
    ("glass_hip_ucb_waveform_wavelet", [_W, _i, _dp, _dp, _dp, _dp, _d, _d, _dp, _ip, _ip, _dp, _dp, _dp], _i),
    ("glass_hip_snr_wavelet", [_dp, _dp, _dp, _dp, _i, _ip, _i], _d),
    ("glass_hip_gaussian_log_likelihood_wavelet", [_W, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _i, _dp, _i], _d),
    ("glass_hip_wdm_transform", [_W, _vp, _vp, _vp], _i),
    ("glass_hip_wdm_batch_create", [_vp, _W, _i, _dp, _dp, _dp, _dp, _d, _d, _dp, _dp, _i, _i], _i),
    ("glass_hip_wdm_batch_create_layers", [_vp, _W, _i, _dp, _dp, _dp, _dp, _d, _d, _dp, _dp, _i, _i, _i], _i),
    ("glass_hip_wdm_batch_capacity", [_vp, _vp], _i),
    ("glass_hip_wdm_batch_destroy", [_vp], _i),
    ("glass_hip_wdm_batch_buffers", [_vp, _vp, _vp, _vp], _i),
    ("glass_hip_wdm_batch_eval", [_vp, _i, _vp, _vp, _vp, _vp], _i),
    ("glass_hip_wdm_batch_templates", [_vp, _i, _ip, _ip, _dp], _i),
    ("glass_hip_wdm_batch_scratch", [_vp, _i, _dp, _i], _i),
    ("glass_hip_wavelet_transform_timefreq_by_layers", [_W, _i, _i, _dp, _dp], _i),
    ("glass_hip_ucb_fisher_wavelet", [_W, _i, _dp, _dp, _dp, _dp, _d, _d, _dp, _dp, _i, _dp], _i),
    ("glass_hip_wdm_batch_fisher", [_vp, _i, _vp, _vp, _vp], _i),
    ("glass_hip_stationary_dft_psd_to_wdm_psd", [_W, _dp, _dp], _i),
    ("glass_hip_stationary_dft_psd_to_wdm_layer_var", [_W, _dp, _dp], _i),
    ("glass_hip_wdm_psd", [_W, _vp, _vp, _vp, _vp], _i),
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


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def interpolated_analytic_orbit(orbit, Tobs, t0=0.0):
    """(t [Norb], x [3,Norb], y [3,Norb], z [3,Norb]) of
    initialize_interpolated_analytic_orbits."""
    L = _L()
    n = ctypes.c_int()
    ob = ctypes.cast(ctypes.byref(orbit), _vp)
    check(L.glass_hip_interpolated_analytic_orbit(ob, float(Tobs), float(t0), ctypes.byref(n), None, None, None,
                                                  None), "interpolated_analytic_orbit")
    N = n.value
    t = np.zeros(N)
    x, y, z = np.zeros((3, N)), np.zeros((3, N)), np.zeros((3, N))
    check(L.glass_hip_interpolated_analytic_orbit(ob, float(Tobs), float(t0), ctypes.byref(n), t.ctypes.data_as(_dp),
                                                  x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                                  z.ctypes.data_as(_dp)), "interpolated_analytic_orbit")
    return t, x, y, z


def wavelet_transform_timefreq(w, timedata):
    """Forward WDM transform of NF*NT samples; returns the NF*NT pixels."""
    out = np.array(timedata, dtype=np.float64, copy=True)
    if out.size != w.NF * w.NT:
        raise ValueError(f"expected {w.NF * w.NT} samples, got {out.size}")
    check(_L().glass_hip_wavelet_transform_timefreq(ctypes.byref(w), out.ctypes.data_as(_dp)),
          "wavelet_transform_timefreq")
    return out


def wavelet_transform_freq(w, freqdata):
    """ND/2+1 complex bins (ND+2 doubles) -> ND pixels."""
    f = _f64(freqdata)
    if f.size != w.NF * w.NT + 2:
        raise ValueError("freqdata must hold ND+2 doubles")
    out = np.zeros(w.NF * w.NT)
    check(_L().glass_hip_wavelet_transform_freq(ctypes.byref(w), f.ctypes.data_as(_dp), out.ctypes.data_as(_dp)),
          "wavelet_transform_freq")
    return out


def wavelet_transform_inverse_freq(w, wdmdata):
    """ND pixels -> ND/2+1 complex bins (ND+2 doubles)."""
    p = _f64(wdmdata)
    if p.size != w.NF * w.NT:
        raise ValueError("wdmdata must hold ND doubles")
    out = np.zeros(w.NF * w.NT + 2)
    check(_L().glass_hip_wavelet_transform_inverse_freq(ctypes.byref(w), p.ctypes.data_as(_dp),
                                                        out.ctypes.data_as(_dp)), "wavelet_transform_inverse_freq")
    return out


def wavelet_transform_timefreq_by_layers(w, jmin, nlayers, window, timedata):
    """wavelet_transform_timefreq_by_layers on a copy of the NT (nlayers+1)
    samples; returns the transformed array."""
    x = np.array(timedata, dtype=np.float64, copy=True)
    win = _f64(window)
    if x.size != w.NT * (nlayers + 1) or win.size != w.NT // 2 + 1:
        raise ValueError("timedata must hold NT (nlayers + 1) samples and window NT/2 + 1")
    check(_L().glass_hip_wavelet_transform_timefreq_by_layers(ctypes.byref(w), int(jmin), int(nlayers),
                                                              win.ctypes.data_as(_dp), x.ctypes.data_as(_dp)),
          "wavelet_transform_timefreq_by_layers")
    return x


def ucb_waveform_wavelet(w, orb, Tobs, params, t0=0.0, X=None, Y=None, Z=None):
    """Drop-in ucb_waveform_wavelet.  orb = (t, x, y, z).  X, Y, Z: arrays of
    w.npix pixels to insert into (allocated, zeroed, when None).  Returns
    (wavelet_list, X, Y, Z)."""
    t, x, y, z = (_f64(a) for a in orb)
    p = _f64(params)
    npix = w.npix
    X = np.zeros(npix) if X is None else X
    Y = np.zeros(npix) if Y is None else Y
    Z = np.zeros(npix) if Z is None else Z
    for a in (X, Y, Z):
        if a.dtype != np.float64 or not a.flags.c_contiguous or a.size < npix:
            raise ValueError("X, Y, Z must be contiguous float64 arrays of npix pixels")
    lst = np.zeros(3 * w.NT, dtype=np.int32)
    nw = ctypes.c_int()
    check(_L().glass_hip_ucb_waveform_wavelet(ctypes.byref(w), len(t), t.ctypes.data_as(_dp),
                                              x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), z.ctypes.data_as(_dp),
                                              float(Tobs), float(t0), p.ctypes.data_as(_dp),
                                              lst.ctypes.data_as(_ip), ctypes.byref(nw), X.ctypes.data_as(_dp),
                                              Y.ctypes.data_as(_dp), Z.ctypes.data_as(_dp)), "ucb_waveform_wavelet")
    return lst[:nw.value].copy(), X, Y, Z


def ucb_fisher_wavelet(w, orb, Tobs, params, invC9, t0=0.0):
    """ucb_fisher_wavelet up to the eigensolve -> Fisher [8, 8].
    orb = (t, x, y, z); invC9 [9, npix] row-major 3x3 planes."""
    t, x, y, z = (_f64(a) for a in orb)
    p, ic = _f64(params), _f64(invC9)
    if ic.shape != (9, w.npix):
        raise ValueError("invC9 must be [9, npix]")
    F = np.zeros((8, 8))
    check(_L().glass_hip_ucb_fisher_wavelet(ctypes.byref(w), len(t), t.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                                            y.ctypes.data_as(_dp), z.ctypes.data_as(_dp), float(Tobs), float(t0),
                                            p.ctypes.data_as(_dp), ic.ctypes.data_as(_dp), w.npix,
                                            F.ctypes.data_as(_dp)), "ucb_fisher_wavelet")
    return F


def snr_wavelet(X, Y, Z, invC9, lst):
    """snr_wavelet: X,Y,Z [N], invC9 [9,N] (row-major 3x3 planes), list."""
    X, Y, Z, ic = _f64(X), _f64(Y), _f64(Z), _f64(invC9)
    lst = _i32(lst)
    return checkd(_L().glass_hip_snr_wavelet(X.ctypes.data_as(_dp), Y.ctypes.data_as(_dp), Z.ctypes.data_as(_dp),
                                             ic.ctypes.data_as(_dp), X.size, lst.ctypes.data_as(_ip), lst.size),
                  "snr_wavelet")


def gaussian_log_likelihood_wavelet(w, data, model, mlist, invC9):
    """gaussian_log_likelhood_wavelet: data, model [3,N]; invC9 [9,N]."""
    d, m, ic = _f64(data), _f64(model), _f64(invC9)
    ml = _i32(mlist)
    N = d.shape[1]
    return checkd(_L().glass_hip_gaussian_log_likelihood_wavelet(
        ctypes.byref(w), d[0].ctypes.data_as(_dp), d[1].ctypes.data_as(_dp), d[2].ctypes.data_as(_dp),
        m[0].ctypes.data_as(_dp), m[1].ctypes.data_as(_dp), m[2].ctypes.data_as(_dp), ml.ctypes.data_as(_ip),
        ml.size, ic.ctypes.data_as(_dp), N), "gaussian_log_likelihood_wavelet")


def stationary_dft_psd_to_wdm_psd(w, dft_psd):
    """stationary_dft_psd_to_wdm_psd (utils/src/glass_wavelet.c:901-911):
    DFT variances at k = 0..ND/2 -> the ND pixel variances."""
    psd = _f64(dft_psd)
    if psd.size != w.NF * w.NT // 2 + 1:
        raise ValueError(f"dft_psd must hold ND/2 + 1 = {w.NF * w.NT // 2 + 1} values")
    out = np.zeros(w.NF * w.NT)
    check(_L().glass_hip_stationary_dft_psd_to_wdm_psd(ctypes.byref(w), psd.ctypes.data_as(_dp),
                                                       out.ctypes.data_as(_dp)), "stationary_dft_psd_to_wdm_psd")
    return out


def stationary_dft_psd_to_wdm_layer_var(w, dft_psd):
    """stationary_dft_psd_to_wdm_layer_var (utils/src/glass_wavelet.c:865-899):
    the NF + 1 layer variances."""
    psd = _f64(dft_psd)
    if psd.size != w.NF * w.NT // 2 + 1:
        raise ValueError(f"dft_psd must hold ND/2 + 1 = {w.NF * w.NT // 2 + 1} values")
    out = np.zeros(w.NF + 1)
    check(_L().glass_hip_stationary_dft_psd_to_wdm_layer_var(ctypes.byref(w), psd.ctypes.data_as(_dp),
                                                             out.ctypes.data_as(_dp)),
          "stationary_dft_psd_to_wdm_layer_var")
    return out


def wdm_transform(w, time_tensor, out=None, stream=None):
    """Forward WDM transform of a float64 device tensor of NF*NT samples."""
    import torch
    x = time_tensor.contiguous().to(torch.float64)
    out = torch.empty_like(x) if out is None else out
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(_L().glass_hip_wdm_transform(ctypes.byref(w), _vp(x.data_ptr()), _vp(out.data_ptr()), _vp(s)),
          "wdm_transform")
    return out


class WdmBatch:
    """Batched wavelet-domain templates with fused inner products.

    data: [3, npix] pixels X,Y,Z (index k - kmin); invC6: [6, npix] planes
    XX, YY, ZZ, XY, XZ, YZ.  eval(params) -> (dlogL, snr) torch tensors,
    where dlogL is the change of the reference's wavelet logL when the
    template is added to an empty model."""

    def __init__(self, w, orb, Tobs, data, invC6, max_templates, t0=0.0, layers=0):
        """layers > 0: every template takes that many frequency layers
        (glass_hip_wdm_batch_create_layers; a timing knob, no reference
        counterpart) instead of ucb_wavelet_layers' 1-3."""
        import torch
        self.w = w
        self.max = int(max_templates)
        self.layers = int(layers)
        t, x, y, z = (_f64(a) for a in orb)
        d, ic = _f64(data), _f64(invC6)
        if d.shape != (3, w.npix) or ic.shape != (6, w.npix):
            raise ValueError("data must be [3, npix] and invC6 [6, npix]")
        self._p = _vp()
        args = (ctypes.byref(self._p), ctypes.byref(w), len(t), t.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                y.ctypes.data_as(_dp), z.ctypes.data_as(_dp), float(Tobs), float(t0), d.ctypes.data_as(_dp),
                ic.ctypes.data_as(_dp), w.npix, self.max)
        if self.layers:
            check(_L().glass_hip_wdm_batch_create_layers(*args, self.layers), "wdm_batch_create_layers")
        else:
            check(_L().glass_hip_wdm_batch_create(*args), "wdm_batch_create")
        vc = ctypes.c_int()
        check(_L().glass_hip_wdm_batch_capacity(self._p, ctypes.byref(vc)), "wdm_batch_capacity")
        self.vcap = vc.value
        dev = torch.device("cuda")
        self.dlogL = torch.empty(self.max, dtype=torch.float64, device=dev)
        self.snr = torch.empty(self.max, dtype=torch.float64, device=dev)

    def eval(self, params, stream=None):
        import torch
        p = params.contiguous().to(torch.float64)
        n = p.shape[0]
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(_L().glass_hip_wdm_batch_eval(self._p, n, _vp(p.data_ptr()), _vp(self.dlogL.data_ptr()),
                                            _vp(self.snr.data_ptr()), _vp(s)), "wdm_batch_eval")
        return self.dlogL[:n], self.snr[:n]

    def fisher(self, params, stream=None):
        """Wavelet-domain FD Fisher matrices [n, 8, 8] (device) of n <=
        max_templates / 16 sources (overwrites the template buffers)."""
        import torch
        p = params.contiguous().to(torch.float64)
        n = p.shape[0]
        out = torch.empty((n, 8, 8), dtype=torch.float64, device=p.device)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(_L().glass_hip_wdm_batch_fisher(self._p, n, _vp(p.data_ptr()), _vp(out.data_ptr()), _vp(s)),
              "wdm_batch_fisher")
        return out

    def templates(self, n):
        """(nwave [n], list [n, vcap], values [n, 3, vcap]) of the last eval,
        copied to the host (synchronous); vcap = 3 NT, or NT x layers."""
        h_nw = np.zeros(n, dtype=np.int32)
        h_l = np.zeros((n, self.vcap), dtype=np.int32)
        h_v = np.zeros((n, 3, self.vcap))
        check(_L().glass_hip_wdm_batch_templates(self._p, n, h_nw.ctypes.data_as(_ip), h_l.ctypes.data_as(_ip),
                                                 h_v.ctypes.data_as(_dp)), "wdm_batch_templates")
        return h_nw, h_l, h_v

    def response(self, e, norb):
        """Diagnostics: (header[64], phase spline [4,N], TDI splines [6,4,N])
        of template e of the last eval."""
        out = np.zeros(64 + 28 * norb)
        check(_L().glass_hip_wdm_batch_scratch(self._p, e, out.ctypes.data_as(_dp), out.size), "wdm_batch_scratch")
        return out[:64], out[64:64 + 4 * norb].reshape(4, norb), out[64 + 4 * norb:].reshape(6, 4, norb)

    def close(self):
        if self._p:
            _L().glass_hip_wdm_batch_destroy(self._p)
            self._p = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
