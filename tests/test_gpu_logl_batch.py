"""GPU parity of the batched full-window likelihood (glass_hip_logl_batch_*)
against the oracle's gaussian_log_likelihood (utils/src/glass_model.c:
726-774), one oracle call per item.

Windows straddle the kernel's 2048-bin chunks (1, 17, 2047, 2048, 2049,
6000 bins) so single-, multi- and tail-chunk items are all covered; items
sharing data and invC through equal offsets; runs are bitwise repeatable.
Bound: the sums differ from the oracle's only in order, so the difference
is rounding, bounded by 1e-12 x sum over bins of |XX|+|YY|+|ZZ|+2(|XY|+
|XZ|+|YZ|) (the analytic XYZ covariance is nearly singular towards low f,
X+Y+Z carrying little noise, so chi^2 cancels by up to ~1e3 between the
diagonal and cross terms); and relative 1e-10, the likelihood gate.
"""
import numpy as np
import pytest
import torch

from glass_amd import synth

pytestmark = pytest.mark.gpu


def chi2_scale(r, e, moff=None):
    """sum_n |XX| + |YY| + |ZZ| + 2(|XY| + |XZ| + |YZ|) of item e's residual."""
    W, o = int(r["W"][e]), int(r["off"][e])
    mo = o if moff is None else moff
    d = r["data"][3 * o:3 * (o + W)].cpu().numpy().reshape(3, W)
    m = r["model"][3 * mo:3 * (mo + W)].cpu().numpy().reshape(3, W)
    ic = r["invC"][6 * o:6 * (o + W)].cpu().numpy().reshape(6, W)
    x = d - m
    re = lambda a, b: np.abs((a * np.conj(b)).real)
    return float(np.sum(re(x[0], x[0]) * np.abs(ic[0]) + re(x[1], x[1]) * np.abs(ic[1]) +
                        re(x[2], x[2]) * np.abs(ic[2]) + 2 * (re(x[0], x[1]) * np.abs(ic[3]) +
                                                              re(x[0], x[2]) * np.abs(ic[4]) +
                                                              re(x[1], x[2]) * np.abs(ic[5]))))


def oracle_logl(orc, r, e, model=None, moff=None):
    W, o = int(r["W"][e]), int(r["off"][e])
    d = r["data"][3 * o:3 * (o + W)].cpu().numpy().reshape(3, W)
    mo = o if moff is None else moff
    m = (r["model"] if model is None else model)[3 * mo:3 * (mo + W)].cpu().numpy().reshape(3, W)
    ic = r["invC"][6 * o:6 * (o + W)].cpu().numpy().reshape(6, W)
    I9 = np.stack([ic[0], ic[3], ic[4], ic[3], ic[1], ic[5], ic[4], ic[5], ic[2]]).reshape(3, 3, W)
    D = np.zeros((5, W), dtype=np.complex128)
    M = np.zeros((5, W), dtype=np.complex128)
    D[:3], M[:3] = d, m
    return orc.gaussian_log_likelihood(D, M, I9, 3)


def test_logl_batch_ragged_vs_oracle(orc):
    from glass_amd.batch import LoglBatch
    windows = (1, 17, 512, 2047, 2048, 2049, 4096, 6000)
    r = synth.logl_batch_inputs(64, seed=21, windows=windows)
    dev = {k: r[k].cuda() for k in ("data", "model", "invC")}
    B = LoglBatch(r["W"], r["off"], r["off"], r["off"])
    got = B.eval(dev["data"], dev["model"], dev["invC"]).cpu().numpy()
    again = B.eval(dev["data"], dev["model"], dev["invC"]).cpu().numpy()
    assert np.array_equal(got, again)
    ref = np.array([oracle_logl(orc, r, e) for e in range(64)])
    scale = np.array([chi2_scale(r, e) for e in range(64)])
    rel = np.abs(got - ref) / np.abs(ref)
    print("W set", sorted(set(r["W"].tolist())), "max rel", rel.max(),
          "max diff/scale", np.max(np.abs(got - ref) / scale))
    assert np.all(np.abs(got - ref) <= 1e-12 * scale)
    assert rel.max() < 1e-10
    B.close()


def test_logl_batch_shared_data(orc):
    """Sampler layout: several models (chains) against one data window."""
    from glass_amd.batch import LoglBatch
    r = synth.logl_batch_inputs(6, seed=5, windows=(3000,))
    W = np.full(5, 3000, dtype=np.int32)
    off0 = np.zeros(5, dtype=np.int64)
    moff = np.asarray(r["off"][1:], dtype=np.int64)        # models of items 1..5
    B = LoglBatch(W, off0, moff, off0)
    got = B.eval(r["data"].cuda(), r["model"].cuda(), r["invC"].cuda()).cpu().numpy()
    for k in range(5):
        ref = oracle_logl(orc, r, 0, moff=int(moff[k]))
        assert abs(got[k] - ref) <= 1e-12 * chi2_scale(r, 0, moff=int(moff[k])), k
        assert abs(got[k] - ref) < 1e-10 * abs(ref), k
    B.close()


def test_logl_batch_matches_dropin(gl):
    """The batch and the per-call drop-in give the same logL."""
    from glass_amd.batch import LoglBatch
    r = synth.logl_batch_inputs(3, seed=9, windows=(1024, 2500))
    B = LoglBatch(r["W"], r["off"], r["off"], r["off"])
    got = B.eval(r["data"].cuda(), r["model"].cuda(), r["invC"].cuda()).cpu().numpy()
    for e in range(3):
        W, o = int(r["W"][e]), int(r["off"][e])
        d = r["data"][3 * o:3 * (o + W)].numpy().reshape(3, W)
        m = r["model"][3 * o:3 * (o + W)].numpy().reshape(3, W)
        ic = r["invC"][6 * o:6 * (o + W)].numpy().reshape(6, W)
        I9 = np.stack([ic[0], ic[3], ic[4], ic[3], ic[1], ic[5], ic[4], ic[5], ic[2]]).reshape(3, 3, W)
        D = np.zeros((5, W), dtype=np.complex128)
        M = np.zeros((5, W), dtype=np.complex128)
        D[:3], M[:3] = d, m
        ref = gl.gaussian_log_likelihood(D, M, I9, 3)
        assert abs(got[e] - ref) <= 1e-12 * chi2_scale(r, e)
    B.close()


def test_logl_batch_bad_args():
    from glass_amd import GlassHipError
    from glass_amd.batch import LoglBatch
    with pytest.raises(GlassHipError):
        LoglBatch([16, 0], [0, 16], [0, 16], [0, 16])
    with pytest.raises(GlassHipError):
        LoglBatch([16], [-1], [0], [0])
