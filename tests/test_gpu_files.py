"""Checkpoint of a batched plan's device state in the reference's
chain_state_<c>.dat format (save_chain_state, ucb/src/glass_ucb_io.c:418-446)
and back: every chain's file holds its live sources and logL as the plan
holds them (to the printed precision), and a plan rebuilt from the restored
parameters (restore_chain_state's regeneration) has the same chain logL to
the precision the files keep."""
import numpy as np
import pytest

from test_gpu_batch import make_cfg, new_plan

pytestmark = pytest.mark.gpu


def test_save_and_restore_chain_states(gl, orc, tmp_path):
    from glass_amd import files
    cfg = make_cfg(orc, seed=23, ngroup=6, ntemp=2, nprop=4)
    E = len(cfg["p_chain"])
    plan = new_plan(gl, cfg, E)
    logl, sp = plan.state()
    nchain = len(cfg["chain_group"])
    lnorm = np.linspace(-5.0, 5.0, nchain)
    files.save_chain_states(plan, tmp_path, step=7, logLmax=float(logl.max()), logLnorm=lnorm, t0=0.0,
                            eta=[1.0, 1.0, 1.0])
    T = cfg["T"]
    restored = np.zeros_like(sp)
    for c in range(nchain):
        r = files.restore_chain_state(tmp_path / f"chain_state_{c}.dat", T, 3, maxsrc=sp.shape[1])
        n = int(cfg["chain_nsrc"][c])
        assert r["step"] == 7 and r["nlive"] == n
        assert r["logL"] == float("%g" % logl[c]) and r["logLnorm"] == float("%g" % lnorm[c])
        # to the file's precision: f0 16 digits, the rest 12 (A through ln A)
        a, b = sp[c, :n], r["params"]
        assert np.all(np.abs(b[:, 0] - a[:, 0]) <= 1e-15 * np.abs(a[:, 0]))
        for k in (1, 2, 4, 5, 6, 7):
            assert np.all(np.abs(b[:, k] - a[:, k]) <= 1e-11 * np.abs(a[:, k]))
        assert np.all(np.abs(np.exp(b[:, 3]) / np.exp(a[:, 3]) - 1.0) <= 1e-11)
        restored[c, :n] = r["params"]
    plan.close()
    # regenerate from the restored parameters, as restore_chain_state does
    cfg2 = dict(cfg)
    cfg2["src_params"] = restored
    plan2 = new_plan(gl, cfg2, E)
    logl2, _ = plan2.state()
    plan2.close()
    rel = np.max(np.abs(logl2 - logl) / np.maximum(np.abs(logl), 1.0))
    print("restored logL rel", rel)
    assert rel < 1e-6


def test_sampler_checkpoint_files_follow_temperature(gl, orc, tmp_path):
    """After ptmcmc swaps have permuted the slot -> chain map, the sampler's
    checkpoint writes file t of ladder l from the chain now at temperature t
    (save_chain_state writes model[chain->index[ic]], glass_ucb_io.c:418-446)."""
    from glass_amd import files
    from test_gpu_sampler import setup_sampler
    cfg, plan, smp, pr, seeds, temps = setup_sampler(gl, orc, seed=61, ngroup=3, ntemp=4)
    smp.fisher()
    K = smp.ntemp
    for _ in range(12):
        smp.sweep(nsteps=4, adapt=False)
        st = smp.state()
        if not np.array_equal(st["index"], np.arange(smp.nslot)):
            break
    idx = st["index"]
    assert not np.array_equal(idx, np.arange(smp.nslot)), "no swap happened"
    logl, sp = plan.state()
    dirs = [tmp_path / f"ladder{l}" for l in range(smp.nladder)]
    for d in dirs:
        d.mkdir()
    lmax = np.arange(smp.nladder, dtype=float)
    smp.save_chain_states([str(d) for d in dirs], step=3, logLmax=lmax, t0=0.0, eta=[1.0, 1.0, 1.0])
    T = cfg["T"]
    for l in range(smp.nladder):
        for t in range(K):
            c = int(idx[l * K + t])
            r = files.restore_chain_state(dirs[l] / f"chain_state_{t}.dat", T, 3, maxsrc=sp.shape[1])
            n = int(cfg["chain_nsrc"][c])
            assert r["nlive"] == n and r["logLmax"] == lmax[l]
            assert r["logL"] == float("%g" % logl[c])
            assert np.all(np.abs(r["params"][:, 0] - sp[c, :n, 0]) <= 1e-15 * np.abs(sp[c, :n, 0]))
    smp.close()
    plan.close()
