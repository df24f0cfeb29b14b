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
        for s in range(n):
            assert files.print_source_params(r["params"][s], T) == files.print_source_params(sp[c, s], T)
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
