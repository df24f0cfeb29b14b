"""BW-class kernels on concurrent side streams against the same kernels in
turn on one stream: eval (eager and graph), accepts and the template
rebuild give bit-identical logL, snr and chain logL."""
import numpy as np
import pytest

from test_gpu_batch import make_cfg, new_plan

pytestmark = pytest.mark.gpu


def run(gl, cfg, E, concurrent):
    import torch
    plan = new_plan(gl, cfg, E)
    plan.set_concurrent(concurrent)
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    out = []
    l0, s0 = plan.eval(E)
    out += [l0.cpu().numpy().copy(), s0.cpu().numpy().copy()]
    for r in range(3):
        plan.logL.zero_()
        l1, s1 = plan.eval(E, graph=True)
        torch.cuda.synchronize()
        torch.randn(1000, device="cuda").sum().item()
        out += [l1.cpu().numpy().copy(), s1.cpu().numpy().copy()]
    # accept the first proposal of every chain (templates rebuilt by the
    # class kernels), evaluate from the new state, rebuild every template
    nchain = len(cfg["chain_group"])
    acc = [int(np.nonzero(cfg["p_chain"] == c)[0][0]) for c in range(nchain)]
    plan.accept(acc)
    lg, sp = plan.state()
    out += [lg, sp]
    l2, s2 = plan.eval(E)
    out += [l2.cpu().numpy().copy(), s2.cpu().numpy().copy()]
    plan.refresh()
    out.append(plan.state()[0])
    plan.close()
    return out


def test_concurrent_matches_serial(gl, orc):
    cfg = make_cfg(orc, seed=19, ngroup=10, ntemp=2, nprop=6)
    E = len(cfg["p_chain"])
    a = run(gl, cfg, E, True)
    b = run(gl, cfg, E, False)
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), i
    # graph replays equal the eager eval
    for i in range(2, 8):
        assert np.array_equal(a[i], a[i % 2]), i
