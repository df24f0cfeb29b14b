"""Launch shapes ("tiles") of the batched eval kernels and the autotuner.

Every compiled tile of every BW class gives the default tile's logL and
snr up to the order of the final sums (relative 1e-12) and the oracle's to
the 1e-10 gate; the autotuner keeps compiled tiles only, and graph replay
after tuning reproduces the eager evaluation bit for bit.
"""
import numpy as np
import pytest

from test_gpu_batch import TOL, check_props, make_cfg, new_plan, oracle_chains

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gl, orc):
    cfg = make_cfg(orc, seed=11, ngroup=12, ntemp=2, nprop=6)
    E = len(cfg["p_chain"])
    plan = new_plan(gl, cfg, E)
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    l0, s0 = plan.eval(E)
    yield cfg, E, plan, l0.cpu().numpy(), s0.cpu().numpy()
    plan.close()


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def test_candidates_defaults_first():
    from glass_amd.batch import NCLASS, tile_candidates
    for k in range(NCLASS):
        c = tile_candidates(k)
        assert c[0][1] == 0 and len(c) >= 4, (k, c)
        assert all(t[0] % 64 == 0 and t[0] <= 1024 for t in c)


def test_every_tile_same_results(setup):
    from glass_amd.batch import NCLASS, tile_candidates
    cfg, E, plan, l0, s0 = setup
    cands = [tile_candidates(k) for k in range(NCLASS)]
    default = plan.tiles()
    assert default == [c[0] for c in cands]
    nmax = max(len(c) for c in cands)
    for i in range(nmax):
        tiles = [c[min(i, len(c) - 1)] for c in cands]
        plan.set_tiles(tiles)
        l1, s1 = plan.eval(E)
        l1, s1 = l1.cpu().numpy(), s1.cpu().numpy()
        print(i, tiles, "logL rel", rel(l1, l0), "snr rel", rel(s1, s0))
        assert rel(l1, l0) <= 1e-12 and rel(s1, s0) <= 1e-12, (i, tiles)
    plan.set_tiles(default)


def test_nondefault_tiles_vs_oracle(setup, orc):
    from glass_amd.batch import NCLASS, tile_candidates
    cfg, E, plan, l0, s0 = setup
    default = plan.tiles()
    plan.set_tiles([tile_candidates(k)[-1] for k in range(NCLASS)])
    l1, s1 = plan.eval(E)
    cs, _ = oracle_chains(orc, cfg, cfg["src_params"])
    worst = check_props(orc, cfg, cs, l1.cpu().numpy(), s1.cpu().numpy(), range(E))
    print("worst logL rel err, last tiles", worst)
    assert worst <= TOL
    plan.set_tiles(default)


def test_autotune_then_graph(setup):
    import torch
    from glass_amd.batch import tile_candidates
    cfg, E, plan, l0, s0 = setup
    tiles, ms = plan.autotune(E, reps=2)
    assert tiles == plan.tiles()
    for k, t in enumerate(tiles):
        assert tuple(t) in tile_candidates(k)
    print("tuned", tiles, ms)
    l1, s1 = plan.eval(E)
    l1, s1 = l1.cpu().numpy().copy(), s1.cpu().numpy().copy()
    assert rel(l1, l0) <= 1e-12 and rel(s1, s0) <= 1e-12
    for r in range(3):
        plan.logL.zero_()
        l2, s2 = plan.eval(E, graph=True)
        torch.cuda.synchronize()
        torch.randn(1000, device="cuda").sum().item()
        assert np.array_equal(l2.cpu().numpy(), l1) and np.array_equal(s2.cpu().numpy(), s1), r


def test_bad_tiles(setup, gl):
    cfg, E, plan, l0, s0 = setup
    t = plan.tiles()
    bad = list(t)
    bad[4] = (96, 0, 1)
    with pytest.raises(gl.GlassHipError):
        plan.set_tiles(bad)
    bad[4] = (512, 0, 3)
    with pytest.raises(gl.GlassHipError):
        plan.set_tiles(bad)
    bad[4] = (512, -1, 1)
    with pytest.raises(gl.GlassHipError):
        plan.set_tiles(bad)
    assert plan.tiles() == t
