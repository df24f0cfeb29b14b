"""Autotuner front end for the batched FD likelihood (glass_hip_fd_batch_*).

The library times every compiled launch shape ("tile": threads per
workgroup, grid multiplier) of each BW class's fused waveform+delta-logL
kernel over a plan's own proposals and keeps the fastest per class
(glass_hip_fd_batch_autotune).  This module adds a JSON cache so a tuned
shape is reused across runs of the same workload on the same device:

    from glass_amd import autotune
    tiles, ms, cached = autotune.tune(plan, n, cache="tiles.json")

The cache key is the device name, the TDI format, the plan's chain count
and the proposal count: a sampler re-creating the same plan shape gets the
tiles without re-timing.

    python -m glass_amd.autotune [--groups 256 --temps 16 --props 80] [--cache FILE]

tunes the batched tempered-sampler workload and prints the chosen tiles.
"""
import argparse
import json
import os
import time


def cache_key(plan, n, fmt="sangria"):
    import torch
    return f"{torch.cuda.get_device_name()}|{fmt}|chains={plan.nchain}|props={int(n)}"


def load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tune(plan, n, cache=None, reps=3, fmt="sangria", retune=False):
    """Tiles for plan over its first n proposals: from the cache when the
    key is there (applied with set_tiles), else measured (and stored).
    Returns (tiles, ms per class or None when cached, from_cache)."""
    key = cache_key(plan, n, fmt)
    db = load(cache)
    if not retune and key in db:
        tiles = [tuple(t) for t in db[key]["tiles"]]
        plan.set_tiles(tiles)
        return tiles, None, True
    tiles, ms = plan.autotune(n, reps=reps)
    if cache:
        db[key] = {"tiles": [list(t) for t in tiles], "ms": [float(x) for x in ms],
                   "tuned": time.strftime("%Y-%m-%dT%H:%M:%S")}
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, cache)
    return tiles, ms, False


def main():
    ap = argparse.ArgumentParser(description="tune the batched FD likelihood tiles")
    ap.add_argument("--groups", type=int, default=256)
    ap.add_argument("--temps", type=int, default=16)
    ap.add_argument("--props", type=int, default=80)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--cache", default=None)
    args = ap.parse_args()
    import glass_amd as gl
    from glass_amd import synth
    from glass_amd.batch import FdBatch, ucb_alignment
    o = gl.analytic_orbit()
    cfg = synth.config1(seed=args.seed, ngroup=args.groups, ntemp=args.temps, nprop_per_chain=args.props,
                        waveform=lambda p, T, BW: gl.ucb_waveform(o, "sangria", T, 0.0, p, 8, BW),
                        alignment=lambda p, T, W, q: ucb_alignment(o, T, W, q, p))
    invC = [gl.invert_noise_covariance_matrix(C)[0] for C in cfg["C"]]
    E = len(cfg["p_chain"])
    plan = FdBatch(o, "sangria", cfg["T"], 0.0, cfg["group_qmin"], cfg["data"], invC, cfg["chain_group"],
                   cfg["chain_nsrc"], cfg["src_params"], E)
    plan.set_proposals(cfg["p_chain"], cfg["p_src"], cfg["p_params"])
    tiles, ms, cached = tune(plan, E, cache=args.cache, reps=args.reps)
    for k, t in enumerate(tiles):
        extra = "" if ms is None else f"  {ms[k]:.4f} ms"
        print(f"BW {32 << k:5d}: {t[0]:5d} threads, grid x{t[1]}{extra}")
    print("from cache" if cached else "measured", flush=True)
    plan.close()


if __name__ == "__main__":
    main()
