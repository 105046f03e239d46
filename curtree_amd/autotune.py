"""Autotuner for the bulk load + window query on one GPU.

Picks the fanout F (tree shape), the frontier tile size (wave-steps per
tile of the chained query kernels) and the grid of the chained launches
(blocks that loop over tiles, i.e. waves per CU) by timing every
combination on a workload, with HIP events around the library calls.

Every configuration returns the same pairs (the knobs change scheduling and
tree shape, not the result set; the pair order is (query, sorted box
position) for a given F): the tuner checks that the pair count agrees across
all of them.

  python -m curtree_amd.autotune --mode batched --boxes 50000000 \\
      --windows 500000 --out tuning.json

Modes:
  plain    objective = build + one plain query into a caller buffer (ms)
  batched  objective = pairs/s of the batched entry point (count pass,
           chunks under a pair budget, one graph replay per chunk); the
           tree is built once per fanout and not timed.
"""
import argparse
import ctypes
import itertools
import json
import sys
import time

import numpy as np

import curtree_amd as ct
from curtree_amd import datagen

FANOUTS = (8, 16, 32, 64)
TILE_STEPS = (4, 8, 16)
GRIDS = (0, 2048, 4096, 8192, 16384)  # 0: by batch size


def _time_ms(fn, repeats):
    ev = [ct.Event(), ct.Event()]
    best = float("inf")
    out = None
    for _ in range(repeats):
        ct.synchronize()
        ev[0].record()
        out = fn()
        ev[1].record()
        ct.synchronize()
        best = min(best, ev[0].elapsed_ms(ev[1]))
    return best, out


def tune(boxes, windows, mode="plain", fanouts=FANOUTS, tile_steps=TILE_STEPS, grids=GRIDS,
         key_mode=None, budget_pairs=256 << 20, repeats=2, log=None):
    """Time every (fanout, tile steps, grid) on host arrays `boxes`
    (xmin, ymin, xmax, ymax) and `windows`; returns a dict with the best
    configuration and all measurements."""
    key_mode = ct.KEY_MORTON64 if key_mode is None else key_mode
    n = boxes[0].size
    ids = np.arange(n, dtype=np.int32)
    pristine = ct.DeviceBoxes.from_host(*boxes, ids)
    work = ct.DeviceBoxes(n)
    q = ct.DeviceBoxes.from_host(*windows)

    def reset():
        for f in ("xmin", "ymin", "xmax", "ymax", "id"):
            ct._check(ct.lib.curtree_memcpy_d2d(getattr(work.s, f), getattr(pristine.s, f),
                                                4 * n, None))

    cus = ct.device_info()["cus"] or 256
    rows = []
    pairs_seen = set()
    try:
        for F in fanouts:
            tree = ct.alloc_tree(n, F)
            build_ms = float("inf")
            for _ in range(2):  # the first build of a shape also sizes the workspace
                reset()
                ms, _ = _time_ms(lambda: ct.build_into(work, tree, key_mode), 1)
                build_ms = min(build_ms, ms)
            H = ct.query_count(tree, work, q)
            pairs_seen.add(H)
            buf = None
            if mode == "plain":
                buf = ctypes.c_void_p()
                ct._check(ct.lib.curtree_malloc(ctypes.byref(buf), 8 * max(H, 1)))
            for st, g in itertools.product(tile_steps, grids):
                ct.set_option("query_tile_steps", st)
                ct.set_option("query_grid", g)
                if mode == "plain":
                    fn = lambda: ct.query_into(tree, work, q, buf.value, buf.value + 4 * H, H)
                    ms, got = _time_ms(fn, repeats)
                    row = {"fanout": F, "tile_steps": st, "grid": g, "build_ms": build_ms,
                           "query_ms": ms, "objective_ms": build_ms + ms, "pairs": got}
                else:
                    fn = lambda: ct.query_batched(tree, work, q, budget_pairs)
                    ms, stt = _time_ms(fn, repeats)
                    got = stt["pairs"]
                    row = {"fanout": F, "tile_steps": st, "grid": g, "build_ms": build_ms,
                           "query_ms": ms, "pairs_per_s": got / (ms * 1e-3), "pairs": got,
                           "chunks": stt["chunks"]}
                row["waves_per_cu"] = g * (ct.kQThreads // 64) / cus if g else "auto"
                if got != H:
                    raise RuntimeError(f"configuration {row} returned {got} pairs, expected {H}")
                rows.append(row)
                if log:
                    log(row)
            if buf is not None:
                ct.lib.curtree_free(buf)
            tree.free()
    finally:
        ct.set_option("query_tile_steps", 8)
        ct.set_option("query_grid", 0)
    if len(pairs_seen) != 1:
        raise RuntimeError(f"pair counts differ between fanouts: {sorted(pairs_seen)}")
    if mode == "plain":
        best = min(rows, key=lambda r: r["objective_ms"])
    else:
        best = max(rows, key=lambda r: r["pairs_per_s"])
    return {"mode": mode, "boxes": n, "windows": int(windows[0].size), "best": best,
            "rows": rows, "device": ct.device_info()}


def apply(result):
    """Set the tuned query knobs (tile steps, grid) process-wide; returns the
    tuned fanout (a tree-build choice the caller makes)."""
    b = result["best"]
    ct.set_option("query_tile_steps", b["tile_steps"])
    ct.set_option("query_grid", b["grid"])
    return b["fanout"]


def _workload(mode, n, nq, seed):
    if mode == "batched":
        return (datagen.clustered_boxes(n, seed=seed, frac=0.1, k=5, sigma=1e-3, wmax=1e-3),
                datagen.loguniform_windows(nq, seed=seed + 1, lo=1e-5, hi=0.05))
    return (datagen.mixture_boxes(n, seed=seed),
            datagen.loguniform_windows(nq, seed=seed + 1, lo=1e-4, hi=1e-2))


def main(argv=None):
    ap = argparse.ArgumentParser(description="Tune fanout, frontier tile size and launch grid.")
    ap.add_argument("--mode", choices=["plain", "batched"], default="batched")
    ap.add_argument("--boxes", type=int, default=50_000_000)
    ap.add_argument("--windows", type=int, default=500_000)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--fanouts", default="8,16,32,64")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    boxes, wins = _workload(a.mode, a.boxes, a.windows, a.seed)
    t0 = time.time()
    res = tune(boxes, wins, mode=a.mode, fanouts=tuple(int(x) for x in a.fanouts.split(",")),
               repeats=a.repeats,
               log=lambda r: print("[autotune]", json.dumps(r), file=sys.stderr, flush=True))
    res["tuning_s"] = time.time() - t0
    print(json.dumps(res["best"]))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
