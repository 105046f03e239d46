#!/usr/bin/env python3
"""Benchmark of the R-tree hot path on one MI355X.

Workload (default): 100,000,000 FP32 boxes drawn from a 64-component Gaussian
mixture (sigma 0.01-0.05) in [0,1]^2 with log-normal(-9, 1) sizes, fanout 16,
64-bit Morton keys; 1,000,000 query windows with side log-uniform
[1e-4, 1e-2].  One step = bulk load of all boxes (keys -> radix sort ->
in-place reorder -> all levels) + one batched window query writing every
(box id, query index) pair.  Inputs are resident in HBM before the timed
region; each step first restores the unsorted boxes with a device copy that
is outside the timed events.

Prints one JSON line.  `value` is bulk-load boxes/s; the query's pairs/s and
per-step latency percentiles are in `query`.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "result pairs/s + p99 latency, served vs plain batch; bulk-load boxes/s"
HBM_PEAK_GBS = 8000.0


def build_bytes_per_box(kbits, fanout):
    """Algorithmic bytes per box of the bulk load, 8-bit-digit LSD sort."""
    kb = kbits // 8
    F = fanout
    return ((16 + kb) + (kb + 2 * kb * (kb + 4)) + (4 + 20 + 20) + (16 + 24 / F)
            + (16 / (F - 1) + 24 / (F * (F - 1))))


def kernel_bytes(name, n, kbits, fanout):
    """Algorithmic bytes of one launch of build kernel `name` over n boxes."""
    kb = kbits // 8
    return {
        "k_keygen": (16 + kb) * n,
        "k_onesweep": 2 * (kb + 4) * n,
        "k_gather_l0": (4 + 20 + 20 + 16 + 24 / fanout) * n,
    }.get(name)


def cpu_baseline(boxes, fanout, key_mode, sample):
    """The oracle's host bulk load (the reference algorithm restated in C),
    single thread, on the first `sample` boxes of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as orc

    sub = tuple(np.ascontiguousarray(a[:sample]) for a in boxes)
    ids = np.arange(sample, dtype=np.int32)
    t0 = time.perf_counter()
    orc.build_rtree(*sub, ids, fanout, key_mode)
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "boxes/s", "cores": 1, "kind": "port",
            "sample": f"oracle host bulk load (Morton key, LSD sort, level reduce) of the first "
                      f"{sample:,} boxes of the same workload, 1 thread, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--boxes", type=int, default=100_000_000)
    ap.add_argument("--windows", type=int, default=1_000_000)
    ap.add_argument("--fanout", type=int, default=16)
    ap.add_argument("--key", choices=["morton64", "ref32"], default="morton64")
    ap.add_argument("--cpu-sample", type=int, default=4_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--query-tile-steps", type=int, default=0,
                    help="wave-steps per query frontier tile (4, 8, 16; 0: library default)")
    ap.add_argument("--query-grid", type=int, default=0,
                    help="blocks per chained query launch (0: library default)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        sys.stderr.write("bench.py measures one GPU only: this build has no multi-GPU path "
                         f"(asked for --gpus {args.gpus}, WORLD_SIZE {world}).\n")
        sys.exit(2)

    import curtree_amd as ct
    from curtree_amd import datagen

    if args.query_tile_steps:
        ct.set_option("query_tile_steps", args.query_tile_steps)
    if args.query_grid:
        ct.set_option("query_grid", args.query_grid)
    key_mode = ct.KEY_MORTON64 if args.key == "morton64" else ct.KEY_REF32
    kbits = 64 if key_mode == ct.KEY_MORTON64 else 32
    n, nq, F = args.boxes, args.windows, args.fanout

    t0 = time.time()
    boxes = datagen.mixture_boxes(n, seed=3)
    wins = datagen.loguniform_windows(nq, seed=4, lo=1e-4, hi=1e-2)
    print(f"[bench] generated {n:,} boxes, {nq:,} windows in {time.time() - t0:.1f} s",
          file=sys.stderr, flush=True)

    ids = np.arange(n, dtype=np.int32)
    pristine = ct.DeviceBoxes.from_host(*boxes, ids)
    work = ct.DeviceBoxes(n)
    q = ct.DeviceBoxes.from_host(*wins)
    nbytes = 4 * n

    def reset():
        for f in ("xmin", "ymin", "xmax", "ymax", "id"):
            ct._check(ct.lib.curtree_memcpy_d2d(getattr(work.s, f), getattr(pristine.s, f),
                                                nbytes, None))

    # size the work buffers and the pair output once, outside the timed region
    reset()
    rt = ct.build_rtree(work, F, key_mode)
    # the reference's traversal (no containment shortcut) gives the work
    # counts of the algorithmic query bytes: X expansions, T tests, H hits
    ct.set_option("containment", 0)
    H = ct.query_count(rt, work, q)
    ref_stats = ct.query_last_stats()
    ct.set_option("containment", -1)
    assert ct.query_count(rt, work, q) == H
    run_stats = ct.query_last_stats()
    rt.free()
    X = sum(ref_stats["entries"][1:])
    # every expanded entry tests all children of its node (the last node of a
    # level may hold fewer than F; the difference is below 1e-6 here)
    T = sum(ref_stats["entries"]) * F
    q_bytes = 16 * nq + 8 * X + 16 * T + 12 * H
    import ctypes
    fbuf, tbuf = ctypes.c_void_p(), ctypes.c_void_p()
    ct._check(ct.lib.curtree_malloc(ctypes.byref(fbuf), 4 * max(H, 1)))
    ct._check(ct.lib.curtree_malloc(ctypes.byref(tbuf), 4 * max(H, 1)))
    print(f"[bench] {H:,} result pairs per query batch", file=sys.stderr, flush=True)

    ev = [ct.Event() for _ in range(3)]
    tree = ct.alloc_tree(n, F)

    def step(record):
        reset()
        ct.synchronize()
        ev[0].record()
        ct.build_into(work, tree, key_mode)
        ev[1].record()
        got = ct.query_into(tree, work, q, fbuf.value, tbuf.value, H)
        ev[2].record()
        ct.synchronize()
        assert got == H, (got, H)
        b, qm = ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])
        if record is not None:
            record.append((b, qm))

    for _ in range(args.warmup):
        step(None)
    print("[bench] warmup done", file=sys.stderr, flush=True)

    ct.Prof.enable(True)
    ct.Prof.reset()
    rec = []
    ct.synchronize()
    w0 = time.perf_counter()
    for i in range(args.steps):
        step(rec)
        print(f"[bench] step {i}: build {rec[-1][0]:.3f} ms, query {rec[-1][1]:.3f} ms",
              file=sys.stderr, flush=True)
    ct.synchronize()
    wall = time.perf_counter() - w0
    ct.Prof.enable(False)

    build_ms = np.array([r[0] for r in rec])
    query_ms = np.array([r[1] for r in rec])
    step_ms = build_ms + query_ms
    kernels = {}
    for name in ("k_build_init", "k_extent", "k_keygen", "k_sort_plan", "k_onesweep",
                 "k_gather_l0", "k_level", "k_safe", "k_run_begin", "k_expand", "k_leaf_count",
                 "k_leaf_pass", "k_write_pieces"):
        tot, cnt = ct.Prof.get(name)
        if cnt:
            kernels[name] = {"avg_ms": tot / cnt, "launches_per_step": cnt / args.steps,
                             "ms_per_step": tot / args.steps}
    # dominant build kernel by time per step, with its algorithmic bytes
    cands = [(v["ms_per_step"], k) for k, v in kernels.items()
             if kernel_bytes(k, n, kbits, F) is not None]
    dom = max(cands)[1]
    dom_bytes = kernel_bytes(dom, n, kbits, F)
    dom_gbs = dom_bytes / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
    bpb = build_bytes_per_box(kbits, F)
    build_gbs = bpb * n / (build_ms.mean() * 1e-3) / 1e9

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        ent = tj.get(dom, {})
        if ent.get("boxes") == n and ent.get("bytes_per_launch"):
            traffic = ent["bytes_per_launch"]

    out = {
        "metric": METRIC,
        "value": n / (build_ms.mean() * 1e-3),
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(step_ms.mean()),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain bulk load + batch window query: {n:,} FP32 boxes "
                               f"(64-component Gaussian mixture, log-normal sizes), fanout {F}, "
                               f"{kbits}-bit Morton keys; {nq:,} windows side log-uniform "
                               f"[1e-4,1e-2]",
                   "boxes": n, "windows": nq, "fanout": F, "key_bits": kbits},
        "build": {"ms": float(build_ms.mean()), "boxes_per_s": n / (build_ms.mean() * 1e-3),
                  "algorithmic_bytes_per_box": bpb, "achieved_GBs": build_gbs,
                  "frac_of_8TBs": build_gbs / HBM_PEAK_GBS},
        "query": {"ms": float(query_ms.mean()), "pairs": int(H),
                  "pairs_per_s": H / (query_ms.mean() * 1e-3),
                  "reference_work": {"X": int(X), "T": int(T), "H": int(H),
                                     "frontier": ref_stats["entries"]},
                  "frontier": run_stats["entries"], "pieces": run_stats["pieces"],
                  "algorithmic_bytes": int(q_bytes),
                  "achieved_GBs": q_bytes / (query_ms.mean() * 1e-3) / 1e9,
                  "step_p50_ms": float(np.percentile(step_ms, 50)),
                  "step_p99_ms": float(np.percentile(step_ms, 99))},
        "kernels": kernels,
        "wall_s": wall,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dom_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(boxes, F, key_mode, min(args.cpu_sample, n))
    ct.lib.curtree_free(fbuf)
    ct.lib.curtree_free(tbuf)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
