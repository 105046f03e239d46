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
per-step latency percentiles are in `query`.  The default run then measures
the rest of the metric, each on its own workload at full size:
  canonical_sort  the device (query id, box id) order of the query's pairs;
  plain_f32  the same boxes and windows at fanout 32 (build + query steps);
  plain_1e6  the 1,000,000-box CPU-runnable workload (build, query, sort);
  batched    the batched entry point on the uneven-window workload below;
  join       the distance self-join of 10,000,000 road-like segment MBRs
             (r = 3e-4);
  serve      served vs plain per-request launches on the 8 resident trees
             below, a shorter request-rate sweep (0.75 s per rate, x1.25 steps);
and `cpu_baseline` times the oracle's host path on this machine (bulk load,
window query, brute force, self-join; 1 thread and every core).
--plain-only skips the legs.

Other modes (same JSON contract, one line):
  --mode batched  50,000,000 boxes (90% uniform + 10% in 5 Gaussian clusters,
                  sigma 1e-3, sides U[0,1e-3]); 4,000,000 windows per launch,
                  side log-uniform [1e-5, 0.05] (0 to ~1e5 hits per window);
                  the batched entry point (count pass, chunks under a pair
                  budget, one graph replay per chunk).  value = pairs/s.
  --mode serve    8 resident trees of 1e7..2e8 boxes (uniform, clustered,
                  road-like segments); Poisson request arrivals, each request
                  on a random tree with 1..65,536 windows (log-uniform count),
                  side log-uniform [1e-5, 2e-3]; 16 GB result budget.  The
                  request rate is swept for plain per-request launches (the
                  plain entry point, one request at a time; and the server
                  with batching off on one stream) and for the serving runtime
                  (continuous batching over 4 streams); value = served pairs/s
                  at the plain entry point's p99 latency (at its highest
                  sustained rate); the comparison with the unbatched server is
                  reported beside it.
  --mode skewed   1,000,000,000 boxes, 70% in 1,000 Zipf(1.2)-weighted hot
                  cells of side 1e-3, 20% points, fanout 16, 64-bit Morton
                  keys (long runs of equal high key halves); value = bulk-load
                  boxes/s; then 10,000,000 queries (half points, half windows
                  of side log-uniform [1e-5, 1e-3], centers uniform) against
                  the tree, in `query`.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "result pairs/s + p99 latency, served vs plain batch; bulk-load boxes/s"
T_START = time.time()
HBM_PEAK_GBS = 8000.0


def build_bytes_per_box(kbits, fanout):
    """Algorithmic bytes per box of the bulk load, 8-bit-digit LSD sort."""
    kb = kbits // 8
    F = fanout
    return ((16 + kb) + (kb + 2 * kb * (kb + 4)) + (4 + 20 + 20) + (16 + 24 / F)
            + (16 / (F - 1) + 24 / (F * (F - 1))))


def kernel_bytes(name, n, kbits, fanout):
    """Algorithmic bytes of one launch of build kernel `name` over n boxes.

    The sort passes carry the payload: each reads and writes the box id and
    the 16-byte box (40 B/box) and computes the sort key from the box it
    read.  The first pass reads the caller's arrays, the last writes them.
    The histogram pass reads the boxes.
    The final pass works in place on the caller's arrays: it reads the sorted
    boxes and ids and writes the leaf-parent nodes (the few boxes that its tie
    sort moves are not counted).  The whole build is also reported against the
    formula of build_bytes_per_box, which prices the reference's algorithm."""
    return {
        "k_extent": 16 * n,
        "k_hist": 16 * n,
        "k_sweep_kv": 40 * n,
        "k_final": (16 + 4 + 24 / fanout) * n,
    }.get(name)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _parallel(fn, nq, threads):
    """fn(q0, q1) over `threads` slices of [0, nq), in threads (the oracle's C
    calls release the GIL); returns the list of results."""
    from concurrent.futures import ThreadPoolExecutor

    cuts = np.linspace(0, nq, threads + 1).astype(np.int64)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda k: fn(int(cuts[k]), int(cuts[k + 1])), range(threads)))


def cpu_baseline(boxes, wins, fanout, key_mode, sample, join_boxes=None, join_r=0.0):
    """The reference's CPU path restated in the oracle (oracle/oracle.c), timed
    on this host: the host R-tree bulk load (1 thread) of the workload, which is
    the headline row; then the host R-tree window query on the same tree and
    windows, the brute-force nested loop (src/utility.h:22-45) on the 1e6-box
    plain workload, and the distance self-join -- each with 1 thread and with
    every core, on bounded samples whose rates extrapolate linearly."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as orc

    nproc = os.cpu_count() or 1
    rows = []
    sub = tuple(np.ascontiguousarray(a[:sample]) for a in boxes)
    ids = np.arange(sample, dtype=np.int32)
    t0 = time.perf_counter()
    sb, tree = orc.build_rtree(*sub, ids, fanout, key_mode)
    dt = time.perf_counter() - t0
    head = {"value": sample / dt, "unit": "boxes/s", "cores": 1, "kind": "port",
            "sample": f"oracle host bulk load (Morton key, LSD sort, level reduce) of the first "
                      f"{sample:,} boxes of the same workload, 1 thread, {dt:.2f} s",
            "nproc": nproc, "cpu_model": _cpu_model()}
    rows.append({"what": "host R-tree bulk load", "threads": 1, "items": sample, "wall_s": dt,
                 "rate": sample / dt, "unit": "boxes/s"})
    del sub
    # host R-tree query (the oracle's traversal) on the same tree
    if sample == boxes[0].shape[0]:
        nq = wins[0].shape[0]
        for threads, cnt in ((1, min(nq, 100_000)), (nproc, nq)):
            w = tuple(a[:cnt] for a in wins[:4])
            t0 = time.perf_counter()
            r = _parallel(lambda q0, q1: orc.bfs_count(tree, sb, tuple(a[q0:q1] for a in w)),
                          cnt, threads)
            dt = time.perf_counter() - t0
            H = sum(x[2] for x in r)
            rows.append({"what": "host R-tree window query (BFS)", "threads": threads,
                         "windows": cnt, "pairs": H, "wall_s": dt, "rate": H / dt,
                         "unit": "pairs/s"})
    del sb, tree
    # brute force on the plain CPU-runnable workload: 1e6 boxes x 1e4 windows
    from curtree_amd import datagen

    b1 = datagen.uniform_boxes(1_000_000, seed=1, wmax=1e-3)
    w1 = datagen.uniform_windows(10_000, seed=2, smax=0.01)
    for threads, cnt in ((1, 2_000), (nproc, 10_000)):
        t0 = time.perf_counter()
        H = orc.brute_count(b1, w1, 0, cnt, threads=threads)
        dt = time.perf_counter() - t0
        rows.append({"what": "brute force nested loop, 1e6 boxes", "threads": threads,
                     "windows": cnt, "tests": cnt * 1_000_000, "pairs": H, "wall_s": dt,
                     "rate": H / dt, "unit": "pairs/s", "tests_per_s": cnt * 1e6 / dt})
    # brute force on the workload's own boxes, a sample of windows, all cores
    if sample == boxes[0].shape[0]:
        cnt = 64
        t0 = time.perf_counter()
        H = orc.brute_count(boxes, wins, 0, cnt, threads=nproc)
        dt = time.perf_counter() - t0
        rows.append({"what": f"brute force nested loop, {sample:,} boxes", "threads": nproc,
                     "windows": cnt, "tests": cnt * sample, "pairs": H, "wall_s": dt,
                     "rate": H / dt, "unit": "pairs/s", "tests_per_s": cnt * sample / dt})
    # distance self-join: host R-tree of the join boxes, queried with every
    # box expanded by r (a sample of the boxes as queries)
    if join_boxes is not None:
        nj = join_boxes[0].shape[0]
        t0 = time.perf_counter()
        jsb, jt = orc.build_rtree(*join_boxes, np.arange(nj, dtype=np.int32), fanout, 1)
        tb = time.perf_counter() - t0
        r32 = np.float32(join_r)
        for threads, cnt in ((1, min(nj, 200_000)), (nproc, nj)):
            jw = (jsb[0][:cnt] - r32, jsb[1][:cnt] - r32, jsb[2][:cnt] + r32, jsb[3][:cnt] + r32)
            t0 = time.perf_counter()
            r = _parallel(lambda q0, q1: orc.bfs_count(jt, jsb, tuple(a[q0:q1] for a in jw)),
                          cnt, threads)
            dt = time.perf_counter() - t0
            H = sum(x[2] for x in r)
            rows.append({"what": f"host R-tree distance self-join, {nj:,} boxes, r={join_r:g}",
                         "threads": threads, "query_boxes": cnt, "pairs": H, "wall_s": dt,
                         "build_s_1thread": tb, "rate": H / dt, "unit": "pairs/s"})
    head["rows"] = rows
    return head


def _batched_tuning(args, ct):
    tf = os.path.join(ROOT, "curtree_amd", "tuning.json")
    if args.no_tuning or not os.path.exists(tf):
        return None
    with open(tf) as f:
        return json.load(f)["tuned"]["batched"]


def batched_leg(args, ct, datagen, n, nq, steps, warmup):
    """The batched entry point on the uneven-window workload: n boxes (90%
    uniform + 10% in 5 Gaussian clusters, sigma 1e-3, sides U[0,1e-3]), nq
    windows of side log-uniform [1e-5, 0.05]; one step = count pass + one
    graph replay per chunk of at most budget_pairs pairs.  Returns a dict."""
    tuning = _batched_tuning(args, ct)
    saved = None
    if tuning:
        ct.set_option("query_tile_steps", tuning["tile_steps"])
        ct.set_option("query_grid", tuning["grid"])
    F = args.fanout if (args.fanout and args.mode == "batched") else (
        tuning["fanout"] if tuning else 16)
    t0 = time.time()
    boxes = datagen.clustered_boxes(n, seed=5, frac=0.1, k=5, sigma=1e-3, wmax=1e-3)
    wins = datagen.loguniform_windows(nq, seed=6, lo=1e-5, hi=0.05)
    print(f"[bench] batched: generated {n:,} boxes, {nq:,} windows in {time.time() - t0:.1f} s",
          file=sys.stderr, flush=True)
    d = ct.DeviceBoxes.from_host(*boxes, np.arange(n, dtype=np.int32))
    del boxes
    rt = ct.build_rtree(d, F, ct.KEY_MORTON64)
    q = ct.DeviceBoxes.from_host(*wins)
    del wins

    def on_chunk(fid, tid, npairs, q0, q1):
        return True  # pairs stay on the device; a consumer would read them here

    def step():
        ct.synchronize()
        t = time.perf_counter()
        st = ct.query_batched(rt, d, q, args.budget_pairs, on_chunk)
        ct.synchronize()
        return time.perf_counter() - t, st

    for _ in range(warmup):
        step()
    rec = []
    for i in range(steps):
        dt, st = step()
        rec.append((dt, st))
        print(f"[bench] batched step {i}: {dt * 1e3:.1f} ms, {st['pairs']:,} pairs in "
              f"{st['chunks']} chunks (count {st['count_ms']:.1f} ms, write "
              f"{st['write_ms']:.1f} ms)", file=sys.stderr, flush=True)
    ms = np.array([r[0] for r in rec]) * 1e3
    st = rec[-1][1]
    total, counts = ct.query_count(rt, d, q, per_query=True)
    assert total == st["pairs"], (total, st["pairs"])
    rt.free()
    d.free()
    q.free()
    if tuning:  # back to the plain workload's knobs
        ct.set_option("query_tile_steps", args.plain_tuning["tile_steps"] if args.plain_tuning
                      else 8)
        ct.set_option("query_grid", args.plain_tuning["grid"] if args.plain_tuning else 0)
    return {"workload": f"batched uneven windows: {n:,} boxes (90% uniform + 10% in 5 Gaussian "
                        f"clusters), {nq:,} windows side log-uniform [1e-5,0.05], fanout {F}, "
                        f"budget {args.budget_pairs:,} pairs per chunk",
            "boxes": n, "windows": nq, "fanout": F, "budget_pairs": args.budget_pairs,
            "tuning": tuning and {k: tuning[k] for k in ("fanout", "tile_steps", "grid")},
            "pairs": int(st["pairs"]), "pairs_per_s": st["pairs"] / (ms.mean() * 1e-3),
            "ms": float(ms.mean()), "chunks": int(st["chunks"]),
            "max_chunk_pairs": int(st["max_chunk_pairs"]),
            "count_ms": float(st["count_ms"]), "write_ms": float(st["write_ms"]),
            "hits_per_window": {"max": int(counts.max()), "mean": float(counts.mean()),
                                "zero_frac": float((counts == 0).mean())},
            "step_p50_ms": float(np.percentile(ms, 50)),
            "step_p99_ms": float(np.percentile(ms, 99)), "steps": steps}


def bench_batched(args, ct, datagen):
    n = args.boxes if args.boxes_set else 50_000_000
    nq = args.windows if args.windows_set else 4_000_000
    b = batched_leg(args, ct, datagen, n, nq, args.steps, args.warmup)
    out = {
        "metric": METRIC, "value": b["pairs_per_s"], "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": b["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": b["workload"], "boxes": n, "windows": nq, "fanout": b["fanout"],
                   "budget_pairs": args.budget_pairs, "tuning": b["tuning"]},
        "batched": b,
    }
    print(json.dumps(out), flush=True)


def _poisson_run(ct, sv, tree_ids, pool, npool, lam, duration, seed, drain_s=20.0):
    """Poisson arrivals at `lam` requests/s for `duration` s; latency from the
    arrival time to the client seeing the completion."""
    r = np.random.default_rng(seed)
    sv.reset_stats()
    t0 = time.perf_counter()
    next_t = t0 + r.exponential(1.0 / lam)
    arrivals = {}
    lat = []
    offered = 0
    t_end = t0
    while True:
        now = time.perf_counter()
        while next_t <= now and next_t - t0 < duration:
            k = int(np.exp(r.uniform(0, np.log(65536))))
            start = int(r.integers(0, npool - k))
            tree = int(r.integers(0, len(tree_ids)))
            rid = sv.submit(tree_ids[tree], ct.window_view(pool, start, k))
            arrivals[rid] = next_t
            offered += 1
            next_t += r.exponential(1.0 / lam)
        done_ids = sv.completed()
        if done_ids:
            t = time.perf_counter()
            for rid in done_ids:
                lat.append(t - arrivals.pop(rid))
                sv.release(rid)  # the pairs were consumed on the device
            t_end = t
        if now - t0 >= duration and not arrivals:
            break
        if now - t0 >= duration + drain_s:
            break  # overloaded: the backlog does not drain
        time.sleep(0)  # poll without a timer sleep (as the plain entry point's loop spins)
    done = len(lat)
    sst = sv.stats()
    pairs = sst["pairs_done"]
    lat_ms = np.array(lat) * 1e3 if lat else np.array([np.inf])
    span = max(t_end - t0, 1e-9)
    return {"rate": lam, "offered": offered, "done": done, "pairs": int(pairs),
            "pairs_per_s": pairs / span, "p50_ms": float(np.percentile(lat_ms, 50)),
            "p99_ms": float(np.percentile(lat_ms, 99)),
            "batches": int(sst["batches"]), "launch_busy": sst["launch_us"] * 1e-6 / span,
            "kept_up": bool(done == offered and span <= duration * 1.2 + 0.5)}


class _Win:
    """A window view in the shape the query wrappers take (`.s`)."""

    def __init__(self, s):
        self.s = s


def _plain_entry_run(ct, trees, pool, npool, lam, duration, seed, buf, cap, drain_s=20.0):
    """The same Poisson request stream served by the plain entry point, one
    request at a time (curtree_query_into into a preallocated buffer),
    latency from the arrival time to the call's return."""
    r = np.random.default_rng(seed)
    t0 = time.perf_counter()
    next_t = t0 + r.exponential(1.0 / lam)
    lat, pairs, offered, t_end = [], 0, 0, t0
    while next_t - t0 < duration:
        while time.perf_counter() < next_t:
            pass
        k = int(np.exp(r.uniform(0, np.log(65536))))
        start = int(r.integers(0, npool - k))
        rt, d, _ = trees[int(r.integers(0, len(trees)))]
        v = _Win(ct.window_view(pool, start, k))
        got = ct.query_into(rt, d, v, buf[0], buf[1], cap)
        if got < 0:  # more pairs than the buffer: an allocating query
            p = ct.query(rt, d, v)
            got = p.n
            p.free()
        t = time.perf_counter()
        lat.append(t - next_t)
        pairs += got
        offered += 1
        t_end = t
        next_t += r.exponential(1.0 / lam)
        if t - t0 > duration + drain_s:
            break
    lat_ms = np.array(lat) * 1e3 if lat else np.array([np.inf])
    span = max(t_end - t0, 1e-9)
    return {"rate": lam, "offered": offered, "done": len(lat), "pairs": int(pairs),
            "pairs_per_s": pairs / span, "p50_ms": float(np.percentile(lat_ms, 50)),
            "p99_ms": float(np.percentile(lat_ms, 99)),
            "kept_up": bool(span <= duration * 1.2 + 0.5 and next_t - t0 >= duration)}


def _knee(points):
    """Highest rate that kept up without a queue building (median latency
    within 3x of the lightest load's)."""
    p50_0 = points[0]["p50_ms"]
    ok = [p for p in points if p["kept_up"] and p["p50_ms"] <= 3 * p50_0]
    return (ok or points[:1])[-1]


def _at_p99(points, p99):
    ok = [p for p in points if p["kept_up"] and p["p99_ms"] <= p99]
    return max(ok, key=lambda p: p["pairs_per_s"]) if ok else None


def bench_skewed(args, ct, datagen):
    """Bulk load of the large/skewed workload: 70% of the boxes in 1,000
    Zipf(1.2)-weighted hot cells of side 1e-3, 20% points, 64-bit Morton keys
    (long runs of equal high key halves); then 1e7 queries, half points and
    half windows of side log-uniform [1e-5, 1e-3], centers uniform, against
    the built tree (`query` in the JSON line; `value` is the bulk load)."""
    n = args.boxes if args.boxes_set else 1_000_000_000
    nq = args.windows if args.windows_set else 10_000_000
    F = args.fanout or 16
    t0 = time.time()

    def progress(done):
        print(f"[bench] generated {done:,} of {n:,} boxes", file=sys.stderr, flush=True)

    boxes = datagen.skewed_boxes_large(n, seed=9, progress=progress)
    pristine = ct.DeviceBoxes.from_host(*boxes, np.arange(n, dtype=np.int32))
    del boxes
    work = ct.DeviceBoxes(n)
    tree = ct.alloc_tree(n, F)
    print(f"[bench] {n:,} skewed boxes resident in {time.time() - t0:.1f} s", file=sys.stderr,
          flush=True)
    ev = [ct.Event() for _ in range(2)]

    def step():
        for f in ("xmin", "ymin", "xmax", "ymax", "id"):
            ct._check(ct.lib.curtree_memcpy_d2d(getattr(work.s, f), getattr(pristine.s, f),
                                                4 * n, None))
        ct.synchronize()
        ev[0].record()
        ct.build_into(work, tree, ct.KEY_MORTON64)
        ev[1].record()
        ct.synchronize()
        return ev[0].elapsed_ms(ev[1])

    for _ in range(args.warmup):
        step()
    ms = []
    for i in range(args.steps):
        ms.append(step())
        print(f"[bench] step {i}: build {ms[-1]:.2f} ms", file=sys.stderr, flush=True)
    ct.Prof.enable(True)
    ct.Prof.reset()
    for _ in range(args.prof_steps):
        step()
    ct.Prof.enable(False)
    kernels = {}
    for name in ("k_build_init", "k_extent", "k_hist", "k_sort_plan", "k_sweep_kv", "k_final",
                 "k_run_len", "k_fixup_short", "k_fixup_long", "k_fixup_huge", "k_fixup_groups",
                 "k_fallback", "k_level"):
        tot, cnt = ct.Prof.get(name)
        if cnt:
            kernels[name] = {"avg_ms": tot / cnt, "ms_per_step": tot / max(args.prof_steps, 1)}
    ms = np.array(ms)
    # the queries: count pass to size the output, then timed query steps
    # into a caller buffer (the tree built by the last step)
    q = ct.DeviceBoxes.from_host(*datagen.mixed_queries(nq, seed=10))
    H = ct.query_count(tree, work, q)
    print(f"[bench] {H:,} result pairs per query batch", file=sys.stderr, flush=True)
    fbuf, tbuf = ctypes.c_void_p(), ctypes.c_void_p()
    ct._check(ct.lib.curtree_malloc(ctypes.byref(fbuf), 4 * max(H, 1)))
    ct._check(ct.lib.curtree_malloc(ctypes.byref(tbuf), 4 * max(H, 1)))
    qms = []
    for i in range(args.warmup + args.steps):
        ct.synchronize()
        ev[0].record()
        got = ct.query_into(tree, work, q, fbuf.value, tbuf.value, H)
        ev[1].record()
        ct.synchronize()
        assert got == H, (got, H)
        if i >= args.warmup:
            qms.append(ev[0].elapsed_ms(ev[1]))
            print(f"[bench] query step {i - args.warmup}: {qms[-1]:.2f} ms", file=sys.stderr,
                  flush=True)
    ct.lib.curtree_free(fbuf)
    ct.lib.curtree_free(tbuf)
    q.free()
    qms = np.array(qms)
    out = {
        "metric": "bulk-load throughput (large/skewed workload)",
        "value": n / (ms.mean() * 1e-3),
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(ms.mean()),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"bulk load: {n:,} FP32 boxes, 70% in 1,000 Zipf(1.2) hot cells "
                               f"of side 1e-3, 20% points, fanout {F}, 64-bit Morton keys; "
                               f"{nq:,} queries, half points, half windows of side log-uniform "
                               f"[1e-5,1e-3], centers uniform",
                   "boxes": n, "queries": nq, "fanout": F, "key_bits": 64},
        "build": {"ms": float(ms.mean()), "ms_p50": float(np.percentile(ms, 50)),
                  "ms_max": float(ms.max())},
        "query": {"ms": float(qms.mean()), "pairs": int(H),
                  "pairs_per_s": H / (qms.mean() * 1e-3),
                  "queries_per_s": nq / (qms.mean() * 1e-3)},
        "kernels": kernels,
    }
    tree.free()
    work.free()
    pristine.free()
    print(json.dumps(out), flush=True)


def serve_trees(args, ct, datagen, scale):
    """The 8 resident trees of the serving workload: 1e7..2e8 boxes (times
    `scale`), uniform, clustered and road-like line-segment MBRs in turn."""
    sizes = [int(x * scale) for x in (1e7, 2e7, 4e7, 6e7, 8e7, 1.2e8, 1.6e8, 2e8)]
    kinds = ["uniform", "clustered", "roadlike"]
    trees = []
    t0 = time.time()
    for i, n in enumerate(sizes):
        kind = kinds[i % 3]
        if kind == "uniform":
            bx = datagen.uniform_boxes(n, seed=70 + i, wmax=1e-3)
        elif kind == "clustered":
            bx = datagen.clustered_boxes(n, seed=70 + i, frac=0.3, k=20, sigma=5e-3, wmax=1e-3)
        else:
            bx = datagen.roadlike_boxes(n, seed=70 + i)
        d = ct.DeviceBoxes.from_host(*bx, np.arange(n, dtype=np.int32))
        del bx
        rt = ct.build_rtree(d, args.serve_fanout, ct.KEY_MORTON64)
        trees.append((rt, d, kind))
        print(f"[bench] serve tree {i}: {n:,} {kind} boxes ({time.time() - t0:.0f} s)",
              file=sys.stderr, flush=True)
    return sizes, trees


def serve_leg(args, ct, datagen, scale, seconds, rate_step, drain_s, full=False):
    """Served vs plain per-request launches on the 8 resident trees: Poisson
    arrivals at a swept rate, requests of 1..65,536 windows (log-uniform
    count), side log-uniform [1e-5, 2e-3], result budget serve_budget_gb.
    The baselines' rates rise by rate_step until they no longer keep up; the
    served pairs/s are compared at each baseline's knee p99.  full: the served
    rate also rises by rate_step from serve_rate0 to serve_rate_max (the whole
    curve); otherwise it starts at the lower baseline knee and rises by
    serve_fine_step until its p99 passes 1.5x the baselines' knee p99 at two
    rates in a row (or 3x at one), which is the part of the curve the
    equal-p99 comparison reads."""
    sizes, trees = serve_trees(args, ct, datagen, scale)
    npool = 4_000_000
    pool = ct.DeviceBoxes.from_host(*datagen.loguniform_windows(npool, seed=80, lo=1e-5, hi=2e-3))
    budget = args.serve_budget_gb << 30
    res = {}

    def sweep(name, run, lam, step, stop):
        pts = []
        while lam <= args.serve_rate_max:
            p = run(lam)
            pts.append(p)
            print(f"[bench] {name} {lam:.0f} req/s: {p['pairs_per_s'] / 1e9:.2f} Gpairs/s, "
                  f"p50 {p['p50_ms']:.2f} ms, p99 {p['p99_ms']:.2f} ms, kept up {p['kept_up']}"
                  + (f", {p['batches']} batches, dispatcher launching {p['launch_busy']:.0%}"
                     if "batches" in p else ""), file=sys.stderr, flush=True)
            if not p["kept_up"] or stop(p):
                break
            lam *= step
        return pts

    def server_sweep(name, lam0, step, stop, **kw):
        sv = ct.Server(budget_bytes=budget, max_batch_windows=args.serve_batch_windows, **kw)
        ids = [sv.add_tree(rt, d) for rt, d, _ in trees]
        # warm-up at a rate that makes batches large enough to grow every buffer
        # once (a growth frees device memory, which waits for the whole device)
        _poisson_run(ct, sv, ids, pool, npool, 2000.0, 1.0, seed=1, drain_s=drain_s)
        pts = sweep(name, lambda lam: _poisson_run(ct, sv, ids, pool, npool, lam, seconds,
                                                   seed=int(lam), drain_s=drain_s),
                    lam0, step, stop)
        sv.close()
        return pts

    # the baselines first (their sweeps start lower, so that the knee lies inside)
    res["plain"] = server_sweep("plain", args.serve_rate0 / 4, rate_step, lambda p: False,
                                slots=1, batching=False)
    # the plain entry point per request (what a caller does without the server)
    cap = 256 << 20
    fb, tb = ctypes.c_void_p(), ctypes.c_void_p()
    ct._check(ct.lib.curtree_malloc(ctypes.byref(fb), 4 * cap))
    ct._check(ct.lib.curtree_malloc(ctypes.byref(tb), 4 * cap))
    _plain_entry_run(ct, trees, pool, npool, 2000.0, 1.0, 1, (fb.value, tb.value), cap,
                     drain_s=drain_s)
    pts = sweep("plain entry point",
                lambda lam: _plain_entry_run(ct, trees, pool, npool, lam, seconds, int(lam),
                                             (fb.value, tb.value), cap, drain_s=drain_s),
                args.serve_rate0 / 4, rate_step, lambda p: False)
    ct.lib.curtree_free(fb)
    ct.lib.curtree_free(tb)
    res["plain_entry"] = pts
    ref = _knee(res["plain"])
    ref_e = _knee(res["plain_entry"])
    if full:
        res["served"] = server_sweep("served", args.serve_rate0, rate_step, lambda p: False,
                                     slots=args.serve_slots, batching=True)
    else:
        # stop past the baselines' knee p99: two points in a row above 1.5x it, or one
        # above 3x (a single hiccup of the host does not end the sweep)
        p99_max = max(ref["p99_ms"], ref_e["p99_ms"])
        over = [0]

        def past_knee(p):
            over[0] = over[0] + 1 if p["p99_ms"] > 1.5 * p99_max else 0
            return over[0] >= 2 or p["p99_ms"] > 3 * p99_max

        res["served"] = server_sweep("served", min(ref["rate"], ref_e["rate"]),
                                     args.serve_fine_step, past_knee,
                                     slots=args.serve_slots, batching=True)
    pool.free()
    for rt, d, _ in trees:
        rt.free()
        d.free()
    # equal-p99 comparisons: served pairs/s at each baseline's knee p99
    best = _at_p99(res["served"], ref["p99_ms"])
    best_e = _at_p99(res["served"], ref_e["p99_ms"])
    value = best_e["pairs_per_s"] if best_e else 0.0
    return {
        "workload": "serving: 8 resident trees of " + "/".join(f"{n / 1e6:g}M" for n in sizes)
                    + " boxes (uniform, clustered, road-like), Poisson arrivals, 1..65536 "
                    "windows per request (log-uniform), side log-uniform [1e-5,2e-3], "
                    f"{args.serve_budget_gb} GB result budget",
        "trees": sizes, "fanout": args.serve_fanout, "seconds_per_rate": seconds,
        "rate_step": rate_step, "served_rate_step": rate_step if full else args.serve_fine_step,
        "served_pairs_per_s": value,
        "plain": res["plain"], "served": res["served"], "plain_entry": res["plain_entry"],
        "vs_plain_entry_point": {
            "baseline": "curtree_query_into per request, one at a time",
            "knee_rate": ref_e["rate"], "p99_ref_ms": ref_e["p99_ms"],
            "plain_pairs_per_s": ref_e["pairs_per_s"],
            "served_pairs_per_s_at_equal_p99": value,
            "served_p99_ms": best_e["p99_ms"] if best_e else None,
            "ratio": value / ref_e["pairs_per_s"] if ref_e["pairs_per_s"] else None},
        "vs_unbatched_server": {
            "baseline": "the server with batching off and one slot",
            "knee_rate": ref["rate"], "p99_ref_ms": ref["p99_ms"],
            "plain_pairs_per_s": ref["pairs_per_s"],
            "served_pairs_per_s_at_equal_p99": best["pairs_per_s"] if best else 0.0,
            "served_p99_ms": best["p99_ms"] if best else None,
            "ratio": (best["pairs_per_s"] / ref["pairs_per_s"])
                     if best and ref["pairs_per_s"] else None}}


def bench_serve(args, ct, datagen):
    sv = serve_leg(args, ct, datagen, args.serve_scale, args.serve_seconds, args.serve_rate_step,
                   20.0, full=True)
    out = {
        "metric": METRIC, "value": sv["served_pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
        "steps": len(sv["plain"]) + len(sv["served"]), "warmup": 1,
        "ms_per_step": args.serve_seconds * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": sv["workload"], "trees": sv["trees"], "fanout": sv["fanout"]},
        "serve": sv,
    }
    print(json.dumps(out), flush=True)


def fanout_leg(ct, pristine, work, q, n, F, key_mode, kbits, H, steps, warmup, fbuf, tbuf):
    """Build + query steps of the plain workload at another fanout (same boxes
    and windows, so the same H pairs)."""
    tree = ct.alloc_tree(n, F)
    ev = [ct.Event() for _ in range(3)]
    rec = []
    for i in range(warmup + steps):
        for f in ("xmin", "ymin", "xmax", "ymax", "id"):
            ct._check(ct.lib.curtree_memcpy_d2d(getattr(work.s, f), getattr(pristine.s, f),
                                                4 * n, None))
        ct.synchronize()
        ev[0].record()
        ct.build_into(work, tree, key_mode)
        ev[1].record()
        got = ct.query_into(tree, work, q, fbuf.value, tbuf.value, H)
        ev[2].record()
        ct.synchronize()
        assert got == H, (got, H)
        if i >= warmup:
            rec.append((ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])))
    tree.free()
    b = np.array([r[0] for r in rec])
    qm = np.array([r[1] for r in rec])
    bpb = build_bytes_per_box(kbits, F)
    gbs = bpb * n / (b.mean() * 1e-3) / 1e9
    print(f"[bench] fanout {F}: build {b.mean():.3f} ms, query {qm.mean():.3f} ms",
          file=sys.stderr, flush=True)
    return {"fanout": F, "steps": steps, "build_ms": float(b.mean()),
            "boxes_per_s": n / (b.mean() * 1e-3), "algorithmic_bytes_per_box": bpb,
            "achieved_GBs": gbs, "frac_of_8TBs": gbs / HBM_PEAK_GBS,
            "query_ms": float(qm.mean()), "pairs": int(H),
            "pairs_per_s": H / (qm.mean() * 1e-3)}


def canonical_leg(ct, tree, work, q, H, fbuf, tbuf, steps):
    """The device canonical (query id, box id) order of a query's H pairs:
    each step writes the pairs again (untimed) and sorts them in place."""
    ev = [ct.Event() for _ in range(2)]
    ms = []
    for _ in range(steps):
        assert ct.query_into(tree, work, q, fbuf.value, tbuf.value, H) == H
        ct.synchronize()
        ev[0].record()
        ct._check(ct.lib.curtree_pairs_canonical_sort(fbuf, tbuf, int(H), None))
        ev[1].record()
        ct.synchronize()
        ms.append(ev[0].elapsed_ms(ev[1]))
    ms = np.array(ms)
    # one more, untimed, with an event pair around every launch
    assert ct.query_into(tree, work, q, fbuf.value, tbuf.value, H) == H
    ct.Prof.enable(True)
    ct.Prof.reset()
    ct._check(ct.lib.curtree_pairs_canonical_sort(fbuf, tbuf, int(H), None))
    ct.Prof.enable(False)
    kernels = {}
    for name in ("k_seg_find", "k_seg_split", "k_seg_wave", "k_seg_block", "k_seg_block16",
                 "k_seg_huge"):
        tot, cnt = ct.Prof.get(name)
        if cnt:
            kernels[name] = tot
    print(f"[bench] canonical sort of {H:,} pairs: {ms.mean():.2f} ms {kernels}", file=sys.stderr,
          flush=True)
    return {"pairs": int(H), "ms": float(ms.mean()), "pairs_per_s": H / (ms.mean() * 1e-3),
            "steps": steps, "kernels_ms": kernels}


def small_leg(ct, datagen, steps):
    """The CPU-runnable plain workload on the GPU: 1,000,000 boxes (centers
    uniform, sides U[0,1e-3]), fanout 16, 1e4 windows of side U[0,0.01];
    bulk load, query and canonical sort timed separately."""
    n, nq, F = 1_000_000, 10_000, 16
    bx = datagen.uniform_boxes(n, seed=1, wmax=1e-3)
    pristine = ct.DeviceBoxes.from_host(*bx, np.arange(n, dtype=np.int32))
    work = ct.DeviceBoxes(n)
    q = ct.DeviceBoxes.from_host(*datagen.uniform_windows(nq, seed=2, smax=0.01))
    tree = ct.alloc_tree(n, F)
    ev = [ct.Event() for _ in range(4)]
    fb, tb = ctypes.c_void_p(), ctypes.c_void_p()
    H = None
    rec = []
    for i in range(steps + 1):
        for f in ("xmin", "ymin", "xmax", "ymax", "id"):
            ct._check(ct.lib.curtree_memcpy_d2d(getattr(work.s, f), getattr(pristine.s, f),
                                                4 * n, None))
        ct.synchronize()
        ev[0].record()
        ct.build_into(work, tree, ct.KEY_MORTON64)
        ev[1].record()
        if H is None:
            H = ct.query_count(tree, work, q)
            ct._check(ct.lib.curtree_malloc(ctypes.byref(fb), 4 * max(H, 1)))
            ct._check(ct.lib.curtree_malloc(ctypes.byref(tb), 4 * max(H, 1)))
            ev[1].record()
        got = ct.query_into(tree, work, q, fb.value, tb.value, H)
        ev[2].record()
        ct._check(ct.lib.curtree_pairs_canonical_sort(fb, tb, int(H), None))
        ev[3].record()
        ct.synchronize()
        assert got == H
        if i:
            rec.append([ev[j].elapsed_ms(ev[j + 1]) for j in range(3)])
    for p_ in (fb, tb):
        ct.lib.curtree_free(p_)
    tree.free()
    work.free()
    pristine.free()
    q.free()
    r = np.array(rec).mean(axis=0)
    return {"workload": f"{n:,} FP32 boxes (centers uniform, sides U[0,1e-3]), fanout {F}, "
                        f"64-bit Morton keys; {nq:,} windows side U[0,0.01]",
            "boxes": n, "windows": nq, "pairs": int(H), "build_ms": float(r[0]),
            "query_ms": float(r[1]), "canonical_sort_ms": float(r[2]),
            "boxes_per_s": n / (r[0] * 1e-3), "pairs_per_s": H / (r[1] * 1e-3), "steps": steps}


def join_leg(ct, datagen, n, r, fanout, steps):
    """Distance self-join on road-like segment MBRs (the reference drivers'
    workload shape: every data box expanded by r queries the tree of the same
    boxes): GPU bulk load once, then timed self-join calls (allocating their
    pairs, as curtree_self_join_f32 does).  Returns the dict and the host boxes
    (for the CPU row)."""
    bx = datagen.roadlike_boxes(n, seed=90)
    d = ct.DeviceBoxes.from_host(*bx, np.arange(n, dtype=np.int32))
    rt = ct.build_rtree(d, fanout, ct.KEY_MORTON64)
    ev = [ct.Event() for _ in range(2)]
    ms, H = [], 0
    for i in range(steps + 1):
        ct.synchronize()
        ev[0].record()
        p = ct.self_join(rt, d, r)
        ev[1].record()
        ct.synchronize()
        H = p.n
        p.free()
        if i:
            ms.append(ev[0].elapsed_ms(ev[1]))
    rt.free()
    d.free()
    ms = np.array(ms)
    print(f"[bench] self-join: {n:,} boxes, r={r:g}: {H:,} pairs in {ms.mean():.2f} ms",
          file=sys.stderr, flush=True)
    return {"workload": f"distance self-join: {n:,} road-like segment MBRs, r = {r:g}, fanout "
                        f"{fanout}", "boxes": n, "r": r, "pairs": int(H), "ms": float(ms.mean()),
            "pairs_per_s": H / (ms.mean() * 1e-3), "steps": steps}, bx


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--prof-steps", type=int, default=3,
                    help="untimed steps (after the warm-up) with per-kernel events for the "
                         "kernel table")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["plain", "batched", "serve", "skewed"], default="plain",
                    help="skewed: bulk load of the large/skewed workload only (1e9 boxes)")
    ap.add_argument("--boxes", type=int, default=None)
    ap.add_argument("--windows", type=int, default=None)
    ap.add_argument("--budget-pairs", type=int, default=256 << 20,
                    help="batched mode: pairs per chunk (8 bytes each)")
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="serve mode: tree sizes relative to 1e7..2e8 boxes")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-rate0", type=float, default=2000.0)
    ap.add_argument("--serve-rate-max", type=float, default=64000.0)
    ap.add_argument("--serve-rate-step", type=float, default=1.2)
    ap.add_argument("--serve-fine-step", type=float, default=1.08,
                    help="served rate step of the default run's serve leg (from the baselines' "
                         "knee up to 1.5x their knee p99)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-batch-windows", type=int, default=1 << 20,
                    help="serve mode: window cap of one batch (and of one request)")
    ap.add_argument("--serve-budget-gb", type=int, default=16)
    ap.add_argument("--fanout", type=int, default=None,
                    help="default: 16 (plain), 32 (serve), the tuned one (batched)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="library defaults instead of curtree_amd/tuning.json (plain, batched)")
    ap.add_argument("--key", choices=["morton64", "ref32"], default="morton64")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="boxes timed in the CPU baseline (0: the whole workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="plain mode: only the 100M-box plain workload (no fanout-32, batched, "
                         "self-join or serving legs)")
    ap.add_argument("--leg-steps", type=int, default=3,
                    help="plain mode: timed steps of the fanout-32, batched and self-join legs")
    ap.add_argument("--leg-serve-seconds", type=float, default=0.75,
                    help="plain mode: seconds per request rate in the serving leg")
    ap.add_argument("--leg-serve-rate-step", type=float, default=1.25,
                    help="plain mode: request rate factor between points of the serving leg")
    ap.add_argument("--query-tile-steps", type=int, default=0,
                    help="wave-steps per query frontier tile (4, 8, 16; 0: library default)")
    ap.add_argument("--query-grid", type=int, default=0,
                    help="blocks per chained query launch (0: library default)")
    ap.add_argument("--box-staging", type=int, default=None, choices=[-1, 0, 1],
                    help="AoS copies of boxes, node MBRs and windows for a query: -1 auto "
                         "(library default), 0 off, 1 on")
    ap.add_argument("--containment", type=int, default=None, choices=[-1, 0, 1],
                    help="the containment shortcut: -1 auto (library default), 0 off, 1 on")
    ap.add_argument("--leaf-fused", type=int, default=None, choices=[-1, 0, 1],
                    help="writing leaf stage: -1 by run size (library default), 0 count/scan/"
                         "write, 1 one chained pass")
    ap.add_argument("--expand-split", type=int, default=-1, choices=[-1, 0, 1],
                    help="internal levels: -1 by run size (library default), 0 chained scan, "
                         "1 count/scan/write")
    ap.add_argument("--window-order", type=int, default=None, choices=[-1, 0, 1],
                    help="windows of a large batch taken in Morton order of their centres: -1 "
                         "by batch size (library default), 0 input order, 1 always")
    ap.add_argument("--leaf-codes", type=int, default=None, choices=[-1, 0, 1],
                    help="leaf codes of the tree (8-bit cells per box in its leaf-parent's "
                         "frame) for the leaf count: -1 built by a query of 65,536 windows or "
                         "more (library default), 0 off, 1 always")
    ap.add_argument("--tile-xcd", type=int, default=None, choices=[0, 1],
                    help="query tiles walked in contiguous eighths, one per XCD: 0 off (library "
                         "default), 1 on")
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
    if args.expand_split != -1:
        ct.set_option("expand_split", args.expand_split)
    if args.leaf_fused is not None:
        ct.set_option("leaf_fused", args.leaf_fused)
    if args.box_staging is not None:
        ct.set_option("box_staging", args.box_staging)
    if args.window_order is not None:
        ct.set_option("window_order", args.window_order)
    if args.tile_xcd is not None:
        ct.set_option("tile_xcd", args.tile_xcd)
    if args.leaf_codes is not None:
        ct.set_option("leaf_codes", args.leaf_codes)
    args.boxes_set = args.boxes is not None
    args.windows_set = args.windows is not None
    if args.mode == "plain" and args.fanout is None:
        args.fanout = 16
    args.plain_tuning = None
    tf = os.path.join(ROOT, "curtree_amd", "tuning.json")
    if args.mode == "plain" and not args.no_tuning and os.path.exists(tf):
        # the tuned query knobs (the fanout stays the workload's)
        with open(tf) as f:
            tp = json.load(f)["tuned"].get("plain")
        if tp:
            args.plain_tuning = {k: tp[k] for k in ("tile_steps", "grid")}
            if not args.query_tile_steps:
                ct.set_option("query_tile_steps", tp["tile_steps"])
            if not args.query_grid:
                ct.set_option("query_grid", tp["grid"])
    args.serve_fanout = args.fanout if (args.mode == "serve" and args.fanout) else 32
    if args.mode == "batched":
        return bench_batched(args, ct, datagen)
    if args.mode == "serve":
        return bench_serve(args, ct, datagen)
    if args.mode == "skewed":
        return bench_skewed(args, ct, datagen)
    if not args.boxes_set:
        args.boxes = 100_000_000
    if not args.windows_set:
        args.windows = 1_000_000
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
    ct.set_option("containment", args.containment if args.containment is not None else -1)
    assert ct.query_count(rt, work, q) == H
    run_stats = ct.query_last_stats()
    rt.free()
    X = sum(ref_stats["entries"][1:])
    # every expanded entry tests all children of its node (the last node of a
    # level may hold fewer than F; the difference is below 1e-6 here)
    T = sum(ref_stats["entries"]) * F
    q_bytes = 16 * nq + 8 * X + 16 * T + 12 * H
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

    # the per-kernel table: untimed steps with an event pair around every launch
    prof_steps = max(1, args.prof_steps)
    ct.Prof.enable(True)
    ct.Prof.reset()
    for _ in range(prof_steps):
        step(None)
    ct.synchronize()
    ct.Prof.enable(False)
    kernels = {}
    for name in ("k_build_init", "k_extent", "k_hist", "k_sort_plan", "k_sweep_kv", "k_final",
                 "k_run_len", "k_fixup_short", "k_fixup_long", "k_fixup_huge", "k_fixup_groups",
                 "k_fallback", "k_level",
                 "k_safe", "k_stage4", "k_win_keys", "k_pair_pack", "k_onesweep_pairs",
                 "k_pair_unpack", "k_win_stage", "k_win_offsets", "k_run_begin", "k_expand", "k_expand_count", "k_expand_write",
                 "k_leaf_codes", "k_leaf_count", "k_tile_scan", "k_leaf_fused",
                 "k_leaf_write", "k_write_pieces"):
        tot, cnt = ct.Prof.get(name)
        if cnt:
            kernels[name] = {"avg_ms": tot / cnt, "launches_per_step": cnt / prof_steps,
                             "ms_per_step": tot / prof_steps}
    # dominant build kernel by time per step, with its algorithmic bytes
    cands = [(v["ms_per_step"], k) for k, v in kernels.items()
             if kernel_bytes(k, n, kbits, F) is not None]
    dom = max(cands)[1]
    dom_bytes = kernel_bytes(dom, n, kbits, F)

    # timed steps: events around the dominant kernel's launches only (its
    # average duration over the timed region gives the roofline)
    ct.Prof.only(dom)
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
    dom_tot, dom_cnt = ct.Prof.get(dom)
    ct.Prof.only(None)
    dom_ms = dom_tot / dom_cnt
    dom_gbs = dom_bytes / (dom_ms * 1e-3) / 1e9

    build_ms = np.array([r[0] for r in rec])
    query_ms = np.array([r[1] for r in rec])
    step_ms = build_ms + query_ms
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
                   "boxes": n, "windows": nq, "fanout": F, "key_bits": kbits,
                   "tuning": args.plain_tuning},
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
                     "algorithmic_bytes_per_launch": dom_bytes,
                     "avg_launch_ms": dom_ms, "launches_timed": dom_cnt},
    }
    # the query's dominant kernel over its average duration in the per-kernel
    # table, priced at what it moves: per (window, leaf-parent) entry of the
    # run's leaf level the entry (8 B), the record it writes (16 B) and, with
    # the tree's leaf codes, the leaf-parent's frame (16 B) and 4 B of cells
    # per child slot (next power of two >= F), else 16 B per box test (F per
    # entry); also priced at the reference traversal's leaf level (8 B per
    # entry, 16 B per box test, containment off), the survey's figure
    lc = kernels.get("k_leaf_count")
    codes_on = kernels.get("k_leaf_codes") is not None or (
        args.leaf_codes == 1 or (args.leaf_codes in (None, -1) and nq >= 65536))
    if lc:
        run_entries = run_stats["entries"][-1]
        g = 1 << max(F - 1, 1).bit_length()
        coded = codes_on and F <= 64 and g >= 4
        lbytes = run_entries * (8 + 16 + (16 + 4 * g if coded else 16 * F))
        lgbs = lbytes / (lc["avg_ms"] * 1e-3) / 1e9
        ref_entries = ref_stats["entries"][-1]
        rbytes = 8 * ref_entries + 16 * F * ref_entries
        ltraffic = None
        if os.path.exists(os.path.join(ROOT, "profiles", "traffic.json")):
            with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
                ent = json.load(f).get("k_leaf_count", {})
            if ent.get("boxes") == n and ent.get("bytes_per_launch"):
                ltraffic = ent["bytes_per_launch"]
        out["roofline_query"] = {"bound": "hbm", "kernel": "k_leaf_count", "achieved": lgbs,
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": lgbs / HBM_PEAK_GBS, "traffic": ltraffic,
                                 "algorithmic_bytes_per_launch": int(lbytes),
                                 "leaf_codes": coded,
                                 "reference_priced_bytes": int(rbytes),
                                 "reference_priced_GBs": rbytes / (lc["avg_ms"] * 1e-3) / 1e9,
                                 "avg_launch_ms": lc["avg_ms"]}
    join_boxes = None
    join_r = 3e-4
    out["canonical_sort"] = canonical_leg(ct, tree, work, q, H, fbuf, tbuf, 2)
    if not args.plain_only:
        # the rest of the metric, each on its own workload at full size
        out["plain_f32"] = fanout_leg(ct, pristine, work, q, n, 32, key_mode, kbits, H,
                                      args.leg_steps, 1, fbuf, tbuf)
    ct.lib.curtree_free(fbuf)
    ct.lib.curtree_free(tbuf)
    tree.free()
    work.free()
    pristine.free()
    q.free()
    if not args.plain_only:
        out["plain_1e6"] = small_leg(ct, datagen, args.leg_steps)
        out["batched"] = batched_leg(args, ct, datagen, 50_000_000, 4_000_000, args.leg_steps, 1)
        out["join"], join_boxes = join_leg(ct, datagen, 10_000_000, join_r, F, args.leg_steps)
        out["serve"] = serve_leg(args, ct, datagen, 1.0, args.leg_serve_seconds,
                                 args.leg_serve_rate_step, 3.0)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(boxes, wins, F, key_mode,
                                           min(args.cpu_sample, n) if args.cpu_sample else n,
                                           join_boxes, join_r)
        if out.get("join"):
            allc = [r for r in out["cpu_baseline"]["rows"]
                    if r["what"].startswith("host R-tree distance self-join")
                    and r["query_boxes"] == out["join"]["boxes"]]
            if allc:  # the host path over every box finds the GPU's pairs
                out["join"]["cpu_pairs_equal"] = allc[-1]["pairs"] == out["join"]["pairs"]
    out["wall_s_total"] = time.time() - T_START
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
