"""GPU: the serving runtime.  Requests of 1..4096 windows on three resident
trees (uniform, clustered and road-like line-segment MBRs) are submitted while
others run; every request's result must equal the same windows queried alone
(plain entry point) and the oracle's traversal of the same sorted boxes and
tree (src/rtree_bfs_frontier.cuh:60-289 semantics), pair for pair, with
batching on (4 slots, continuous batching) and off (one request at a time),
under a result budget small enough that batches wait for released results."""
import time

import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

pytestmark = pytest.mark.gpu


class _View:
    def __init__(self, s):
        self.s = s


def make_trees(ct):
    sets = [datagen.uniform_boxes(50_000, seed=130, wmax=2e-3),
            datagen.clustered_boxes(80_000, seed=131, frac=0.2, sigma=2e-3, wmax=1e-3),
            datagen.roadlike_boxes(120_000, seed=132)]
    out = []
    for i, bx in enumerate(sets):
        d = ct.DeviceBoxes.from_host(*bx)
        rt = ct.build_rtree(d, (16, 8, 32)[i], ct.KEY_MORTON64)
        out.append((rt, d))
    return out


@pytest.mark.parametrize("batching,slots", [(True, 4), (False, 1)])
def test_results_identical_alone_and_in_the_mix(batching, slots):
    import curtree_amd as ct

    trees = make_trees(ct)
    hpool = datagen.loguniform_windows(200_000, seed=133, lo=1e-4, hi=3e-2)
    pool = ct.DeviceBoxes.from_host(*hpool)
    # host copies of every tree as built (sorted boxes and node arrays)
    host = [(d.to_host(), rt.to_host()) for rt, d in trees]
    r = np.random.default_rng(134)
    nreq = 240
    sizes = np.exp(r.uniform(0, np.log(4096), nreq)).astype(np.int64)
    starts = r.integers(0, 200_000 - 4096, nreq)
    which = r.integers(0, 3, nreq)
    sv = ct.Server(slots=slots, budget_bytes=4 << 20, max_batch_windows=1 << 16,
                   batching=batching)
    ids = [sv.add_tree(rt, d) for rt, d in trees]
    pending = {}
    got = {}
    reported = set()
    k = 0
    t_end = time.time() + 120
    while (k < nreq or pending) and time.time() < t_end:
        # submit in bursts so that requests arrive while others are running
        for _ in range(int(r.integers(1, 12))):
            if k < nreq:
                v = ct.window_view(pool, starts[k], sizes[k])
                pending[sv.submit(ids[which[k]], v)] = k
                k += 1
        # finished requests come from the completion queue, each id once
        for rid in sv.completed():
            assert rid in pending and rid not in reported
            reported.add(rid)
            st, res, lat = sv.poll(rid)
            assert st == ct.REQ_DONE
            n, fp, tp = res
            got[pending.pop(rid)] = (ct._host_from_dev(fp, n, np.int32),
                                     ct._host_from_dev(tp, n, np.int32), lat)
            sv.release(rid)
        time.sleep(0.0005)
    stats = sv.stats()
    sv.close()
    assert len(got) == nreq, (len(got), stats)
    assert stats["requests_done"] == nreq and stats["p99_us"] >= stats["p50_us"] > 0
    if batching:
        assert stats["batches"] < nreq  # requests were merged
    for i in range(nreq):
        rt, d = trees[which[i]]
        fid, tid = ct.bfs_frontier(rt, d, _View(ct.window_view(pool, starts[i], sizes[i])))
        gf, gt, _ = got[i]
        assert np.array_equal(gf, fid) and np.array_equal(gt, tid), i
        sb, t = host[which[i]]
        win = tuple(w[starts[i]:starts[i] + sizes[i]] for w in hpool)
        ofid, otid, _ = orc.bfs(t, sb, win)
        assert np.array_equal(gf, ofid) and np.array_equal(gt, otid), ("oracle", i)


def test_served_roadlike_tree_matches_oracle():
    # the serving workload's tree kinds and window sizes at oracle-feasible
    # sizes: a 2M-box road-like tree at fanout 32 and a 1M-box clustered tree
    # at fanout 100 (one wave per frontier entry), windows of side
    # log-uniform [1e-5, 2e-3]; 150 requests of 1..65,536 windows (log-uniform
    # count) submitted at once, each checked against the oracle's traversal
    import curtree_amd as ct

    sets = [datagen.roadlike_boxes(2_000_000, seed=140),
            datagen.clustered_boxes(1_000_000, seed=141, frac=0.3, k=20, sigma=5e-3, wmax=1e-3)]
    trees, host = [], []
    for bx, F in zip(sets, (32, 100)):
        d = ct.DeviceBoxes.from_host(*bx)
        rt = ct.build_rtree(d, F, ct.KEY_MORTON64)
        trees.append((rt, d))
        host.append((d.to_host(), rt.to_host()))
    npool = 400_000
    hpool = datagen.loguniform_windows(npool, seed=142, lo=1e-5, hi=2e-3)
    pool = ct.DeviceBoxes.from_host(*hpool)
    r = np.random.default_rng(143)
    nreq = 150
    sizes = np.exp(r.uniform(0, np.log(65536), nreq)).astype(np.int64)
    starts = np.array([r.integers(0, npool - k) for k in sizes])
    which = r.integers(0, 2, nreq)
    sv = ct.Server(slots=4, budget_bytes=64 << 20, max_batch_windows=1 << 17, batching=True)
    ids = [sv.add_tree(rt, d) for rt, d in trees]
    req = {sv.submit(ids[which[k]], ct.window_view(pool, starts[k], sizes[k])): k
           for k in range(nreq)}
    got = {}
    t_end = time.time() + 120
    while len(got) < nreq and time.time() < t_end:
        for rid in sv.completed():
            st, res, _ = sv.poll(rid)
            assert st == ct.REQ_DONE
            n, fp, tp = res
            got[req[rid]] = (ct._host_from_dev(fp, n, np.int32), ct._host_from_dev(tp, n, np.int32))
            sv.release(rid)
        time.sleep(0.0005)
    sv.close()
    assert len(got) == nreq
    total = 0
    for k in range(nreq):
        sb, t = host[which[k]]
        win = tuple(w[starts[k]:starts[k] + sizes[k]] for w in hpool)
        ofid, otid, _ = orc.bfs(t, sb, win)
        gf, gt = got[k]
        assert np.array_equal(gf, ofid) and np.array_equal(gt, otid), k
        total += gf.shape[0]
    print(f"[case] {nreq} requests, {int(sizes.sum())} windows, {total} pairs", flush=True)
    assert total > 0
    pool.free()
    for rt, d in trees:
        rt.free()
        d.free()
