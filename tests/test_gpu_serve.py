"""GPU: the serving runtime.  Requests of 1..4096 windows on three resident
trees are submitted while others run; every request's result must equal the
same windows queried alone (plain entry point), pair for pair, with batching
on (4 slots, continuous batching) and off (one request at a time), under a
result budget small enough that batches wait for released results."""
import time

import numpy as np
import pytest

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
    pool = ct.DeviceBoxes.from_host(*datagen.loguniform_windows(200_000, seed=133, lo=1e-4,
                                                                  hi=3e-2))
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
