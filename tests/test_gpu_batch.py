"""GPU: the batched path (count pass, chunk plan under a pair budget, one
captured hipGraph replayed per chunk) delivers exactly the plain query's pairs,
in the same order, for windows with wildly uneven result counts.  Every replay
is checked against the eager result, with eager work between replays."""
import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

pytestmark = pytest.mark.gpu


def collect(ct, rt, d, q, budget):
    got = []

    def on_chunk(fid, tid, n, q0, q1):
        f = ct._host_from_dev(fid, n, np.int32)
        t = ct._host_from_dev(tid, n, np.int32)
        got.append((f, t, q0, q1))
        # eager work between replays: a small plain query on the same tree
        ct.query_count(rt, d, q)
        return True

    st = ct.query_batched(rt, d, q, budget, on_chunk)
    return got, st


@pytest.mark.parametrize("fanout", [16, 8])
def test_batched_equals_plain(fanout):
    import curtree_amd as ct

    n = 200_000
    bx = datagen.clustered_boxes(n, seed=120 + fanout, frac=0.1, sigma=1e-3, wmax=1e-3)
    qs = datagen.loguniform_windows(20_000, seed=121, lo=1e-5, hi=0.05)
    d = ct.DeviceBoxes.from_host(*bx, np.arange(n, dtype=np.int32))
    rt = ct.build_rtree(d, fanout, ct.KEY_MORTON64)
    q = ct.DeviceBoxes.from_host(*qs)
    fid, tid = ct.bfs_frontier(rt, d, q)
    total, counts = ct.query_count(rt, d, q, per_query=True)
    assert total == fid.size and counts.max() > 1000 and (counts == 0).any()
    for budget in (max(fid.size // 7, 1), int(counts.max()) // 2, fid.size + 5):
        got, st = collect(ct, rt, d, q, budget)
        assert st["pairs"] == fid.size and st["chunks"] == len(got)
        # consecutive window ranges covering all windows
        assert got[0][2] == 0 and got[-1][3] == 20_000
        assert all(a[3] == b[2] for a, b in zip(got, got[1:]))
        for f, t, q0, q1 in got:
            assert f.size <= max(budget, counts[q0:q1].max())
            assert (t >= q0).all() and (t < q1).all()
        gf = np.concatenate([g[0] for g in got])
        gt = np.concatenate([g[1] for g in got])
        assert np.array_equal(gf, fid) and np.array_equal(gt, tid)
    # against the oracle on a window subset
    sb, t = orc.build_rtree(*bx, np.arange(n, dtype=np.int32), fanout, 1)
    sub = tuple(a[:2000] for a in qs)
    ofid, otid, _ = orc.bfs(t, sb, sub)
    m = tid < 2000
    assert np.array_equal(fid[m], ofid) and np.array_equal(tid[m], otid)
