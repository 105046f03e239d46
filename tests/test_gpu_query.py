"""GPU parity of the window query (through the C ABI).

Bar: bit-exact integer results.  The raw pair list must equal the oracle's
BFS output element for element (same (query, sorted box position) order the
reference produces before its host sort), and the canonical (query id, box id)
sorted pair set must equal brute force (src/utility.h:22-45) for every query
checked."""
import ctypes

import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

pytestmark = pytest.mark.gpu


def setup(boxes, ids, windows, fanout, key_mode):
    import curtree_amd as ct

    d = ct.DeviceBoxes.from_host(*boxes, ids)
    rt = ct.build_rtree(d, fanout, key_mode)
    q = ct.DeviceBoxes.from_host(*windows[:4])
    return d, rt, q


def oracle_tree(boxes, ids, fanout, key_mode):
    return orc.build_rtree(*boxes, ids, fanout, key_mode)


def run_case(boxes, ids, windows, fanout, key_mode, brute_sample=None):
    import curtree_amd as ct

    d, rt, q = setup(boxes, ids, windows, fanout, key_mode)
    fid, tid = ct.bfs_frontier(rt, d, q)
    sb, t = oracle_tree(boxes, ids, fanout, key_mode)
    ofid, otid, cnt = orc.bfs(t, sb, windows)
    assert fid.shape == ofid.shape, (fid.shape, ofid.shape)
    assert np.array_equal(tid, otid) and np.array_equal(fid, ofid)
    if brute_sample == 0:
        # NaN coordinates: a NaN box "intersects" every window in brute force, but
        # the tree reaches it only through its parent's MBR (src/rtree_bfs_frontier.cuh
        # semantics), so parity is against the oracle's traversal only
        return fid, tid, cnt
    nq = windows[0].shape[0]
    sel = np.arange(nq) if brute_sample is None else np.sort(
        np.random.default_rng(0).choice(nq, min(brute_sample, nq), replace=False))
    sub = tuple(w[sel] for w in windows[:4])
    bf, bt = orc.brute_force(tuple(boxes) + (ids,), sub)
    bt = sel[bt].astype(np.int32)
    m = np.isin(tid, sel)
    gf, gt = orc.canonical(fid[m], tid[m])
    cf, ct_ = orc.canonical(bf, bt)
    assert np.array_equal(gt, ct_) and np.array_equal(gf, cf)
    rt.free()
    d.free()
    q.free()
    return fid, tid, cnt


def test_plain_workload_one_million_boxes():
    # 1e6 uniform boxes, F=16, 1e4 windows: the full CPU-runnable workload
    n, nq = 1_000_000, 10_000
    bx = datagen.uniform_boxes(n, seed=1)
    qs = datagen.uniform_windows(nq, seed=2)
    fid, tid, (X, T, H) = run_case(bx, np.arange(n, dtype=np.int32), qs, 16, 0, brute_sample=1000)
    assert H == fid.shape[0] and H > 300_000
    # same pair set with 64-bit Morton keys (different tree, same answer)
    fid2, tid2, _ = run_case(bx, np.arange(n, dtype=np.int32), qs, 16, 1, brute_sample=200)
    a = orc.canonical(fid, tid)
    b = orc.canonical(fid2, tid2)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("fanout", [2, 3, 4, 8, 10, 16, 32, 64])
def test_fanouts_clustered(fanout):
    n = 60_000
    bx = datagen.clustered_boxes(n, seed=30 + fanout, frac=0.3, sigma=0.01, wmax=0.005)
    ids = np.random.default_rng(fanout).permutation(n).astype(np.int32)
    qs = datagen.loguniform_windows(2000, seed=40 + fanout, lo=1e-4, hi=0.1)
    run_case(bx, ids, qs, fanout, 1)


def test_huge_windows_never_overflow():
    # windows that would exceed the reference's 1024-slot LDS queue (it aborts)
    n = 100_000
    bx = datagen.uniform_boxes(n, seed=50)
    qs = datagen.uniform_windows(64, seed=51, smax=1.0)
    qs = (qs[0] - 0.3, qs[1] - 0.3, qs[2] + 0.3, qs[3] + 0.3)
    fid, tid, (X, T, H) = run_case(bx, np.arange(n, dtype=np.int32), qs, 16, 1, brute_sample=8)
    assert H > 64 * 1024


def test_touching_points_and_specials():
    n = 20_000
    r = np.random.default_rng(60)
    x = r.integers(0, 100, n).astype(np.float32) / 100
    y = r.integers(0, 100, n).astype(np.float32) / 100
    bx = (x, y, x, y)  # points on a grid: windows touch them exactly
    g = np.arange(0, 100, 7, dtype=np.float32) / 100
    qx, qy = np.meshgrid(g, g)
    qx, qy = qx.ravel(), qy.ravel()
    qs = (qx, qy, qx + np.float32(0.05), qy + np.float32(0.05))
    for km in (0, 1):
        run_case(bx, np.arange(n, dtype=np.int32), qs, 8, km)


def test_empty_cases():
    import curtree_amd as ct

    bx = datagen.uniform_boxes(1000, seed=61)
    d = ct.DeviceBoxes.from_host(*bx)
    rt = ct.build_rtree(d, 16, 1)
    q0 = ct.DeviceBoxes(0)
    fid, tid = ct.bfs_frontier(rt, d, q0)
    assert fid.size == 0
    far = ct.DeviceBoxes.from_host(*(np.full(5, v, np.float32) for v in (5, 5, 6, 6)))
    fid, tid = ct.bfs_frontier(rt, d, far)
    assert fid.size == 0
    one = ct.DeviceBoxes.from_host(*(np.array([v], np.float32) for v in (0.1, 0.1, 0.2, 0.2)),
                                   np.array([42], np.int32))
    rt1 = ct.build_rtree(one, 16, 0)
    assert rt1.sz == 1
    qa = ct.DeviceBoxes.from_host(*(np.array([v, 9], np.float32) for v in (0.0, 0.0, 0.1, 0.1)))
    fid, tid = ct.bfs_frontier(rt1, one, qa)
    assert fid.tolist() == [42] and tid.tolist() == [0]


def test_query_into_and_counts():
    import curtree_amd as ct

    n = 50_000
    bx = datagen.uniform_boxes(n, seed=70, wmax=0.01)
    qs = datagen.uniform_windows(3000, seed=71, smax=0.05)
    d, rt, q = setup(bx, np.arange(n, dtype=np.int32), qs, 16, 1)
    total, counts = ct.query_count(rt, d, q, per_query=True)
    fid, tid = ct.bfs_frontier(rt, d, q)
    assert total == fid.size
    assert np.array_equal(counts, np.bincount(tid, minlength=3000))
    # too-small caller buffer: nothing written, the needed size is reported
    p = ctypes.c_void_p()
    ct._check(ct.lib.curtree_malloc(ctypes.byref(p), 8 * (total + 1)))
    need = ct.query_into(rt, d, q, p.value, p.value + 4 * (total + 1), total - 1)
    assert need == -total
    got = ct.query_into(rt, d, q, p.value, p.value + 4 * (total + 1), total + 1)
    assert got == total
    f2 = ct._host_from_dev(p.value, total, np.int32)
    t2 = ct._host_from_dev(p.value + 4 * (total + 1), total, np.int32)
    assert np.array_equal(f2, fid) and np.array_equal(t2, tid)
    ct.lib.curtree_free(p)


def test_device_canonical_sort():
    import curtree_amd as ct

    n = 200_000
    bx = datagen.mixture_boxes(n, seed=80)
    ids = np.random.default_rng(80).permutation(n).astype(np.int32) - n // 2  # signed ids
    qs = datagen.loguniform_windows(5000, seed=81, lo=1e-3, hi=0.05)
    d, rt, q = setup(bx, ids, qs, 16, 1)
    p = ct.query(rt, d, q)
    raw_f, raw_t = p.to_host()
    ct.canonical_sort(p)
    f, t = p.to_host()
    cf, ctt = orc.canonical(raw_f, raw_t)
    assert np.array_equal(t, ctt) and np.array_equal(f, cf)
    p.free()


def test_repeated_calls_cycle_chained_scan_epochs():
    # > 127 chained launches: the status words must be recycled correctly
    import curtree_amd as ct

    n = 5000
    bx = datagen.uniform_boxes(n, seed=90, wmax=0.02)
    qs = datagen.uniform_windows(100, seed=91, smax=0.05)
    d, rt, q = setup(bx, np.arange(n, dtype=np.int32), qs, 4, 1)
    ref = ct.bfs_frontier(rt, d, q)
    for _ in range(60):  # 5 levels + 1 leaf pass per call
        got = ct.bfs_frontier(rt, d, q)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_tree_save_load_round_trip(tmp_path):
    import curtree_amd as ct

    n = 30_000
    bx = datagen.mixture_boxes(n, seed=95)
    qs = datagen.loguniform_windows(500, seed=96, lo=1e-3, hi=0.05)
    d, rt, q = setup(bx, np.arange(n, dtype=np.int32), qs, 16, 1)
    ref = ct.bfs_frontier(rt, d, q)
    path = str(tmp_path / "tree.ctree")
    ct.save_tree(path, rt, d)
    rt2, d2 = ct.load_tree(path)
    assert rt2.sz == rt.sz and rt2.fanout == 16 and d2.n == n
    a, b = rt.to_host(), rt2.to_host()
    for k in a:
        assert np.array_equal(a[k], b[k])
    got = ct.bfs_frontier(rt2, d2, q)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("contain,staging", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_containment_shortcut_exact_with_malformed_boxes(contain, staging):
    # inverted and NaN boxes make a subtree unsafe for the shortcut; the
    # result must equal the oracle (which tests every child) either way, with
    # boxes read from the caller's SoA arrays or from the staged AoS copy
    import curtree_amd as ct

    n = 80_000
    bx = [a.copy() for a in datagen.clustered_boxes(n, seed=97, frac=0.5, sigma=0.02, wmax=0.003)]
    r = np.random.default_rng(97)
    inv = r.integers(0, n, 200)
    bx[0][inv], bx[2][inv] = bx[2][inv] + np.float32(0.002), bx[0][inv]  # xmin > xmax
    bx[1][r.integers(0, n, 50)] = np.nan
    qs = datagen.loguniform_windows(3000, seed=98, lo=1e-3, hi=0.3)
    ct.set_option("containment", contain)
    ct.set_option("box_staging", staging)
    try:
        for f, km in ((16, 1), (10, 1), (8, 0)):
            run_case(tuple(bx), np.arange(n, dtype=np.int32), qs, f, km, brute_sample=0)
        # the single-pass leaf stage into a caller buffer gives the same pairs
        d, rt, q = setup(tuple(bx), np.arange(n, dtype=np.int32), qs, 16, 1)
        fid, tid = ct.bfs_frontier(rt, d, q)
        p = ctypes.c_void_p()
        ct._check(ct.lib.curtree_malloc(ctypes.byref(p), 8 * (fid.size + 1)))
        got = ct.query_into(rt, d, q, p.value, p.value + 4 * (fid.size + 1), fid.size + 1)
        assert got == fid.size
        assert np.array_equal(ct._host_from_dev(p.value, got, np.int32), fid)
        assert np.array_equal(ct._host_from_dev(p.value + 4 * (fid.size + 1), got, np.int32), tid)
        ct.lib.curtree_free(p)
    finally:
        ct.set_option("containment", -1)
        ct.set_option("box_staging", -1)


@pytest.mark.parametrize("contain,staging", [(0, 0), (1, 1)])
def test_non_bulk_load_layout_follows_pos_len(contain, staging):
    # Two leaf-parents trade places (pos, len and MBR travel together), so the
    # nodes are no longer in bulk-load order.  The computed-child fast path
    # must notice and follow pos/len, as the oracle does, with the containment
    # shortcut off for such a tree.
    import curtree_amd as ct

    n = 60_000
    bx = datagen.uniform_boxes(n, seed=141, wmax=2e-3)
    qs = datagen.loguniform_windows(2000, seed=142, lo=1e-3, hi=0.3)
    d, rt, q = setup(bx, np.arange(n, dtype=np.int32), qs, 16, 1)
    h = rt.to_host()
    r_sz = rt.sz
    n0 = (n + 15) // 16
    a, b = r_sz - n0 + 3, r_sz - n0 + 1000
    for k in h:
        h[k][a], h[k][b] = h[k][b].copy(), h[k][a].copy()
    rt.upload(h)
    ct.set_option("containment", contain)
    ct.set_option("box_staging", staging)
    try:
        fid, tid = ct.bfs_frontier(rt, d, q)
    finally:
        ct.set_option("containment", -1)
        ct.set_option("box_staging", -1)
    sb = d.to_host()
    ofid, otid, _ = orc.bfs(h, sb, qs)
    assert np.array_equal(fid, ofid) and np.array_equal(tid, otid)


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("fanout,key_mode", [(16, 1), (10, 1), (4, 0), (64, 1)])
def test_expansion_paths(split, fanout, key_mode):
    # internal levels by the chained-scan kernel (0) or by count / tile scan /
    # write with no chain between tiles (1): the same pairs in the same order
    import curtree_amd as ct

    n = 80_000
    bx = datagen.clustered_boxes(n, seed=150 + fanout, frac=0.3, sigma=0.01, wmax=0.005)
    qs = datagen.loguniform_windows(3000, seed=151, lo=1e-4, hi=0.2)
    big = datagen.uniform_windows(64, seed=152, smax=1.0)
    big = (big[0] - 0.3, big[1] - 0.3, big[2] + 0.3, big[3] + 0.3)
    ct.set_option("expand_split", split)
    try:
        run_case(bx, np.arange(n, dtype=np.int32), qs, fanout, key_mode, brute_sample=300)
        # windows over most of the domain: with F=4 the leaf-parent frontier
        # outgrows the first buffers and the run is repeated with larger ones
        run_case(bx, np.arange(n, dtype=np.int32), big, fanout, key_mode, brute_sample=4)
    finally:
        ct.set_option("expand_split", -1)


def test_skewed_mixed_point_and_window_queries():
    # the large/skewed workload's shape at a parity size: skewed boxes (hot
    # cells, 20% points) and a mix of point queries and windows, half of
    # them centred in the hot cells
    n = 300_000
    bx = datagen.skewed_boxes(n, seed=195)
    r = np.random.default_rng(196)
    nq = 4000
    pick = r.integers(0, n, nq // 2)
    cx = np.concatenate([(bx[0][pick] + bx[2][pick]) / 2, r.random(nq // 2).astype(np.float32)])
    cy = np.concatenate([(bx[1][pick] + bx[3][pick]) / 2, r.random(nq // 2).astype(np.float32)])
    side = np.where(r.random(nq) < 0.5, 0.0, 10 ** r.uniform(-5, -2, nq)).astype(np.float32)
    qs = (cx.astype(np.float32) - side / 2, cy.astype(np.float32) - side / 2,
          cx.astype(np.float32) + side / 2, cy.astype(np.float32) + side / 2)
    for fanout, km in ((16, 1), (32, 1)):
        run_case(bx, np.arange(n, dtype=np.int32), qs, fanout, km, brute_sample=400)


@pytest.mark.parametrize("fanout", [16, 32])
def test_mixture_loguniform_query_shape(fanout):
    # the query shape of the 100M-box plain workload at a size the oracle
    # finishes in seconds: 64-component Gaussian-mixture boxes with log-normal
    # sizes, windows of side log-uniform [1e-4, 1e-2], 64-bit Morton keys; the
    # raw pairs equal the oracle's traversal element for element, the
    # canonical pairs of 1,000 sampled windows equal brute force
    n, nq = 1_500_000, 40_000
    bx = datagen.mixture_boxes(n, seed=401)
    ws = datagen.loguniform_windows(nq, seed=402, lo=1e-4, hi=1e-2)
    ids = np.random.default_rng(403).permutation(n).astype(np.int32)
    fid, tid, (X, T, H) = run_case(bx, ids, ws, fanout, 1, brute_sample=1000)
    print(f"[case] F={fanout}: {H} pairs, X={X}, T={T}", flush=True)
    assert H == fid.shape[0] and H > 10 * nq


def test_work_counts_match_oracle_plain_1e6():
    # the CPU-runnable plain workload in full (1e6 boxes, fanout 16, 1e4
    # windows): with the containment shortcut off the GPU frontier holds the
    # same entries at every depth as the oracle's traversal, whose counts are
    # pinned by a closed form; X and T (the query's algorithmic bytes) follow
    import curtree_amd as ct

    n, nq, F = 1_000_000, 10_000, 16
    bx = datagen.uniform_boxes(n, seed=1, wmax=1e-3)
    ws = datagen.uniform_windows(nq, seed=2, smax=0.01)
    ids = np.arange(n, dtype=np.int32)
    d, rt, q = setup(bx, ids, ws, F, 1)
    ct.set_option("containment", 0)
    try:
        H = ct.query_count(rt, d, q)
        st = ct.query_last_stats()
    finally:
        ct.set_option("containment", -1)
    sb = d.to_host()
    t = rt.to_host()
    _, _, (X, T, Ho), depths = orc.bfs(t, sb, ws, depths=True)
    print(f"[case] depths {depths}, X={X}, T={T}, H={Ho}", flush=True)
    assert st["entries"] == depths and H == Ho == st["pairs"]
    assert X == sum(st["entries"][1:])
    # T: every entry tests all of its node's children, F but for the last
    # node of a level (the bench prices T as F per entry)
    Tf = F * sum(st["entries"])
    assert T <= Tf and (Tf - T) < 1e-3 * T
    rt.free()
    d.free()
    q.free()


@pytest.mark.parametrize("fanout", [8, 16, 64])
def test_persistent_algorithm_same_pairs(fanout):
    # CURTREE_ALGO_PERSISTENT: every level one chained launch of ticketed
    # blocks (the leaf stage included); windows large enough for contained
    # ranges longer than a flat tile (pieces) and for frontier regrowth
    import curtree_amd as ct

    n, nq = 400_000, 3_000
    bx = datagen.mixture_boxes(n, seed=410 + fanout)
    ws = datagen.loguniform_windows(nq, seed=420 + fanout, lo=1e-4, hi=0.2)
    ids = np.random.default_rng(fanout).permutation(n).astype(np.int32)
    d, rt, q = setup(bx, ids, ws, fanout, 1)
    p = ct.query(rt, d, q, algo=ct.ALGO_PERSISTENT)
    fid, tid = p.to_host()
    p.free()
    sb = d.to_host()
    ofid, otid, _ = orc.bfs(rt.to_host(), sb, ws)
    print(f"[case] F={fanout}: {ofid.shape[0]} pairs", flush=True)
    assert np.array_equal(fid, ofid) and np.array_equal(tid, otid)
    rt.free()
    d.free()
    q.free()


@pytest.mark.parametrize("fused", [0, 1])
def test_leaf_stage_paths_large_batch(fused):
    # the writing leaf stage as one chained pass (1) or count / scan / write
    # (0), forced on a batch above the automatic cut (32,768 windows), with
    # contained ranges of every length and a caller buffer
    import curtree_amd as ct

    n, nq = 600_000, 40_000
    bx = datagen.clustered_boxes(n, seed=430, frac=0.3, sigma=2e-3, wmax=1e-3)
    ws = datagen.loguniform_windows(nq, seed=431, lo=1e-4, hi=0.1)
    d, rt, q = setup(bx, np.arange(n, dtype=np.int32), ws, 16, 1)
    ct.set_option("leaf_fused", fused)
    try:
        H = ct.query_count(rt, d, q)
        f = ctypes.c_void_p()
        t = ctypes.c_void_p()
        ct._check(ct.lib.curtree_malloc(ctypes.byref(f), 4 * H))
        ct._check(ct.lib.curtree_malloc(ctypes.byref(t), 4 * H))
        assert ct.query_into(rt, d, q, f.value, t.value, H) == H
        st = ct.query_last_stats()
        fid = ct._host_from_dev(f.value, H, np.int32)
        tid = ct._host_from_dev(t.value, H, np.int32)
        ct.lib.curtree_free(f)
        ct.lib.curtree_free(t)
    finally:
        ct.set_option("leaf_fused", -1)
    print(f"[case] fused={fused}: {H} pairs, {st['pieces']} pieces", flush=True)
    assert st["pieces"] > 0
    ofid, otid, _ = orc.bfs(rt.to_host(), d.to_host(), ws)
    assert np.array_equal(fid, ofid) and np.array_equal(tid, otid)
    rt.free()
    d.free()
    q.free()


@pytest.mark.parametrize("fanout", [65, 100, 257])
def test_fanout_above_64(fanout):
    # the reference takes any fanout (src/rtree_test_frontier.cu:57 reads it
    # from argv): trees of fanout above 64 are queried one wave per entry,
    # children 64 at a time -- raw pairs equal the oracle's traversal, the
    # caller-buffer and count entry points agree, and so does the batched one
    import curtree_amd as ct

    n, nq = 300_000, 4_000
    bx = datagen.mixture_boxes(n, seed=440 + fanout)
    ws = datagen.loguniform_windows(nq, seed=450 + fanout, lo=1e-4, hi=0.05)
    ids = np.random.default_rng(fanout).permutation(n).astype(np.int32)
    fid, tid, (X, T, H) = run_case(bx, ids, ws, fanout, 1, brute_sample=300)
    print(f"[case] F={fanout}: {H} pairs", flush=True)
    d, rt, q = setup(bx, ids, ws, fanout, 1)
    total, counts = ct.query_count(rt, d, q, per_query=True)
    assert total == H and np.array_equal(counts, np.bincount(tid, minlength=nq))
    f = ctypes.c_void_p()
    t = ctypes.c_void_p()
    ct._check(ct.lib.curtree_malloc(ctypes.byref(f), 4 * H))
    ct._check(ct.lib.curtree_malloc(ctypes.byref(t), 4 * H))
    assert ct.query_into(rt, d, q, f.value, t.value, H) == H
    assert np.array_equal(ct._host_from_dev(f.value, H, np.int32), fid)
    assert np.array_equal(ct._host_from_dev(t.value, H, np.int32), tid)
    ct.lib.curtree_free(f)
    ct.lib.curtree_free(t)
    got = []
    st = ct.query_batched(rt, d, q, max(H // 5, 1),
                          lambda fp, tp, m, q0, q1: got.append(
                              (ct._host_from_dev(fp, m, np.int32),
                               ct._host_from_dev(tp, m, np.int32))))
    assert st["pairs"] == H and st["chunks"] > 1
    assert np.array_equal(np.concatenate([g[0] for g in got]), fid)
    assert np.array_equal(np.concatenate([g[1] for g in got]), tid)
    rt.free()
    d.free()
    q.free()


def test_kept_tree_state_skips_flag_pass_and_forgets_changed_boxes():
    # A tree built by the library records whether all its boxes are
    # well-formed, so a large batch on it runs no per-node flag pass.  Boxes
    # rewritten through the library afterwards drop that record: the query
    # then computes the flags again.  The rewrite inverts one box of a
    # leaf-parent that a window contains, so a stale record would report it
    # through the containment shortcut although it does not intersect.
    import curtree_amd as ct

    n = 50_000
    bx = datagen.uniform_boxes(n, seed=151, wmax=2e-3)
    ids = np.arange(n, dtype=np.int32)
    d = ct.DeviceBoxes.from_host(*bx, ids)
    rt = ct.build_rtree(d, 16, 1)
    h = rt.to_host()
    sb = list(d.to_host())
    l0 = rt.sz - (n + 15) // 16
    g = 7  # leaf-parent whose MBR the first window contains
    gx0, gy0, gx1, gy1 = (float(h[k][l0 + g]) for k in ("xmin", "ymin", "xmax", "ymax"))
    ws = datagen.loguniform_windows(2000, seed=152, lo=1e-3, hi=0.2)
    ws = [a.copy() for a in ws[:4]]
    ws[0][0], ws[1][0], ws[2][0], ws[3][0] = gx0 - 1e-4, gy0 - 1e-4, gx1 + 1e-4, gy1 + 1e-4
    q = ct.DeviceBoxes.from_host(*ws)
    ct.set_option("containment", 1)
    ct.Prof.enable(True)
    try:
        ct.Prof.reset()
        fid, tid = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_safe")[1] == 0  # the build's record was used
        ofid, otid, _ = orc.bfs(h, tuple(sb), ws)
        assert np.array_equal(fid, ofid) and np.array_equal(tid, otid)
        # invert box g*16+3 so that it misses the first window: xmax left of it
        b = g * 16 + 3
        sb[0][b], sb[2][b] = np.float32(gx1), np.float32(gx0 - 0.01)
        for k, a in ((0, "xmin"), (2, "xmax")):
            ct._check(ct.lib.curtree_memcpy_h2d(ctypes.c_void_p(getattr(d.s, a)),
                                                sb[k].ctypes.data, sb[k].nbytes))
        ct.Prof.reset()
        fid2, tid2 = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_safe")[1] == 1  # flags computed again
        ofid2, otid2, _ = orc.bfs(h, tuple(sb), ws)
        assert np.array_equal(fid2, ofid2) and np.array_equal(tid2, otid2)
        assert not np.any((tid2 == 0) & (fid2 == sb[4][b]))
        # a rebuild over the changed boxes records them as not all well-formed
        rt.free()
        rt = ct.build_rtree(d, 16, 1)
        ct.Prof.reset()
        fid3, tid3 = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_safe")[1] == 0
        sb3 = d.to_host()
        ofid3, otid3, _ = orc.bfs(rt.to_host(), sb3, ws)
        assert np.array_equal(fid3, ofid3) and np.array_equal(tid3, otid3)
    finally:
        ct.Prof.enable(False)
        ct.set_option("containment", -1)
    rt.free()
    d.free()
    q.free()


@pytest.mark.parametrize("contain", [0, 1])
def test_window_order_same_pairs(contain):
    # a batch whose windows are processed in Morton order of their centres
    # must give the same raw pair list (input order of the windows, then
    # sorted box position) as input order and as the oracle's traversal:
    # through the allocating query (count, then write), a caller buffer (one
    # writing run) and the per-window counts; NaN and point windows included
    import curtree_amd as ct

    n, nq = 200_000, 70_000
    bx = datagen.mixture_boxes(n, seed=161)
    ids = np.arange(n, dtype=np.int32)
    qs = [a.copy() for a in datagen.loguniform_windows(nq, seed=162, lo=1e-4, hi=1e-2)[:4]]
    qs[0][:7] = np.nan
    qs[2][7:20] = qs[0][7:20]
    qs[3][7:20] = qs[1][7:20]
    d, rt, q = setup(bx, ids, qs, 16, 1)
    ct.set_option("containment", contain)
    try:
        ct.set_option("window_order", 0)
        f0, t0 = ct.bfs_frontier(rt, d, q)
        ct.set_option("window_order", 1)
        f1, t1 = ct.bfs_frontier(rt, d, q)
        H = f1.size
        p = ctypes.c_void_p()
        ct._check(ct.lib.curtree_malloc(ctypes.byref(p), 8 * (H + 1)))
        got = ct.query_into(rt, d, q, p.value, p.value + 4 * (H + 1), H + 1)
        f2 = ct._host_from_dev(p.value, got, np.int32)
        t2 = ct._host_from_dev(p.value + 4 * (H + 1), got, np.int32)
        ct.lib.curtree_free(p)
        n3, c3 = ct.query_count(rt, d, q, per_query=True)
    finally:
        ct.set_option("window_order", -1)
        ct.set_option("containment", -1)
    assert np.array_equal(f0, f1) and np.array_equal(t0, t1)
    assert got == H and np.array_equal(f2, f0) and np.array_equal(t2, t0)
    assert n3 == H and np.array_equal(c3, np.bincount(t0, minlength=nq))
    sb = d.to_host()
    of, ot, _ = orc.bfs(rt.to_host(), sb, qs)
    assert np.array_equal(f0, of) and np.array_equal(t0, ot)
    rt.free()
    d.free()
    q.free()


@pytest.mark.parametrize("contain", [0, 1])
def test_leaf_codes_same_pairs(contain):
    # Leaf codes (8-bit cells of each box in its leaf-parent's frame) decide
    # most leaf tests without reading the box; a test the cells leave open
    # reads the box.  The raw pair list and the per-window counts must equal
    # those of the direct tests and the oracle's traversal, with windows whose
    # edges fall exactly on box edges (open cells), NaN, inverted, point and
    # infinite boxes, NaN and point windows, and leaf-parents of zero or
    # infinite extent (no frame: every test reads the box).
    import curtree_amd as ct

    n, nq = 200_000, 70_000
    bx = [a.copy() for a in datagen.mixture_boxes(n, seed=171)]
    r = np.random.default_rng(171)
    bx[1][r.integers(0, n, 40)] = np.nan
    bx[2][r.integers(0, n, 40)] = np.nan
    inv = r.integers(0, n, 100)
    bx[0][inv], bx[2][inv] = bx[2][inv] + np.float32(1e-4), bx[0][inv]
    pts = r.integers(0, n, 2000)
    bx[2][pts], bx[3][pts] = bx[0][pts], bx[1][pts]
    bx[2][r.integers(0, n, 5)] = np.inf
    same = np.arange(1000, 1200)  # 200 copies of one box: leaf-parents of zero extent
    for a in bx:
        a[same] = a[1000]
    ids = np.arange(n, dtype=np.int32)
    qs = [a.copy() for a in datagen.loguniform_windows(nq, seed=172, lo=1e-4, hi=1e-2)[:4]]
    # windows whose edges are box edges (shifted by whole boxes)
    k = r.integers(0, n, 20_000)
    qs[0][:20_000], qs[1][:20_000] = bx[2][k], bx[3][k]
    qs[2][:20_000], qs[3][:20_000] = bx[2][k] + np.float32(3e-3), bx[3][k] + np.float32(3e-3)
    qs[0][20_000:20_500] = np.nan
    qs[2][20_500:21_000], qs[3][20_500:21_000] = qs[0][20_500:21_000], qs[1][20_500:21_000]
    d, rt, q = setup(tuple(bx), ids, qs, 16, 1)
    ct.set_option("containment", contain)
    ct.Prof.enable(True)
    try:
        ct.set_option("leaf_codes", 0)
        f0, t0 = ct.bfs_frontier(rt, d, q)
        ct.set_option("leaf_codes", 1)
        ct.Prof.reset()
        f1, t1 = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_leaf_codes")[1] == 1  # built once
        f2, t2 = ct.bfs_frontier(rt, d, q)
        n3, c3 = ct.query_count(rt, d, q, per_query=True)
        assert ct.Prof.get("k_leaf_codes")[1] == 1  # then kept with the tree
    finally:
        ct.Prof.enable(False)
        ct.set_option("leaf_codes", -1)
        ct.set_option("containment", -1)
    assert np.array_equal(f0, f1) and np.array_equal(t0, t1)
    assert np.array_equal(f0, f2) and np.array_equal(t0, t2)
    assert n3 == f0.size and np.array_equal(c3, np.bincount(t0, minlength=nq))
    of, ot, _ = orc.bfs(rt.to_host(), d.to_host(), qs)
    assert np.array_equal(f0, of) and np.array_equal(t0, ot)
    rt.free()
    d.free()
    q.free()


@pytest.mark.parametrize("fanout", [3, 4, 8, 10, 16, 32, 33, 64])
def test_leaf_codes_fanouts(fanout):
    # every lane shape of the coded kernels (G/4 lanes per entry for G = 4 ..
    # 64, fanouts below G with padded slots) against the direct tests and the
    # oracle, with the containment shortcut on; codes forced for any query
    import curtree_amd as ct

    n, nq = 60_000, 3_000
    bx = [a.copy() for a in datagen.clustered_boxes(n, seed=181, frac=0.4, sigma=0.01, wmax=2e-3)]
    bx[0][:50] = np.nan
    ids = np.arange(n, dtype=np.int32)
    qs = [a.copy() for a in datagen.loguniform_windows(nq, seed=182, lo=1e-3, hi=0.2)[:4]]
    k = np.random.default_rng(183).integers(0, n, 500)
    qs[0][:500], qs[1][:500], qs[2][:500], qs[3][:500] = bx[2][k], bx[3][k], bx[2][k] + 0.01, bx[3][k] + 0.01
    qs[1][500:520] = np.nan
    d, rt, q = setup(tuple(bx), ids, qs, fanout, 1)
    ct.set_option("containment", 1)
    ct.set_option("expand_split", 1)  # the count / scan / write levels (the coded kernels)
    ct.Prof.enable(True)
    try:
        ct.set_option("leaf_codes", 0)
        f0, t0 = ct.bfs_frontier(rt, d, q)
        ct.set_option("leaf_codes", 1)
        ct.Prof.reset()
        f1, t1 = ct.bfs_frontier(rt, d, q)
        built = ct.Prof.get("k_leaf_codes")[1]
        n1, c1 = ct.query_count(rt, d, q, per_query=True)
    finally:
        ct.Prof.enable(False)
        ct.set_option("leaf_codes", -1)
        ct.set_option("expand_split", -1)
        ct.set_option("containment", -1)
    assert built == (1 if fanout >= 3 else 0)
    assert np.array_equal(f0, f1) and np.array_equal(t0, t1)
    assert n1 == f0.size and np.array_equal(c1, np.bincount(t0, minlength=nq))
    of, ot, _ = orc.bfs(rt.to_host(), d.to_host(), qs)
    assert np.array_equal(f0, of) and np.array_equal(t0, ot)
    rt.free()
    d.free()
    q.free()


def test_leaf_codes_forgotten_when_boxes_change():
    # The leaf codes are kept with a tree built here; boxes rewritten through
    # the library afterwards forget them with the rest of the tree's record (a
    # stale code would decide a test from the old box).  One box moves after
    # the codes are built: the next query (no record, so no codes) must see it
    # where it now is, and a rebuild keeps the tree again with new codes.
    import curtree_amd as ct

    n, nq = 40_000, 2_000
    bx = datagen.uniform_boxes(n, seed=191, wmax=2e-3)
    ids = np.arange(n, dtype=np.int32)
    d = ct.DeviceBoxes.from_host(*bx, ids)
    rt = ct.build_rtree(d, 16, 1)
    sb = list(d.to_host())
    h = rt.to_host()
    l0 = rt.sz - (n + 15) // 16
    b = 123  # the box (sorted position) that moves
    gx0, gy0 = float(h["xmin"][l0 + b // 16]), float(h["ymin"][l0 + b // 16])
    ex = float(h["xmax"][l0 + b // 16]) - gx0
    ws = [a.copy() for a in datagen.loguniform_windows(nq, seed=192, lo=1e-3, hi=0.1)[:4]]
    # window 0 a little larger than the box (its cells prove a hit with the
    # old box), window 1 a speck at the leaf-parent's low corner, where the
    # box goes (its cells prove a miss with the old box)
    m = ex / 50
    ws[0][0], ws[1][0], ws[2][0], ws[3][0] = (sb[0][b] - m, sb[1][b] - m, sb[2][b] + m,
                                              sb[3][b] + m)
    ws[0][1], ws[1][1], ws[2][1], ws[3][1] = gx0 - 1e-7, gy0 - 1e-7, gx0 + 1e-7, gy0 + 1e-7
    q = ct.DeviceBoxes.from_host(*ws)
    ct.set_option("leaf_codes", 1)
    ct.Prof.enable(True)
    try:
        ct.Prof.reset()
        f0, t0 = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_leaf_codes")[1] == 1
        of, ot, _ = orc.bfs(h, tuple(sb), ws)
        assert np.array_equal(f0, of) and np.array_equal(t0, ot)
        assert np.any((t0 == 0) & (f0 == sb[4][b])) and not np.any((t0 == 1) & (f0 == sb[4][b]))
        # shrink the box to a point at its leaf-parent MBR's low corner: still
        # inside the MBR, so the node MBRs stay valid, but its cells change
        for k, a, v in ((0, "xmin", gx0), (1, "ymin", gy0), (2, "xmax", gx0), (3, "ymax", gy0)):
            sb[k][b] = np.float32(v)
            ct._check(ct.lib.curtree_memcpy_h2d(ctypes.c_void_p(getattr(d.s, a)),
                                                sb[k].ctypes.data, sb[k].nbytes))
        ct.Prof.reset()
        f1, t1 = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_leaf_codes")[1] == 0  # the record, codes included, is gone
        of1, ot1, _ = orc.bfs(h, tuple(sb), ws)
        assert np.array_equal(f1, of1) and np.array_equal(t1, ot1)
        assert not np.any((t1 == 0) & (f1 == sb[4][b])) and np.any((t1 == 1) & (f1 == sb[4][b]))
        # a rebuild keeps the tree again, with codes from its first query
        rt.free()
        rt = ct.build_rtree(d, 16, 1)
        ct.Prof.reset()
        f2, t2 = ct.bfs_frontier(rt, d, q)
        assert ct.Prof.get("k_leaf_codes")[1] == 1
        of2, ot2, _ = orc.bfs(rt.to_host(), d.to_host(), ws)
        assert np.array_equal(f2, of2) and np.array_equal(t2, ot2)
    finally:
        ct.Prof.enable(False)
        ct.set_option("leaf_codes", -1)
    rt.free()
    d.free()
    q.free()
