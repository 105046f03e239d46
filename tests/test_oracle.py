"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement (dropped
term, swapped axis, wrong sign or index) fails one of them:
  - the reference's own dilation table (tests/golden/dilate_table.json, parsed
    from src/z_order.cuh:20-42) over all 65,536 inputs;
  - closed forms (textbook bit-by-bit Morton on integer data, exact grids);
  - library routines (numpy stable argsort, numpy reduceat min/max);
  - brute force (numpy broadcasting) for the query results.
"""
import json
import os

import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def morton_bits(x, y, bits):
    """Textbook interleave: bit i of x -> bit 2i, bit i of y -> bit 2i+1."""
    z = 0
    for i in range(bits):
        z |= ((x >> i) & 1) << (2 * i)
        z |= ((y >> i) & 1) << (2 * i + 1)
    return z


def test_dilation_matches_reference_table_exhaustively():
    with open(os.path.join(GOLDEN, "dilate_table.json")) as f:
        tab = json.load(f)["dilate_table"]
    assert len(tab) == 256
    L = orc.lib()
    for x in range(65536):
        # dilate_2: src/z_order.cuh:63-65
        want = tab[x & 0xFF] | (tab[x >> 8] << 16)
        assert L.or_dilate16(x) == want, x


def test_undilate_table_inverts_dilation():
    with open(os.path.join(GOLDEN, "dilate_table.json")) as f:
        g = json.load(f)
    dil, und = g["dilate_table"], g["undilate_table"]
    # undilate_2 (src/z_order.cuh:72-75) applied to a dilated byte returns it
    for b in range(256):
        d = dil[b]
        assert und[0xFF & ((d >> 7) | d)] == b


def test_zorder_textbook_on_integer_coordinates():
    L = orc.lib()
    r = np.random.default_rng(0)
    for x, y in r.integers(0, 65536, (2000, 2)):
        assert L.or_zorder16(int(x), int(y)) == morton_bits(int(x), int(y), 16)
    assert L.or_zorder16(1, 0) == 1 and L.or_zorder16(0, 1) == 2


def test_ref32_key_integer_boxes_is_textbook_morton_signed():
    r = np.random.default_rng(1)
    x = r.integers(0, 65536, 5000).astype(np.float32)
    y = r.integers(0, 65536, 5000).astype(np.float32)
    # box = [x-1, x+1]: centroid exactly x (exact in FP32 for these magnitudes)
    k = orc.keys_ref32(x - 1, y - 1, x + 1, y + 1)
    want = np.array([morton_bits(int(a), int(b), 16) for a, b in zip(x, y)], np.uint32)
    assert np.array_equal(k.view(np.uint32), want)
    # stored as int32: keys with bit 31 (y >= 32768) are negative
    assert np.array_equal(k < 0, y >= 32768)


def test_ref32_key_special_cases():
    f = np.float32
    nan, inf = f(np.nan), f(np.inf)
    xmin = np.array([0.1, 0.5, 0.999, 1.0, -5, nan, 70000, 65535.9, 2.9, -0.4], f)
    xmax = xmin.copy()
    y = np.zeros_like(xmin)
    k = orc.keys_ref32(xmin, y, xmax, y)
    # [0,1) centroids truncate to 0; 1.0 -> x=1 -> key 1; negative/NaN -> 0;
    # >= 65535 saturates; truncation toward zero
    assert list(k) == [0, 0, 0, 1, 0, 0, morton_bits(65535, 0, 16), morton_bits(65535, 0, 16),
                       morton_bits(2, 0, 16), 0]
    # the configured [0,1)^2 data gives key 0 for every box
    bx = datagen.uniform_boxes(20000, seed=11)
    assert not np.any(orc.keys_ref32(*bx))
    _ = inf


def test_fmin_fmax_definitions():
    L = orc.lib()
    nan = float("nan")
    assert L.or_fmin_one(1.0, 2.0) == 1.0 and L.or_fmax_one(1.0, 2.0) == 2.0
    assert L.or_fmin_one(nan, 3.0) == 3.0 and L.or_fmin_one(3.0, nan) == 3.0
    assert L.or_fmax_one(nan, 3.0) == 3.0 and L.or_fmax_one(3.0, nan) == 3.0
    assert np.isnan(L.or_fmin_one(nan, nan))
    # -0 orders below +0, whichever operand comes first
    for a, b in ((-0.0, 0.0), (0.0, -0.0)):
        assert np.signbit(L.or_fmin_one(a, b))
        assert not np.signbit(L.or_fmax_one(a, b))


def test_intersect_closed_intervals():
    L = orc.lib()
    # touching edges and corners count
    assert L.or_intersect_one(0, 0, 1, 1, 1, 1, 2, 2) == 1
    assert L.or_intersect_one(0, 0, 1, 1, 1.0001, 0, 2, 1) == 0
    assert L.or_intersect_one(0, 0, 1, 1, 0.5, 0.5, 0.5, 0.5) == 1
    # NaN makes every comparison false -> "intersects" (src/utility.h:16-20)
    assert L.or_intersect_one(0, 0, 1, 1, float("nan"), 0, 6, 1) == 1
    assert L.or_intersect_one(0, 0, 1, 1, float("nan"), 5, 6, 7) == 0
    assert L.or_intersect_one(float("nan"), float("nan"), float("nan"), float("nan"),
                              5, 5, 6, 6) == 1


def test_morton64_exact_grid_closed_form():
    # centroids on i/1024 over the extent [0, 1]: scale = 2^32 exactly,
    # so q = i * 2^22 and the key is the textbook 32-bit interleave
    r = np.random.default_rng(2)
    i = r.integers(0, 1025, 3000)
    j = r.integers(0, 1025, 3000)
    i[0], j[0], i[1], j[1] = 0, 0, 1024, 1024
    cx = (i / 1024.0).astype(np.float32)
    cy = (j / 1024.0).astype(np.float32)
    k = orc.keys_morton64(cx, cy, cx, cy)
    ext = orc.centroid_extent(cx, cy, cx, cy)
    assert list(ext) == [0.0, 0.0, 1.0, 1.0]
    qi = np.minimum(i.astype(np.uint64) << np.uint64(22), np.uint64(0xFFFFFFFF))
    qj = np.minimum(j.astype(np.uint64) << np.uint64(22), np.uint64(0xFFFFFFFF))
    want = np.array([morton_bits(int(a), int(b), 32) for a, b in zip(qi, qj)], np.uint64)
    assert np.array_equal(k, want)


def test_centroid_extent_matches_numpy():
    bx = datagen.mixture_boxes(50000, seed=12)
    cx = (bx[0] + bx[2]) * np.float32(0.5)
    cy = (bx[1] + bx[3]) * np.float32(0.5)
    ext = orc.centroid_extent(*bx)
    assert list(ext) == [cx.min(), cy.min(), cx.max(), cy.max()]
    # NaN centroids are skipped
    b2 = [a.copy() for a in bx]
    b2[0][7] = np.nan
    ext2 = orc.centroid_extent(*b2)
    m = np.ones(cx.shape, bool)
    m[7] = False
    assert ext2[0] == cx[m].min() and ext2[2] == cx[m].max()


def test_morton64_scale_and_quantisation_monotone():
    bx = datagen.uniform_boxes(20000, seed=13)
    k = orc.keys_morton64(*bx)
    cx = (bx[0] + bx[2]) * np.float32(0.5)
    # x occupies the even bits: extracting them must be monotone in cx
    xe = np.zeros(k.shape, np.uint64)
    for b in range(32):
        xe |= ((k >> np.uint64(2 * b)) & np.uint64(1)) << np.uint64(b)
    o = np.argsort(cx, kind="stable")
    assert np.all(np.diff(xe[o].astype(np.int64)) >= 0)
    assert xe.min() == 0 and xe.max() >= 0xFFFFFF00


def test_sort_perm_is_numpy_stable_argsort():
    r = np.random.default_rng(3)
    k32 = r.integers(-50, 50, 100000).astype(np.int32)  # many ties, both signs
    assert np.array_equal(orc.sort_perm_i32(k32), np.argsort(k32, kind="stable"))
    k64 = r.integers(0, 2**63, 50000, dtype=np.uint64) >> r.integers(0, 63, 50000).astype(np.uint64)
    k64[::7] = k64[3]
    assert np.array_equal(orc.sort_perm_u64(k64), np.argsort(k64, kind="stable"))
    # all-equal keys: identity permutation
    assert np.array_equal(orc.sort_perm_i32(np.zeros(1000, np.int32)), np.arange(1000))


LEVELS = {
    (1_000_000, 16): ([62500, 3907, 245, 16, 1], 66669),
}


def test_level_plan_closed_form_and_workload_sizes():
    assert orc.level_plan(1_000_000, 16) == LEVELS[(1_000_000, 16)]
    assert orc.level_plan(100_000_000, 16)[1] == 6_666_669
    assert orc.level_plan(100_000_000, 32)[1] == 3_225_809
    assert orc.level_plan(1_000_000_000, 16)[1] == 66_666_669
    for n in list(range(2, 300)) + [4095, 4096, 4097, 65536, 1 << 20, 5 * 10**7, 10**8]:
        for f in (2, 3, 8, 10, 16, 32, 64):
            lv, rs = orc.level_plan(n, f)
            # closed form: repeated integer ceil-division down to one node
            want, m = [], n
            while True:
                m = -(-m // f)
                want.append(m)
                if m == 1:
                    break
            assert lv == want and rs == sum(want)
            if n <= 1 << 24:
                assert orc.level_plan(n, f, use_float=True) == (lv, rs)
    assert orc.level_plan(1, 16) == ([1], 1)
    assert orc.level_plan(0, 16) == ([], 0)


def test_level_plan_float_hazard_documented():
    # the reference's float ceil undercounts here (src/rtree_bulkload_thrust.cuh:109)
    assert orc.level_plan(16_777_217, 16, use_float=True)[0][0] == 1_048_576
    assert orc.level_plan(16_777_217, 16)[0][0] == 1_048_577


@pytest.mark.parametrize("fanout", [2, 3, 10, 16, 64])
def test_build_levels_match_numpy_reduce(fanout):
    n = 12345
    bx = datagen.uniform_boxes(n, seed=14)
    t = orc.build_levels(bx, fanout)
    lv, rs = orc.level_plan(n, fanout)
    assert t["pos"].shape[0] == rs
    start = rs - lv[0]
    starts = np.arange(0, n, fanout)
    # level 0 against numpy reduceat over consecutive groups of F boxes
    assert np.array_equal(t["xmin"][start:], np.minimum.reduceat(bx[0], starts))
    assert np.array_equal(t["ymin"][start:], np.minimum.reduceat(bx[1], starts))
    assert np.array_equal(t["xmax"][start:], np.maximum.reduceat(bx[2], starts))
    assert np.array_equal(t["ymax"][start:], np.maximum.reduceat(bx[3], starts))
    assert np.array_equal(t["pos"][start:], rs + starts)
    assert np.array_equal(t["len"][start:], np.minimum(fanout, n - starts))
    # upper levels: pos/len tile the child level exactly and MBRs contain children
    child, cn = start, lv[0]
    for L in range(1, len(lv)):
        ps = child - lv[L]
        pos, ln = t["pos"][ps:child], t["len"][ps:child]
        assert pos[0] == child and np.all(pos[1:] == pos[:-1] + ln[:-1])
        assert pos[-1] + ln[-1] == child + cn
        cstarts = np.arange(0, cn, fanout)
        assert np.array_equal(t["xmin"][ps:child],
                              np.minimum.reduceat(t["xmin"][child:child + cn], cstarts))
        assert np.array_equal(t["ymax"][ps:child],
                              np.maximum.reduceat(t["ymax"][child:child + cn], cstarts))
        child, cn = ps, lv[L]
    assert child == 0 and cn == 1


def numpy_brute(boxes, windows):
    bx0, by0, bx1, by1, ids = boxes
    q0, r0, q1, r1 = windows[:4]
    fid, tid = [], []
    for i in range(q0.shape[0]):
        hit = ~((q0[i] > bx1) | (q1[i] < bx0) | (r0[i] > by1) | (r1[i] < by0))
        j = np.nonzero(hit)[0]
        fid.append(ids[j])
        tid.append(np.full(j.shape, i, np.int32))
    return np.concatenate(fid), np.concatenate(tid)


def test_brute_force_matches_numpy():
    bx = datagen.uniform_boxes(3000, seed=15, wmax=0.05)
    ids = np.arange(3000, dtype=np.int32)[::-1].copy()
    qs = datagen.uniform_windows(200, seed=16, smax=0.1)
    f1, t1 = orc.brute_force(bx + (ids,), qs)
    f2, t2 = numpy_brute(bx + (ids,), qs)
    assert np.array_equal(f1, f2) and np.array_equal(t1, t2)


@pytest.mark.parametrize("fanout,key_mode", [(2, 0), (3, 1), (16, 1), (16, 0), (64, 1)])
def test_bfs_equals_brute_force(fanout, key_mode):
    n = 20000
    bx = datagen.clustered_boxes(n, seed=17, frac=0.3, sigma=0.01, wmax=0.01)
    ids = np.random.default_rng(5).permutation(n).astype(np.int32)
    qs = datagen.loguniform_windows(300, seed=18, lo=1e-4, hi=0.2)
    sb, t = orc.build_rtree(*bx, ids, fanout, key_mode)
    fid, tid, (X, T, H) = orc.bfs(t, sb, qs)
    assert H == fid.shape[0]
    # order: by query, then by sorted box position
    pos_of = np.empty(n, np.int64)
    pos_of[sb[4]] = np.arange(n)
    key = tid.astype(np.int64) * n + pos_of[fid]
    assert np.all(np.diff(key) > 0)
    bf, bt = orc.brute_force(bx + (ids,), qs)
    assert np.array_equal(orc.canonical(fid, tid)[0], orc.canonical(bf, bt)[0])
    assert np.array_equal(orc.canonical(fid, tid)[1], orc.canonical(bf, bt)[1])
    assert T >= H and X > 0


def test_bfs_edge_cases():
    # one box: a single root level holding it
    bx = tuple(np.array([v], np.float32) for v in (0.1, 0.1, 0.2, 0.2))
    sb, t = orc.build_rtree(*bx, np.array([7], np.int32), 16)
    assert t["pos"].tolist() == [1] and t["len"].tolist() == [1]
    qs = tuple(np.array(v, np.float32) for v in ([0, 0.3], [0, 0.3], [0.1, 0.4], [0.1, 0.4]))
    fid, tid, _ = orc.bfs(t, sb, qs)
    assert fid.tolist() == [7] and tid.tolist() == [0]  # touching at the corner
    # zero queries
    fid, tid, c = orc.bfs(t, sb, tuple(np.zeros(0, np.float32) for _ in range(4)))
    assert fid.size == 0 and c == (0, 0, 0)


def test_skewed_large_generator():
    # chunked generation: deterministic, independent of nothing but (n, seed,
    # chunk), 70% of the boxes inside the 1,000 hot cells of side 1e-3,
    # about 20% points, every box inside [0, 1]^2 up to its half-size
    a = datagen.skewed_boxes_large(30_000, seed=3, chunk=7_000)
    b = datagen.skewed_boxes_large(30_000, seed=3, chunk=7_000)
    for u, v in zip(a, b):
        assert u.dtype == np.float32 and np.array_equal(u, v)
    xmin, ymin, xmax, ymax = (x.astype(np.float64) for x in a)
    assert np.all(xmax >= xmin) and np.all(ymax >= ymin)
    pts = np.mean((xmax == xmin) & (ymax == ymin))
    assert 0.17 < pts < 0.23
    cx, cy = (xmin + xmax) / 2, (ymin + ymax) / 2
    assert cx.min() > -1e-4 and cx.max() < 1 + 1e-4 and cy.min() > -1e-4 and cy.max() < 1 + 1e-4
    # hot-cell occupancy: the 1,000 most occupied cells of side 1e-3 (on a
    # 1e-3 grid a hot cell spans up to four) hold well over half the boxes
    g = np.floor(cx * 1000).astype(np.int64) * 1001 + np.floor(cy * 1000).astype(np.int64)
    _, cnt = np.unique(g, return_counts=True)
    assert np.sort(cnt)[::-1][:4000].sum() / cx.size > 0.6


def closed_form_work(t, windows, n_levels_nodes):
    """Frontier sizes of the BFS from its closed form: in a bulk-loaded tree
    every parent MBR contains its children's (min/max of the children), so a
    window that meets a node meets all its ancestors, and the depth-d frontier
    is exactly the set of (window, node at depth d) pairs that intersect.
    Returns (sizes per depth, X, T): X sums the depths >= 1, T the child
    counts of every frontier entry (the tests the traversal makes)."""
    q0, r0, q1, r1 = (np.asarray(a, np.float32) for a in windows[:4])
    nq = q0.shape[0]
    sizes = [nq]
    T = nq * int(t["len"][0])
    start = 0
    for d, cnt in enumerate(n_levels_nodes):  # root level first
        if d == 0:
            start += cnt
            continue
        sl = slice(start, start + cnt)
        x0, y0, x1, y1 = t["xmin"][sl], t["ymin"][sl], t["xmax"][sl], t["ymax"][sl]
        ln = t["len"][sl].astype(np.int64)
        hits = 0
        tests = 0
        for i in range(0, nq, 256):  # windows in blocks: (256 x nodes) masks
            a = slice(i, min(i + 256, nq))
            m = ~((q0[a, None] > x1[None]) | (q1[a, None] < x0[None]) |
                  (r0[a, None] > y1[None]) | (r1[a, None] < y0[None]))
            hits += int(m.sum())
            tests += int((m * ln[None]).sum())
        sizes.append(hits)
        T += tests
        start += cnt
    return sizes, sum(sizes[1:]), T


@pytest.mark.parametrize("n,fanout,nq", [(200_000, 16, 2000), (50_000, 5, 700), (30_000, 64, 500)])
def test_bfs_work_counters_closed_form(n, fanout, nq):
    # the workload shape of the plain CPU-runnable case, scaled down
    bx = datagen.uniform_boxes(n, seed=21, wmax=1e-3)
    qs = datagen.uniform_windows(nq, seed=22, smax=0.01)
    ids = np.arange(n, dtype=np.int32)
    sb, t = orc.build_rtree(*bx, ids, fanout, 1)
    fid, tid, (X, T, H), depths = orc.bfs(t, sb, qs, depths=True)
    lv, _ = orc.level_plan(n, fanout)
    sizes, Xc, Tc = closed_form_work(t, qs, lv[::-1])
    assert depths == sizes
    assert (X, T) == (Xc, Tc)
    assert H == orc.brute_count(bx, qs) == fid.shape[0]


def test_canonical_sort_c_matches_lexsort():
    r = np.random.default_rng(23)
    n = 50_000
    tid = np.sort(r.integers(-5, 300, n)).astype(np.int32)[r.permutation(n)]
    fid = r.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    fid[: n // 4] = fid[n // 4: n // 2]  # duplicate (query, box) pairs too
    tid[: n // 4] = tid[n // 4: n // 2]
    a = orc.canonical_c(fid, tid)
    o = np.lexsort((fid, tid))
    assert np.array_equal(a[0], fid[o]) and np.array_equal(a[1], tid[o])
