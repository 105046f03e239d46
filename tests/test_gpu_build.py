"""GPU parity of the bulk load (through the C ABI) against the oracle.

Bar: bit-exact.  The caller's box arrays after the in-place sort and all six
node arrays must equal the oracle's byte for byte, in both key modes, for
ragged sizes and every fanout the query supports, plus special values."""
import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

pytestmark = pytest.mark.gpu


def gpu_build(boxes, ids, fanout, key_mode):
    import curtree_amd as ct

    d = ct.DeviceBoxes.from_host(*boxes, ids)
    rt = ct.build_rtree(d, fanout, key_mode)
    sb = d.to_host()
    t = rt.to_host()
    sz = rt.sz
    rt.free()
    d.free()
    return sb, t, sz


def assert_bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, what
    bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}"


def check_against_oracle(boxes, ids, fanout, key_mode):
    sb, t, sz = gpu_build(boxes, ids, fanout, key_mode)
    osb, ot = orc.build_rtree(*boxes, ids, fanout, key_mode)
    assert sz == ot["pos"].shape[0]
    for i, name in enumerate(("xmin", "ymin", "xmax", "ymax", "id")):
        assert_bits_equal(sb[i], osb[i], f"sorted {name}")
    for name in ("xmin", "ymin", "xmax", "ymax", "pos", "len"):
        assert_bits_equal(t[name], ot[name], f"node {name}")
    return sb, t


def test_ref32_unit_square_is_input_order():
    # the configured [0,1)^2 data: every reference key is 0, so the tree is
    # built over the input order (stable sort of equal keys)
    n = 1_000_000
    bx = datagen.uniform_boxes(n, seed=1)
    ids = np.arange(n, dtype=np.int32)
    sb, _ = check_against_oracle(bx, ids, 16, 0)
    assert np.array_equal(sb[4], ids)


@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 4095, 4097, 100_003])
def test_ref32_integer_coordinates_signed_order(n):
    r = np.random.default_rng(n)
    x = r.integers(0, 65536, n).astype(np.float32)
    y = r.integers(0, 65536, n).astype(np.float32)
    x[: n // 3] = x[0]  # many ties
    bx = (x - 0.5, y - 0.5, x + 0.5, y + 0.5)
    ids = r.permutation(n).astype(np.int32)
    check_against_oracle(bx, ids, 16, 0)


@pytest.mark.parametrize("fanout", [2, 3, 4, 8, 10, 16, 32, 64])
def test_morton64_fanouts(fanout):
    n = 200_001
    bx = datagen.mixture_boxes(n, seed=fanout)
    ids = np.random.default_rng(fanout).permutation(n).astype(np.int32)
    check_against_oracle(bx, ids, fanout, 1)


@pytest.mark.parametrize("n", [2047, 2048, 2049, 10241, 12287, 12288, 12289, 24577])
def test_tile_boundaries(n):
    # sizes around the tiles of the final pass (2048 positions), the 64-bit
    # LSD passes (10240 keys) and the payload passes (12288 items), with runs
    # of equal high halves crossing tile edges
    r = np.random.default_rng(n)
    g = r.integers(0, 64, n).astype(np.float32) / np.float32(64)  # 64 x 64 cells
    h = r.integers(0, 64, n).astype(np.float32) / np.float32(64)
    bx = (g, h, g + np.float32(1e-3), h + np.float32(1e-3))
    ids = r.permutation(n).astype(np.int32)
    for f, km in ((16, 1), (10, 1), (8, 0)):
        check_against_oracle(bx, ids, f, km)


def test_morton64_one_million_uniform():
    n = 1_000_000
    bx = datagen.uniform_boxes(n, seed=21)
    check_against_oracle(bx, np.arange(n, dtype=np.int32), 16, 1)


def test_special_values():
    n = 50_000
    bx = [a.copy() for a in datagen.uniform_boxes(n, seed=22, wmax=1e-2)]
    r = np.random.default_rng(22)
    for k, v in ((0, np.nan), (2, np.nan), (1, -0.0), (3, 0.0), (0, -0.0), (2, np.inf)):
        idx = r.integers(0, n, 300)
        bx[k][idx] = v
    # degenerate points and identical boxes
    bx[2][:1000] = bx[0][:1000]
    bx[3][:1000] = bx[1][:1000]
    for km in (0, 1):
        for f in (16, 10):
            check_against_oracle(tuple(bx), np.arange(n, dtype=np.int32), f, km)


def test_all_equal_keys_and_tiny():
    for n in (1, 2, 3, 64, 65, 4096, 4097):
        bx = tuple(np.full(n, v, np.float32) for v in (0.25, 0.25, 0.5, 0.5))
        for km in (0, 1):
            for f in (2, 16, 64):
                sb, _ = check_against_oracle(bx, np.arange(n, dtype=np.int32)[::-1].copy(), f, km)


def test_empty_input():
    import curtree_amd as ct

    d = ct.DeviceBoxes(0)
    rt = ct.build_rtree(d, 16, 1)
    assert rt.sz == 0


def test_large_ragged_properties():
    # 2^24 + 1 boxes: beyond the reference's float level-count range
    n = (1 << 24) + 1
    bx = datagen.mixture_boxes(n, seed=23)
    ids = np.arange(n, dtype=np.int32)
    sb, t, sz = gpu_build(bx, ids, 16, 1)
    lv, rs = orc.level_plan(n, 16)
    assert sz == rs and lv[0] == 1_048_577
    # the ids are a permutation and the keys of the sorted boxes are sorted
    assert np.array_equal(np.sort(sb[4]), ids)
    k = orc.keys_morton64(*sb[:4], ext=orc.centroid_extent(*bx))
    assert np.all(k[1:] >= k[:-1])
    # leaf-parent MBRs are the min/max of their boxes
    start = rs - lv[0]
    g = np.arange(0, n, 16)
    assert np.array_equal(t["xmin"][start:], np.minimum.reduceat(sb[0], g))
    assert np.array_equal(t["ymax"][start:], np.maximum.reduceat(sb[3], g))
    # the permuted payload travels with its box
    o = np.argsort(sb[4])
    for i in range(4):
        assert np.array_equal(sb[i][o], bx[i])


def _clustered_points(sizes, seed, background):
    """Point boxes: clusters whose centroids share one 1/65536 cell per axis
    (so their 64-bit Morton keys agree in the high 32 bits and differ below),
    plus uniform background points and two corners fixing the extent."""
    r = np.random.default_rng(seed)
    xs, ys = [r.random(background)], [r.random(background)]
    for s in sizes:
        cx, cy = r.integers(1000, 64000, 2)
        xs.append((cx + 0.5 + r.uniform(-0.45, 0.45, s)) / 65536.0)
        ys.append((cy + 0.5 + r.uniform(-0.45, 0.45, s)) / 65536.0)
    xs.append(np.array([0.0, 1.0]))
    ys.append(np.array([0.0, 1.0]))
    x = np.concatenate(xs).astype(np.float32)
    y = np.concatenate(ys).astype(np.float32)
    p = r.permutation(x.size)
    x, y = x[p], y[p]
    return (x, y, x.copy(), y.copy())


@pytest.mark.parametrize("case", ["short_runs", "wave_runs", "long_runs", "huge_runs",
                                  "limit_run", "too_long_runs"])
def test_morton64_equal_high_half_runs(case):
    # the 64-bit sort orders by the high half, then sorts runs of equal high
    # halves by (low half, input position) in place: runs inside one chunk of
    # 64 positions by its wave, other runs of up to 64 by one wave each, up to
    # 4096 by a block network, up to max(65536, n / 64) by a block LSD sort; a
    # longer run takes the full eight-pass path (the other runs may or may not
    # have been settled before).  Every case must equal the oracle's stable sort.
    r = np.random.default_rng(7)
    sizes = {"short_runs": r.integers(2, 33, 3000),
             "wave_runs": r.integers(48, 81, 400),
             "long_runs": r.integers(33, 1026, 40),
             "huge_runs": list(r.integers(4097, 20000, 12)) + [4097, 8192, 8193, 5000],
             "limit_run": [65536, 4500, 70],
             "too_long_runs": [65537, 9000, 5000, 40, 3]}[case]
    bx = _clustered_points(sizes, seed=len(case), background=100_000)
    n = bx[0].size
    for f in (16, 10):
        check_against_oracle(bx, np.arange(n, dtype=np.int32), f, 1)


@pytest.mark.parametrize("dups", [9000, 20000])
def test_morton64_duplicate_heavy_run(dups):
    # a run of equal high halves longer than 8192 whose low halves are mostly
    # one value: after the split by the top digit of the low half one digit
    # still holds more than 8192 boxes (the run then takes four LSD passes)
    x, y, _, _ = _clustered_points([3000], seed=dups, background=50_000)
    # the cluster's points are the ones sharing the most common cell
    cell = (np.floor(x * 65536).astype(np.int64) << 20) + np.floor(y * 65536).astype(np.int64)
    vals, cnt = np.unique(cell, return_counts=True)
    j = np.nonzero(cell == vals[np.argmax(cnt)])[0][0]
    x = np.concatenate([x, np.full(dups, x[j], np.float32)])
    y = np.concatenate([y, np.full(dups, y[j], np.float32)])
    p = np.random.default_rng(dups).permutation(x.size)
    bx = (x[p], y[p], x[p].copy(), y[p].copy())
    n = x.size
    for f in (16, 10):
        check_against_oracle(bx, np.arange(n, dtype=np.int32), f, 1)


@pytest.mark.parametrize("fanout,key_mode", [(16, 1), (32, 1), (8, 0)])
def test_skewed_hot_cells(fanout, key_mode):
    # the large/skewed workload's shape at a parity size: 70% of the boxes in
    # 1,000 Zipf-weighted hot cells of side 1e-3, 20% points; in the hottest
    # cells many boxes share the high half of their 64-bit key (runs settled
    # in the final pass, by the wave fixups and by the block fixup)
    n = 400_000
    bx = datagen.skewed_boxes(n, seed=190)
    ids = np.random.default_rng(190).permutation(n).astype(np.int32)
    check_against_oracle(bx, ids, fanout, key_mode)


def test_skewed_large_properties():
    # 2^23 boxes of the skewed shape: thousands of runs of hundreds of equal
    # high halves; the sorted keys must be sorted, ties in input order, and
    # every leaf-parent the min/max of its boxes
    n = 1 << 23
    bx = datagen.skewed_boxes(n, seed=191)
    ids = np.arange(n, dtype=np.int32)
    sb, t, sz = gpu_build(bx, ids, 16, 1)
    lv, rs = orc.level_plan(n, 16)
    assert sz == rs
    assert np.array_equal(np.sort(sb[4]), ids)
    k = orc.keys_morton64(*sb[:4], ext=orc.centroid_extent(*bx))
    assert np.all(k[1:] >= k[:-1])
    # equal full keys keep their input order (ids are input positions)
    same = k[1:] == k[:-1]
    assert np.all(sb[4][1:][same] > sb[4][:-1][same])
    start = rs - lv[0]
    g = np.arange(0, n, 16)
    assert np.array_equal(t["xmin"][start:], np.minimum.reduceat(sb[0], g))
    assert np.array_equal(t["ymin"][start:], np.minimum.reduceat(sb[1], g))
    assert np.array_equal(t["xmax"][start:], np.maximum.reduceat(sb[2], g))
    assert np.array_equal(t["ymax"][start:], np.maximum.reduceat(sb[3], g))
    o = np.argsort(sb[4])
    for i in range(4):
        assert np.array_equal(sb[i][o], bx[i])


def test_over_a_million_listed_runs():
    # 1.1M runs of 65 equal high halves (each crosses a chunk of 64, so every
    # run is listed and sorted by one wave): the run-length pass takes more
    # than one round over the list (1,048,576 runs per round); keys sorted,
    # ties in input order, leaf-parents the min/max of their boxes
    r = np.random.default_rng(77)
    m, L = 1_100_000, 65
    cells = np.unique(r.integers(0, 65536, (m + m // 8, 2)), axis=0)[:m]
    r.shuffle(cells)
    cx = np.repeat(cells[:, 0], L) + 0.5 + r.uniform(-0.45, 0.45, m * L)
    cy = np.repeat(cells[:, 1], L) + 0.5 + r.uniform(-0.45, 0.45, m * L)
    x = np.concatenate([cx / 65536.0, [0.0, 1.0]]).astype(np.float32)
    y = np.concatenate([cy / 65536.0, [0.0, 1.0]]).astype(np.float32)
    p = r.permutation(x.size)
    bx = (x[p], y[p], x[p].copy(), y[p].copy())
    n = x.size
    ids = np.arange(n, dtype=np.int32)
    sb, t, sz = gpu_build(bx, ids, 16, 1)
    lv, rs = orc.level_plan(n, 16)
    assert sz == rs
    k = orc.keys_morton64(*sb[:4], ext=orc.centroid_extent(*bx))
    assert np.all(k[1:] >= k[:-1])
    same = k[1:] == k[:-1]
    assert np.all(sb[4][1:][same] > sb[4][:-1][same])
    hi = k >> np.uint64(32)
    assert np.count_nonzero(hi[1:] != hi[:-1]) >= m - 1  # the runs are there
    start = rs - lv[0]
    g = np.arange(0, n, 16)
    assert np.array_equal(t["xmin"][start:], np.minimum.reduceat(sb[0], g))
    assert np.array_equal(t["ymax"][start:], np.maximum.reduceat(sb[3], g))
    o = np.argsort(sb[4])
    for i in range(4):
        assert np.array_equal(sb[i][o], bx[i])


def test_async_builds_on_two_streams():
    # two asynchronous builds share the library's work buffers; enqueued on
    # two streams back to back, both trees must still equal the oracle's
    import curtree_amd as ct

    sets = [(datagen.mixture_boxes(300_001, seed=41), 16),
            (datagen.uniform_boxes(250_000, seed=42, wmax=1e-3), 8)]
    streams = [ct.Stream(), ct.Stream()]
    devs, trees = [], []
    for (bx, F), st in zip(sets, streams):
        ids = np.arange(bx[0].shape[0], dtype=np.int32)
        d = ct.DeviceBoxes.from_host(*bx, ids)
        ct.synchronize()
        devs.append(d)
    for d, (bx, F), st in zip(devs, sets, streams):
        trees.append(ct.build_rtree(d, F, ct.KEY_MORTON64, stream=st.p, sync=False))
    for st in streams:
        ct.synchronize(st.p)
    for d, rt, (bx, F) in zip(devs, trees, sets):
        ids = np.arange(bx[0].shape[0], dtype=np.int32)
        osb, ot = orc.build_rtree(*bx, ids, F, 1)
        sb = d.to_host()
        t = rt.to_host()
        for i, name in enumerate(("xmin", "ymin", "xmax", "ymax", "id")):
            assert_bits_equal(sb[i], osb[i], f"sorted {name}")
        for name in ("xmin", "ymin", "xmax", "ymax", "pos", "len"):
            assert_bits_equal(t[name], ot[name], f"node {name}")
        rt.free()
        d.free()
    for st in streams:
        st.destroy()
