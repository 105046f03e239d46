"""GPU: the allocation-free query (curtree_query_into_noalloc_f32), called
through the C ABI, against the allocating curtree_query_into_f32 on the same
inputs: identical pair counts and pairs, element for element, on two calls in
a row (the second with a smaller tree and fewer windows, so that it reuses the
scratch the first one left behind), plus its limits: too many windows is
CURTREE_EINVAL, a frontier beyond the scratch is CURTREE_ERANGE, and a small
caller buffer reports the pair count as query_into does."""
import ctypes

import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

pytestmark = pytest.mark.gpu


def _buf(ct, n):
    p = ctypes.c_void_p()
    ct._check(ct.lib.curtree_malloc(ctypes.byref(p), 4 * max(int(n), 1)))
    return p.value


def _pairs(ct, f, t, n):
    return ct._host_from_dev(f, n, np.int32), ct._host_from_dev(t, n, np.int32)


def _both(ct, rt, d, q, cap):
    bufs = [_buf(ct, cap) for _ in range(4)]
    try:
        n_ref = ct.query_into(rt, d, q, bufs[0], bufs[1], cap)
        n_na = ct.query_into_noalloc(rt, d, q, bufs[2], bufs[3], cap)
        ref = _pairs(ct, bufs[0], bufs[1], n_ref) if n_ref > 0 else None
        got = _pairs(ct, bufs[2], bufs[3], n_na) if n_na > 0 else None
    finally:
        for b in bufs:
            ct.lib.curtree_free(ctypes.c_void_p(b))
    return n_ref, n_na, ref, got


def test_noalloc_equals_query_into_twice_then_smaller():
    import curtree_amd as ct

    cases = [  # (boxes, windows, fanout): the second call is smaller in both
        (datagen.mixture_boxes(400_000, seed=301), datagen.loguniform_windows(
            20_000, seed=302, lo=1e-4, hi=2e-2), 16),
        (datagen.uniform_boxes(60_000, seed=303, wmax=2e-3), datagen.uniform_windows(
            3_000, seed=304, smax=0.02), 8),
    ]
    for bx, ws, F in cases:
        d = ct.DeviceBoxes.from_host(*bx)
        rt = ct.build_rtree(d, F, ct.KEY_MORTON64)
        q = ct.DeviceBoxes.from_host(*ws)
        H = ct.query_count(rt, d, q)
        n_ref, n_na, ref, got = _both(ct, rt, d, q, H)
        print(f"[case] n={d.n} q={q.n} F={F}: {n_ref} pairs", flush=True)
        assert n_ref == H and n_na == H
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
        # and against the oracle's traversal of the same sorted boxes and tree
        sb = d.to_host()
        ofid, otid, _ = orc.bfs(rt.to_host(), sb, ws)
        assert np.array_equal(got[0], ofid) and np.array_equal(got[1], otid)
        q.free()
        rt.free()
        d.free()


def test_noalloc_small_buffer_reports_count():
    import curtree_amd as ct

    bx = datagen.uniform_boxes(50_000, seed=305, wmax=2e-3)
    d = ct.DeviceBoxes.from_host(*bx)
    rt = ct.build_rtree(d, 16, ct.KEY_MORTON64)
    q = ct.DeviceBoxes.from_host(*datagen.uniform_windows(2_000, seed=306, smax=0.02))
    H = ct.query_count(rt, d, q)
    assert H > 10
    n_ref, n_na, _, _ = _both(ct, rt, d, q, H // 2)
    assert n_ref == -H and n_na == -H
    q.free()
    rt.free()
    d.free()


def test_noalloc_limits():
    import curtree_amd as ct

    bx = datagen.uniform_boxes(1_000_000, seed=307, wmax=1e-3)
    d = ct.DeviceBoxes.from_host(*bx)
    rt = ct.build_rtree(d, 16, ct.KEY_MORTON64)
    # 100 windows over the whole square: 100 x 62,500 leaf-parent entries,
    # beyond the 3,145,728 entries of the scratch
    big = tuple(np.full(100, v, np.float32) for v in (-1.0, -1.0, 2.0, 2.0))
    q = ct.DeviceBoxes.from_host(*big)
    f, t = _buf(ct, 16), _buf(ct, 16)
    n = ctypes.c_int64()
    rc = ct.lib.curtree_query_into_noalloc_f32(ctypes.byref(rt.s), ctypes.byref(d.s),
                                               ctypes.byref(q.s), f, t, 16, ctypes.byref(n),
                                               None)
    assert rc == ct.ERANGE and b"scratch" in ct.lib.curtree_last_error(), rc
    # too many windows for one call: refused before any device work
    v = ct.BBox()
    v.sz = (3 << 20) // 2 + 1
    v.xmin = v.ymin = v.xmax = v.ymax = q.s.xmin
    rc = ct.lib.curtree_query_into_noalloc_f32(ctypes.byref(rt.s), ctypes.byref(d.s),
                                               ctypes.byref(v), f, t, 16, ctypes.byref(n), None)
    assert rc == ct.EINVAL and b"at most" in ct.lib.curtree_last_error(), rc
    # the scratch still answers correctly after the refused calls
    small = ct.DeviceBoxes.from_host(*datagen.uniform_windows(500, seed=308, smax=0.01))
    H = ct.query_count(rt, d, small)
    n_ref, n_na, ref, got = _both(ct, rt, d, small, H)
    assert n_ref == n_na == H
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    for b in (f, t):
        ct.lib.curtree_free(ctypes.c_void_p(b))
    small.free()
    q.free()
    rt.free()
    d.free()
