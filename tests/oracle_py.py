"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        _lib = ctypes.CDLL(ORACLE_SO)
        I64, P = ctypes.c_int64, ctypes.c_void_p
        _lib.or_dilate16.restype = ctypes.c_uint32
        _lib.or_dilate16.argtypes = [ctypes.c_uint32]
        _lib.or_dilate32.restype = ctypes.c_uint64
        _lib.or_dilate32.argtypes = [ctypes.c_uint64]
        _lib.or_zorder16.restype = ctypes.c_uint32
        _lib.or_zorder16.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        _lib.or_fmin_one.restype = ctypes.c_float
        _lib.or_fmin_one.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.or_fmax_one.restype = ctypes.c_float
        _lib.or_fmax_one.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.or_intersect_one.restype = ctypes.c_int
        _lib.or_intersect_one.argtypes = [ctypes.c_float] * 8
        _lib.or_m64_scale.restype = ctypes.c_float
        _lib.or_m64_scale.argtypes = [ctypes.c_float, ctypes.c_float]
        for name in ("or_keys_ref32", "or_keys_morton64", "or_centroid_extent", "or_sort_perm_i32",
                     "or_sort_perm_u64", "or_permute_boxes", "or_canonical_sort"):
            getattr(_lib, name).restype = None
        _lib.or_level_plan.restype = ctypes.c_int
        _lib.or_level_plan.argtypes = [I64, ctypes.c_int, P, ctypes.c_int, P]
        _lib.or_level_plan_float.restype = ctypes.c_int
        _lib.or_level_plan_float.argtypes = [I64, ctypes.c_int, P, ctypes.c_int, P]
        _lib.or_build_levels.restype = ctypes.c_int
        _lib.or_bfs.restype = I64
        _lib.or_brute_force.restype = I64
        _lib.or_brute_count.restype = I64
        _lib.or_build_rtree.restype = I64
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def keys_ref32(xmin, ymin, xmax, ymax):
    xmin, ymin, xmax, ymax = map(f32, (xmin, ymin, xmax, ymax))
    n = xmin.shape[0]
    k = np.empty(n, np.int32)
    lib().or_keys_ref32(ctypes.c_int64(n), _p(xmin), _p(ymin), _p(xmax), _p(ymax), _p(k))
    return k


def centroid_extent(xmin, ymin, xmax, ymax):
    xmin, ymin, xmax, ymax = map(f32, (xmin, ymin, xmax, ymax))
    ext = np.empty(4, np.float32)
    lib().or_centroid_extent(ctypes.c_int64(xmin.shape[0]), _p(xmin), _p(ymin), _p(xmax), _p(ymax),
                             _p(ext))
    return ext


def keys_morton64(xmin, ymin, xmax, ymax, ext=None):
    xmin, ymin, xmax, ymax = map(f32, (xmin, ymin, xmax, ymax))
    n = xmin.shape[0]
    k = np.empty(n, np.uint64)
    e = None if ext is None else _p(f32(ext))
    lib().or_keys_morton64(ctypes.c_int64(n), _p(xmin), _p(ymin), _p(xmax), _p(ymax), e, _p(k))
    return k


def sort_perm_i32(keys):
    keys = i32(keys)
    perm = np.empty(keys.shape[0], np.uint32)
    lib().or_sort_perm_i32(ctypes.c_int64(keys.shape[0]), _p(keys), _p(perm))
    return perm


def sort_perm_u64(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.empty(keys.shape[0], np.uint32)
    lib().or_sort_perm_u64(ctypes.c_int64(keys.shape[0]), _p(keys), _p(perm))
    return perm


def level_plan(n, fanout, use_float=False):
    lv = np.zeros(64, np.int64)
    rs = np.zeros(1, np.int64)
    fn = lib().or_level_plan_float if use_float else lib().or_level_plan
    nl = fn(n, fanout, _p(lv), 64, _p(rs))
    return [int(x) for x in lv[:nl]], int(rs[0])


def build_rtree(xmin, ymin, xmax, ymax, ids, fanout, key_mode=0):
    """Host bulk load.  Returns (sorted boxes tuple, tree dict)."""
    xmin, ymin, xmax, ymax = (f32(a).copy() for a in (xmin, ymin, xmax, ymax))
    ids = i32(ids).copy()
    n = xmin.shape[0]
    _, r_sz = level_plan(n, fanout)
    t = {k: np.empty(r_sz, np.float32) for k in ("xmin", "ymin", "xmax", "ymax")}
    t["pos"] = np.empty(r_sz, np.int32)
    t["len"] = np.empty(r_sz, np.int32)
    rs = lib().or_build_rtree(ctypes.c_int64(n), ctypes.c_int(fanout), ctypes.c_int(key_mode),
                              _p(xmin), _p(ymin), _p(xmax), _p(ymax), _p(ids), _p(t["xmin"]),
                              _p(t["ymin"]), _p(t["xmax"]), _p(t["ymax"]), _p(t["pos"]),
                              _p(t["len"]))
    assert rs == r_sz
    return (xmin, ymin, xmax, ymax, ids), t


def build_levels(sorted_boxes, fanout):
    xmin, ymin, xmax, ymax = map(f32, sorted_boxes[:4])
    n = xmin.shape[0]
    _, r_sz = level_plan(n, fanout)
    t = {k: np.empty(r_sz, np.float32) for k in ("xmin", "ymin", "xmax", "ymax")}
    t["pos"] = np.empty(r_sz, np.int32)
    t["len"] = np.empty(r_sz, np.int32)
    lib().or_build_levels(ctypes.c_int64(n), ctypes.c_int(fanout), _p(xmin), _p(ymin), _p(xmax),
                          _p(ymax), _p(t["xmin"]), _p(t["ymin"]), _p(t["xmax"]), _p(t["ymax"]),
                          _p(t["pos"]), _p(t["len"]))
    return t


def bfs(tree, sorted_boxes, windows, cap=None, depths=False):
    """Oracle window query: (fid, tid) in (query, sorted position) order and
    counters (X, T, H); with depths=True also the frontier size at every depth
    (depth 0: one entry per window at the root)."""
    xmin, ymin, xmax, ymax = map(f32, sorted_boxes[:4])
    ids = i32(sorted_boxes[4])
    qx0, qy0, qx1, qy1 = map(f32, windows[:4])
    r_sz = tree["pos"].shape[0]
    cnt = np.zeros(3, np.int64)
    L = lib()
    tx0, ty0, tx1, ty1 = (f32(tree[k]) for k in ("xmin", "ymin", "xmax", "ymax"))
    tpos, tlen = i32(tree["pos"]), i32(tree["len"])
    args = [ctypes.c_int64(r_sz), _p(tx0), _p(ty0), _p(tx1), _p(ty1), _p(tpos),
            _p(tlen), _p(xmin), _p(ymin), _p(xmax), _p(ymax), _p(ids),
            ctypes.c_int64(qx0.shape[0]), _p(qx0), _p(qy0), _p(qx1), _p(qy1)]
    dsz = np.zeros(64, np.int64)
    if cap is None:
        L.or_bfs(*args, None, None, ctypes.c_int64(0), _p(cnt), None, ctypes.c_int(0))
        cap = int(cnt[2])
    fid = np.empty(max(cap, 1), np.int32)
    tid = np.empty(max(cap, 1), np.int32)
    h = L.or_bfs(*args, _p(fid), _p(tid), ctypes.c_int64(cap), _p(cnt), _p(dsz),
                 ctypes.c_int(64))
    assert h >= 0
    res = (fid[:h], tid[:h], tuple(int(c) for c in cnt))
    if depths:
        nz = np.nonzero(dsz)[0]
        res = res + ([int(x) for x in dsz[: nz[-1] + 1]] if nz.size else [],)
    return res


def bfs_count(tree, sorted_boxes, windows):
    """Oracle window query, counters only (X, T, H); releases the GIL, so
    calls on disjoint window slices run in parallel threads."""
    xmin, ymin, xmax, ymax = map(f32, sorted_boxes[:4])
    ids = i32(sorted_boxes[4])
    qx0, qy0, qx1, qy1 = map(f32, windows[:4])
    tx0, ty0, tx1, ty1 = (f32(tree[k]) for k in ("xmin", "ymin", "xmax", "ymax"))
    tpos, tlen = i32(tree["pos"]), i32(tree["len"])
    cnt = np.zeros(3, np.int64)
    lib().or_bfs(ctypes.c_int64(tpos.shape[0]), _p(tx0), _p(ty0), _p(tx1), _p(ty1), _p(tpos),
                 _p(tlen), _p(xmin), _p(ymin), _p(xmax), _p(ymax), _p(ids),
                 ctypes.c_int64(qx0.shape[0]), _p(qx0), _p(qy0), _p(qx1), _p(qy1), None, None,
                 ctypes.c_int64(0), _p(cnt), None, ctypes.c_int(0))
    return tuple(int(c) for c in cnt)


def brute_force(boxes, windows):
    """nested_loop_query semantics: (fid, tid) in (query, input order) order."""
    xmin, ymin, xmax, ymax = map(f32, boxes[:4])
    ids = i32(boxes[4])
    qx0, qy0, qx1, qy1 = map(f32, windows[:4])
    args = [ctypes.c_int64(qx0.shape[0]), _p(qx0), _p(qy0), _p(qx1), _p(qy1),
            ctypes.c_int64(xmin.shape[0]), _p(xmin), _p(ymin), _p(xmax), _p(ymax), _p(ids)]
    L = lib()
    hc = brute_count(boxes, windows)
    fid = np.empty(max(hc, 1), np.int32)
    tid = np.empty(max(hc, 1), np.int32)
    h = L.or_brute_force(*args, _p(fid), _p(tid), ctypes.c_int64(hc))
    assert h == hc
    return fid[:h], tid[:h]


def brute_count(boxes, windows, q0=0, q1=None, threads=1):
    xmin, ymin, xmax, ymax = map(f32, boxes[:4])
    qx0, qy0, qx1, qy1 = map(f32, windows[:4])
    if q1 is None:
        q1 = qx0.shape[0]
    return int(lib().or_brute_count(ctypes.c_int64(q0), ctypes.c_int64(q1), _p(qx0), _p(qy0),
                                    _p(qx1), _p(qy1), ctypes.c_int64(xmin.shape[0]), _p(xmin),
                                    _p(ymin), _p(xmax), _p(ymax), ctypes.c_int(threads)))


def canonical_c(fid, tid):
    """The oracle's canonical ordering (or_canonical_sort) on copies."""
    fid = i32(fid).copy()
    tid = i32(tid).copy()
    lib().or_canonical_sort(ctypes.c_int64(fid.shape[0]), _p(fid), _p(tid))
    return fid, tid


def canonical(fid, tid):
    """(query id, box id) lexicographic order."""
    fid = np.asarray(fid)
    tid = np.asarray(tid)
    o = np.lexsort((fid, tid))
    return fid[o], tid[o]
