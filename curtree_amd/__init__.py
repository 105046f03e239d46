"""curtree_amd -- MI355X (gfx950) R-tree bulk load and window query.

Python host mirror of the reference's template API (src/rtree.h,
src/rtree_bulkload_thrust.cuh, src/rtree_bfs_frontier.cuh,
src/rtree_bfs_thrust.cuh, src/utility.cuh) over the C ABI of
libcurtree_amd.so (include/curtree/curtree_c.h).  All device work runs in the
HIP library; this module only moves arrays and calls entry points.  There is no
CPU fallback: importing this module fails if the library is missing.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "lib", "CurtreeError", "DeviceBoxes", "RTree", "DevicePairs", "build_rtree",
    "bfs_frontier", "bfs_thrust", "query", "query_into", "query_count", "canonical_sort",
    "KEY_REF32", "KEY_MORTON64", "ALGO_FRONTIER", "ALGO_PERSISTENT", "synchronize",
    "device_info", "Prof", "Event", "Stream",
]

HERE = os.path.dirname(os.path.abspath(__file__))
# CURTREE_AMD_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("CURTREE_AMD_LIB") or os.path.join(HERE, "libcurtree_amd.so")

KEY_REF32 = 0
KEY_MORTON64 = 1
ALGO_FRONTIER = 0
ALGO_PERSISTENT = 1

OK, EINVAL, ENOMEM, EHIP, ERANGE, EBUSY = 0, 1, 2, 3, 4, 5


class CurtreeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"curtree error {code}: {msg}")
        self.code = code


class BBox(ctypes.Structure):
    """curtree_bbox_f32 == BBOX<float> (src/rtree.h:3-11)."""
    _fields_ = [("sz", ctypes.c_int), ("xmin", ctypes.c_void_p), ("ymin", ctypes.c_void_p),
                ("xmax", ctypes.c_void_p), ("ymax", ctypes.c_void_p), ("id", ctypes.c_void_p)]


class RTreeStruct(ctypes.Structure):
    """curtree_rtree_f32 == RTREE<float> (src/rtree.h:13-23)."""
    _fields_ = [("sz", ctypes.c_int), ("fanout", ctypes.c_int), ("xmin", ctypes.c_void_p),
                ("ymin", ctypes.c_void_p), ("xmax", ctypes.c_void_p), ("ymax", ctypes.c_void_p),
                ("pos", ctypes.c_void_p), ("len", ctypes.c_void_p)]


class IdPair64(ctypes.Structure):
    """curtree_idpair64: IDPAIR (src/rtree.h:25-29) with a 64-bit count."""
    _fields_ = [("sz", ctypes.c_int64), ("fid", ctypes.c_void_p), ("tid", ctypes.c_void_p)]


class BatchStats(ctypes.Structure):
    """curtree_batch_stats."""
    _fields_ = [("pairs", ctypes.c_int64), ("chunks", ctypes.c_int64),
                ("max_chunk_pairs", ctypes.c_int64), ("count_ms", ctypes.c_double),
                ("write_ms", ctypes.c_double)]


class ServerOpts(ctypes.Structure):
    """curtree_server_opts."""
    _fields_ = [("slots", ctypes.c_int), ("budget_bytes", ctypes.c_int64),
                ("max_batch_windows", ctypes.c_int64), ("batching", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """curtree_server_stats."""
    _fields_ = [("requests_done", ctypes.c_int64), ("pairs_done", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("queued", ctypes.c_int64),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("span_us", ctypes.c_double), ("launch_us", ctypes.c_double)]


CHUNK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_int64, ctypes.c_int64, ctypes.c_int64)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
            "g.build()' or make -C curtree_amd)")
    l = ctypes.CDLL(LIB_PATH)
    P, I, I64, V = ctypes.POINTER, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
    sig = {
        "curtree_build_f32": (I, [P(BBox), P(RTreeStruct), I, V]),
        "curtree_build_f32_async": (I, [P(BBox), P(RTreeStruct), I, V]),
        "curtree_build_into_f32": (I, [P(BBox), P(RTreeStruct), I, V]),
        "curtree_tree_size": (I, [I64, I, P(I64), P(I)]),
        "curtree_rt_d_alloc": (I, [P(RTreeStruct), I]),
        "curtree_read_mbr_csv": (I, [ctypes.c_char_p, I, ctypes.c_float, V, V, V, V, V]),
        "curtree_tree_save": (I, [ctypes.c_char_p, P(RTreeStruct), P(BBox)]),
        "curtree_tree_load": (I, [ctypes.c_char_p, P(RTreeStruct), P(BBox)]),
        "curtree_query_f32": (I, [P(RTreeStruct), P(BBox), P(BBox), P(IdPair64), I, V]),
        "curtree_query_into_f32": (I, [P(RTreeStruct), P(BBox), P(BBox), V, V, I64, P(I64), I, V]),
        "curtree_query_count_f32": (I, [P(RTreeStruct), P(BBox), P(BBox), V, P(I64), V]),
        "curtree_query_into_noalloc_f32": (I, [P(RTreeStruct), P(BBox), P(BBox), V, V, I64,
                                               P(I64), V]),
        "curtree_pairs_canonical_sort": (I, [V, V, I64, V]),
        "curtree_self_join_f32": (I, [P(RTreeStruct), P(BBox), ctypes.c_float, P(IdPair64), V]),
        "curtree_query_batched_f32": (I, [P(RTreeStruct), P(BBox), P(BBox), I64, V, V,
                                          P(BatchStats)]),
        "curtree_server_create": (I, [P(ServerOpts), P(V)]),
        "curtree_server_add_tree": (I, [V, P(RTreeStruct), P(BBox), P(I)]),
        "curtree_server_submit": (I, [V, I, P(BBox), P(I64)]),
        "curtree_server_poll": (I, [V, I64, P(I), P(IdPair64), P(ctypes.c_double)]),
        "curtree_server_completed": (I, [V, P(I64), I, P(I)]),
        "curtree_server_release": (I, [V, I64]),
        "curtree_server_get_stats": (I, [V, P(ServerStats)]),
        "curtree_server_reset_stats": (I, [V]),
        "curtree_server_destroy": (I, [V]),
        "curtree_box_d_alloc": (I, [P(BBox), I]),
        "curtree_box_d_free": (I, [P(BBox)]),
        "curtree_rt_d_free": (I, [P(RTreeStruct)]),
        "curtree_idpair64_d_free": (I, [P(IdPair64)]),
        "curtree_malloc": (I, [P(V), I64]),
        "curtree_free": (I, [V]),
        "curtree_memcpy_h2d": (I, [V, V, I64]),
        "curtree_memcpy_d2h": (I, [V, V, I64]),
        "curtree_memcpy_d2d": (I, [V, V, I64, V]),
        "curtree_memset": (I, [V, I, I64]),
        "curtree_synchronize": (I, [V]),
        "curtree_stream_create": (I, [P(V)]),
        "curtree_stream_destroy": (I, [V]),
        "curtree_last_error": (ctypes.c_char_p, []),
        "curtree_version": (ctypes.c_char_p, []),
        "curtree_device_info": (I, [P(I), ctypes.c_char_p, I, P(I), P(I64)]),
        "curtree_release_workspace": (I, []),
        "curtree_query_last_stats": (I, [P(I64), I, P(I), P(I64), P(I64)]),
        "curtree_set_option": (I, [ctypes.c_char_p, I64]),
        "curtree_prof_enable": (I, [I]),
        "curtree_prof_only": (I, [ctypes.c_char_p]),
        "curtree_prof_reset": (I, []),
        "curtree_prof_get": (I, [ctypes.c_char_p, P(ctypes.c_double), P(I64)]),
        "curtree_event_create": (I, [P(V)]),
        "curtree_event_destroy": (I, [V]),
        "curtree_event_record": (I, [V, V]),
        "curtree_event_elapsed": (I, [V, V, P(ctypes.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    return l


lib = _load()


def _check(rc):
    if rc != OK:
        raise CurtreeError(rc, lib.curtree_last_error().decode(errors="replace"))


def _dev_from_host(arr):
    arr = np.ascontiguousarray(arr)
    p = ctypes.c_void_p()
    _check(lib.curtree_malloc(ctypes.byref(p), arr.nbytes))
    if arr.nbytes:
        _check(lib.curtree_memcpy_h2d(p, arr.ctypes.data, arr.nbytes))
    return p.value


def _host_from_dev(ptr, n, dtype):
    out = np.empty(int(n), dtype=dtype)
    if out.nbytes:
        _check(lib.curtree_memcpy_d2h(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes))
    return out


def query_last_stats():
    """Frontier sizes of the last plain query (see curtree_query_last_stats)."""
    ent = (ctypes.c_int64 * 64)()
    nl, pairs, pieces = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
    _check(lib.curtree_query_last_stats(ent, 64, ctypes.byref(nl), ctypes.byref(pairs),
                                        ctypes.byref(pieces)))
    return {"entries": [int(ent[i]) for i in range(nl.value)], "pairs": pairs.value,
            "pieces": pieces.value}


def set_option(name, value):
    """Process-wide tuning knob (see curtree_set_option in curtree_c.h)."""
    _check(lib.curtree_set_option(name.encode(), int(value)))


def synchronize(stream=None):
    _check(lib.curtree_synchronize(stream))


def device_info():
    cnt, cus, hbm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    name = ctypes.create_string_buffer(64)
    _check(lib.curtree_device_info(ctypes.byref(cnt), name, 64, ctypes.byref(cus), ctypes.byref(hbm)))
    return {"count": cnt.value, "arch": name.value.decode(), "cus": cus.value, "hbm_bytes": hbm.value}


class DeviceBoxes:
    """A BBOX<float> in device memory (box_h2d / box_d2h / box_d_free,
    src/utility.cuh:16-112)."""

    def __init__(self, n=0):
        self.s = BBox()
        _check(lib.curtree_box_d_alloc(ctypes.byref(self.s), int(n)))

    @classmethod
    def from_host(cls, xmin, ymin, xmax, ymax, ids=None):
        xmin = np.ascontiguousarray(xmin, dtype=np.float32)
        n = xmin.shape[0]
        if ids is None:
            ids = np.arange(n, dtype=np.int32)
        self = cls(n)
        for name, a, dt in (("xmin", xmin, np.float32), ("ymin", ymin, np.float32),
                            ("xmax", xmax, np.float32), ("ymax", ymax, np.float32),
                            ("id", ids, np.int32)):
            a = np.ascontiguousarray(a, dtype=dt)
            if a.shape != (n,):
                raise ValueError(f"{name}: expected shape ({n},), got {a.shape}")
            if n:
                _check(lib.curtree_memcpy_h2d(getattr(self.s, name), a.ctypes.data, a.nbytes))
        return self

    @property
    def n(self):
        return self.s.sz

    def to_host(self):
        n = self.n
        return (_host_from_dev(self.s.xmin, n, np.float32), _host_from_dev(self.s.ymin, n, np.float32),
                _host_from_dev(self.s.xmax, n, np.float32), _host_from_dev(self.s.ymax, n, np.float32),
                _host_from_dev(self.s.id, n, np.int32))

    def free(self):
        _check(lib.curtree_box_d_free(ctypes.byref(self.s)))

    def __del__(self):
        try:
            if self.s.xmin or self.s.id:
                lib.curtree_box_d_free(ctypes.byref(self.s))
        except Exception:
            pass


class RTree:
    """An RTREE<float> in device memory; nodes root-level first."""

    def __init__(self, fanout):
        self.s = RTreeStruct()
        self.s.fanout = int(fanout)

    @property
    def sz(self):
        return self.s.sz

    @property
    def fanout(self):
        return self.s.fanout

    def to_host(self):
        n = self.sz
        return {
            "xmin": _host_from_dev(self.s.xmin, n, np.float32),
            "ymin": _host_from_dev(self.s.ymin, n, np.float32),
            "xmax": _host_from_dev(self.s.xmax, n, np.float32),
            "ymax": _host_from_dev(self.s.ymax, n, np.float32),
            "pos": _host_from_dev(self.s.pos, n, np.int32),
            "len": _host_from_dev(self.s.len, n, np.int32),
        }

    def upload(self, host):
        """Overwrite the node arrays with host arrays of the same size (the
        dict layout of to_host)."""
        for k, t in (("xmin", np.float32), ("ymin", np.float32), ("xmax", np.float32),
                     ("ymax", np.float32), ("pos", np.int32), ("len", np.int32)):
            a = np.ascontiguousarray(host[k], dtype=t)
            assert a.shape == (self.sz,)
            if a.nbytes:
                _check(lib.curtree_memcpy_h2d(ctypes.c_void_p(getattr(self.s, k)), a.ctypes.data,
                                              a.nbytes))

    def free(self):
        _check(lib.curtree_rt_d_free(ctypes.byref(self.s)))

    def __del__(self):
        try:
            if self.s.pos:
                lib.curtree_rt_d_free(ctypes.byref(self.s))
        except Exception:
            pass


class DevicePairs:
    """Result pairs in device memory: fid = box id, tid = query index."""

    def __init__(self, s=None):
        self.s = s if s is not None else IdPair64()

    @property
    def n(self):
        return int(self.s.sz)

    def to_host(self):
        return _host_from_dev(self.s.fid, self.n, np.int32), _host_from_dev(self.s.tid, self.n, np.int32)

    def free(self):
        _check(lib.curtree_idpair64_d_free(ctypes.byref(self.s)))

    def __del__(self):
        try:
            if self.s.fid:
                lib.curtree_idpair64_d_free(ctypes.byref(self.s))
        except Exception:
            pass


def tree_size(n, fanout):
    """(R_sz, levels) of a bulk-loaded tree of n boxes."""
    rs, nl = ctypes.c_int64(), ctypes.c_int()
    _check(lib.curtree_tree_size(int(n), int(fanout), ctypes.byref(rs), ctypes.byref(nl)))
    return rs.value, nl.value


def build_rtree(boxes, fanout, key_mode=KEY_REF32, stream=None, sync=True):
    """build_rtree<float> (src/rtree_bulkload_thrust.cuh:62-213): sorts `boxes`
    in place and returns the tree (its .sz is the reference's return value)."""
    rt = RTree(fanout)
    fn = lib.curtree_build_f32 if sync else lib.curtree_build_f32_async
    _check(fn(ctypes.byref(boxes.s), ctypes.byref(rt.s), int(key_mode), stream))
    return rt


def alloc_tree(n, fanout):
    """Node arrays for n boxes (rt_d_alloc, src/utility.cuh:114-135)."""
    rt = RTree(fanout)
    _check(lib.curtree_rt_d_alloc(ctypes.byref(rt.s), tree_size(n, fanout)[0]))
    rt.s.fanout = int(fanout)
    return rt


def build_into(boxes, rt, key_mode=KEY_REF32, stream=None):
    """Asynchronous bulk load into preallocated node arrays (no allocation)."""
    _check(lib.curtree_build_into_f32(ctypes.byref(boxes.s), ctypes.byref(rt.s), int(key_mode),
                                      stream))
    return rt


def query(rt, rbox, qbox, algo=ALGO_FRONTIER, stream=None):
    """Window query; pairs in (query, sorted box position) order, on device."""
    out = IdPair64()
    _check(lib.curtree_query_f32(ctypes.byref(rt.s), ctypes.byref(rbox.s), ctypes.byref(qbox.s),
                                 ctypes.byref(out), int(algo), stream))
    return DevicePairs(out)


def query_into(rt, rbox, qbox, fid_ptr, tid_ptr, cap, algo=ALGO_FRONTIER, stream=None):
    n = ctypes.c_int64()
    rc = lib.curtree_query_into_f32(ctypes.byref(rt.s), ctypes.byref(rbox.s), ctypes.byref(qbox.s),
                                    fid_ptr, tid_ptr, int(cap), ctypes.byref(n), int(algo), stream)
    if rc == ERANGE:
        return -n.value
    _check(rc)
    return n.value


def query_into_noalloc(rt, rbox, qbox, fid_ptr, tid_ptr, cap, stream=None):
    """query_into without any device allocation (the library's fixed scratch);
    returns the pair count, negated when it exceeds `cap`."""
    n = ctypes.c_int64()
    rc = lib.curtree_query_into_noalloc_f32(ctypes.byref(rt.s), ctypes.byref(rbox.s),
                                            ctypes.byref(qbox.s), fid_ptr, tid_ptr, int(cap),
                                            ctypes.byref(n), stream)
    if rc == ERANGE and n.value > cap:
        return -n.value
    _check(rc)
    return n.value


def query_count(rt, rbox, qbox, per_query=False, stream=None):
    """Count pass only: total pairs, and per-query counts if asked."""
    n = ctypes.c_int64()
    cptr = None
    if per_query and qbox.n:
        p = ctypes.c_void_p()
        _check(lib.curtree_malloc(ctypes.byref(p), 8 * qbox.n))
        cptr = p.value
    try:
        _check(lib.curtree_query_count_f32(ctypes.byref(rt.s), ctypes.byref(rbox.s),
                                           ctypes.byref(qbox.s), cptr, ctypes.byref(n), stream))
        counts = _host_from_dev(cptr, qbox.n, np.int64) if cptr else None
    finally:
        if cptr:
            lib.curtree_free(ctypes.c_void_p(cptr))
    return (n.value, counts) if per_query else n.value


def query_batched(rt, rbox, qbox, budget_pairs, on_chunk=None):
    """Batched query under a result budget: count pass, then one graph replay
    per chunk of windows.  on_chunk(fid_ptr, tid_ptr, n, q_first, q_end) gets
    each chunk's device pairs; return True (or None) to continue.  Returns the
    curtree_batch_stats as a dict."""
    def cb(_user, fid, tid, n, q0, q1):
        if on_chunk is None:
            return 0
        try:
            r = on_chunk(fid, tid, n, q0, q1)
            return 0 if (r is None or r) else 1
        except Exception:  # stop the batch; the error is reported below
            import traceback
            traceback.print_exc()
            return 1
    fn = CHUNK_FN(cb)
    st = BatchStats()
    _check(lib.curtree_query_batched_f32(ctypes.byref(rt.s), ctypes.byref(rbox.s),
                                         ctypes.byref(qbox.s), int(budget_pairs),
                                         ctypes.cast(fn, ctypes.c_void_p), None, ctypes.byref(st)))
    return {k: getattr(st, k) for k, _ in BatchStats._fields_}


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
kQThreads = 256  # threads per block of the chained query kernels


class Server:
    """Serving runtime: resident trees, continuous batching of window requests
    (see curtree_server_* in curtree_c.h)."""

    def __init__(self, slots=4, budget_bytes=16 << 30, max_batch_windows=1 << 20, batching=True):
        o = ServerOpts(int(slots), int(budget_bytes), int(max_batch_windows), 1 if batching else 0)
        p = ctypes.c_void_p()
        _check(lib.curtree_server_create(ctypes.byref(o), ctypes.byref(p)))
        self.p = p.value
        self._keep = {}

    def add_tree(self, rt, rbox):
        tid = ctypes.c_int()
        _check(lib.curtree_server_add_tree(self.p, ctypes.byref(rt.s), ctypes.byref(rbox.s),
                                           ctypes.byref(tid)))
        return tid.value

    def submit(self, tree_id, windows):
        """windows: a BBox struct of device arrays (kept alive until released)."""
        rid = ctypes.c_int64()
        _check(lib.curtree_server_submit(self.p, int(tree_id), ctypes.byref(windows),
                                         ctypes.byref(rid)))
        self._keep[rid.value] = windows
        return rid.value

    def poll(self, rid):
        """(state, DevicePairs-like (n, fid_ptr, tid_ptr) or None, latency_us)."""
        st = ctypes.c_int()
        out = IdPair64()
        lat = ctypes.c_double()
        rc = lib.curtree_server_poll(self.p, int(rid), ctypes.byref(st), ctypes.byref(out),
                                     ctypes.byref(lat))
        if st.value == REQ_FAILED:
            raise CurtreeError(rc, lib.curtree_last_error().decode(errors="replace"))
        _check(rc)
        res = (int(out.sz), out.fid, out.tid) if st.value == REQ_DONE else None
        return st.value, res, lat.value

    def completed(self, cap=4096):
        """Ids of requests finished since the last call (completion queue)."""
        if getattr(self, "_ids_cap", 0) < cap:  # one buffer per server: cheap to poll
            self._ids = (ctypes.c_int64 * cap)()
            self._ids_cap = cap
            self._n = ctypes.c_int()
        _check(lib.curtree_server_completed(self.p, self._ids, cap, ctypes.byref(self._n)))
        return self._ids[:self._n.value]

    def release(self, rid):
        _check(lib.curtree_server_release(self.p, int(rid)))
        self._keep.pop(rid, None)

    def stats(self):
        s = ServerStats()
        _check(lib.curtree_server_get_stats(self.p, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib.curtree_server_reset_stats(self.p))

    def close(self):
        if self.p:
            _check(lib.curtree_server_destroy(self.p))
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def window_view(boxes, start, count):
    """A BBox struct viewing windows [start, start+count) of device boxes."""
    v = BBox()
    v.sz = int(count)
    for f in ("xmin", "ymin", "xmax", "ymax"):
        base = getattr(boxes.s, f)
        setattr(v, f, (base or 0) + 4 * int(start))
    v.id = None
    return v


def self_join(rt, rbox, r, stream=None):
    """Distance self-join: pairs (box id, querying box id) of boxes within r of
    each other (the reference drivers' queries = data boxes expanded by r)."""
    out = IdPair64()
    _check(lib.curtree_self_join_f32(ctypes.byref(rt.s), ctypes.byref(rbox.s), float(r),
                                     ctypes.byref(out), stream))
    return DevicePairs(out)


def canonical_sort(pairs, stream=None):
    """Sort device pairs into (query id, box id) order in place
    (the host sorts of src/rtree_test_frontier.cu:102-105, on the device)."""
    _check(lib.curtree_pairs_canonical_sort(pairs.s.fid, pairs.s.tid, pairs.n, stream))
    return pairs


def read_mbr(path, size, r=0.0):
    """readMBR (src/utility.cpp:5-40): host arrays (xmin, ymin, xmax, ymax, id)."""
    out = [np.empty(int(size), np.float32) for _ in range(4)] + [np.empty(int(size), np.int32)]
    _check(lib.curtree_read_mbr_csv(os.fsencode(path), int(size), float(r),
                                    *[a.ctypes.data for a in out]))
    return tuple(out)


def save_tree(path, rt, rbox):
    _check(lib.curtree_tree_save(os.fsencode(path), ctypes.byref(rt.s), ctypes.byref(rbox.s)))


def load_tree(path):
    """Returns (RTree, DeviceBoxes) read from a file written by save_tree."""
    rt = RTree(2)
    boxes = DeviceBoxes(0)
    _check(lib.curtree_tree_load(os.fsencode(path), ctypes.byref(rt.s), ctypes.byref(boxes.s)))
    return rt, boxes


def bfs_frontier(rt, rbox, qbox):
    """bfs_frontier<float> (src/rtree_bfs_frontier.cuh:292-529) followed by
    idpair_d2h: returns host arrays (fid, tid) in the reference's pre-sort
    order.  Never aborts on large frontiers."""
    p = query(rt, rbox, qbox, ALGO_FRONTIER)
    try:
        return p.to_host()
    finally:
        p.free()


def bfs_thrust(rt, rbox, qbox):
    """bfs_thrust<float> (src/rtree_bfs_thrust.cuh:58-256): same pair set."""
    return bfs_frontier(rt, rbox, qbox)


class Prof:
    """Per-kernel HIP-event timing inside the library."""

    @staticmethod
    def enable(on=True):
        _check(lib.curtree_prof_enable(1 if on else 0))

    @staticmethod
    def only(name=None):
        """Bracket only the launches of kernel `name` (None: every kernel)."""
        _check(lib.curtree_prof_only(name.encode() if name else None))

    @staticmethod
    def reset():
        _check(lib.curtree_prof_reset())

    @staticmethod
    def get(name):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(lib.curtree_prof_get(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib.curtree_event_create(ctypes.byref(p)))
        self.p = p.value

    def record(self, stream=None):
        _check(lib.curtree_event_record(self.p, stream))

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _check(lib.curtree_event_elapsed(self.p, later.p, ctypes.byref(ms)))
        return ms.value


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib.curtree_stream_create(ctypes.byref(p)))
        self.p = p.value

    def destroy(self):
        if self.p:
            lib.curtree_stream_destroy(self.p)
            self.p = None
