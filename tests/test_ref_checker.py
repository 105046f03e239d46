"""CPU: pin the oracle against the reference's own CPU code, compiled from the
curtree sources by oracle/ref.mk into oracle/_ref/libref.so (skipped where that
library was not built)."""
import ctypes
import os

import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

REF_SO = os.path.join(orc.ORACLE_DIR, "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built")


def ref():
    L = ctypes.CDLL(REF_SO)
    L.ref_nested_loop_query.restype = ctypes.c_int64
    L.ref_read_mbr.restype = ctypes.c_int
    return L


def P(a):
    return ctypes.c_void_p(a.ctypes.data)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_brute_force_equals_reference_nested_loop(seed):
    n, nq = 4000, 300
    bx = datagen.clustered_boxes(n, seed=seed, frac=0.5, sigma=0.02, wmax=0.02)
    ids = np.random.default_rng(seed).permutation(n).astype(np.int32)
    qs = datagen.loguniform_windows(nq, seed=seed + 10, lo=1e-3, hi=0.3)
    qid = np.arange(nq, dtype=np.int32)
    # special values the predicate must treat the same way
    bx = [a.copy() for a in bx]
    bx[0][5] = np.nan
    bx[3][6] = -0.0
    ofid, otid = orc.brute_force(tuple(bx) + (ids,), qs)
    cap = ofid.size + 1
    rf = np.empty(cap, np.int32)
    rt = np.empty(cap, np.int32)
    q = [np.ascontiguousarray(a, np.float32) for a in qs]
    b = [np.ascontiguousarray(a, np.float32) for a in bx]
    h = ref().ref_nested_loop_query(nq, P(q[0]), P(q[1]), P(q[2]), P(q[3]), P(qid), n, P(b[0]),
                                    P(b[1]), P(b[2]), P(b[3]), P(ids), P(rf), P(rt),
                                    ctypes.c_int64(cap))
    assert h == ofid.size
    assert np.array_equal(rf[:h], ofid) and np.array_equal(rt[:h], otid)


def test_reference_csv_reader_semantics(tmp_path):
    # readMBR: 7 comma-separated fields, min/max normalised, expanded by r
    rows = [(1.5, 2.0, 0.5, 3.0, 10, 0, 0), (-1, -1, -2, 5, 11, 1, 2), (0, 0, 0, 0, 12, 3, 4)]
    f = tmp_path / "edges.txt"
    f.write_text("".join(",".join(str(v) for v in r) + "\n" for r in rows))
    for r in (0.0, 10.0):
        out = [np.empty(3, np.float32) for _ in range(4)] + [np.empty(3, np.int32)]
        rc = ref().ref_read_mbr(str(f).encode(), 3, ctypes.c_float(r), *[P(a) for a in out])
        assert rc == 0
        x1 = np.array([v[0] for v in rows], np.float32)
        y1 = np.array([v[1] for v in rows], np.float32)
        x2 = np.array([v[2] for v in rows], np.float32)
        y2 = np.array([v[3] for v in rows], np.float32)
        assert np.array_equal(out[0], np.minimum(x1, x2) - np.float32(r))
        assert np.array_equal(out[1], np.minimum(y1, y2) - np.float32(r))
        assert np.array_equal(out[2], np.maximum(x1, x2) + np.float32(r))
        assert np.array_equal(out[3], np.maximum(y1, y2) + np.float32(r))
        assert out[4].tolist() == [10, 11, 12]
    bad = tmp_path / "bad.txt"
    bad.write_text("1,2,3\n")
    out = [np.empty(1, np.float32) for _ in range(4)] + [np.empty(1, np.int32)]
    assert ref().ref_read_mbr(str(bad).encode(), 1, ctypes.c_float(0), *[P(a) for a in out]) == -1


def test_library_csv_reader_equals_reference_reader(tmp_path):
    import curtree_amd as ct

    r = np.random.default_rng(7)
    n = 2000
    x1, y1, x2, y2 = (r.uniform(-1000, 1000, n).astype(np.float32) for _ in range(4))
    idv = r.integers(0, 10**6, n)
    f = tmp_path / "edges.txt"
    f.write_text("".join(f"{a!r},{b!r},{c!r},{d!r},{i},{i % 7},{i % 3}\n"
                         for a, b, c, d, i in zip(x1.tolist(), y1.tolist(), x2.tolist(),
                                                  y2.tolist(), idv.tolist())))
    for rr in (0.0, 10.0):
        mine = ct.read_mbr(str(f), n, rr)
        out = [np.empty(n, np.float32) for _ in range(4)] + [np.empty(n, np.int32)]
        assert ref().ref_read_mbr(str(f).encode(), n, ctypes.c_float(rr), *[P(a) for a in out]) == 0
        for a, b in zip(mine, out):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    bad = tmp_path / "bad.txt"
    bad.write_text("1,2,3\n")
    with pytest.raises(ct.CurtreeError):
        ct.read_mbr(str(bad), 1)
    with pytest.raises(ct.CurtreeError):
        ct.read_mbr(str(tmp_path / "missing.txt"), 1)
