"""CPU tests of the tree file reader: a file whose node ranges point outside
the tree, or whose size does not match its header, is refused with
CURTREE_EINVAL before any device memory is touched (no GPU here)."""
import ctypes
import struct

import numpy as np
import pytest


def _layout(n, F):
    """Bulk-load node arrays of n boxes at fanout F (root level first)."""
    levels = []
    m = n
    while True:
        m = (m + F - 1) // F
        levels.append(m)
        if m == 1:
            break
    r_sz = sum(levels)
    pos = np.empty(r_sz, np.int32)
    ln = np.empty(r_sz, np.int32)
    start = r_sz
    for li, cnt in enumerate(levels):
        start -= cnt
        child0 = r_sz if li == 0 else start + cnt
        nchild = n if li == 0 else levels[li - 1]
        for i in range(cnt):
            pos[start + i] = child0 + i * F
            ln[start + i] = min(F, nchild - i * F)
    return r_sz, pos, ln


def _write(path, n, F, pos, ln, truncate=0):
    r_sz = pos.shape[0]
    with open(path, "wb") as f:
        f.write(b"CTREE001" + struct.pack("<qqii", n, r_sz, F, 0))
        for _ in range(4):
            f.write(np.zeros(n, np.float32).tobytes())
        f.write(np.arange(n, dtype=np.int32).tobytes())
        for _ in range(4):
            f.write(np.zeros(r_sz, np.float32).tobytes())
        f.write(pos.tobytes())
        f.write(ln.tobytes()[: len(ln.tobytes()) - truncate])


def _load_rc(path):
    import curtree_amd as ct

    rt = ct.RTreeStruct()
    b = ct.BBox()
    rc = ct.lib.curtree_tree_load(str(path).encode(), ctypes.byref(rt), ctypes.byref(b))
    return rc, ct.lib.curtree_last_error().decode()


def test_layout_helper_matches_library_plan():
    import curtree_amd as ct

    for n, F in ((40, 4), (1000, 16), (17, 2), (1, 8)):
        r_sz, pos, ln = _layout(n, F)
        assert ct.tree_size(n, F)[0] == r_sz
        # every box is the child of exactly one leaf-parent
        lp = pos >= r_sz
        assert ln[lp].sum() == n


@pytest.mark.parametrize("which,value", [("pos", -5), ("pos", 10**6), ("len", -1),
                                         ("len", 10**6)])
def test_corrupt_node_range_is_refused(tmp_path, which, value):
    n, F = 40, 4
    r_sz, pos, ln = _layout(n, F)
    (pos if which == "pos" else ln)[3] = value
    p = tmp_path / "bad.ctree"
    _write(p, n, F, pos, ln)
    rc, msg = _load_rc(p)
    assert rc == 1 and "outside the tree" in msg, (rc, msg)


def test_range_straddling_nodes_and_boxes_is_refused(tmp_path):
    n, F = 40, 4
    r_sz, pos, ln = _layout(n, F)
    pos[0] = r_sz - 1  # children would run from the last node into the boxes
    ln[0] = 3
    p = tmp_path / "bad.ctree"
    _write(p, n, F, pos, ln)
    rc, msg = _load_rc(p)
    assert rc == 1 and "outside the tree" in msg, (rc, msg)


def test_short_file_is_refused(tmp_path):
    n, F = 40, 4
    r_sz, pos, ln = _layout(n, F)
    p = tmp_path / "short.ctree"
    _write(p, n, F, pos, ln, truncate=4)
    rc, msg = _load_rc(p)
    assert rc == 1 and "size" in msg, (rc, msg)


def test_not_a_tree_file_is_refused(tmp_path):
    p = tmp_path / "junk.ctree"
    p.write_bytes(b"not a tree file at all, definitely not" * 4)
    rc, msg = _load_rc(p)
    assert rc == 1 and "not a curtree file" in msg, (rc, msg)
