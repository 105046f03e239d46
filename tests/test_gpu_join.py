"""GPU parity of the distance self-join (the reference drivers' workload:
queries = the data boxes expanded by r, src/rtree_test_frontier.cu:84-85,
src/utility.cpp:34-37).  Bar: the canonical pair set equals the reference's
brute force (src/utility.h:22-45) over the expanded boxes, bit-exact."""
import numpy as np
import pytest

import oracle_py as orc
from curtree_amd import datagen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fanout,r", [(16, 0.0), (10, 2e-3), (3, 0.01)])
def test_self_join_equals_brute_force(fanout, r):
    import curtree_amd as ct

    n = 20_000
    bx = datagen.clustered_boxes(n, seed=110 + fanout, frac=0.3, sigma=0.01, wmax=2e-3)
    ids = np.arange(n, dtype=np.int32)  # file positions, as the reference's query ids are
    d = ct.DeviceBoxes.from_host(*bx, ids)
    rt = ct.build_rtree(d, fanout, ct.KEY_MORTON64)
    p = ct.self_join(rt, d, r)
    fid, tid = p.to_host()
    p.free()
    f32 = np.float32
    wins = (bx[0] - f32(r), bx[1] - f32(r), bx[2] + f32(r), bx[3] + f32(r))
    bf, bt = orc.brute_force(tuple(bx) + (ids,), wins)
    a = orc.canonical(fid, tid)
    b = orc.canonical(bf, bt)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert fid.size >= n  # every box meets itself
