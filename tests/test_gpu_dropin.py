"""GPU: a C++ caller written against the reference's template API
(tests/dropin/dropin_frontier.cpp, in the shape of
src/rtree_test_frontier.cu:37-137) runs unchanged on the library: readMBR of
an edge file and a query file (queries expanded by r = 10), box_h2d,
build_rtree, bfs_frontier and bfs_thrust, idpair_d2h, the two host stable
sorts, and a host nested loop that every pair must match."""
import os
import subprocess

import numpy as np
import pytest

from curtree_amd import datagen

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "dropin", "dropin_frontier")


def write_edges(path, n, seed, scale):
    """Road-like edges in [0, scale]^2 in the readMBR format
    (x1,y1,x2,y2,id,pid,nume), endpoints in either order."""
    x0, y0, x1, y1 = (a.astype(np.float64) * scale for a in datagen.roadlike_boxes(n, seed=seed))
    flip = np.random.default_rng(seed).random(n) < 0.5
    x0[flip], x1[flip] = x1[flip].copy(), x0[flip].copy()
    ids = np.random.default_rng(seed + 1).permutation(n) + 1000
    with open(path, "w") as f:
        for i in range(n):
            f.write(f"{x0[i]:.3f},{y0[i]:.3f},{x1[i]:.3f},{y1[i]:.3f},{ids[i]},{i % 7},{i % 3}\n")


@pytest.mark.parametrize("fanout,key_mode", [(16, 0), (10, 1), (32, 1)])
def test_reference_shaped_driver(tmp_path, fanout, key_mode):
    assert os.path.exists(EXE), "build the drop-in caller first (build() or make -C tests/dropin)"
    data, query = tmp_path / "edges.csv", tmp_path / "queries.csv"
    write_edges(data, 200_000, 501, 3e3)
    write_edges(query, 2_000, 502, 3e3)
    r = subprocess.run([EXE, str(data), "200000", str(query), "2000", str(fanout), str(key_mode)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "verified successfully" in r.stdout
    nh = int(r.stdout.split("num_h=")[1].split()[0])
    assert nh > 10_000  # about 11 edges within r of each query at this density
