"""GPU: the autotuner times every (fanout, tile steps, grid) combination on a
small workload, checks that all return the same pair count, and its choice
can be applied."""
import pytest

from curtree_amd import datagen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_autotune_small(mode):
    import curtree_amd as ct
    from curtree_amd import autotune

    boxes = datagen.clustered_boxes(150_000, seed=150, frac=0.1, k=5, sigma=1e-3, wmax=1e-3)
    wins = datagen.loguniform_windows(10_000, seed=151, lo=1e-4, hi=0.02)
    res = autotune.tune(boxes, wins, mode=mode, fanouts=(8, 16), tile_steps=(4, 8),
                        grids=(0, 1024), budget_pairs=1 << 20, repeats=1)
    rows = res["rows"]
    assert len(rows) == 2 * 2 * 2
    assert len({r["pairs"] for r in rows}) == 1 and rows[0]["pairs"] > 0
    assert res["best"] in rows
    try:
        f = autotune.apply(res)
        assert f in (8, 16)
        d = ct.DeviceBoxes.from_host(*boxes)
        rt = ct.build_rtree(d, f, ct.KEY_MORTON64)
        q = ct.DeviceBoxes.from_host(*wins)
        assert ct.query_count(rt, d, q) == rows[0]["pairs"]
    finally:
        ct.set_option("query_tile_steps", 8)
        ct.set_option("query_grid", 0)
