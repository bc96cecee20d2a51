"""The oracle's local deepening against numpy.

Deepening splits the densest leaves (>= min_points stored points) before they
reach 50,000 and moves their points to the children. Checked from the outside:
which leaves were chosen (numpy ranking of the leaf table), that each child now
holds exactly its octant of the former leaf's points (counters included), that
the former leaf's occupied cells are its points' cells, and that nothing else
in the tree changed.
"""
from __future__ import annotations

import numpy as np

from oracle.closed_form import quantise
from oracle.oracle import OracleTree
from tests.cases import lidar_scene
from tests.helpers import batch_arrays, cell_colours, keyed, point_rows


def test_deepen_against_numpy():
    points, meta = lidar_scene(600_000, seed=17)
    size = float(max(meta.box_size))
    slots, sizes, nb = batch_arrays(points, 150_000)
    o = OracleTree(meta.box_size)
    o.build_step(slots, sizes, nb, 150_000, nb)
    before = o.nodes()
    kb = keyed(before)
    leaf = (before["first_child"] < 0) & (before["num_points"] >= 20_000) & (before["level"] < 20)
    ids = np.nonzero(leaf)[0]
    order = sorted(ids, key=lambda i: (-int(before[i]["num_points"]), int(before[i]["level"]), int(before[i]["x"]),
                                       int(before[i]["y"]), int(before[i]["z"])))
    want = order[:5]
    assert len(order) > 5
    stored = {int(i): o.samples(int(i)) for i in want}
    res = o.deepen(5, 20_000)
    assert res["candidates"] == len(order) and res["splits"] == 5
    assert res["points_moved"] == sum(len(v) for v in stored.values())
    after = o.nodes()
    assert len(after) == len(before) + 40
    ka = keyed(after)
    for i in want:
        r = after[i]
        assert r["first_child"] >= 0 and r["num_points"] == 0 and r["grid"] >= 0
        pts = stored[int(i)]
        L = int(r["level"])
        q20, _ = quantise(pts, size)
        q20 = q20 & np.uint64((1 << 20) - 1)
        octant = (((q20[:, 0] >> np.uint64(19 - L)) & np.uint64(1)) << np.uint64(2)) | \
                 (((q20[:, 1] >> np.uint64(19 - L)) & np.uint64(1)) << np.uint64(1)) | \
                 ((q20[:, 2] >> np.uint64(19 - L)) & np.uint64(1))
        for c in range(8):
            child = int(r["first_child"]) + c
            part = pts[octant == c]
            assert after[child]["counter"] == after[child]["num_points"] == len(part)
            assert np.array_equal(point_rows(o.samples(child)), point_rows(part))
        cells, _ = cell_colours(pts, size, L)
        bits = np.unpackbits(o.grid(int(i)).view(np.uint8), bitorder="little")
        assert np.array_equal(np.nonzero(bits)[0], np.unique(cells))
    # every other node is as it was
    for k, j in kb.items():
        if j in want:
            continue
        ra, rb = after[ka[k]], before[j]
        for f in ("first_child", "counter", "num_points", "num_voxels", "grid"):
            assert ra[f] == rb[f], (k, f)
    assert o.stats()["num_points"] == len(points)
