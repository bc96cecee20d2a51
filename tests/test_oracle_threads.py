"""The oracle's threaded mode (the all-core CPU baseline) against its serial mode.

The threaded mode runs the per-point descents, voxel sampling and the per-leaf
folds on several threads. It claims the serial results exactly, in the order
the serial code produces them. That includes which point's colour a voxel takes
(the first point to reach the cell) and the order of each node's points and
voxels. Checked at every step of a build that splits and spills, with every
pass large enough to run threaded.
"""
from __future__ import annotations

from oracle.oracle import OracleTree
from tests.cases import lidar_scene
from tests.helpers import batch_arrays


def _state(o: OracleTree):
    nodes = o.nodes()
    per_node = []
    for i in range(len(nodes)):
        g = o.grid(i) if nodes[i]["grid"] >= 0 else None
        per_node.append((o.samples(i).tobytes(), o.samples(i, voxels=True).tobytes(),
                         None if g is None else g.tobytes()))
    return o.stats(), nodes.tobytes(), o.accums().tobytes(), per_node


def test_threaded_oracle_matches_serial():
    points, meta = lidar_scene(800_000, seed=23)
    batch = 100_000
    slots, sizes, nb = batch_arrays(points, batch)
    serial = OracleTree(meta.box_size, batches_per_step=1)
    threaded = OracleTree(meta.box_size, batches_per_step=1, threads=4)
    done = 0
    while done < nb:
        done = serial.build_step(slots, sizes, nb, batch, nb)
        assert threaded.build_step(slots, sizes, nb, batch, nb) == done
        serial.accumulate(done)
        threaded.accumulate(done)
        a, b = _state(serial), _state(threaded)
        assert a[0] == b[0], f"stats differ after batch {done}"
        assert a[1] == b[1], f"node table differs after batch {done}"
        assert a[2] == b[2], f"accumulators differ after batch {done}"
        for i, (x, y) in enumerate(zip(a[3], b[3])):
            assert x == y, f"node {i}: points, voxels or grid differ after batch {done}"
    # the build split nodes, spilled points and sampled voxels
    s = serial.stats()
    assert s["total_nodes_split"] > 8 and s["total_spilled_points"] > 0 and s["total_new_voxels"] > 0

    def same(what):
        a, b = _state(serial), _state(threaded)
        assert a[:3] == b[:3], f"stats, node table or accumulators differ after {what}"
        for i, (x, y) in enumerate(zip(a[3], b[3])):
            assert x == y, f"node {i}: points, voxels or grid differ after {what}"

    # Refinement's deepening and collapse run the same passes (the spill list
    # descended, sampled and inserted with an empty batch)
    d1, d2 = serial.deepen(8, 20_000), threaded.deepen(8, 20_000)
    assert d1 == d2 and d1["splits"] > 0 and d1["points_moved"] >= 16_384
    same("deepening")
    c1, c2 = serial.collapse(4, 50_000), threaded.collapse(4, 50_000)
    assert c1 == c2 and c1["collapsed"] > 0 and c1["points_moved"] >= 16_384
    same("collapse")
