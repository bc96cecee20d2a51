"""The oracle's analytic-bucketing expand arm (fixed depth) against a closed form.

With the depth D fixed and no node-pool pressure, the tree after any sequence of
batches depends only on the set of points: the root, plus the 8 children of
every cell that holds points at each level L < D (a cell at level L is the top L
bits of the 20-bit coordinate); every leaf at level D counts exactly the points
of its cell and stores them; leaves above D are empty; and a gridded node's
occupied 128^3 cells are its region's points' cells, as in the iterative arm.
numpy derives all of it from the quantised coordinates (oracle/closed_form.py),
sharing no code with the C++ restatement.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.closed_form import quantise
from oracle.oracle import OracleTree
from tests.cases import lidar_scene, tiny_cases
from tests.helpers import batch_arrays, cell_colours, keyed, point_rows


def build(points, meta, batch, depth, **kw):
    slots, sizes, nb = batch_arrays(points, batch)
    o = OracleTree(meta.box_size, fixed_depth=depth, **kw)
    done = 0
    while done < nb:
        done = o.build_step(slots, sizes, nb, batch, nb)
    return o, nb


def closed_form(points, size, depth):
    q20, _ = quantise(points, size)
    q20 = (q20 & np.uint64((1 << 20) - 1)).astype(np.int64)
    nodes = {(0, 0, 0, 0)}
    for L in range(depth):
        c = np.unique(q20 >> (20 - L), axis=0)
        for x, y, z in c:
            for i in range(8):
                nodes.add((L + 1, 2 * int(x) + ((i >> 2) & 1), 2 * int(y) + ((i >> 1) & 1), 2 * int(z) + (i & 1)))
    leaf_cells, inv, counts = np.unique(q20 >> (20 - depth), axis=0, return_inverse=True, return_counts=True)
    leaf_count = {(depth, int(x), int(y), int(z)): int(n) for (x, y, z), n in zip(leaf_cells, counts)}
    return nodes, leaf_count, q20, inv.ravel(), leaf_cells


def grid_cells(words):
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    return np.nonzero(bits)[0].astype(np.int64)


def check(o, nb, points, size, depth):
    nodes = o.nodes()
    k = keyed(nodes)
    want_nodes, leaf_count, q20, inv, leaf_cells = closed_form(points, size, depth)
    assert set(k) == want_nodes
    parents = {(kk[0] - 1, kk[1] >> 1, kk[2] >> 1, kk[3] >> 1) for kk in want_nodes if kk[0] > 0}
    for key, i in k.items():
        r = nodes[i]
        leaf = r["first_child"] < 0
        assert leaf == (key not in parents), key
        if leaf:
            n = leaf_count.get(key, 0)
            assert r["counter"] == r["num_points"] == n, key
        else:
            # occupied cells = the region's points' cells at this level
            L = key[0]
            m = np.all((q20 >> (20 - L)) == np.array(key[1:], dtype=np.int64), axis=1) if L else np.ones(len(q20), bool)
            cells, _ = cell_colours(points[m], size, L)
            assert np.array_equal(np.unique(cells), grid_cells(o.grid(i))), key
    # per-leaf point multisets at depth D
    for j, (x, y, z) in enumerate(leaf_cells[:50]):
        i = k[(depth, int(x), int(y), int(z))]
        assert np.array_equal(point_rows(o.samples(i)), point_rows(points[inv == j]))
    s = o.stats()
    assert s["total_expand_iters"] == nb and s["total_spilled_points"] == 0
    assert s["num_points"] == len(points)


@pytest.mark.parametrize("depth", [3, 4])
def test_fixed_depth_lidar_closed_form(depth):
    points, meta = lidar_scene(400_000, seed=8)
    size = float(max(meta.box_size))
    o, nb = build(points, meta, 100_000, depth)
    check(o, nb, points, size, depth)


@pytest.mark.parametrize("name,depth", [("box_max_wrap", 3), ("skewed_ragged", 3), ("single_point", 2),
                                        ("root_leaf_voxels", 2)])
def test_fixed_depth_tiny_closed_form(name, depth):
    points, meta, batch = tiny_cases()[name]
    size = float(max(meta.box_size))
    o, nb = build(points, meta, batch, depth)
    check(o, nb, points, size, depth)


def test_fixed_depth_leaves_exceed_the_split_threshold():
    """Leaves at D never split: at depth 2 the 230,017-point cluster stays in one leaf."""
    points, meta, batch = tiny_cases()["skewed_ragged"]
    o, _ = build(points, meta, batch, 2)
    n = o.nodes()
    assert n["level"].max() == 2 and n["num_points"].max() > 230_000


def test_fixed_depth_pool_overflow_keeps_every_point():
    """Under node-pool pressure a refused split stays a leaf above D and keeps
    collecting its region's points: nothing is dropped."""
    points, meta = lidar_scene(400_000, seed=9)
    o, _ = build(points, meta, 100_000, 5, max_nodes=200)
    s = o.stats()
    assert s["node_pool_overflow"] == 1 and s["num_nodes"] <= 200
    assert s["num_points"] == len(points)
    n = o.nodes()
    leaves = n[n["first_child"] < 0]
    assert (leaves["counter"] >= leaves["num_points"]).all()
    assert int(leaves["num_points"].sum()) == len(points)
