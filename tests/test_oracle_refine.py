"""The oracle's Refinement (colour filtering) pinned by a numpy box filter.

For a gridded node at level l with region (X, Y, Z), the box filter takes every
point of the cloud whose 20-bit coordinate lies in the region (all points at
level 0), groups them by their 128^3 cell at level l (numpy, from the float32
quantisation of oracle/closed_form.py), and averages each RGBA channel with
(sum + n // 2) // n; positions are the cell centres written out in float32. The
oracle gets there another way (every stored leaf point walked up its root path,
per-node hash maps), so a dropped point, a wrong cell shift, truncation instead
of rounding or a transposed cell order fails here.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.closed_form import quantise
from oracle.oracle import OracleTree
from tests.cases import lidar_scene, tiny_cases
from tests.helpers import batch_arrays, cell_colours


def oracle_build(points, meta, batch):
    slots, sizes, nb = batch_arrays(points, batch)
    o = OracleTree(meta.box_size)
    o.build_step(slots, sizes, nb, batch, nb)
    return o


def region_points(points, size, rec):
    lvl = int(rec["level"])
    if lvl == 0:
        return points
    q20, _ = quantise(points, size)
    q20 = q20 & np.uint64((1 << 20) - 1)
    sh = np.uint64(20 - lvl)
    m = (((q20[:, 0] >> sh) == rec["x"]) & ((q20[:, 1] >> sh) == rec["y"]) & ((q20[:, 2] >> sh) == rec["z"]))
    return points[m]


def box_filter(points, size, rec):
    """(cells ascending, colours, truncated colours, float32 centres) of a node."""
    lvl = int(rec["level"])
    cells, cols = cell_colours(region_points(points, size, rec), size, lvl)
    uniq, inv = np.unique(cells, return_inverse=True)
    n = np.bincount(inv).astype(np.int64)
    colour = np.zeros(len(uniq), dtype=np.int64)
    trunc = np.zeros(len(uniq), dtype=np.int64)
    for ch in range(4):
        s = np.bincount(inv, weights=((cols >> (8 * ch)) & 255).astype(np.float64)).astype(np.int64)
        colour |= ((s + n // 2) // n) << (8 * ch)
        trunc |= (s // n) << (8 * ch)
    ns = np.float32(size) / np.float32(2.0 ** lvl)
    cxyz = [uniq % 128, (uniq // 128) % 128, uniq // 16384]
    centre = []
    for a, ax in enumerate(("x", "y", "z")):
        mn = np.float32(rec[ax]) * ns
        t = ns * (cxyz[a].astype(np.float32) + np.float32(0.5))
        centre.append(mn + t / np.float32(128.0))
    return uniq, colour, trunc, centre


def grid_cells(words):
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    return np.nonzero(bits)[0].astype(np.int64)


def check_refined(o, points, size):
    nodes = o.nodes()
    rounding_matters = False
    checked = 0
    for i, rec in enumerate(nodes):
        if rec["grid"] < 0:
            continue
        vox = o.samples(i, True)
        cells = grid_cells(o.grid(i))
        assert rec["num_voxels"] == rec["num_voxels_stored"] == len(cells) == len(vox)
        uniq, colour, trunc, centre = box_filter(points, size, rec)
        assert np.array_equal(uniq, cells), f"occupied cells differ from the box filter at node {i}"
        assert np.array_equal(vox["color"].astype(np.int64), colour), f"colours differ at node {i}"
        for a, ax in enumerate(("x", "y", "z")):
            assert np.array_equal(vox[ax].view(np.uint32), centre[a].view(np.uint32)), f"{ax} differs at {i}"
        rounding_matters |= bool((colour != trunc).any())
        checked += 1
    assert checked > 0
    # the scene exercises the rounding: truncation would have failed above
    assert rounding_matters


def test_refine_equals_the_numpy_box_filter_lidar():
    points, meta = lidar_scene(400_000, seed=5)
    size = float(max(meta.box_size))
    o = oracle_build(points, meta, 100_000)
    nodes = o.nodes()
    gridded = int((nodes["grid"] >= 0).sum())
    assert gridded >= 5 and nodes["level"][nodes["grid"] >= 0].max() >= 1  # gridded nodes with gridded children
    refined, skipped, empty = o.refine(0)
    assert (refined, skipped, empty) == (gridded, 0, 0)
    check_refined(o, points, size)


@pytest.mark.parametrize("name", ["root_leaf_voxels", "skewed_ragged", "box_max_wrap", "one_over_one_octant"])
def test_refine_equals_the_numpy_box_filter_tiny(name):
    points, meta, batch = tiny_cases()[name]
    size = float(max(meta.box_size))
    o = oracle_build(points, meta, batch)
    before = o.nodes()
    refined, skipped, empty = o.refine(0)
    assert skipped == 0 and empty == 0
    assert refined == int((before["grid"] >= 0).sum())
    check_refined(o, points, size)
    if name == "root_leaf_voxels":
        # the root sampled voxels as a leaf, then again into its re-zeroed grid:
        # refinement leaves one voxel per occupied cell
        assert before[0]["num_voxels"] > o.nodes()[0]["num_voxels"]


def test_refine_is_idempotent_and_keeps_positions():
    points, meta, batch = tiny_cases()["skewed_ragged"]
    o = oracle_build(points, meta, batch)
    nodes = o.nodes()
    gridded = [i for i in range(len(nodes)) if nodes[i]["grid"] >= 0]
    pos_before = {i: np.sort(o.samples(i, True)[["x", "y", "z"]]) for i in gridded}
    o.refine(0)
    first = {i: o.samples(i, True).copy() for i in gridded}
    o.refine(0)
    for i in gridded:
        assert first[i].tobytes() == o.samples(i, True).tobytes()
        # without the root quirk the voxel positions are the ones sampling wrote
        assert np.array_equal(np.sort(first[i][["x", "y", "z"]]), pos_before[i])


def test_refine_min_level_leaves_shallower_nodes_alone():
    points, meta = lidar_scene(400_000, seed=6)
    o = oracle_build(points, meta, 100_000)
    nodes = o.nodes()
    gridded = [i for i in range(len(nodes)) if nodes[i]["grid"] >= 0]
    deepest = max(int(nodes[i]["level"]) for i in gridded)
    before = {i: o.samples(i, True).copy() for i in gridded}
    refined, skipped, _ = o.refine(deepest)
    assert refined == sum(1 for i in gridded if nodes[i]["level"] == deepest) and skipped == 0
    for i in gridded:
        after = o.samples(i, True)
        if nodes[i]["level"] < deepest:
            assert after.tobytes() == before[i].tobytes()
        else:
            assert np.all(np.diff(grid_cells(o.grid(i))) > 0)
