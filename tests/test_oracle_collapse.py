"""The oracle's collapsing (Refinement) against numpy and against deepening.

A node Update built always held more than 50,000 points when it split, so in a
tree Update alone built no inner node qualifies for a collapse; the nodes that
do are those local deepening split early. That makes deepening and collapsing
inverses, which pins collapse from the outside:

* deepen k leaves, then collapse: exactly those k nodes are the candidates and
  the tree is again the one before deepening -- node set, counters, per-leaf
  point multisets, every remaining grid and voxel count (numpy on the dumps);
  the collapse allocates no store bytes and its pools hold k blocks and k grids;
* deepen again: the same leaves split into the pooled blocks and grids, so the
  node pool does not grow and the tree equals the one deepening gave first;
* keep building after a collapse: the final tree is the closed form of all the
  points (numpy, oracle/closed_form.py), as for a tree Update alone built.

The chunk bookkeeping is held to its closed form: chunks in use equal the sum of
the leaves' chunk counts, and the chunk pool grows by the voxel chunks of the
freed grids.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import closed_form as CF
from oracle.oracle import OracleTree
from tests.cases import lidar_scene
from tests.helpers import batch_arrays, keyed, point_rows


def snapshot(o):
    nodes = o.nodes()
    k = keyed(nodes)
    leaves, grids = {}, {}
    for key, i in k.items():
        r = nodes[i]
        if r["first_child"] < 0:
            leaves[key] = point_rows(o.samples(i)) if r["num_points"] else None
        if r["grid"] >= 0:
            grids[key] = o.grid(i).copy()
    rec = {key: tuple(int(nodes[i][f]) for f in ("counter", "num_points", "num_voxels", "num_voxels_stored"))
           + (bool(nodes[i]["first_child"] < 0), bool(nodes[i]["grid"] >= 0)) for key, i in k.items()}
    return rec, leaves, grids


def assert_same_tree(a, b):
    ra, la, ga = a
    rb, lb, gb = b
    assert ra.keys() == rb.keys()
    for key in ra:
        assert ra[key] == rb[key], (key, ra[key], rb[key])
    for key in la:
        if la[key] is None:
            assert lb[key] is None, key
        else:
            assert np.array_equal(la[key], lb[key]), key
    assert ga.keys() == gb.keys()
    for key in ga:
        assert np.array_equal(ga[key], gb[key]), key


def chunk_closed_form(o):
    nodes = o.nodes()
    live = nodes[nodes["free_slot"] == 0]
    leaf = live["first_child"] < 0
    return int(((live["num_points"][leaf].astype(np.int64) + 999) // 1000).sum())


def built(n=600_000, seed=17, batch=150_000, **kw):
    points, meta = lidar_scene(n, seed=seed)
    slots, sizes, nb = batch_arrays(points, batch)
    o = OracleTree(meta.box_size, **kw)
    o.build_step(slots, sizes, nb, batch, nb)
    return o, points, meta


def test_update_alone_leaves_no_candidates():
    o, _, _ = built()
    res = o.collapse(100, 50_000)
    assert res == {"candidates": 0, "collapsed": 0, "points_moved": 0}


def test_collapse_undoes_deepening():
    o, points, meta = built()
    before = snapshot(o)
    s0 = o.stats()
    dp = o.deepen(5, 20_000)
    assert dp["splits"] == 5
    deepened = snapshot(o)
    s1 = o.stats()
    nodes1 = o.nodes()
    k1 = keyed(nodes1)
    # the deepened nodes: inner now, leaves before
    split_keys = [key for key, v in deepened[0].items() if not v[4] and before[0][key][4]]
    assert len(split_keys) == 5
    voxel_chunks = sum((int(nodes1[k1[key]]["num_voxels"]) + 999) // 1000 for key in split_keys)
    moved = sum(before[0][key][1] for key in split_keys)

    res = o.collapse(100, 50_000)
    assert res == {"candidates": 5, "collapsed": 5, "points_moved": moved}
    after = snapshot(o)
    assert_same_tree(after, before)
    s2 = o.stats()
    assert s2["allocated_bytes_persistent"] == s1["allocated_bytes_persistent"]
    assert s2["num_nodes"] == s0["num_nodes"] and s2["free_node_blocks"] == 5 and s2["pooled_grids"] == 5
    assert s2["num_allocated_chunks"] == chunk_closed_form(o) == s2["num_chunks_points"]
    assert s2["chunk_pool_size"] == s1["chunk_pool_size"] + voxel_chunks
    nodes2 = o.nodes()
    assert int(nodes2["free_slot"].sum()) == 40
    assert len(nodes2) == len(nodes1)  # the pool's high-water mark stays
    live = nodes2[nodes2["free_slot"] == 0]
    assert len(live) == 1 + 8 * int((live["first_child"] >= 0).sum())

    # deepening again takes the pooled blocks and grids
    dp2 = o.deepen(5, 20_000)
    assert dp2["splits"] == 5 and dp2["points_moved"] == dp["points_moved"]
    assert_same_tree(snapshot(o), deepened)
    s3 = o.stats()
    assert s3["free_node_blocks"] == 0 and s3["pooled_grids"] == 0
    assert s3["num_nodes"] == s1["num_nodes"] and len(o.nodes()) == len(nodes1)
    # grids and blocks came from the pools and point chunks from the queue: only
    # the voxel chunks (always new from the store, as in Update) cost bytes
    assert s3["allocated_bytes_persistent"] - s2["allocated_bytes_persistent"] == voxel_chunks * 16_032


def test_collapse_choice_sparsest_first_and_limits():
    o, _, _ = built()
    o.deepen(12, 12_000)
    nodes = o.nodes()
    k = keyed(nodes)
    cands = []
    for key, i in k.items():
        r = nodes[i]
        if r["first_child"] < 0 or key[0] == 0:
            continue
        kids = nodes[int(r["first_child"]) + np.arange(8)]
        if (kids["first_child"] >= 0).any() or int(kids["num_points"].sum()) > 50_000:
            continue
        cands.append((int(kids["num_points"].sum()), key))
    cands.sort()
    assert len(cands) == 12
    # a max_points below the sparsest candidate leaves nothing to do
    res = o.collapse(100, cands[0][0] - 1)
    assert res["candidates"] == 0 and res["collapsed"] == 0
    # three at most, and the point budget cuts the fourth's way
    budget = cands[0][0] + cands[1][0] + cands[2][0]
    res = o.collapse(4, 50_000, point_budget=budget)
    assert res["candidates"] == 12 and res["collapsed"] == 3
    after = keyed(o.nodes())
    for _, key in cands[:3]:
        assert o.nodes()[after[key]]["first_child"] < 0, key
    for _, key in cands[3:]:
        assert o.nodes()[after[key]]["first_child"] >= 0, key


@pytest.mark.parametrize("seed", [17, 23])
def test_growth_after_collapse_matches_closed_form(seed):
    """Half the cloud, deepen, collapse, the other half: the closed-form tree."""
    points, meta = lidar_scene(800_000, seed=seed)
    size = float(max(meta.box_size))
    batch = 100_000
    slots, sizes, nb = batch_arrays(points, batch)
    o = OracleTree(meta.box_size, batches_per_step=nb // 2)
    o.build_step(slots, sizes, nb, batch, nb)
    o.deepen(20, 8_000)
    res = o.collapse(20, 50_000)
    assert res["collapsed"] > 0
    o.build_step(slots, sizes, nb, batch, nb)
    assert o.stats()["batchlet_index"] == nb
    nodes = o.nodes()
    k = keyed(nodes)
    cf = CF.build(points, size)
    assert set(k) == set(cf)
    for key, i in k.items():
        r, e = nodes[i], cf[key]
        assert (r["first_child"] < 0) == e["leaf"], key
        if e["leaf"]:
            assert r["counter"] == r["num_points"] == e["count"], key
            if e["count"]:
                assert np.array_equal(point_rows(o.samples(i)), point_rows(points[e["indices"]])), key
        else:
            bits = np.nonzero(np.unpackbits(o.grid(i).view(np.uint8), bitorder="little"))[0]
            assert np.array_equal(bits, e["cells"]), key
            assert r["num_voxels"] == len(e["cells"]), key
    s = o.stats()
    assert s["num_allocated_chunks"] == chunk_closed_form(o)
    assert s["num_points"] == len(points)


def test_merged_leaf_accumulates_from_scratch():
    """Accumulate after a collapse: the merged leaf's fixed-point sums are those
    of its points in its own frame (numpy), not its children's."""
    o, points, meta = built()
    size = float(max(meta.box_size))
    o.accumulate(4)
    o.deepen(3, 20_000)
    o.accumulate(4)
    o.collapse(3, 50_000)
    o.accumulate(5)
    nodes, acc = o.nodes(), o.accums()
    k = keyed(nodes)
    checked = 0
    for key, i in k.items():
        r = nodes[i]
        if r["first_child"] >= 0 or r["num_points"] == 0:
            continue
        a = acc[i]
        assert a["count"] == r["num_points"], key
        pts = o.samples(i)
        L = key[0]
        inv = 1.0 / np.float64(np.float32(size))
        q = []
        for ax, c in zip(("x", "y", "z"), key[1:]):
            l = pts[ax].astype(np.float64) * inv * float(1 << L) - c
            l = np.clip(l, 0.0, 1.0)
            q.append(np.floor(l * 1048576.0).astype(np.int64) - 524288)
        assert [int(v) for v in a["q1"]] == [int(v.sum()) for v in q], key
        checked += 1
    assert checked > 10
