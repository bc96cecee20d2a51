"""GPU parity of the analytic-bucketing expand arm (fixed depth) against the oracle."""
from __future__ import annotations

import pytest

from .cases import lidar_scene, tiny_cases
from .helpers import compare_trees, stats_subset
from .test_gpu_construct import build_hip, build_oracle, compare_accums


def check(points, meta, batch, depth, check_colours=True, **kw):
    size = max(meta.box_size)
    tree, _ = build_hip(points, meta, batch, fixed_depth=depth, **kw)
    o = build_oracle(points, meta, batch, fixed_depth=depth, **kw)
    # the colour-membership check quantises the whole cloud once per inner node
    compare_trees(tree, o, size, points, check_colours=check_colours)
    s, so = tree.stats(), o.stats()
    assert stats_subset(s) == stats_subset(so)
    for k in ("total_expand_iters", "total_spilled_points", "total_nodes_split", "node_pool_overflow"):
        assert s[k] == so[k], (k, s[k], so[k])
    compare_accums(tree, o)
    assert s["sum_leaf_counts"] == s["num_points"] == len(points)
    assert s["mem_capacity_reached"] == 0
    return tree, o


@pytest.mark.gpu
@pytest.mark.parametrize("name,depth", [("single_point", 2), ("box_max_wrap", 3), ("root_leaf_voxels", 2),
                                        ("skewed_ragged", 3), ("one_over_one_octant", 4)])
def test_fixed_depth_tiny_parity(gpu, name, depth):
    points, meta, batch = tiny_cases()[name]
    print(f"fixed depth {depth}, case {name}", flush=True)
    check(points, meta, batch, depth)


@pytest.mark.gpu
@pytest.mark.parametrize("depth", [4, 5, 6])
def test_fixed_depth_lidar_2m_parity(gpu, depth):
    points, meta = lidar_scene()
    print(f"fixed depth {depth}, lidar 2M", flush=True)
    # deep fixed trees hold many more grids than the default 512 MiB store allows
    tree, o = check(points, meta, 262_144, depth, check_colours=depth < 6, persistent_bytes=4 << 30)
    s = tree.stats()
    assert s["total_expand_iters"] == 8 and s["total_spilled_points"] == 0
    assert s["max_depth"] == depth


@pytest.mark.gpu
def test_fixed_depth_under_node_pool_pressure(gpu):
    points, meta = lidar_scene(400_000, seed=9)
    tree, o = check(points, meta, 100_000, 5, max_nodes=200)
    assert tree.stats()["node_pool_overflow"] == 1


@pytest.mark.gpu
def test_fixed_depth_tree_refines_like_the_oracle(gpu):
    from .test_gpu_refine import clean, compare_refined
    points, meta = lidar_scene(2_097_152, seed=5)
    tree, o = check(points, meta, 262_144, 5)
    st = tree.refine()
    o.refine(0)
    assert clean(st) and st["complete"] == 1
    compare_refined(tree, o)
