"""Seeded input cases shared by the oracle tests and the GPU parity tests."""
from __future__ import annotations

import numpy as np

from remobench_amd import lib as L


def make_meta(points: np.ndarray, box_size=None) -> L.CloudMeta:
    meta = L.CloudMeta()
    meta.num_points = len(points)
    if box_size is None:
        box_size = [float(points[a].max()) if len(points) else 1.0 for a in ("x", "y", "z")]
    for a in range(3):
        meta.box_size[a] = float(box_size[a])
        meta.translation[a] = 0.0
    return meta


def uniform_in_box(n, lo, hi, seed, color=0xFF808080):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, dtype=L.POINT_DTYPE)
    for a, ax in enumerate(("x", "y", "z")):
        p[ax] = rng.uniform(lo[a], hi[a], n).astype(np.float32)
    p["color"] = (rng.integers(0, 1 << 24, n, dtype=np.uint32) | np.uint32(0xFF000000)) if color is None else color
    return p


def tiny_cases():
    """name -> (points, CloudMeta, batch size). Sizes the oracle finishes in seconds."""
    cases = {}
    # one point at the origin
    p = np.zeros(1, dtype=L.POINT_DTYPE)
    p["color"] = 0xFF0000FF
    cases["single_point"] = (p, make_meta(p, [1.0, 1.0, 1.0]), 1)
    # exactly MAX_POINTS_PER_NODE in the root: no split
    p = uniform_in_box(50_000, (0, 0, 0), (1, 1, 1), 11, None)
    cases["at_capacity"] = (p, make_meta(p, [1.0, 1.0, 1.0]), 50_000)
    # one more: the root splits; all in one octant -> that child splits again
    p = uniform_in_box(50_001, (0, 0, 0), (0.5, 0.5, 0.5), 12, None)
    cases["one_over_one_octant"] = (p, make_meta(p, [1.0, 1.0, 1.0]), 50_001)
    # points exactly on boxMax quantise to 2^20 and wrap into child 0
    p = uniform_in_box(120_000, (0, 0, 0), (1, 1, 1), 13, None)
    p["x"][:3000] = 1.0
    p["y"][1000:4000] = 1.0
    p["z"][2000:5000] = 1.0
    cases["box_max_wrap"] = (p, make_meta(p, [1.0, 1.0, 1.0]), 40_000)
    # first batches hold <= 50,000 points: the root samples voxels while a leaf
    p = uniform_in_box(150_000, (0, 0, 0), (2, 1, 1), 14, None)
    cases["root_leaf_voxels"] = (p, make_meta(p, [2.0, 1.0, 1.0]), 30_000)
    # ragged last batch, deep skew (one tight cluster forces several levels in a batch)
    rng = np.random.default_rng(15)
    a = uniform_in_box(200_000, (0, 0, 0), (10, 10, 10), 16, None)
    b = uniform_in_box(230_017, (3.0, 3.0, 3.0), (3.02, 3.02, 3.02), 17, None)
    p = np.concatenate([a, b])
    p = p[rng.permutation(len(p))]
    cases["skewed_ragged"] = (p, make_meta(p, [10.0, 10.0, 10.0]), 100_000)
    return cases


def lidar_scene(n=2_097_152, seed=1, kind="lidar"):
    return L.generate_scene(n, seed, kind)
