"""Closed-form octree of a point set -- an independent pin of the oracle.

TEST INFRASTRUCTURE ONLY (imported by tests/ alone).

After all points are in, the reference's tree depends only on the point set:
a node exists iff every ancestor's region holds more than 50,000 points
(MAX_POINTS_PER_NODE, kernels/remolod/remolod_structures.cuh:12), every split
creates all 8 children (kernels/remolod/remolod_octree.cu:323-347), a leaf's
counter and numPoints equal its region's point count, and an inner node's
occupied cells are exactly the 128^3 cells its region's points fall in
(remolod_octree.cu:95-140, :677-740).

This module derives that structure with numpy alone, from the quantisation
written out in float32 (remolod_octree.cu:163-169: X = u32(2^20 * x / size),
pX = u32(2^28 * x / size)) and the per-level bit walk with its wrap of a
coordinate equal to 2^20 into child 0 (:182-199). It shares no code with the
C++ transcription in remolod_oracle.cpp.
"""
from __future__ import annotations

import numpy as np

MAX_POINTS_PER_NODE = 50_000
MAX_DEPTH = 20
GRID = 128


def quantise(points: np.ndarray, size: float):
    """(q20, q28) per axis as uint64, computed with float32 IEEE arithmetic.

    numpy evaluates float32 * float32 and float32 / float32 with a single
    rounding each, which is the reference's unfused expression order.
    """
    size32 = np.float32(size)
    out20, out28 = [], []
    for axis in ("x", "y", "z"):
        d = points[axis].astype(np.float32) - np.float32(0.0)
        with np.errstate(invalid="ignore", over="ignore"):
            f20 = (np.float32(1048576.0) * d) / size32
            f28 = (np.float32(268435456.0) * d) / size32
        out20.append(_f2u_sat(f20))
        out28.append(_f2u_sat(f28))
    return np.stack(out20, 1), np.stack(out28, 1)


def _f2u_sat(f: np.ndarray) -> np.ndarray:
    f = f.astype(np.float64)
    out = np.zeros(f.shape, dtype=np.uint64)
    pos = f > 0
    big = f >= 4294967296.0
    ok = pos & ~big
    out[ok] = np.floor(f[ok]).astype(np.uint64)
    out[big] = 0xFFFFFFFF
    return out


def build(points: np.ndarray, size: float, with_cells: bool = True):
    """Closed-form tree.

    Returns a dict keyed by (level, X, Y, Z) with entries
    {"leaf": bool, "count": int, "cells": set or None, "indices": ndarray (leaves)}.
    """
    q20, q28 = quantise(points, size)
    # the descent reads bits 19..0; a coordinate of 2^20 wraps into child 0
    q20 = q20 & np.uint64((1 << MAX_DEPTH) - 1)
    q28 = q28 & np.uint64((1 << 28) - 1)
    nodes = {}
    n = len(points)
    # frontier: list of (level, key tuple, index array)
    frontier = [(0, (0, 0, 0), np.arange(n, dtype=np.int64))]
    while frontier:
        nxt = []
        for level, (X, Y, Z), idx in frontier:
            count = len(idx)
            split = count > MAX_POINTS_PER_NODE and level < MAX_DEPTH
            entry = {"leaf": not split, "count": count, "cells": None, "indices": None}
            if split:
                if with_cells:
                    sh = np.uint64(MAX_DEPTH + 1 - level)
                    c = (q28[idx] >> sh) & np.uint64(GRID - 1)
                    lin = c[:, 0] + c[:, 1] * GRID + c[:, 2] * GRID * GRID
                    entry["cells"] = np.unique(lin)
                sh = np.uint64(MAX_DEPTH - level - 1)
                bits = (q20[idx] >> sh) & np.uint64(1)
                child = (bits[:, 0] << np.uint64(2)) | (bits[:, 1] << np.uint64(1)) | bits[:, 2]
                order = np.argsort(child, kind="stable")
                sorted_child = child[order]
                bounds = np.searchsorted(sorted_child, np.arange(9, dtype=np.uint64))
                for ci in range(8):
                    sub = idx[order[bounds[ci]:bounds[ci + 1]]]
                    cx, cy, cz = (ci >> 2) & 1, (ci >> 1) & 1, ci & 1
                    nxt.append((level + 1, (2 * X + cx, 2 * Y + cy, 2 * Z + cz), sub))
            else:
                entry["indices"] = idx
            nodes[(level, X, Y, Z)] = entry
        frontier = nxt
    return nodes
