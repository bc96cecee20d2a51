"""The allocation-free deepening (remo_deepen_noalloc) against remo_deepen,
through the C ABI: two calls in a row, the second with fewer candidates than
the first (so records of the first call are still in the shared array), each
giving the same choice, splits, counters, tree and accumulated sums, and
leaving free device memory unchanged."""
from __future__ import annotations

import ctypes as C

import pytest

from remobench_amd import lib as L

from .cases import lidar_scene
from .helpers import compare_trees, stats_subset
from .test_gpu_construct import build_hip, compare_accums


def _free_bytes():
    free, total = C.c_size_t(0), C.c_size_t(0)
    L.check(L.load().remo_device_mem_info(C.byref(free), C.byref(total)), "mem info")
    return free.value


@pytest.mark.gpu
def test_deepen_noalloc_matches_deepen(gpu):
    points, meta = lidar_scene(2_097_152, seed=11)
    size = max(meta.box_size)
    print("no-alloc deepening: two trees of the 2M scene", flush=True)
    a, ring_a = build_hip(points, meta, 262_144)
    b, ring_b = build_hip(points, meta, 262_144)
    prev = None
    for max_splits, min_points in [(6, 20_000), (2, 30_000)]:
        print(f"deepen({max_splits}, {min_points})", flush=True)
        before = _free_bytes()
        got = b.deepen(max_splits, min_points, noalloc=True)
        after = _free_bytes()
        want = a.deepen(max_splits, min_points)
        assert got == want, (got, want)
        assert got["splits"] == max_splits and got["refused"] == 0
        assert after == before
        if prev is not None:
            assert got["candidates"] < prev["candidates"]  # the second input is smaller
        prev = got
        compare_trees(a, b, size, points)
        sa, sb = a.stats(), b.stats()
        assert stats_subset(sa) == stats_subset(sb)
        for k in ("total_nodes_split", "total_new_voxels", "spill_overflow", "backlog_overflow"):
            assert sa[k] == sb[k], k
    a.accumulate(8)
    b.accumulate(8)
    compare_accums(a, b)
    for t in (a, b):
        t.close()
    ring_a.close()
    ring_b.close()
