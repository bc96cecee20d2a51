"""The captured Analysis + Rasterize graph (remo_frame) and the launch-shape
autotuner: every graph replay and every tuned shape gives the eager result."""
from __future__ import annotations

import numpy as np
import pytest

from remobench_amd import camera
from remobench_amd import lib as L
from remobench_amd.score_model import DEFAULT_MODEL, ScoreMLP

from .cases import lidar_scene
from .test_gpu_construct import build_hip

FIELDS = ("lambda", "surface_variation", "planarity", "linearity", "coverage", "score", "n", "state", "flags")


@pytest.fixture(scope="module")
def scene(gpu):
    points, meta = lidar_scene()
    tree, ring = build_hip(points, meta, 262_144, closure_batches=3)
    fb = L.DeviceBuffer(1920 * 1080 * 8)
    poses = camera.poses(points, meta.box_size)
    yield tree, fb, poses
    fb.free()


def _eager(tree, u, fb, mode):
    tree.analysis(u, mode)
    tree.render(u, fb.ptr)
    tree.sync()
    return tree.analysis_results(), fb.download(np.uint64, (1080 * 1920,)), tree.stats()["draw_samples"]


def _graph(tree, u, fb, mode):
    tree.frame(u, fb.ptr, mode)
    tree.sync()
    return tree.analysis_results(), fb.download(np.uint64, (1080 * 1920,)), tree.stats()["draw_samples"]


def _same(a, b):
    ra, fa, sa = a
    rb, fb_, sb = b
    for k in FIELDS:
        np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)
    assert np.array_equal(fa, fb_)
    assert sa == sb


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [L.REMO_ANALYSIS_ANALYTIC, L.REMO_ANALYSIS_LEARNED])
def test_graph_replays_match_eager(scene, mode):
    tree, fb, poses = scene
    if mode == L.REMO_ANALYSIS_LEARNED:
        if not DEFAULT_MODEL.exists():
            pytest.skip("no trained score model committed")
        tree.set_score_model(ScoreMLP.load(DEFAULT_MODEL))
    eager = {name: _eager(tree, u, fb, mode) for name, u in poses.items()}
    # replay the same graph for every pose, twice, with eager work in between
    for rnd in range(2):
        for name, u in poses.items():
            print(f"graph replay {rnd} pose {name} mode {mode}", flush=True)
            got = _graph(tree, u, fb, mode)
            _same(eager[name], got)
            _eager(tree, poses["street"], fb, L.REMO_ANALYSIS_ANALYTIC)  # eager work between replays


@pytest.mark.gpu
def test_autotune_keeps_results(scene):
    tree, fb, poses = scene
    u = poses["overview"]
    before = _eager(tree, u, fb, L.REMO_ANALYSIS_ANALYTIC)
    table = tree.autotune_frame(u, fb.ptr)
    assert len(table) == 27 and all(ms > 0 for *_, ms in table)
    best = min(table, key=lambda e: e[3])
    assert tree.frame_shape() == best[:3]
    _same(before, _eager(tree, u, fb, L.REMO_ANALYSIS_ANALYTIC))
    _same(before, _graph(tree, u, fb, L.REMO_ANALYSIS_ANALYTIC))
