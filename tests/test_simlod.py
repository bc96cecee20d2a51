"""The .simlod reader, C++ (remobench_amd/host/SimlodReader.cpp) and Python
(remobench_amd/simlod.py), against each other and against what the format and
the reference's translation fix (src/io/RawReader.cpp:14-138,
src/io/PointSource.cpp:360-366): record count from the file size, observed
bounds in double, translation -min, box size float32(max + translation),
coordinates translated by a float32 addition; and the failures."""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from remobench_amd import build as B
from remobench_amd import lib as L
from remobench_amd import simlod

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("simlod") / "simlod_check"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", str(ROOT / "include"), "-I", str(B.HOST),
           str(ROOT / "tests" / "native" / "simlod_check.cpp"), str(B.HOST / "SimlodReader.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True)
    return out


def run(checker, path, *rng):
    r = subprocess.run([str(checker), str(path), *map(str, rng)], stdout=subprocess.PIPE, timeout=60)
    return r.returncode, json.loads(r.stdout.decode().strip())


def fnv(points: np.ndarray) -> str:
    h = 1469598103934665603
    for b in np.ascontiguousarray(points).tobytes():
        h = ((h ^ b) * 1099511628211) & ((1 << 64) - 1)
    return f"{h:016x}"


def offset_cloud(n=5000, seed=3):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, dtype=L.POINT_DTYPE)
    p["x"] = (1000.25 + rng.random(n) * 37.5).astype(np.float32)
    p["y"] = (-50.5 + rng.random(n) * 12.0).astype(np.float32)
    p["z"] = (3.0 + rng.random(n) * 4.0).astype(np.float32)
    p["color"] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    return p


def test_header_bounds_and_translation(tmp_path, checker):
    p = offset_cloud()
    f = tmp_path / "c.simlod"
    simlod.write(f, p, header_box=[9, 9, 9, 1, 1, 1])  # a box no reader trusts
    n, box = simlod.read_header(f)
    assert n == len(p) and f.stat().st_size == 24 + 16 * n
    assert np.array_equal(box, np.float32([9, 9, 9, 1, 1, 1]))
    lo, hi = simlod.observed_bounds(f)
    for k, ax in enumerate("xyz"):
        assert lo[k] == float(p[ax].min()) and hi[k] == float(p[ax].max())
    pts, meta = simlod.open_cloud(f)
    for k, ax in enumerate("xyz"):
        t = -float(p[ax].min())
        assert meta.translation[k] == t
        assert meta.box_size[k] == np.float32(float(p[ax].max()) + t)
        # a float32 addition of float32(translation), record by record
        assert np.array_equal(pts[ax], (p[ax] + np.float32(t)).astype(np.float32))
    assert np.array_equal(pts["color"], p["color"])
    assert pts["x"].min() == 0.0 and pts["y"].min() == 0.0 and pts["z"].min() == 0.0
    # the C++ reader: the same count, bounds, box and translated records
    rc, out = run(checker, f)
    assert rc == 0, out
    assert out["num_points"] == n
    assert out["bounds"] == [float(v) for v in (*lo, *hi)]
    # printed with 9 significant digits: exact once read back as float32
    assert [np.float32(v) for v in out["box_size"]] == [np.float32(meta.box_size[k]) for k in range(3)]
    assert out["fnv"] == fnv(pts)
    # a range
    rc, out = run(checker, f, 1234, 777)
    assert rc == 0 and out["fnv"] == fnv(pts[1234:1234 + 777])


def test_untranslated_cloud_keeps_its_bytes(tmp_path, checker):
    """A cloud whose minimum is already 0 (the generated scenes) reads back as
    the same bytes and the same CloudMeta box."""
    p, meta = L.generate_scene(20_000, 5, "lidar")
    f = tmp_path / "s.simlod"
    simlod.write(f, p)
    pts, m2 = simlod.open_cloud(f)
    assert pts.tobytes() == p.tobytes()
    assert [m2.box_size[k] for k in range(3)] == [meta.box_size[k] for k in range(3)]
    rc, out = run(checker, f)
    assert rc == 0 and out["fnv"] == fnv(p)


@pytest.mark.parametrize("case", ["short", "ragged", "empty", "range"])
def test_failures(tmp_path, checker, case):
    f = tmp_path / "bad.simlod"
    p = offset_cloud(10)
    if case == "short":
        f.write_bytes(b"\0" * 20)
        msg = "smaller than a .simlod header"
    elif case == "ragged":
        simlod.write(f, p)
        f.write_bytes(f.read_bytes()[:-3])
        msg = "whole number of 16-byte records"
    elif case == "empty":
        simlod.write(f, p[:0])
        msg = "holds no points"
    else:
        simlod.write(f, p)
        rc, out = run(checker, f, 5, 6)
        assert rc == 1 and "past the end" in out["error"]
        return
    with pytest.raises(simlod.SimlodError, match=msg):
        simlod.read_header(f)
    rc, out = run(checker, f)
    assert rc == 1 and msg in out["error"]


@pytest.mark.gpu
def test_headless_builds_a_simlod_file_like_the_oracle(gpu, tmp_path):
    """An offset cloud in a .simlod file, streamed batch by batch from the file
    through a wrapping ring by the headless driver: the tree of the oracle built
    from the Python reader's points and box."""
    from .test_gpu_construct import build_oracle

    p, _ = L.generate_scene(600_000, 11, "lidar")
    for ax, off in (("x", 512.125), ("y", -64.0), ("z", 7.5)):
        p[ax] = (p[ax] + np.float32(off)).astype(np.float32)
    f = tmp_path / "cloud.simlod"
    simlod.write(f, p)
    slot = 10_000  # 60 batches through the pipeline's 50-slot ring: it wraps
    r = subprocess.run([str(B.HEADLESS), "--simlod", str(f), "--slot", str(slot), "--budget-ms", "-1",
                        "--batches-per-step", "2"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=240)
    assert r.returncode == 0, r.stderr.decode()
    out = json.loads(r.stdout.decode().strip().splitlines()[-1])
    pts, meta = simlod.open_cloud(f)
    assert out["points"] == len(p) and out["num_points_ingested"] == len(p)
    assert [np.float32(v) for v in out["box_size"]] == [np.float32(meta.box_size[k]) for k in range(3)]
    o = build_oracle(pts, meta, slot)
    s = o.stats()
    for k in ("num_nodes", "num_inner", "num_leaves", "num_points", "num_voxels", "max_depth"):
        assert out[k] == s[k], (k, out[k], s[k])
    assert out["sum_leaf_counts"] == len(p)
    o.close()
