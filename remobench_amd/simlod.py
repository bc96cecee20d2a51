"""The .simlod point file: a 24-byte header (six floats of a box that readers
ignore) followed by 16-byte records in remo_point layout (x, y, z as float32,
colour as uint32), so record i starts at byte 24 + 16 i.

The Python side of remobench_amd/host/SimlodReader.cpp and the C++
SimlodFileSource: the same failures and the same translation (the reference's
src/io/RawReader.cpp:14-138 and deriveTranslation, src/io/PointSource.cpp:360-366).

    read_header(path)      -> (num_points, header_box)
    observed_bounds(path)  -> (lo[3], hi[3]) in float64 over every record
    open_cloud(path)       -> (points, CloudMeta): translated so the box minimum is 0
    write(path, points)    -> a file holding `points` as they are
"""
from __future__ import annotations

import os

import numpy as np

from . import lib as L

HEADER_BYTES = 24
POINT_BYTES = 16
SCAN_RECORDS = 1 << 16


class SimlodError(ValueError):
    pass


def read_header(path) -> tuple[int, np.ndarray]:
    size = os.stat(path).st_size
    if size < HEADER_BYTES:
        raise SimlodError(f"file is smaller than a .simlod header: {path}")
    if (size - HEADER_BYTES) % POINT_BYTES:
        raise SimlodError(f"point data is not a whole number of 16-byte records (truncated?): {path}")
    n = (size - HEADER_BYTES) // POINT_BYTES
    if n == 0:
        raise SimlodError(f".simlod file holds no points: {path}")
    box = np.fromfile(path, dtype="<f4", count=6)
    return int(n), box


def records(path, n: int) -> np.ndarray:
    """The file's records, memory-mapped (read-only)."""
    return np.memmap(path, dtype=L.POINT_DTYPE, mode="r", offset=HEADER_BYTES, shape=(n,))


def observed_bounds(path) -> tuple[np.ndarray, np.ndarray]:
    n, _ = read_header(path)
    r = records(path, n)
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for s in range(0, n, SCAN_RECORDS):
        blk = r[s:s + SCAN_RECORDS]
        for k, ax in enumerate("xyz"):
            c = blk[ax]
            # NaN never moves a bound (comparisons with it are false)
            lo[k] = min(lo[k], float(np.fmin.reduce(c)))
            hi[k] = max(hi[k], float(np.fmax.reduce(c)))
    return lo, hi


def open_cloud(path) -> tuple[np.ndarray, L.CloudMeta]:
    """All records, each coordinate plus float32(-min) in float32 arithmetic,
    and the cloud's CloudMeta (translation -min, box size float32(max - min)
    computed as max + translation in double)."""
    n, _ = read_header(path)
    lo, hi = observed_bounds(path)
    meta = L.CloudMeta()
    meta.num_points = n
    for k in range(3):
        meta.translation[k] = -lo[k]
        meta.box_size[k] = np.float32(hi[k] + meta.translation[k])
    pts = np.array(records(path, n))
    t = np.array([meta.translation[k] for k in range(3)]).astype(np.float32)
    if np.any(t != 0):
        for k, ax in enumerate("xyz"):
            pts[ax] = pts[ax] + t[k]
    return pts, meta


def write(path, points: np.ndarray, header_box=None) -> None:
    points = np.ascontiguousarray(points, dtype=L.POINT_DTYPE)
    box = np.zeros(6, dtype="<f4") if header_box is None else np.asarray(header_box, dtype="<f4")
    with open(path, "wb") as f:
        f.write(box.tobytes())
        f.write(points.tobytes())
