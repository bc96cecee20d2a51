"""ctypes binding of the CPU oracle (oracle/remolod_oracle.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from remobench_amd.lib import ACCUM_DTYPE, ANALYSIS_DTYPE, NODE_DTYPE, POINT_DTYPE, SharedUniforms, Stats

ANALYSIS_OUT_DTYPE = np.dtype([
    ("r", ANALYSIS_DTYPE), ("lambda64", "<f8", 3), ("cov", "<f8", 6), ("mean", "<f8", 3),
    ("s1_hi", "<i8", 3), ("s1_lo", "<i8", 3), ("s2_hi", "<i8", 6), ("s2_lo", "<i8", 6),
    ("n", "<u8"), ("e", "<u4"), ("exact", "<u4"),
])

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(r.stdout.decode(errors="replace"))
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    sig = {
        "oracle_create": (C.c_void_p, [P(C.c_float), C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
        "oracle_destroy": (None, [C.c_void_p]),
        "oracle_set_threads": (None, [C.c_void_p, C.c_uint32]),
        "oracle_reset": (None, [C.c_void_p]),
        "oracle_build_step": (C.c_uint32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.c_uint32, C.c_uint32]),
        "oracle_accumulate": (None, [C.c_void_p, C.c_uint32]),
        "oracle_stats": (None, [C.c_void_p, P(Stats)]),
        "oracle_num_nodes": (C.c_uint32, [C.c_void_p]),
        "oracle_dump_nodes": (None, [C.c_void_p, C.c_void_p]),
        "oracle_node_samples": (C.c_uint32, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_uint32]),
        "oracle_grid": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
        "oracle_dump_accums": (None, [C.c_void_p, C.c_void_p]),
        "oracle_quantise": (None, [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]),
        "oracle_analysis_record_size": (C.c_uint32, []),
        "oracle_analyse": (None, [C.c_void_p, P(SharedUniforms), C.c_uint32, C.c_void_p]),
        "oracle_render": (None, [C.c_void_p, P(SharedUniforms), C.c_void_p, P(C.c_uint32), P(C.c_uint64)]),
        "oracle_jacobi": (None, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "oracle_morton": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "oracle_refine": (C.c_uint32, [C.c_void_p, C.c_uint32, P(C.c_uint32), P(C.c_uint32)]),
        "oracle_deepen": (None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, P(C.c_uint32), P(C.c_uint32),
                                 P(C.c_uint64), P(C.c_uint64)]),
        "oracle_collapse": (None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, P(C.c_uint32), P(C.c_uint32),
                                   P(C.c_uint64)]),
        "oracle_render_flat": (None, [C.c_void_p, C.c_uint64, P(SharedUniforms), C.c_void_p]),
        "oracle_apply_edl": (None, [P(SharedUniforms), C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    assert lib.oracle_analysis_record_size() == ANALYSIS_OUT_DTYPE.itemsize
    _lib = lib
    return lib


class OracleTree:
    """The reference's construct/accumulate path, restated on the CPU."""

    def __init__(self, box_size, max_nodes=200_000, persistent_bytes=1 << 62, batches_per_step=20, fixed_depth=0,
                 threads=1):
        """threads: threads for the per-point passes and the per-leaf folds; every
        result is the serial one."""
        self.lib = load()
        bs = (C.c_float * 3)(*[float(v) for v in box_size])
        self.h = self.lib.oracle_create(bs, max_nodes, persistent_bytes, batches_per_step, fixed_depth)
        if threads != 1:
            self.lib.oracle_set_threads(self.h, threads)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def build_step(self, slots: np.ndarray, batch_sizes: np.ndarray, num_slots: int,
                   slot_capacity: int, num_uploaded: int) -> int:
        slots = np.ascontiguousarray(slots, dtype=POINT_DTYPE)
        sizes = np.ascontiguousarray(batch_sizes, dtype=np.uint32)
        return self.lib.oracle_build_step(self.h, slots.ctypes.data, sizes.ctypes.data, num_slots,
                                          slot_capacity, num_uploaded)

    def accumulate(self, batch_index: int):
        self.lib.oracle_accumulate(self.h, batch_index)

    def stats(self) -> dict:
        s = Stats()
        self.lib.oracle_stats(self.h, C.byref(s))
        return s.as_dict()

    def nodes(self) -> np.ndarray:
        n = self.lib.oracle_num_nodes(self.h)
        out = np.zeros(n, dtype=NODE_DTYPE)
        self.lib.oracle_dump_nodes(self.h, out.ctypes.data)
        return out

    def samples(self, node: int, voxels: bool = False) -> np.ndarray:
        n = self.lib.oracle_node_samples(self.h, node, int(voxels), None, 0)
        out = np.zeros(n, dtype=POINT_DTYPE)
        if n:
            self.lib.oracle_node_samples(self.h, node, int(voxels), out.ctypes.data, n)
        return out

    def grid(self, node: int):
        out = np.zeros(65536, dtype=np.uint32)
        if self.lib.oracle_grid(self.h, node, out.ctypes.data) != 0:
            return None
        return out

    def accums(self) -> np.ndarray:
        n = self.lib.oracle_num_nodes(self.h)
        out = np.zeros(n, dtype=ACCUM_DTYPE)
        self.lib.oracle_dump_accums(self.h, out.ctypes.data)
        return out


def _analyse(self, uniforms, closure_batches: int = 8) -> np.ndarray:
    out = np.zeros(self.lib.oracle_num_nodes(self.h), dtype=ANALYSIS_OUT_DTYPE)
    self.lib.oracle_analyse(self.h, C.byref(uniforms), closure_batches, out.ctypes.data)
    return out


def _render(self, uniforms):
    w, h = int(uniforms.width), int(uniforms.height)
    fb = np.zeros(w * h, dtype=np.uint64)
    items = C.c_uint32(0)
    samples = C.c_uint64(0)
    self.lib.oracle_render(self.h, C.byref(uniforms), fb.ctypes.data, C.byref(items), C.byref(samples))
    return fb, items.value, samples.value


def _refine(self, min_level: int = 0):
    """Colour filtering of every gridded node at level >= min_level; returns
    (rewritten, skipped, empty cells)."""
    sk = C.c_uint32(0)
    em = C.c_uint32(0)
    n = self.lib.oracle_refine(self.h, int(min_level), C.byref(sk), C.byref(em))
    return n, sk.value, em.value


def _deepen(self, max_splits: int, min_points: int = 25_000, point_budget: int = 10_000_000) -> dict:
    """Refinement's local deepening (the HIP path's choice and split path);
    point_budget: the chosen leaves' points fit the spill list and voxel backlog
    (the HIP path's default capacities are 10,000,000 each)."""
    c, sp = C.c_uint32(0), C.c_uint32(0)
    mv, vx = C.c_uint64(0), C.c_uint64(0)
    self.lib.oracle_deepen(self.h, int(max_splits), int(min_points), int(point_budget), C.byref(c), C.byref(sp),
                           C.byref(mv), C.byref(vx))
    return {"candidates": c.value, "splits": sp.value, "points_moved": mv.value, "new_voxels": vx.value}


def _collapse(self, max_collapses: int, max_points: int = 50_000, point_budget: int = 10_000_000) -> dict:
    """Refinement's collapsing (the HIP path's choice and merge); point_budget:
    the merged nodes' points fit the spill list (the HIP path's default
    capacity is 10,000,000)."""
    c, n = C.c_uint32(0), C.c_uint32(0)
    mv = C.c_uint64(0)
    self.lib.oracle_collapse(self.h, int(max_collapses), int(max_points), int(point_budget), C.byref(c),
                             C.byref(n), C.byref(mv))
    return {"candidates": c.value, "collapsed": n.value, "points_moved": mv.value}


OracleTree.analyse = _analyse
OracleTree.collapse = _collapse
OracleTree.deepen = _deepen
OracleTree.refine = _refine
OracleTree.render = _render


def jacobi(a6):
    a = np.ascontiguousarray(a6, dtype=np.float64)
    ev64 = np.zeros(3, dtype=np.float64)
    ev32 = np.zeros(3, dtype=np.float32)
    load().oracle_jacobi(a.ctypes.data, ev64.ctypes.data, ev32.ctypes.data)
    return ev64, ev32


def quantise(point, size: float):
    p = np.zeros(1, dtype=POINT_DTYPE)
    p[0] = point
    q20 = np.zeros(3, dtype=np.uint32)
    q28 = np.zeros(3, dtype=np.uint32)
    load().oracle_quantise(p.ctypes.data, float(size), q20.ctypes.data, q28.ctypes.data)
    return q20, q28


def morton(x: int, y: int, z: int) -> int:
    return int(load().oracle_morton(x, y, z))


def render_flat(points: np.ndarray, uniforms) -> np.ndarray:
    """Every point splatted, no LOD (the flat pipeline's frame)."""
    pts = np.ascontiguousarray(points, dtype=POINT_DTYPE)
    fb = np.zeros(int(uniforms.width) * int(uniforms.height), dtype=np.uint64)
    load().oracle_render_flat(pts.ctypes.data, len(pts), C.byref(uniforms), fb.ctypes.data)
    return fb


def apply_edl(fb: np.ndarray, uniforms) -> np.ndarray:
    out = np.ascontiguousarray(fb, dtype=np.uint64).copy()
    load().oracle_apply_edl(C.byref(uniforms), out.ctypes.data)
    return out
