"""ctypes binding of the C ABI in include/remolod.h.

Loads remobench_amd/_lib/libremolod.so (built in-tree by remobench_amd.build).
There is no fallback: if the library is missing or a call fails, this raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libremolod.so"

REMO_OK = 0
REMO_DONE = 1
ERRORS = {
    -1: "REMO_ERR_INVALID", -2: "REMO_ERR_OOM", -3: "REMO_ERR_DEVICE",
    -4: "REMO_ERR_RING", -5: "REMO_ERR_STATE", -6: "REMO_ERR_CAPACITY",
}
REMO_FLAG_NO_ACCUM = 1
REMO_ANALYSIS_ANALYTIC = 0
REMO_ANALYSIS_LEARNED = 1
MLP_INPUTS, MLP_HIDDEN, MLP_OUTPUTS, MLP_PARAMS = 13, 64, 2, 5186

POINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("color", "<u4")])


class RemoError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {ERRORS.get(code, code)}")
        self.code = code


class Point(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("color", C.c_uint32)]


class Mat4(C.Structure):
    _fields_ = [("rows", (C.c_float * 4) * 4)]


class Vec3f(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class SharedUniforms(C.Structure):
    _fields_ = [
        ("width", C.c_float), ("height", C.c_float), ("fovy_rad", C.c_float), ("time", C.c_float),
        ("transform", Mat4), ("view", Mat4), ("proj", Mat4),
        ("transform_frozen", Mat4), ("transform_frozen_inv", Mat4),
        ("box_min", Vec3f), ("box_max", Vec3f),
        ("frame_counter", C.c_uint64),
        ("lod_pixel_budget", C.c_float), ("min_node_size", C.c_float), ("lod_scale", C.c_float),
        ("point_size", C.c_int32), ("color_mode", C.c_int32), ("edl_strength", C.c_float),
        ("do_update_visibility", C.c_int32), ("use_high_quality_shading", C.c_int32),
        ("enable_edl", C.c_int32), ("show_bounding_box", C.c_int32), ("show_points", C.c_int32),
        ("pad0", C.c_int32),
    ]


class CloudMeta(C.Structure):
    _fields_ = [("num_points", C.c_uint64), ("translation", C.c_double * 3),
                ("box_size", C.c_float * 3), ("pad0", C.c_uint32)]


class BatchView(C.Structure):
    _fields_ = [("slots", C.c_void_p), ("batch_sizes", C.c_void_p),
                ("num_batches_uploaded", C.c_void_p), ("slot_capacity", C.c_uint32),
                ("num_slots", C.c_uint32), ("num_batches_total", C.c_uint32), ("pad0", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [("persistent_bytes", C.c_uint64), ("max_nodes", C.c_uint32),
                ("max_batches_per_step", C.c_uint32), ("step_budget_ms", C.c_float),
                ("slot_capacity", C.c_uint32), ("ring_slots", C.c_uint32),
                ("spill_capacity", C.c_uint32), ("backlog_capacity", C.c_uint32),
                ("flags", C.c_uint32), ("closure_batches", C.c_uint32), ("fixed_depth", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_uint32), ("num_inner", C.c_uint32), ("num_leaves", C.c_uint32),
        ("num_nonempty_leaves", C.c_uint32),
        ("num_points", C.c_uint64), ("num_voxels", C.c_uint64), ("num_points_processed", C.c_uint64),
        ("batchlet_index", C.c_uint32), ("num_chunks_points", C.c_uint32),
        ("num_chunks_voxels", C.c_uint32), ("max_points_per_node", C.c_uint32),
        ("counter_underflows", C.c_uint32), ("max_depth", C.c_uint32),
        ("num_allocated_chunks", C.c_uint64), ("chunk_pool_size", C.c_uint64),
        ("allocated_bytes_persistent", C.c_uint64), ("persistent_capacity", C.c_uint64),
        ("mem_capacity_reached", C.c_uint32), ("node_capacity_reached", C.c_uint32),
        ("node_pool_overflow", C.c_uint32), ("alloc_overflow", C.c_uint32),
        ("spill_overflow", C.c_uint32), ("backlog_overflow", C.c_uint32),
        ("dir_overflow", C.c_uint32), ("free_node_blocks", C.c_uint32),
        ("batches", C.c_uint64), ("expand_iters", C.c_uint64), ("nodes_split", C.c_uint64),
        ("spilled_points", C.c_uint64), ("new_voxels", C.c_uint64),
        ("total_batches", C.c_uint64), ("total_expand_iters", C.c_uint64),
        ("total_nodes_split", C.c_uint64), ("total_spilled_points", C.c_uint64),
        ("total_new_voxels", C.c_uint64),
        ("morton_watermark", C.c_uint64), ("num_accumulated", C.c_uint64),
        ("sum_leaf_counts", C.c_uint64),
        ("points_folded", C.c_uint32), ("leaves_folded", C.c_uint32),
        ("inner_with_sums", C.c_uint32), ("pooled_grids", C.c_uint32),
        ("draw_items", C.c_uint32), ("draw_list_overflow", C.c_uint32), ("draw_samples", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("pad")}


NODE_DTYPE = np.dtype([
    ("level", "<u4"), ("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("first_child", "<i4"),
    ("counter", "<u4"), ("num_points", "<u4"), ("num_voxels", "<u4"),
    ("num_voxels_stored", "<u4"), ("grid", "<i4"), ("free_slot", "<u4"), ("pad1", "<u4"),
])

ACCUM_DTYPE = np.dtype([
    ("q1", "<i8", 3), ("q2", "<i8", 6), ("g", "<f8", 9), ("c", "<f4", 6),
    ("count", "<u4"), ("last_touched_batch", "<u4"), ("state", "<u4"), ("score", "<f4"),
    ("morton_max", "<u8"),
])

ANALYSIS_DTYPE = np.dtype([
    ("lambda", "<f4", 3), ("surface_variation", "<f4"), ("planarity", "<f4"),
    ("linearity", "<f4"), ("coverage", "<f4"), ("score", "<f4"), ("n", "<u8"),
    ("state", "<u4"), ("flags", "<u4"),
])


ROLLUP_DTYPE = np.dtype([
    ("s1", "<i8", (3, 2)), ("s2", "<i8", (6, 2)), ("n", "<u8"), ("e", "<u4"), ("closed", "<u4"),
])


def int128(pairs: np.ndarray) -> list:
    """[..., 2] (low, high) int64 halves -> Python ints (two's complement 128-bit)."""
    a = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    return [(int(hi) << 64) | (int(lo) & 0xFFFFFFFFFFFFFFFF) for lo, hi in a]


SCOPES = ["remolod.construct.count", "remolod.construct.split", "remolod.construct.voxelSampling",
          "remolod.construct.allocChunks", "remolod.construct.insertPoints",
          "remolod.construct.insertVoxels", "remolod.accumulate", "remolod.analysis", "remolod.render",
          "remolod.frame", "remolod.refine", "flat.render"]
PROFILE_ALL = 0xFFFFFFFF


class ProfileEntry(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("total_ms", C.c_double), ("launches", C.c_uint64)]


class TuneEntry(C.Structure):
    _fields_ = [("node_blocks", C.c_uint32), ("raster_blocks", C.c_uint32), ("raster_tile", C.c_uint32),
                ("ms", C.c_float)]


class RefineStats(C.Structure):
    """remo_refine_stats (Refinement, colour filtering)."""
    _fields_ = [("levels_pending", C.c_uint32), ("levels_processed", C.c_uint32),
                ("nodes_dirty", C.c_uint32), ("nodes_refined", C.c_uint32), ("nodes_skipped", C.c_uint32),
                ("complete", C.c_uint32), ("points_folded", C.c_uint64), ("child_sums", C.c_uint64),
                ("voxels_written", C.c_uint64), ("missing_cells", C.c_uint32), ("empty_cells", C.c_uint32),
                ("arena_full", C.c_uint32), ("cell_overflow", C.c_uint32), ("estimate_ms", C.c_float),
                ("device_ms", C.c_float)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class DeepenStats(C.Structure):
    """remo_deepen_stats (Refinement, local deepening)."""
    _fields_ = [("candidates", C.c_uint32), ("chosen", C.c_uint32), ("splits", C.c_uint32),
                ("refused", C.c_uint32), ("points_moved", C.c_uint64), ("new_voxels", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class CollapseStats(C.Structure):
    """remo_collapse_stats (Refinement, collapsing)."""
    _fields_ = [("candidates", C.c_uint32), ("collapsed", C.c_uint32), ("points_moved", C.c_uint64),
                ("grids_pooled", C.c_uint32), ("voxels_dropped", C.c_uint32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class SceneSpec(C.Structure):
    _fields_ = [("extent", C.c_double), ("num_boxes", C.c_uint32), ("num_cylinders", C.c_uint32),
                ("cluster_fraction", C.c_double), ("cluster_volume_fraction", C.c_double),
                ("cluster_sigma", C.c_double), ("ground_sigma", C.c_double),
                ("scanline", C.c_double), ("shuffle_window", C.c_uint32), ("threads", C.c_uint32)]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m remobench_amd.build` "
                           "(the HIP path has no fallback)")
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    sig = {
        "remo_create": (C.c_int, [P(Config), P(CloudMeta), P(C.c_void_p)]),
        "remo_destroy": (None, [C.c_void_p]),
        "remo_reset": (C.c_int, [C.c_void_p]),
        "remo_build_step": (C.c_int, [C.c_void_p, P(BatchView), P(C.c_uint32)]),
        "remo_accumulate": (C.c_int, [C.c_void_p, C.c_uint32]),
        "remo_analysis": (C.c_int, [C.c_void_p, P(SharedUniforms), C.c_int]),
        "remo_render": (C.c_int, [C.c_void_p, P(SharedUniforms), C.c_void_p]),
        "remo_get_stats": (C.c_int, [C.c_void_p, P(Stats)]),
        "remo_request_stats": (C.c_int, [C.c_void_p]),
        "remo_collect_stats": (C.c_int, [C.c_void_p, P(Stats)]),
        "remo_get_render_diagnostics": (C.c_int, [C.c_void_p, P(C.c_uint32), P(C.c_uint64), P(C.c_uint32)]),
        "remo_last_error": (C.c_char_p, [C.c_void_p]),
        "remo_sync": (C.c_int, [C.c_void_p]),
        "remo_dump_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_dump_node_samples": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                             C.c_uint32, P(C.c_uint32)]),
        "remo_dump_grid": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
        "remo_dump_accums": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_dump_analysis": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_dump_rollup": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_ring_create": (C.c_int, [C.c_uint32, C.c_uint32, P(C.c_void_p)]),
        "remo_ring_destroy": (None, [C.c_void_p]),
        "remo_ring_upload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
        "remo_ring_view": (C.c_int, [C.c_void_p, C.c_uint32, P(BatchView)]),
        "remo_ring_rewind": (C.c_int, [C.c_void_p]),
        "remo_ring_push": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]),
        "remo_ring_order_before": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
        "remo_ring_sync": (C.c_int, [C.c_void_p]),
        "remo_host_alloc": (C.c_int, [C.c_size_t, P(C.c_void_p)]),
        "remo_host_free": (C.c_int, [C.c_void_p]),
        "remo_dev_alloc": (C.c_int, [C.c_size_t, P(C.c_void_p)]),
        "remo_dev_free": (C.c_int, [C.c_void_p]),
        "remo_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "remo_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "remo_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "remo_atomic_rate": (C.c_int, [C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint32, P(C.c_double)]),
        "remo_device_count": (C.c_int, [P(C.c_int)]),
        "remo_device_mem_info": (C.c_int, [P(C.c_size_t), P(C.c_size_t)]),
        "remo_profile_enable": (C.c_int, [C.c_void_p, C.c_uint32]),
        "remo_profile_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_profile_clear": (C.c_int, [C.c_void_p]),
        "remo_mlp_create": (C.c_int, [C.c_uint64, C.c_uint32, P(C.c_void_p)]),
        "remo_mlp_destroy": (None, [C.c_void_p]),
        "remo_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
        "remo_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
        "remo_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
        "remo_mlp_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_float),
                                        C.c_void_p]),
        "remo_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                          P(C.c_float)]),
        "remo_set_score_model": (C.c_int, [C.c_void_p, C.c_void_p]),
        "remo_set_rollup_limits": (C.c_int, [C.c_void_p, C.c_uint32]),
        "remo_accum_fp32_atomics": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_frame": (C.c_int, [C.c_void_p, P(SharedUniforms), C.c_int, C.c_void_p]),
        "remo_autotune_frame": (C.c_int, [C.c_void_p, P(SharedUniforms), C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_uint32, P(C.c_uint32)]),
        "remo_get_frame_shape": (C.c_int, [C.c_void_p, P(TuneEntry)]),
        "remo_refine": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32, P(RefineStats)]),
        "remo_render_flat": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, P(SharedUniforms), C.c_void_p]),
        "remo_deepen": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(DeepenStats)]),
        "remo_deepen_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(DeepenStats)]),
        "remo_collapse": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(CollapseStats)]),
        "remo_apply_edl": (C.c_int, [C.c_void_p, P(SharedUniforms), C.c_void_p]),
        "remo_mlp_features": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32)]),
        "remo_scene_default": (None, [P(SceneSpec)]),
        "remo_scene_generate": (C.c_int, [P(SceneSpec), C.c_uint64, C.c_uint64, C.c_void_p,
                                          P(CloudMeta)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, what: str, ctx=None) -> int:
    if code < 0:
        detail = ""
        if ctx:
            msg = load().remo_last_error(ctx)
            detail = f" ({msg.decode(errors='replace')})" if msg else ""
        raise RemoError(code, what + detail)
    return code


# ---- synthetic scenes ---------------------------------------------------------

SCENES = {
    # ground 40x40 m, 20 boxes, 50 cylinders (the 2M-point and 100M-point scenes)
    "lidar": dict(extent=40.0, num_boxes=20, num_cylinders=50, cluster_fraction=0.0),
    # 70% of the points in Gaussian clusters covering 5% of the volume
    "skewed": dict(extent=40.0, num_boxes=20, num_cylinders=50, cluster_fraction=0.7),
    # the same generator tiled over 400x400 m
    "city": dict(extent=400.0, num_boxes=2000, num_cylinders=5000, cluster_fraction=0.0),
}


def scene_spec(kind: str = "lidar", **overrides) -> SceneSpec:
    spec = SceneSpec()
    load().remo_scene_default(C.byref(spec))
    for k, v in {**SCENES[kind], **overrides}.items():
        setattr(spec, k, v)
    return spec


def generate_scene(num_points: int, seed: int, kind: str = "lidar", threads: int = 0, **overrides):
    """Returns (points structured array, CloudMeta)."""
    spec = scene_spec(kind, threads=threads, **overrides)
    pts = np.zeros(num_points, dtype=POINT_DTYPE)
    meta = CloudMeta()
    check(load().remo_scene_generate(C.byref(spec), seed, num_points,
                                     pts.ctypes.data_as(C.c_void_p), C.byref(meta)), "scene")
    return pts, meta


class DeviceBuffer:
    """A device allocation from remo_dev_alloc with numpy upload / download."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().remo_dev_alloc(max(16, self.nbytes), C.byref(p)), "device alloc")
        self.ptr = p.value

    @classmethod
    def of(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(load().remo_memcpy_h2d(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes), "upload")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            check(load().remo_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes),
                  "download")
        return out

    def free(self):
        if self.ptr:
            load().remo_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int(0)
    load().remo_device_count(C.byref(n))
    return n.value
