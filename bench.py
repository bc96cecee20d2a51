#!/usr/bin/env python3
"""Ingest benchmark of the HIP RemoLOD path on one MI355X.

A step is one pass of the hot path over one batch: Update (the octree build of
one 1,000,000-point batch), Accumulate and Analysis. The workload is the
100,000,000-point seed-generated lidar-like scene (ground 40 x 40 m with 2 cm
noise, 20 boxes, 50 cylinders, scan-line order, seed 2) in 100 batches of
1,000,000, resident in HBM before timing starts. `value` is ingested points per
second over the K timed steps (steps run from batch 0 of a fresh tree).

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "M pts/s ingest (Update+Accumulate+Analysis), plain vs learned score; % HBM BW"
HBM_PEAK_GBS = 8000.0


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(scope: str, t: dict) -> float:
    """Compulsory HBM bytes of one scope over the timed steps.

    t: totals over the timed region -- points (batch points ingested), spilled,
    voxels (new voxels), splits, rounds (count launches), folded (points folded by
    Accumulate), nodes (tree size at the end).
    """
    pts, sp, vox = t["points"], t["spilled"], t["voxels"]
    if scope == "remolod.construct.count":
        # round 0 reads every batch point once; later rounds read batch + spill again
        extra = max(0, t["rounds"] - t["steps"])
        return 16.0 * (pts + extra * (pts / max(1, t["steps"]) + sp / max(1, t["steps"])))
    if scope == "remolod.construct.split":
        return 262144.0 * t["splits"] + 32.0 * sp
    if scope == "remolod.construct.voxelSampling":
        return 16.0 * (pts + sp) + 20.0 * vox
    if scope == "remolod.construct.insertPoints":
        return 32.0 * (pts + sp)
    if scope == "remolod.construct.insertVoxels":
        return 36.0 * vox
    if scope == "remolod.construct.allocChunks":
        return 64.0 * t["nodes"] * t["steps"]
    if scope == "remolod.accumulate":
        return 16.0 * t["folded"] + 192.0 * t["nodes"] * t["steps"]
    if scope == "remolod.analysis":
        # the finalize of every node: its cached shape, closing batch and coordinates
        # in (56 B), its result, closure flag and accumulator state out (57 B); the
        # incremental roll-up adds some 0.1 MB a step (a few hundred changed records)
        return 113.0 * t["nodes"] * t["steps"]
    return 0.0


def cpu_baseline(points, meta, batch: int, sample_batches: int) -> dict:
    """The CPU oracle (a transcription of the reference kernels) on the first
    `sample_batches` batches of the same workload: on one core, and threaded
    over the cores this process may use (its per-point passes and per-leaf
    folds; same results)."""
    import numpy as np

    from oracle.oracle import OracleTree

    n = min(len(points), sample_batches * batch)
    nb = (n + batch - 1) // batch
    sizes = np.array([min(batch, n - i * batch) for i in range(nb)], dtype=np.uint32)
    from remobench_amd import lib as L

    slots = points[:n]
    uniforms = L.SharedUniforms()

    def timed(threads: int) -> float:
        o = OracleTree(meta.box_size, batches_per_step=1, threads=threads)
        t0 = time.perf_counter()
        done = 0
        while done < nb:
            done = o.build_step(slots, sizes, nb, batch, nb)
            o.accumulate(done)
            o.analyse(uniforms)
        dt = time.perf_counter() - t0
        o.close()
        return dt

    cores = min(16, len(os.sched_getaffinity(0)))
    dt1 = timed(1)
    dtn = timed(cores)
    sample = f"first {nb} of the workload's batches ({n:,} points), one step per batch (Update+Accumulate+Analysis)"
    return {"value": round(n / dt1 / 1e6, 3), "unit": "M pts/s", "cores": 1, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{sample}, serial C++ oracle, {dt1:.1f} s",
            "all_cores": {"value": round(n / dtn / 1e6, 3), "unit": "M pts/s", "cores": cores, "kind": "port",
                          "sample": f"{sample}, the oracle threaded over {cores} cores "
                                    f"(descents, voxel sampling, leaf folds; Analysis serial), {dtn:.1f} s"}}


# Kernels behind each timing scope (a scope launch may run several kernels).
SCOPE_KERNELS = {
    "remolod.construct.count": ["kCount"],
    "remolod.construct.split": ["kSplitDecide", "kSplitApply", "kSpillAndZero"],
    "remolod.construct.voxelSampling": ["kVoxelSample"],
    "remolod.construct.allocChunks": ["kAllocChunks"],
    "remolod.construct.insertPoints": ["kInsertPoints"],
    "remolod.construct.insertVoxels": ["kInsertVoxels"],
    "remolod.accumulate": ["kAccumPlan", "kAccumTiles", "kAccumFinish"],
    "remolod.analysis": ["kAnalysisStep", "kFinalizeLearned", "kAnalysisLeaves", "kAnalysisLevel", "kAnalysisLevel2",
                         "kShapes", "kSetFrame"],
}
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_lidar100M.txt"


def measured_traffic(scope: str, workload: str):
    """HBM bytes per launch of `scope` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, calibrated against a copy kernel of
    known size), when it was measured on this same workload; else None."""
    if not TRAFFIC_FILE.exists():
        return None
    rows = {}
    wl = None
    for line in TRAFFIC_FILE.read_text().splitlines():
        f = line.split()
        if len(f) >= 2 and f[0] == "workload":
            wl = line.split(None, 1)[1].strip()
        elif len(f) == 4 and f[0] == "scope":
            rows[f[1]] = (float(f[2]), float(f[3]))
    if wl != workload or scope not in rows:
        return None
    return rows[scope][0] + rows[scope][1]


ATOMICS_FILE = ROOT / "profiles" / "atomics_lidar100M.txt"


def measured_atomics(kernels, workload: str):
    """Memory-side atomics per dispatch (rocprofv3 TCC_ATOMIC_sum) of the first of
    `kernels` in the committed summary, when measured on this same workload."""
    if not ATOMICS_FILE.exists():
        return None
    wl, rows = None, {}
    for line in ATOMICS_FILE.read_text().splitlines():
        f = line.split()
        if len(f) >= 2 and f[0] == "workload":
            wl = line.split(None, 1)[1].strip()
        elif len(f) >= 3 and f[0] == "kernel":
            rows[f[1]] = float(f[2])
    if wl != workload:
        return None
    for k in kernels:
        if k in rows:
            return {"kernel": k, "tcc_atomic": rows[k]}
    return None


def atomic_rates(L, mib: int = 256, per_lane: int = 64):
    """Device-scope atomicOr rates of the microbenchmark kernel (remo_atomic_rate)
    on a `mib` MiB buffer: scattered returning (the peak the voxel climb's
    atomics are held against), scattered non-returning, and one word per wave."""
    import ctypes as C
    lib = L.load()
    buf = C.c_void_p()
    if lib.remo_dev_alloc(mib << 20, C.byref(buf)) != 0:
        return None
    out = {"buffer_mib": mib, "atomics_per_launch": 4096 * 256 * per_lane}
    try:
        for mode, name in ((0, "scattered_returning"), (1, "scattered_nonreturning"), (2, "one_word_per_wave")):
            r = C.c_double(0.0)
            L.check(lib.remo_atomic_rate(buf, mib << 20, mode, per_lane if mode < 2 else 4, C.byref(r)),
                    "atomic rate")
            out[f"{name}_g_per_s"] = round(r.value / 1e9, 3)
    finally:
        lib.remo_dev_free(buf)
    return out


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine() or "host CPU"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=100_000_000)
    ap.add_argument("--batch", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--scene", default="lidar")
    ap.add_argument("--simlod", default="", help="bench a .simlod point file instead of a generated scene")
    ap.add_argument("--cpu-sample-batches", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--raster", default="1920x1080", help="framebuffer size of the Rasterize timing, WxH")
    ap.add_argument("--no-streaming", action="store_true",
                    help="skip the PCIe-inclusive run (cloud in pinned host memory, 50-slot ring)")
    ap.add_argument("--analysis", choices=["analytic", "learned", "off"], default="analytic")
    ap.add_argument("--refine-budget-ms", type=float, default=2.0,
                    help="Refinement budget per step of the interleaved refinement pass")
    ap.add_argument("--no-refine", action="store_true", help="skip the Refinement passes")
    ap.add_argument("--fixed-depth", type=int, default=0,
                    help="the analytic-bucketing expand arm at this depth (0: the iterative expand)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run remo_copy_kernel over this many MiB 5 times (the known-size "
                         "reference for HBM byte counters) and exit")
    ap.add_argument("--atomic-microbench", action="store_true",
                    help="only run the atomicOr microbenchmark (remo_atomic_kernel) and exit")
    args = ap.parse_args()

    if args.atomic_microbench:
        from remobench_amd import lib as L
        print(json.dumps({"atomic_microbench": atomic_rates(L)}), flush=True)
        return 0

    if args.calibrate_copy_mib:
        import ctypes as C

        from remobench_amd import lib as L
        n = args.calibrate_copy_mib << 20
        src, dst = C.c_void_p(), C.c_void_p()
        L.check(L.load().remo_dev_alloc(n, C.byref(src)), "alloc")
        L.check(L.load().remo_dev_alloc(n, C.byref(dst)), "alloc")
        for _ in range(5):
            L.check(L.load().remo_dev_copy(dst, src, n), "copy")
        print(json.dumps({"calibration": "remo_copy_kernel", "bytes_read": n, "bytes_written": n,
                          "launches": 5}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        log("bench.py measures one GPU: there is no multi-GPU path (several GPUs would each own the "
            "subtrees below a fixed octree level by Morton prefix); run with --gpus 1")
        return 2

    from remobench_amd import lib as L
    from remobench_amd.pipeline import PointRing, Tree

    if L.device_count() < 1:
        log("no GPU visible")
        return 1

    t = time.perf_counter()
    threads = min(16, os.cpu_count() or 1)
    if args.simlod:
        from remobench_amd import simlod
        points, meta = simlod.open_cloud(args.simlod)
        args.scene = f"simlod:{Path(args.simlod).name}"
    else:
        points, meta = L.generate_scene(args.points, args.seed, args.scene, threads=threads)
    nb = (len(points) + args.batch - 1) // args.batch
    log(f"scene: {len(points):,} points, {nb} batches, box {list(meta.box_size)} "
        f"({time.perf_counter() - t:.1f} s)")
    steps = args.steps
    if steps > nb:
        log(f"--steps {steps} > {nb} batches in the scene: timing {nb}")
        steps = nb

    ring = PointRing(args.batch, nb)
    ring.set_cloud(points, args.batch)
    ring.set_batches_consumed(0)
    t_up = time.perf_counter()
    ring.pump()
    upload_s = time.perf_counter() - t_up
    view = ring.view()
    tree = Tree(meta, slot_capacity=args.batch, ring_slots=nb, batches_per_step=1,
                step_budget_ms=-1.0, fixed_depth=args.fixed_depth)
    log(f"uploaded; tree context created ({time.perf_counter() - t:.1f} s)")

    uniforms = L.SharedUniforms()
    analysis_mode = {"analytic": L.REMO_ANALYSIS_ANALYTIC, "learned": L.REMO_ANALYSIS_LEARNED}.get(
        args.analysis)
    analysis_ok = analysis_mode is not None
    # the learned score, when a trained model is committed beside the package
    from remobench_amd.score_model import DEFAULT_MODEL, ScoreMLP
    model = None
    if DEFAULT_MODEL.exists():
        model = ScoreMLP.load(DEFAULT_MODEL)
        tree.set_score_model(model)
        log(f"learned score model: {DEFAULT_MODEL.relative_to(ROOT)}")

    def step():
        nonlocal analysis_ok, analysis_mode
        _, consumed = tree.build_step(view)
        tree.accumulate(consumed)
        if analysis_ok:
            try:
                tree.analysis(uniforms, analysis_mode)
            except L.RemoError as e:
                if e.code != -5:
                    raise
                analysis_ok = False
        return consumed

    for i in range(min(args.warmup, nb)):
        step()
    tree.sync()

    def run(scopes):
        tree.reset()
        tree.profile_enable(True, scopes)
        tree.profile_clear()
        tree.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        tree.sync()
        return time.perf_counter() - t0

    # breakdown pass: every scope timed (its event pairs slow the steps a little)
    dt_all = run(None)
    prof = tree.profile()
    dominant_name = max(prof, key=lambda k: prof[k][0])
    # timed pass: only the dominant kernel carries HIP events
    dt = run([dominant_name])
    prof_dom = tree.profile()
    tree.profile_enable(False)
    st = tree.stats()
    learned = None
    if model is not None and analysis_ok and analysis_mode == L.REMO_ANALYSIS_ANALYTIC:
        # the same steps with the learned score (the metric's "plain vs learned")
        analysis_mode = L.REMO_ANALYSIS_LEARNED
        dt_learned = run(["remolod.analysis"])
        ms_l, n_l = tree.profile()["remolod.analysis"]
        analysis_mode = L.REMO_ANALYSIS_ANALYTIC
        tree.profile_enable(False)
        ms_a, n_a = prof["remolod.analysis"]
        learned = {"value": round(tree.stats()["num_points_processed"] / dt_learned / 1e6, 2),
                   "ms_per_step": round(dt_learned * 1e3 / steps, 4),
                   "analysis_avg_us": round(1e3 * ms_l / max(1, n_l), 2),
                   "analysis_avg_us_analytic": round(1e3 * ms_a / max(1, n_a), 2),
                   "analysis_speedup": round((ms_a / max(1, n_a)) / max(1e-9, ms_l / max(1, n_l)), 3),
                   "model": str(DEFAULT_MODEL.relative_to(ROOT))}
        log(f"learned score: {learned}")
    # The same steps in the sequence the C++ plugin's build runs per frame at one
    # batch per step (RemolodHipPipeline::build, host/RemolodHipPipeline.cpp:110-175:
    # build step, Accumulate, the stats request (the per-node stats sweep the C-ABI
    # step leaves out, and the read of the counters), Analysis, then the stats)
    plugin_path = None
    if analysis_ok and analysis_mode == L.REMO_ANALYSIS_ANALYTIC:
        import ctypes as C
        raw = L.Stats()

        def plugin_step():
            _, consumed = tree.build_step(view)
            tree.accumulate(consumed)
            tree._ck(tree.lib.remo_request_stats(tree.h), "stats request")
            tree.analysis(uniforms, analysis_mode)
            tree._ck(tree.lib.remo_collect_stats(tree.h, C.byref(raw)), "stats")

        tree.reset()
        tree.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            plugin_step()
        tree.sync()
        dt_plugin = time.perf_counter() - t0
        plugin_path = {"ms_per_step": round(dt_plugin * 1e3 / steps, 4),
                       "vs_c_abi_step": round(dt_plugin / dt, 4),
                       "sequence": "remo_build_step, remo_accumulate, remo_request_stats, remo_analysis, "
                                   "remo_collect_stats (RemolodHipPipeline::build)"}
        log(f"plugin build path: {plugin_path}")
    ingested = st["num_points_processed"]
    value = ingested / dt / 1e6
    log(f"timed {steps} steps: {dt * 1e3:.1f} ms, {value:.1f} M pts/s ({ingested / dt_all / 1e6:.1f} with "
        f"every scope timed); nodes {st['num_nodes']}, voxels {st['num_voxels']:,}")

    totals = {"points": ingested, "spilled": st["total_spilled_points"],
              "voxels": st["total_new_voxels"], "splits": st["total_nodes_split"],
              "rounds": st["total_expand_iters"], "folded": st["num_accumulated"],
              "nodes": st["num_nodes"], "steps": steps}
    scopes = {}
    for name, (ms, launches) in prof.items():
        if launches:
            scopes[name] = {"ms": round(ms, 3), "launches": launches,
                            "avg_us": round(1e3 * ms / launches, 2),
                            "algorithmic_bytes": algorithmic_bytes(name, totals)}
    dominant = dominant_name if dominant_name in scopes else None
    workload = (f"{args.scene}-{len(points) // 1_000_000}M: {len(points):,} points in {nb} "
                f"batches of {args.batch:,}; Update+Accumulate"
                + ("+Analysis (analytic)" if analysis_ok and args.analysis == "analytic" else
                   "+Analysis (learned)" if analysis_ok else "")
                + (f"; fixed-depth expand D={args.fixed_depth}" if args.fixed_depth else ""))
    roofline = None
    if dominant:
        d = scopes[dominant]
        ms_dom, launches_dom = prof_dom[dominant]
        avg_us = 1e3 * ms_dom / max(1, launches_dom)  # from the timed pass
        per_launch = d["algorithmic_bytes"] / d["launches"]
        achieved = per_launch / (avg_us * 1e-6) / 1e9
        traffic = measured_traffic(dominant, workload)
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": round(traffic) if traffic else None, "bytes_per_launch": round(per_launch),
                    "avg_launch_us": round(avg_us, 2),
                    "kernels": SCOPE_KERNELS.get(dominant, [])}
        if traffic:
            roofline["traffic_source"] = f"{TRAFFIC_FILE.relative_to(ROOT)} (rocprofv3 --pmc)"
    # The path's compulsory HBM bytes per point: read the input, store it in its
    # leaf, store the new voxels, Accumulate's re-read, zero the new grids
    # (16 + 16 + 16 v + 16 + 262,144 s, v new voxels and s splits per point over
    # the timed steps). The per-scope figure above also counts re-reads (count
    # rounds, spill re-descents, the backlog) and is reported beside it.
    v_pp = totals["voxels"] / max(1, ingested)
    s_pp = totals["splits"] / max(1, ingested)
    compulsory_pp = 48.0 + 16.0 * v_pp + 262144.0 * s_pp
    path_gbs = compulsory_pp * ingested / dt / 1e9
    scopes_bytes = sum(s["algorithmic_bytes"] for s in scopes.values())
    scopes_gbs = scopes_bytes / dt / 1e9

    # Atomic rate of the dominant kernel (memory-side atomics per dispatch from the
    # committed rocprofv3 TCC_ATOMIC summary over its live average launch time)
    # against the microbenchmark kernel's scattered atomicOr rate, measured here.
    atomics = atomic_rates(L)
    if roofline and atomics:
        per = measured_atomics(roofline["kernels"], workload)
        if per:
            atomics["kernel"] = per["kernel"]
            atomics["tcc_atomic_per_launch"] = per["tcc_atomic"]
            atomics["achieved_g_per_s"] = round(per["tcc_atomic"] / (roofline["avg_launch_us"] * 1e-6) / 1e9, 3)
            atomics["frac_of_scattered_returning"] = round(
                atomics["achieved_g_per_s"] / atomics["scattered_returning_g_per_s"], 4)
            atomics["source"] = f"{ATOMICS_FILE.relative_to(ROOT)} (rocprofv3 --pmc TCC_ATOMIC_sum)"
    log(f"atomics: {atomics}")

    # the rest of the scene, untimed: the Rasterize, image-quality, Refinement and
    # streaming legs run on (or rebuild) the whole cloud, whatever --steps is
    consumed = st["batchlet_index"]
    t0 = time.perf_counter()
    while consumed < nb:
        _, consumed = tree.build_step(view)
        tree.accumulate(consumed)
    tree.sync()
    st_full = tree.stats()
    log(f"whole scene ingested ({nb} batches, {time.perf_counter() - t0:.2f} s): nodes {st_full['num_nodes']}, "
        f"voxels {st_full['num_voxels']:,}")

    # Rasterize, and Analysis + Rasterize as one captured graph (remo_frame), three
    # poses, on the full tree
    from remobench_amd import camera
    import ctypes as C
    rw, rh = (int(v) for v in args.raster.lower().split("x"))
    raster = {"resolution": f"{rw}x{rh}"}
    fbp = C.c_void_p()
    L.check(L.load().remo_dev_alloc(rw * rh * 8, C.byref(fbp)), "framebuffer")
    sample_pts = points[:: max(1, len(points) // 2_000_000)]
    tree.profile_enable(True, ["remolod.render", "remolod.frame"])
    for pose, u in camera.poses(sample_pts, meta.box_size, rw, rh).items():
        tree.render(u, fbp.value)  # warm
        tree.frame(u, fbp.value)
        tree.sync()
        tree.profile_clear()
        reps = 10
        for _ in range(reps):
            tree.render(u, fbp.value)
        for _ in range(reps):
            tree.frame(u, fbp.value)
        tree.sync()
        prof_r = tree.profile()
        ms, launches = prof_r["remolod.render"]
        fms, flaunches = prof_r["remolod.frame"]
        rs = tree.stats()
        raster[pose] = {"ms": round(ms / max(1, launches), 4),
                        "frame_graph_ms": round(fms / max(1, flaunches), 4),
                        "draw_items": rs["draw_items"], "samples": rs["draw_samples"]}
        log(f"render {pose}: {raster[pose]}")
    tree.profile_enable(False)
    L.load().remo_dev_free(fbp)

    # Refinement, colour filtering: one complete pass over the finished tree, an
    # idle call (nothing dirty), then the same steps again with a budgeted
    # refinement after each one (the per-frame budget of the 1B-point config)
    refinement = None
    image_quality = None
    if not args.no_refine:
        # image ground truth: the flat (LOD-free) frame of the resident cloud,
        # against the LOD frame before and after the complete pass below
        from remobench_amd import image_quality as IQ
        iq_frames = IQ.Frames(rw, rh)
        iq_poses = camera.poses(sample_pts, meta.box_size, rw, rh)
        flat_fb, before_fb = {}, {}
        tree.profile_enable(True, ["flat.render"])
        tree.profile_clear()
        for pose, u in iq_poses.items():
            flat_fb[pose] = IQ.flat_frame(tree, int(view.slots), len(points), u, iq_frames)
            before_fb[pose] = IQ.lod_frame(tree, u, iq_frames)
        flat_ms, flat_n = tree.profile()["flat.render"]
        tree.profile_enable(False)
        tree.sync()
        t0 = time.perf_counter()
        full = tree.refine()
        full_ms = (time.perf_counter() - t0) * 1e3
        image_quality = {"flat_render_ms": round(flat_ms / max(1, flat_n), 3), "resolution": f"{rw}x{rh}"}
        for pose, u in iq_poses.items():
            after = IQ.lod_frame(tree, u, iq_frames)
            a = IQ.compare(before_fb[pose], flat_fb[pose], rw, rh)
            b = IQ.compare(after, flat_fb[pose], rw, rh)
            image_quality[pose] = {"coverage": round(a["coverage"], 4), "extra": round(a["extra"], 4),
                                   "depth_rel_p50": a.get("depth_rel_p50"),
                                   "psnr_rgb_first_come": round(a.get("psnr_rgb", 0.0), 3),
                                   "psnr_rgb_refined": round(b.get("psnr_rgb", 0.0), 3),
                                   "mean_abs_rgb_first_come": round(a.get("mean_abs_rgb", 0.0), 3),
                                   "mean_abs_rgb_refined": round(b.get("mean_abs_rgb", 0.0), 3)}
        iq_frames.free()
        log(f"image quality against the flat frame: {image_quality}")
        t0 = time.perf_counter()
        idle = tree.refine()
        idle_ms = (time.perf_counter() - t0) * 1e3
        refinement = {"full_pass_ms": round(full_ms, 3), "full_pass_device_ms": round(full["device_ms"], 3),
                      "nodes_refined": full["nodes_refined"], "points_folded": full["points_folded"],
                      "voxels_written": full["voxels_written"], "levels": full["levels_processed"],
                      "complete": full["complete"], "skipped": full["nodes_skipped"],
                      "estimate_ms": round(full["estimate_ms"], 3),
                      "points_plus_voxels_per_s_g": round((full["points_folded"] + full["voxels_written"]) /
                                                          max(1e-9, full["device_ms"]) / 1e6, 3),
                      "idle_call_ms": round(idle_ms, 3), "idle_dirty": idle["nodes_dirty"]}
        # local deepening of the 64 densest leaves, then Refinement of what it dirtied
        tree.sync()
        t0 = time.perf_counter()
        dp = tree.deepen(64, 25_000)
        deepen_ms = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        after = tree.refine()
        refinement["deepening"] = {"max_splits": 64, "min_points": 25_000, "splits": dp["splits"],
                                   "candidates": dp["candidates"], "points_moved": dp["points_moved"],
                                   "new_voxels": dp["new_voxels"], "ms": round(deepen_ms, 3),
                                   "refine_after_ms": round((time.perf_counter() - t0) * 1e3, 3),
                                   "refine_after_nodes": after["nodes_refined"]}
        log(f"refinement, full pass: {refinement}")
        budget = args.refine_budget_ms
        tree.reset()
        tree.sync()
        calls = []
        t0 = time.perf_counter()
        for _ in range(nb):
            step()
            c0 = time.perf_counter()
            rs = tree.refine(budget)
            calls.append(((time.perf_counter() - c0) * 1e3, rs))
        tree.sync()
        dt_ref = time.perf_counter() - t0
        ms = sorted(c[0] for c in calls)
        drain = 0
        while not tree.refine(budget)["complete"] and drain < 10_000:
            drain += 1
        refinement["interleaved"] = {
            "budget_ms": budget, "calls": len(calls), "avg_ms": round(sum(ms) / len(ms), 3),
            "p95_ms": round(ms[int(0.95 * (len(ms) - 1))], 3), "max_ms": round(ms[-1], 3),
            "avg_device_ms": round(sum(c[1]["device_ms"] for c in calls) / len(calls), 3),
            "nodes_refined": sum(c[1]["nodes_refined"] for c in calls),
            "ingest_with_refinement_mpts": round(tree.stats()["num_points_processed"] / dt_ref / 1e6, 2),
            "calls_to_finish_after_ingest": drain + 1}
        log(f"refinement, interleaved: {refinement['interleaved']}")

    # The full per-frame loop of the 1B-point configuration at this workload's
    # size: per step one batch of Update + Accumulate, Analysis + Rasterize as the
    # captured frame graph at the Rasterize resolution (street pose), Refinement
    # within the budget, and local deepening of the two densest leaves; analytic
    # then learned score on the same steps. Score error: the learned path's scores
    # against the analytic ones at half the steps (mean |difference| over the mean
    # analytic score, nodes with a positive analytic score).
    full_loop = None
    if not args.no_refine and analysis_mode is not None:
        import numpy as np
        full_loop = {"budget_ms": args.refine_budget_ms, "deepen_per_frame": 2, "pose": "street",
                     "resolution": f"{rw}x{rh}"}
        u_loop = camera.poses(sample_pts, meta.box_size, rw, rh)["street"]
        fbl = C.c_void_p()
        L.check(L.load().remo_dev_alloc(rw * rh * 8, C.byref(fbl)), "framebuffer")
        modes = [("analytic", L.REMO_ANALYSIS_ANALYTIC)] + (
            [("learned", L.REMO_ANALYSIS_LEARNED)] if model is not None else [])
        for name, mode in modes:
            tree.reset()
            tree.sync()
            frames_ms = []
            t0 = time.perf_counter()
            for k in range(nb):
                f0 = time.perf_counter()
                _, consumed = tree.build_step(view)
                tree.accumulate(consumed)
                tree.frame(u_loop, fbl.value, mode)
                tree.refine(args.refine_budget_ms)
                tree.deepen(2, 25_000)
                frames_ms.append((time.perf_counter() - f0) * 1e3)
                if name == "learned" and k == nb // 2:
                    tree.analysis(u_loop, L.REMO_ANALYSIS_ANALYTIC)
                    sa = tree.analysis_results()["score"].astype(np.float64)
                    tree.analysis(u_loop, L.REMO_ANALYSIS_LEARNED)
                    sl = tree.analysis_results()["score"].astype(np.float64)
                    m = sa > 0
                    full_loop["score_error_learned"] = round(float(np.abs(sl[m] - sa[m]).mean() /
                                                                   max(1e-12, sa[m].mean())), 5)
            tree.sync()
            dtl = time.perf_counter() - t0
            fs = sorted(frames_ms)
            full_loop[name] = {"ingest_mpts": round(tree.stats()["num_points_processed"] / dtl / 1e6, 2),
                               "frame_ms_avg": round(sum(fs) / len(fs), 3),
                               "frame_ms_p95": round(fs[int(0.95 * (len(fs) - 1))], 3),
                               "nodes": tree.stats()["num_nodes"]}
        L.load().remo_dev_free(fbl)
        log(f"full loop: {full_loop}")

    # PCIe-inclusive streaming: the cloud in pinned host memory (one untimed copy,
    # as a reader would leave it), the reference's 50-slot ring refilled on a copy
    # stream while the same steps run; each step waits only for its own batch
    streaming = None
    if not args.no_streaming:
        depth = min(50, nb)
        sring = PointRing(args.batch, depth)
        sring.set_cloud_pinned(points, args.batch)
        stree = Tree(meta, slot_capacity=args.batch, ring_slots=depth, batches_per_step=1, step_budget_ms=-1.0,
                     fixed_depth=args.fixed_depth)
        stree.sync()
        t0 = time.perf_counter()
        consumed = 0
        while consumed < nb:
            sring.push_async(stree)
            sring.order_before(stree, consumed, 1)
            _, consumed = stree.build_step(sring.view())
            sring.set_batches_consumed(consumed)
            stree.accumulate(consumed)
            stree.analysis(uniforms, L.REMO_ANALYSIS_ANALYTIC)
        stree.sync()
        sring.sync()
        dts = time.perf_counter() - t0
        sst = stree.stats()
        streaming = {"value": round(sst["num_points_processed"] / dts / 1e6, 2), "unit": "M pts/s",
                     "ring_slots": depth, "batches": nb, "pushes": sring.produced,
                     "ms_per_step": round(dts * 1e3 / nb, 4),
                     "same_tree": sst["num_nodes"] == st_full["num_nodes"] and
                     sst["num_voxels"] == st_full["num_voxels"]}
        log(f"streaming from pinned host memory: {streaming}")
        stree.close()
        sring.close()

    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle on {args.cpu_sample_batches} batches")
        cpu = cpu_baseline(points, meta, args.batch, args.cpu_sample_batches)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "M pts/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / steps, 4),
        "ms_per_step_all_scopes_timed": round(dt_all * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32+int64",
        "data": (f"file: {args.simlod}, resident in HBM" if args.simlod else
                 f"synthetic: seed-generated {args.scene} scene (seed {args.seed}), resident in HBM"),
        "config": {
            "workload": workload,
            "points": len(points), "batch": args.batch, "batches_timed": steps,
            "expand": f"fixed depth {args.fixed_depth}" if args.fixed_depth else "iterative",
            "analysis": args.analysis if analysis_ok else "unavailable",
        },
        "roofline": roofline,
        "learned_score": learned,
        "cpu_baseline": cpu,
        "streaming": streaming,
        "upload_gbs": round(16.0 * len(points) / upload_s / 1e9, 2),
        "pcie_inclusive_mpts": round(ingested / (dt + upload_s * ingested / len(points)) / 1e6, 2),
        "path_compulsory_bytes_per_point": round(compulsory_pp, 2),
        "path_hbm_gbs": round(path_gbs, 2),
        "path_hbm_frac": round(path_gbs / HBM_PEAK_GBS, 5),
        "scopes_hbm_gbs": round(scopes_gbs, 2),
        "scopes_hbm_frac": round(scopes_gbs / HBM_PEAK_GBS, 5),
        "atomics": atomics,
        "plugin_build_path": plugin_path,
        "scopes": scopes,
        "rasterize": raster,
        "refinement": refinement,
        "image_quality": image_quality,
        "full_loop": full_loop,
        "tree": {k: st_full[k] for k in ("num_nodes", "num_inner", "num_leaves", "num_points", "max_points_per_node",
                                    "total_expand_iters", "total_spilled_points",
                                    "num_voxels", "max_depth", "allocated_bytes_persistent",
                                    "mem_capacity_reached", "node_pool_overflow", "spill_overflow",
                                    "backlog_overflow", "dir_overflow")},
    }
    print(json.dumps(out), flush=True)
    tree.close()
    ring.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
