"""Precision of the per-leaf covariance: int64 fixed point vs FP32 atomics vs FP64.

Builds a scene on the GPU, then for a random sample of leaves compares three ways
of accumulating a leaf's second moments against an FP64 reference computed here
from the leaf's stored points (numpy, float64, node-local coordinates):

  int64   the path's node-centred fixed point (q = floor(l * 2^20) - 2^19, exact
          integer sums; what Analysis rolls up);
  fp64    the reference's FP64 sums (kernel_accumulate's g[9], order-dependent);
  fp32    naive FP32 atomics of the same node-local coordinates
          (remo_accum_fp32_atomics).

Errors are max |cov - cov_ref| over the six entries, relative to the trace of
cov_ref. Also reports the Analysis eigenvalues (FP32 Jacobi) against
numpy.linalg.eigvalsh of the FP64 reference covariance, for the same leaves.

    python -m remobench_amd.precision --scene skewed --points 300000000
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import lib as L
from .pipeline import PointRing, Tree

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def cov_from_sums(n, s1, s2):
    """Biased covariance (6 entries) from sums of l and of l_a l_b."""
    m = np.asarray(s1, dtype=np.float64) / n
    return np.array([s2[k] / n - m[a] * m[b] for k, (a, b) in enumerate(PAIRS)])


def cov_int64(acc):
    """Exact covariance in node-extent units from the int64 fixed-point sums."""
    n = int(acc["count"])
    q1 = [int(v) for v in acc["q1"]]
    q2 = [int(v) for v in acc["q2"]]
    # n^2 cov = n * S2 - S1 S1 (exact integers), units 2^-40
    return np.array([(n * q2[k] - q1[a] * q1[b]) / (n * n) / 2.0 ** 40 for k, (a, b) in enumerate(PAIRS)])


def run(scene: str, num_points: int, batch: int, seed: int, sample: int, log=print) -> dict:
    t0 = time.perf_counter()
    points, meta = L.generate_scene(num_points, seed, scene)
    nb = (len(points) + batch - 1) // batch
    ring = PointRing(batch, nb)
    ring.set_cloud(points, batch)
    ring.pump()
    tree = Tree(meta, slot_capacity=batch, ring_slots=nb, batches_per_step=20, step_budget_ms=-1.0)
    consumed = 0
    while consumed < nb:
        _, consumed = tree.build_step(ring.view())
        tree.accumulate(consumed)
    tree.analysis(L.SharedUniforms(), L.REMO_ANALYSIS_ANALYTIC)
    log(f"built {scene} {len(points):,} points in {time.perf_counter() - t0:.1f} s")
    nodes, acc = tree.nodes(), tree.accums()
    f32 = tree.fp32_atomic_sums()
    res = tree.analysis_results()
    size = float(max(meta.box_size))
    leaves = np.nonzero((nodes["first_child"] < 0) & (acc["count"] >= 3))[0]
    rng = np.random.default_rng(seed)
    pick = rng.choice(leaves, min(sample, len(leaves)), replace=False)
    err = {"int64": [], "fp64": [], "fp32": [], "eig": []}
    counts = []
    for i in pick:
        rec = nodes[i]
        pts = tree.samples(int(i))
        ns = size / 2.0 ** int(rec["level"])
        mn = np.array([rec["x"], rec["y"], rec["z"]], dtype=np.float64) * ns
        l = np.clip((np.stack([pts["x"], pts["y"], pts["z"]], 1).astype(np.float64) - mn) / ns, 0.0, 1.0)
        c = np.cov(l.T, bias=True)
        ref = np.array([c[a, b] for a, b in PAIRS])
        tr = float(np.trace(c))
        if not tr > 0:
            continue
        n = len(pts)
        counts.append(n)
        a = acc[i]
        err["int64"].append(np.abs(cov_int64(a) - ref).max() / tr)
        g = a["g"]
        err["fp64"].append(np.abs(cov_from_sums(n, g[:3], g[3:]) - ref).max() / tr)
        s = f32[i]
        err["fp32"].append(np.abs(cov_from_sums(float(s[0]), s[1:4], s[4:10]) - ref).max() / tr)
        err["eig"].append(np.abs(res[i]["lambda"].astype(np.float64) - np.linalg.eigvalsh(c)).max() / tr)
    tree.close()
    ring.close()

    def summary(v):
        v = np.asarray(v)
        return {"median": float(np.median(v)), "p99": float(np.quantile(v, 0.99)), "max": float(v.max())}

    return {
        "scene": scene, "points": len(points), "batch": batch, "leaves_sampled": len(counts),
        "leaf_points": {"min": int(min(counts)), "median": int(np.median(counts)), "max": int(max(counts))},
        "cov_error_rel_trace": {k: summary(err[k]) for k in ("int64", "fp64", "fp32")},
        "eigen_error_rel_trace": summary(err["eig"]),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--scene", default="skewed")
    ap.add_argument("--points", type=int, default=300_000_000)
    ap.add_argument("--batch", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--sample", type=int, default=400)
    args = ap.parse_args(argv)
    if L.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    out = run(args.scene, args.points, args.batch, args.seed, args.sample,
              log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
