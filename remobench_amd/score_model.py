"""The learned score: training-pair logging, training and evaluation on the GPU.

A 13 -> 64 -> 64 -> 2 MLP (BF16 weights, FP32 accumulation, exact GELU) predicts
a node's final (surface variation, planarity) from its partial accumulators, so
Analysis can score open nodes before they close; the analytic eigen-solve stays
the fallback (predicted surface variation above 0.05, and every closed node) and
the judge of its error. All arithmetic runs in the HIP kernels of
`csrc/mlp.hip` through the C ABI; this module only moves arrays.

Pairs are logged from plain-pipeline runs of the 2,097,152-point scene (8 batches
of 262,144): after 1, 2, 4 and 6 batches (12.5 / 25 / 50 / 75 % arrival) every
node's 13 inputs are recorded, and paired with the same node's surface variation
and planarity once the whole cloud has arrived.

    python -m remobench_amd.score_model --seeds 100 --out remobench_amd/models/score_mlp.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import lib as L
from .pipeline import PointRing, Tree

DEFAULT_MODEL = Path(__file__).resolve().parent / "models" / "score_mlp.npy"
SNAPSHOT_BATCHES = (1, 2, 4, 6)
FALLBACK_SV = 0.05


class ScoreMLP:
    """A remo_mlp training object."""

    def __init__(self, seed: int = 1, max_batch: int = 8192):
        self.lib = L.load()
        self.h = C.c_void_p()
        L.check(self.lib.remo_mlp_create(seed, max_batch, C.byref(self.h)), "remo_mlp_create")
        self.max_batch = max_batch

    def close(self):
        if self.h:
            self.lib.remo_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def params(self) -> np.ndarray:
        out = np.zeros(L.MLP_PARAMS, dtype=np.float32)
        L.check(self.lib.remo_mlp_get_params(self.h, out.ctypes.data_as(C.c_void_p), L.MLP_PARAMS), "get_params")
        return out

    def set_params(self, theta: np.ndarray):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.shape != (L.MLP_PARAMS,):
            raise ValueError(f"expected {L.MLP_PARAMS} parameters")
        L.check(self.lib.remo_mlp_set_params(self.h, theta.ctypes.data_as(C.c_void_p), L.MLP_PARAMS),
                "set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.MLP_INPUTS)
        bx, by = L.DeviceBuffer.of(x), L.DeviceBuffer(8 * len(x))
        try:
            L.check(self.lib.remo_mlp_forward(self.h, C.c_void_p(bx.ptr), len(x), C.c_void_p(by.ptr)), "forward")
            return by.download(np.float32, (len(x), L.MLP_OUTPUTS))
        finally:
            bx.free()
            by.free()

    def gradient(self, x: np.ndarray, t: np.ndarray):
        """(loss, gradient) on one batch, without an update."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.MLP_INPUTS)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, L.MLP_OUTPUTS)
        bx, bt = L.DeviceBuffer.of(x), L.DeviceBuffer.of(t)
        g = np.zeros(L.MLP_PARAMS, dtype=np.float32)
        loss = C.c_float(0)
        try:
            L.check(self.lib.remo_mlp_gradient(self.h, C.c_void_p(bx.ptr), C.c_void_p(bt.ptr), len(x),
                                               C.byref(loss), g.ctypes.data_as(C.c_void_p)), "gradient")
        finally:
            bx.free()
            bt.free()
        return loss.value, g

    def train_step_device(self, x_ptr: int, t_ptr: int, n: int, lr: float) -> float:
        loss = C.c_float(0)
        L.check(self.lib.remo_mlp_train_step(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), n, lr, C.byref(loss)),
                "train_step")
        return loss.value

    def save(self, path: Path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.save(path, self.params())

    @classmethod
    def load(cls, path: Path = DEFAULT_MODEL) -> "ScoreMLP":
        theta = np.load(path, allow_pickle=False)
        m = cls()
        m.set_params(theta)
        return m


def log_pairs(seed: int, num_points: int = 2_097_152, batch: int = 262_144, snapshots=SNAPSHOT_BATCHES,
              kind: str = "lidar"):
    """Runs the plain pipeline on one seed and returns (X, Y, partial) where X are
    node inputs at the snapshots, Y the same nodes' final (surface variation,
    planarity), partial their analytic values at the snapshot."""
    points, meta = L.generate_scene(num_points, seed, kind)
    nb = (len(points) + batch - 1) // batch
    ring = PointRing(batch, nb)
    ring.set_cloud(points, batch)
    ring.pump()
    tree = Tree(meta, slot_capacity=batch, ring_slots=nb, batches_per_step=1, step_budget_ms=-1.0)
    u = L.SharedUniforms()
    snaps = []
    try:
        consumed = 0
        while consumed < nb:
            _, consumed = tree.build_step(ring.view())
            tree.accumulate(consumed)
            if consumed in snapshots:
                tree.analysis(u, L.REMO_ANALYSIS_ANALYTIC)
                r = tree.analysis_results()
                snaps.append((tree.features(), r["surface_variation"].copy(), r["planarity"].copy(), r["n"].copy()))
        tree.analysis(u, L.REMO_ANALYSIS_ANALYTIC)
        final = tree.analysis_results()
    finally:
        tree.close()
        ring.close()
    X, Y, P = [], [], []
    for f, sv, pl, n in snaps:
        k = len(f)  # node indices are stable: nodes present at the snapshot keep theirs
        ok = (n >= 3) & (final["n"][:k] >= 3)
        X.append(f[ok])
        Y.append(np.stack([final["surface_variation"][:k][ok], final["planarity"][:k][ok]], axis=1))
        P.append(np.stack([sv[ok], pl[ok]], axis=1))
    return (np.concatenate(X).astype(np.float32), np.concatenate(Y).astype(np.float32),
            np.concatenate(P).astype(np.float32))


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 60, batch: int = 8192, lr: float = 1e-3, seed: int = 1,
          log=None) -> ScoreMLP:
    """Adam on shuffled minibatches of (X, Y), all arithmetic on the GPU."""
    rng = np.random.default_rng(seed)
    m = ScoreMLP(seed=seed, max_batch=batch)
    bx, bt = L.DeviceBuffer(4 * L.MLP_INPUTS * batch), L.DeviceBuffer(4 * L.MLP_OUTPUTS * batch)
    try:
        for ep in range(epochs):
            order = rng.permutation(len(X))
            losses = []
            for s in range(0, len(X), batch):
                idx = order[s: s + batch]
                bx.upload(X[idx])
                bt.upload(Y[idx])
                losses.append(m.train_step_device(bx.ptr, bt.ptr, len(idx), lr) * len(idx))
            if log and (ep % 10 == 0 or ep == epochs - 1):
                log(f"epoch {ep}: loss {sum(losses) / len(X):.3e}")
    finally:
        bx.free()
        bt.free()
    return m


def evaluate(m: ScoreMLP, X: np.ndarray, Y: np.ndarray, P: np.ndarray) -> dict:
    """Errors of the learned prediction and of the analytic value on partial data,
    both against the final analytic (surface variation, planarity). "score error"
    is the mean absolute surface-variation error over the mean final surface
    variation (the score is 3 * surface variation * coverage, coverage shared).
    The learned path is what Analysis uses: the prediction where it is <= 0.05,
    the analytic value otherwise."""
    pred = m.forward(X)
    used = np.where((pred[:, 0] <= FALLBACK_SV)[:, None], pred, P)
    msv = float(np.mean(np.abs(Y[:, 0])))
    return {
        "pairs": int(len(X)),
        "learned_fraction": float(np.mean(pred[:, 0] <= FALLBACK_SV)),
        "sv_mae_learned_path": float(np.mean(np.abs(used[:, 0] - Y[:, 0]))),
        "sv_mae_prediction": float(np.mean(np.abs(pred[:, 0] - Y[:, 0]))),
        "sv_mae_partial_analytic": float(np.mean(np.abs(P[:, 0] - Y[:, 0]))),
        "planarity_mae_prediction": float(np.mean(np.abs(pred[:, 1] - Y[:, 1]))),
        "planarity_mae_partial_analytic": float(np.mean(np.abs(P[:, 1] - Y[:, 1]))),
        "score_error_learned_path": float(np.mean(np.abs(used[:, 0] - Y[:, 0])) / max(msv, 1e-12)),
        "score_error_partial_analytic": float(np.mean(np.abs(P[:, 0] - Y[:, 0])) / max(msv, 1e-12)),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--seeds", type=int, default=100, help="training scenes (seeds 1000..)")
    ap.add_argument("--eval-seeds", type=int, default=10, help="held-out scenes (seeds 5000..)")
    ap.add_argument("--points", type=int, default=2_097_152)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--out", type=Path, default=DEFAULT_MODEL)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if L.device_count() < 1:
        log("no GPU visible")
        return 1
    t0 = time.perf_counter()
    parts = [log_pairs(1000 + s, args.points) for s in range(args.seeds)]
    X = np.concatenate([p[0] for p in parts])
    Y = np.concatenate([p[1] for p in parts])
    log(f"logged {len(X):,} pairs from {args.seeds} scenes in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    m = train(X, Y, epochs=args.epochs, log=log)
    log(f"trained in {time.perf_counter() - t1:.1f} s")
    ev = [log_pairs(5000 + s, args.points) for s in range(args.eval_seeds)]
    res = evaluate(m, np.concatenate([e[0] for e in ev]), np.concatenate([e[1] for e in ev]),
                   np.concatenate([e[2] for e in ev]))
    res["train"] = evaluate(m, X, Y, np.concatenate([p[2] for p in parts]))
    m.save(args.out)
    res["model"] = str(args.out)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
