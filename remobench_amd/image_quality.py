"""Image quality of LOD frames against the flat (LOD-free) frame.

The flat frame splats every point of the cloud (remo_render_flat, the
reference's `flat` pipeline); a LOD frame splats the frontier Rasterize selects
(leaf points and inner-node voxels). Both are u64 buffers of
(float_bits(depth) << 32) | rgba. Per pose this module reports, over the pixels
the flat frame covers:

  coverage        fraction also covered by the LOD frame
  extra           LOD-covered pixels the flat frame leaves empty, over flat-covered
  psnr_rgb        PSNR (dB, peak 255) of the RGB channels over pixels both cover
  mean_abs_rgb    mean absolute channel difference there
  depth_rel_p50 / depth_rel_p95   |d_lod - d_flat| / d_flat there

and, for the same tree before and after Refinement's colour filtering, how the
voxel colours change those numbers.

    python -m remobench_amd.image_quality --points 20000000
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import camera
from . import lib as L
from .pipeline import PointRing, Tree

EMPTY_DEPTH_BITS = 0x7F800000  # the cleared buffer's depth (+inf)


def resolve(fb: np.ndarray, width: int, height: int):
    """(rgb uint8 [h, w, 3], depth float32 [h, w], covered bool [h, w])."""
    fb = np.asarray(fb, dtype=np.uint64).reshape(height, width)
    depth_bits = (fb >> np.uint64(32)).astype(np.uint32)
    covered = depth_bits != EMPTY_DEPTH_BITS
    depth = depth_bits.view(np.float32)
    c = (fb & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    rgb = np.stack([(c >> s) & 0xFF for s in (0, 8, 16)], axis=-1).astype(np.uint8)
    return rgb, depth, covered


def compare(lod: np.ndarray, flat: np.ndarray, width: int, height: int) -> dict:
    rl, dl, cl = resolve(lod, width, height)
    rf, df, cf = resolve(flat, width, height)
    nf = int(cf.sum())
    both = cl & cf
    nb = int(both.sum())
    out = {"flat_pixels": nf, "coverage": nb / max(1, nf), "extra": int((cl & ~cf).sum()) / max(1, nf)}
    if nb:
        diff = rl[both].astype(np.float64) - rf[both].astype(np.float64)
        mse = float(np.mean(diff * diff))
        out["psnr_rgb"] = float(10.0 * np.log10(255.0 ** 2 / mse)) if mse > 0 else float("inf")
        out["mean_abs_rgb"] = float(np.mean(np.abs(diff)))
        rel = np.abs(dl[both].astype(np.float64) - df[both]) / np.maximum(df[both].astype(np.float64), 1e-30)
        out["depth_rel_p50"] = float(np.quantile(rel, 0.5))
        out["depth_rel_p95"] = float(np.quantile(rel, 0.95))
    return out


class Frames:
    """Device framebuffers of one size and their downloads."""

    def __init__(self, width: int, height: int):
        self.width, self.height = width, height
        self.buf = L.DeviceBuffer(8 * width * height)

    def download(self) -> np.ndarray:
        return self.buf.download(np.uint64, (self.width * self.height,))

    def free(self):
        self.buf.free()


def lod_frame(tree: Tree, u: L.SharedUniforms, frames: Frames) -> np.ndarray:
    tree.render(u, frames.buf.ptr)
    tree.sync()
    return frames.download()


def flat_frame(tree: Tree, points_ptr: int, n: int, u: L.SharedUniforms, frames: Frames) -> np.ndarray:
    tree.render_flat(points_ptr, n, u, frames.buf.ptr)
    tree.sync()
    return frames.download()


def measure(tree: Tree, points_ptr: int, n: int, poses: dict, width: int, height: int) -> dict:
    """Per pose: the LOD frame as built, then after one complete Refinement
    pass, both against the flat frame. Leaves the tree refined."""
    frames = Frames(width, height)
    try:
        flat, before = {}, {}
        for name, u in poses.items():
            flat[name] = flat_frame(tree, points_ptr, n, u, frames)
            before[name] = lod_frame(tree, u, frames)
        st = tree.refine()
        out = {"refine_complete": st["complete"]}
        for name, u in poses.items():
            after = lod_frame(tree, u, frames)
            out[name] = {"first_come": compare(before[name], flat[name], width, height),
                         "refined": compare(after, flat[name], width, height)}
        return out
    finally:
        frames.free()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--scene", default="lidar")
    ap.add_argument("--points", type=int, default=20_000_000)
    ap.add_argument("--batch", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--raster", default="1920x1080")
    args = ap.parse_args(argv)
    if L.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    w, h = (int(v) for v in args.raster.lower().split("x"))
    points, meta = L.generate_scene(args.points, args.seed, args.scene)
    nb = (len(points) + args.batch - 1) // args.batch
    ring = PointRing(args.batch, nb)
    ring.set_cloud(points, args.batch)
    ring.pump()
    tree = Tree(meta, slot_capacity=args.batch, ring_slots=nb, batches_per_step=20, step_budget_ms=-1.0)
    consumed = 0
    while consumed < nb:
        _, consumed = tree.build_step(ring.view())
        tree.accumulate(consumed)
    sample = points[:: max(1, len(points) // 2_000_000)]
    poses = camera.poses(sample, meta.box_size, w, h)
    view = ring.view()
    out = {"scene": args.scene, "points": len(points), "raster": args.raster,
           **measure(tree, int(view.slots), len(points), poses, w, h)}
    print(json.dumps(out), flush=True)
    tree.close()
    ring.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
