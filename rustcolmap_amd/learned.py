"""Training data, training loop and weight files of the plane-proposal network.

Training scenes are seeded random variants of the synthetic room (random
camera arc, distance, height and target, 3-10 seeded spheres/boxes, random
texture scale and noise), 8 views each.  Targets are the plain patch-match's
own 8-sweep output on those scenes (normalised inverse depth, accepted
normal, validity = the reference's gate) at every second pixel, computed on
the GPU; the network input is the 32-plane census sweep volume over 4
sources plus the gray level on the same coarse grid.  The confidence learns
whether a proposal lies within 2 % of a valid target.  (Counting pixels
where the plain method finds no valid depth as positives too was tried: it
accepts almost every proposal and lowers the agreement with the 8-sweep
run below that of two plain sweeps.)

    python -m rustcolmap_amd.learned --scenes 2000 --steps 3000 --out weights.npz
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time
from typing import Callable, List, Optional

import numpy as np

from .mvs import NET_STRIDE, MvsContext, MvsOptions, ProposalNet, ViewData
from .scene import SceneSpec, render_scene, select_sources

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "proposal_unet.npz")


def training_spec(seed: int, size=(640, 480), n_views: int = 8) -> SceneSpec:
    """Seeded random scene for training (seeds 1000-2999 in the default set)."""
    r = np.random.default_rng(seed)
    w, h = size
    return SceneSpec(seed=seed, n_views=n_views, sizes=((w, h),), focal=float(w * r.uniform(0.75, 1.0)),
                     arc_deg=float(r.uniform(20.0, 70.0)), radius=float(r.uniform(3.2, 5.0)),
                     arc_center=float(r.uniform(-0.8, 1.0)), cam_h=float(r.uniform(0.9, 2.4)),
                     target=(float(r.uniform(0.3, 1.2)), float(r.uniform(0.5, 1.5)), float(r.uniform(0.3, 1.2))),
                     n_objects=int(r.integers(3, 11)), obj_region=(0.3, 2.0), flat_patch=bool(r.random() < 0.5),
                     noise_sigma=float(r.choice([0.0, 1.0, 2.0])), tex_freq=float(r.uniform(10.0, 24.0)),
                     persistence=0.65)


def targets_from_depth(dm, zmin: float, zmax: float) -> np.ndarray:
    """(h, w, 5) target: normalised inverse depth, normal, validity."""
    inv_min, inv_max = 1.0 / zmax, 1.0 / zmin
    z = dm.depth
    valid = z > 0
    t = np.zeros(z.shape + (5,), np.float32)
    t[..., 0] = np.where(valid, (1.0 / np.where(valid, z, 1.0) - inv_min) / (inv_max - inv_min), 0.0)
    t[..., 1:4] = dm.normals
    t[..., 4] = valid
    return t


def build_training_set(ctx: MvsContext, net: ProposalNet, n_scenes: int, size=(640, 480), views_per_scene: int = 2,
                       seed0: int = 1000, sweeps: int = 8, k: int = 4, zmin: float = 1.0, zmax: float = 10.0,
                       log: Optional[Callable[[str], None]] = None) -> int:
    """Renders n_scenes training scenes, runs the plain method on
    `views_per_scene` reference views of each, and stores (sweep volume,
    target) samples on the device.  Returns the number of samples."""
    opts = MvsOptions(num_sources=k, sweeps=sweeps)
    n = 0
    t0 = time.time()
    for si in range(n_scenes):
        spec = training_spec(seed0 + si, size)
        sc = render_scene(spec, with_gt=False)
        views = [ViewData.from_scene_view(v) for v in sc.views]
        ctx.upload_views(views)
        srcs = select_sources(sc.views, k)
        step = max(1, len(views) // views_per_scene)
        refs = list(range(step // 2, len(views), step))[:views_per_scene]
        jobs = [(r, srcs[r], zmin, zmax) for r in refs]
        out = ctx.estimate_depth_batch(jobs, opts, mode=2 | 4)
        tg = [targets_from_depth(out[r], zmin, zmax)[::NET_STRIDE, ::NET_STRIDE] for r in refs]
        n = net.add_samples(jobs, tg)
        if log and (si + 1) % max(1, n_scenes // 10) == 0:
            log("training set: %d/%d scenes, %d samples, %.1fs" % (si + 1, n_scenes, n, time.time() - t0))
    return n


def train(net: ProposalNet, n_samples: int, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 0,
          log: Optional[Callable[[str], None]] = None) -> List[np.ndarray]:
    """Adam steps over random batches, cosine-decayed learning rate."""
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for s in range(steps):
        idx = rng.choice(n_samples, size=min(batch, n_samples), replace=False)
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * s / max(1, steps)))
        hist.append(net.train_step(idx, cur))
        if log and (s + 1) % max(1, steps // 20) == 0:
            m = np.mean(hist[-max(1, steps // 20):], axis=0)
            log("step %d/%d lr %.2e depthL1 %.4f normal %.4f bce %.4f valid %.3f (%.1fs)"
                % (s + 1, steps, cur, m[0], m[1], m[2], m[3], time.time() - t0))
    return hist


def to_bf16_bits(p: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even bfloat16 bit patterns of float32 values."""
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def save_weights(net: ProposalNet, path: str, meta: dict):
    """Weights as bfloat16 bit patterns (the convolutions run on bf16
    weights, so inference is unchanged by the rounding)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, params_bf16=to_bf16_bits(net.get_params()),
                        **{k: np.asarray(v) for k, v in meta.items()})


def load_weights(net: ProposalNet, path: str = DEFAULT_WEIGHTS) -> dict:
    with np.load(path, allow_pickle=False) as z:
        net.set_params(from_bf16_bits(z["params_bf16"]) if "params_bf16" in z.files else z["params"])
        return {k: z[k] for k in z.files if k not in ("params", "params_bf16")}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--views-per-scene", type=int, default=1)
    ap.add_argument("--size", default="640x480")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    w, h = (int(v) for v in a.size.split("x"))

    def log(m):
        print("[learned] " + m, file=sys.stderr, flush=True)

    ctx = MvsContext(0)
    net = ProposalNet(ctx, seed=a.seed)
    t0 = time.time()
    n = build_training_set(ctx, net, a.scenes, (w, h), a.views_per_scene, log=log)
    t1 = time.time()
    hist = train(net, n, a.steps, a.batch, a.lr, a.seed, log=log)
    t2 = time.time()
    save_weights(net, a.out, {"scenes": a.scenes, "samples": n, "steps": a.steps, "batch": a.batch, "lr": a.lr,
                              "size": [w, h], "data_seconds": t1 - t0, "train_seconds": t2 - t1,
                              "final_loss": np.mean(hist[-50:], axis=0)})
    log("saved %s (%d samples, data %.0fs, train %.0fs)" % (a.out, n, t1 - t0, t2 - t1))


if __name__ == "__main__":
    main()
