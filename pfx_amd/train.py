"""Training pairs from the tracer, and training of the learned denoiser.

Pairs are cut from frames of seeded terrain scenes (the procedural terrain of
`scenes.terrain`: fBm heightfield, instanced meshes, 128 point and spot
lights; scene seed 2, the 4K benchmark scene, is never used): per frame the
tracer renders a low-spp input (colour, albedo and normal as the still
develops them, depth from the probe) and, from an independent sample
sequence, a high-spp target; both are cut into square crops and go to the
device-resident training set. Training then runs Adam on the GPU through
`Learner.train`.

    python -m pfx_amd.train --crops 1600 --out weights.pfxl [--init weights.pfxl] [--loss rel_l2]
"""
from __future__ import annotations

import argparse
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import Bounces, Camera, Detail, Trace, scenes
from . import learn as Lm

BENCH_SEED = 2


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def views(rng: np.random.Generator, scene, count: int, aspect: float) -> list[Camera]:
    """Cameras over the terrain: half wide shots like the benchmark view,
    half closer views of the ground, instances and lights."""
    top = float(scene.triangles["vertices"][:, 1::3].max())
    out = []
    for k in range(count):
        if k % 2 == 0:
            eye = (rng.uniform(-40, 40), top + rng.uniform(6, 22), rng.uniform(-50, -25))
            target = (eye[0] * 0.3 + rng.uniform(-10, 10), rng.uniform(-2, 2), rng.uniform(-5, 15))
            fov = rng.uniform(45, 65)
        else:
            eye = (rng.uniform(-55, 55), top + rng.uniform(1, 8), rng.uniform(-28, 28))
            yaw = rng.uniform(0, 2 * math.pi)
            dist = rng.uniform(8, 30)
            target = (eye[0] + dist * math.cos(yaw), rng.uniform(-4, 3), eye[2] + dist * math.sin(yaw))
            fov = rng.uniform(35, 70)
        out.append(Camera.look(eye, target, fov, aspect))
    return out


def frame_pairs(trace: Trace, camera: Camera, input_spp: int, target_spp: int, seed: int):
    """(color, albedo, normal, depth, deviation) at input_spp and the target
    colour at target_spp, from independent sample sequences."""
    trace.set_camera(camera)
    trace.sample(input_spp, seed)
    out = trace.readback()
    deviation = trace.readback_deviation()
    depth = trace.probe().depth.astype(np.float32)
    trace.set_camera(camera)
    left, chunk = target_spp, 64
    while left > 0:
        trace.sample(min(chunk, left), seed + 7919)
        left -= chunk
    target = trace.readback_color()
    return out.color, out.albedo, out.normal, depth, deviation, target


def generate(learner: Lm.Learner, crops: int, side: int, input_spp, target_spp: int, seed: int,
             frame=(1024, 512), views_per_scene: int = 20, deadline: float | None = None) -> int:
    """Fill the learner's training set with up to `crops` crops; returns how
    many. `input_spp` is one count or a list drawn from per view."""
    spps = list(input_spp) if isinstance(input_spp, (list, tuple)) else [int(input_spp)]
    fw, fh = frame
    per_frame = (fw // side) * (fh // side)
    learner.reserve(side, crops)
    rng = np.random.default_rng(seed)
    added, scene_seed = 0, 1000 * seed + 1
    while added < crops:
        if scene_seed == BENCH_SEED:
            scene_seed += 1
        t0 = time.perf_counter()
        scene = scenes.terrain(seed=scene_seed, width=fw, height=fh)
        with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(6)), fw, fh) as trace:
            log(f"scene {scene_seed}: built in {time.perf_counter() - t0:.1f} s")
            for k, cam in enumerate(views(rng, scene, views_per_scene, fw / fh)):
                t1 = time.perf_counter()
                spp = int(spps[int(rng.integers(0, len(spps)))])
                color, albedo, normal, depth, deviation, target = frame_pairs(trace, cam, spp, target_spp,
                                                                              int(rng.integers(1, 2**31)))
                for cy in range(fh // side):
                    for cx in range(fw // side):
                        if added >= crops:
                            break
                        sl = (slice(cy * side, (cy + 1) * side), slice(cx * side, (cx + 1) * side))
                        learner.add(color[sl], albedo[sl], normal[sl], depth[sl], target[sl], deviation[sl])
                        added += 1
                log(f"  view {k} ({spp} spp): {per_frame} crops in {time.perf_counter() - t1:.1f} s, {added}/{crops}")
                if added >= crops or (deadline and time.perf_counter() > deadline):
                    break
        scene_seed += 1
        if deadline and time.perf_counter() > deadline:
            break
    return added


def synthetic(learner: Lm.Learner, crops: int, side: int, seed: int) -> int:
    """Random crops of the right shapes, for timing the training step alone."""
    rng = np.random.default_rng(seed)
    learner.reserve(side, crops)
    for _ in range(crops):
        shape = (side, side, 4)
        learner.add(rng.exponential(0.5, shape).astype(np.float32), rng.uniform(0, 1, shape).astype(np.float32),
                    rng.uniform(-1, 1, shape).astype(np.float32), rng.uniform(0, 50, (side, side)).astype(np.float32),
                    rng.exponential(0.5, shape).astype(np.float32), rng.exponential(0.1, (side, side)).astype(np.float32))
    return crops


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--crops", type=int, default=1600)
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--input-spp", default="4", help="spp of the inputs: one count or a comma list drawn per view")
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4, help="cosine decay to this learning rate")
    ap.add_argument("--loss", default="l1_log", choices=["l1_log", "rel_l2"],
                    help="L1 on log radiance, or relative L2 of linear radiance (noisy targets allowed)")
    ap.add_argument("--init", type=Path, help="start from these weights (else He init)")
    ap.add_argument("--out", type=Path, required=True)
    ap.add_argument("--gen-seconds", type=float, default=0, help="stop generating after this long (0: no limit)")
    ap.add_argument("--train-seconds", type=float, default=0, help="stop training after this long (0: no limit)")
    ap.add_argument("--synthetic", action="store_true", help="random crops instead of renders (timing only)")
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    params = Lm.load_params(args.init) if args.init else Lm.init_params(args.seed)
    with Lm.Learner(params=params) as learner:
        learner.set_loss(args.loss)
        deadline = t0 + args.gen_seconds if args.gen_seconds else None
        if args.synthetic:
            n = synthetic(learner, args.crops, args.side, args.seed)
        else:
            spps = [int(v) for v in str(args.input_spp).split(",")]
            n = generate(learner, args.crops, args.side, spps, args.target_spp, args.seed, deadline=deadline)
        log(f"training set: {n} crops of {args.side}^2 in {time.perf_counter() - t0:.0f} s")
        t1 = time.perf_counter()
        chunk, done, history = 100, 0, []
        while done < args.steps:
            k = min(chunk, args.steps - done)
            frac = done / max(args.steps, 1)
            lr = args.lr_final + 0.5 * (args.lr - args.lr_final) * (1 + math.cos(math.pi * frac))
            losses = learner.train(k, batch=args.batch, lr=lr, seed=args.seed * 1_000_003 + done)
            done += k
            history.extend(float(v) for v in losses)
            el = time.perf_counter() - t1
            log(f"step {done}: loss {np.mean(losses):.4f} lr {lr:.2e} ({el / done * 1e3:.1f} ms/step)")
            learner.save(args.out)
            if args.train_seconds and el > args.train_seconds:
                log(f"stopping at the time limit after {done} steps")
                break
        learner.save(args.out)
    log(f"saved {args.out}; total {time.perf_counter() - t0:.0f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
