#!/usr/bin/env python3
"""Throughput of the LSD-SLAM front end on one MI355X.

Metric: SE3 tracked frames/s (and depth-map updates/s) at 640x480, batched vs
plain.  A step is one pass of the hot path over one batch of synthetic
640x480 alignments (keyframe with ~25% semi-dense points, new frame at a
2 cm / 1 deg perturbation with sigma-2 noise, 5-level pyramid, levels 4..1,
up to 20 LM iterations per level): every frame's pyramid is rebuilt from its
resident level-0 image and the frame is tracked.  Inputs are resident in HBM
when the timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lsd-slam-pangolin-gui_amd"))

import numpy as np  # noqa: E402

import lsdg  # noqa: E402
import posemath  # noqa: E402
import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0


def tracking_params(its=20):
    p = lsdg.default_params()
    for l in range(5):
        p.max_its_per_lvl[l] = its
    return p


def make_instances(n, seed0=1001):
    return [synth.tracking_instance(seed0 + 17 * i) for i in range(n)]


class PlainPath:
    """One alignment per call: lsdg_frame_build_pyramid + lsdg_se3_track."""

    def __init__(self, ctx, insts, precision=lsdg.FP32):
        self.ctx = ctx
        self.items = []
        for i, inst in enumerate(insts):
            kf = ctx.frame(2 * i, inst["cam"], inst["kf_img"], precision)
            kf.set_depth(inst["idepth"], inst["var"])
            fr = ctx.frame(2 * i + 1, inst["cam"], inst["fr_img"], precision)
            ref = ctx.ref(kf)
            for l in range(1, 5):
                ref.num_points(l)  # point clouds are per keyframe, built outside the timed region
            self.items.append((kf, fr, ref))
        self.poses = [None] * len(self.items)

    def step(self):
        for i, (kf, fr, ref) in enumerate(self.items):
            fr.build_pyramid()
            pose, st, status = self.ctx.se3_track(ref, fr, synth.IDENTITY7)
            self.poses[i] = pose


def timed(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.sync()
    return time.perf_counter() - t0


def cpu_baseline(insts, params, seconds, gpu_poses):
    """The oracle's FP32 restatement of the reference CPU path (timing build),
    one thread, over a bounded sample of the same alignments; plus the FP64
    oracle's poses for the pose-error report."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(0, params=params, fast=True)
    pairs = []
    for i, inst in enumerate(insts):
        kf = o.frame(2 * i, inst["cam"], inst["kf_img"].astype(np.float32))
        kf.set_depth(inst["idepth"], inst["var"])
        ref = o.ref(kf)
        for l in range(1, 5):
            ref.num_points(l)
        pairs.append((kf, ref, inst))
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        kf, ref, inst = pairs[n % len(pairs)]
        fr = o.frame(10000 + n, inst["cam"], inst["fr_img"].astype(np.float32))  # pyramid build is per frame
        o.se3_track(ref, fr, synth.IDENTITY7)
        n += 1
    dt = time.perf_counter() - t0
    o64 = pyoracle.Oracle(1, params=params)
    errs = []
    for i in range(min(4, len(insts))):
        inst = insts[i]
        kf = o64.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
        kf.set_depth(inst["idepth"], inst["var"])
        fr = o64.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
        pose64 = o64.se3_track(o64.ref(kf), fr, synth.IDENTITY7)[0]
        errs.append(posemath.twist_error(gpu_poses[i], pose64))
    return dict(value=n / dt, unit="frames/s", cores=1, kind="port",
                sample=f"{n} plain SE3 alignments at 640x480 (FP32 restatement, -O3 x86-64-v3, 1 thread, "
                       f"{dt:.1f} s)"), max(errs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note, not a measured path)"}))
        sys.exit(2)

    params = tracking_params(20)
    ctx = lsdg.Context(0, params)
    insts = make_instances(args.batch)

    plain = PlainPath(ctx, insts)
    t = timed(ctx, plain.step, args.steps, args.warmup)
    plain_fps = args.batch * args.steps / t

    # roofline of the dominant kernel (fused evaluation), one profiled step
    ctx.profile(True)
    plain.step()
    ks = ctx.kernel_stats(lsdg.KERNEL_EVAL)
    ctx.profile(False)
    achieved = ks["bytes"] / (ks["ms"] * 1e-3) / 1e9

    value, ms_per_step = plain_fps, 1e3 * t / args.steps
    out = {
        "metric": "SE3 track frames/s + depth updates/s @640x480, batched vs plain, 1 GPU",
        "value": round(value, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain SE3 alignment, {args.batch} x 640x480 frames per step, ~25% semi-dense "
                               "points, 5-level pyramid, levels 4..1, <=20 LM iterations/level",
                   "batch": args.batch, "resolution": "640x480", "precision": "fp32 pyramids"},
        "plain_frames_per_s": round(plain_fps, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "k_se3_eval (fused residual + weights + LGS6)",
                     "avg_launch_us": round(1e3 * ks["ms"] / max(ks["launches"], 1), 3),
                     "launches": ks["launches"]},
    }
    if not args.no_cpu:
        cb, err = cpu_baseline(insts, params, args.cpu_seconds, plain.poses)
        out["cpu_baseline"] = cb
        out["pose_error_vs_fp64_oracle"] = err
    print(json.dumps(out))


if __name__ == "__main__":
    main()
