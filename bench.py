#!/usr/bin/env python3
"""Throughput of the LSD-SLAM front end on one MI355X.

Metric: SE3 tracked frames/s (plus depth-map updates/s) at 640x480, batched vs
plain.  A step is one pass of the hot path over one batch of 64 synthetic
640x480 alignments (keyframe with ~25% semi-dense points, new frame at a 2 cm /
1 deg perturbation with sigma-2 noise, 5-level pyramid, levels 4..1, up to 20
LM iterations per level): every frame's pyramid is rebuilt from its resident
level-0 image and every frame is tracked.  `value` is the batched path with
FP32 pyramids (the reference's image precision; one persistent launch per
batch, hipGraph-replayed, in the shape the plan tuner picked); FP16 and BF16
pyramids, the plain path (one alignment per call) and the CPU baseline are
reported beside it, with every path's pose error against the FP64 oracle
over all 64 instances.  Inputs are resident in HBM when the timed region
starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import math
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


def _pmc_pass(counter, env, timeout_s=120):
    """One rocprofv3 counter pass over lsdg_prof (the headline workload through
    the C-ABI, after a float4 copy of 256 MiB): per-dispatch values of
    `counter` for the copy kernel and the last k_batch_coop dispatch."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = os.path.join(ROOT, "lsd-slam-pangolin-gui_amd", "lsdg_prof")
    if not shutil.which("rocprofv3") or not os.path.exists(prof):
        return None, "rocprofv3 or lsdg_prof not available"
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "c", "--", prof, "2", "256", "0"]
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=timeout_s)
        except subprocess.TimeoutExpired:
            return None, f"{counter} pass timed out"
        path = None
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    path = os.path.join(root, f)
        if r.returncode != 0 or path is None:
            return None, f"{counter} pass failed (exit {r.returncode})"
        copy, coop = [], []
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                if "copy_f4" in row["Kernel_Name"]:
                    copy.append(float(row["Counter_Value"]))
                elif "k_batch_coop" in row["Kernel_Name"]:
                    coop.append(float(row["Counter_Value"]))
        if not copy or not coop:
            return None, f"{counter} pass: kernels not found"
        return (float(np.median(copy)), coop[-1]), None


def measured_traffic(shape):
    """HBM bytes per launch of k_batch_coop, measured now: FETCH_SIZE and
    WRITE_SIZE passes (KiB per dispatch) over the same workload in the plan's
    shape, each scaled by what the counter reads for a copy kernel that reads
    and writes 256 MiB.  (bytes, description) or (None, reason)."""
    env = dict(os.environ, TMPDIR="/tmp", LSDG_BATCH_COOP="1", LSDG_BATCH_COOP_BS=str(shape["coop_block"]),
               LSDG_BATCH_COOP_K=str(shape["coop_k"]), LSDG_BATCH_SOLO=str(shape["solo_max"]),
               LSDG_BATCH_XCDS=str(shape["xcd_groups"]))
    f, why = _pmc_pass("FETCH_SIZE", env)
    if f is None:
        return None, why
    w, why = _pmc_pass("WRITE_SIZE", env)
    if w is None:
        return None, why
    mib256 = 256.0 * (1 << 20)
    fetch_scale = mib256 / (f[0] * 1024.0)
    write_scale = mib256 / (w[0] * 1024.0)
    traffic = f[1] * 1024.0 * fetch_scale + w[1] * 1024.0 * write_scale
    return traffic, (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over lsdg_prof 2 256 0 in the same shape: "
                     f"FETCH {f[1]:.0f} KiB x {fetch_scale:.3f}, WRITE {w[1]:.0f} KiB x {write_scale:.3f} "
                     f"(copy kernel of 256 MiB: FETCH {f[0]:.0f} KiB, WRITE {w[0]:.0f} KiB)")


def cpu_model():
    """The host CPU as /proc/cpuinfo names it (what lscpu prints as the model
    name), with the family / model numbers and the widest SIMD extension,
    since virtualised hosts often report only a generic name."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    info.setdefault(k.strip(), v.strip())
                if not line.strip() and info:
                    break
    except OSError:
        return "unknown"
    flags = info.get("flags", "").split()
    simd = "avx512f" if "avx512f" in flags else "avx2" if "avx2" in flags else "sse4_2" if "sse4_2" in flags else "?"
    return (f"{info.get('model name', 'unknown')} (family {info.get('cpu family', '?')}, model "
            f"{info.get('model', '?')}, widest SIMD {simd})")


def cpu_threads():
    """The host threads a CPU baseline may use: the process's CPU set, at most
    16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def tracking_params(its=20, crop=None):
    p = lsdg.default_params()
    for l in range(5):
        p.max_its_per_lvl[l] = its
    if crop is not None:
        p.max_epl_length_crop = float(crop)
    return p


class Tracking:
    """64 alignments on the device: plain calls or one batch plan."""

    def __init__(self, ctx, insts, precision):
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
        self.batch = None
        self.poses = [None] * len(self.items)

    def plain_step(self):
        for i, (kf, fr, ref) in enumerate(self.items):
            fr.invalidate()  # the tracker rebuilds the levels it reads, as the batch does
            self.poses[i] = self.ctx.se3_track(ref, fr, synth.IDENTITY7)[0]

    def jobs(self):
        return [lsdg.Job(ref.h, fr.h, lsdg.pose8(synth.IDENTITY7), lsdg.JOB_SE3, 4, 1, 1)
                for kf, fr, ref in self.items]

    def tune(self, runs=5):
        """Searches the plan shapes for these jobs (lsdg_batch_tune); later
        plans of this class take the winner."""
        best, default_ms, best_ms = self.ctx.batch_tune(self.jobs(), runs)
        return {"shape": best, "default_ms": round(default_ms, 4), "tuned_ms": round(best_ms, 4),
                "gain": round(default_ms / best_ms, 3)}

    def make_batch(self):
        self.batch = self.ctx.batch(self.jobs())

    def batch_step(self):
        """One batch run; the C call copies every job's pose and statistics
        into the result array (host memory) before it returns."""
        self.results = self.batch.run()
        return self.results

    def collect(self):
        self.poses = [np.array(r.frame_to_ref[:7]) for r in self.results]
        return self.poses


def sim3_leg(ctx, steps, warmup, pairs=16, per_pair=8, check_fp64=True):
    """Keyframe-to-keyframe Sim3 at 1920x1080 (~500k points per keyframe,
    scale drift U[0.9, 1.1], up to 10 cm / 5 deg, default LM settings): 16
    keyframe pairs x 8 jobs per batch, FP16 and BF16 pyramids, the FP16 pairs
    through the plain path beside them; each pair's pose, scale and final
    depth residual against the FP64 oracle, next to the FP32 oracle twin's."""
    ident = list(synth.IDENTITY7) + [1.0]
    data = [synth.sim3_instance(1005 + k) for k in range(pairs)]
    out = {"jobs_per_batch": pairs * per_pair}
    results = {}
    for name, prec in (("fp16", lsdg.FP16), ("bf16", lsdg.BF16)):
        items = []
        for k, d in enumerate(data):
            a = ctx.frame(2 * k, d["cam"], d["a_img"], prec)
            a.set_depth(d["a_idepth"], d["a_var"])
            b = ctx.frame(2 * k + 1, d["cam"], d["b_img"], prec)
            b.set_depth(d["b_idepth"], d["b_var"])
            items.append((a, b, ctx.ref(a)))
        for a, b, ref in items:
            for l in range(1, 5):
                ref.num_points(l)
        jobs = [lsdg.Job(ref.h, b.h, lsdg.pose8(ident), lsdg.JOB_SIM3, 4, 1, 0) for a, b, ref in items
                for _ in range(per_pair)]
        bt = ctx.batch(jobs)
        t = timed(ctx, bt.run, steps, warmup)
        out[f"batched_{name}_alignments_per_s"] = round(len(jobs) * steps / t, 1)
        res = bt.run()
        results[name] = [(list(res[per_pair * k].frame_to_ref), res[per_pair * k].stats) for k in range(pairs)]
        bt.destroy()
        if prec == lsdg.FP16:
            t = timed(ctx, lambda: [ctx.sim3_track(ref, b, ident, 4, 1) for a, b, ref in items], 1, 1)
            out["plain_fp16_alignments_per_s"] = round(pairs / t, 1)
        for a, b, ref in items:
            ref.destroy()
            b.destroy()
            a.destroy()
    if check_fp64:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        ref64, scenes = {}, []
        for prec in (1, 2):
            o = pyoracle.Oracle(prec)
            ref64[prec] = []
            for d in data:
                a = o.frame(0, d["cam"], d["a_img"].astype(np.float32))
                a.set_depth(d["a_idepth"], d["a_var"])
                b = o.frame(1, d["cam"], d["b_img"].astype(np.float32))
                b.set_depth(d["b_idepth"], d["b_var"])
                r = o.ref(a)
                q, st, _ = o.sim3_track(r, b, ident, 4, 1)
                ref64[prec].append((q, st))
                if prec == 1:
                    scenes.append((o, r, b, a))

        def mean_d(k, pose8):
            """The FP64 depth residual (mean over terms) at a pose, level 1."""
            o, r, b, _ = scenes[k]
            ev = o.sim3_evaluate(r, b, 1, pyoracle.sim3_inv(np.asarray(pose8, np.float64)))
            return ev.sum_res_d / ev.num_terms_d

        def errs(rows):
            rig, sc, dr = [], [], []
            for k, ((p, st), (q, qst)) in enumerate(zip(rows, ref64[1])):
                e = posemath.sim3_error(p, q)
                rig.append(e[0])
                sc.append(e[1])
                d64 = mean_d(k, q)
                dr.append(abs(mean_d(k, p) - d64) / d64)
            return {"pose_twist_max": float(np.max(rig)), "pose_twist_median": float(np.median(rig)),
                    "scale_max": float(np.max(sc)), "depth_residual_at_pose_rel_max": float(np.max(dr))}
        out["error_vs_fp64_oracle"] = {"fp16_batched": errs(results["fp16"]), "bf16_batched": errs(results["bf16"]),
                                       "fp32_oracle_twin": errs(ref64[2]), "pairs": pairs}
    return out


def adjoint_leg(ctx, steps, warmup, cpu_seconds):
    """Gradient of the Huber-weighted photometric energy (lsdg_energy_grad)
    at a 1280x960 keyframe with ~300k semi-dense points (level 0) against a
    frame at a 3 cm / 2 deg perturbation, FP32: evaluations per second of the
    C call (launch, reduction, copy of the ~300k per-point gradients to host
    memory), the adjoint kernel's device time and algorithmic bytes
    (B_eval + 4 N per launch); the oracle's FP32 restatement, one thread,
    beside it."""
    inst = synth.tracking_instance(1004, cam=synth.camera(1280, 960), trans_m=0.03, rot_deg=2.0, threshold=None,
                                   fraction=0.25)
    kf = ctx.frame(0, inst["cam"], inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    fr = ctx.frame(1, inst["cam"], inst["fr_img"])
    ref = ctx.ref(kf)
    n = ref.num_points(0)
    fr.build_pyramid()
    pose = list(synth.IDENTITY7)
    t = timed(ctx, lambda: ctx.energy_grad(ref, fr, 0, pose), steps, warmup)
    achieved, ks = kernel_roofline(ctx, lambda: [ctx.energy_grad(ref, fr, 0, pose) for _ in range(20)],
                                   lsdg.KERNEL_ADJOINT)
    out = {"gradient_evaluations_per_s": round(steps / t, 1), "points": n,
           "kernel_us": round(1e3 * ks["ms"] / max(ks["launches"], 1), 3), "kernel_gbs": round(achieved, 2),
           "kernel_bytes_per_launch": round(ks["bytes"] / max(ks["launches"], 1), 1),
           "kernel_frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 5)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(0, fast=True)
        okf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
        okf.set_depth(inst["idepth"], inst["var"])
        ofr = o.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
        oref = o.ref(okf)
        o.energy_grad(oref, ofr, 0, pose)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            o.energy_grad(oref, ofr, 0, pose)
            k += 1
        out["cpu_gradient_evaluations_per_s"] = round(k / (time.perf_counter() - t0), 2)
    for x in (ref, fr, kf):
        x.destroy()
    return out


def timed(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.sync()
    return time.perf_counter() - t0


def kernel_roofline(ctx, fn, kind, reps=1):
    ctx.profile(True)
    for _ in range(reps):
        fn()
    ks = ctx.kernel_stats(kind)
    ctx.profile(False)
    achieved = ks["bytes"] / (ks["ms"] * 1e-3) / 1e9 if ks["ms"] > 0 else 0.0
    return achieved, ks


def depth_leg(ctx, steps, warmup):
    """One DepthMap::updateKeyframe with one reference frame at a 5 cm
    baseline, 640x480, epipolar search up to 100 px; the prior state is
    restored (untimed) before every update."""
    d = synth.depth_instance()
    cam = d["cam"]
    ctx.set_params(tracking_params(20, crop=100))
    kf = ctx.frame(0, cam, d["kf_img"])
    ref = ctx.frame(1, cam, d["ref_img"])
    ref.set_tracking(kf, list(d["ref_pose"]) + [1.0], 0.4)
    dm = ctx.depthmap(cam)
    dm.set_prior(kf, d["prior_idepth"], d["prior_var"], 20)
    state = dm.read()
    times = []
    for k in range(warmup + steps):
        dm.write(state)
        ctx.sync()
        t0 = time.perf_counter()
        dm.update([ref])
        ctx.sync()
        if k >= warmup:
            times.append(time.perf_counter() - t0)
    def restored_update():
        # the restore is a copy, not one of the update's kernels: only the
        # update is inside the library's event timing
        dm.write(state)
        dm.update([ref])

    # ten updates from the same prior state, so the device time is an average
    # rather than one sample
    achieved, ks = kernel_roofline(ctx, restored_update, lsdg.KERNEL_DEPTH, reps=10)
    out = dict(updates_per_s=len(times) / sum(times), ms_per_update=1e3 * float(np.median(times)),
               roofline_gbs=round(achieved, 2), device_ms_per_update=ks["ms"] / max(ks["launches"], 1))
    return out, d


def export_leg(ctx, steps, warmup, cpu_seconds):
    """Keyframe export for display at 640x480: publishKeyframe's records and
    Keyframe::computeVbo's filter (lib/Pangolin_IOWrapper/) on the device,
    results copied to host memory; the oracle's restatement of the same two
    loops, one thread, beside it."""
    d = synth.depth_instance()
    rng = np.random.default_rng(11)
    idp = np.where(d["prior_idepth"] > 0, 1.0 / d["kf_depth"], -1).astype(np.float32)
    var = np.where(idp > 0, rng.uniform(1e-7, 3e-6, idp.shape), -1).astype(np.float32)
    kf = ctx.frame(0, d["cam"], d["kf_img"])
    kf.set_depth(idp, var)
    p = lsdg.vbo_params(1.0)
    for _ in range(warmup):
        kf.keyframe_vbo(0, p)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        v = kf.keyframe_vbo(0, p)
    t = (time.perf_counter() - t0) / steps
    out = {"keyframe_vbo_ms": round(1e3 * t, 4), "keyframe_vbo_vertices": int(len(v))}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(0, fast=True)
        of = o.frame(0, d["cam"], d["kf_img"].astype(np.float32))
        of.set_depth(idp, var)
        w, h = of.size(0)
        cam4 = of.camera(0)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            o.compute_vbo(of.keyframe_points(0), w, h, cam4, p)
            n += 1
        out["cpu_keyframe_vbo_ms"] = round(1e3 * (time.perf_counter() - t0) / n, 4)
    return out


def ingest_leg(ctx, steps, warmup, cpu_seconds):
    """InputThread's ingest of one 640x480 8-bit camera image (FOV lens,
    omega 0.92) into a 640x480 pinhole frame: upload, undistortion, float
    conversion and image pyramid in one launch, into a reused frame and into
    a new frame per image (as upstream allocates); the oracle's undistort +
    Frame::buildImage beside it."""
    in_k, out_k, omega = (330.0, 330.0, 319.5, 239.5), (270.0, 270.0, 319.5, 239.5), 0.92
    cam = dict(w=640, h=480, fx=out_k[0], fy=out_k[1], cx=out_k[2], cy=out_k[3])
    src, _ = synth.Scene(21).render(synth.IDENTITY7, synth.camera(640, 480), 2.0, 21)
    table = lsdg.fov_undistort_table(in_k, omega, (640, 480), out_k, (640, 480))
    und = ctx.undistorter((640, 480), table)
    for _ in range(warmup):
        ctx.frame_undistorted(0, cam, und, src).destroy()
    ctx.sync()
    t_new = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        f = ctx.frame_undistorted(0, cam, und, src)
        t_new += time.perf_counter() - t0
        f.destroy()
    # a stream of images into one reused frame (no allocation per image)
    f = ctx.frame_undistorted(0, cam, und, src)
    t0 = time.perf_counter()
    for k in range(steps):
        f.ingest(k, und, src)
    t_reuse = time.perf_counter() - t0
    achieved, ks = kernel_roofline(ctx, lambda: [f.ingest(k, und, src) for k in range(20)], lsdg.KERNEL_INGEST)
    f.destroy()
    und.close()
    out = {"ingest_call_ms": round(1e3 * t_reuse / steps, 4), "ingest_new_frame_call_ms": round(1e3 * t_new / steps, 4),
           "ingest_kernel_us": round(1e3 * ks["ms"] / max(ks["launches"], 1), 3),
           "ingest_kernel_gbs": round(achieved, 2)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(0, fast=True)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            of = o.frame(0, cam, o.undistort(src, table).astype(np.float32))
            of.read(lsdg.WHAT_IMAGE, 4)
            n += 1
        out["cpu_ingest_ms"] = round(1e3 * (time.perf_counter() - t0) / n, 4)
    return out


def constraint_leg(ctx, steps, warmup, cpu_seconds, pairs=16):
    """SlamSystem::testConstraint for 16 keyframe pairs of 640x480 (a third of
    them of unrelated scenes), strictness 1.5: three stages of batched Sim3
    alignments in both directions plus the consistency test; the oracle's
    sequential restatement (FP32, timing build, one thread) beside it."""
    init = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    cam = synth.camera(640, 480)
    data = []
    for k in range(pairs):
        d = synth.sim3_instance(2000 + k, cam=cam, max_trans=0.04, max_rot_deg=2.0, scale_range=(0.95, 1.05),
                                fraction=0.3)
        if k % 3 == 2:
            e = synth.sim3_instance(2500 + k, cam=cam, fraction=0.3)
            d = dict(d, b_img=e["b_img"], b_idepth=e["b_idepth"], b_var=e["b_var"])
        data.append(d)
    items = []
    for k, d in enumerate(data):
        a = ctx.frame(2 * k, cam, d["a_img"], lsdg.FP16)
        a.set_depth(d["a_idepth"], d["a_var"])
        b = ctx.frame(2 * k + 1, cam, d["b_img"], lsdg.FP16)
        b.set_depth(d["b_idepth"], d["b_var"])
        items.append((a, b, ctx.ref(a), ctx.ref(b)))
    cands = [(ra, rb, init) for a, b, ra, rb in items]
    t = timed(ctx, lambda: ctx.constraint_search(cands, 1.5), steps, warmup)
    res = ctx.constraint_search(cands, 1.5)
    out = {"pairs_per_s": round(pairs * steps / t, 1), "pairs": pairs,
           "accepted": int(sum(r.accepted for r in res))}
    for a, b, ra, rb in items:
        for x in (ra, rb, a, b):
            x.destroy()
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(0, fast=True)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            d = data[n % pairs]
            a = o.frame(0, cam, d["a_img"].astype(np.float32))
            a.set_depth(d["a_idepth"], d["a_var"])
            b = o.frame(1, cam, d["b_img"].astype(np.float32))
            b.set_depth(d["b_idepth"], d["b_var"])
            o.constraint_test(o.ref(a), o.ref(b), init, 1.5)
            n += 1
        out["cpu_pairs_per_s"] = round(n / (time.perf_counter() - t0), 2)
    return out


def system_leg(ctx, cpu_seconds, n=60):
    """lsd-slam's SlamSystem front end (trackFrame + the mapping it waits
    for) over a 60-frame 640x480 synthetic sequence (1.3 cm and 0.15 deg per
    frame), FP16 pyramids, GT-depth initialisation untimed; the oracle's
    restatement (FP32, timing build, one thread) on the first frames beside it."""
    d = synth.vo_sequence(7, n, cam=synth.camera(640, 480))
    s = ctx.system(d["cam"], lsdg.FP16)
    s.init_gt(0, d["images"][0], d["idepth"], d["var"])
    ctx.sync()
    t0 = time.perf_counter()
    infos = [s.track(k, d["images"][k]) for k in range(1, n)]
    t = time.perf_counter() - t0
    err = max(float(np.linalg.norm(np.array(i.cam_to_world[4:7]) - d["poses"][k + 1][4:7]))
              for k, i in enumerate(infos))
    ids, _ = s.keyframes()
    s.destroy()
    out = {"frames_per_s": round((n - 1) / t, 1), "frames": n - 1, "keyframes": len(ids),
           "max_translation_error_m": round(err, 5)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.OSystem(pyoracle.Oracle(0, fast=True), d["cam"])
        o.init_gt(0, d["images"][0], d["idepth"], d["var"])
        k, t0 = 1, time.perf_counter()
        while k < n and time.perf_counter() - t0 < cpu_seconds:
            o.track(k, d["images"][k])
            k += 1
        out["cpu_frames_per_s"] = round((k - 1) / (time.perf_counter() - t0), 2)
    return out


def cpu_tracking_baseline(insts, params, seconds_per_run, runs=5):
    """The oracle's FP32 restatement of the reference CPU path (timing build,
    upstream FP32 sums) on the same 64 alignments, per-frame pyramid build
    included as on the GPU: one thread, and one alignment per thread on every
    host thread this process may use; each figure the median of `runs` runs
    of about `seconds_per_run` after a warm-up run."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    def workers(nthreads):
        ws = []
        for t in range(nthreads):
            o = pyoracle.Oracle(0, params=params, fast=True)
            pairs = []
            for inst in insts[t::nthreads] or insts:
                kf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
                kf.set_depth(inst["idepth"], inst["var"])
                ref = o.ref(kf)
                for l in range(1, 5):
                    ref.num_points(l)
                pairs.append((kf, ref, inst["cam"], inst["fr_img"].astype(np.float32)))
            ws.append((o, pairs))
        return ws

    def rate(ws, seconds):
        counts = [0] * len(ws)
        stop = time.perf_counter() + seconds

        def work(k):
            o, pairs = ws[k]
            n = 0
            while time.perf_counter() < stop:
                kf, ref, cam, img = pairs[n % len(pairs)]
                o.se3_track(ref, o.frame(100000 + n, cam, img), synth.IDENTITY7)
                n += 1
            counts[k] = n

        th = [threading.Thread(target=work, args=(k,)) for k in range(len(ws))]
        t0 = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        return sum(counts) / (time.perf_counter() - t0), sum(counts)

    out = {}
    for nthreads in sorted({1, cpu_threads()}):
        ws = workers(nthreads)
        rate(ws, min(1.0, seconds_per_run))
        rs = [rate(ws, seconds_per_run) for _ in range(runs)]
        out[nthreads] = (float(np.median([r[0] for r in rs])), sum(r[1] for r in rs))
    n_max = max(out)
    return dict(value=round(out[n_max][0], 2), unit="frames/s", cores=n_max, kind="port",
                single_thread_value=round(out[1][0], 2), cpu_model=cpu_model(),
                sample=f"plain SE3 alignments of the 64 640x480 instances (oracle FP32 restatement of the "
                       f"reference CPU path, -O3 -march=x86-64-v3 where the reference's Release build uses "
                       f"-march=native -DENABLE_SSE, per-frame pyramid build included): median of {runs} "
                       f"runs of {seconds_per_run:.1f} s each on {n_max} threads ({out[n_max][1]} alignments) and "
                       f"on 1 thread ({out[1][1]} alignments)")


def cpu_depth_baseline(d, params, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(0, params=params, fast=True)
    cam = d["cam"]
    kf = o.frame(0, cam, d["kf_img"].astype(np.float32))
    ref = o.frame(1, cam, d["ref_img"].astype(np.float32))
    ref.set_tracking(kf, list(d["ref_pose"]) + [1.0], 0.4)
    out = {}
    for threads in sorted({1, cpu_threads()}):
        dm = o.depthmap(cam, threads=threads)
        n, t = 0, 0.0
        while t < seconds / 2:
            dm.set_prior(kf, d["prior_idepth"], d["prior_var"], 20)
            t0 = time.perf_counter()
            dm.update([ref])
            t += time.perf_counter() - t0
            n += 1
        out[threads] = n / t
    return out


def oracle_poses(insts, params, prec=1, init_eps=0.0, seed=0):
    """Every instance's pose from the oracle (prec 1: FP64 throughout),
    tracked from the identity or, with init_eps > 0, from an initial pose
    perturbed by init_eps in translation (m) and rotation (rad)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(prec, params=params)
    out = []
    for k, inst in enumerate(insts):
        kf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
        kf.set_depth(inst["idepth"], inst["var"])
        fr = o.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
        init = synth.IDENTITY7 if init_eps <= 0 else list(synth.sample_pose(seed + k, init_eps, math.degrees(init_eps)))
        out.append(o.se3_track(o.ref(kf), fr, init)[0])
    return out


def fp64_self_spread(insts, params, ref64, eps=1e-9, seeds=(11, 12, 13)):
    """How far the FP64 oracle's own result moves when its initial pose is
    perturbed by eps (1e-9 m / 1e-9 rad): per instance, the largest twist
    between the pose tracked from the identity and from each perturbed start.
    The LM's accept / reject and stop decisions and the in-image / isGood
    sets are discontinuous, so on some instances a 1e-9 perturbation of the
    start moves the FP64 answer by 1e-5 or more: no finite-precision path can
    be held closer to the FP64 oracle than this spread."""
    spread = np.zeros(len(insts))
    for sd in seeds:
        alt = oracle_poses(insts, params, 1, eps, 1000 * sd)
        spread = np.maximum(spread, [posemath.twist_error(p, q) for p, q in zip(alt, ref64)])
    return spread


def within_spread(poses, ref64, spread, floor=1e-5):
    e = np.array([posemath.twist_error(p, q) for p, q in zip(poses, ref64)])
    return int((e <= np.maximum(spread, floor)).sum())


def error_stats(poses, ref):
    e = np.array([posemath.twist_error(p, q) for p, q in zip(poses, ref)])
    return {"max": float(e.max()), "median": float(np.median(e)), "p90": float(np.quantile(e, 0.9)),
            "above_1e-5": int((e > 1e-5).sum()), "instances": int(len(e)), "argmax": int(e.argmax())}


def converged_params(its=100):
    """The LM run to its minimum: the error-ratio stop off (eps 1), steps
    down to 1e-6 in twist norm, up to `its` iterations per level."""
    p = tracking_params(its)
    for l in range(5):
        p.convergence_eps[l] = 1.0
        p.step_size_min[l] = 1e-12
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="plans in their default shape")
    ap.add_argument("--no-traffic", action="store_true", help="skip the FETCH_SIZE / WRITE_SIZE counter passes")
    ap.add_argument("--headline-only", action="store_true",
                    help="run only the batched FP32 leg (for kernel profiling)")
    ap.add_argument("--precision", choices=("fp32", "fp16", "bf16"), default="fp32",
                    help="pyramid storage of the --headline-only leg")
    ap.add_argument("--depth-only", action="store_true",
                    help="run only the depth-map update leg (for kernel profiling)")
    ap.add_argument("--ingest-only", action="store_true",
                    help="run only the ingest leg (for kernel profiling)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg (for kernel profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="also run a plain device copy of MB MiB (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note, not a measured path)"}))
        sys.exit(2)

    params = tracking_params(20)
    ctx = lsdg.Context(0, params)
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    if args.depth_only:
        depth, _ = depth_leg(ctx, args.steps, args.warmup)
        print(json.dumps({"depth_updates_per_s": round(depth["updates_per_s"], 2),
                          "depth_update_ms": round(depth["ms_per_update"], 4),
                          "depth_device_ms_per_update": round(depth["device_ms_per_update"], 4),
                          "depth_roofline_gbs": depth["roofline_gbs"]}))
        return
    if args.ingest_only:
        print(json.dumps(ingest_leg(ctx, args.steps, args.warmup, 0.0)))
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(ctx, args.steps, args.warmup, 0.0)))
        return
    insts = [synth.tracking_instance(1001 + 17 * i) for i in range(args.batch)]
    B = args.batch
    prec_of = {"fp32": lsdg.FP32, "fp16": lsdg.FP16, "bf16": lsdg.BF16}

    def batched(precision, steps, warmup):
        """The 64-alignment batch in the shape the tuner picks (tuned outside
        the timed region): frames/s, the plan's shape, the poses."""
        tr = Tracking(ctx, insts, precision)
        tune = None if args.no_tune else tr.tune(runs=5)
        tr.make_batch()
        t = timed(ctx, tr.batch_step, steps, warmup)
        tr.batch_step()
        return tr, B * steps / t, 1e3 * t / steps, tune

    # headline: batched, FP32 pyramids (or the --precision of --headline-only)
    head_prec = prec_of[args.precision] if args.headline_only else lsdg.FP32
    t32, batched32_fps, ms_per_step, tune32 = batched(head_prec, args.steps, args.warmup)
    poses32 = list(t32.collect())
    # HIP events around every launch, over a profiled repeat of the timed steps
    achieved, ks = kernel_roofline(ctx, t32.batch_step, lsdg.KERNEL_BATCH_EVAL, reps=min(args.steps, 20))
    dom = ctx.kernel_stats(lsdg.KERNEL_BATCH_COOP)
    shape = t32.batch.config()
    dom_name = "k_batch_coop"
    dom_desc = (f"k_batch_coop (every job's whole LM loop, {shape['coop_k']} blocks of {shape['coop_block']} threads "
                "per job: fused residual + weights + LGS, block reductions, row exchange, LM step; FP32 taps)")
    if dom["launches"] == 0:
        dom, dom_name = ks, "k_batch_eval"
        dom_desc = "k_batch_eval (fused residual + weights + LGS over all jobs' tiles, FP32 taps)"
    dom_gbs = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
    if args.headline_only:
        print(json.dumps({"value": round(batched32_fps, 2), "unit": "frames/s", "precision": args.precision,
                          "ms_per_step": round(ms_per_step, 3), "shape": shape, "tune": tune32,
                          "dominant_kernel": dom_name, "dominant_avg_launch_us":
                          round(1e3 * dom["ms"] / max(dom["launches"], 1), 3), "dominant_gbs": round(dom_gbs, 2),
                          "dominant_bytes_per_launch": round(dom["bytes"] / max(dom["launches"], 1), 1)}))
        return
    if dom_name == "k_batch_coop" and not args.no_traffic:
        traffic, traffic_src = measured_traffic(shape)
    else:
        traffic, traffic_src = None, "not measured (--no-traffic or not the cooperative schedule)"

    # reduced-precision pyramids beside it
    t16, batched16_fps, _, tune16 = batched(lsdg.FP16, args.steps, args.warmup)
    poses16 = list(t16.collect())
    tb16, batchedbf_fps, _, tunebf = batched(lsdg.BF16, args.steps, args.warmup)
    posesbf = list(tb16.collect())

    # compensated FP32 sums (Kahan per thread, TwoSum trees, FP64 across
    # blocks) beside the FP64 per-thread default: FP32 pyramids, same batch
    ctx.set_sums(lsdg.SUMS_KAHAN)
    try:
        tk, batchedk_fps, _, tunek = batched(lsdg.FP32, args.steps, args.warmup)
        posesk = list(tk.collect())
        kahan = {"batched_fp32_frames_per_s": round(batchedk_fps, 2),
                 "vs_fp64_sums": round(batchedk_fps / batched32_fps, 3), "tune": tunek,
                 "max_twist_vs_fp64_sums": max(posemath.twist_error(a, b) for a, b in zip(posesk, poses32))}
        if not args.no_plain:
            _, k_ks = kernel_roofline(ctx, tk.plain_step, lsdg.KERNEL_EVAL)
            kahan["plain_eval_avg_launch_us"] = round(1e3 * k_ks["ms"] / max(k_ks["launches"], 1), 3)
    finally:
        ctx.set_sums(lsdg.SUMS_FP64)

    out_plain = {}
    if not args.no_plain:
        t = timed(ctx, t32.plain_step, max(1, args.steps // 2), 1)
        plain_fps = B * max(1, args.steps // 2) / t
        plain_ach, plain_ks = kernel_roofline(ctx, t32.plain_step, lsdg.KERNEL_EVAL)
        out_plain = {"plain_fp32_frames_per_s": round(plain_fps, 2),
                     "batched_fp32_vs_plain_speedup": round(batched32_fps / plain_fps, 1),
                     "batched_fp16_vs_plain_speedup": round(batched16_fps / plain_fps, 1),
                     "plain_eval_kernel_gbs": round(plain_ach, 2),
                     "plain_eval_avg_launch_us": round(1e3 * plain_ks["ms"] / max(plain_ks["launches"], 1), 3),
                     "max_twist_batched_fp32_vs_plain": max(posemath.twist_error(a, b)
                                                            for a, b in zip(poses32, t32.poses))}

    # converged variant (error-ratio stop off): FP32 and FP16 batches
    conv = {}
    ctx.set_params(converged_params())
    try:
        for name, prec in (("fp32", lsdg.FP32), ("fp16", lsdg.FP16)):
            tr = Tracking(ctx, insts, prec)
            tr.make_batch()
            tr.batch_step()
            conv[name] = list(tr.collect())
    finally:
        ctx.set_params(params)

    # mixed-resolution batch (320x240 / 640x480 / 1280x960, up to 5 cm / 3 deg), FP32
    mixed = synth.batch_instances(64, seed=1003)
    tm = Tracking(ctx, mixed, lsdg.FP32)
    tune_mixed = None if args.no_tune else tm.tune(runs=5)
    tm.make_batch()
    t = timed(ctx, tm.batch_step, args.steps, args.warmup)
    mixed_aps = 64 * args.steps / t

    sim3 = sim3_leg(ctx, args.steps, args.warmup, check_fp64=not args.no_cpu)
    adjoint = adjoint_leg(ctx, max(args.steps, 20), args.warmup, 0.0 if args.no_cpu else 2.0)

    depth, dinst = depth_leg(ctx, args.steps, args.warmup)
    ctx.set_params(params)
    export = export_leg(ctx, 20, 3, 0.0 if args.no_cpu else 1.0)
    ingest = ingest_leg(ctx, 20, 3, 0.0 if args.no_cpu else 1.0)
    constraint = constraint_leg(ctx, 3, 1, 0.0 if args.no_cpu else 4.0)
    system = system_leg(ctx, 0.0 if args.no_cpu else 6.0)

    out = {
        "metric": "SE3 track frames/s + depth updates/s @640x480, batched vs plain, 1 GPU",
        "value": round(batched32_fps, 2),
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
        "config": {"workload": f"batched SE3 alignment: {B} x 640x480 frames per step (pyramid rebuild + LM "
                               "levels 4..1, <=20 iterations/level), ~25% semi-dense points, 2 cm / 1 deg, "
                               "FP32 (I, gx, gy) pyramids, FP32 residuals, FP64 sums, FP64 merge and LDLT",
                   "batch": B, "resolution": "640x480", "storage": "fp32"},
        "plan_shape": shape,
        "tune_fp32": tune32,
        "batched_fp16_frames_per_s": round(batched16_fps, 2),
        "batched_bf16_frames_per_s": round(batchedbf_fps, 2),
        "tune_fp16": tune16,
        "tune_bf16": tunebf,
        "sums_kahan": kahan,
        **out_plain,
        "mixed_resolution_batch_alignments_per_s": round(mixed_aps, 2),
        "tune_mixed_resolution": tune_mixed,
        "sim3_1920x1080": sim3,
        "adjoint_1280x960": adjoint,
        "depth_updates_per_s": round(depth["updates_per_s"], 2),
        "depth_update_ms": round(depth["ms_per_update"], 4),
        "depth_device_ms_per_update": round(depth["device_ms_per_update"], 4),
        "depth_roofline_gbs": depth["roofline_gbs"],
        "keyframe_export_640x480": export,
        "ingest_640x480": ingest,
        "constraint_search_640x480": constraint,
        "slam_system_640x480": system,
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 5),
                     "traffic": None if traffic is None else round(traffic, 0),
                     "traffic_over_algorithmic": None if traffic is None else
                     round(traffic / max(dom["bytes"] / max(dom["launches"], 1), 1.0), 3),
                     "traffic_source": traffic_src, "kernel": dom_desc,
                     "avg_launch_us": round(1e3 * dom["ms"] / max(dom["launches"], 1), 3),
                     "launches": dom["launches"],
                     "bytes_per_launch": round(dom["bytes"] / max(dom["launches"], 1), 1)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_tracking_baseline(insts, params, args.cpu_seconds * 0.1)
        cd = cpu_depth_baseline(dinst, tracking_params(20, crop=100), args.cpu_seconds * 0.4)
        out["cpu_depth_updates_per_s"] = {str(k): round(v, 2) for k, v in cd.items()}
        ref64 = oracle_poses(insts, params)
        ref64c = oracle_poses(insts, converged_params())
        spread = fp64_self_spread(insts, params, ref64)
        out["pose_error_vs_fp64_oracle"] = {
            "fp32_batched_and_plain": error_stats(poses32, ref64),
            "fp16_batched": error_stats(poses16, ref64),
            "bf16_batched": error_stats(posesbf, ref64),
            "kahan_fp32_batched": error_stats(posesk, ref64),
            "converged_fp32": error_stats(conv["fp32"], ref64c),
            "converged_fp16": error_stats(conv["fp16"], ref64c),
            "fp64_self_spread": {"max": float(spread.max()), "median": float(np.median(spread)),
                                 "above_1e-5": int((spread > 1e-5).sum()),
                                 "fp32_within_max_1e-5_spread": within_spread(poses32, ref64, spread),
                                 "fp16_within_max_1e-5_spread": within_spread(poses16, ref64, spread),
                                 "bf16_within_max_1e-5_spread": within_spread(posesbf, ref64, spread),
                                 "kahan_within_max_1e-5_spread": within_spread(posesk, ref64, spread)},
            "note": "twist norm of every instance's frameToRef against the FP64 oracle's; the FP32 batch equals the "
                    "plain path to 1e-16 and the oracle's FP32 twin to 1e-9; fp64_self_spread is how far the FP64 "
                    "oracle itself moves when its start pose is perturbed by 1e-9 (3 seeds), the floor under "
                    "any path's error against it"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
