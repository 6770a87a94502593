"""Edge cases of the tracking and depth paths, against the oracle or the plain
GPU path:

* a keyframe whose depth has no valid pixel (no points at any level):
  SE3Tracker::trackFrame diverges at the first level; plain, batched (both
  schedules) and the oracle agree, and the batch's other jobs are untouched;
* the smallest frame the C-ABI takes (64x64, 4x4 at the coarsest level):
  plain tracking within 1e-9 twist of the oracle's FP64-sum twin with the
  same evaluation counts, the batch equal to the plain path, and a depth
  update bit-exact against the FP32 restatement;
* a batch of one job, and a batch of more jobs than the cooperative kernel
  holds at once (the plan takes the step schedule), each job equal to its
  plain alignment; one job more than a batch holds is refused;
* a depth update whose reference frame is the keyframe itself (zero
  baseline: every epipolar line is degenerate), bit-exact against the FP32
  restatement;
* a depth update on a ragged 200x150 map (partial gate blocks and fill
  tiles on the right and bottom edges, a tile count that does not divide
  over the XCDs), bit-exact against the FP32 restatement."""
import ctypes as C

import numpy as np
import pytest

import lsdg
import pyoracle
import synth
from helpers import tracking_params, twist

pytestmark = pytest.mark.gpu

IDENT8 = list(synth.IDENTITY7) + [1.0]
DEPTH_FIELDS = ["is_valid", "blacklisted", "next_stereo_frame_min_id", "validity_counter", "idepth", "idepth_var",
                "idepth_smoothed", "idepth_var_smoothed", "observe_code"]


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _depth_diff(a, b):
    return {k: int((_bits(a[k]) != _bits(b[k])).sum()) for k in DEPTH_FIELDS if (_bits(a[k]) != _bits(b[k])).any()}


def _gpu_pair(ctx, inst, fid=0, idepth=None, var=None):
    kf = ctx.frame(2 * fid, inst["cam"], inst["kf_img"])
    kf.set_depth(inst["idepth"] if idepth is None else idepth, inst["var"] if var is None else var)
    fr = ctx.frame(2 * fid + 1, inst["cam"], inst["fr_img"])
    return kf, fr, ctx.ref(kf)


def _oracle_track(inst, p, idepth=None, var=None):
    o = pyoracle.Oracle(2, params=p)
    kf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
    kf.set_depth(inst["idepth"] if idepth is None else idepth, inst["var"] if var is None else var)
    fr = o.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
    return o.se3_track(o.ref(kf), fr, synth.IDENTITY7)


def _batch(ctx, items, env, monkeypatch):
    for var, val in env.items():
        monkeypatch.setenv(var, val)
    try:
        b = ctx.batch([lsdg.Job(ref.h, fr.h, lsdg.pose8(IDENT8), lsdg.JOB_SE3, 4, 1, 1) for kf, fr, ref in items])
        cfg = b.config()
        r = b.run()
        b.destroy()
    finally:
        for var in env:
            monkeypatch.delenv(var, raising=False)
    return cfg, r


SCHEDULES = {"coop": {"LSDG_BATCH_COOP": "1", "LSDG_BATCH_SOLO": "0"}, "steps": {"LSDG_BATCH_COOP": "0"}}


@pytest.mark.parametrize("schedule", list(SCHEDULES))
def test_keyframe_without_points(ctx, monkeypatch, schedule):
    p = tracking_params(20)
    ctx.set_params(p)
    try:
        inst = synth.tracking_instance(1201, cam=synth.camera(320, 240))
        none = np.full_like(inst["idepth"], -1.0)
        empty = _gpu_pair(ctx, inst, 0, none, none)
        assert all(empty[2].num_points(lv) == 0 for lv in range(5))
        pose, st, status = ctx.se3_track(empty[2], empty[1], synth.IDENTITY7)
        opose, ost, ostatus = _oracle_track(inst, p, none, none)
        assert status == ostatus == lsdg.DIVERGED
        assert np.array_equal(pose, opose)
        assert list(st.num_residual_calls) == list(ost.num_residual_calls)
        assert st.tracking_was_good == ost.tracking_was_good == 0
        good = _gpu_pair(ctx, inst, 1)
        gpose, _, gstatus = ctx.se3_track(good[2], good[1], synth.IDENTITY7)
        cfg, r = _batch(ctx, [empty, good, empty], SCHEDULES[schedule], monkeypatch)
        assert cfg["schedule"] == (lsdg.SCHEDULE_COOP if schedule == "coop" else lsdg.SCHEDULE_STEPS)
        for k in (0, 2):
            assert r[k].status == status and np.array_equal(np.array(r[k].frame_to_ref[:7]), pose), k
        assert r[1].status == gstatus and twist(r[1].frame_to_ref[:7], gpose) <= 1e-9
    finally:
        ctx.set_params(lsdg.default_params())


def test_smallest_frame_tracks_like_the_oracle(ctx, monkeypatch):
    p = tracking_params(20)
    ctx.set_params(p)
    try:
        cam = synth.camera(64, 64)
        L = ctx.L
        out = C.c_void_p()
        small = lsdg.Camera(48, 64, 40.0, 40.0, 24.0, 32.0)  # 3 px wide at the coarsest level: refused
        assert L.lsdg_frame_create(ctx.h, 0, C.byref(small), np.zeros((64, 48), np.float32).ctypes.data, lsdg.FP32,
                                   C.byref(out)) == lsdg.EINVAL
        items, plain = [], []
        for k, seed in enumerate((1211, 1212, 1213)):
            inst = synth.tracking_instance(seed, cam=cam, threshold=None, fraction=0.3)
            it = _gpu_pair(ctx, inst, k)
            pose, st, status = ctx.se3_track(it[2], it[1], synth.IDENTITY7)
            opose, ost, ostatus = _oracle_track(inst, p)
            assert status == ostatus and twist(pose, opose) <= 1e-9, (seed, status, ostatus)
            assert list(st.num_residual_calls) == list(ost.num_residual_calls)
            items.append(it)
            plain.append((pose, status))
        for schedule in SCHEDULES:
            _, r = _batch(ctx, items, SCHEDULES[schedule], monkeypatch)
            for k, (pose, status) in enumerate(plain):
                assert r[k].status == status and twist(r[k].frame_to_ref[:7], pose) <= 1e-9, (schedule, k)
    finally:
        ctx.set_params(lsdg.default_params())


def _depth_run(sys_, d, same_frame=False):
    cam = d["cam"]
    f32 = (lambda a: a.astype(np.float32)) if isinstance(sys_, pyoracle.Oracle) else (lambda a: a)
    kf = sys_.frame(0, cam, f32(d["kf_img"]))
    img, pose = (d["kf_img"], list(synth.IDENTITY7)) if same_frame else (d["ref_img"], list(d["ref_pose"]))
    r = sys_.frame(1, cam, f32(img))
    r.set_tracking(kf, pose + [1.0], 0.4)
    dm = sys_.depthmap(cam)
    dm.set_prior(kf, d["prior_idepth"], d["prior_var"], 20)
    dm.update([r])
    return dm.read()


DEPTH_EDGE_CAMS = {"smallest-frame": (1221, 64, 64), "zero-baseline": (1222, 160, 120), "ragged": (1223, 200, 150)}


@pytest.mark.parametrize("case", list(DEPTH_EDGE_CAMS))
def test_depth_update_edges_bit_exact(ctx, case):
    p = pyoracle.default_params()
    p.max_epl_length_crop = 100.0
    seed, w, h = DEPTH_EDGE_CAMS[case]
    d = synth.depth_instance(seed, cam=synth.camera(w, h))
    same = case == "zero-baseline"
    ctx.set_params(p)
    try:
        g = _depth_run(ctx, d, same)
    finally:
        ctx.set_params(lsdg.default_params())
    o = _depth_run(pyoracle.Oracle(0, params=p), d, same)
    assert not _depth_diff(g, o), _depth_diff(g, o)
    codes = np.bincount(g["observe_code"], minlength=15)
    if same:  # no baseline: no pixel gets past the epipolar-line checks
        assert codes[lsdg.OBS["CREATED"]] == 0 and codes[lsdg.OBS["UPDATED"]] == 0
    else:
        assert codes[lsdg.OBS["CREATED"]] + codes[lsdg.OBS["UPDATED"]] > 0


def test_batch_of_one_and_of_many(ctx, monkeypatch):
    p = tracking_params(20)
    ctx.set_params(p)
    try:
        cam = synth.camera(64, 64)
        insts = [synth.tracking_instance(1300 + k, cam=cam, threshold=None, fraction=0.3) for k in range(8)]
        items = [_gpu_pair(ctx, insts[k % 8], k) for k in range(600)]
        plain = [ctx.se3_track(ref, fr, synth.IDENTITY7) for kf, fr, ref in items[:8]]
        cfg, r = _batch(ctx, items[:1], {}, monkeypatch)
        assert r[0].status == plain[0][2] and twist(r[0].frame_to_ref[:7], plain[0][0]) <= 1e-9
        cfg, r = _batch(ctx, items, {}, monkeypatch)
        assert cfg["schedule"] == lsdg.SCHEDULE_STEPS  # 600 jobs: more than the cooperative kernel holds at once
        for k in range(600):
            pose, st, status = plain[k % 8]
            assert r[k].status == status and twist(r[k].frame_to_ref[:7], pose) <= 1e-9, k
        kf, fr, ref = items[0]
        jobs = (lsdg.Job * 2049)(*[lsdg.Job(ref.h, fr.h, lsdg.pose8(IDENT8), lsdg.JOB_SE3, 4, 1, 1)] * 2049)
        out = C.c_void_p()
        assert ctx.L.lsdg_batch_create(ctx.h, jobs, 2049, C.byref(out)) == lsdg.EINVAL and not out.value
    finally:
        ctx.set_params(lsdg.default_params())
