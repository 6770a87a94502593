"""Batched tracking (hipGraph-replayed) against the plain path, per instance.

Every test runs in the schedules a plan can take: every level in the
whole-job kernel (one block per job), every level in the multi-block steps
(tile evaluation + per-job solve launches), the hybrid (levels up to 3000
points in the whole-job kernel, larger ones in the steps), the defaults, and
the hybrid with the jobs split over 3 chains that step as parallel graph
branches.

Bars: FP32 pyramids with FP64 sums take the plain path's LM path (poses within
1e-9 twist, identical call counts); FP16 pyramids stay within 1e-5 of the
plain FP16 path; replays are bit-identical."""
import numpy as np
import pytest

import lsdg
import posemath
import synth
from helpers import tracking_params, twist

pytestmark = pytest.mark.gpu


SCHEDULES = {
    # name: (LSDG_BATCH_COOP, LSDG_BATCH_JOB_POINTS, LSDG_BATCH_CHAINS, LSDG_BATCH_COOP_K, LSDG_BATCH_SOLO)
    "job": ("0", "1000000000", "1", None, None),
    "steps": ("0", "0", "1", None, None),
    "hybrid": ("0", "3000", "1", None, None),
    "chains": ("0", "3000", "3", None, None),
    "default": (None, None, None, None, None),
    "coop-k2": (None, None, None, "2", None),
    "coop-split-all": (None, None, None, None, "0"),
}


@pytest.fixture(autouse=True, params=list(SCHEDULES))
def schedule(request, monkeypatch):
    """The cooperative kernel (several blocks per job, every level, levels
    above LSDG_BATCH_SOLO points split over the blocks) unless
    LSDG_BATCH_COOP=0; then levels up to LSDG_BATCH_JOB_POINTS points run in
    the whole-job kernel and LSDG_BATCH_CHAINS groups of jobs step as
    parallel graph branches (all read when a plan is created)."""
    names = ("LSDG_BATCH_COOP", "LSDG_BATCH_JOB_POINTS", "LSDG_BATCH_CHAINS", "LSDG_BATCH_COOP_K", "LSDG_BATCH_SOLO")
    for var, val in zip(names, SCHEDULES[request.param]):
        if val is None:
            monkeypatch.delenv(var, raising=False)
        else:
            monkeypatch.setenv(var, val)
    return request.param


def _instances(n, seed, resolutions):
    return synth.batch_instances(n, seed=seed, resolutions=resolutions)


def _setup(ctx, insts, precision):
    items = []
    for i, inst in enumerate(insts):
        kf = ctx.frame(2 * i, inst["cam"], inst["kf_img"], precision)
        kf.set_depth(inst["idepth"], inst["var"])
        fr = ctx.frame(2 * i + 1, inst["cam"], inst["fr_img"], precision)
        items.append((kf, fr, ctx.ref(kf)))
    return items


def _jobs(items, kind=lsdg.JOB_SE3, init=None, rebuild=1, levels=(4, 1)):
    jobs = []
    for i, (kf, fr, ref) in enumerate(items):
        p = init[i] if init is not None else list(synth.IDENTITY7) + [1.0]
        jobs.append(lsdg.Job(ref.h, fr.h, lsdg.pose8(p), kind, levels[0], levels[1], rebuild))
    return jobs


@pytest.mark.parametrize("precision", [lsdg.FP32, lsdg.FP16])
def test_batch_se3_matches_plain(ctx, precision):
    insts = _instances(12, 2031, [(320, 240), (640, 480)])
    ctx.set_params(tracking_params(20))
    try:
        items = _setup(ctx, insts, precision)
        plain = [ctx.se3_track(ref, fr, synth.IDENTITY7) for kf, fr, ref in items]
        b = ctx.batch(_jobs(items))
        r1 = b.run()
        r2 = b.run()
        b.destroy()
    finally:
        ctx.set_params(lsdg.default_params())
    for i, (pose, st, status) in enumerate(plain):
        res = r1[i]
        assert res.status == status, i
        e = twist(res.frame_to_ref[:7], pose)
        if precision == lsdg.FP32:
            assert e <= 1e-9, (i, e)
            assert list(res.stats.num_residual_calls) == list(st.num_residual_calls), i
            assert res.stats.last_good_count == st.last_good_count
        else:
            assert e <= 1e-5, (i, e)
        assert twist(res.frame_to_ref[:7], insts[i]["gt"]) < 2e-2
        # replay of the same graph is bit-identical
        assert list(r2[i].frame_to_ref) == list(res.frame_to_ref)


def test_batch_schedule_kernels(ctx, schedule):
    """The profiled run launches the kernels of the schedule the plan chose."""
    insts = _instances(4, 2051, [(320, 240)])
    items = _setup(ctx, insts, lsdg.FP32)
    b = ctx.batch(_jobs(items))
    ctx.profile(True)
    try:
        b.run()
        coop = ctx.kernel_stats(lsdg.KERNEL_BATCH_COOP)["launches"]
        job = ctx.kernel_stats(lsdg.KERNEL_BATCH_JOB)["launches"]
    finally:
        ctx.profile(False)
        b.destroy()
    if schedule.startswith("coop") or schedule == "default":
        assert coop == 1 and job == 0
    else:
        assert coop == 0


def test_batch_divergent_job(ctx):
    insts = _instances(3, 2077, [(320, 240)])
    items = _setup(ctx, insts, lsdg.FP32)
    init = [list(synth.IDENTITY7) + [1.0]] * 3
    init[1] = [0, 0, 0, 1, 50.0, 0, 0, 1.0]
    b = ctx.batch(_jobs(items, init=init))
    r = b.run()
    b.destroy()
    assert r[1].status == lsdg.DIVERGED and r[1].stats.diverged == 1
    assert np.allclose(r[1].frame_to_ref[:7], synth.IDENTITY7)
    assert r[0].status == lsdg.OK and r[2].status == lsdg.OK


def test_batch_sim3_matches_plain(ctx):
    cam = synth.camera(320, 240)
    items, inits = [], []
    params = tracking_params(100)
    for l in range(5):
        params.convergence_eps[l] = 0.9999
    ctx.set_params(params)
    try:
        for i in range(6):
            sc = synth.Scene(3100 + i)
            kf_img, kf_dep = sc.render(synth.IDENTITY7, cam)
            pose = synth.sample_pose(3200 + i, 0.02, 1.0)
            fr_img, fr_dep = sc.render(pose, cam, 2.0, 3300 + i)
            s = 0.95 + 0.02 * i
            idk, vk, _, _ = synth.semi_dense_depth(kf_img, kf_dep, 3400 + i, fraction=0.3)
            idf, vf, _, _ = synth.semi_dense_depth(fr_img, fr_dep * s, 3500 + i, fraction=0.3)
            kf = ctx.frame(2 * i, cam, kf_img)
            kf.set_depth(idk, vk)
            fr = ctx.frame(2 * i + 1, cam, fr_img)
            fr.set_depth(idf, vf)
            items.append((kf, fr, ctx.ref(kf)))
            inits.append(list(synth.IDENTITY7) + [1.0])
        plain = [ctx.sim3_track(ref, fr, synth.IDENTITY7, 4, 1) for kf, fr, ref in items]
        b = ctx.batch(_jobs(items, kind=lsdg.JOB_SIM3, init=inits, rebuild=0))
        r = b.run()
        b.destroy()
    finally:
        ctx.set_params(lsdg.default_params())
    for i, (pose, st, status) in enumerate(plain):
        assert r[i].status == status
        rig, sc = posemath.sim3_error(list(r[i].frame_to_ref), pose)
        assert rig <= 1e-9 and sc <= 1e-9, (i, rig, sc)
        H, Hp = np.array(r[i].stats.hessian[:]), np.array(st.hessian[:])
        assert np.abs(H - Hp).max() <= 1e-9 * np.abs(Hp).max()


@pytest.mark.parametrize("precision", [lsdg.FP32, lsdg.FP16])
def test_batch_sim3_1920x1080(ctx, schedule, precision):
    """Batched Sim3 at 1920x1080: 4 keyframe pairs, two jobs each (shared
    frames), against the plain Sim3 path of the same precision."""
    insts = [synth.sim3_instance(1005 + k) for k in range(4)]
    items = []
    for i, d in enumerate(insts):
        a = ctx.frame(2 * i, d["cam"], d["a_img"], precision)
        a.set_depth(d["a_idepth"], d["a_var"])
        b = ctx.frame(2 * i + 1, d["cam"], d["b_img"], precision)
        b.set_depth(d["b_idepth"], d["b_var"])
        items.append((a, b, ctx.ref(a)))
    ident = list(synth.IDENTITY7) + [1.0]
    plain = [ctx.sim3_track(ref, b, ident, 4, 1) for a, b, ref in items]
    jobs = [lsdg.Job(ref.h, b.h, lsdg.pose8(ident), lsdg.JOB_SIM3, 4, 1, 0) for a, b, ref in items for _ in range(2)]
    bt = ctx.batch(jobs)
    res = bt.run()
    bt.destroy()
    for j, r in enumerate(res):
        pose, st, status = plain[j // 2]
        assert r.status == status
        rig, sc = posemath.sim3_error(list(r.frame_to_ref), pose)
        tol = 1e-9 if precision == lsdg.FP32 else 1e-5
        assert rig <= tol and sc <= tol, (j, rig, sc)
    assert list(res[0].frame_to_ref) == list(res[1].frame_to_ref)


def _identity_table(w, h):
    """Undistortion table that copies the image (integer source positions);
    the last column and row have no two-tap source and stay 0."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    t = np.stack([x, y], -1)
    t[:, -1] = -1.0
    t[-1, :] = -1.0
    return t


def test_batch_rerun_after_new_frame_content(ctx, schedule):
    """A plan run again after its frames were given new images
    (lsdg_frame_ingest, no pyramid rebuild inside the batch) tracks the new
    content: its poses equal a fresh plan's on frames made from the new
    images, bit for bit.  A Sim3 plan whose frame lost its depth that way
    reports LSDG_ESTATE."""
    insts = _instances(3, 2091, [(320, 240)])
    cam = insts[0]["cam"]
    und = ctx.undistorter((cam["w"], cam["h"]), _identity_table(cam["w"], cam["h"]))
    ctx.set_params(tracking_params(20))
    try:
        items = []
        for i, inst in enumerate(insts):
            kf = ctx.frame_undistorted(2 * i, cam, und, inst["kf_img"])
            kf.set_depth(inst["idepth"], inst["var"])
            fr = ctx.frame_undistorted(2 * i + 1, cam, und, insts[(i + 1) % 3]["fr_img"])
            items.append((kf, fr, ctx.ref(kf)))
        b = ctx.batch(_jobs(items, rebuild=0))
        r_old = b.run()
        for i, (kf, fr, ref) in enumerate(items):
            fr.ingest(2 * i + 1, und, insts[i]["fr_img"])
        r_new = b.run()
        b.destroy()
        fresh = []
        for i, (kf, fr, ref) in enumerate(items):
            fresh.append((kf, ctx.frame_undistorted(100 + i, cam, und, insts[i]["fr_img"]), ref))
        bf = ctx.batch(_jobs(fresh, rebuild=0))
        r_fresh = bf.run()
        bf.destroy()
        for i in range(3):
            assert list(r_new[i].frame_to_ref) == list(r_fresh[i].frame_to_ref), i
            assert list(r_new[i].stats.num_residual_calls) == list(r_fresh[i].stats.num_residual_calls)
            assert list(r_new[i].frame_to_ref) != list(r_old[i].frame_to_ref), i
        # Sim3: the frame's depth is gone after the new image
        kf, fr, ref = items[0]
        fr.set_depth(insts[0]["idepth"], insts[0]["var"])
        bs = ctx.batch([lsdg.Job(ref.h, fr.h, lsdg.pose8(list(synth.IDENTITY7) + [1.0]), lsdg.JOB_SIM3, 4, 1, 0)])
        bs.run()
        fr.ingest(1, und, insts[0]["fr_img"])
        with pytest.raises(lsdg.LsdgError, match="status -4"):
            bs.run()
        bs.destroy()
    finally:
        ctx.set_params(lsdg.default_params())
        und.close()


def test_cooperative_wait_gives_up_cleanly(ctx, schedule, monkeypatch):
    """The cooperative kernel's row-exchange wait is bounded: with the poll
    limit forced to 0 (LSDG_BATCH_SPIN_MAX, every level split over the job's
    blocks) the first block of a job to arrive at an exchange gives up, every
    block of the launch exits, and the run reports LSDG_EDEVICE instead of
    hanging.  The context stays usable: a plan with the normal limit then
    tracks as before."""
    if not (schedule.startswith("coop") or schedule == "default"):
        pytest.skip("the bounded wait belongs to the cooperative schedule")
    insts = _instances(8, 2101, [(320, 240)])
    items = _setup(ctx, insts, lsdg.FP32)
    monkeypatch.setenv("LSDG_BATCH_SOLO", "0")
    monkeypatch.setenv("LSDG_BATCH_SPIN_MAX", "0")
    b = ctx.batch(_jobs(items))
    assert b.config()["schedule"] == lsdg.SCHEDULE_COOP
    with pytest.raises(lsdg.LsdgError, match="status -2"):
        b.run()
    b.destroy()
    monkeypatch.delenv("LSDG_BATCH_SPIN_MAX")
    b = ctx.batch(_jobs(items))
    r = b.run()
    b.destroy()
    plain = [ctx.se3_track(ref, fr, synth.IDENTITY7) for kf, fr, ref in items]
    for i, (pose, st, status) in enumerate(plain):
        assert r[i].status == status and twist(r[i].frame_to_ref[:7], pose) <= 1e-9, i
