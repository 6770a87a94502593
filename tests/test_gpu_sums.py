"""Compensated FP32 sums (lsdg_ctx_set_sums(LSDG_SUMS_KAHAN)) against the
oracle's twin (prec 2: the same FP32 per-point terms, summed in FP64).

The per-point terms are the same FP32 values in both modes, so a sum differs
from the twin's only by the summation error.  Kahan's bound for n terms is
(2u + O(n u^2)) * sum |x_i| with u = 2^-24, and the TwoSum merges of the wave
and block trees add only O(u^2) terms; the bars below use 8u times a bound on
sum |x_i| that the outputs themselves give (Cauchy-Schwarz with w >= 0):
  A_ij = sum w J_i J_j:   sum |w J_i J_j| <= sqrt(A_ii A_jj)
  b_i  = -sum w J_i r:    sum |w J_i r|  <= sqrt(2 A_ii E)   (E = sum w r^2;
                          the factor 2 covers Sim3's photometric + depth parts)
  positive sums S:        sum |x| = S
  sum_signed_res = sum r: sum |r| <= sqrt(good_count * sum r^2)
A dropped correction term (plain FP32 sums) misses these bars by orders of
magnitude at the larger levels; an error in the pair merges shows in A.

Tracking: Kahan sums differ from FP64 ones by ~1e-10 of the bound above
(measured), which can still flip a borderline LM accept; the bar is 1e-5
twist against the twin on the 12 seeded 640x480 instances and a
mixed-resolution batch, and for Sim3 1e-5 plus the FP32-sum restatement's own
spread."""
import numpy as np
import pytest

import lsdg
import posemath
import pyoracle
import synth
from helpers import gpu_pair, oracle_pair, tracking_params, twist

pytestmark = pytest.mark.gpu

U = 2.0 ** -24
BAR = 8.0 * U


@pytest.fixture
def kahan(ctx):
    ctx.set_sums(lsdg.SUMS_KAHAN)
    try:
        yield ctx
    finally:
        ctx.set_sums(lsdg.SUMS_FP64)
        ctx.set_params(lsdg.default_params())


def _check_sums(ev, e2, n, tag):
    A, A2 = ev.A7()[:n, :n], e2.A7()[:n, :n]
    d = np.sqrt(np.maximum(np.diag(A2), 0.0))
    bound_A = np.outer(d, d)
    assert np.all(np.abs(A - A2) <= BAR * bound_A + 1e-300), (tag, np.max(np.abs(A - A2) / (bound_A + 1e-300)))
    E = e2.error
    b, b2 = np.array(ev.b[:n]), np.array(e2.b[:n])
    bound_b = np.sqrt(2.0 * np.diag(A2) * E)
    assert np.all(np.abs(b - b2) <= BAR * bound_b + 1e-300), (tag, np.max(np.abs(b - b2) / (bound_b + 1e-300)))
    assert abs(ev.error - E) <= BAR * E, tag
    return np.max(np.abs(A - A2) / (bound_A + 1e-300))


@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_kahan_se3_evaluate(kahan, level):
    inst = synth.workload1()
    kf, fr, ref = gpu_pair(kahan, inst)
    pose = synth.sample_pose(77, 0.01, 0.5)
    ev, mask = kahan.se3_evaluate(ref, fr, level, pose, want_mask=True)
    o = pyoracle.Oracle(2)
    _, ofr, oref = oracle_pair(o, inst)
    e2, m2 = o.se3_evaluate(oref, ofr, level, pose, want_mask=True)
    assert (ev.warped_count, ev.good_count, ev.bad_count) == (e2.warped_count, e2.good_count, e2.bad_count)
    if level == 1:
        assert np.array_equal(mask, m2)
    rel = _check_sums(ev, e2, 6, f"se3 level {level}")
    for k in ("sum_res_unweighted", "usage", "sxx", "syy", "sx", "sy", "sw"):
        v, v2 = getattr(ev, k), getattr(e2, k)
        assert v2 >= 0 and abs(v - v2) <= BAR * v2, k
    bound = np.sqrt(e2.good_count * e2.sum_res_unweighted)
    assert abs(ev.sum_signed_res - e2.sum_signed_res) <= BAR * bound
    # the compensation is what keeps the sums this close: the same terms
    # summed by the oracle in plain FP32 (prec 0 sums its FP32 terms in FP32)
    # are further off at the level with the most points
    if level == 1:
        o0 = pyoracle.Oracle(0)
        _, ofr0, oref0 = oracle_pair(o0, inst)
        e0, _ = o0.se3_evaluate(oref0, ofr0, level, pose)
        A0 = e0.A7()[:6, :6]
        d = np.sqrt(np.diag(e2.A7()[:6, :6]))
        rel0 = np.max(np.abs(A0 - e2.A7()[:6, :6]) / np.outer(d, d))
        print(f"level 1: Kahan max |dA|/bound {rel:.2e}, plain FP32 sums {rel0:.2e}")
        assert rel < rel0


@pytest.mark.parametrize("level", [1, 2, 3])
def test_kahan_sim3_evaluate(kahan, level):
    cam = synth.camera(320, 240)
    sc = synth.Scene(501)
    kf_img, kf_dep = sc.render(synth.IDENTITY7, cam)
    pose = synth.sample_pose(502, 0.02, 1.0)
    fr_img, fr_dep = sc.render(pose, cam, 2.0, 503)
    idk, vk, _, _ = synth.semi_dense_depth(kf_img, kf_dep, 504, fraction=0.3)
    idf, vf, _, _ = synth.semi_dense_depth(fr_img, fr_dep * 1.05, 505, fraction=0.3)
    p8 = list(synth.sample_pose(506, 0.01, 0.5)) + [1.02]
    out = []
    for sys_ in (kahan, pyoracle.Oracle(2)):
        f32 = (lambda a: a.astype(np.float32)) if isinstance(sys_, pyoracle.Oracle) else (lambda a: a)
        kf = sys_.frame(0, cam, f32(kf_img))
        kf.set_depth(idk, vk)
        fr = sys_.frame(1, cam, f32(fr_img))
        fr.set_depth(idf, vf)
        out.append(sys_.sim3_evaluate(sys_.ref(kf), fr, level, p8))
    ev, e2 = out
    assert ev.warped_count == e2.warped_count
    assert (ev.num_terms_p, ev.num_terms_d) == (e2.num_terms_p, e2.num_terms_d)
    _check_sums(ev, e2, 7, f"sim3 level {level}")
    for k in ("sum_res_p", "sum_res_d"):
        assert abs(getattr(ev, k) - getattr(e2, k)) <= BAR * getattr(e2, k), k


@pytest.mark.parametrize("seed", [1001 + 17 * i for i in range(12)])
def test_kahan_se3_track_seeds(kahan, seed):
    inst = synth.tracking_instance(seed)
    p = tracking_params(20)
    kahan.set_params(p)
    kf, fr, ref = gpu_pair(kahan, inst)
    pose, st, status = kahan.se3_track(ref, fr, synth.IDENTITY7)
    o = pyoracle.Oracle(2, params=p)
    _, ofr, oref = oracle_pair(o, inst)
    opose, ost, ostatus = o.se3_track(oref, ofr, synth.IDENTITY7)
    assert status == ostatus == lsdg.OK
    t = twist(pose, opose)
    same = list(st.num_residual_calls) == list(ost.num_residual_calls)
    print(f"seed {seed}: twist vs twin {t:.2e}, same evaluation counts {same}")
    assert t <= 1e-5


def test_kahan_sim3_track(kahan):
    """1920x1080 keyframe pair.  Sim3's scale is weakly constrained and the
    default stop rule fires while it still moves, so summation differences
    alone move the result: the bar is 1e-5 plus twice the spread between the
    twin and the FP32-sum restatement (prec 0, the reference's arithmetic)."""
    d = synth.sim3_instance(1005)
    out = []
    for sys_ in (kahan, pyoracle.Oracle(2), pyoracle.Oracle(0)):
        f32 = (lambda a: a.astype(np.float32)) if isinstance(sys_, pyoracle.Oracle) else (lambda a: a)
        a = sys_.frame(0, d["cam"], f32(d["a_img"]))
        a.set_depth(d["a_idepth"], d["a_var"])
        b = sys_.frame(1, d["cam"], f32(d["b_img"]))
        b.set_depth(d["b_idepth"], d["b_var"])
        out.append(sys_.sim3_track(sys_.ref(a), b, list(synth.IDENTITY7) + [1.0], 4, 1))
    (p, st, s), (q, qst, qs), (q0, _, qs0) = out
    assert s == qs == qs0 == lsdg.OK
    rig, sc = posemath.sim3_error(p, q)
    rig0, sc0 = posemath.sim3_error(q0, q)
    print(f"Sim3 1920x1080 Kahan vs twin: rigid {rig:.2e} scale {sc:.2e}; FP32 sums vs twin: {rig0:.2e} {sc0:.2e}")
    assert rig <= 1e-5 + 2 * rig0 and sc <= 1e-5 + 2 * sc0


def test_kahan_batch(kahan):
    """Mixed-resolution batch of 16 under Kahan sums, cooperative and step
    schedules: every job within 1e-5 of the twin; the tuning key carries the
    summation mode."""
    insts = synth.batch_instances(16, seed=1003)
    p = tracking_params(20)
    kahan.set_params(p)
    items = []
    for i, inst in enumerate(insts):
        kf = kahan.frame(2 * i, inst["cam"], inst["kf_img"])
        kf.set_depth(inst["idepth"], inst["var"])
        fr = kahan.frame(2 * i + 1, inst["cam"], inst["fr_img"])
        items.append((kf, fr, kahan.ref(kf)))
    jobs = [lsdg.Job(ref.h, fr.h, lsdg.pose8(list(synth.IDENTITY7) + [1.0]), lsdg.JOB_SE3, 4, 1, 1)
            for kf, fr, ref in items]
    key = kahan.batch_key(jobs)
    assert "-kahan/" in key
    b = kahan.batch(jobs)
    lsdg.tune_table_put(key, b.config())  # the key is one the table accepts
    b.destroy()
    assert lsdg.tune_table_get(key) is not None
    o = pyoracle.Oracle(2, params=p)
    twin = []
    for inst in insts:
        kf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
        kf.set_depth(inst["idepth"], inst["var"])
        fr = o.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
        twin.append(o.se3_track(o.ref(kf), fr, synth.IDENTITY7)[0])
    for schedule in (lsdg.SCHEDULE_COOP, lsdg.SCHEDULE_STEPS):
        b = kahan.batch(jobs)
        cfg = dict(b.config())
        b.destroy()
        cfg["schedule"] = schedule
        b = kahan.batch_with_config(jobs, cfg)
        res = b.run()
        b.destroy()
        errs = [twist(r.frame_to_ref[:7], q) for r, q in zip(res, twin)]
        print(f"schedule {schedule}: max twist vs twin {max(errs):.2e}")
        assert all(r.status == lsdg.OK for r in res)
        assert max(errs) <= 1e-5


def test_sums_mode_round_trip(ctx):
    assert ctx.sums() == lsdg.SUMS_FP64
    ctx.set_sums(lsdg.SUMS_KAHAN)
    assert ctx.sums() == lsdg.SUMS_KAHAN
    ctx.set_sums(lsdg.SUMS_FP64)
    with pytest.raises(lsdg.LsdgError):
        ctx.set_sums(7)
    assert ctx.sums() == lsdg.SUMS_FP64
