"""GPU parity of the keyframe constraint search (lsdg_constraint_search:
SlamSystem::testConstraint / tryTrackSim3 over batched Sim3 alignments)
against the oracle's sequential restatement with FP64 sums (prec 2), on
320x240 keyframe pairs of one scene and of unrelated scenes.

Bars: accept / reject decisions and the rejecting stage identical; stage
errors within 1e-4 relative; final alignments of both directions within
1e-7 (translation, quaternion, scale); edge informations within 1e-6
relative; edge residuals and usage within 1e-5 relative."""
import numpy as np
import pytest

import lsdg
import pyoracle
import synth

pytestmark = pytest.mark.gpu

PAIRS = [(41, False), (42, False), (45, False), (43, True), (44, True)]


def _pair(sys_, seed, other):
    d = synth.sim3_instance(seed, cam=synth.camera(320, 240), max_trans=0.04, max_rot_deg=2.0,
                            scale_range=(0.95, 1.05), fraction=0.3)
    if other:
        e = synth.sim3_instance(seed + 500, cam=synth.camera(320, 240), fraction=0.3)
        b_img, b_id, b_var = e["b_img"], e["b_idepth"], e["b_var"]
    else:
        b_img, b_id, b_var = d["b_img"], d["b_idepth"], d["b_var"]
    f32 = (lambda a: a.astype(np.float32)) if isinstance(sys_, pyoracle.Oracle) else (lambda a: a)
    a = sys_.frame(2 * seed, d["cam"], f32(d["a_img"]))
    a.set_depth(d["a_idepth"], d["a_var"])
    b = sys_.frame(2 * seed + 1, d["cam"], f32(b_img))
    b.set_depth(b_id, b_var)
    return a, b, sys_.ref(a), sys_.ref(b)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 + 1e-6 * np.max(np.abs(b))))


@pytest.mark.parametrize("strictness", [1.5, 0.002])
def test_constraint_search_matches_oracle(ctx, strictness):
    o = pyoracle.Oracle(2)
    init = [0.004, -0.003, 0.002, 1.0, 0.01, -0.01, 0.005, 1.02]
    gpu_items = [_pair(ctx, s, other) for s, other in PAIRS]
    res = ctx.constraint_search([(ra, rb, init) for a, b, ra, rb in gpu_items], strictness)
    decisions = []
    for (seed, other), r in zip(PAIRS, res):
        oa, ob, ora, orb = _pair(o, seed, other)
        e = o.constraint_test(ora, orb, init, strictness)
        assert (r.accepted, r.rejected_at) == (e.accepted, e.rejected_at), (seed, list(r.error), list(e.error))
        decisions.append(r.accepted)
        ran = 3 if r.accepted else r.rejected_at
        for s in range(ran):
            if e.error[s] >= 1e19:
                assert r.error[s] >= 1e19
            else:
                assert abs(r.error[s] - e.error[s]) <= 1e-4 * max(abs(e.error[s]), 1.0), (seed, s, r.error[s], e.error[s])
        assert np.max(np.abs(np.array(r.candidate_to_frame) - np.array(e.candidate_to_frame))) < 1e-7
        assert np.max(np.abs(np.array(r.frame_to_candidate) - np.array(e.frame_to_candidate))) < 1e-7
        if r.accepted:
            for g, x in ((r.candidate_on_frame, e.candidate_on_frame), (r.frame_on_candidate, e.frame_on_candidate)):
                assert (g.first_id, g.second_id) == (x.first_id, x.second_id)
                assert np.max(np.abs(np.array(g.second_to_first) - np.array(x.second_to_first))) < 1e-7
                assert _rel(g.information, x.information) < 1e-6
                for k in ("mean_residual", "mean_residual_d", "mean_residual_p", "usage"):
                    assert abs(getattr(g, k) - getattr(x, k)) <= 1e-5 * max(abs(getattr(x, k)), 1e-3), k
    if strictness == 1.5:
        # the same-scene pairs are accepted, the unrelated ones are not
        assert decisions == [1, 1, 1, 0, 0]
    else:
        assert not any(decisions)
    for a, b, ra, rb in gpu_items:
        for x in (ra, rb, a, b):
            x.destroy()


def test_constraint_search_empty_and_invalid(ctx):
    assert ctx.constraint_search([], 1.5) == []
    arr = (lsdg.ConstraintCandidate * 1)()
    out = (lsdg.ConstraintResult * 1)()
    assert ctx.L.lsdg_constraint_search(ctx.h, arr, 1, 1.5, out) == lsdg.EINVAL
