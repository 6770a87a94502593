"""The SlamSystem relocaliser on the GPU (lsdg_system_track after a loss:
one batched SE3 plan of quick checks against every keyframe, the chain
neighbours' confirmation, the keyframe's depth reloaded) against the
oracle's restatement (pyoracle.OSystem on prec 2) on the two sequences of
relocalise_scenarios.py.

Bars: the same status, keyframe and new-keyframe decision for every image
(tracked, lost, not tracked, relocalised); frame-to-keyframe and camToWorld
within 1e-6 while the depth maps agree bit for bit (up to the relocalised
image and the next two), then within 1e-4; the relocalised camToWorld within
millimetres of the ground truth."""
import ctypes as C

import numpy as np
import pytest

import lsdg
import pyoracle
import relocalise_scenarios as rs

pytestmark = pytest.mark.gpu


def _params(relocalise=1):
    p = lsdg.SystemParams()
    lsdg.load().lsdg_system_params_default(C.byref(p))
    p.initialization_phase_count = 1
    p.relocalise = relocalise
    return p


@pytest.mark.parametrize("kind", ["resume", "kidnap"])
def test_relocalisation_matches_the_restatement(ctx, kind):
    seq = rs.sequence()
    s = ctx.system(seq["cam"], params=_params())
    s.init_gt(0, seq["images"][0], seq["idepth"], seq["var"])
    o = pyoracle.OSystem(pyoracle.Oracle(2), seq["cam"])
    o.INIT_PHASE = 1
    o.init_gt(0, seq["images"][0], seq["idepth"], seq["var"])
    k_rel = None
    for k, idx in enumerate(rs.order(kind)):
        img = rs.image(seq, idx)
        g = s.track(k + 1, img)
        e = o.track(k + 1, img)
        assert (g.status, g.keyframe_id, g.new_keyframe) == (e["status"], e["keyframe_id"], e["new_keyframe"]), \
            (k, idx, g.status, e["status"])
        if g.status in (lsdg.SYSTEM_LOST, lsdg.SYSTEM_NOT_TRACKED):
            continue
        tol = 1e-6 if k_rel is None or k <= k_rel + 2 else 1e-4
        assert np.max(np.abs(np.array(g.frame_to_keyframe) - np.asarray(e["frame_to_keyframe"]))) < tol, k
        assert np.max(np.abs(np.array(g.cam_to_world) - np.asarray(e["cam_to_world"]))) < tol, k
        if g.status == lsdg.SYSTEM_RELOCALISED:
            k_rel = k
            assert np.linalg.norm(np.array(g.cam_to_world[4:7]) - seq["poses"][idx][4:7]) < 5e-3
    assert k_rel is not None
    ids, _ = s.keyframes()
    assert ids == [f.id for f, _ in o.keyframes]
    s.destroy()


def test_without_relocalisation_the_system_stays_lost(ctx):
    seq = rs.sequence()
    s = ctx.system(seq["cam"], params=_params(relocalise=0))
    s.init_gt(0, seq["images"][0], seq["idepth"], seq["var"])
    statuses = [s.track(k + 1, rs.image(seq, idx)).status for k, idx in enumerate(rs.order("resume"))]
    k_lost = statuses.index(lsdg.SYSTEM_LOST)
    assert all(st == lsdg.SYSTEM_NOT_TRACKED for st in statuses[k_lost + 1:])
    s.destroy()


def test_find_constraints_matches_the_restatement(ctx):
    """findConstraintsForNewKeyframes on every keyframe of the sequence:
    the same candidates (position and viewing-direction overlap, minus the
    keyframe's parent and children) and the same constraint decisions as the
    restatement; accepted alignments close to the restatement's."""
    seq = rs.sequence()
    s = ctx.system(seq["cam"], params=_params())
    s.init_gt(0, seq["images"][0], seq["idepth"], seq["var"])
    o = pyoracle.OSystem(pyoracle.Oracle(2), seq["cam"])
    o.INIT_PHASE = 1
    o.init_gt(0, seq["images"][0], seq["idepth"], seq["var"])
    for k in range(1, rs.N):
        s.track(k, seq["images"][k])
        o.track(k, seq["images"][k])
    ids, _ = s.keyframes()
    assert ids == [f.id for f, _ in o.keyframes] and len(ids) >= 4
    total = accepted = 0
    for i in range(len(ids)):
        g = s.find_constraints(i)
        e = o.find_constraints(i)
        assert [c for c, _ in g] == [c for c, _ in e], i
        for (c, rg), (_, re) in zip(g, e):
            total += 1
            assert (rg.accepted, rg.rejected_at) == (re.accepted, re.rejected_at), (i, c)
            if rg.accepted:
                accepted += 1
                assert np.max(np.abs(np.array(rg.candidate_to_frame) - np.array(re.candidate_to_frame))) < 1e-4
    assert total >= 2 and accepted >= 1
    s.destroy()
