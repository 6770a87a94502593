"""The oracle's restatement of lsd-slam's keyframe constraint test
(SlamSystem::testConstraint / tryTrackSim3), pinned on what the mathematics
fixes: the Sim3 log inverts the (separately pinned) exp, the adjoint
satisfies T exp(x) T^-1 = exp(Adj_T x), the reciprocal consistency reduces to
the harmonic combination of two diagonal informations at the identity and
vanishes for exactly inverse alignments; and end to end on the CPU a pair of
keyframes of one scene is accepted with its true relative Sim3 while a pair
of unrelated scenes is not.  The stage thresholds (3000 / 4000 / 6000 x
strictness) and levels follow the published lsd-slam, which is not in the
tree: parity unpinned for those constants."""
import numpy as np
import pytest

import lsdg
import posemath
import pyoracle
import synth


@pytest.fixture(scope="module")
def o():
    return pyoracle.Oracle(2)


def _rand_xi(rng, rot=1.0, sig=0.3, tr=1.0):
    return np.r_[rng.uniform(-tr, tr, 3), rng.uniform(-rot, rot, 3), rng.uniform(-sig, sig)]


def _close_pose(a8, b8, tol):
    d = pyoracle.sim3_mul(a8, pyoracle.sim3_inv(b8))
    return np.all(np.abs(np.asarray(d[4:7])) < tol) and abs(abs(d[3]) - 1.0) < tol and abs(d[7] - 1.0) < tol


def test_log_inverts_exp(o):
    rng = np.random.default_rng(3)
    cases = [_rand_xi(rng) for _ in range(30)]
    cases += [np.r_[0.3, -0.2, 0.1, 1e-9, 0, 0, 1e-8], np.r_[0.1, 0.2, 0.3, 0, 0, 0, 0], np.r_[1, 2, 3, 0.5, 0, 0, 0]]
    for xi in cases:
        back = o.sim3_log(pyoracle.sim3_exp(xi))
        assert np.allclose(back, xi, atol=1e-9), (xi, back)


def test_adjoint_conjugates_the_exponential(o):
    rng = np.random.default_rng(4)
    for _ in range(20):
        T = pyoracle.sim3_exp(_rand_xi(rng))
        x = _rand_xi(rng, 0.5, 0.2, 0.5)
        lhs = pyoracle.sim3_mul(pyoracle.sim3_mul(T, pyoracle.sim3_exp(x)), pyoracle.sim3_inv(T))
        rhs = pyoracle.sim3_exp(o.sim3_adj(T) @ x)
        assert _close_pose(lhs, rhs, 1e-10)


def test_consistency_of_inverse_pairs_is_zero(o):
    rng = np.random.default_rng(5)
    for _ in range(5):
        A = pyoracle.sim3_exp(_rand_xi(rng))
        M = rng.normal(size=(7, 7))
        info = M @ M.T + 7 * np.eye(7)
        assert abs(o.reciprocal_consistency(A, info, pyoracle.sim3_inv(A), info)) < 1e-16


def test_consistency_at_identity_is_the_harmonic_combination(o):
    # AtoB = identity (Adj = I): diff = log(BtoA), information (Sa^-1 + Sb^-1)^-1
    rng = np.random.default_rng(6)
    a, b = rng.uniform(1, 100, 7), rng.uniform(1, 100, 7)
    x = _rand_xi(rng, 0.2, 0.1, 0.3)
    ident = list(synth.IDENTITY7) + [1.0]
    e = o.reciprocal_consistency(ident, np.diag(a), pyoracle.sim3_exp(x), np.diag(b))
    assert np.isclose(e, np.sum(x * x * a * b / (a + b)), rtol=1e-12)


def _keyframes(o, seed, w=320, h=240, other_scene=False):
    d = synth.sim3_instance(seed, cam=synth.camera(w, h), max_trans=0.04, max_rot_deg=2.0, scale_range=(0.95, 1.05),
                            fraction=0.3)
    a = o.frame(2 * seed, d["cam"], d["a_img"].astype(np.float32))
    a.set_depth(d["a_idepth"], d["a_var"])
    if other_scene:
        e = synth.sim3_instance(seed + 500, cam=synth.camera(w, h), fraction=0.3)
        b_img, b_id, b_var = e["b_img"], e["b_idepth"], e["b_var"]
    else:
        b_img, b_id, b_var = d["b_img"], d["b_idepth"], d["b_var"]
    b = o.frame(2 * seed + 1, d["cam"], b_img.astype(np.float32))
    b.set_depth(b_id, b_var)
    return d, a, b


def test_consistent_pair_is_accepted_with_its_relative_sim3(o):
    d, a, b = _keyframes(o, 41)
    ra, rb = o.ref(a), o.ref(b)
    ident = list(synth.IDENTITY7) + [1.0]
    r = o.constraint_test(ra, rb, ident, 1.5)
    assert r.accepted == 1 and r.rejected_at == 0, (r.rejected_at, list(r.error))
    # B's keyframe on A is the instance's frame-to-ref Sim3
    rot, sc = posemath.sim3_error(list(r.candidate_to_frame), d["gt"])
    assert rot < 2e-2 and sc < 2e-2, (rot, sc)
    # the two directions are inverse to each other
    assert _close_pose(pyoracle.sim3_mul(r.frame_to_candidate, r.candidate_to_frame), ident, 5e-3)
    e1, e2 = r.candidate_on_frame, r.frame_on_candidate
    assert (e1.first_id, e1.second_id, e2.first_id, e2.second_id) == (a.id, b.id, b.id, a.id)
    assert list(e1.second_to_first) == list(r.candidate_to_frame)
    H = np.array(e1.information).reshape(7, 7)
    assert np.allclose(H, H.T) and np.all(np.linalg.eigvalsh(H) > 0)
    assert 0 <= r.error[2] <= 6000 * 1.5


def test_unrelated_scenes_are_rejected(o):
    d, a, b = _keyframes(o, 43, other_scene=True)
    ra, rb = o.ref(a), o.ref(b)
    r = o.constraint_test(ra, rb, list(synth.IDENTITY7) + [1.0], 1.5)
    assert r.accepted == 0 and r.rejected_at >= 1
    assert r.error[r.rejected_at - 1] > [3000, 4000, 6000][r.rejected_at - 1] * 1.5


# ----------------------------------------------------- keyframe candidates
def _yaw(deg):
    a = np.deg2rad(deg) / 2.0
    return [0.0, np.sin(a), 0.0, np.cos(a)]


def _cam():
    return {"w": 320, "h": 240, "fx": 262.5, "fy": 262.5, "cx": 159.5, "cy": 119.5}


def test_keyframe_candidates_closed_forms():
    """findEuclideanOverlapFrames' two tests on their boundaries: squared
    distance closeness * 15 / KFDistWeight^2 after scaling by meanIdepth /
    scale (the smaller factor of the two keyframes), and the optical axes
    within (1 - 0.25 closeness) of the mean field of view."""
    import lsdg
    cam = _cam()
    me = (0, 0.5, [0, 0, 0, 1, 0, 0, 0, 1.0])
    dist_th = 15.0 / 16.0
    t_edge = np.sqrt(dist_th) / 0.5            # |t * 0.5|^2 = dist_th
    fov = 2 * np.arctan(320 / 262.5 / 2) + 2 * np.arctan(240 / 262.5 / 2)  # fov_x + fov_y
    ang_edge = np.rad2deg(0.75 * 0.5 * fov)    # closeness 1: angle_th = 0.75
    others = [
        (1, 0.5, [0, 0, 0, 1, 0, 0, 0, 1.0]),                        # same pose
        (2, 0.5, [0, 0, 0, 1, 0.99 * t_edge, 0, 0, 1.0]),             # just inside
        (3, 0.5, [0, 0, 0, 1, 1.01 * t_edge, 0, 0, 1.0]),             # just outside
        (4, 0.5, _yaw(0.98 * ang_edge) + [0, 0, 0, 1.0]),             # axis just inside
        (5, 0.5, _yaw(1.02 * ang_edge) + [0, 0, 0, 1.0]),             # axis just outside
        (6, 0.5, [0, 0, 0, 1, 1.5 * t_edge, 0, 0, 2.0]),              # scale 2: factor 0.25 -> inside
        (7, 2.0, [0, 0, 0, 1, 1.5 * t_edge, 0, 0, 1.0]),              # larger own factor: min keeps 0.5 -> outside
    ]
    got = lsdg.keyframe_candidates(cam, me, others)
    assert [i for i, _, _ in got] == [0, 1, 3, 5]
    d2 = {i: d for i, d, _ in got}
    assert d2[0] == 0.0 and abs(d2[1] - 0.99 ** 2 * dist_th) < 1e-12 and abs(d2[5] - (1.5 * t_edge * 0.25) ** 2) < 1e-12
    assert abs(dict((i, c) for i, _, c in got)[3] - np.cos(np.deg2rad(0.98 * ang_edge))) < 1e-12
    assert [i for i, _, _ in pyoracle.keyframe_candidates(cam, me, others)] == [0, 1, 3, 5]


def test_keyframe_candidates_match_the_restatement():
    import lsdg
    rng = np.random.default_rng(5)
    cam = _cam()
    for trial in range(20):
        def pose():
            q = rng.normal(size=4) * [0.1, 0.3, 0.1, 1.0]
            q /= np.linalg.norm(q)
            return list(q) + list(rng.normal(size=3) * 1.5) + [float(rng.uniform(0.5, 2.0))]
        def mean():  # the C-ABI carries meanIdepth as a float
            return float(np.float32(rng.uniform(0.2, 1.0)))
        me = (0, mean(), pose())
        others = [(k, mean(), pose()) for k in range(1, 30)]
        closeness = float(rng.uniform(0.5, 2.0))
        a = lsdg.keyframe_candidates(cam, me, others, closeness)
        b = pyoracle.keyframe_candidates(cam, me, others, closeness)
        assert [i for i, _, _ in a] == [i for i, _, _ in b], trial
        for (_, d1, c1), (_, d2, c2) in zip(a, b):
            assert abs(d1 - d2) <= 1e-12 * max(1.0, d2) and abs(c1 - c2) <= 1e-12
