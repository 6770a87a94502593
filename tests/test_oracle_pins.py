"""Pins the oracle (the CPU restatement the GPU path is checked against) to
mathematics it must satisfy, since the reference ships no fixtures for this
path: closed forms, library routines on the same inputs, brute force, finite
differences and conservation identities.  Each check is chosen so that a
plausible slip (a transposed rotation, a wrong sign of a twist component, a
dropped series term, a swapped tap weight, an off-by-one border) fails it.

CPU only; the oracle's parity build (liboracle.so) is used throughout."""
import ctypes as C

import numpy as np
import pytest
from scipy.linalg import expm

import lsdg
import posemath
import synth
from helpers import frozen_energy, frozen_energy_sim3
from pyoracle import Oracle, ldlt, load, se3_exp, sim3_exp, sim3_inv, sim3_mul


def hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]], np.float64)


def mat8(p8):
    """4x4 matrix of a Sim3 pose (qx, qy, qz, qw, tx, ty, tz, s): X -> s R X + t."""
    M = np.eye(4)
    M[:3, :3] = p8[7] * posemath.quat_to_R(p8[:4])
    M[:3, 3] = p8[4:7]
    return M


# ------------------------------------------------------------------ Lie maps
TWISTS = [np.array(v, np.float64) for v in (
    [0.1, -0.2, 0.3, 0.4, -0.5, 0.6],
    [1.0, 2.0, -1.0, 2.5, 0.3, -1.1],        # large rotation (> 2.7 rad)
    [0.01, 0.0, 0.02, 3e-6, -2e-6, 1e-6],    # second series branch (theta < 1e-5)
    [0.3, 0.1, -0.2, 1e-12, 0.0, 0.0],       # first-order branch (theta < 1e-10)
    [0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
)]


@pytest.mark.parametrize("xi", TWISTS)
def test_se3_exp_is_matrix_exponential(xi):
    G = np.zeros((4, 4))
    G[:3, :3] = hat(xi[3:])
    G[:3, 3] = xi[:3]
    ref = expm(G)
    T = se3_exp(xi)
    M = mat8(np.r_[T, 1.0])
    assert np.abs(M - ref).max() < 1e-12


SIM3_TWISTS = [np.array(v, np.float64) for v in (
    [0.1, -0.2, 0.3, 0.4, -0.5, 0.6, 0.2],
    [0.5, 0.1, -0.4, 1.2, 0.7, -0.3, -0.7],
    [0.1, 0.2, 0.3, 0.2, -0.1, 0.05, 3e-7],   # |sigma| < 1e-6 series
    [0.1, 0.2, 0.3, 2e-6, 1e-6, 0.0, 0.3],    # theta < 1e-5, sigma large
    [0.1, 0.2, 0.3, 2e-6, 1e-6, 0.0, 4e-7],   # both small
)]


@pytest.mark.parametrize("xi", SIM3_TWISTS)
def test_sim3_exp_is_matrix_exponential(xi):
    G = np.zeros((4, 4))
    G[:3, :3] = hat(xi[3:6]) + xi[6] * np.eye(3)
    G[:3, 3] = xi[:3]
    ref = expm(G)
    T = sim3_exp(xi)
    assert abs(T[7] - np.exp(xi[6])) < 1e-14
    assert np.abs(mat8(T) - ref).max() < 1e-10


def test_sim3_product_and_inverse_are_matrix_algebra():
    rng = np.random.default_rng(7)
    for _ in range(20):
        a = sim3_exp(rng.normal(0, 0.5, 7))
        b = sim3_exp(rng.normal(0, 0.5, 7))
        assert np.abs(mat8(sim3_mul(a, b)) - mat8(a) @ mat8(b)).max() < 1e-12
        assert np.abs(mat8(sim3_inv(a)) - np.linalg.inv(mat8(a))).max() < 1e-12


# ------------------------------------------------------------------ LDL^T
@pytest.mark.parametrize("n", [6, 7])
def test_ldlt_solves_spd_systems(n):
    rng = np.random.default_rng(11 + n)
    for _ in range(50):
        J = rng.normal(size=(3 * n, n))
        A = J.T @ J + 1e-3 * np.eye(n)
        b = rng.normal(size=n)
        x = ldlt(A.ravel(), b)
        ref = np.linalg.solve(A, b)
        assert np.abs(x - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


# ------------------------------------------------------------------ bilinear
def test_bilinear_is_exact_on_affine_images():
    w, h = 17, 13
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    img = (7.0 + 0.75 * x - 0.5 * y).astype(np.float32)
    fn = load().orc_interp_f32
    fn.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int]
    fn.restype = C.c_float
    rng = np.random.default_rng(3)
    for _ in range(200):
        u, v = rng.uniform(0, w - 1.001), rng.uniform(0, h - 1.001)
        got = fn(img.ctypes.data, u, v, w)
        assert abs(got - (7.0 + 0.75 * u - 0.5 * v)) < 2e-5
    # the tap weights are not symmetric in x and y: a transposed image breaks it
    assert abs(fn(img.ctypes.data, 3.25, 5.5, w) - (7.0 + 0.75 * 3.25 - 0.5 * 5.5)) < 1e-5


# ------------------------------------------------------------------ pyramids
def _oframe(img, cam=None, prec=0):
    h, w = img.shape
    cam = cam or dict(w=w, h=h, fx=100.0, fy=100.0, cx=(w - 1) / 2.0, cy=(h - 1) / 2.0)
    o = Oracle(prec)
    return o, o.frame(0, cam, img.astype(np.float32)), cam


def test_pyramid_of_a_ramp():
    w, h = 64, 48
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = 10.0 + 3.0 * x + 2.0 * y
    o, f, cam = _oframe(img)
    for lvl in range(5):
        s = 1 << lvl
        I = f.read(lsdg.WHAT_IMAGE, lvl)
        hh, ww = I.shape
        yy, xx = np.mgrid[0:hh, 0:ww].astype(np.float64)
        # 2x2 box means of a ramp: the ramp sampled at the block centres
        expect = 10.0 + 3.0 * (s * xx + (s - 1) / 2.0) + 2.0 * (s * yy + (s - 1) / 2.0)
        assert np.abs(I - expect).max() < 1e-3, lvl
        G = f.read(lsdg.WHAT_GRADIENTS, lvl)
        # central half differences: slope x level scale, interior only
        assert np.abs(G[1:-1, 1:-1, 0] - 3.0 * s).max() < 1e-3
        assert np.abs(G[1:-1, 1:-1, 1] - 2.0 * s).max() < 1e-3
        # the intensity lane carries the image on rows 1..h-2 (the rows
        # buildGradients writes); the first and last rows stay zero
        assert np.array_equal(G[1:-1, :, 2], I[1:-1, :])
        assert np.all(G[0] == 0) and np.all(G[-1] == 0)


def test_constant_image_has_no_gradient():
    o, f, _ = _oframe(np.full((32, 40), 77.0))
    for lvl in range(5):
        assert np.all(f.read(lsdg.WHAT_IMAGE, lvl) == 77.0)
        assert np.all(f.read(lsdg.WHAT_GRADIENTS, lvl)[..., :2] == 0.0)
        assert np.all(f.read(lsdg.WHAT_MAXGRAD, lvl) == 0.0)


def test_max_gradient_is_3x3_max_of_gradient_norm():
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 255, (30, 40)).astype(np.float32)
    o, f, _ = _oframe(img)
    for lvl in range(2):
        I = f.read(lsdg.WHAT_IMAGE, lvl).astype(np.float32)
        gx = np.float32(0.5) * (I[1:-1, 2:] - I[1:-1, :-2])
        gy = np.float32(0.5) * (I[2:, 1:-1] - I[:-2, 1:-1])
        g = np.zeros_like(I)
        g[1:-1, 1:-1] = np.sqrt(gx * gx + gy * gy)
        M = f.read(lsdg.WHAT_MAXGRAD, lvl)
        hh, ww = I.shape
        for yy in range(2, hh - 2):
            for xx in range(2, ww - 2):
                assert M[yy, xx] == g[yy - 1:yy + 2, xx - 1:xx + 2].max()


# ------------------------------------------------------------ point cloud
def test_point_cloud_reprojects_onto_its_pixels():
    inst = synth.tracking_instance(1001, cam=synth.camera(160, 120))
    o = Oracle(0)
    kf = o.frame(0, inst["cam"], inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    ref = o.ref(kf)
    cam = inst["cam"]
    for lvl in range(4):
        pts = ref.points(lvl)
        assert len(pts) > 0
        s = 1 << lvl
        fx, fy = cam["fx"] / s, cam["fy"] / s
        cx, cy = (cam["cx"] + 0.5) / s - 0.5, (cam["cy"] + 0.5) / s - 0.5
        w = cam["w"] >> lvl
        idx = pts[:, 7].astype(np.int64)
        u = pts[:, 0] / pts[:, 2] * fx + cx
        v = pts[:, 1] / pts[:, 2] * fy + cy
        assert np.abs(u - idx % w).max() < 1e-3
        assert np.abs(v - idx // w).max() < 1e-3
        # every point is a distinct pixel with valid depth, emitted column by
        # column (x outer, y inner), the order of makePointCloud's loops
        xs, ys = idx % w, idx // w
        assert np.array_equal(np.lexsort((ys, xs)), np.arange(len(idx)))
        assert len(np.unique(idx)) == len(idx)
        assert np.all(pts[:, 2] > 0)


# ------------------------------------------------- warp / residual convention
def _plane_pair(tx, z0=2.0, w=96, h=72):
    """Keyframe and frame of an intensity ramp painted on a fronto-parallel
    plane at depth z0; the frame's camera sits at +tx along x (frameToRef).
    The ramp is affine in the image, so bilinear sampling is exact and the
    photometric residual vanishes at the true pose."""
    fx = fy = 80.0
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    kf_img = 40.0 + 1.5 * x + 0.8 * y
    # the frame pixel u sees X = (u - cx) z0 / fx + tx, i.e. keyframe pixel u + fx tx / z0
    fr_img = 40.0 + 1.5 * (x + fx * tx / z0) + 0.8 * y
    cam = dict(w=w, h=h, fx=fx, fy=fy, cx=cx, cy=cy)
    idepth = np.full((h, w), 1.0 / z0, np.float32)
    var = np.full((h, w), 1e-3, np.float32)
    return cam, kf_img.astype(np.float32), fr_img.astype(np.float32), idepth, var


def test_residual_vanishes_at_the_true_pose_only():
    tx = 0.05
    cam, kfi, fri, idepth, var = _plane_pair(tx)
    o = Oracle(1)
    kf = o.frame(0, cam, kfi)
    kf.set_depth(idepth, var)
    fr = o.frame(1, cam, fri)
    ref = o.ref(kf)
    # evaluations take referenceToFrame, the inverse of frameToRef = (I, +tx)
    ev, _ = o.se3_evaluate(ref, fr, 1, [0, 0, 0, 1, -tx, 0, 0])
    assert ev.warped_count > 0 and ev.error < 1e-18
    ev2, _ = o.se3_evaluate(ref, fr, 1, [0, 0, 0, 1, tx, 0, 0])  # the sign slip
    assert ev2.error > 1.0


def _textured_plane_pair(tx, z0=2.0, w=160, h=120):
    """A smooth texture painted on a fronto-parallel plane at depth z0, seen
    from the keyframe and from a frame whose camera sits at +tx along x."""
    fx = fy = 120.0
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)

    def tex(X, Y):
        return (128 + 40 * np.sin(2 * np.pi * X / 0.4) + 30 * np.cos(2 * np.pi * Y / 0.3)
                + 20 * np.sin(2 * np.pi * (X + Y) / 0.25))

    X, Y = (x - cx) * z0 / fx, (y - cy) * z0 / fy
    cam = dict(w=w, h=h, fx=fx, fy=fy, cx=cx, cy=cy)
    return (cam, tex(X, Y).astype(np.float32), tex(X + tx, Y).astype(np.float32),
            np.full((h, w), 1.0 / z0, np.float32), np.full((h, w), 1e-3, np.float32))


@pytest.mark.parametrize("prec", [0, 1])
def test_tracker_recovers_the_true_pose_of_a_noise_free_pair(prec):
    tx = 0.03
    cam, kfi, fri, idepth, var = _textured_plane_pair(tx)
    o = Oracle(prec)
    kf = o.frame(0, cam, kfi)
    kf.set_depth(idepth, var)
    fr = o.frame(1, cam, fri)
    pose, st, status = o.se3_track(o.ref(kf), fr, synth.IDENTITY7)
    assert status == lsdg.OK
    # frameToRef = (I, +tx); the residual left is the bilinear sampling of a
    # non-affine texture and the planar rotation / translation ambiguity
    assert posemath.twist_error(pose, [0, 0, 0, 1, tx, 0, 0]) < 5e-4
    assert abs(pose[4] - tx) < 5e-4


# ------------------------------------------------------------------ adjoint
def _adjoint_setup():
    """Keyframe points of a textured synthetic scene against a frame whose
    image is affine: there the central-difference gradient the adjoint uses
    is the exact derivative of the bilinear sample, so finite differences of
    the frozen-weight energy must reproduce it (on a textured frame the two
    differ by the interpolation's kinks, by design of the reference)."""
    inst = synth.tracking_instance(1007, cam=synth.camera(160, 120))
    cam = inst["cam"]
    o = Oracle(1)
    kf = o.frame(0, cam, inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    y, x = np.mgrid[0:cam["h"], 0:cam["w"]].astype(np.float64)
    fr = o.frame(1, cam, (60.0 + 0.9 * x + 0.4 * y).astype(np.float32))
    ref = o.ref(kf)
    ref_to_frame = sim3_exp(np.array([0.01, -0.008, 0.005, 0.004, -0.003, 0.002, 0.0]))[:7]
    return o, ref, fr, ref_to_frame


@pytest.mark.parametrize("lvl", [1, 2])
def test_adjoint_matches_central_differences(lvl):
    o, ref, fr, pose = _adjoint_setup()
    n = ref.num_points(lvl)
    E, g_xi, g_d = o.energy_grad(ref, fr, lvl, pose)
    wf = np.zeros(n, np.float64)
    mf = np.zeros(n, np.uint8)
    zero_d = np.zeros(n)
    E0 = frozen_energy(o, ref, fr, lvl, pose, np.zeros(6), zero_d, wf, mf, True)
    assert mf.sum() > 0.9 * n
    assert abs(E0 - E) <= 1e-12 * abs(E)
    h = 1e-6
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        fd = (frozen_energy(o, ref, fr, lvl, pose, e, zero_d, wf, mf, False)
              - frozen_energy(o, ref, fr, lvl, pose, -e, zero_d, wf, mf, False)) / (2 * h)
        assert abs(fd - g_xi[k]) <= 1e-6 * np.abs(g_xi).max(), (k, fd, g_xi[k])
    # E is separable in the point inverse depths: one directional derivative
    # along a random direction checks every dE/dd_i at once
    rng = np.random.default_rng(1)
    dirn = rng.normal(size=n)
    fd = (frozen_energy(o, ref, fr, lvl, pose, np.zeros(6), h * dirn, wf, mf, False)
          - frozen_energy(o, ref, fr, lvl, pose, np.zeros(6), -h * dirn, wf, mf, False)) / (2 * h)
    assert abs(fd - g_d @ dirn) <= 1e-6 * np.abs(g_d).sum()


def _adjoint_setup_sim3(prec=1):
    """Sim3 adjoint scene: textured keyframe points against an affine frame
    image (exact central-difference gradients, as for SE3) whose inverse
    depth is a smooth ramp with a hole of invalid pixels (no depth residual
    there), at a Sim3 pose with a 3% scale.  ESM off: the adjoint's exact
    derivative and the tracker's LGS then use the same gradient."""
    inst = synth.tracking_instance(1011, cam=synth.camera(160, 120))
    cam = inst["cam"]
    p = lsdg.default_params()
    p.use_esm_tracking = 0
    o = Oracle(prec, params=p)
    kf = o.frame(0, cam, inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    y, x = np.mgrid[0:cam["h"], 0:cam["w"]].astype(np.float64)
    fr = o.frame(1, cam, (60.0 + 0.9 * x + 0.4 * y).astype(np.float32))
    fid = (0.35 + 0.002 * x - 0.001 * y).astype(np.float32)
    fvar = np.full(fid.shape, 2e-3, np.float32)
    fvar[40:70, 60:100] = -1.0  # no frame depth: the point keeps only its photometric term
    fr.set_depth(fid, fvar)
    ref = o.ref(kf)
    ref_to_frame = sim3_exp(np.array([0.01, -0.008, 0.005, 0.004, -0.003, 0.002, 0.03]))
    return o, ref, fr, ref_to_frame


@pytest.mark.parametrize("lvl", [1, 2])
def test_sim3_adjoint_matches_central_differences(lvl):
    o, ref, fr, pose = _adjoint_setup_sim3()
    n = ref.num_points(lvl)
    E, g_xi, g_d = o.energy_grad_sim3(ref, fr, lvl, pose)
    fz = np.zeros(3 * n, np.float64)
    mf = np.zeros(n, np.uint8)
    zero_d = np.zeros(n)
    E0 = frozen_energy_sim3(o, ref, fr, lvl, pose, np.zeros(7), zero_d, fz, mf, True)
    assert mf.sum() > 0.9 * n
    # both kinds of point are present: with and without a depth residual
    wd = fz[1::3][mf.astype(bool)]
    assert (wd > 0).sum() > 0.5 * mf.sum() and (wd == 0).sum() > 10
    assert abs(E0 - E) <= 1e-12 * abs(E)
    h = 1e-6
    for k in range(7):
        e = np.zeros(7)
        e[k] = h
        fd = (frozen_energy_sim3(o, ref, fr, lvl, pose, e, zero_d, fz, mf, False)
              - frozen_energy_sim3(o, ref, fr, lvl, pose, -e, zero_d, fz, mf, False)) / (2 * h)
        assert abs(fd - g_xi[k]) <= 1e-6 * np.abs(g_xi).max(), (k, fd, g_xi[k])
    # the scale component is carried by the depth residual alone (the
    # projection of a scaled point does not move): a real, non-zero term
    assert abs(g_xi[6]) > 1e-3 * np.abs(g_xi).max()
    rng = np.random.default_rng(2)
    dirn = rng.normal(size=n)
    fd = (frozen_energy_sim3(o, ref, fr, lvl, pose, np.zeros(7), h * dirn, fz, mf, False)
          - frozen_energy_sim3(o, ref, fr, lvl, pose, np.zeros(7), -h * dirn, fz, mf, False)) / (2 * h)
    assert abs(fd - g_d @ dirn) <= 1e-6 * np.abs(g_d).sum()


def test_sim3_adjoint_gradient_is_the_lgs7_right_hand_side():
    """dE/dxi is calcSim3LGS's b (LGS6 + LGS4 remapped to {2, 3, 4, 6}, not
    divided by n) and E is half the weighted residual sums, at prec 1 (FP64)
    and prec 2 (FP32 per-point terms, FP64 sums: the GPU's twin)."""
    for prec in (1, 2):
        o, ref, fr, pose = _adjoint_setup_sim3(prec)
        for lvl in (1, 2, 3):
            E, g_xi, _ = o.energy_grad_sim3(ref, fr, lvl, pose)
            ev = o.sim3_evaluate(ref, fr, lvl, pose)
            assert np.allclose(np.array(ev.b[:7]), g_xi, rtol=1e-12, atol=0), (prec, lvl)
            assert abs(E - 0.5 * (ev.sum_res_p + ev.sum_res_d)) <= 1e-12 * abs(E)


def test_adjoint_gradient_is_the_lgs_right_hand_side():
    """E's pose gradient is LGS6's b (the unnormalised sum): the identity
    that ties the adjoint to the tracker's normal equations."""
    o, ref, fr, pose = _adjoint_setup()
    for lvl in (1, 2):
        E, g_xi, _ = o.energy_grad(ref, fr, lvl, pose)
        ev, _ = o.se3_evaluate(ref, fr, lvl, pose)
        assert np.allclose(np.array(ev.b[:6]), g_xi, rtol=1e-12, atol=0)


# ------------------------------------------------------------------ depth map
def _plane_stereo(baseline, z0=2.0, w=160, h=120):
    """Keyframe and a reference frame displaced by `baseline` along x, both
    looking at a smooth texture on a fronto-parallel plane at depth z0."""
    fx = fy = 120.0
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)

    # a sum of sinusoids with incommensurate wavelengths (0.05 - 0.5 m): no
    # repeat within the epipolar search range, so the match is unique
    rng = np.random.default_rng(21)
    freq = rng.uniform(2.0, 20.0, (10, 2)) * rng.choice([-1.0, 1.0], (10, 2))
    amp = rng.uniform(6.0, 14.0, 10)
    phase = rng.uniform(0, 2 * np.pi, 10)

    def tex(X, Y):
        return 128 + sum(a * np.sin(2 * np.pi * (f[0] * X + f[1] * Y) + p) for a, f, p in zip(amp, freq, phase))

    X, Y = (x - cx) * z0 / fx, (y - cy) * z0 / fy
    cam = dict(w=w, h=h, fx=fx, fy=fy, cx=cx, cy=cy)
    return cam, tex(X, Y).astype(np.float32), tex(X + baseline, Y).astype(np.float32)


def _depth_system(prec, cam, kf_img, refs, prior_idepth, prior_var, crop=100.0):
    p = pyoracle_params()
    p.max_epl_length_crop = crop
    o = Oracle(prec, p)
    kf = o.frame(0, cam, kf_img)
    frames = []
    for k, (img, pose8) in enumerate(refs):
        f = o.frame(1 + k, cam, img)
        f.set_tracking(kf, pose8, 0.3, None)
        frames.append(f)
    dm = o.depthmap(cam)
    dm.set_prior(kf, prior_idepth, prior_var, 20)
    return o, kf, frames, dm


def pyoracle_params():
    import pyoracle
    return pyoracle.default_params()


@pytest.mark.parametrize("prec", [0, 1])
def test_rectified_stereo_recovers_the_plane_depth(prec):
    z0, b = 2.0, 0.06
    cam, kfi, refi = _plane_stereo(b, z0)
    n = cam["w"] * cam["h"]
    # epipolar search over 30 px (the upstream crop): a longer line would be
    # clipped at the image border, where doLineStereo gives up (-1)
    o, kf, refs, dm = _depth_system(prec, cam, kfi, [(refi, [0, 0, 0, 1, b, 0, 0, 1.0])],
                                    np.full(n, -1, np.float32), np.full(n, -1, np.float32), crop=30.0)
    dm.update(refs)
    h = dm.read()
    created = h["observe_code"] == lsdg.OBS["CREATED"]
    assert created.sum() > 0.2 * n
    idepth = h["idepth"][created]
    err = np.abs(idepth - 1.0 / z0) * z0
    # the disparity fx * b / z0 = 3.6 px is found to a few hundredths of a
    # pixel on most pixels (a sign or scale slip in the epipolar geometry
    # puts every estimate far off); the rest are ambiguous texture matches
    assert np.median(err) < 0.03
    assert np.mean(err < 0.05) > 0.5


def test_regularisation_keeps_a_constant_patch_constant():
    cam, kfi, refi = _plane_stereo(0.05)
    n = cam["w"] * cam["h"]
    o, kf, refs, dm = _depth_system(0, cam, kfi, [(refi, [0, 0, 0, 1, 0.05, 0, 0, 1.0])],
                                    np.full(n, 0.7, np.float32), np.full(n, 0.01, np.float32))
    for remove_occlusions in (0, 1):
        dm.step(2, remove_occlusions)
        h = dm.read()
        v = h["is_valid"].astype(bool)
        assert v.sum() > 0.9 * n
        # a weighted mean of equal values is that value (to FP32 rounding)
        assert np.abs(h["idepth_smoothed"][v] - 0.7).max() < 1e-6
        assert np.all(h["idepth"][v] == np.float32(0.7))


def test_fill_holes_rule_by_brute_force():
    """Holes fill exactly where the 5x5 sum of validity counters clears the
    threshold (brute-force box sums instead of the integral image) and the
    keyframe max-gradient is large enough; the new hypothesis is the
    inverse-variance mean of its valid neighbours."""
    cam, kfi, refi = _plane_stereo(0.05)
    w, h = cam["w"], cam["h"]
    rng = np.random.default_rng(9)
    valid = rng.uniform(size=(h, w)) < 0.06
    prior = np.where(valid, np.float32(0.5), np.float32(-1)).astype(np.float32).ravel()
    var = np.where(valid, np.float32(0.01), np.float32(-1)).astype(np.float32).ravel()
    o, kf, refs, dm = _depth_system(0, cam, kfi, [(refi, [0, 0, 0, 1, 0.05, 0, 0, 1.0])], prior, var)
    h0 = dm.read()
    before = h0["is_valid"].reshape(h, w).astype(bool)
    counter = np.where(before, h0["validity_counter"].reshape(h, w), 0).astype(np.int64)
    dm.step(1)
    h1 = dm.read()
    after = h1["is_valid"].reshape(h, w).astype(bool)
    p = pyoracle_params()
    mg = kf.read(lsdg.WHAT_MAXGRAD, 0)
    expect = np.zeros((h, w), bool)
    for yy in range(3, h - 2):
        for xx in range(3, w - 2):
            if before[yy, xx] or mg[yy, xx] < p.min_abs_grad_decrease:
                continue
            val = counter[yy - 2:yy + 3, xx - 2:xx + 3].sum()
            # prior hypotheses are not blacklisted (0 >= min_blacklist)
            expect[yy, xx] = (0 >= p.min_blacklist and val > p.val_sum_min_for_create) or \
                val > p.val_sum_min_for_unblacklist
    new = after & ~before
    assert new.any() and (~new & ~before & (mg >= p.min_abs_grad_decrease)).any()
    assert np.array_equal(new, expect)
    assert np.array_equal(after & before, before)     # never removes
    idn = h1["idepth"].reshape(h, w)[new]
    assert np.abs(idn - 0.5).max() < 1e-6
    assert np.all(h1["idepth_var"].reshape(h, w)[new] == np.float32(p.var_random_init_initial))


def test_propagation_along_the_optical_axis():
    """Moving the camera t_z forward maps inverse depth d to 1/(1/d - t_z)
    before the keyframe rescale."""
    cam, kfi, refi = _plane_stereo(0.05)
    n = cam["w"] * cam["h"]
    z0, tz = 2.0, 0.1
    o, kf, refs, dm = _depth_system(1, cam, kfi, [], np.full(n, 1 / z0, np.float32),
                                    np.full(n, 1e-4, np.float32))
    nkf = o.frame(9, cam, kfi)
    nkf.set_tracking(kf, [0, 0, 0, 1, 0, 0, tz, 1.0], 0.3, None)
    dm.create_keyframe(nkf)
    hyp = dm.read_f64()
    pid, t2p, _ = nkf.get_tracking()
    s = t2p[7]                                   # rescale factor of the new keyframe
    v = hyp["idepth"] > 0
    assert v.sum() > 0.5 * n
    # after the rescale every propagated idepth is (1 / (z0 - tz)) * s = 1
    assert np.abs(hyp["idepth"][v] / s - 1.0 / (z0 - tz)).max() < 1e-6


def test_random_init_draws_are_splitmix64():
    # published SplitMix64 outputs (seed 0: first two; seed 1234567: first)
    L = load()
    assert L.orc_splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert L.orc_splitmix64(0, 1) == 0x6E789E6AA1B965F4
    assert L.orc_splitmix64(1234567, 0) == 6457827717110365317


def test_random_init_rule():
    """DepthMap::initializeRandomly: interior pixels above MIN_ABS_GRAD_CREATE
    get idepth = smoothed in [0.5, 1.5], VAR_RANDOM_INIT_INITIAL and
    validity 20; a seed reproduces the map and another seed changes it."""
    cam, kfi, _ = _plane_stereo(0.05)
    w, h = cam["w"], cam["h"]
    o = Oracle(0)
    p = pyoracle_params()
    kf = o.frame(0, cam, kfi)
    maps = []
    for seed in (7, 7, 8):
        dm = o.depthmap(cam)
        dm.init_random(kf, seed)
        maps.append(dm.read())
    a = maps[0]
    v = a["is_valid"].reshape(h, w).astype(bool)
    mg = kf.read(lsdg.WHAT_MAXGRAD, 0)
    expect = np.zeros((h, w), bool)
    expect[1:-1, 1:-1] = mg[1:-1, 1:-1] > p.min_abs_grad_create
    assert np.array_equal(v, expect) and v.any()
    idp = a["idepth"].reshape(h, w)[v]
    assert idp.min() >= 0.5 and idp.max() <= 1.5 and abs(idp.mean() - 1.0) < 0.02
    assert np.array_equal(a["idepth_smoothed"].reshape(h, w)[v], idp)
    assert np.all(a["idepth_var"].reshape(h, w)[v] == np.float32(p.var_random_init_initial))
    assert np.all(a["validity_counter"].reshape(h, w)[v] == 20)
    assert np.array_equal(maps[1]["idepth"], a["idepth"])
    assert not np.array_equal(maps[2]["idepth"], a["idepth"])
