"""Closed-form pins of the oracle functions that the GPU matches bit for bit
but that no other check fixes (the reference ships no fixtures for this
path).  Each compares the oracle with an independent computation chosen so
that a plausible slip fails it:

  * buildIDepthAndIDepthVar: numpy.average with inverse-variance weights and
    scipy.stats.hmean of the valid children's variances (the spec's fused
    inverse depth and output variance), invalid children excluded;
  * calcWeightsAndResidual: the weighted error sum from raw images at the
    identity pose (w_p = 1/sigma_pix^2, Huber knee at huber_d / 2), and with a
    translation, w_p = 1/(sigma_pix^2 + var * (dr/dd)^2) with dr/dd from
    central differences of the projection on an affine frame;
  * doLineStereo's variance on a rectified pair (alpha = 1 / (f b), the
    geometric and photometric disparity errors from the keyframe image) and
    observeDepthUpdate's EKF step (two identical observations leave
    1.01 / 2.01 of one observation's variance);
  * Sim3's depth residual r_d = 1/z' - idepth_frame and the LGS4 -> LGS7
    remap on fronto-parallel planes, where every point shares z'.

CPU only."""
import numpy as np
import pytest
from scipy.stats import hmean

import lsdg
import synth
from pyoracle import Oracle, default_params, sim3_exp, sim3_inv


# ------------------------------------------------- buildIDepthAndIDepthVar
def _fused_level(idepth, var):
    """Level l+1 from level l by the spec's fusion, with library routines."""
    h, w = idepth.shape
    out_i = np.full((h // 2, w // 2), -1.0)
    out_v = np.full((h // 2, w // 2), -1.0)
    for y in range(h // 2):
        for x in range(w // 2):
            d = idepth[2 * y:2 * y + 2, 2 * x:2 * x + 2].ravel()
            v = var[2 * y:2 * y + 2, 2 * x:2 * x + 2].ravel()
            ok = v > 0
            if ok.any():
                out_i[y, x] = np.average(d[ok], weights=1.0 / v[ok])
                out_v[y, x] = hmean(v[ok])
    return out_i, out_v


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 2e-6)])
def test_depth_pyramid_is_inverse_variance_fusion(prec, tol):
    w, h = 64, 48
    rng = np.random.default_rng(31)
    idepth = rng.uniform(0.2, 2.0, (h, w))
    var = rng.uniform(1e-3, 1e-1, (h, w))
    var[rng.uniform(size=(h, w)) < 0.35] = -1.0      # invalid children
    var[0:2, 0:2] = -1.0                              # a block with no valid child
    idepth[var <= 0] = rng.uniform(5.0, 9.0, (var <= 0).sum())  # garbage that must not leak in
    o = Oracle(prec)
    cam = dict(w=w, h=h, fx=60.0, fy=60.0, cx=31.5, cy=23.5)
    f = o.frame(0, cam, rng.uniform(0, 255, (h, w)).astype(np.float32))
    f.set_depth(idepth.astype(np.float32), var.astype(np.float32))
    cur_i = f.read(lsdg.WHAT_IDEPTH, 0, f64=True).astype(np.float64)
    cur_v = f.read(lsdg.WHAT_IDEPTH_VAR, 0, f64=True).astype(np.float64)
    for lvl in range(1, 5):
        exp_i, exp_v = _fused_level(cur_i, cur_v)
        got_i = f.read(lsdg.WHAT_IDEPTH, lvl, f64=True)
        got_v = f.read(lsdg.WHAT_IDEPTH_VAR, lvl, f64=True)
        none = exp_v < 0
        assert np.all(got_i[none] == -1) and np.all(got_v[none] == -1), lvl
        assert np.all(np.abs(got_i[~none] - exp_i[~none]) <= tol * exp_i[~none]), lvl
        assert np.all(np.abs(got_v[~none] - exp_v[~none]) <= tol * exp_v[~none]), lvl
        cur_i, cur_v = got_i.astype(np.float64), got_v.astype(np.float64)
    assert none.sum() == 0 or lvl == 4
    # equal variances: the fused variance is that variance (n / sum(1/v)), not
    # a quarter of it (1 / sum(1/v)); a constant map stays constant
    f2 = o.frame(1, cam, np.zeros((h, w), np.float32))
    f2.set_depth(np.full((h, w), 0.7, np.float32), np.full((h, w), 0.02, np.float32))
    for lvl in range(5):
        assert np.all(np.abs(f2.read(lsdg.WHAT_IDEPTH, lvl, f64=True) - 0.7) <= 1e-6)
        assert np.all(np.abs(f2.read(lsdg.WHAT_IDEPTH_VAR, lvl, f64=True) - 0.02) <= 1e-8)


# ------------------------------------------------- calcWeightsAndResidual
def _huber_error(r, w_p, k):
    """sum wh * w_p * r^2 with wh = min(1, k / |r sqrt(w_p)|)."""
    e = np.abs(r) * np.sqrt(w_p)
    wh = np.where(e < k, 1.0, k / np.maximum(e, 1e-300))
    return float(np.sum(wh * w_p * r * r))


def test_weights_at_identity_pose_from_raw_images():
    """At the identity pose every point projects onto its own pixel (bilinear
    taps at integer positions), the translation is zero so dr/dd = 0 and
    w_p = 1 / cameraPixelNoise2 = 1/16; the error is then a closed form of
    the two raw images, with the Huber knee at huber_d / 2 = 1.5 (|r| = 6)."""
    inst = synth.tracking_instance(1013, cam=synth.camera(160, 120), trans_m=0.05, rot_deg=2.0)
    p = default_params()
    o = Oracle(1, params=p)
    kf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
    kf.set_depth(inst["idepth"], inst["var"])
    fr = o.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
    ref = o.ref(kf)
    for lvl in (1, 2):
        pts = ref.points(lvl)
        w, h = kf.size(lvl)
        idx = pts[:, 7].astype(np.int64)
        x, y = idx % w, idx // w
        inside = (x > 1) & (y > 1) & (x < w - 2) & (y < h - 2)  # u > 1 && u < w - 2, strict
        I_fr = fr.read(lsdg.WHAT_IMAGE, lvl, f64=True).ravel()
        r = pts[inside, 3].astype(np.float64) - I_fr[idx[inside]]
        expect = _huber_error(r, np.full(r.shape, 1.0 / p.camera_pixel_noise2), p.huber_d / 2.0)
        ev, _ = o.se3_evaluate(ref, fr, lvl, synth.IDENTITY7)
        assert ev.warped_count == inside.sum()
        assert (np.abs(r) >= 6.0).any() and (np.abs(r) < 6.0).any()  # both sides of the knee
        assert abs(ev.error - expect) <= 1e-9 * expect, (lvl, ev.error, expect)


def test_weights_with_translation_use_the_inverse_depth_derivative():
    """With a translation the weight is w_p = 1/(sigma^2 + var * (dr/dd)^2),
    dr/dd the derivative of the residual with respect to the point's inverse
    depth.  On an affine frame image that derivative is the image slope
    times d(u, v)/dd, taken here by central differences of the projection of
    the point moved along its ray."""
    inst = synth.tracking_instance(1014, cam=synth.camera(160, 120))
    cam = inst["cam"]
    p = default_params()
    p.var_weight = 300.0  # large enough that the variance term rivals sigma^2
    o = Oracle(1, params=p)
    kf = o.frame(0, cam, inst["kf_img"].astype(np.float32))
    kf.set_depth(inst["idepth"], inst["var"])
    a, bx, by = 50.0, 0.8, -0.35
    yy, xx = np.mgrid[0:cam["h"], 0:cam["w"]].astype(np.float64)
    fr = o.frame(1, cam, (a + bx * xx + by * yy + 60.0).astype(np.float32))
    ref = o.ref(kf)
    T = sim3_exp(np.array([0.03, -0.02, 0.015, 0.004, -0.002, 0.003, 0.0]))[:7]
    from posemath import quat_to_R
    R, t = quat_to_R(T[:4]), np.array(T[4:7])
    for lvl in (1, 2):
        s = 1 << lvl
        fx, fy = cam["fx"] / s, cam["fy"] / s
        cx, cy = (cam["cx"] + 0.5) / s - 0.5, (cam["cy"] + 0.5) / s - 0.5
        w, h = kf.size(lvl)
        pts = ref.points(lvl).astype(np.float64)
        P = pts[:, :3]

        def project(d):
            q = (P * (pts[:, 2] / (1.0 / d))[:, None]) @ R.T + t  # the point at inverse depth d on its ray
            return q[:, 0] / q[:, 2] * fx + cx, q[:, 1] / q[:, 2] * fy + cy

        d0 = 1.0 / pts[:, 2]
        u, v = project(d0)
        inside = (u > 1) & (v > 1) & (u < w - 2) & (v < h - 2)
        # level-l image of the affine frame: the same plane with the 2x2
        # mean's half-pixel shift, slope s times the level-0 slope
        I_l = a + 60.0 + bx * (s * u + (s - 1) / 2.0) + by * (s * v + (s - 1) / 2.0)
        r = pts[:, 3] - I_l
        hd = 1e-7
        up, vp = project(d0 + hd)
        um, vm = project(d0 - hd)
        drdd = -(bx * s * (up - um) + by * s * (vp - vm)) / (2 * hd)
        w_p = 1.0 / (p.camera_pixel_noise2 + p.var_weight * pts[:, 4] * drdd * drdd)
        expect = _huber_error(r[inside], w_p[inside], p.huber_d / 2.0)
        ev, _ = o.se3_evaluate(ref, fr, lvl, T)
        assert ev.warped_count == inside.sum()
        assert np.median(p.var_weight * pts[inside, 4] * drdd[inside] ** 2) > 0.5 * p.camera_pixel_noise2
        assert abs(ev.error - expect) <= 1e-6 * expect, (lvl, ev.error, expect)


# ------------------------------------------------ Sim3 depth residual
def test_sim3_depth_residual_on_fronto_parallel_planes():
    """Keyframe and frame both see constant-intensity fronto-parallel planes
    (no photometric term: r_p = 0 and zero image gradient); a Sim3 of scale
    s and translation t maps every keyframe point to the same depth
    z' = s Z_k + t_z, so r_d = 1/z' - 1/Z_f and w_d = 1/(var_f + g2^2 var_k),
    g2 = s Z_k^2 / z'^2, are the same for every point, and the LGS4 rows of
    v4 = (1/z'^2, y'/z'^2, -x'/z'^2, 1/z') land at twist rows 2, 3, 4, 6."""
    w, h = 64, 48
    cam = dict(w=w, h=h, fx=60.0, fy=60.0, cx=31.5, cy=23.5)
    Zk, Zf, vk, vf = 2.0, 2.4, 4e-3, 6e-3
    p = default_params()
    o = Oracle(1, params=p)
    img = np.full((h, w), 100.0, np.float32)
    kf = o.frame(0, cam, img)
    kf.set_depth(np.full((h, w), 1.0 / Zk, np.float32), np.full((h, w), vk, np.float32))
    fr = o.frame(1, cam, img)
    fr.set_depth(np.full((h, w), 1.0 / Zf, np.float32), np.full((h, w), vf, np.float32))
    ref = o.ref(kf)
    s, tx, ty, tz = 1.05, 0.01, -0.008, 0.06
    pose = np.array([0, 0, 0, 1, tx, ty, tz, s])
    lvl = 1
    ev = o.sim3_evaluate(ref, fr, lvl, pose)
    pts = ref.points(lvl).astype(np.float64)
    sc = 1 << lvl
    fx, cx = cam["fx"] / sc, (cam["cx"] + 0.5) / sc - 0.5
    fy, cy = cam["fy"] / sc, (cam["cy"] + 0.5) / sc - 0.5
    Wx, Wy, Wz = s * pts[:, 0] + tx, s * pts[:, 1] + ty, s * pts[:, 2] + tz
    u, v = Wx / Wz * fx + cx, Wy / Wz * fy + cy
    inside = (u > 1) & (v > 1) & (u < (w >> lvl) - 2) & (v < (h >> lvl) - 2)
    assert inside.sum() > 0.5 * len(pts)
    # the maps are handed over as float32: the closed form uses those values
    idk, idf = float(np.float32(1.0 / Zk)), float(np.float32(1.0 / Zf))
    vk, vf = float(np.float32(vk)), float(np.float32(vf))
    zp = s / idk + tz
    assert np.allclose(Wz, zp)
    rd = 1.0 / zp - idf
    g2 = s / (idk * idk * zp * zp)
    wd = 1.0 / (vf + g2 * g2 * vk)
    assert abs(rd) * np.sqrt(wd) < p.huber_d  # inside the Huber knee: wh = 1
    n = inside.sum()
    assert ev.num_terms_d == n and ev.warped_count == n
    assert abs(ev.sum_res_d - n * wd * rd * rd) <= 1e-9 * n * wd * rd * rd
    assert ev.sum_res_p < 1e-20  # bilinear weights of a constant image sum to 1 within rounding
    X, Y = Wx[inside], Wy[inside]
    v4 = np.stack([np.full(n, 1 / zp ** 2), Y / zp ** 2, -X / zp ** 2, np.full(n, 1 / zp)], 1)
    b = np.zeros(7)
    b[[2, 3, 4, 6]] = -(wd * rd) * v4.sum(0)
    A = np.zeros((7, 7))
    A[np.ix_([2, 3, 4, 6], [2, 3, 4, 6])] = wd * v4.T @ v4
    # x', y' come from the point records read back as float32: 1e-6
    assert np.abs(np.array(ev.b[:7]) - b).max() <= 1e-6 * np.abs(b).max(), (np.array(ev.b[:7]), b)
    assert np.abs(ev.A7() - A).max() <= 1e-6 * np.abs(A).max()


# ------------------------------------- doLineStereo variance and EKF update
def _rectified_pair(baseline, z0=2.0, w=160, h=120):
    """A keyframe and a reference frame displaced by `baseline` along x, both
    looking at a fronto-parallel plane at z0 with a sum-of-sinusoids texture
    (no repeat within the search range)."""
    fx = fy = 120.0
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    rng = np.random.default_rng(21)
    freq = rng.uniform(2.0, 20.0, (10, 2)) * rng.choice([-1.0, 1.0], (10, 2))
    amp = rng.uniform(6.0, 14.0, 10)
    phase = rng.uniform(0, 2 * np.pi, 10)

    def tex(X, Y):
        return 128 + sum(a * np.sin(2 * np.pi * (f[0] * X + f[1] * Y) + p) for a, f, p in zip(amp, freq, phase))

    X, Y = (x - cx) * z0 / fx, (y - cy) * z0 / fy
    cam = dict(w=w, h=h, fx=fx, fy=fy, cx=cx, cy=cy)
    return cam, tex(X, Y).astype(np.float32), tex(X + baseline, Y).astype(np.float32)


def test_stereo_variance_and_ekf_update_on_a_rectified_pair():
    """Rectified stereo (R = I, t = (b, 0, 0)), sub-pixel step off: the
    search steps are exactly +-1 px along x and the sampling rescale is 1, so
    a created hypothesis' variance is
        alpha^2 (0.5 + geo + photo),  alpha = 1 / (f_x b),
        geo   = tf^2 |g|^2 / g_x^2,   tf = (1 + initialTrackedResidual) / 4,
        photo = 4 sigma_pix^2 / sum_k (I[x+k+1] - I[x+k])^2  (k = -2..1),
    g the keyframe's central-difference gradient at the pixel: all from the
    keyframe image.  A second, identical observation then goes through the
    EKF update: prior variance inflated by 1.01, Kalman gain
    K = 1.01 v / (1.01 v + v), posterior (1 - K) 1.01 v = 1.01 / 2.01 v."""
    z0, b, itr = 2.0, 0.06, 0.3
    cam, kfi, refi = _rectified_pair(b, z0)
    w, h = cam["w"], cam["h"]
    n = w * h
    p = default_params()
    p.max_epl_length_crop = 30.0
    p.use_subpixel_stereo = 0
    o = Oracle(1, params=p)
    kf = o.frame(0, cam, kfi)
    refs = []
    for k in (1, 2):
        f = o.frame(k, cam, refi)
        f.set_tracking(kf, [0, 0, 0, 1, b, 0, 0, 1.0], itr, None)
        refs.append(f)
    dm = o.depthmap(cam)
    dm.set_prior(kf, np.full(n, -1, np.float32), np.full(n, -1, np.float32), 20)
    dm.update([refs[0]])
    h1 = dm.read_f64()
    codes = dm.read()["observe_code"]
    created = codes == lsdg.OBS["CREATED"]
    assert created.sum() > 0.2 * n
    I = kfi.astype(np.float64)
    ys, xs = np.divmod(np.nonzero(created)[0], w)
    gx = 0.5 * (I[ys, xs + 1] - I[ys, xs - 1])
    gy = 0.5 * (I[ys + 1, xs] - I[ys - 1, xs])
    gal = sum((I[ys, xs + k + 1] - I[ys, xs + k]) ** 2 for k in range(-2, 2))
    tf = 0.25 * (1.0 + np.float64(np.float32(itr)))
    geo = tf * tf * (gx * gx + gy * gy) / (np.abs(gx) + p.division_eps) ** 2
    photo = 4.0 * p.camera_pixel_noise2 / (gal + p.division_eps)
    alpha = 1.0 / (cam["fx"] * b)
    expect = alpha * alpha * (0.5 + geo + photo)
    got = h1["idepth_var"][created]
    assert np.all(np.abs(got - expect) <= 1e-6 * expect), np.abs(got / expect - 1).max()
    # the second observation: EKF update of every created hypothesis that the
    # new search confirms
    dm.update([refs[1]])
    codes2 = dm.read()["observe_code"]
    upd = created & (codes2 == lsdg.OBS["UPDATED"])
    assert upd.sum() > 0.5 * created.sum()
    var2 = dm.read_f64()["idepth_var"][upd]
    v0 = h1["idepth_var"][upd]
    q = float(np.float32(p.succ_var_inc_fac))  # 1.01 as the float parameter holds it
    assert np.all(np.abs(var2 - v0 * q / (1.0 + q)) <= 1e-12 * v0), np.abs(var2 / v0 - q / (1.0 + q)).max()
