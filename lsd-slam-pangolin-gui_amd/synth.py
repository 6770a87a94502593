"""Seeded synthetic LSD-SLAM workloads (textured-plane scenes, poses, noise).

Thin numpy layer over liblsdg_synth.so (csrc/synth.cpp).  Every builder is a
pure function of its seed, so the GPU box regenerates the same inputs.
Workload seeds are 1000 + workload index.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

IDENTITY7 = (0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)
# texture contrast giving ~25% of 640x480 pixels with |grad I| > 5
CONTRAST = 0.53


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liblsdg_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.synth_scene_size.restype = C.c_size_t
        L.synth_scene_init.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double]
        L.synth_render.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_int] + [C.c_double] * 5 + [
            C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.synth_sample_pose.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.synth_uniform.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, C.c_void_p]
        L.synth_gaussian.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_void_p]
        _LIB = L
    return _LIB


def camera(w, h):
    """Pinhole camera of the 640x480 workloads (fx=fy=525, cx=319.5, cy=239.5), scaled by w/640."""
    s = w / 640.0
    return dict(w=w, h=h, fx=525.0 * s, fy=525.0 * s, cx=(319.5 + 0.5) * s - 0.5, cy=(239.5 + 0.5) * s - 0.5)


def camera_1080():
    """1920x1080 camera; f from the reference's sample calibration (d2_camera.xml:6)."""
    return dict(w=1920, h=1080, fx=1430.15, fy=1430.15, cx=959.5, cy=539.5)


class Scene:
    def __init__(self, seed, contrast=CONTRAST, freq_scale=1.0):
        L = lib()
        self.buf = C.create_string_buffer(L.synth_scene_size())
        L.synth_scene_init(self.buf, C.c_uint64(seed), contrast, freq_scale)

    def render(self, cam_to_world, cam, noise_sigma=0.0, noise_seed=0, intensity=False):
        w, h = cam["w"], cam["h"]
        img = np.zeros((h, w), np.uint8)
        dep = np.zeros((h, w), np.float32)
        inten = np.zeros((h, w), np.float32) if intensity else None
        P = (C.c_double * 7)(*[float(v) for v in cam_to_world])
        lib().synth_render(self.buf, P, w, h, cam["fx"], cam["fy"], cam["cx"], cam["cy"], float(noise_sigma),
                           C.c_uint64(noise_seed), None if inten is None else inten.ctypes.data,
                           img.ctypes.data, dep.ctypes.data)
        if intensity:
            return img, dep, inten
        return img, dep


def sample_pose(seed, trans_m, rot_deg, uniform=False):
    out = (C.c_double * 7)()
    lib().synth_sample_pose(C.c_uint64(seed), trans_m, np.deg2rad(rot_deg), int(uniform), out)
    return np.array(out[:], np.float64)


def uniform_field(seed, n, lo, hi):
    out = np.zeros(n, np.float32)
    lib().synth_uniform(C.c_uint64(seed), n, lo, hi, out.ctypes.data)
    return out


def gaussian_field(seed, n, sigma):
    out = np.zeros(n, np.float32)
    lib().synth_gaussian(C.c_uint64(seed), n, sigma, out.ctypes.data)
    return out


def gradient_magnitude(img):
    """|grad I| with the pyramid's central half differences (0 on the border)."""
    I = img.astype(np.float32)
    g = np.zeros_like(I)
    gx = np.float32(0.5) * (I[1:-1, 2:] - I[1:-1, :-2])
    gy = np.float32(0.5) * (I[2:, 1:-1] - I[:-2, 1:-1])
    g[1:-1, 1:-1] = np.sqrt(gx * gx + gy * gy)
    return g


def semi_dense_depth(img, depth, seed, threshold=None, fraction=None, var_range=(1e-3, 1e-2)):
    """Keyframe inverse depth on pixels with |grad I| > threshold (or the
    `fraction` highest-gradient pixels); variance ~ U[var_range]; -1 elsewhere."""
    g = gradient_magnitude(img)
    if threshold is None:
        threshold = float(np.quantile(g[1:-1, 1:-1], 1.0 - fraction))
    mask = g > threshold
    var = uniform_field(seed, img.size, var_range[0], var_range[1]).reshape(img.shape)
    idepth = np.where(mask, (1.0 / depth).astype(np.float32), np.float32(-1))
    var = np.where(mask, var, np.float32(-1)).astype(np.float32)
    return idepth.astype(np.float32), var, float(mask.mean()), threshold


def tracking_instance(seed, cam=None, trans_m=0.02, rot_deg=1.0, uniform=False, noise=2.0, threshold=5.0,
                      fraction=None, contrast=CONTRAST):
    """Keyframe (noise-free render, semi-dense idepth) + new frame at a seeded
    perturbation with N(0, noise^2) sensor noise.  gt is frameToRef."""
    cam = cam or camera(640, 480)
    sc = Scene(seed, contrast=contrast)
    kf_img, kf_dep = sc.render(IDENTITY7, cam)
    pose = sample_pose(seed * 7919 + 1, trans_m, rot_deg, uniform)
    fr_img, _ = sc.render(pose, cam, noise, seed * 7919 + 2)
    idepth, var, frac, thr = semi_dense_depth(kf_img, kf_dep, seed * 7919 + 3, threshold, fraction)
    return dict(cam=cam, kf_img=kf_img, fr_img=fr_img, kf_depth=kf_dep, idepth=idepth, var=var, gt=pose,
                fraction=frac, threshold=thr, seed=seed)


def workload1(seed=1001):
    """Plain SE3 alignment at 640x480: ~25% points, 2 cm / 1 deg, noise sigma 2."""
    return tracking_instance(seed)


def depth_instance(seed=1002, cam=None, baseline=0.05, prior_fraction=0.30, noise=2.0):
    """DepthMap workload: keyframe with a prior on the 30% highest-maxGrad
    pixels (true idepth + N(0, 0.1^2), clamped to [0.2, 2.0], var 1e-2), a
    reference frame at a `baseline` translation, and a second frame further
    along for propagation.  Poses are camToWorld with the keyframe at the origin."""
    cam = cam or camera(640, 480)
    sc = Scene(seed)
    kf_img, kf_dep = sc.render(IDENTITY7, cam)
    d = sample_pose(seed * 7919 + 11, 1.0, 0.0)[4:7]  # a unit direction
    d = d / np.linalg.norm(d)
    d[2] *= 0.3  # keep the baseline mostly lateral, like a moving camera
    d = d / np.linalg.norm(d)
    rot = sample_pose(seed * 7919 + 12, 0.0, 0.5)
    ref_pose = np.array([rot[0], rot[1], rot[2], rot[3], *(baseline * d)])
    ref_img, _ = sc.render(ref_pose, cam, noise, seed * 7919 + 13)
    rot2 = sample_pose(seed * 7919 + 14, 0.0, 1.0)
    kf2_pose = np.array([rot2[0], rot2[1], rot2[2], rot2[3], *(2.0 * baseline * d)])
    kf2_img, kf2_dep = sc.render(kf2_pose, cam, noise, seed * 7919 + 15)
    # maxGrad-ranked prior
    g = gradient_magnitude(kf_img)
    from_max = _max3x3(g)
    thr = float(np.quantile(from_max[3:-3, 3:-3], 1.0 - prior_fraction))
    mask = from_max > thr
    mask[:3, :] = mask[-3:, :] = False
    mask[:, :3] = mask[:, -3:] = False
    noise_id = gaussian_field(seed * 7919 + 16, kf_img.size, 0.1).reshape(kf_img.shape)
    prior = np.clip(1.0 / kf_dep + noise_id, 0.2, 2.0).astype(np.float32)
    prior = np.where(mask, prior, np.float32(-1)).astype(np.float32)
    pvar = np.where(mask, np.float32(1e-2), np.float32(-1)).astype(np.float32)
    return dict(cam=cam, kf_img=kf_img, kf_depth=kf_dep, ref_img=ref_img, ref_pose=ref_pose, kf2_img=kf2_img,
                kf2_pose=kf2_pose, kf2_depth=kf2_dep, prior_idepth=prior, prior_var=pvar,
                prior_fraction=float(mask.mean()), seed=seed)


def _max3x3(g):
    m = g.copy()
    m[1:-1, :] = np.maximum(np.maximum(g[:-2, :], g[1:-1, :]), g[2:, :])
    m2 = m.copy()
    m2[:, 1:-1] = np.maximum(np.maximum(m[:, :-2], m[:, 1:-1]), m[:, 2:])
    return m2


def sim3_instance(seed, cam=None, max_trans=0.10, max_rot_deg=5.0, scale_range=(0.9, 1.1), fraction=0.25,
                  noise=2.0):
    """Keyframe-to-keyframe Sim3 alignment (the constraint search's job): two
    keyframes of one scene, each with semi-dense inverse depth on its
    `fraction` highest-gradient pixels; the second keyframe's depth map is in
    its own scale, depth / s.  gt is frameToRef (qx, qy, qz, qw, tx, ty, tz, s):
    X_ref = s R X_frame + t."""
    cam = cam or camera_1080()
    sc = Scene(seed)
    a_img, a_dep = sc.render(IDENTITY7, cam)
    pose = sample_pose(seed * 7919 + 31, max_trans, max_rot_deg, uniform=True)
    s = float(uniform_field(seed * 7919 + 32, 1, scale_range[0], scale_range[1])[0])
    b_img, b_dep = sc.render(pose, cam, noise, seed * 7919 + 33)
    a_id, a_var, a_frac, _ = semi_dense_depth(a_img, a_dep, seed * 7919 + 34, fraction=fraction)
    b_id, b_var, b_frac, _ = semi_dense_depth(b_img, (b_dep / s).astype(np.float32), seed * 7919 + 35,
                                              fraction=fraction)
    return dict(cam=cam, a_img=a_img, a_idepth=a_id, a_var=a_var, b_img=b_img, b_idepth=b_id, b_var=b_var,
                gt=np.r_[pose, s], fraction=(a_frac, b_frac), seed=seed)


RESOLUTIONS = [(320, 240), (640, 480), (1280, 960)]


def batch_instances(n=64, seed=1003, max_trans=0.05, max_rot_deg=3.0, resolutions=None, fractions=(0.15, 0.35)):
    """Batched-tracking workload: n independent alignments with a seeded split
    over `resolutions`, point fraction ~ U[fractions], perturbations up to
    max_trans / max_rot_deg."""
    resolutions = resolutions or RESOLUTIONS
    out = []
    split = uniform_field(seed * 7919 + 21, n, 0.0, float(len(resolutions)))
    fr = uniform_field(seed * 7919 + 22, n, fractions[0], fractions[1])
    for i in range(n):
        w, h = resolutions[min(int(split[i]), len(resolutions) - 1)]
        inst = tracking_instance(seed * 1000 + i, cam=camera(w, h), trans_m=max_trans, rot_deg=max_rot_deg,
                                 uniform=True, threshold=None, fraction=float(fr[i]))
        out.append(inst)
    return out


def vo_sequence(seed, n, cam=None, step=(0.004, 0.001, 0.012), yaw_deg=0.15, noise=1.0, fraction=0.3):
    """A camera moving steadily through one scene (the SlamSystem workload):
    frame k at camToWorld (0, sin(k a / 2), 0, cos(k a / 2), k * step) with
    a = yaw_deg per frame; 8-bit images with sensor noise, the first frame's
    semi-dense inverse depth (the gtDepthInit input) and every camToWorld."""
    cam = cam or camera(320, 240)
    sc = Scene(seed)
    imgs, poses = [], []
    for k in range(n):
        a = np.deg2rad(yaw_deg) * k
        pose = np.array([0.0, np.sin(a / 2), 0.0, np.cos(a / 2), k * step[0], k * step[1], k * step[2]])
        img, dep = sc.render(pose, cam, noise, seed * 7919 + k)
        if k == 0:
            idepth, var, _, _ = semi_dense_depth(img, dep, seed * 7919 + 999, fraction=fraction)
        imgs.append(img)
        poses.append(pose)
    return dict(cam=cam, images=imgs, poses=np.array(poses), idepth=idepth, var=var)
