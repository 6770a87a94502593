"""GPU parity of the adjoint kernels (lsdg_energy_grad, lsdg_energy_grad_sim3)
on the adjoint workload: a 1280x960 keyframe with ~300k semi-dense points and a
frame at a 3 cm / 2 deg perturbation, plus Sim3 keyframe pairs.

Bars (FP32 pyramids and per-point terms, FP64 sums):
  * against the oracle's twin (prec 2: FP32 terms, FP64 sums), at every
    pose: dE/dd_i bit for bit, E and dE/dxi within 1e-9 relative;
  * against the FP64 oracle (prec 1), at poses off the optimum and off the
    identity: SE3 E and dE/dxi within 1e-5 relative (dE/dxi against its
    largest component), dE/dd_i within 1e-4 in relative L1 norm; Sim3 within
    1e-4 and 5e-4.  These are the floors of FP32 per-point arithmetic, which
    the reference computes in: the projection u of a 1280-pixel image carries
    ~3e-5 px of rounding, times the image gradient in every residual, and
    Sim3's depth residual 1/z' - idepth_frame cancels.  At the identity pose
    points project exactly onto the strict in-image test (u > 1), and at the
    optimum dE/dxi is a near-total cancellation: there only the twin applies;
  * against FP64 central finite differences (h = 1e-6) of the frozen-weight
    energy, on a frame whose image is affine (where the pyramid's central
    differences are the exact derivative of the bilinear sample): within
    1e-5 of the largest component (twist) and of sum |dE/dd_i d_i| (one
    separable direction d covers every dE/dd_i);
  * the tracker's normal equations: dE/dxi equals lsdg_se3_evaluate's b
    (LGS6) and lsdg_sim3_evaluate's b (LGS7 with ESM off) to 1e-12."""
import numpy as np
import pytest

import lsdg
import pyoracle
import synth
from helpers import central_differences, frozen_energy, frozen_energy_sim3, gpu_pair, oracle_pair

pytestmark = pytest.mark.gpu


def _config4():
    return synth.tracking_instance(1004, cam=synth.camera(1280, 960), trans_m=0.03, rot_deg=2.0, threshold=None,
                                   fraction=0.25)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _ref_to_frame(inst):
    """The perturbation the frame was rendered at, as referenceToFrame."""
    return pyoracle.sim3_inv(np.r_[inst["gt"], 1.0])[:7]


def _by_pixel(ref, level, gd):
    """Per-point values in pixel-index order: the GPU point cloud is
    row-major, the oracle's is column by column (makePointCloud's loops)."""
    idx = ref.points(level)[:, 7].astype(np.int64)
    return np.asarray(gd)[np.argsort(idx, kind="stable")]


def _l1(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / np.abs(b).sum())


@pytest.mark.parametrize("level", [0, 1, 2])
def test_se3_adjoint_parity_1280x960(ctx, level):
    inst = _config4()
    kf, fr, ref = gpu_pair(ctx, inst)
    n = ref.num_points(level)
    if level == 0:
        assert 250000 < n < 350000, n
    o2, o1 = pyoracle.Oracle(2), pyoracle.Oracle(1)
    _, ofr2, oref2 = oracle_pair(o2, inst)
    _, ofr1, oref1 = oracle_pair(o1, inst)
    true = pyoracle.sim3_inv(np.r_[inst["gt"], 1.0])
    generic = [pyoracle.sim3_mul(pyoracle.sim3_exp(np.r_[0.004, -0.003, 0.002, 0.003, 0.002, -0.001, 0]), true)[:7],
               pyoracle.sim3_exp(np.r_[0.002, -0.001, 0.0015, 0.001, -0.0012, 0.0007, 0])[:7]]
    for k, pose in enumerate([synth.IDENTITY7, true[:7]] + generic):
        E, g, gd = ctx.energy_grad(ref, fr, level, pose)
        gd = _by_pixel(ref, level, gd)
        E2, g2, gd2 = o2.energy_grad(oref2, ofr2, level, pose)
        gd2 = _by_pixel(oref2, level, gd2)
        assert abs(E - E2) <= 1e-9 * abs(E2)
        assert _rel(g, g2) <= 1e-9, (g, g2)
        assert np.array_equal(gd, gd2.astype(np.float32)), np.abs(gd - gd2).max()
        # the tracker's unnormalised LGS6 right-hand side is the same reduction
        ev, _ = ctx.se3_evaluate(ref, fr, level, pose)
        assert _rel(g, ev.b[:6]) <= 1e-12, (g, ev.b[:6])
        if k >= 2:
            E1, g1, gd1 = o1.energy_grad(oref1, ofr1, level, pose)
            gd1 = _by_pixel(oref1, level, gd1)
            assert abs(E - E1) <= 1e-5 * abs(E1), (k, E, E1)
            assert _rel(g, g1) <= 1e-5, (k, g, g1)
            assert _l1(gd, gd1) <= 1e-4, (k, _l1(gd, gd1))


def _affine_frame(cam):
    y, x = np.mgrid[0:cam["h"], 0:cam["w"]].astype(np.float64)
    return (60.0 + 0.9 * x + 0.4 * y).astype(np.float32)


@pytest.mark.parametrize("level", [0, 1])
def test_se3_adjoint_matches_fp64_central_differences(ctx, level):
    """dE/dxi and every dE/dd_i of the GPU kernel (FP32 terms) against FP64
    central differences of the frozen-weight energy at 1280x960."""
    inst = _config4()
    cam = inst["cam"]
    img = _affine_frame(cam)
    kf = ctx.frame(0, cam, inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    fr = ctx.frame(1, cam, img)
    ref = ctx.ref(kf)
    pose = _ref_to_frame(inst)
    E, g, gd = ctx.energy_grad(ref, fr, level, pose)
    o = pyoracle.Oracle(1)
    okf = o.frame(0, cam, inst["kf_img"].astype(np.float32))
    okf.set_depth(inst["idepth"], inst["var"])
    ofr = o.frame(1, cam, img)
    oref = o.ref(okf)
    n = oref.num_points(level)
    assert n == ref.num_points(level)
    wf, mf = np.zeros(n), np.zeros(n, np.uint8)
    zero_d = np.zeros(n)
    E0 = frozen_energy(o, oref, ofr, level, pose, np.zeros(6), zero_d, wf, mf, True)
    assert mf.sum() > 0.8 * n
    assert abs(E - E0) <= 1e-5 * abs(E0)
    fd = central_differences(lambda e: frozen_energy(o, oref, ofr, level, pose, e, zero_d, wf, mf, False), 6)
    assert np.abs(g - fd).max() <= 1e-5 * np.abs(fd).max(), (g, fd)
    # one direction d over the oracle's points; the GPU's dE/dd in that order
    dirn = np.random.default_rng(4).normal(size=n)
    gdo = np.empty(n, np.float64)
    gdo[np.argsort(oref.points(level)[:, 7].astype(np.int64), kind="stable")] = _by_pixel(ref, level, gd)
    h = 1e-6
    fdd = (frozen_energy(o, oref, ofr, level, pose, np.zeros(6), h * dirn, wf, mf, False)
           - frozen_energy(o, oref, ofr, level, pose, np.zeros(6), -h * dirn, wf, mf, False)) / (2 * h)
    assert abs(fdd - gdo @ dirn) <= 1e-5 * np.abs(gdo * dirn).sum(), (fdd, gdo @ dirn)


def _sim3_gpu_oracle(ctx, d, prec):
    a = ctx.frame(0, d["cam"], d["a_img"])
    a.set_depth(d["a_idepth"], d["a_var"])
    b = ctx.frame(1, d["cam"], d["b_img"])
    b.set_depth(d["b_idepth"], d["b_var"])
    o = pyoracle.Oracle(prec)
    oa = o.frame(0, d["cam"], d["a_img"].astype(np.float32))
    oa.set_depth(d["a_idepth"], d["a_var"])
    ob = o.frame(1, d["cam"], d["b_img"].astype(np.float32))
    ob.set_depth(d["b_idepth"], d["b_var"])
    return (ctx.ref(a), b), (o, o.ref(oa), ob)


@pytest.mark.parametrize("level", [0, 2])
def test_sim3_adjoint_parity_1920x1080(ctx, level):
    """Sim3 adjoint on a keyframe pair of the Sim3 workload (1920x1080, ~500k
    points per keyframe, scale drift): twin bit for bit on dE/dd at the
    solution and off it, FP64 off it, and dE/dxi equal to the LGS7
    right-hand side with ESM off."""
    d = synth.sim3_instance(1005)
    true = pyoracle.sim3_inv(d["gt"])
    off = pyoracle.sim3_mul(pyoracle.sim3_exp(np.r_[0.004, -0.003, 0.002, 0.003, 0.002, -0.001, 0.01]), true)
    (ref, fb), (o2, oref2, ob2) = _sim3_gpu_oracle(ctx, d, 2)
    _, (o1, oref1, ob1) = _sim3_gpu_oracle(ctx, d, 1)
    assert ref.num_points(0) > 400000
    p = lsdg.default_params()
    p.use_esm_tracking = 0
    for k, pose in enumerate((true, off)):
        E, g, gd = ctx.energy_grad_sim3(ref, fb, level, pose)
        gd = _by_pixel(ref, level, gd)
        E2, g2, gd2 = o2.energy_grad_sim3(oref2, ob2, level, pose)
        gd2 = _by_pixel(oref2, level, gd2)
        assert abs(E - E2) <= 1e-9 * abs(E2)
        assert _rel(g, g2) <= 1e-9, (g, g2)
        assert np.array_equal(gd, gd2.astype(np.float32)), np.abs(gd - gd2).max()
        ctx.set_params(p)
        try:
            ev = ctx.sim3_evaluate(ref, fb, level, pose)
        finally:
            ctx.set_params(lsdg.default_params())
        assert _rel(g, ev.b[:7]) <= 1e-12, (g, ev.b[:7])
        assert abs(E - 0.5 * (ev.sum_res_p + ev.sum_res_d)) <= 1e-12 * abs(E)
        if k == 1:
            E1, g1, gd1 = o1.energy_grad_sim3(oref1, ob1, level, pose)
            gd1 = _by_pixel(oref1, level, gd1)
            assert abs(E - E1) <= 1e-4 * abs(E1), (E, E1)
            assert _rel(g, g1) <= 1e-4, (g, g1)
            assert _l1(gd, gd1) <= 5e-4, _l1(gd, gd1)


def test_sim3_adjoint_matches_fp64_central_differences(ctx):
    """Sim3 adjoint (7 twist components including scale, and every dE/dd_i)
    against FP64 central differences, 640x480, affine frame image with an
    inverse-depth ramp that has a hole (points without a depth residual)."""
    cam = synth.camera(640, 480)
    inst = synth.tracking_instance(1011, cam=cam, threshold=None, fraction=0.25)
    img = _affine_frame(cam)
    y, x = np.mgrid[0:cam["h"], 0:cam["w"]].astype(np.float64)
    fid = (0.35 + 5e-4 * x - 2.5e-4 * y).astype(np.float32)
    fvar = np.full(fid.shape, 2e-3, np.float32)
    fvar[160:280, 240:400] = -1.0
    pose = pyoracle.sim3_exp(np.array([0.01, -0.008, 0.005, 0.004, -0.003, 0.002, 0.03]))
    kf = ctx.frame(0, cam, inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    fr = ctx.frame(1, cam, img)
    fr.set_depth(fid, fvar)
    ref = ctx.ref(kf)
    o = pyoracle.Oracle(1)
    okf = o.frame(0, cam, inst["kf_img"].astype(np.float32))
    okf.set_depth(inst["idepth"], inst["var"])
    ofr = o.frame(1, cam, img)
    ofr.set_depth(fid, fvar)
    oref = o.ref(okf)
    for level in (0, 1):
        E, g, gd = ctx.energy_grad_sim3(ref, fr, level, pose)
        n = oref.num_points(level)
        fz, mf = np.zeros(3 * n), np.zeros(n, np.uint8)
        zero_d = np.zeros(n)
        E0 = frozen_energy_sim3(o, oref, ofr, level, pose, np.zeros(7), zero_d, fz, mf, True)
        assert abs(E - E0) <= 1e-5 * abs(E0)
        fd = central_differences(
            lambda e: frozen_energy_sim3(o, oref, ofr, level, pose, e, zero_d, fz, mf, False), 7)
        assert np.abs(g - fd).max() <= 1e-5 * np.abs(fd).max(), (level, g, fd)
        # every component against its own size, the scale one included (it
        # is carried by the depth residual alone and is ~1e-3 of the largest)
        assert np.all(np.abs(g - fd) <= 5e-5 * np.abs(fd)), (level, np.abs(g - fd) / np.abs(fd))
        dirn = np.random.default_rng(5).normal(size=n)
        gdo = np.empty(n, np.float64)
        gdo[np.argsort(oref.points(level)[:, 7].astype(np.int64), kind="stable")] = _by_pixel(ref, level, gd)
        h = 1e-6
        fdd = (frozen_energy_sim3(o, oref, ofr, level, pose, np.zeros(7), h * dirn, fz, mf, False)
               - frozen_energy_sim3(o, oref, ofr, level, pose, np.zeros(7), -h * dirn, fz, mf, False)) / (2 * h)
        assert abs(fdd - gdo @ dirn) <= 1e-5 * np.abs(gdo * dirn).sum(), (level, fdd, gdo @ dirn)


def test_adjoint_rejects_bad_arguments(ctx):
    inst = synth.tracking_instance(7, cam=synth.camera(160, 120))
    kf, fr, ref = gpu_pair(ctx, inst)
    n = ref.num_points(1)
    g = (lsdg.C.c_double * 7)()
    e = lsdg.C.c_double()
    gd = np.zeros(max(n - 1, 1), np.float32)
    L = ctx.L
    assert L.lsdg_energy_grad(ctx.h, ref.h, fr.h, 1, lsdg.pose7(synth.IDENTITY7), lsdg.C.byref(e), g,
                              gd.ctypes.data, n - 1) == lsdg.EINVAL
    assert L.lsdg_energy_grad(ctx.h, ref.h, fr.h, 5, lsdg.pose7(synth.IDENTITY7), lsdg.C.byref(e), g,
                              None, 0) == lsdg.EINVAL
    # the Sim3 adjoint needs frame depth
    assert L.lsdg_energy_grad_sim3(ctx.h, ref.h, fr.h, 1, lsdg.pose8(list(synth.IDENTITY7) + [1.0]),
                                   lsdg.C.byref(e), g, None, 0) == lsdg.ESTATE
