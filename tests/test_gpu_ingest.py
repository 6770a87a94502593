"""GPU parity of the ingest step (undistortion + uint8 -> float + image
pyramid in one launch, lib/App/InputThread.cpp:59-71) and of the pose.txt
line (TextOutputIOWrapper.cpp:100-120) against the oracle.

Bars: the 8-bit undistorted image and every pyramid level bit-exact against
oracle/ingest.hpp's undistort followed by the oracle's Frame::buildImage;
the gradients the tracker builds later from the ingested frame bit-exact too;
the pose line byte-identical to the oracle's ostream output."""
import numpy as np
import pytest

import lsdg
import pyoracle
import synth

pytestmark = pytest.mark.gpu


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _distorted(w, h, seed):
    sc = synth.Scene(seed)
    img, _ = sc.render(synth.IDENTITY7, synth.camera(w, h), 2.0, seed)
    return img


CASES = {
    # in size, in camera (fx, fy, cx, cy), omega, out camera, out size
    "vga-fov": ((640, 480), (330.0, 330.0, 319.5, 239.5), 0.92, (270.0, 270.0, 319.5, 239.5), (640, 480)),
    "wvga-to-vga": ((752, 480), (360.0, 358.0, 375.5, 241.0), 0.85, (300.0, 300.0, 319.5, 239.5), (640, 480)),
    "pinhole-crop": ((400, 300), (300.0, 300.0, 199.5, 149.5), 0.0, (300.0, 300.0, 159.5, 119.5), (320, 240)),
    # output tiles cut at the right and bottom edges
    "ragged": ((230, 171), (150.0, 150.0, 114.5, 85.0), 0.7, (130.0, 130.0, 104.5, 74.5), (210, 150)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_ingest_bit_exact(ctx, case):
    in_size, ik, omega, ok, out_size = CASES[case]
    src = _distorted(*in_size, seed=len(case))
    table = lsdg.fov_undistort_table(ik, omega, in_size, ok, out_size)
    assert (table[..., 0] >= 0).mean() > 0.5
    cam = dict(w=out_size[0], h=out_size[1], fx=ok[0], fy=ok[1], cx=ok[2], cy=ok[3])
    und = ctx.undistorter(in_size, table)
    f, u8 = ctx.frame_undistorted(3, cam, und, src, want_undistorted=True)
    o = pyoracle.Oracle(0)
    ou8 = o.undistort(src, table)
    assert np.array_equal(u8, ou8), np.argwhere(u8 != ou8)[:5]
    of = o.frame(3, cam, ou8.astype(np.float32))
    for l in range(5):
        a, b = f.read(lsdg.WHAT_IMAGE, l), of.read(lsdg.WHAT_IMAGE, l)
        assert np.array_equal(_u32(a), _u32(b)), (l, np.argwhere(_u32(a) != _u32(b))[:5])
    # what the tracker builds on first use from the ingested levels
    for l in (0, 2, 4):
        for what in (lsdg.WHAT_GRADIENTS, lsdg.WHAT_MAXGRAD):
            a, b = f.read(what, l), of.read(what, l)
            assert np.array_equal(_u32(a), _u32(b)), (what, l)
    f.destroy()
    und.close()


def test_ingest_without_copy_out_and_padded_rows(ctx):
    in_size, ik, omega, ok, out_size = CASES["vga-fov"]
    src = _distorted(*in_size, seed=77)
    padded = np.zeros((in_size[1], in_size[0] + 24), np.uint8)
    padded[:, :in_size[0]] = src
    table = lsdg.fov_undistort_table(ik, omega, in_size, ok, out_size)
    cam = dict(w=out_size[0], h=out_size[1], fx=ok[0], fy=ok[1], cx=ok[2], cy=ok[3])
    und = ctx.undistorter(in_size, table)
    # the binding passes the row stride of the array it is given
    h = lsdg.C.c_void_p()
    cc = lsdg.Camera.of(cam)
    ctx.check(ctx.L.lsdg_frame_create_undistorted(ctx.h, 5, lsdg.C.byref(cc), und.h, padded.ctypes.data,
                                                  padded.shape[1], lsdg.FP32, None, lsdg.C.byref(h)),
              "lsdg_frame_create_undistorted")
    f = lsdg.Frame(ctx, h, 5, cam)
    ref = pyoracle.Oracle(0).undistort(src, table).astype(np.float32)
    assert np.array_equal(_u32(f.read(lsdg.WHAT_IMAGE, 0)), _u32(ref))
    # the same undistorter again, now with the copy-out
    f2, u8 = ctx.frame_undistorted(6, cam, und, src, want_undistorted=True)
    assert np.array_equal(u8.astype(np.float32), ref)
    f.destroy()
    f2.destroy()
    und.close()


def test_reused_frame_is_a_fresh_frame(ctx):
    in_size, ik, omega, ok, out_size = CASES["vga-fov"]
    table = lsdg.fov_undistort_table(ik, omega, in_size, ok, out_size)
    cam = dict(w=out_size[0], h=out_size[1], fx=ok[0], fy=ok[1], cx=ok[2], cy=ok[3])
    und = ctx.undistorter(in_size, table)
    a, b = _distorted(*in_size, seed=31), _distorted(*in_size, seed=32)
    kf = ctx.frame_undistorted(0, cam, und, a)
    f = ctx.frame_undistorted(1, cam, und, a)
    f.read(lsdg.WHAT_GRADIENTS, 1)
    f.set_depth(np.full((cam["h"], cam["w"]), 0.5, np.float32), np.full((cam["h"], cam["w"]), 0.01, np.float32))
    f.set_tracking(kf, [0.0, 0.0, 0.0, 1.0, 0.1, 0.2, 0.3, 1.0], 0.7)
    f.set_counters(4, 2)
    r = ctx.ref(f)
    n_before = r.num_points(2)
    u8 = f.ingest(9, und, b, want_undistorted=True)
    fresh = ctx.frame_undistorted(9, cam, und, b)
    assert np.array_equal(u8, pyoracle.Oracle(0).undistort(b, table))
    for l in range(5):
        for what in (lsdg.WHAT_IMAGE, lsdg.WHAT_GRADIENTS, lsdg.WHAT_MAXGRAD):
            assert np.array_equal(_u32(f.read(what, l)), _u32(fresh.read(what, l))), (what, l)
    pid, tp, itr = f.get_tracking()
    assert pid == -1 and np.allclose(tp, [0, 0, 0, 1, 0, 0, 0, 1]) and itr == 0.0
    assert f.counters() == (0, 0)
    assert f.pose_txt() == "9,0,0,0,0,0,0\n"
    # no depth any more: a tracking reference on it has no points to take
    with pytest.raises(lsdg.LsdgError):
        r.num_points(2)
    assert n_before > 0
    for x in (r, f, fresh, kf):
        x.destroy()
    und.close()


def test_identity_table_matches_the_u8_frame(ctx):
    cam = synth.camera(320, 240)
    src = _distorted(320, 240, seed=5)
    y, x = np.mgrid[0:240, 0:320].astype(np.float32)
    table = np.stack([x, y], -1)
    table[(x >= 319) | (y >= 239)] = -1.0
    und = ctx.undistorter((320, 240), table)
    f = ctx.frame_undistorted(1, cam, und, src)
    expect = src.copy()
    expect[-1, :] = 0
    expect[:, -1] = 0
    g = ctx.frame(1, cam, expect)
    for l in range(5):
        assert np.array_equal(_u32(f.read(lsdg.WHAT_IMAGE, l)), _u32(g.read(lsdg.WHAT_IMAGE, l))), l
    f.destroy()
    g.destroy()
    und.close()


def test_ingest_rejects_bad_tables_and_sizes(ctx):
    table = lsdg.fov_undistort_table((300.0, 300.0, 159.5, 119.5), 0.5, (320, 240), (250.0, 250.0, 159.5, 119.5),
                                     (320, 240))
    for bad in ((5, 7, 319.0, 10.0), (9, 9, 10.0, 239.5), (3, 4, np.nan, 1.0), (0, 0, 2.0, -0.5)):
        t = table.copy()
        t[bad[0], bad[1]] = bad[2:]
        with pytest.raises(lsdg.LsdgError):
            ctx.undistorter((320, 240), t)
    und = ctx.undistorter((320, 240), table)
    with pytest.raises(lsdg.LsdgError):
        ctx.frame_undistorted(0, synth.camera(640, 480), und, np.zeros((240, 320), np.uint8))
    und.close()


def test_pose_txt_line(ctx):
    cam = synth.camera(320, 240)
    img = _distorted(320, 240, seed=9)
    kf = ctx.frame(0, cam, img)
    fr = ctx.frame(17, cam, img)
    c2w = [0.0, 0.0, 0.0, 1.0, 1.25, -0.5e-7, 3.0 / 7.0, 1.0]
    t2p = [0.1, 0.0, 0.0, np.sqrt(0.99), 123456.789, 0.0, -2.5, 1.0]
    fr.set_cam_to_world(c2w)
    fr.set_tracking(kf, t2p, 0.3)
    line = fr.pose_txt()
    assert line == pyoracle.Oracle(0).pose_txt_line(17, c2w[4:7], t2p[4:7])
    assert line == "17,1.25,-5e-08,0.428571,123457,0,-2.5\n"
    fr.destroy()
    kf.destroy()
