"""Pins of the oracle's keyframe export restatement (publishKeyframe's
records, lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:81-89, and
Keyframe::computeVbo, lib/Pangolin_IOWrapper/Keyframe.h:76-142) against what
the mathematics fixes: a fronto-parallel plane keeps every interior pixel and
back-projects onto z = depth; each filter threshold acts on its side of the
boundary; an isolated pixel lacks near support; the color bytes are
swizzled with alpha 100; raster order."""
import numpy as np
import pytest

import lsdg
import pyoracle


def _records(idepth, var, color=(10, 20, 30, 40)):
    h, w = idepth.shape
    r = np.zeros(w * h, lsdg.POINT_DENSE_DTYPE)
    r["idepth"] = idepth.ravel()
    r["idepth_var"] = var.ravel()
    r["color"] = np.array(color, np.uint8)
    return r


CAM = (200.0, 180.0, 15.5, 11.5)


def _vbo(rec, w, h, **kw):
    o = pyoracle.Oracle(0)
    return o.compute_vbo(rec, w, h, CAM, lsdg.vbo_params(**kw))


def test_plane_keeps_interior_and_back_projects():
    w, h = 32, 24
    rec = _records(np.full((h, w), 0.5, np.float32), np.full((h, w), 1e-5, np.float32))
    v = _vbo(rec, w, h)
    assert len(v) == (w - 2) * (h - 2)
    ys, xs = np.mgrid[1:h - 1, 1:w - 1]
    fx, fy, cx, cy = CAM
    assert np.all(v["point"][:, 2] == 2.0)
    np.testing.assert_allclose(v["point"][:, 0], ((xs - cx) / fx * 2.0).ravel(), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(v["point"][:, 1], ((ys - cy) / fy * 2.0).ravel(), rtol=1e-6, atol=1e-6)
    # swizzle: (c0, c1, c2) -> (c2, c1, c0), alpha 100
    assert np.all(v["color"] == np.array([30, 20, 10, 100], np.uint8))


@pytest.mark.parametrize("var,kept", [(0.9e-3, True), (1.1e-3, False)])
def test_scaled_threshold(var, kept):
    w, h = 5, 5
    rec = _records(np.ones((h, w), np.float32), np.full((h, w), var, np.float32))
    v = _vbo(rec, w, h, min_near_support=1)
    assert (len(v) == 9) == kept and (len(v) == 0) == (not kept)


@pytest.mark.parametrize("scale,kept", [(14.0, True), (15.0, False)])
def test_absolute_threshold_uses_scale_squared(scale, kept):
    w, h = 5, 5
    rec = _records(np.ones((h, w), np.float32), np.full((h, w), 5e-4, np.float32))
    v = _vbo(rec, w, h, min_near_support=1, scale=scale)  # 5e-4 * s^2 > 0.1 <=> s > 14.14
    assert (len(v) == 9) == kept


def test_near_support():
    w, h = 7, 7
    idp = np.zeros((h, w), np.float32)
    idp[3, 3] = 1.0
    var = np.full((h, w), 1e-4, np.float32)
    rec = _records(idp, var)
    assert len(_vbo(rec, w, h)) == 0            # alone: support 1 < 9
    assert len(_vbo(rec, w, h, min_near_support=1)) == 1
    idp[2:5, 2:5] = 1.0
    assert len(_vbo(_records(idp, var), w, h)) == 1     # the centre has all 9
    idp[2, 2] = 1.5                                     # one neighbour off by 0.5 > sqrt(2e-4)
    assert len(_vbo(_records(idp, var), w, h)) == 0


def test_invalid_pixels_and_raster_order():
    w, h = 9, 6
    rng = np.random.default_rng(3)
    idp = rng.uniform(0.5, 1.5, (h, w)).astype(np.float32)
    idp[rng.random((h, w)) < 0.3] = -1.0
    rec = _records(idp, np.full((h, w), 1e-6, np.float32))
    v = _vbo(rec, w, h, min_near_support=1)
    fx, fy, cx, cy = CAM
    xs = np.rint(v["point"][:, 0] / v["point"][:, 2] * fx + cx).astype(int)
    ys = np.rint(v["point"][:, 1] / v["point"][:, 2] * fy + cy).astype(int)
    keep = [(y, x) for y in range(1, h - 1) for x in range(1, w - 1) if idp[y, x] > 0]
    assert list(zip(ys, xs)) == keep


def test_publish_points_color_conversion():
    o = pyoracle.Oracle(0)
    w, h = 8, 8
    img = np.full((h, w), 3.7, np.float32)
    img[0, :4] = [255.9, -1.0, np.nan, 0.5]
    cam = dict(w=w, h=h, fx=10.0, fy=10.0, cx=3.5, cy=3.5)
    f = o.frame(0, cam, img)
    f.set_depth(np.full((h, w), 0.5, np.float32), np.full((h, w), 1e-3, np.float32))
    r = f.keyframe_points(0)
    assert list(r["color"][:4, 0]) == [255, 0, 0, 0] and np.all(r["color"][4:] == 3)
    assert np.all(r["color"][:, 0:1] == r["color"])
    assert np.all(r["idepth"] == np.float32(0.5)) and np.all(r["idepth_var"] == np.float32(1e-3))
