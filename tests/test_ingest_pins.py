"""The oracle's restatement of the ingest step (libvideoio's
Undistorter::undistort as lib/App/InputThread.cpp:59-62 applies it, the FOV
lens model's table) and of TextOutputIOWrapper's pose.txt line, pinned on
closed forms; plus the product's host-side table builder against it.

libvideoio is not in the tree (no golden vectors exist for it), so every pin
here is mathematical: integer source positions copy, half-pixel positions on
even ramps average exactly, truncation stays within one grey level below the
exact bilinear value, omega = 0 is a pinhole rescale, and the table inverts
the FOV model's radial law."""
import numpy as np
import pytest

import lsdg
import pyoracle


@pytest.fixture(scope="module")
def o():
    return pyoracle.Oracle(0)


def _identity_table(w, h, dx=0.0, dy=0.0):
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    t = np.stack([x + np.float32(dx), y + np.float32(dy)], -1)
    bad = (t[..., 0] >= w - 1) | (t[..., 1] >= h - 1)
    t[bad] = -1.0
    return t


def test_integer_positions_copy_the_source(o):
    rng = np.random.default_rng(1)
    src = rng.integers(0, 256, (40, 56), dtype=np.uint8)
    out = o.undistort(src, _identity_table(56, 40))
    assert np.array_equal(out[:-1, :-1], src[:-1, :-1])
    # the last row and column have no second tap: no source, output 0
    assert not out[-1].any() and not out[:, -1].any()


def test_no_source_gives_zero(o):
    src = np.full((20, 20), 200, np.uint8)
    t = _identity_table(20, 20, 0.25, 0.25)
    t[5:9, 3:11] = -1.0
    out = o.undistort(src, t)
    assert not out[5:9, 3:11].any()


def test_half_pixel_positions_average_even_ramps(o):
    # ramps of even values: the midpoint of two taps is an exact integer
    w, h = 64, 16
    src = np.tile((4 * np.arange(w) % 250).astype(np.uint8), (h, 1))
    out = o.undistort(src, _identity_table(w, h, 0.5, 0.0))
    exp = (src[:, :-1].astype(int) + src[:, 1:].astype(int)) // 2
    assert np.array_equal(out[:h - 1, :w - 2], exp[:h - 1, :w - 2])
    src_v = np.tile((6 * np.arange(h)).astype(np.uint8)[:, None], (1, w))
    out_v = o.undistort(src_v, _identity_table(w, h, 0.0, 0.5))
    exp_v = (src_v[:-1].astype(int) + src_v[1:].astype(int)) // 2
    assert np.array_equal(out_v[:h - 2, :w - 1], exp_v[:h - 2, :w - 1])


def test_truncation_is_within_one_grey_level_of_exact_bilinear(o):
    rng = np.random.default_rng(2)
    w, h = 48, 36
    src = rng.integers(0, 256, (h, w), dtype=np.uint8)
    tab = np.stack([rng.uniform(0, w - 1.001, (h, w)), rng.uniform(0, h - 1.001, (h, w))], -1).astype(np.float32)
    out = o.undistort(src, tab).astype(np.float64)
    x, y = tab[..., 0].astype(np.float64), tab[..., 1].astype(np.float64)
    xi, yi = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - xi, y - yi
    s = src.astype(np.float64)
    exact = ((1 - fx) * (1 - fy) * s[yi, xi] + fx * (1 - fy) * s[yi, xi + 1] + (1 - fx) * fy * s[yi + 1, xi] +
             fx * fy * s[yi + 1, xi + 1])
    d = exact - out
    # truncation: never above the exact value by more than FP32 rounding, and
    # less than one grey level below it (a swapped tap or weight fails this)
    assert d.min() > -1e-3 and d.max() < 1.0 + 1e-3


def test_fov_omega_zero_is_a_pinhole_rescale(o):
    ik, ok = [300.0, 280.0, 160.0, 120.0], [250.0, 240.0, 150.5, 110.5]
    t = o.fov_table(ik, 0.0, (320, 240), ok, (300, 220))
    y, x = np.mgrid[0:220, 0:300].astype(np.float64)
    ex = 300.0 * ((x - 150.5) / 250.0) + 160.0
    ey = 280.0 * ((y - 110.5) / 240.0) + 120.0
    inside = (ex > 0) & (ey > 0) & (ex < 319) & (ey < 239)
    assert np.allclose(t[..., 0][inside], ex[inside], atol=1e-4) and np.allclose(t[..., 1][inside], ey[inside], atol=1e-4)
    assert np.all(t[~inside] == -1.0)
    # identical cameras: the identity map
    t2 = o.fov_table(ik, 0.0, (320, 240), ik, (320, 240))
    y2, x2 = np.mgrid[0:240, 0:320].astype(np.float32)
    m = t2[..., 0] >= 0
    assert np.array_equal(t2[..., 0][m], x2[m]) and np.array_equal(t2[..., 1][m], y2[m])


def test_fov_table_inverts_the_radial_law(o):
    omega = 0.93
    ik, ok = [260.0, 260.0, 319.5, 239.5], [200.0, 200.0, 319.5, 239.5]
    t = o.fov_table(ik, omega, (640, 480), ok, (640, 480)).astype(np.float64)
    y, x = np.mgrid[0:480, 0:640].astype(np.float64)
    ux, uy = (x - 319.5) / 200.0, (y - 239.5) / 200.0
    m = t[..., 0] >= 0
    dx, dy = (t[..., 0] - 319.5) / 260.0, (t[..., 1] - 239.5) / 260.0
    ru, rd = np.hypot(ux, uy)[m], np.hypot(dx, dy)[m]
    # undistorted radius from the distorted one: tan(omega rd) / (2 tan(omega / 2))
    back = np.tan(omega * rd) / (2 * np.tan(omega / 2))
    assert np.max(np.abs(back - ru) / np.maximum(ru, 1e-3)) < 1e-4
    # the direction is kept, and the distortion pulls points inwards
    c = (ux[m] * dy[m] - uy[m] * dx[m]) / np.maximum(ru * rd, 1e-9)
    assert np.max(np.abs(c)) < 1e-4
    # the distorted radius grows with the undistorted one (along the centre row)
    row = t[240, :, 0]
    k = row >= 0
    assert np.all(np.diff(row[k]) > 0)
    assert 0.5 < m.mean() < 1.0


def test_product_table_matches_the_oracle(o):
    for omega in (0.0, 0.5, 0.93):
        args = ([300.0, 300.0, 376.0, 240.0], omega, (752, 480), [260.0, 258.0, 319.5, 239.5], (640, 480))
        a = lsdg.fov_undistort_table(*args)
        b = o.fov_table(*args)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), omega


@pytest.mark.parametrize("vals", [
    ((1.5, -2e-7, 123456789.0), (0.1, 0.0, 1.0 / 3.0)),
    ((0.0, -0.0, 1e-300), (1e21, -12.5, 7.0)),
    ((2.0 / 3.0, -1.0 / 7.0, 100000.0), (999999.5, 1234567.0, -0.000123456789)),
])
def test_pose_line_formats_like_an_ostream(o, vals):
    a, b = vals
    line = o.pose_txt_line(42, a, b)
    # an ostream prints a double with %g at precision 6
    exp = "42," + ",".join("%g" % v for v in (*a, *b)) + "\n"
    assert line == exp
    assert line.count(",") == 6 and line.endswith("\n")


def test_product_pose_line_matches_the_oracle(o):
    rng = np.random.default_rng(8)
    for k in range(20):
        a = rng.normal(size=3) * 10.0 ** rng.integers(-8, 8, 3)
        b = rng.normal(size=3) * 10.0 ** rng.integers(-8, 8, 3)
        line = lsdg.pose_txt_line(k, [0, 0, 0, 1, *a, 1.0], [0, 0, 0, 1, *b, 1.0])
        assert line == o.pose_txt_line(k, a, b)
