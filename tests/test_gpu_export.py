"""GPU parity of the keyframe export (publishKeyframe's InputPointDense
records and Keyframe::computeVbo's filter, lib/Pangolin_IOWrapper/) against
the oracle's restatement: records and vertices bit-exact, vertex order
included."""
import numpy as np
import pytest

import lsdg
import pyoracle
import synth

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.fixture(scope="module")
def keyframes(ctx):
    """The depth workload's keyframe with its true inverse depth on the prior's
    30% of pixels (small variances, so that the filter keeps a good part)."""
    d = synth.depth_instance()
    cam = d["cam"]
    rng = np.random.default_rng(11)
    idp = np.where(d["prior_idepth"] > 0, 1.0 / d["kf_depth"], -1).astype(np.float32)
    var = np.where(idp > 0, rng.uniform(1e-7, 3e-6, idp.shape), -1).astype(np.float32)
    g = ctx.frame(0, cam, d["kf_img"])
    g.set_depth(idp, var)
    o = pyoracle.Oracle(0)
    of = o.frame(0, cam, d["kf_img"].astype(np.float32))
    of.set_depth(idp, var)
    return g, of, o


@pytest.mark.parametrize("level", [0, 1, 3])
def test_keyframe_points_bit_exact(keyframes, level):
    g, of, _ = keyframes
    a, b = g.keyframe_points(level), of.keyframe_points(level)
    assert np.array_equal(_bits(a), _bits(b))
    assert (a["idepth"] > 0).sum() > 100


@pytest.mark.parametrize("level,scale", [(0, 1.0), (0, 30.0), (1, 1.0), (2, 0.5)])
def test_keyframe_vbo_bit_exact(keyframes, level, scale):
    g, of, o = keyframes
    p = lsdg.vbo_params(scale)
    got = g.keyframe_vbo(level, p)
    w, h = of.size(level)
    want = o.compute_vbo(of.keyframe_points(level), w, h, of.camera(level), p)
    assert len(got) == len(want) and len(want) > 50
    assert np.array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("near", [1, 9])
def test_points_vbo_edge_records(ctx, near):
    """Host records with invalid, NaN, zero-variance and threshold-boundary
    entries, odd sizes."""
    w, h = 37, 29
    rng = np.random.default_rng(5 + near)
    rec = np.zeros(w * h, lsdg.POINT_DENSE_DTYPE)
    ys, xs = np.mgrid[0:h, 0:w]
    idp = (0.8 + 0.002 * xs + 0.001 * ys).astype(np.float32).ravel()  # a slanted plane
    idp[rng.random(w * h) < 0.04] = -1.0
    idp[rng.random(w * h) < 0.02] = 0.0
    idp[rng.random(w * h) < 0.01] = np.nan
    var = rng.choice(np.array([0.0, 1e-6, 1e-5, 1e-4, 2.4e-4, 5e-4, 2e-3], np.float32), w * h)
    rec["idepth"], rec["idepth_var"] = idp, var
    rec["color"] = rng.integers(0, 256, (w * h, 4), dtype=np.uint8)
    cam4 = (300.0, 280.0, 18.0, 14.0)
    p = lsdg.vbo_params(1.7, min_near_support=near)
    got = ctx.points_vbo(rec, w, h, cam4, p)
    want = pyoracle.Oracle(0).compute_vbo(rec, w, h, cam4, p)
    assert len(got) == len(want) and len(want) > 0
    assert np.array_equal(_bits(got), _bits(want))


def test_no_depth_is_an_error(ctx):
    cam = synth.camera(128, 96)
    f = ctx.frame(0, cam, np.zeros((96, 128), np.uint8))
    with pytest.raises(RuntimeError):
        f.keyframe_points(0)
