"""GPU parity of the depth map (DepthMap::updateKeyframe / createKeyFrame /
finalizeKeyFrame / initializeFromGTDepth) against the CPU oracle on the same
seeded 640x480 scene.

Bars: every hypothesis field of every pixel and the per-pixel observe outcome
codes bit-exact against the FP32 restatement (updateKeyframe has no
reduction).  createKeyFrame's mean-idepth rescale is an FP64 sum on the GPU:
bit-exact against the oracle's FP64-accumulation mode, within 1e-4 relative of
the FP32-sum restatement with masks still bit-exact."""
import numpy as np
import pytest

import lsdg
import pyoracle
import synth

pytestmark = pytest.mark.gpu

FIELDS = ["is_valid", "blacklisted", "next_stereo_frame_min_id", "validity_counter", "idepth", "idepth_var",
          "idepth_smoothed", "idepth_var_smoothed", "observe_code"]


def _params(crop):
    p = pyoracle.default_params()
    p.max_epl_length_crop = float(crop)
    return p


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _diff(a, b):
    out = {}
    for k in FIELDS:
        ne = _bits(a[k]) != _bits(b[k])
        if ne.any():
            out[k] = (int(ne.sum()), np.argwhere(ne)[:3].ravel().tolist())
    return out


def _scenario(sys_, d, masks=None, n_refs=1, ids=(1, 3, 4), updates=1):
    """prior on the keyframe -> update(s) with reference frame(s) -> createKeyFrame."""
    cam = d["cam"]
    f32 = (lambda a: a.astype(np.float32)) if isinstance(sys_, pyoracle.Oracle) else (lambda a: a)
    kf = sys_.frame(0, cam, f32(d["kf_img"]))
    refs = []
    ids = list(ids)[:n_refs]
    for k, fid in enumerate(ids):
        img = d["ref_img"] if k == 0 else d["kf2_img"]
        pose = d["ref_pose"] if k == 0 else d["kf2_pose"]
        r = sys_.frame(fid, cam, f32(img))
        r.set_tracking(kf, list(pose) + [1.0], 0.4 + 0.1 * k, None if masks is None else masks[k])
        refs.append(r)
    kf2 = sys_.frame(9, cam, f32(d["kf2_img"]))
    kf2.set_tracking(kf, list(d["kf2_pose"]) + [1.0], 0.3, None if masks is None else masks[-1])
    dm = sys_.depthmap(cam)
    dm.set_prior(kf, d["prior_idepth"], d["prior_var"], 20)
    kf.set_counters(7, 2)
    for u in range(updates):
        dm.update(refs[u % len(refs):] if updates > 1 else refs)
    h1 = dm.read()
    # the keyframe's whole depth pyramid after the update (setDepth and
    # buildIDepthAndIDepthVar, written by the update's last pass)
    kfd = tuple(kf.read(w, l) for l in range(5) for w in (lsdg.WHAT_IDEPTH, lsdg.WHAT_IDEPTH_VAR))
    dm.create_keyframe(kf2)
    h2 = dm.read()
    kf2d = [kf2.read(lsdg.WHAT_IDEPTH, l) for l in range(5)] + [kf2.read(lsdg.WHAT_IDEPTH_VAR, 0)]
    t2 = kf2.get_tracking()
    counters = kf.counters()
    return h1, h2, kfd, kf2d, t2, counters


@pytest.fixture(scope="module")
def depth_inst():
    return synth.depth_instance()


@pytest.mark.parametrize("crop", [100, 30])
def test_update_and_create_keyframe_bit_exact(ctx, depth_inst, crop):
    p = _params(crop)
    ctx.set_params(p)
    try:
        g = _scenario(ctx, depth_inst)
    finally:
        ctx.set_params(lsdg.default_params())
    o0 = _scenario(pyoracle.Oracle(0, params=p), depth_inst)
    o2 = _scenario(pyoracle.Oracle(2, params=p), depth_inst)
    # updateKeyframe: everything bit-exact against the FP32 restatement
    assert not _diff(g[0], o0[0]), _diff(g[0], o0[0])
    for a, b in zip(g[2], o0[2]):
        assert np.array_equal(_bits(a), _bits(b))
    assert g[5] == o0[5]
    codes = np.bincount(g[0]["observe_code"], minlength=15)
    assert codes[lsdg.OBS["UPDATED"]] > 1000 and codes[lsdg.OBS["CREATED"]] > 1000
    # createKeyFrame: bit-exact against the FP64-accumulation restatement
    assert not _diff(g[1], o2[1]), _diff(g[1], o2[1])
    for a, b in zip(g[3], o2[3]):
        assert np.array_equal(_bits(a), _bits(b))
    assert np.allclose(g[4][1], o2[4][1], rtol=0, atol=0)
    # against the FP32-sum restatement: masks exact, values within 1e-4 relative
    for k in ("is_valid", "blacklisted", "validity_counter", "observe_code"):
        assert np.array_equal(g[1][k], o0[1][k]), k
    v = g[1]["is_valid"].astype(bool)
    for k in ("idepth", "idepth_var", "idepth_smoothed", "idepth_var_smoothed"):
        a, b = g[1][k][v].astype(np.float64), o0[1][k][v].astype(np.float64)
        assert np.all(np.abs(a - b) <= 1e-4 * np.abs(b)), k
    assert abs(g[4][1][7] - o0[4][1][7]) <= 1e-4 * abs(o0[4][1][7])


@pytest.mark.parametrize("size", [(200, 136), (328, 250)])
def test_ragged_sizes_bit_exact(ctx, size):
    """Image sizes that are not multiples of the 16x16 tiles (partial tiles
    at the right and bottom edges of every per-pixel and tile pass)."""
    inst = synth.depth_instance(1017, cam=synth.camera(*size))
    p = _params(100)
    ctx.set_params(p)
    try:
        g = _scenario(ctx, inst, n_refs=2, updates=2)
    finally:
        ctx.set_params(lsdg.default_params())
    o2 = _scenario(pyoracle.Oracle(2, params=p), inst, n_refs=2, updates=2)
    o0 = _scenario(pyoracle.Oracle(0, params=p), inst, n_refs=2, updates=2)
    assert not _diff(g[0], o0[0]), _diff(g[0], o0[0])
    # the keyframe's depth pyramid, partial tiles and odd level sizes included
    for a, b in zip(g[2], o0[2]):
        assert np.array_equal(_bits(a), _bits(b))
    assert not _diff(g[1], o2[1]), _diff(g[1], o2[1])
    for a, b in zip(g[3], o2[3]):
        assert np.array_equal(_bits(a), _bits(b))
    codes = np.bincount(g[0]["observe_code"], minlength=15)
    assert codes[lsdg.OBS["UPDATED"]] > 50 and codes[lsdg.OBS["CREATED"]] > 50


@pytest.mark.parametrize("ids", [(1, 3, 4), (1, 40, 41)], ids=["inline", "byid-in-device-memory"])
def test_update_with_tracking_masks_and_several_refs(ctx, depth_inst, ids):
    """refPixelWasGood gates (from a seeded mask), three reference frames with
    id gaps (referenceFrameByID), and the propagation gate on the new keyframe's
    mask.  The second id set spans more frame ids than the kernel arguments
    hold, so the reference table goes through device memory."""
    w1, h1 = 320, 240
    rng = np.random.default_rng(5)
    masks = [(rng.random((h1, w1)) > 0.1).astype(np.uint8) for _ in range(3)]
    p = _params(100)
    ctx.set_params(p)
    try:
        g = _scenario(ctx, depth_inst, masks=masks, n_refs=3, ids=ids)
    finally:
        ctx.set_params(lsdg.default_params())
    o2 = _scenario(pyoracle.Oracle(2, params=p), depth_inst, masks=masks, n_refs=3, ids=ids)
    o0 = _scenario(pyoracle.Oracle(0, params=p), depth_inst, masks=masks, n_refs=3, ids=ids)
    assert not _diff(g[0], o0[0]), _diff(g[0], o0[0])
    assert np.bincount(g[0]["observe_code"], minlength=15)[lsdg.OBS["SKIP_NOT_GOOD_TRACKED"]] > 100
    assert not _diff(g[1], o2[1]), _diff(g[1], o2[1])


def test_consecutive_updates_bit_exact(ctx, depth_inst):
    """Three updates in a row on one depth map (the work-list counters
    alternate between calls), with the reference set shrinking each time."""
    p = _params(100)
    ctx.set_params(p)
    try:
        g = _scenario(ctx, depth_inst, n_refs=3, updates=3)
    finally:
        ctx.set_params(lsdg.default_params())
    o0 = _scenario(pyoracle.Oracle(0, params=p), depth_inst, n_refs=3, updates=3)
    assert not _diff(g[0], o0[0]), _diff(g[0], o0[0])
    codes = np.bincount(g[0]["observe_code"], minlength=15)
    assert codes[lsdg.OBS["UPDATED"]] > 1000


def test_update_posterized_images_bit_exact(ctx, depth_inst):
    """Images quantised to 9 grey levels: many exactly equal matching errors
    along a line, so the first-minimum rules for the best and second match
    decide the outcome."""
    d = dict(depth_inst)
    for k in ("kf_img", "ref_img", "kf2_img"):
        a = d[k].astype(np.float32)
        d[k] = np.clip(np.round(a / 32.0) * 32.0, 0, 255).astype(d[k].dtype)
    p = _params(100)
    ctx.set_params(p)
    try:
        g = _scenario(ctx, d)
    finally:
        ctx.set_params(lsdg.default_params())
    o0 = _scenario(pyoracle.Oracle(0, params=p), d)
    assert not _diff(g[0], o0[0]), _diff(g[0], o0[0])
    codes = np.bincount(g[0]["observe_code"], minlength=15)
    assert codes[lsdg.OBS["CREATED"]] > 100 and codes[lsdg.OBS["STEREO_FAIL_BASE"] + 2] > 100


def test_initialize_from_gt_and_finalize(ctx, depth_inst):
    d = depth_inst
    cam = d["cam"]
    out = []
    for sys_ in (ctx, pyoracle.Oracle(0)):
        f32 = (lambda a: a.astype(np.float32)) if sys_ is not ctx else (lambda a: a)
        kf = sys_.frame(0, cam, f32(d["kf_img"]))
        gt = np.where(d["prior_idepth"] > 0, (1.0 / d["kf_depth"]).astype(np.float32), np.float32(-1))
        kf.set_depth(gt, np.where(gt > 0, np.float32(1e-3), np.float32(-1)))
        dm = sys_.depthmap(cam)
        dm.init_gt(kf)
        a = dm.read()
        dm.finalize()
        out.append((a, dm.read(), kf.read(lsdg.WHAT_IDEPTH, 1)))
    assert not _diff(out[0][0], out[1][0])
    assert not _diff(out[0][1], out[1][1])
    assert np.array_equal(_bits(out[0][2]), _bits(out[1][2]))


@pytest.mark.parametrize("size", [(640, 480), (200, 136)])
def test_initialize_randomly_then_update_bit_exact(ctx, size):
    """DepthMap::initializeRandomly (seeded draws) and an update on top of it,
    every field of every pixel against the FP32 restatement."""
    inst = synth.depth_instance(1023, cam=synth.camera(*size))
    p = _params(100)

    def run(sys_):
        cam = inst["cam"]
        f32 = (lambda a: a.astype(np.float32)) if isinstance(sys_, pyoracle.Oracle) else (lambda a: a)
        kf = sys_.frame(0, cam, f32(inst["kf_img"]))
        r = sys_.frame(1, cam, f32(inst["ref_img"]))
        r.set_tracking(kf, list(inst["ref_pose"]) + [1.0], 0.4)
        dm = sys_.depthmap(cam)
        dm.init_random(kf, 99)
        h0 = dm.read()
        d0 = kf.read(lsdg.WHAT_IDEPTH, 0)
        dm.update([r])
        return h0, d0, dm.read()

    ctx.set_params(p)
    try:
        g = run(ctx)
    finally:
        ctx.set_params(lsdg.default_params())
    o = run(pyoracle.Oracle(0, params=p))
    assert not _diff(g[0], o[0]), _diff(g[0], o[0])
    assert np.array_equal(_bits(g[1]), _bits(o[1]))
    assert not _diff(g[2], o[2]), _diff(g[2], o[2])
    assert g[0]["is_valid"].mean() > 0.05
