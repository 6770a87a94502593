"""The lsd_slam shim classes (SE3Tracker::trackFrame, DepthMap::updateKeyframe
on top of the C ABI) give the same results as the C ABI called directly: the
shim example program runs in a child process on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import lsdg
import synth
from conftest import PKG

pytestmark = pytest.mark.gpu


def test_shim_track_and_depth_update_match_the_c_abi(ctx, tmp_path):
    exe = os.path.join(PKG, "shim_track")
    if not os.path.exists(exe):
        pytest.fail("shim_track is not built; run __graft_entry__.build()")
    inst = synth.tracking_instance(1011, cam=synth.camera(320, 240))
    cam = inst["cam"]
    inst["kf_img"].astype(np.uint8).tofile(tmp_path / "kf.u8")
    inst["fr_img"].astype(np.uint8).tofile(tmp_path / "fr.u8")
    inst["idepth"].astype(np.float32).tofile(tmp_path / "idepth.f32")
    inst["var"].astype(np.float32).tofile(tmp_path / "var.f32")
    args = [exe, str(tmp_path)] + [repr(float(cam[k])) if k not in ("w", "h") else str(cam[k])
                                   for k in ("w", "h", "fx", "fy", "cx", "cy")]
    print("shim_track", cam, flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = {l.split()[0]: l.split()[1:] for l in out.stdout.strip().splitlines()}
    pose = np.array([float(v) for v in lines["pose"]])

    # the same calls through the C ABI in this process (default parameters)
    kf = ctx.frame(0, cam, inst["kf_img"])
    kf.set_depth(inst["idepth"], inst["var"])
    fr = ctx.frame(1, cam, inst["fr_img"])
    ref = ctx.ref(kf)
    p, st, status = ctx.se3_track(ref, fr, synth.IDENTITY7)
    assert status == lsdg.OK
    assert np.array_equal(pose, p)
    kv = dict(zip(lines["residual"][1::2], lines["residual"][2::2]))
    assert int(kv["good"]) == st.last_good_count and int(kv["bad"]) == st.last_bad_count
    assert int(kv["diverged"]) == 0 and int(kv["good_tracking"]) == st.tracking_was_good

    dm = ctx.depthmap(cam)
    dm.init_gt(kf)
    dm.update([fr])
    n_valid = int(dm.read()["is_valid"].sum())
    assert int(lines["depth_valid"][0]) == n_valid
