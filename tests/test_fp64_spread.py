"""How stable the FP64 oracle's own tracking result is (CPU).

The LM's accept / reject and stop decisions and the in-image / isGood sets
are discontinuous in the pose, so a perturbation far below any rounding error
of the GPU path can move the FP64 answer a long way.  On bench instance 54
(seed 1001 + 17 * 54) a 1e-9 change of the start pose moves the FP64 result
by about 1e-3 twist, while on instance 0 it moves it by about 2e-6.  This is
the floor under any finite-precision path's error against the FP64 oracle on
such instances; bench.py reports it per instance as fp64_self_spread."""
import numpy as np

import pyoracle
import synth
from helpers import tracking_params, twist


def _fp64_pose(inst, init):
    o = pyoracle.Oracle(1, params=tracking_params(20))
    kf = o.frame(0, inst["cam"], inst["kf_img"].astype(np.float32))
    kf.set_depth(inst["idepth"], inst["var"])
    fr = o.frame(1, inst["cam"], inst["fr_img"].astype(np.float32))
    pose, _, status = o.se3_track(o.ref(kf), fr, init)
    assert status == 0
    return pose


def _spread(idx):
    inst = synth.tracking_instance(1001 + 17 * idx)
    base = _fp64_pose(inst, list(synth.IDENTITY7))
    eps = 1e-9
    moved = [twist(base, _fp64_pose(inst, list(synth.sample_pose(11000 + k, eps, np.degrees(eps)))))
             for k in range(3)]
    return base, max(moved)


def test_fp64_result_jumps_under_a_1e9_start_perturbation():
    _, s54 = _spread(54)
    _, s0 = _spread(0)
    print(f"FP64 self-spread under 1e-9 start perturbations: instance 54 {s54:.2e}, instance 0 {s0:.2e}")
    assert s54 > 1e-4
    assert s0 < 1e-5
