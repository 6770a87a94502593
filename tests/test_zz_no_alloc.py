"""lsdg_energy_grad_noalloc, the adjoint entry point that allocates nothing
(its scratch is a fixed device array), against lsdg_energy_grad through the
C-ABI: two calls in a row, the second on a smaller input, each bit for bit
equal to the allocating entry point; its preconditions (point cloud and
pyramid level already built, at most LSDG_NOALLOC_MAX_POINTS points) are
reported, not worked around."""
import ctypes as C

import numpy as np
import pytest

import lsdg
import pyoracle
import synth
from helpers import gpu_pair

pytestmark = pytest.mark.gpu


def _call(ctx, fn, ref, fr, level, pose, n):
    e = C.c_double()
    g = (C.c_double * 6)()
    gd = np.zeros(max(n, 1), np.float32)
    st = fn(ctx.h, ref.h, fr.h, level, lsdg.pose7(pose), C.byref(e), g, gd.ctypes.data, n)
    return st, e.value, np.array(g[:]), gd[:n]


def test_noalloc_adjoint_equals_energy_grad(ctx):
    big = synth.tracking_instance(1004, cam=synth.camera(1280, 960), trans_m=0.03, rot_deg=2.0, threshold=None,
                                  fraction=0.25)
    small = synth.tracking_instance(7, cam=synth.camera(320, 240))
    sizes = []
    for inst, level in ((big, 0), (small, 2)):
        kf, fr, ref = gpu_pair(ctx, inst)
        n = ref.num_points(level)
        sizes.append(n)
        fr.build_pyramid()
        pose = pyoracle.sim3_mul(pyoracle.sim3_exp(np.r_[0.004, -0.003, 0.002, 0.003, 0.002, -0.001, 0.0]),
                                 pyoracle.sim3_inv(np.r_[inst["gt"], 1.0]))[:7]
        st, E, g, gd = _call(ctx, ctx.L.lsdg_energy_grad_noalloc, ref, fr, level, pose, n)
        assert st == lsdg.OK
        st2, E2, g2, gd2 = _call(ctx, ctx.L.lsdg_energy_grad, ref, fr, level, pose, n)
        assert st2 == lsdg.OK
        assert E == E2 and np.array_equal(g, g2) and np.array_equal(gd, gd2), (E, E2, g, g2)
        assert np.abs(g).max() > 0 and np.count_nonzero(gd) > 0.5 * n
        # results only (no per-point output)
        e = C.c_double()
        gx = (C.c_double * 6)()
        assert ctx.L.lsdg_energy_grad_noalloc(ctx.h, ref.h, fr.h, level, lsdg.pose7(pose), C.byref(e), gx, None,
                                              0) == lsdg.OK
        assert e.value == E and np.array_equal(np.array(gx[:]), g)
    assert 250000 < sizes[0] and sizes[1] < sizes[0] // 10


def test_noalloc_adjoint_preconditions(ctx):
    inst = synth.tracking_instance(8, cam=synth.camera(320, 240))
    kf, fr, ref = gpu_pair(ctx, inst)
    # neither the point cloud of level 1 nor the frame's gradients exist yet
    st, *_ = _call(ctx, ctx.L.lsdg_energy_grad_noalloc, ref, fr, 1, synth.IDENTITY7, 0)
    assert st == lsdg.ESTATE
    n = ref.num_points(1)
    st, *_ = _call(ctx, ctx.L.lsdg_energy_grad_noalloc, ref, fr, 1, synth.IDENTITY7, n)
    assert st == lsdg.ESTATE  # the frame level is still missing
    fr.build_pyramid()
    st, *_ = _call(ctx, ctx.L.lsdg_energy_grad_noalloc, ref, fr, 1, synth.IDENTITY7, n - 1)
    assert st == lsdg.EINVAL  # output buffer too small
    st, *_ = _call(ctx, ctx.L.lsdg_energy_grad_noalloc, ref, fr, 1, synth.IDENTITY7, n)
    assert st == lsdg.OK
    # more points than the fixed array holds
    huge = synth.tracking_instance(9, cam=synth.camera_1080(), threshold=None, fraction=0.6)
    hk, hf, hr = gpu_pair(ctx, huge)
    m = hr.num_points(0)
    assert m > lsdg.NOALLOC_MAX_POINTS
    hf.build_pyramid()
    st, *_ = _call(ctx, ctx.L.lsdg_energy_grad_noalloc, hr, hf, 0, synth.IDENTITY7, m)
    assert st == lsdg.EINVAL
