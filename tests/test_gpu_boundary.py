"""C-ABI argument checks that need a device context: odd image sizes are
refused (LSDG_EINVAL) by every entry point that creates level-0 storage, since
propagateDepth indexes the level-1 tracking mask with (x >> 1, y >> 1) of
level-0 pixels; even sizes that are not multiples of 16 are accepted and run
bit-exact against the oracle (test_gpu_depthmap.py::test_ragged_sizes_bit_exact)."""
import ctypes as C

import numpy as np
import pytest

import lsdg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h", [(201, 138), (202, 137), (641, 481)])
def test_odd_sizes_are_refused(ctx, w, h):
    L = ctx.L
    cam = lsdg.Camera(w, h, 200.0, 200.0, w / 2.0, h / 2.0)
    out = C.c_void_p()
    img_f = np.zeros((h, w), np.float32)
    img_u8 = np.zeros((h, w), np.uint8)
    assert L.lsdg_frame_create(ctx.h, 0, C.byref(cam), img_f.ctypes.data, lsdg.FP32, C.byref(out)) == lsdg.EINVAL
    assert L.lsdg_frame_create_u8(ctx.h, 0, C.byref(cam), img_u8.ctypes.data, w, lsdg.FP16,
                                  C.byref(out)) == lsdg.EINVAL
    assert L.lsdg_depthmap_create(ctx.h, C.byref(cam), C.byref(out)) == lsdg.EINVAL
    table = np.full((h, w, 2), -1.0, np.float32)
    assert L.lsdg_undistorter_create(ctx.h, 64, 64, w, h, table.ctypes.data, C.byref(out)) == lsdg.EINVAL
    assert not out.value
    # the even neighbour is accepted
    ew, eh = w & ~1, h & ~1
    f = ctx.frame(0, dict(w=ew, h=eh, fx=200.0, fy=200.0, cx=ew / 2.0, cy=eh / 2.0), np.zeros((eh, ew), np.uint8))
    f.destroy()
