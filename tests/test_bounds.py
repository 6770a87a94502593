"""Image sizes at the boundary: the level-1 tracking mask that
DepthMap::propagateDepth reads is indexed with (x >> 1, y >> 1) of level-0
pixels, so an odd width or height would read one column or row past it.  The
C-ABI and the oracle both refuse odd sizes; the oracle's depth pipeline runs
under AddressSanitizer at even sizes that are not multiples of 16, with depth
on the last row and column and a keyframe move that carries those pixels past
the propagation gate (the case that overruns the mask at odd sizes).

CPU only."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOUNDS = os.path.join(ROOT, "oracle", "bounds_check")


def test_oracle_depth_pipeline_in_bounds_under_asan():
    if not os.path.exists(BOUNDS):
        pytest.fail(f"{BOUNDS} is not built; run `make -C oracle`")
    r = subprocess.run([BOUNDS, "202", "138", "330", "250", "66", "40", "642", "482"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.split() == ["ok", "202x138", "ok", "330x250", "ok", "66x40", "ok", "642x482"]


@pytest.mark.parametrize("w,h", [(201, 138), (202, 137), (641, 481)])
def test_oracle_refuses_odd_sizes(w, h):
    o = pyoracle.Oracle(0)
    cam = dict(w=w, h=h, fx=200.0, fy=200.0, cx=w / 2.0, cy=h / 2.0)
    with pytest.raises(ValueError):
        o.frame(0, cam, np.zeros((h, w), np.float32))
    with pytest.raises(ValueError):
        o.depthmap(cam)
