"""Shared builders for parity tests: the same seeded instance on the HIP
library (lsdg) and on the oracle (pyoracle), with identical parameters."""
import numpy as np

import lsdg
import posemath
import pyoracle
import synth


def tracking_params(its=20):
    p = pyoracle.default_params()
    for l in range(5):
        p.max_its_per_lvl[l] = its
    return p


def oracle_pair(o, inst, kf_id=0, fr_id=1):
    kf = o.frame(kf_id, inst["cam"], inst["kf_img"].astype(np.float32))
    kf.set_depth(inst["idepth"], inst["var"])
    fr = o.frame(fr_id, inst["cam"], inst["fr_img"].astype(np.float32))
    return kf, fr, o.ref(kf)


def gpu_pair(ctx, inst, precision=lsdg.FP32, kf_id=0, fr_id=1):
    kf = ctx.frame(kf_id, inst["cam"], inst["kf_img"], precision)
    kf.set_depth(inst["idepth"], inst["var"])
    fr = ctx.frame(fr_id, inst["cam"], inst["fr_img"], precision)
    return kf, fr, ctx.ref(kf)


def twist(a, b):
    return posemath.twist_error(a, b)


def sorted_points(pts):
    return pts[np.argsort(pts[:, 7], kind="stable")]


def rel_close(a, b, rel, absol=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rel * np.abs(b) + absol)


def small_instance(seed=7, w=160, h=120, **kw):
    return synth.tracking_instance(seed, cam=synth.camera(w, h), **kw)
