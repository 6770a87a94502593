"""Shared builders for parity tests: the same seeded instance on the HIP
library (lsdg) and on the oracle (pyoracle), with identical parameters."""
import ctypes as C

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


def frozen_energy(o, ref, fr, lvl, pose, xi, delta, wf, mf, compute):
    """FP64 SE3 adjoint energy with w_p and the in-image set frozen (wf, mf),
    at exp(xi) * pose and point inverse depths d_i + delta_i (oracle prec 1)."""
    e = C.c_double()
    dp = C.POINTER(C.c_double)
    st = o.L.orc_energy_frozen(ref.h, fr.h, C.byref(o.params), lvl, lsdg.pose7(pose),
                               np.ascontiguousarray(xi, np.float64).ctypes.data_as(dp),
                               np.ascontiguousarray(delta, np.float64).ctypes.data_as(dp),
                               wf.ctypes.data_as(dp), mf.ctypes.data, int(compute), C.byref(e))
    assert st == 0, "a frozen point left the image"
    return e.value


def frozen_energy_sim3(o, ref, fr, lvl, pose8, xi, delta, fz, mf, compute):
    """FP64 Sim3 adjoint energy with (wp, wd, frame idepth sample) per point
    and the in-image set frozen (fz, mf), at exp(xi) * pose8 and point inverse
    depths d_i + delta_i (oracle prec 1)."""
    e = C.c_double()
    dp = C.POINTER(C.c_double)
    st = o.L.orc_energy_frozen_sim3(ref.h, fr.h, C.byref(o.params), lvl, lsdg.pose8(pose8),
                                    np.ascontiguousarray(xi, np.float64).ctypes.data_as(dp),
                                    np.ascontiguousarray(delta, np.float64).ctypes.data_as(dp),
                                    fz.ctypes.data_as(dp), mf.ctypes.data, int(compute), C.byref(e))
    assert st == 0, "a frozen point left the image"
    return e.value


def central_differences(energy, dim, h=1e-6):
    """Central differences of energy(xi) at xi = 0 along each axis."""
    out = np.zeros(dim)
    for k in range(dim):
        e = np.zeros(dim)
        e[k] = h
        out[k] = (energy(e) - energy(-e)) / (2 * h)
    return out
