"""ctypes mirror of include/lsdg.h and a thin Python façade over liblsdg.so.

Structure layouts here must match the C header field for field (a CPU test
checks sizes against the library).  The façade raises if the HIP library is
missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PYR = 5

OK, DIVERGED, EINVAL, EDEVICE, ENOMEM, ESTATE = 0, 1, -1, -2, -3, -4
NOALLOC_MAX_POINTS = 1 << 20  # LSDG_NOALLOC_MAX_POINTS
FP32, FP16, BF16 = 0, 1, 2
SUMS_FP64, SUMS_KAHAN = 0, 1
WHAT_IMAGE, WHAT_GRADIENTS, WHAT_MAXGRAD, WHAT_IDEPTH, WHAT_IDEPTH_VAR, WHAT_REF_PIXEL_WAS_GOOD = range(6)
JOB_SE3, JOB_SIM3 = 0, 1

OBS = dict(NONE=0, LOWGRAD_INVALIDATE=1, SKIP_ALREADY_GOOD=2, SKIP_NOT_GOOD_TRACKED=3, EPL_FAIL=4,
           CREATE_VAR_TOO_BIG=5, CREATED=6, INCONSISTENT=7, UPDATED=8, STEREO_FAIL_BASE=10)


class Camera(C.Structure):
    _fields_ = [("w", C.c_int32), ("h", C.c_int32), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float),
                ("cy", C.c_float)]

    @staticmethod
    def of(cam):
        return Camera(cam["w"], cam["h"], cam["fx"], cam["fy"], cam["cx"], cam["cy"])


_PARAM_FIELDS = [
    ("lambda_initial", C.c_float * PYR), ("lambda_success_fac", C.c_float * PYR),
    ("lambda_fail_fac", C.c_float * PYR), ("step_size_min", C.c_float * PYR),
    ("convergence_eps", C.c_float * PYR), ("max_its_per_lvl", C.c_int32 * PYR),
    ("var_weight", C.c_float), ("huber_d", C.c_float), ("camera_pixel_noise2", C.c_float),
    ("se3_min_level", C.c_int32), ("se3_max_level", C.c_int32),
    ("min_goodperall_pixel", C.c_float), ("min_goodperall_pixel_absmin", C.c_float),
    ("min_goodpergoodbad_pixel", C.c_float), ("max_diff_constant", C.c_float),
    ("max_diff_grad_mult", C.c_float), ("use_affine_lightning_estimation", C.c_int32),
    ("use_esm_tracking", C.c_int32),
    ("min_abs_grad_create", C.c_float), ("min_abs_grad_decrease", C.c_float),
    ("min_epl_length_squared", C.c_float), ("min_epl_grad_squared", C.c_float),
    ("min_epl_angle_squared", C.c_float), ("gradient_sample_dist", C.c_float),
    ("sample_point_to_border", C.c_float), ("max_epl_length_crop", C.c_float),
    ("min_epl_length_crop", C.c_float), ("max_error_stereo", C.c_float),
    ("min_distance_error_stereo", C.c_float), ("stereo_epl_var_fac", C.c_float),
    ("min_depth", C.c_float), ("max_var", C.c_float), ("var_random_init_initial", C.c_float),
    ("var_gt_init_initial", C.c_float), ("succ_var_inc_fac", C.c_float), ("fail_var_inc_fac", C.c_float),
    ("diff_fac_observe", C.c_float), ("diff_fac_prop_merge", C.c_float), ("diff_fac_smoothing", C.c_float),
    ("depth_smoothing_factor", C.c_float), ("division_eps", C.c_float),
    ("validity_counter_inc", C.c_int32), ("validity_counter_dec", C.c_int32),
    ("validity_counter_initial_observe", C.c_int32), ("validity_counter_max", C.c_int32),
    ("validity_counter_max_variable", C.c_int32), ("val_sum_min_for_create", C.c_int32),
    ("val_sum_min_for_keep", C.c_int32), ("val_sum_min_for_unblacklist", C.c_int32),
    ("min_blacklist", C.c_int32), ("allow_negative_idepths", C.c_int32), ("use_subpixel_stereo", C.c_int32),
    ("stereo_subpixel_oob_guard", C.c_int32), ("stereo_max_steps", C.c_int32),
]


class Params(C.Structure):
    _fields_ = _PARAM_FIELDS

    def as_dict(self):
        d = {}
        for name, _ in _PARAM_FIELDS:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class TrackStats(C.Structure):
    _fields_ = [("last_residual", C.c_float), ("last_depth_residual", C.c_float),
                ("last_photometric_residual", C.c_float), ("last_mean_res", C.c_float),
                ("point_usage", C.c_float), ("affine_a", C.c_float), ("affine_b", C.c_float),
                ("last_good_count", C.c_int32), ("last_bad_count", C.c_int32), ("diverged", C.c_int32),
                ("tracking_was_good", C.c_int32), ("num_residual_calls", C.c_int32 * PYR),
                ("num_warp_update_calls", C.c_int32 * PYR), ("hessian", C.c_double * 49)]


class Eval(C.Structure):
    _fields_ = [("A", C.c_double * 49), ("b", C.c_double * 7), ("error", C.c_double), ("sum_res_p", C.c_double),
                ("sum_res_d", C.c_double), ("num_terms_p", C.c_double), ("num_terms_d", C.c_double),
                ("n_constraints", C.c_double), ("warped_count", C.c_int32), ("good_count", C.c_int32),
                ("bad_count", C.c_int32), ("sum_res_unweighted", C.c_double), ("sum_signed_res", C.c_double),
                ("usage", C.c_double), ("sxx", C.c_double), ("syy", C.c_double), ("sx", C.c_double),
                ("sy", C.c_double), ("sw", C.c_double)]

    def A7(self):
        return np.array(self.A[:], np.float64).reshape(7, 7)


class Hypotheses(C.Structure):
    _fields_ = [("is_valid", C.c_void_p), ("blacklisted", C.c_void_p), ("next_stereo_frame_min_id", C.c_void_p),
                ("validity_counter", C.c_void_p), ("idepth", C.c_void_p), ("idepth_var", C.c_void_p),
                ("idepth_smoothed", C.c_void_p), ("idepth_var_smoothed", C.c_void_p),
                ("observe_code", C.c_void_p)]


HYP_FIELDS = [("is_valid", np.uint8), ("blacklisted", np.int32), ("next_stereo_frame_min_id", np.float32),
              ("validity_counter", np.int32), ("idepth", np.float32), ("idepth_var", np.float32),
              ("idepth_smoothed", np.float32), ("idepth_var_smoothed", np.float32), ("observe_code", np.uint8)]


def hyp_arrays(n):
    arrs = {k: np.zeros(n, dt) for k, dt in HYP_FIELDS}
    h = Hypotheses(*[arrs[k].ctypes.data for k, _ in HYP_FIELDS])
    return h, arrs


def hyp_struct(arrs):
    """Hypotheses struct pointing at (possibly partial) numpy arrays of the right dtypes."""
    vals = []
    for k, dt in HYP_FIELDS:
        a = arrs.get(k)
        if a is not None:
            assert a.dtype == dt and a.flags.c_contiguous
        vals.append(None if a is None else a.ctypes.data)
    return Hypotheses(*vals)


KERNEL_EVAL, KERNEL_PYRAMID, KERNEL_DEPTH, KERNEL_BATCH_EVAL, KERNEL_BATCH_SOLVE, KERNEL_BATCH_JOB, KERNEL_BATCH_COOP, \
    KERNEL_INGEST, KERNEL_ADJOINT = range(9)


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("ms", C.c_double), ("bytes", C.c_double)]


SCHEDULE_COOP, SCHEDULE_STEPS = 0, 1
_CONFIG_FIELDS = ["schedule", "coop_block", "coop_k", "solo_max", "job_max_points", "chains", "points_per_thread",
                  "xcd_groups"]


class BatchConfig(C.Structure):
    """lsdg_batch_config (include/lsdg_tune.h): the shape of a batch plan."""
    _fields_ = [(n, C.c_int32) for n in _CONFIG_FIELDS]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n in _CONFIG_FIELDS}

    @staticmethod
    def of(d):
        return BatchConfig(*[int(d[n]) for n in _CONFIG_FIELDS])


class Job(C.Structure):
    _fields_ = [("ref", C.c_void_p), ("frame", C.c_void_p), ("frame_to_ref_init", C.c_double * 8),
                ("kind", C.c_int32), ("start_level", C.c_int32), ("final_level", C.c_int32),
                ("rebuild_pyramid", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [("frame_to_ref", C.c_double * 8), ("stats", TrackStats), ("status", C.c_int32)]


class ConstraintCandidate(C.Structure):
    _fields_ = [("frame_ref", C.c_void_p), ("candidate_ref", C.c_void_p),
                ("candidate_to_frame_init", C.c_double * 8)]


class ConstraintEdge(C.Structure):
    """KFConstraintStruct: second_to_first, information and residuals."""
    _fields_ = [("first_id", C.c_int32), ("second_id", C.c_int32), ("second_to_first", C.c_double * 8),
                ("information", C.c_double * 49), ("mean_residual", C.c_float), ("mean_residual_d", C.c_float),
                ("mean_residual_p", C.c_float), ("usage", C.c_float)]


class ConstraintResult(C.Structure):
    _fields_ = [("accepted", C.c_int32), ("rejected_at", C.c_int32), ("error", C.c_float * 3),
                ("candidate_to_frame", C.c_double * 8), ("frame_to_candidate", C.c_double * 8),
                ("candidate_on_frame", ConstraintEdge), ("frame_on_candidate", ConstraintEdge)]


class KeyframePose(C.Structure):
    _fields_ = [("id", C.c_int32), ("mean_idepth", C.c_float), ("cam_to_world", C.c_double * 8)]


class SystemParams(C.Structure):
    _fields_ = [("kf_dist_weight", C.c_float), ("kf_usage_weight", C.c_float), ("min_num_mapped", C.c_int32),
                ("initialization_phase_count", C.c_int32), ("relocalise", C.c_int32), ("reloc_level", C.c_int32)]


class SystemFrame(C.Structure):
    _fields_ = [("id", C.c_int32), ("status", C.c_int32), ("keyframe_id", C.c_int32), ("new_keyframe", C.c_int32),
                ("closeness_score", C.c_float), ("keyframe_mean_idepth", C.c_float),
                ("frame_to_keyframe", C.c_double * 7), ("cam_to_world", C.c_double * 8), ("stats", TrackStats)]


SYSTEM_TRACKED, SYSTEM_LOST, SYSTEM_NOT_TRACKED, SYSTEM_RELOCALISED = 0, 1, 2, 3


class PointDense(C.Structure):
    """InputPointDense (lib/Pangolin_IOWrapper/Keyframe.h:16-21)."""
    _fields_ = [("idepth", C.c_float), ("idepth_var", C.c_float), ("color", C.c_uint8 * 4)]


class Vertex(C.Structure):
    """Keyframe::MyVertex (lib/Pangolin_IOWrapper/Keyframe.h:48-52)."""
    _fields_ = [("point", C.c_float * 3), ("color", C.c_uint8 * 4)]


class VboParams(C.Structure):
    _fields_ = [("scaled_th", C.c_float), ("abs_th", C.c_float), ("scale", C.c_float),
                ("min_near_support", C.c_int32)]


POINT_DENSE_DTYPE = np.dtype([("idepth", "<f4"), ("idepth_var", "<f4"), ("color", "u1", (4,))])
VERTEX_DTYPE = np.dtype([("point", "<f4", (3,)), ("color", "u1", (4,))])


def vbo_params(scale=1.0, **kw):
    """Keyframe::computeVbo's constants (Keyframe.h:78-82) with the keyframe's
    camToWorld scale; keyword arguments override fields."""
    p = VboParams()
    load().lsdg_vbo_params_default(C.byref(p))
    p.scale = scale
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def pose7(p):
    return (C.c_double * 7)(*[float(v) for v in list(p)[:7]])


def pose8(p):
    p = list(p)
    if len(p) == 7:
        p = p + [1.0]
    return (C.c_double * 8)(*[float(v) for v in p])


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data


_LIB = None


def lib_path():
    return os.path.join(HERE, "liblsdg.so")


def load():
    """Loads liblsdg.so (the HIP build).  Raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    vp, i32, f32, dp, fp = C.c_void_p, C.c_int32, C.c_float, C.POINTER(C.c_double), C.c_void_p
    sig = {
        "lsdg_params_default": (None, [C.POINTER(Params)]),
        "lsdg_ctx_create": (i32, [i32, C.POINTER(Params), C.POINTER(vp)]),
        "lsdg_ctx_destroy": (i32, [vp]),
        "lsdg_ctx_set_params": (i32, [vp, C.POINTER(Params)]),
        "lsdg_ctx_synchronize": (i32, [vp]),
        "lsdg_ctx_set_sums": (i32, [vp, i32]),
        "lsdg_ctx_get_sums": (i32, [vp, C.POINTER(i32)]),
        "lsdg_ctx_last_error": (C.c_char_p, [vp]),
        "lsdg_frame_create": (i32, [vp, i32, C.POINTER(Camera), fp, i32, C.POINTER(vp)]),
        "lsdg_frame_create_u8": (i32, [vp, i32, C.POINTER(Camera), fp, i32, i32, C.POINTER(vp)]),
        "lsdg_frame_destroy": (i32, [vp]),
        "lsdg_frame_build_pyramid": (i32, [vp]),
        "lsdg_frame_invalidate": (i32, [vp]),
        "lsdg_frame_set_depth": (i32, [vp, fp, fp]),
        "lsdg_frame_read": (i32, [vp, i32, i32, fp, C.c_size_t]),
        "lsdg_frame_set_tracking": (i32, [vp, vp, dp, f32, fp]),
        "lsdg_frame_get_tracking": (i32, [vp, C.POINTER(i32), dp, C.POINTER(f32)]),
        "lsdg_frame_set_cam_to_world": (i32, [vp, dp]),
        "lsdg_frame_counters": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
        "lsdg_frame_set_counters": (i32, [vp, i32, i32]),
        "lsdg_ref_create": (i32, [vp, vp, C.POINTER(vp)]),
        "lsdg_ref_destroy": (i32, [vp]),
        "lsdg_ref_make_point_cloud": (i32, [vp, i32]),
        "lsdg_ref_num_points": (i32, [vp, i32, C.POINTER(i32)]),
        "lsdg_ref_read_points": (i32, [vp, i32, fp, i32]),
        "lsdg_se3_track": (i32, [vp, vp, vp, dp, dp, C.POINTER(TrackStats)]),
        "lsdg_sim3_track": (i32, [vp, vp, vp, dp, i32, i32, dp, C.POINTER(TrackStats)]),
        "lsdg_se3_evaluate": (i32, [vp, vp, vp, i32, dp, f32, f32, C.POINTER(Eval), fp]),
        "lsdg_sim3_evaluate": (i32, [vp, vp, vp, i32, dp, C.POINTER(Eval)]),
        "lsdg_energy_grad": (i32, [vp, vp, vp, i32, dp, dp, dp, fp, i32]),
        "lsdg_energy_grad_sim3": (i32, [vp, vp, vp, i32, dp, dp, dp, fp, i32]),
        "lsdg_energy_grad_noalloc": (i32, [vp, vp, vp, i32, dp, dp, dp, fp, i32]),
        "lsdg_batch_tune_key": (i32, [vp, vp, i32, C.c_char_p]),
        "lsdg_batch_create_with_config": (i32, [vp, vp, i32, C.POINTER(BatchConfig), C.POINTER(vp)]),
        "lsdg_batch_get_config": (i32, [vp, C.POINTER(BatchConfig)]),
        "lsdg_batch_tune": (i32, [vp, vp, i32, i32, C.POINTER(BatchConfig), dp, dp]),
        "lsdg_tune_table_put": (i32, [C.c_char_p, C.POINTER(BatchConfig)]),
        "lsdg_tune_table_get": (i32, [C.c_char_p, C.POINTER(BatchConfig)]),
        "lsdg_tune_table_clear": (i32, []),
        "lsdg_tune_table_size": (i32, []),
        "lsdg_tune_table_save": (i32, [C.c_char_p]),
        "lsdg_tune_table_load": (i32, [C.c_char_p]),
        "lsdg_depthmap_create": (i32, [vp, C.POINTER(Camera), C.POINTER(vp)]),
        "lsdg_depthmap_destroy": (i32, [vp]),
        "lsdg_depth_reset": (i32, [vp]),
        "lsdg_depth_is_valid": (i32, [vp]),
        "lsdg_depth_initialize_from_gt": (i32, [vp, vp]),
        "lsdg_depth_initialize_randomly": (i32, [vp, vp, C.c_uint64]),
        "lsdg_depth_set_prior": (i32, [vp, vp, fp, fp, i32]),
        "lsdg_depth_update": (i32, [vp, C.POINTER(vp), i32]),
        "lsdg_depth_create_keyframe": (i32, [vp, vp]),
        "lsdg_depth_finalize_keyframe": (i32, [vp]),
        "lsdg_depth_read": (i32, [vp, C.POINTER(Hypotheses)]),
        "lsdg_depth_write": (i32, [vp, C.POINTER(Hypotheses)]),
        "lsdg_depth_active_keyframe": (i32, [vp, C.POINTER(i32)]),
        "lsdg_ctx_profile": (i32, [vp, i32]),
        "lsdg_ctx_kernel_stats": (i32, [vp, i32, C.POINTER(KernelStats)]),
        "lsdg_batch_create": (i32, [vp, C.POINTER(Job), i32, C.POINTER(vp)]),
        "lsdg_batch_run": (i32, [vp, C.POINTER(JobResult)]),
        "lsdg_batch_destroy": (i32, [vp]),
        "lsdg_batch_track": (i32, [vp, C.POINTER(Job), i32, C.POINTER(JobResult)]),
        "lsdg_vbo_params_default": (None, [C.POINTER(VboParams)]),
        "lsdg_keyframe_points": (i32, [vp, i32, fp]),
        "lsdg_points_vbo": (i32, [vp, fp, i32, i32, f32, f32, f32, f32, C.POINTER(VboParams), fp, C.POINTER(i32)]),
        "lsdg_keyframe_vbo": (i32, [vp, i32, C.POINTER(VboParams), fp, C.POINTER(i32)]),
        "lsdg_undistorter_create": (i32, [vp, i32, i32, i32, i32, fp, C.POINTER(vp)]),
        "lsdg_undistorter_destroy": (i32, [vp]),
        "lsdg_fov_undistort_table": (i32, [fp, f32, i32, i32, fp, i32, i32, fp]),
        "lsdg_frame_create_undistorted": (i32, [vp, i32, C.POINTER(Camera), vp, fp, i32, i32, fp, C.POINTER(vp)]),
        "lsdg_frame_pose_txt": (i32, [vp, C.c_char_p, C.c_size_t]),
        "lsdg_pose_txt_line": (i32, [i32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p, C.c_size_t]),
        "lsdg_frame_ingest": (i32, [vp, i32, vp, fp, i32, fp]),
        "lsdg_system_params_default": (None, [C.POINTER(SystemParams)]),
        "lsdg_system_create": (i32, [vp, C.POINTER(Camera), i32, C.POINTER(SystemParams), C.POINTER(vp)]),
        "lsdg_system_destroy": (i32, [vp]),
        "lsdg_system_init_gt": (i32, [vp, i32, fp, i32, fp, fp]),
        "lsdg_system_init_random": (i32, [vp, i32, fp, i32, C.c_uint64]),
        "lsdg_system_track": (i32, [vp, i32, fp, i32, C.POINTER(SystemFrame)]),
        "lsdg_system_keyframes": (i32, [vp, C.POINTER(i32), C.POINTER(C.c_double), i32, C.POINTER(i32)]),
        "lsdg_system_depth_read": (i32, [vp, C.POINTER(Hypotheses)]),
        "lsdg_system_find_constraints": (i32, [vp, i32, f32, f32, C.POINTER(ConstraintResult), C.POINTER(i32), i32,
                                               C.POINTER(i32)]),
        "lsdg_keyframe_candidates": (i32, [C.POINTER(Camera), C.POINTER(KeyframePose), C.POINTER(KeyframePose), i32,
                                           f32, f32, C.POINTER(i32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           i32, C.POINTER(i32)]),
        "lsdg_constraint_search": (i32, [vp, C.POINTER(ConstraintCandidate), i32, f32,
                                         C.POINTER(ConstraintResult)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def default_params():
    p = Params()
    load().lsdg_params_default(C.byref(p))
    return p


class LsdgError(RuntimeError):
    pass


class Context:
    """One lsdg_ctx (device + HIP stream + parameters)."""

    def __init__(self, device=0, params=None):
        self.L = load()
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        self._check(self.L.lsdg_ctx_create(device, C.byref(self.params), C.byref(h)), "lsdg_ctx_create", None)
        self.h = h

    def _check(self, st, what, ctx=None):
        if st < 0:
            msg = self.L.lsdg_ctx_last_error(ctx) if ctx is not None else b""
            raise LsdgError(f"{what} failed with status {st}: {msg.decode() if msg else ''}")
        return st

    def check(self, st, what):
        return self._check(st, what, self.h)

    def set_params(self, params):
        self.params = params
        self.check(self.L.lsdg_ctx_set_params(self.h, C.byref(params)), "lsdg_ctx_set_params")

    def sync(self):
        self.check(self.L.lsdg_ctx_synchronize(self.h), "lsdg_ctx_synchronize")

    def set_sums(self, mode):
        """SUMS_FP64 (default) or SUMS_KAHAN: how tracker evaluations and
        batch plans created afterwards sum their FP32 per-point terms."""
        self.check(self.L.lsdg_ctx_set_sums(self.h, int(mode)), "lsdg_ctx_set_sums")

    def sums(self):
        m = C.c_int32()
        self.check(self.L.lsdg_ctx_get_sums(self.h, C.byref(m)), "lsdg_ctx_get_sums")
        return m.value

    def profile(self, enable):
        self.check(self.L.lsdg_ctx_profile(self.h, int(enable)), "lsdg_ctx_profile")

    def kernel_stats(self, kind):
        st = KernelStats()
        self.check(self.L.lsdg_ctx_kernel_stats(self.h, kind, C.byref(st)), "lsdg_ctx_kernel_stats")
        return dict(launches=st.launches, ms=st.ms, bytes=st.bytes)

    def close(self):
        if self.h:
            self.L.lsdg_ctx_destroy(self.h)
            self.h = None

    def frame(self, fid, cam, image, precision=FP32):
        h = C.c_void_p()
        cc = Camera.of(cam)
        if image.dtype == np.uint8:
            img = np.ascontiguousarray(image)
            self.check(self.L.lsdg_frame_create_u8(self.h, fid, C.byref(cc), img.ctypes.data, cam["w"], precision,
                                                   C.byref(h)), "lsdg_frame_create_u8")
        else:
            img, ptr = _f32(image)
            self.check(self.L.lsdg_frame_create(self.h, fid, C.byref(cc), ptr, precision, C.byref(h)),
                       "lsdg_frame_create")
        return Frame(self, h, fid, cam)

    def ref(self, keyframe):
        h = C.c_void_p()
        self.check(self.L.lsdg_ref_create(self.h, keyframe.h, C.byref(h)), "lsdg_ref_create")
        return Ref(self, h, keyframe)

    def se3_track(self, ref, frame, init7):
        out = (C.c_double * 7)()
        st = TrackStats()
        status = self.check(self.L.lsdg_se3_track(self.h, ref.h, frame.h, pose7(init7), out, C.byref(st)),
                            "lsdg_se3_track")
        return np.array(out[:]), st, status

    def sim3_track(self, ref, frame, init8, start_level, final_level):
        out = (C.c_double * 8)()
        st = TrackStats()
        status = self.check(self.L.lsdg_sim3_track(self.h, ref.h, frame.h, pose8(init8), start_level, final_level,
                                                   out, C.byref(st)), "lsdg_sim3_track")
        return np.array(out[:]), st, status

    def se3_evaluate(self, ref, frame, level, pose, a=1.0, b=0.0, want_mask=False):
        ev = Eval()
        mask = None
        if want_mask:
            mask = np.full((frame.cam["h"] >> 1) * (frame.cam["w"] >> 1), 1, np.uint8)
        self.check(self.L.lsdg_se3_evaluate(self.h, ref.h, frame.h, level, pose7(pose), a, b, C.byref(ev),
                                            None if mask is None else mask.ctypes.data), "lsdg_se3_evaluate")
        return ev, mask

    def sim3_evaluate(self, ref, frame, level, pose8_):
        ev = Eval()
        self.check(self.L.lsdg_sim3_evaluate(self.h, ref.h, frame.h, level, pose8(pose8_), C.byref(ev)),
                   "lsdg_sim3_evaluate")
        return ev

    def energy_grad(self, ref, frame, level, pose):
        n = ref.num_points(level)
        e = C.c_double()
        g = (C.c_double * 6)()
        gd = np.zeros(n, np.float32)
        self.check(self.L.lsdg_energy_grad(self.h, ref.h, frame.h, level, pose7(pose), C.byref(e), g,
                                           gd.ctypes.data, n), "lsdg_energy_grad")
        return e.value, np.array(g[:]), gd

    def energy_grad_sim3(self, ref, frame, level, pose8_):
        """Sim3 adjoint: (E, dE/dxi[7], dE/dd per point)."""
        n = ref.num_points(level)
        e = C.c_double()
        g = (C.c_double * 7)()
        gd = np.zeros(n, np.float32)
        self.check(self.L.lsdg_energy_grad_sim3(self.h, ref.h, frame.h, level, pose8(pose8_), C.byref(e), g,
                                                gd.ctypes.data, n), "lsdg_energy_grad_sim3")
        return e.value, np.array(g[:]), gd

    def points_vbo(self, records, w, h, cam4, params=None):
        """Keyframe::computeVbo over InputPointDense records (w x h, raster
        order) of a camera (fx, fy, cx, cy): the kept vertices."""
        rec = np.ascontiguousarray(records, POINT_DENSE_DTYPE)
        assert rec.size == w * h
        out = np.zeros(max(w * h, 1), VERTEX_DTYPE)
        n = C.c_int32(0)
        p = params if params is not None else vbo_params()
        fx, fy, cx, cy = (float(v) for v in cam4)
        self.check(self.L.lsdg_points_vbo(self.h, rec.ctypes.data, w, h, fx, fy, cx, cy, C.byref(p), out.ctypes.data,
                                          C.byref(n)), "lsdg_points_vbo")
        return out[:n.value]

    def system(self, cam, precision=FP32, params=None):
        """lsd-slam's SlamSystem front end (tracking + mapping) on this context."""
        h = C.c_void_p()
        cc = Camera.of(cam)
        self.check(self.L.lsdg_system_create(self.h, C.byref(cc), precision, None if params is None else C.byref(params),
                                             C.byref(h)), "lsdg_system_create")
        return System(self, h, cam)

    def constraint_search(self, candidates, strictness=1.5):
        """SlamSystem::testConstraint for [(frame_ref, candidate_ref, candidate_to_frame_init8), ...]."""
        n = len(candidates)
        arr = (ConstraintCandidate * max(n, 1))()
        for i, (a, b, init) in enumerate(candidates):
            arr[i].frame_ref = a.h
            arr[i].candidate_ref = b.h
            arr[i].candidate_to_frame_init[:] = [float(v) for v in init]
        out = (ConstraintResult * max(n, 1))()
        self.check(self.L.lsdg_constraint_search(self.h, arr, n, float(strictness), out), "lsdg_constraint_search")
        return list(out[:n])

    def undistorter(self, in_size, table):
        """libvideoio's Undistorter from a (out_h, out_w, 2) table of source
        positions into the in_size = (w, h) distorted image (x < 0: no source)."""
        tab = np.ascontiguousarray(table, np.float32)
        h = C.c_void_p()
        self.check(self.L.lsdg_undistorter_create(self.h, in_size[0], in_size[1], tab.shape[1], tab.shape[0],
                                                  tab.ctypes.data, C.byref(h)), "lsdg_undistorter_create")
        return Undistorter(self, h, in_size, (tab.shape[1], tab.shape[0]))

    def frame_undistorted(self, fid, cam, und, image, precision=FP32, want_undistorted=False):
        """InputThread's ingest: undistort the 8-bit image, convert, build the
        image pyramid; returns the frame (and the 8-bit undistorted image)."""
        img = np.ascontiguousarray(image, np.uint8)
        assert img.shape == (und.in_size[1], und.in_size[0])
        cc = Camera.of(cam)
        out = np.zeros((cam["h"], cam["w"]), np.uint8) if want_undistorted else None
        h = C.c_void_p()
        self.check(self.L.lsdg_frame_create_undistorted(self.h, fid, C.byref(cc), und.h, img.ctypes.data,
                                                        img.shape[1], precision,
                                                        None if out is None else out.ctypes.data, C.byref(h)),
                   "lsdg_frame_create_undistorted")
        f = Frame(self, h, fid, cam)
        return (f, out) if want_undistorted else f

    def depthmap(self, cam):
        h = C.c_void_p()
        cc = Camera.of(cam)
        self.check(self.L.lsdg_depthmap_create(self.h, C.byref(cc), C.byref(h)), "lsdg_depthmap_create")
        return DepthMapH(self, h, cam)

    def batch_tune(self, jobs, runs=5):
        """lsdg_batch_tune: (best shape, default median ms, best median ms);
        later plans of this class of batch take the best shape."""
        arr = (Job * len(jobs))(*jobs)
        best = BatchConfig()
        d, b = C.c_double(), C.c_double()
        self.check(self.L.lsdg_batch_tune(self.h, arr, len(jobs), runs, C.byref(best), C.byref(d), C.byref(b)),
                   "lsdg_batch_tune")
        return best.as_dict(), d.value, b.value

    def batch_key(self, jobs):
        arr = (Job * len(jobs))(*jobs)
        buf = C.create_string_buffer(64)
        self.check(self.L.lsdg_batch_tune_key(self.h, arr, len(jobs), buf), "lsdg_batch_tune_key")
        return buf.value.decode()

    def batch_with_config(self, jobs, config):
        arr = (Job * len(jobs))(*jobs)
        cfg = BatchConfig.of(config)
        h = C.c_void_p()
        self.check(self.L.lsdg_batch_create_with_config(self.h, arr, len(jobs), C.byref(cfg), C.byref(h)),
                   "lsdg_batch_create_with_config")
        return Batch(self, h, len(jobs))

    def batch(self, jobs):
        arr = (Job * len(jobs))(*jobs)
        h = C.c_void_p()
        self.check(self.L.lsdg_batch_create(self.h, arr, len(jobs), C.byref(h)), "lsdg_batch_create")
        return Batch(self, h, len(jobs))


class System:
    def __init__(self, ctx, h, cam):
        self.ctx, self.h, self.cam = ctx, h, cam

    def init_gt(self, fid, image, idepth, var):
        img = np.ascontiguousarray(image, np.uint8)
        a, pa = _f32(idepth)
        b, pb = _f32(var)
        self.ctx.check(self.ctx.L.lsdg_system_init_gt(self.h, fid, img.ctypes.data, img.shape[1], pa, pb),
                       "lsdg_system_init_gt")

    def init_random(self, fid, image, seed):
        img = np.ascontiguousarray(image, np.uint8)
        self.ctx.check(self.ctx.L.lsdg_system_init_random(self.h, fid, img.ctypes.data, img.shape[1], C.c_uint64(seed)),
                       "lsdg_system_init_random")

    def track(self, fid, image):
        img = np.ascontiguousarray(image, np.uint8)
        out = SystemFrame()
        self.ctx.check(self.ctx.L.lsdg_system_track(self.h, fid, img.ctypes.data, img.shape[1], C.byref(out)),
                       "lsdg_system_track")
        return out

    def keyframes(self):
        n = C.c_int32()
        self.ctx.check(self.ctx.L.lsdg_system_keyframes(self.h, None, None, 0, C.byref(n)), "lsdg_system_keyframes")
        ids = (C.c_int32 * max(n.value, 1))()
        c2w = (C.c_double * (8 * max(n.value, 1)))()
        self.ctx.check(self.ctx.L.lsdg_system_keyframes(self.h, ids, c2w, n.value, C.byref(n)), "lsdg_system_keyframes")
        return list(ids[:n.value]), np.array(c2w[:8 * n.value]).reshape(-1, 8)

    def depth_read(self):
        n = self.cam["w"] * self.cam["h"]
        h, arrs = hyp_arrays(n)
        self.ctx.check(self.ctx.L.lsdg_system_depth_read(self.h, C.byref(h)), "lsdg_system_depth_read")
        return arrs

    def find_constraints(self, index, closeness_th=1.0, strictness=1.5):
        """findConstraintsForNewKeyframes for keyframe `index`: [(candidate id, ConstraintResult), ...]."""
        cap = 64
        res = (ConstraintResult * cap)()
        ids = (C.c_int32 * cap)()
        n = C.c_int32()
        self.ctx.check(self.ctx.L.lsdg_system_find_constraints(self.h, index, closeness_th, strictness, res, ids, cap,
                                                               C.byref(n)), "lsdg_system_find_constraints")
        return [(ids[i], res[i]) for i in range(min(n.value, cap))]

    def destroy(self):
        if self.h:
            self.ctx.L.lsdg_system_destroy(self.h)
            self.h = None


def fov_undistort_table(in_k, omega, in_size, out_k, out_size):
    """The FOV lens model's undistortion table, (out_h, out_w, 2) float32."""
    L = load()
    ik = (C.c_float * 4)(*in_k)
    ok = (C.c_float * 4)(*out_k)
    out = np.zeros((out_size[1], out_size[0], 2), np.float32)
    st = L.lsdg_fov_undistort_table(ik, float(omega), in_size[0], in_size[1], ok, out_size[0], out_size[1],
                                    out.ctypes.data)
    if st < 0:
        raise LsdgError(f"lsdg_fov_undistort_table failed with status {st}")
    return out


def pose_txt_line(fid, cam_to_world8, this_to_parent8):
    """TextOutputIOWrapper::publishTrackedFrame's pose.txt line for an id and two poses."""
    a = (C.c_double * 8)(*[float(v) for v in cam_to_world8])
    b = (C.c_double * 8)(*[float(v) for v in this_to_parent8])
    buf = C.create_string_buffer(512)
    st = load().lsdg_pose_txt_line(int(fid), a, b, buf, 512)
    if st < 0:
        raise LsdgError(f"lsdg_pose_txt_line failed with status {st}")
    return buf.value.decode()


class Undistorter:
    def __init__(self, ctx, h, in_size, out_size):
        self.ctx, self.h, self.in_size, self.out_size = ctx, h, in_size, out_size

    def close(self):
        if self.h:
            self.ctx.L.lsdg_undistorter_destroy(self.h)
            self.h = None


class Frame:
    def __init__(self, ctx, h, fid, cam):
        self.ctx, self.h, self.id, self.cam = ctx, h, fid, cam

    def size(self, level):
        return self.cam["w"] >> level, self.cam["h"] >> level

    def read(self, what, level=0):
        w, h = self.size(level)
        if what == WHAT_GRADIENTS:
            out = np.zeros((h, w, 4), np.float32)
        elif what == WHAT_REF_PIXEL_WAS_GOOD:
            w, h = self.size(1)
            out = np.zeros((h, w), np.uint8)
        else:
            out = np.zeros((h, w), np.float32)
        self.ctx.check(self.ctx.L.lsdg_frame_read(self.h, what, level, out.ctypes.data, out.nbytes),
                       "lsdg_frame_read")
        return out

    def build_pyramid(self):
        self.ctx.check(self.ctx.L.lsdg_frame_build_pyramid(self.h), "lsdg_frame_build_pyramid")

    def keyframe_points(self, level=0):
        """publishKeyframe's InputPointDense records of `level` (raster order)."""
        w, h = self.size(level)
        out = np.zeros(w * h, POINT_DENSE_DTYPE)
        self.ctx.check(self.ctx.L.lsdg_keyframe_points(self.h, level, out.ctypes.data), "lsdg_keyframe_points")
        return out

    def keyframe_vbo(self, level=0, params=None):
        """publishKeyframe + Keyframe::computeVbo on the device: the kept vertices."""
        w, h = self.size(level)
        out = np.zeros(max(w * h, 1), VERTEX_DTYPE)
        n = C.c_int32(0)
        p = params if params is not None else vbo_params()
        self.ctx.check(self.ctx.L.lsdg_keyframe_vbo(self.h, level, C.byref(p), out.ctypes.data, C.byref(n)),
                       "lsdg_keyframe_vbo")
        return out[:n.value]

    def ingest(self, fid, und, image, want_undistorted=False):
        """Reuse this frame for a new camera image (lsdg_frame_ingest)."""
        img = np.ascontiguousarray(image, np.uint8)
        assert img.shape == (und.in_size[1], und.in_size[0])
        out = np.zeros((self.cam["h"], self.cam["w"]), np.uint8) if want_undistorted else None
        self.ctx.check(self.ctx.L.lsdg_frame_ingest(self.h, fid, und.h, img.ctypes.data, img.shape[1],
                                                    None if out is None else out.ctypes.data), "lsdg_frame_ingest")
        self.id = fid
        return out

    def pose_txt(self):
        """TextOutputIOWrapper::publishTrackedFrame's pose.txt line."""
        buf = C.create_string_buffer(512)
        self.ctx.check(self.ctx.L.lsdg_frame_pose_txt(self.h, buf, 512), "lsdg_frame_pose_txt")
        return buf.value.decode()

    def invalidate(self):
        """Drop the derived pyramid levels; the next reader rebuilds what it needs."""
        self.ctx.check(self.ctx.L.lsdg_frame_invalidate(self.h), "lsdg_frame_invalidate")

    def set_depth(self, idepth, var):
        a, pa = _f32(idepth)
        b, pb = _f32(var)
        self.ctx.check(self.ctx.L.lsdg_frame_set_depth(self.h, pa, pb), "lsdg_frame_set_depth")

    def set_tracking(self, parent, this_to_parent8, initial_tracked_residual=0.0, was_good=None):
        wg = None
        if was_good is not None:
            wg = np.ascontiguousarray(was_good, np.uint8)
        self.ctx.check(self.ctx.L.lsdg_frame_set_tracking(self.h, None if parent is None else parent.h,
                                                          pose8(this_to_parent8), initial_tracked_residual,
                                                          None if wg is None else wg.ctypes.data),
                       "lsdg_frame_set_tracking")

    def get_tracking(self):
        pid = C.c_int32()
        tp = (C.c_double * 8)()
        itr = C.c_float()
        self.ctx.check(self.ctx.L.lsdg_frame_get_tracking(self.h, C.byref(pid), tp, C.byref(itr)),
                       "lsdg_frame_get_tracking")
        return pid.value, np.array(tp[:]), itr.value

    def set_cam_to_world(self, c2w8):
        self.ctx.check(self.ctx.L.lsdg_frame_set_cam_to_world(self.h, pose8(c2w8)), "lsdg_frame_set_cam_to_world")

    def counters(self):
        a, b = C.c_int32(), C.c_int32()
        self.ctx.check(self.ctx.L.lsdg_frame_counters(self.h, C.byref(a), C.byref(b)), "lsdg_frame_counters")
        return a.value, b.value

    def set_counters(self, tracked, mapped):
        self.ctx.check(self.ctx.L.lsdg_frame_set_counters(self.h, tracked, mapped), "lsdg_frame_set_counters")

    def destroy(self):
        if self.h:
            self.ctx.L.lsdg_frame_destroy(self.h)
            self.h = None


class Ref:
    def __init__(self, ctx, h, kf):
        self.ctx, self.h, self.kf = ctx, h, kf

    def num_points(self, level):
        n = C.c_int32()
        self.ctx.check(self.ctx.L.lsdg_ref_num_points(self.h, level, C.byref(n)), "lsdg_ref_num_points")
        return n.value

    def points(self, level):
        n = self.num_points(level)
        out = np.zeros((n, 8), np.float32)
        self.ctx.check(self.ctx.L.lsdg_ref_read_points(self.h, level, out.ctypes.data, n), "lsdg_ref_read_points")
        return out

    def destroy(self):
        if self.h:
            self.ctx.L.lsdg_ref_destroy(self.h)
            self.h = None


class DepthMapH:
    def __init__(self, ctx, h, cam):
        self.ctx, self.h, self.cam = ctx, h, cam

    def set_prior(self, kf, idepth, var, validity):
        a, pa = _f32(idepth)
        b, pb = _f32(var)
        self.ctx.check(self.ctx.L.lsdg_depth_set_prior(self.h, kf.h, pa, pb, validity), "lsdg_depth_set_prior")

    def init_gt(self, kf):
        self.ctx.check(self.ctx.L.lsdg_depth_initialize_from_gt(self.h, kf.h), "lsdg_depth_initialize_from_gt")

    def init_random(self, kf, seed):
        self.ctx.check(self.ctx.L.lsdg_depth_initialize_randomly(self.h, kf.h, C.c_uint64(seed)),
                       "lsdg_depth_initialize_randomly")

    def update(self, refs):
        arr = (C.c_void_p * len(refs))(*[r.h for r in refs])
        self.ctx.check(self.ctx.L.lsdg_depth_update(self.h, arr, len(refs)), "lsdg_depth_update")

    def create_keyframe(self, nkf):
        self.ctx.check(self.ctx.L.lsdg_depth_create_keyframe(self.h, nkf.h), "lsdg_depth_create_keyframe")

    def finalize(self):
        self.ctx.check(self.ctx.L.lsdg_depth_finalize_keyframe(self.h), "lsdg_depth_finalize_keyframe")

    def read(self):
        n = self.cam["w"] * self.cam["h"]
        h, arrs = hyp_arrays(n)
        self.ctx.check(self.ctx.L.lsdg_depth_read(self.h, C.byref(h)), "lsdg_depth_read")
        return arrs

    def write(self, arrs):
        h = hyp_struct(arrs)
        self.ctx.check(self.ctx.L.lsdg_depth_write(self.h, C.byref(h)), "lsdg_depth_write")

    def destroy(self):
        if self.h:
            self.ctx.L.lsdg_depthmap_destroy(self.h)
            self.h = None


class Batch:
    def __init__(self, ctx, h, n):
        self.ctx, self.h, self.n = ctx, h, n

    def run(self):
        res = (JobResult * self.n)()
        self.ctx.check(self.ctx.L.lsdg_batch_run(self.h, res), "lsdg_batch_run")
        return res

    def config(self):
        """The shape this plan took."""
        c = BatchConfig()
        self.ctx.check(self.ctx.L.lsdg_batch_get_config(self.h, C.byref(c)), "lsdg_batch_get_config")
        return c.as_dict()

    def destroy(self):
        if self.h:
            self.ctx.L.lsdg_batch_destroy(self.h)
            self.h = None


def tune_table_put(key, config):
    st = load().lsdg_tune_table_put(key.encode(), C.byref(BatchConfig.of(config)))
    if st != OK:
        raise LsdgError(f"lsdg_tune_table_put failed with status {st}")


def tune_table_get(key):
    c = BatchConfig()
    st = load().lsdg_tune_table_get(key.encode(), C.byref(c))
    return c.as_dict() if st == OK else None


def tune_table_clear():
    load().lsdg_tune_table_clear()


def tune_table_size():
    return int(load().lsdg_tune_table_size())


def tune_table_save(path):
    st = load().lsdg_tune_table_save(os.fsencode(path))
    if st != OK:
        raise LsdgError(f"lsdg_tune_table_save({path}) failed with status {st}")


def tune_table_load(path):
    st = load().lsdg_tune_table_load(os.fsencode(path))
    if st != OK:
        raise LsdgError(f"lsdg_tune_table_load({path}) failed with status {st}")


def keyframe_candidates(cam, keyframe, others, closeness_th=1.0, kf_dist_weight=4.0):
    """lsdg_keyframe_candidates (host computation): keyframe and others are
    (id, mean_idepth, cam_to_world8) tuples; returns [(index, dist2, cos), ...]."""
    L = load()
    n = len(others)
    arr = (KeyframePose * max(n, 1))(*[KeyframePose(i, m, (C.c_double * 8)(*p)) for i, m, p in others])
    kf = KeyframePose(keyframe[0], keyframe[1], (C.c_double * 8)(*keyframe[2]))
    idx = (C.c_int32 * max(n, 1))()
    d2 = (C.c_double * max(n, 1))()
    cs = (C.c_double * max(n, 1))()
    cnt = C.c_int32()
    cc = Camera.of(cam)
    st = L.lsdg_keyframe_candidates(C.byref(cc), C.byref(kf), arr, n, closeness_th, kf_dist_weight, idx, d2, cs, n,
                                    C.byref(cnt))
    if st != OK:
        raise LsdgError(f"lsdg_keyframe_candidates failed with status {st}")
    return [(idx[i], d2[i], cs[i]) for i in range(cnt.value)]
