/*
 * lsdg.h — C-ABI of the MI355X (gfx950) LSD-SLAM per-pixel front end.
 *
 * The reference wrapper (apl-ocean-engineering/lsd-slam-pangolin-gui) links
 * the LSD-SLAM core as the external fips module `lsdslam`
 * (fips.yml:2-4, lib/CMakeLists.txt:16, tools/CMakeLists.txt:6) and reaches
 * the hot path only through SlamSystem::nextImage
 * (lib/App/InputThread.cpp:71).  Inside that core the hot path is the C++
 * classes SE3Tracker, Sim3Tracker, TrackingReference, Frame and DepthMap.
 * They have no FFI of their own; the entry points below are what such an FFI
 * binds, one per public method, with plain pointers and sizes.  The C++ shim
 * in lsd-slam-pangolin-gui_amd/shim/ wraps them back into those classes.
 *
 * Conventions
 *   - images and per-pixel maps are row-major, idx = x + y*w;
 *   - SE3 poses are double[7]  = (qx, qy, qz, qw, tx, ty, tz), unit quaternion;
 *   - Sim3 poses are double[8] = (qx, qy, qz, qw, tx, ty, tz, scale);
 *   - twists are (v1, v2, v3, w1, w2, w3[, sigma]); updates are T <- exp(d) * T;
 *   - pyramid level l has size (w >> l) x (h >> l), fx_l = fx / 2^l,
 *     cx_l = (cx + 0.5) / 2^l - 0.5 (likewise fy, cy);
 *   - inverse depth <= 0 or variance <= 0 marks a pixel without depth, as in the
 *     keyframe arrays the wrapper publishes
 *     (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:77-89).
 *
 * Errors are returned, never thrown.  A tracker that diverges returns
 * LSDG_DIVERGED and writes the identity pose, as the reference trackers
 * return SE3() / Sim3() with `diverged` set.
 *
 * Threading: one lsdg_ctx per host thread (each owns one HIP stream).  Frames,
 * references and depth maps belong to the context that created them.
 */
#ifndef LSDG_H
#define LSDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSDG_PYRAMID_LEVELS 5

typedef int32_t lsdg_status;
#define LSDG_OK 0
#define LSDG_DIVERGED 1
#define LSDG_EINVAL (-1)
#define LSDG_EDEVICE (-2)
#define LSDG_ENOMEM (-3)
#define LSDG_ESTATE (-4)

typedef struct lsdg_ctx_s* lsdg_ctx;
typedef struct lsdg_frame_s* lsdg_frame;
typedef struct lsdg_ref_s* lsdg_ref;
typedef struct lsdg_depthmap_s* lsdg_depthmap;
typedef struct lsdg_batch_s* lsdg_batch;

/* Pyramid storage precision of (I, gx, gy).  Residuals, weights and sums are
 * always computed in FP32; normal equations are merged and solved in FP64. */
#define LSDG_FP32 0
#define LSDG_FP16 1
#define LSDG_BF16 2

typedef struct lsdg_camera {
  int32_t w, h;
  float fx, fy, cx, cy;
} lsdg_camera;

/* Tuning constants of the LSD-SLAM core (settings.h, DenseDepthTrackerSettings
 * and the Conf() singleton the wrapper configures at tools/LSD.cpp:97-100).
 * lsdg_params_default() fills the upstream defaults. */
typedef struct lsdg_params {
  /* tracker (per pyramid level) */
  float lambda_initial[LSDG_PYRAMID_LEVELS];
  float lambda_success_fac[LSDG_PYRAMID_LEVELS];
  float lambda_fail_fac[LSDG_PYRAMID_LEVELS];
  float step_size_min[LSDG_PYRAMID_LEVELS];
  float convergence_eps[LSDG_PYRAMID_LEVELS];
  int32_t max_its_per_lvl[LSDG_PYRAMID_LEVELS];
  float var_weight;
  float huber_d;
  float camera_pixel_noise2;
  int32_t se3_min_level;          /* finest tracked level (1) */
  int32_t se3_max_level;          /* one past the coarsest tracked level (5) */
  float min_goodperall_pixel;
  float min_goodperall_pixel_absmin;
  float min_goodpergoodbad_pixel;
  float max_diff_constant;
  float max_diff_grad_mult;
  int32_t use_affine_lightning_estimation;
  int32_t use_esm_tracking;       /* Sim3: average frame and rotated reference gradients */
  /* depth map */
  float min_abs_grad_create;
  float min_abs_grad_decrease;
  float min_epl_length_squared;
  float min_epl_grad_squared;
  float min_epl_angle_squared;
  float gradient_sample_dist;
  float sample_point_to_border;
  float max_epl_length_crop;
  float min_epl_length_crop;
  float max_error_stereo;
  float min_distance_error_stereo;
  float stereo_epl_var_fac;
  float min_depth;
  float max_var;
  float var_random_init_initial;
  float var_gt_init_initial;
  float succ_var_inc_fac;
  float fail_var_inc_fac;
  float diff_fac_observe;
  float diff_fac_prop_merge;
  float diff_fac_smoothing;
  float depth_smoothing_factor;
  float division_eps;
  int32_t validity_counter_inc;
  int32_t validity_counter_dec;
  int32_t validity_counter_initial_observe;
  int32_t validity_counter_max;
  int32_t validity_counter_max_variable;
  int32_t val_sum_min_for_create;
  int32_t val_sum_min_for_keep;
  int32_t val_sum_min_for_unblacklist;
  int32_t min_blacklist;
  int32_t allow_negative_idepths;
  int32_t use_subpixel_stereo;
  int32_t stereo_subpixel_oob_guard; /* skip the sub-pixel step when the best match has no pre/post sample */
  int32_t stereo_max_steps;          /* safety cap on epipolar search steps (>= crop + 16) */
} lsdg_params;

/* Per-tracking statistics: the public result fields of SE3Tracker / Sim3Tracker. */
typedef struct lsdg_track_stats {
  float last_residual;          /* SE3: final lastErr; Sim3: mean */
  float last_depth_residual;    /* Sim3 meanD */
  float last_photometric_residual; /* Sim3 meanP */
  float last_mean_res;          /* SE3 lastMeanRes */
  float point_usage;
  float affine_a, affine_b;
  int32_t last_good_count;
  int32_t last_bad_count;
  int32_t diverged;
  int32_t tracking_was_good;
  int32_t num_residual_calls[LSDG_PYRAMID_LEVELS];
  int32_t num_warp_update_calls[LSDG_PYRAMID_LEVELS];
  double hessian[49];           /* Sim3: lastSim3Hessian (row-major 7x7) */
} lsdg_track_stats;

/* One fused evaluation of the tracker at a fixed pose and level: the sums of
 * calcResidualAndBuffers + calcWeightsAndResidual + calculateWarpUpdate
 * (SE3) or calcSim3Buffers + calcSim3WeightsAndResidual + calcSim3LGS (Sim3).
 * Normal equations are NOT divided by n. */
typedef struct lsdg_eval {
  double A[49];        /* SE3: top-left 6x6 of a 7-stride matrix; Sim3: full 7x7 */
  double b[7];
  double error;        /* sum W r^2 (SE3); Sim3: sumResP + sumResD */
  double sum_res_p, sum_res_d;
  double num_terms_p, num_terms_d;
  double n_constraints;
  int32_t warped_count;      /* buf_warped_size */
  int32_t good_count, bad_count;
  double sum_res_unweighted; /* sum over good points of r^2 */
  double sum_signed_res;     /* sum over good points of r */
  double usage;              /* sum min(1, z / z') */
  double sxx, syy, sx, sy, sw; /* affine brightness statistics */
} lsdg_eval;

/* Per-pixel depth hypotheses (DepthMapPixelHypothesis), structure of arrays.
 * Any NULL member is skipped on read and left unchanged on write. */
typedef struct lsdg_hypotheses {
  uint8_t* is_valid;
  int32_t* blacklisted;
  float* next_stereo_frame_min_id;
  int32_t* validity_counter;
  float* idepth;
  float* idepth_var;
  float* idepth_smoothed;
  float* idepth_var_smoothed;
  uint8_t* observe_code;  /* outcome of the last observeDepth per pixel (LSDG_OBS_*) */
} lsdg_hypotheses;

/* observeDepth outcome codes (one per pixel, from the last lsdg_depth_update). */
#define LSDG_OBS_NONE 0            /* outside [3,w-3)x[3,h-3), low gradient or blacklisted */
#define LSDG_OBS_LOWGRAD_INVALIDATE 1
#define LSDG_OBS_SKIP_ALREADY_GOOD 2
#define LSDG_OBS_SKIP_NOT_GOOD_TRACKED 3
#define LSDG_OBS_EPL_FAIL 4
#define LSDG_OBS_CREATE_VAR_TOO_BIG 5
#define LSDG_OBS_CREATED 6
#define LSDG_OBS_INCONSISTENT 7
#define LSDG_OBS_UPDATED 8
#define LSDG_OBS_STEREO_FAIL_BASE 10  /* 10 - code for doLineStereo codes -1..-4 (11..14) */

/* What lsdg_frame_read returns. */
#define LSDG_WHAT_IMAGE 0           /* float[w_l*h_l] */
#define LSDG_WHAT_GRADIENTS 1       /* float[4*w_l*h_l]: gx, gy, I, 0 (FP32 view) */
#define LSDG_WHAT_MAXGRAD 2         /* float[w_l*h_l] */
#define LSDG_WHAT_IDEPTH 3          /* float[w_l*h_l] */
#define LSDG_WHAT_IDEPTH_VAR 4      /* float[w_l*h_l] */
#define LSDG_WHAT_REF_PIXEL_WAS_GOOD 5 /* uint8[w_1*h_1] (level ignored) */

/* ---------------------------------------------------------------- context */

/* Fills the upstream defaults (settings.cpp / settings.h). */
void lsdg_params_default(lsdg_params* p);

/* One context per host thread: device, stream and the parameters that
 * settings.h / Conf() hold globally in the reference (tools/LSD.cpp:97-100). */
lsdg_status lsdg_ctx_create(int32_t device, const lsdg_params* params, lsdg_ctx* out);
lsdg_status lsdg_ctx_destroy(lsdg_ctx ctx);
lsdg_status lsdg_ctx_set_params(lsdg_ctx ctx, const lsdg_params* params);
lsdg_status lsdg_ctx_synchronize(lsdg_ctx ctx);

/* How the tracker evaluations sum their FP32 per-point terms.
 *   LSDG_SUMS_FP64  (default) each term is widened to FP64 and summed in FP64
 *                   per thread and through every tree: the CPU restatement's
 *                   sums to ~1e-16, so the GPU takes its LM path.
 *   LSDG_SUMS_KAHAN compensated FP32: Kahan per-thread sums, TwoSum-merged
 *                   wave and block trees, FP64 only in the cross-block merge
 *                   and the solve.  Sums within ~1e-7 relative of the FP64
 *                   ones, which can flip a borderline LM accept.
 * Applies to the SE3 and Sim3 evaluations (lsdg_se3_*, lsdg_sim3_*) and to
 * batch plans created afterwards; the adjoint kernels always sum in FP64.
 * No counterpart in the reference, whose LGS6/LGS7 accumulations (upstream
 * SE3Tracker::calculateWarpUpdate, Sim3Tracker::calcSim3LGS) sum in FP32 on
 * the CPU. */
#define LSDG_SUMS_FP64 0
#define LSDG_SUMS_KAHAN 1
lsdg_status lsdg_ctx_set_sums(lsdg_ctx ctx, int32_t mode);
lsdg_status lsdg_ctx_get_sums(lsdg_ctx ctx, int32_t* mode);
/* Last error message of this context (static storage, valid until next call). */
const char* lsdg_ctx_last_error(lsdg_ctx ctx);

/* ------------------------------------------------------------------ frames */

/* Frame(id, width, height, K, timestamp, image): uploads the level-0 image
 * (float, w*h) from host memory.  Image, gradient and max-gradient levels
 * (Frame::buildImage / buildGradients / buildMaxGradients) are built on the
 * device the first time a call reads them, per level and kind, as
 * Frame::require does: a tracker builds the gradients of the levels it
 * tracks, a depth map the keyframe's level-0 gradients and max-gradients.
 * The wrapper's input is an undistorted CV_8UC1 image
 * (lib/App/InputThread.cpp:59-65); lsdg_frame_create_u8 takes that form. */
lsdg_status lsdg_frame_create(lsdg_ctx ctx, int32_t id, const lsdg_camera* cam,
                              const float* image, int32_t precision, lsdg_frame* out);
lsdg_status lsdg_frame_create_u8(lsdg_ctx ctx, int32_t id, const lsdg_camera* cam,
                                 const uint8_t* image, int32_t stride, int32_t precision,
                                 lsdg_frame* out);
lsdg_status lsdg_frame_destroy(lsdg_frame f);
/* Rebuilds every pyramid level of every kind from the resident level-0 image. */
lsdg_status lsdg_frame_build_pyramid(lsdg_frame f);
/* Drops every pyramid level above the level-0 image; the next call that reads
 * one rebuilds it (what a new frame costs the tracker). */
lsdg_status lsdg_frame_invalidate(lsdg_frame f);
/* Frame::setDepth from plain arrays (idepth <= 0 or var <= 0: no depth), then
 * the inverse-variance depth pyramid (Frame::buildIDepthAndIDepthVar). */
lsdg_status lsdg_frame_set_depth(lsdg_frame f, const float* idepth, const float* idepth_var);
lsdg_status lsdg_frame_read(lsdg_frame f, int32_t what, int32_t level, void* out, size_t bytes);
/* Frame pose fields used by mapping: trackingParent, thisToParent_raw,
 * initialTrackedResidual and refPixelWasGood (NULL keeps the current mask;
 * a fresh frame's mask is all true).  The reference writes these at the end
 * of SE3Tracker::trackFrame; lsdg_se3_track does the same. */
lsdg_status lsdg_frame_set_tracking(lsdg_frame f, lsdg_frame parent, const double this_to_parent[8],
                                    float initial_tracked_residual, const uint8_t* ref_pixel_was_good);
lsdg_status lsdg_frame_get_tracking(lsdg_frame f, int32_t* parent_id, double this_to_parent[8],
                                    float* initial_tracked_residual);
/* Frame::getScaledCamToWorld, used by updateKeyframe when a reference frame was
 * tracked on another keyframe. */
lsdg_status lsdg_frame_set_cam_to_world(lsdg_frame f, const double cam_to_world[8]);
/* Keyframe counters numFramesTrackedOnThis / numMappedOnThis. */
lsdg_status lsdg_frame_counters(lsdg_frame f, int32_t* num_frames_tracked_on_this,
                                int32_t* num_mapped_on_this);
lsdg_status lsdg_frame_set_counters(lsdg_frame f, int32_t num_frames_tracked_on_this,
                                    int32_t num_mapped_on_this);

/* ------------------------------------------------------- tracking reference */

/* TrackingReference::importFrame(keyframe). */
lsdg_status lsdg_ref_create(lsdg_ctx ctx, lsdg_frame keyframe, lsdg_ref* out);
lsdg_status lsdg_ref_destroy(lsdg_ref r);
/* TrackingReference::makePointCloud(level): stream-compacts the keyframe's
 * pixels with depth into points (x, y, z, I, var, idx).  Idempotent. */
lsdg_status lsdg_ref_make_point_cloud(lsdg_ref r, int32_t level);
lsdg_status lsdg_ref_num_points(lsdg_ref r, int32_t level, int32_t* n);
/* Reads points as float[8] records: x, y, z, I, var, gx, gy, idx (as float). */
lsdg_status lsdg_ref_read_points(lsdg_ref r, int32_t level, float* out, int32_t max_points);

/* --------------------------------------------------------------- trackers */

/* SE3Tracker::trackFrame(TrackingReference*, Frame*, const SE3& frameToReference_initialEstimate).
 * Returns frameToReference.  Writes the frame's tracking fields (trackingParent,
 * thisToParent_raw, initialTrackedResidual, refPixelWasGood) as the reference does. */
lsdg_status lsdg_se3_track(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame,
                           const double frame_to_ref_init[7], double frame_to_ref_out[7],
                           lsdg_track_stats* stats);

/* Sim3Tracker::trackFrameSim3(TrackingReference*, Frame*, const Sim3&, int startLevel, int finalLevel).
 * The frame must carry depth (lsdg_frame_set_depth). */
lsdg_status lsdg_sim3_track(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame,
                            const double frame_to_ref_init[8], int32_t start_level,
                            int32_t final_level, double frame_to_ref_out[8],
                            lsdg_track_stats* stats);

/* One fused evaluation at referenceToFrame (SE3: double[7], scale ignored;
 * Sim3: double[8]).  is_good (optional, uint8[w_1*h_1]) receives the
 * refPixelWasGood writes of that evaluation (only when level == se3_min_level). */
lsdg_status lsdg_se3_evaluate(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame, int32_t level,
                              const double ref_to_frame[7], float affine_a, float affine_b,
                              lsdg_eval* out, uint8_t* is_good);
lsdg_status lsdg_sim3_evaluate(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame, int32_t level,
                               const double ref_to_frame[8], lsdg_eval* out);

/* Adjoint of the Huber-weighted photometric energy at referenceToFrame, with
 * the depth-variance weights w_p and the in-image mask frozen at that pose:
 *   E = sum_i rho(sqrt(w_p,i) r_i),  rho Huber at huber_d/2,
 *   grad_xi[6]     = dE/dxi for T <- exp(xi) T,
 *   grad_idepth[i] = dE/d(idepth of point i), point order of lsdg_ref_read_points
 *                    (0 for points outside the image).
 * No counterpart in the reference; it is the reduction behind LGS6::b. */
lsdg_status lsdg_energy_grad(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame, int32_t level,
                             const double ref_to_frame[7], double* energy, double grad_xi[6],
                             float* grad_idepth, int32_t max_points);

/* Sim3 adjoint at referenceToFrame (double[8]): the gradient of the Sim3
 * tracker's reweighted least-squares energy with its weights frozen at that
 * pose,
 *   E = 1/2 sum_i wh_i (w_p,i r_p,i^2 + w_d,i r_d,i^2),
 * r_p the photometric and r_d = 1/z' - idepth_frame the depth residual of
 * Sim3Tracker::calcSim3Buffers; wh the joint Huber weight at huber_d and
 * w_p / w_d the variance weights of calcSim3WeightsAndResidual; the weights,
 * the nearest-pixel frame depth sample and the in-image set are held fixed.
 *   grad_xi[7]     = dE/dxi (v, w, sigma) for T <- exp(xi) T,
 *   grad_idepth[i] = dE/d(idepth of point i) (0 outside the image).
 * grad_xi equals calcSim3LGS's b (LGS7 of LGS6 + LGS4, not divided by n)
 * with the frame's own gradient in place of the ESM average.  The frame must
 * carry depth.  No counterpart in the reference; it is the reduction behind
 * Sim3Tracker's normal equations. */
lsdg_status lsdg_energy_grad_sim3(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame, int32_t level,
                                  const double ref_to_frame[8], double* energy, double grad_xi[7],
                                  float* grad_idepth, int32_t max_points);

/* lsdg_energy_grad with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Its reduction rows, merged row and
 * per-point dE/dd staging live in a fixed device array of about 4.1 MB per
 * GPU (held until the process ends), so:
 *   - the reference's point cloud of `level` and the frame's pyramid level
 *     must already be built (lsdg_ref_make_point_cloud,
 *     lsdg_frame_build_pyramid), else LSDG_ESTATE;
 *   - at most 1048576 points (LSDG_EINVAL beyond);
 *   - every call on one GPU shares that GPU's array: calls that overlap on
 *     one GPU (several contexts or threads) race.
 * Results equal lsdg_energy_grad's bit for bit. */
#define LSDG_NOALLOC_MAX_POINTS 1048576
lsdg_status lsdg_energy_grad_noalloc(lsdg_ctx ctx, lsdg_ref ref, lsdg_frame frame, int32_t level,
                                     const double ref_to_frame[7], double* energy, double grad_xi[6],
                                     float* grad_idepth, int32_t max_points);

/* ------------------------------------------------------------- depth map */

lsdg_status lsdg_depthmap_create(lsdg_ctx ctx, const lsdg_camera* cam, lsdg_depthmap* out);
lsdg_status lsdg_depthmap_destroy(lsdg_depthmap dm);
/* DepthMap::reset / isValid. */
lsdg_status lsdg_depth_reset(lsdg_depthmap dm);
int32_t lsdg_depth_is_valid(lsdg_depthmap dm);
/* DepthMap::initializeFromGTDepth(Frame*): the keyframe's idepth becomes the
 * hypotheses (var VAR_GT_INIT_INITIAL, validity 20). */
lsdg_status lsdg_depth_initialize_from_gt(lsdg_depthmap dm, lsdg_frame keyframe);
/* DepthMap::initializeRandomly(Frame*): every interior pixel whose level-0
 * max-gradient exceeds MIN_ABS_GRAD_CREATE gets inverse depth
 * 0.5 + (draw % 100001) / 100000, variance VAR_RANDOM_INIT_INITIAL and
 * validity 20; the others lose their hypothesis.  The draws come from
 * SplitMix64 seeded with `seed` (output i + 1 for pixel i, top 31 bits)
 * instead of the reference's rand(), so a seed reproduces the map. */
lsdg_status lsdg_depth_initialize_randomly(lsdg_depthmap dm, lsdg_frame keyframe, uint64_t seed);
/* Prior hypotheses from arrays (idepth, var, validity) on the keyframe, then
 * setDepth.  Pixels with var <= 0 or non-finite idepth are invalid. */
lsdg_status lsdg_depth_set_prior(lsdg_depthmap dm, lsdg_frame keyframe, const float* idepth,
                                 const float* idepth_var, int32_t validity);
/* DepthMap::updateKeyframe(std::deque<std::shared_ptr<Frame>> referenceFrames),
 * frames ordered oldest first. */
lsdg_status lsdg_depth_update(lsdg_depthmap dm, const lsdg_frame* refs, int32_t n);
/* DepthMap::createKeyFrame(Frame* new_keyframe). */
lsdg_status lsdg_depth_create_keyframe(lsdg_depthmap dm, lsdg_frame new_keyframe);
/* DepthMap::finalizeKeyFrame(): fill holes, regularize, setDepth. */
lsdg_status lsdg_depth_finalize_keyframe(lsdg_depthmap dm);
lsdg_status lsdg_depth_read(lsdg_depthmap dm, lsdg_hypotheses* out);
lsdg_status lsdg_depth_write(lsdg_depthmap dm, const lsdg_hypotheses* in);
/* Keyframe the map currently belongs to (activeKeyFrame) and its id. */
lsdg_status lsdg_depth_active_keyframe(lsdg_depthmap dm, int32_t* id);

/* ------------------------------------------------------------ measurement */

/* Per-kernel device time, recorded with a hipEvent pair around every launch
 * on the context's stream, and the algorithmic HBM bytes each launch is
 * charged (tracker evaluation at level l: N_l * 21 + P_l * b_px, b_px = 12
 * for FP32 and 6 for FP16/BF16 pyramids; Sim3 adds N_l * 9 + P_l * 8).
 * No counterpart in the reference, whose core only keeps gettimeofday
 * averages (msTrackFrame, msObserve, ...).  Off by default. */
#define LSDG_KERNEL_EVAL 0        /* fused SE3 / Sim3 evaluation (plain path) */
#define LSDG_KERNEL_PYRAMID 1     /* frame pyramid build */
#define LSDG_KERNEL_DEPTH 2       /* depth-map kernels of one update */
#define LSDG_KERNEL_BATCH_EVAL 3  /* batched evaluation sweep */
#define LSDG_KERNEL_BATCH_SOLVE 4 /* batched LM step (merge, solve, accept) */
#define LSDG_KERNEL_BATCH_JOB 5   /* batched whole-job LM, one block per job */
#define LSDG_KERNEL_BATCH_COOP 6  /* batched whole LM, several cooperating blocks per job */
#define LSDG_KERNEL_INGEST 7      /* undistortion + conversion + image pyramid of one camera image */
#define LSDG_KERNEL_ADJOINT 8     /* SE3 / Sim3 adjoint (lsdg_energy_grad, lsdg_energy_grad_sim3) */
#define LSDG_KERNEL_KINDS 9

typedef struct lsdg_kernel_stats {
  int64_t launches;
  double ms;     /* summed device time */
  double bytes;  /* summed algorithmic bytes */
} lsdg_kernel_stats;

/* enable != 0 resets the statistics and starts recording. */
lsdg_status lsdg_ctx_profile(lsdg_ctx ctx, int32_t enable);
lsdg_status lsdg_ctx_kernel_stats(lsdg_ctx ctx, int32_t kernel, lsdg_kernel_stats* out);

/* ------------------------------------------------------- batched tracking */

#define LSDG_JOB_SE3 0
#define LSDG_JOB_SIM3 1

/* One frame-vs-keyframe alignment of a batch. */
typedef struct lsdg_job {
  lsdg_ref ref;
  lsdg_frame frame;
  double frame_to_ref_init[8];
  int32_t kind;         /* LSDG_JOB_SE3 / LSDG_JOB_SIM3 */
  int32_t start_level;  /* Sim3 only; SE3 uses se3_max_level-1 .. se3_min_level */
  int32_t final_level;
  int32_t rebuild_pyramid; /* rebuild the frame pyramid inside the batch */
} lsdg_job;

typedef struct lsdg_job_result {
  double frame_to_ref[8];
  lsdg_track_stats stats;
  lsdg_status status;
} lsdg_job_result;

/* Plans a batch: many alignments of uneven size run in one persistent launch
 * per LM step, the steps captured in a hipGraph.  The plan may be run many
 * times.  Frames may change content between runs, not shape: each run
 * rebuilds the pyramid levels a re-ingested frame's jobs track on (and resets
 * its refPixelWasGood mask); a keyframe whose depth changed gets new point
 * clouds; a Sim3 job's frame must carry depth when the plan runs
 * (LSDG_ESTATE otherwise). */
lsdg_status lsdg_batch_create(lsdg_ctx ctx, const lsdg_job* jobs, int32_t n, lsdg_batch* out);
lsdg_status lsdg_batch_run(lsdg_batch b, lsdg_job_result* results);
lsdg_status lsdg_batch_destroy(lsdg_batch b);
/* Convenience: create, run once, destroy. */
lsdg_status lsdg_batch_track(lsdg_ctx ctx, const lsdg_job* jobs, int32_t n, lsdg_job_result* results);

/* ------------------------------------------------------ keyframe export */

/* InputPointDense (lib/Pangolin_IOWrapper/Keyframe.h:16-21): one per pixel. */
typedef struct lsdg_point_dense {
  float idepth;
  float idepth_var;
  uint8_t color[4];
} lsdg_point_dense;

/* Keyframe::MyVertex (lib/Pangolin_IOWrapper/Keyframe.h:48-52). */
typedef struct lsdg_vertex {
  float point[3];
  uint8_t color[4];
} lsdg_vertex;

/* The constants of Keyframe::computeVbo (Keyframe.h:78-82) and the
 * keyframe's camToWorld scale; my_sparsifyFactor (rand()-based, 1 there) is
 * not supported. */
typedef struct lsdg_vbo_params {
  float scaled_th;           /* my_scaledTH = 1e-3 */
  float abs_th;              /* my_absTH = 1e-1 */
  float scale;               /* camToWorld.scale() */
  int32_t min_near_support;  /* my_minNearSupport = 9 */
} lsdg_vbo_params;

void lsdg_vbo_params_default(lsdg_vbo_params* p);

/* PangolinOutputIOWrapper::publishKeyframe (PangolinOutputIOWrapper.cpp:56-89):
 * the keyframe's level-`level` idepth, idepth variance and image as
 * w_l * h_l InputPointDense records in raster order, into host memory `out`.
 * The image value goes to all four color bytes, truncated (values the
 * reference's float-to-byte conversion leaves undefined are clamped to
 * 0..255, NaN to 0).  LSDG_ESTATE when the frame has no depth (the reference
 * then publishes an uninitialised buffer). */
lsdg_status lsdg_keyframe_points(lsdg_frame kf, int32_t level, lsdg_point_dense* out);

/* Keyframe::computeVbo (Keyframe.h:66-142) on InputPointDense records `in`
 * (host memory, w x h, raster order) of a camera (fx, fy, cx, cy): the
 * points that pass the variance and near-support filters, back-projected,
 * color bytes swizzled, alpha 100, in raster order, into host memory `out`
 * (capacity (w - 2) * (h - 2)); *count receives their number. */
lsdg_status lsdg_points_vbo(lsdg_ctx ctx, const lsdg_point_dense* in, int32_t w, int32_t h, float fx, float fy,
                            float cx, float cy, const lsdg_vbo_params* p, lsdg_vertex* out, int32_t* count);

/* Both on the device in one call: publishKeyframe at `level`, then
 * computeVbo with that level's camera. */
lsdg_status lsdg_keyframe_vbo(lsdg_frame kf, int32_t level, const lsdg_vbo_params* p, lsdg_vertex* out,
                              int32_t* count);

/* ----------------------------------------------------- constraint search */

/* One keyframe pair for lsd-slam's SlamSystem::testConstraint, the consumer
 * of keyframe-to-keyframe Sim3 alignment that feeds the pose graph whose
 * poses PangolinOutputIOWrapper::publishKeyframeGraph shows
 * (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:148-169): the new
 * keyframe's tracking reference (A), a candidate keyframe's (B), and the
 * initial candidate-to-keyframe Sim3 (qx, qy, qz, qw, tx, ty, tz, s).  Both
 * keyframes must carry depth. */
typedef struct lsdg_constraint_candidate {
  lsdg_ref frame_ref;
  lsdg_ref candidate_ref;
  double candidate_to_frame_init[8];
} lsdg_constraint_candidate;

/* KFConstraintStruct: one directed edge of an accepted pair. */
typedef struct lsdg_constraint_edge {
  int32_t first_id, second_id;  /* keyframe ids */
  double second_to_first[8];    /* the Sim3 alignment of `second` on `first` */
  double information[49];       /* its lastSim3Hessian (row-major 7x7) */
  float mean_residual, mean_residual_d, mean_residual_p, usage;
} lsdg_constraint_edge;

typedef struct lsdg_constraint_result {
  int32_t accepted;
  int32_t rejected_at;          /* 0 if accepted, else the stage (1: levels 4-3, 2: level 2, 3: level 1) */
  float error[3];               /* reciprocal consistency of each stage run; 1e20 when an alignment failed */
  double candidate_to_frame[8]; /* the last alignments of the two directions */
  double frame_to_candidate[8];
  lsdg_constraint_edge candidate_on_frame;  /* first = A, second = B (when accepted) */
  lsdg_constraint_edge frame_on_candidate;  /* first = B, second = A */
} lsdg_constraint_result;

/* testConstraint for n pairs: per stage (levels 4..3, 2, 1) both directions
 * of every pair still in the running are tracked in one batch
 * (trackFrameSim3 of B's keyframe on A and of A's keyframe on B, each from
 * its previous result), a direction that diverges or whose scale or
 * Hessian degenerates fails the pair's stage, and otherwise the reciprocal
 * consistency diff' (Sigma_AB + Adj_AB Sigma_BA Adj_AB')^-1 diff with
 * diff = log(AtoB * BtoA) is compared against 3000, 4000 and 6000 times
 * `strictness`.  The 7x7 algebra is double precision here (float upstream). */
lsdg_status lsdg_constraint_search(lsdg_ctx ctx, const lsdg_constraint_candidate* candidates, int32_t n,
                                   float strictness, lsdg_constraint_result* results);

/* The keyframe candidates of a constraint search (upstream's
 * TrackableKeyFrameSearch::findCandidates and findEuclideanOverlapFrames,
 * without FAB-MAP's appearance candidates): of the keyframes `others`, those
 * whose camera position lies near the keyframe's and whose optical axis
 * points the same way.  With t the camToWorld translation, s its scale, d
 * the mean inverse depth and z the optical axis (third column of the
 * rotation), a keyframe k is a candidate when
 *   |(t_k - t) * min(d_k / s_k, d / s)|^2 <= closeness_th * 15 / kf_dist_weight^2
 *   z_k . z >= cos((1 - 0.25 closeness_th) * (fov_x + fov_y) / 2),
 * fov = 2 atan(size / (2 f)) of `cam`.  Host computation, no device work.
 * out_index[i] indexes `others`; dist2 / cos_angle receive the two measures
 * (may be null); *count receives the number of candidates (at most cap are
 * written). */
typedef struct lsdg_keyframe_pose {
  int32_t id;
  float mean_idepth;        /* Frame::meanIdepth */
  double cam_to_world[8];   /* Sim3 (qx, qy, qz, qw, tx, ty, tz, s) */
} lsdg_keyframe_pose;

lsdg_status lsdg_keyframe_candidates(const lsdg_camera* cam, const lsdg_keyframe_pose* keyframe,
                                     const lsdg_keyframe_pose* others, int32_t n, float closeness_th,
                                     float kf_dist_weight, int32_t* out_index, double* dist2, double* cos_angle,
                                     int32_t cap, int32_t* count);

/* ----------------------------------------------------------- the system */

/* lsd-slam's SlamSystem front end, the direct caller of the path
 * (lib/App/InputThread.cpp:71 hands it every image; tools/LSD.cpp:102,139
 * builds it): per image trackFrame (SE3 tracking on the current keyframe,
 * loss detection, keyframe selection) followed, as with blockUntilMapped, by
 * the mapping iterations it waits for (updateKeyframe with the queued
 * frames, or finalizeKeyFrame + createKeyFrame on the latest tracked frame).
 * One thread, deterministic.  The pose graph and loop closures are not part
 * of it.  After a loss every image goes to the relocaliser (upstream's
 * Relocalizer, which SlamSystem::trackFrame feeds while tracking is lost):
 * one batched SE3 plan tracks the image against every keyframe from the
 * identity at reloc_level only (trackFrameOnPermaref's quick check); a
 * keyframe is a candidate when the check does not diverge and good / (good
 * + bad) of its last evaluation exceeds 1 - 0.75 (1 - MIN_GOODPERGOODBAD_PIXEL).
 * The candidate with the lowest residual is confirmed when at least half of
 * its neighbours in the keyframe chain (the keyframes before and after it)
 * also pass the check from the pose it implies; the lowest-residual one of
 * the candidate and its passing neighbours becomes the current keyframe (its
 * stored depth reloaded into the map, loadNewCurrentKeyframe), and the image
 * is then tracked on it as usual (LSDG_SYSTEM_RELOCALISED).  Upstream runs
 * the checks on several threads and takes the first confirmed candidate;
 * here the choice is deterministic. */
typedef struct lsdg_system_s* lsdg_system;

/* SlamSystem's keyframe-selection constants (lsd-slam settings) and the
 * relocaliser's. */
typedef struct lsdg_system_params {
  float kf_dist_weight;                /* KFDistWeight = 4 */
  float kf_usage_weight;               /* KFUsageWeight = 3 */
  int32_t min_num_mapped;              /* MIN_NUM_MAPPED = 5 */
  int32_t initialization_phase_count;  /* INITIALIZATION_PHASE_COUNT = 5 */
  int32_t relocalise;                  /* 1: relocalise after a loss (default); 0: stay lost */
  int32_t reloc_level;                 /* QUICK_KF_CHECK_LVL = 4: the level of the candidate checks */
} lsdg_system_params;
void lsdg_system_params_default(lsdg_system_params* p);

#define LSDG_SYSTEM_TRACKED 0
#define LSDG_SYSTEM_LOST 1        /* tracking lost on this image */
#define LSDG_SYSTEM_NOT_TRACKED 2 /* tracking was lost before and this image did not relocalise */
#define LSDG_SYSTEM_RELOCALISED 3 /* tracking was lost before; this image relocalised and was tracked */

typedef struct lsdg_system_frame {
  int32_t id;
  int32_t status;             /* LSDG_SYSTEM_* */
  int32_t keyframe_id;        /* the keyframe it was tracked on; -1 when not tracked */
  int32_t new_keyframe;       /* 1 when the mapping made it the new keyframe */
  float closeness_score;      /* lastTrackingClosenessScore; -1 when not evaluated */
  float keyframe_mean_idepth; /* the keyframe's meanIdepth the score used */
  double frame_to_keyframe[7];
  double cam_to_world[8];
  lsdg_track_stats stats;
} lsdg_system_frame;

lsdg_status lsdg_system_create(lsdg_ctx ctx, const lsdg_camera* cam, int32_t precision,
                               const lsdg_system_params* params, lsdg_system* out);
lsdg_status lsdg_system_destroy(lsdg_system s);
/* SlamSystem::gtDepthInit: the first image (8-bit, row stride `stride`) with
 * its inverse depth and variance (idepth <= 0: none) becomes the first
 * keyframe, its depth map initialised from them (initializeFromGTDepth). */
lsdg_status lsdg_system_init_gt(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride,
                                const float* idepth, const float* idepth_var);
/* SlamSystem::randomInit: the first image becomes the first keyframe with a
 * random depth map (lsdg_depth_initialize_randomly with `seed`). */
lsdg_status lsdg_system_init_random(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride, uint64_t seed);
/* SlamSystem::trackFrame(image, id, blockUntilMapped = true). */
lsdg_status lsdg_system_track(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride,
                              lsdg_system_frame* out);
/* The keyframes so far, oldest first: ids and camToWorld (8 doubles each);
 * *count receives their number (at most cap are written). */
lsdg_status lsdg_system_keyframes(lsdg_system s, int32_t* ids, double* cam_to_world, int32_t cap, int32_t* count);
/* The current keyframe's depth map hypotheses (lsdg_depth_read). */
lsdg_status lsdg_system_depth_read(lsdg_system s, lsdg_hypotheses* out);
/* SlamSystem::findConstraintsForNewKeyframes for the system's keyframe
 * `index` (0 = the first): lsdg_keyframe_candidates over the other
 * keyframes (meanIdepth of each from its level-0 depth), minus its
 * neighbours (its tracking parent and the keyframes tracked on it), then
 * lsdg_constraint_search of every candidate against it (frame = the
 * keyframe, initial candidate-to-frame from the two camToWorld) at
 * `strictness`.  Writes up to cap results with their candidates' keyframe
 * ids; *count receives the number of candidates. */
lsdg_status lsdg_system_find_constraints(lsdg_system s, int32_t index, float closeness_th, float strictness,
                                         lsdg_constraint_result* results, int32_t* candidate_ids, int32_t cap,
                                         int32_t* count);

/* ----------------------------------------------------------------- ingest */

/* libvideoio's Undistorter, which tools/LSD.cpp:88 builds from the
 * calibration file and InputThread applies to every camera image
 * (lib/App/InputThread.cpp:59-62): a table of one source position (x, y) in
 * the distorted input per output pixel, x < 0 meaning no source (output 0).
 * map_xy holds 2 * out_w * out_h floats in raster order; every entry with
 * x >= 0 must satisfy 0 <= x < in_w - 1 and 0 <= y < in_h - 1 (LSDG_EINVAL
 * otherwise), which keeps both bilinear taps inside the input.  The table is
 * copied to device memory; calls that use one undistorter must not overlap. */
typedef struct lsdg_undistorter_s* lsdg_undistorter;
lsdg_status lsdg_undistorter_create(lsdg_ctx ctx, int32_t in_w, int32_t in_h, int32_t out_w, int32_t out_h,
                                    const float* map_xy, lsdg_undistorter* out);
lsdg_status lsdg_undistorter_destroy(lsdg_undistorter u);

/* The undistortion table of the FOV (ATAN) lens model of the calibration
 * files LSD-SLAM ships (fx fy cx cy omega, Devernay and Faugeras): an output
 * pixel (x, y) of the pinhole camera out_k = (fx, fy, cx, cy) at normalised
 * radius r takes its value from the input camera in_k at radius
 * atan(2 r tan(omega / 2)) / omega (r itself when omega or r is 0).  Sources
 * outside 0 < x < in_w - 1, 0 < y < in_h - 1 are marked -1.  Host-only. */
lsdg_status lsdg_fov_undistort_table(const float in_k[4], float omega, int32_t in_w, int32_t in_h,
                                     const float out_k[4], int32_t out_w, int32_t out_h, float* map_xy);

/* One camera image through InputThread and Frame's constructor: the
 * distorted 8-bit image (host memory, in_w x in_h, row stride `stride`) is
 * undistorted (Undistorter::undistort: bilinear interpolation of the 8-bit
 * input, truncated to 8 bits), converted to float and reduced to the image
 * pyramid (Frame::buildImage, levels 0-4) in one launch, into a new frame of
 * the undistorter's output size (cam->w and cam->h must equal it).  When
 * `undistorted` is not NULL it receives the 8-bit undistorted image
 * (out_w * out_h, the live image InputThread.cpp:80 hands the output). */
lsdg_status lsdg_frame_create_undistorted(lsdg_ctx ctx, int32_t id, const lsdg_camera* cam, lsdg_undistorter u,
                                          const uint8_t* image, int32_t stride, int32_t precision,
                                          uint8_t* undistorted, lsdg_frame* out);
/* The same into an existing frame of the undistorter's output size, for a
 * stream of camera images without an allocation per image: every field of f
 * is reset as in a new frame with this id (no depth, identity poses, no
 * tracking parent, zero counters), its arrays and camera are kept, and point
 * clouds built from its earlier content are rebuilt when next used.  No
 * counterpart upstream, which allocates a Frame per image. */
lsdg_status lsdg_frame_ingest(lsdg_frame f, int32_t id, lsdg_undistorter u, const uint8_t* image, int32_t stride,
                              uint8_t* undistorted);

/* TextOutputIOWrapper::publishTrackedFrame's pose.txt line
 * (lib/Pangolin_IOWrapper/TextOutputIOWrapper.cpp:100-120): the frame id, the
 * camToWorld translation and the thisToParent_raw translation, comma
 * separated, each number as an ostream prints a double by default (six
 * significant digits), ending in a newline; NUL-terminated into buf.
 * LSDG_EINVAL when it does not fit in cap bytes. */
lsdg_status lsdg_frame_pose_txt(lsdg_frame f, char* buf, size_t cap);
/* The same line from an id and two poses (qx, qy, qz, qw, tx, ty, tz, s):
 * for frames the caller does not hold as lsdg_frame, such as the system's
 * (lsdg_system_frame: cam_to_world, and frame_to_keyframe with s = 1). */
lsdg_status lsdg_pose_txt_line(int32_t id, const double cam_to_world[8], const double this_to_parent[8], char* buf,
                               size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* LSDG_H */
