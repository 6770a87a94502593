// Helpers shared by the C-ABI translation units.
#pragma once
#include "lsdg_internal.h"

namespace lsdg {

// Records `what` (and the HIP error) as the context's last error; returns st.
lsdg_status fail(lsdg_ctx c, lsdg_status st, const char* what, hipError_t e = hipSuccess);
lsdg_status frame_depth_from_level0(lsdg_frame f);
// Marks a frame's depth pyramid as written (every level) by a kernel already
// queued, e.g. launch_fill_regularize with the frame as setDepth.
void frame_depth_written(lsdg_frame f);
// A frame of camera cam with every array allocated and nothing valid yet.
lsdg_status frame_alloc(lsdg_ctx c, int32_t id, const lsdg_camera* cam, int32_t precision, lsdg_frame* out);
// Builds the pyramid data in `need` (fd_bit flags) that the frame does not
// hold yet, on the context stream (Frame::require).
lsdg_status frame_require(lsdg_frame f, unsigned need);
// The (I, gx, gy) image a tracker samples at levels lo..hi: FP32 float4 or
// the packed reduced-precision copy.
unsigned fd_track(lsdg_frame f, int lo, int hi);
// Algorithmic bytes of a pyramid build (image stage if `images`, gradient
// stage for the fd_bit data in `grads`).
double pyramid_bytes(lsdg_frame f, bool images, unsigned grads);
lsdg_status frame_ensure_was_good(lsdg_frame f);
lsdg_status ref_points(lsdg_ref r, int level);
Xf pose_xf(const Pose& T, bool scaled);
TrackKParams track_kparams(const lsdg_params& P, bool sim3);
void esm_roll(const Pose& T, float* r4);
double eval_bytes(int n, lsdg_frame f, int lvl, bool sim3);
double adjoint_bytes(int n, lsdg_frame f, int lvl, bool sim3);

// Kernel timing: begin() records the start event (if profiling is on) and
// returns a handle for end(), which records the stop event and the bytes.
struct ProfMark {
  int kind = -1;
  hipEvent_t a = nullptr;
};
ProfMark prof_begin(lsdg_ctx c, int kind);
void prof_end(lsdg_ctx c, ProfMark& m, double bytes);
// Synchronises the stream and folds finished event pairs into the totals.
lsdg_status prof_resolve(lsdg_ctx c);
// A kernel's own start / stop events (for launch_k) when profiling is on,
// null events otherwise; prof_timer_done files the pair under `kind`.
KTimer prof_timer(lsdg_ctx c);
void prof_timer_done(lsdg_ctx c, int kind, const KTimer& t, double bytes);

}  // namespace lsdg

#define CK(ctx, expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) return ::lsdg::fail((ctx), LSDG_EDEVICE, #expr, e_); \
  } while (0)
