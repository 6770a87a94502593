// lsd_slam::SE3Tracker / Sim3Tracker / DepthMap / SlamSystem on top of liblsdg.so.
//
// The upstream classes (lsd-slam's Tracking/SE3Tracker.h, Tracking/
// Sim3Tracker.h, DepthEstimation/DepthMap.h, linked by the reference as the
// `lsdslam` dependency: fips.yml:2-4, lib/CMakeLists.txt:16) keep their method
// names, argument meaning and public result fields here, and forward to the
// C ABI in include/lsdg.h.  Upstream's Frame / TrackingReference / Sophus
// types are not available in this repository, so the shim carries small
// stand-ins with the members the trackers and the depth map use
// (`lsd_slam::Frame`, `TrackingReference`, `SE3`, `Sim3`); an integration into
// a real lsd-slam build maps its Frame onto an `lsdg_frame` once per frame
// (Frame(id, w, h, K, timestamp, image) -> lsdg_frame_create_u8) and its
// Sophus poses onto the 7 / 8-double layout below.
//
// Errors: like upstream, nothing throws.  A diverged track sets `diverged`,
// clears `trackingWasGood` and returns the identity; device failures set
// `diverged` too and leave the message in lastError().
#pragma once

#include <array>
#include <cstdint>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lsdg.h"

namespace lsd_slam {

// Rigid / similarity transforms as (qx, qy, qz, qw, tx, ty, tz[, s]),
// X -> s R X + t; the layout Sophus::Sim3f::data() gives the reference's GUI
// (lib/GUI.cpp:148, lib/GUI/Keyframe.h:23-27) plus the scale.
struct SE3 {
  std::array<double, 7> v{0, 0, 0, 1, 0, 0, 0};
};
struct Sim3 {
  std::array<double, 8> v{0, 0, 0, 1, 0, 0, 0, 1};
  double scale() const { return v[7]; }
};

// One GPU context (stream, parameters) per host thread, as upstream keeps
// one tracker per SlamSystem thread.
class Context {
 public:
  explicit Context(int device = 0) {
    lsdg_params p;
    lsdg_params_default(&p);
    ok_ = lsdg_ctx_create(device, &p, &ctx_) == LSDG_OK;
  }
  ~Context() {
    if (ctx_) lsdg_ctx_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  bool ok() const { return ok_; }
  lsdg_ctx get() const { return ctx_; }
  std::string lastError() const { return ctx_ ? lsdg_ctx_last_error(ctx_) : "no context"; }
  // Settings (settings.h / DenseDepthTrackerSettings), same names as upstream.
  bool setParams(const lsdg_params& p) { return lsdg_ctx_set_params(ctx_, &p) == LSDG_OK; }

 private:
  lsdg_ctx ctx_ = nullptr;
  bool ok_ = false;
};

// Frame(id, width, height, K, timestamp, image) with the pose fields mapping
// reads and writes (trackingParent, thisToParent_raw, initialTrackedResidual,
// refPixelWasGood); pyramids are built on the device on first use.
class Frame {
 public:
  typedef std::shared_ptr<Frame> SharedPtr;

  Frame(Context& ctx, int id, int width, int height, float fx, float fy, float cx, float cy, double timestamp,
        const unsigned char* image, int32_t precision = LSDG_FP32)
      : id_(id), timestamp_(timestamp) {
    cam_ = lsdg_camera{width, height, fx, fy, cx, cy};
    if (lsdg_frame_create_u8(ctx.get(), id, &cam_, image, width, precision, &f_) != LSDG_OK) f_ = nullptr;
  }
  Frame(Context& ctx, int id, int width, int height, float fx, float fy, float cx, float cy, double timestamp,
        const float* image, int32_t precision = LSDG_FP32)
      : id_(id), timestamp_(timestamp) {
    cam_ = lsdg_camera{width, height, fx, fy, cx, cy};
    if (lsdg_frame_create(ctx.get(), id, &cam_, image, precision, &f_) != LSDG_OK) f_ = nullptr;
  }
  ~Frame() {
    if (f_) lsdg_frame_destroy(f_);
  }
  Frame(const Frame&) = delete;
  Frame& operator=(const Frame&) = delete;

  int id() const { return id_; }
  int width(int level = 0) const { return cam_.w >> level; }
  int height(int level = 0) const { return cam_.h >> level; }
  double timestamp() const { return timestamp_; }
  bool valid() const { return f_ != nullptr; }
  lsdg_frame handle() const { return f_; }

  // Frame::setDepth from level-0 arrays (idepth <= 0 invalid).
  bool setDepth(const float* idepth, const float* idepthVar) {
    return lsdg_frame_set_depth(f_, idepth, idepthVar) == LSDG_OK;
  }
  // image(level), gradients(level) (4 floats per pixel), maxGradients(level),
  // idepth(level), idepthVar(level) copied to host memory.
  std::vector<float> read(int32_t what, int level) const {
    const size_t n = (size_t)width(level) * height(level) * (what == LSDG_WHAT_GRADIENTS ? 4 : 1);
    std::vector<float> out(n);
    if (lsdg_frame_read(f_, what, level, out.data(), n * sizeof(float)) != LSDG_OK) out.clear();
    return out;
  }
  // pose->thisToParent_raw and the tracking parent's id.
  Sim3 thisToParent() const {
    Sim3 s;
    int32_t pid = -1;
    float itr = 0.f;
    lsdg_frame_get_tracking(f_, &pid, s.v.data(), &itr);
    return s;
  }
  void setCamToWorld(const Sim3& camToWorld) { lsdg_frame_set_cam_to_world(f_, camToWorld.v.data()); }

 private:
  lsdg_frame f_ = nullptr;
  lsdg_camera cam_{};
  int id_ = 0;
  double timestamp_ = 0;
};

// TrackingReference::importFrame(keyframe); point clouds per level on the
// device, rebuilt when the keyframe's depth changes.
class TrackingReference {
 public:
  explicit TrackingReference(Context& ctx) : ctx_(ctx.get()) {}
  ~TrackingReference() { release(); }
  TrackingReference(const TrackingReference&) = delete;
  TrackingReference& operator=(const TrackingReference&) = delete;
  void importFrame(Frame* keyframe) {
    release();
    keyframe_ = keyframe;
    if (keyframe && lsdg_ref_create(ctx_, keyframe->handle(), &r_) != LSDG_OK) r_ = nullptr;
  }
  void invalidate() { release(); }
  void makePointCloud(int level) { lsdg_ref_make_point_cloud(r_, level); }
  int numData(int level) const {
    int32_t n = 0;
    lsdg_ref_num_points(r_, level, &n);
    return n;
  }
  Frame* keyframe() const { return keyframe_; }
  lsdg_ref handle() const { return r_; }

 private:
  void release() {
    if (r_) lsdg_ref_destroy(r_);
    r_ = nullptr;
    keyframe_ = nullptr;
  }
  lsdg_ctx ctx_;
  lsdg_ref r_ = nullptr;
  Frame* keyframe_ = nullptr;
};

// SE3Tracker: trackFrame and its public result fields.
class SE3Tracker {
 public:
  SE3Tracker(Context& ctx) : ctx_(ctx) {}

  float lastResidual = 0, lastMeanRes = 0, pointUsage = 0;
  int lastGoodCount = 0, lastBadCount = 0;
  float affineEstimation_a = 1, affineEstimation_b = 0;
  bool diverged = false, trackingWasGood = true;
  std::array<int, LSDG_PYRAMID_LEVELS> numResidualCalls{}, numWarpUpdateCalls{};

  SE3 trackFrame(TrackingReference* reference, Frame* frame, const SE3& frameToReference_initialEstimate) {
    lsdg_track_stats st{};
    SE3 out;
    const lsdg_status s = lsdg_se3_track(ctx_.get(), reference->handle(), frame->handle(),
                                         frameToReference_initialEstimate.v.data(), out.v.data(), &st);
    take(st, s);
    if (s != LSDG_OK) return SE3();
    return out;
  }
  SE3 trackFrame(TrackingReference* reference, const Frame::SharedPtr& frame,
                 const SE3& frameToReference_initialEstimate) {
    return trackFrame(reference, frame.get(), frameToReference_initialEstimate);
  }

 private:
  void take(const lsdg_track_stats& st, lsdg_status s) {
    lastResidual = st.last_residual;
    lastMeanRes = st.last_mean_res;
    pointUsage = st.point_usage;
    lastGoodCount = st.last_good_count;
    lastBadCount = st.last_bad_count;
    affineEstimation_a = st.affine_a;
    affineEstimation_b = st.affine_b;
    diverged = s != LSDG_OK || st.diverged != 0;
    trackingWasGood = s == LSDG_OK && st.tracking_was_good != 0;
    for (int l = 0; l < LSDG_PYRAMID_LEVELS; ++l) {
      numResidualCalls[l] = st.num_residual_calls[l];
      numWarpUpdateCalls[l] = st.num_warp_update_calls[l];
    }
  }
  Context& ctx_;
};

// Sim3Tracker: trackFrameSim3 and its public result fields.
class Sim3Tracker {
 public:
  Sim3Tracker(Context& ctx) : ctx_(ctx) {}

  float lastResidual = 0, lastDepthResidual = 0, lastPhotometricResidual = 0, pointUsage = 0;
  bool diverged = false;
  std::array<double, 49> lastSim3Hessian{};  // row-major 7x7

  Sim3 trackFrameSim3(TrackingReference* reference, Frame* frame, const Sim3& frameToReference_initialEstimate,
                      int startLevel, int finalLevel) {
    lsdg_track_stats st{};
    Sim3 out;
    const lsdg_status s =
        lsdg_sim3_track(ctx_.get(), reference->handle(), frame->handle(), frameToReference_initialEstimate.v.data(),
                        startLevel, finalLevel, out.v.data(), &st);
    lastResidual = st.last_residual;
    lastDepthResidual = st.last_depth_residual;
    lastPhotometricResidual = st.last_photometric_residual;
    pointUsage = st.point_usage;
    diverged = s != LSDG_OK || st.diverged != 0;
    for (int k = 0; k < 49; ++k) lastSim3Hessian[k] = st.hessian[k];
    if (s != LSDG_OK) return Sim3();
    return out;
  }
  Sim3 trackFrameSim3(TrackingReference* reference, const Frame::SharedPtr& frame,
                      const Sim3& frameToReference_initialEstimate, int startLevel, int finalLevel) {
    return trackFrameSim3(reference, frame.get(), frameToReference_initialEstimate, startLevel, finalLevel);
  }

 private:
  Context& ctx_;
};

// DepthMap: the semi-dense inverse-depth filter of the active keyframe.
class DepthMap {
 public:
  DepthMap(Context& ctx, int w, int h, float fx, float fy, float cx, float cy) {
    lsdg_camera cam{w, h, fx, fy, cx, cy};
    if (lsdg_depthmap_create(ctx.get(), &cam, &dm_) != LSDG_OK) dm_ = nullptr;
  }
  ~DepthMap() {
    if (dm_) lsdg_depthmap_destroy(dm_);
  }
  DepthMap(const DepthMap&) = delete;
  DepthMap& operator=(const DepthMap&) = delete;

  void reset() { lsdg_depth_reset(dm_); }
  bool isValid() const { return lsdg_depth_is_valid(dm_) != 0; }
  void initializeFromGTDepth(Frame* new_frame) { lsdg_depth_initialize_from_gt(dm_, new_frame->handle()); }
  void initializeFromGTDepth(const Frame::SharedPtr& new_frame) { initializeFromGTDepth(new_frame.get()); }
  // referenceFrames oldest first, as upstream's deque.
  void updateKeyframe(const std::deque<std::shared_ptr<Frame>>& referenceFrames) {
    std::vector<lsdg_frame> refs;
    refs.reserve(referenceFrames.size());
    for (const auto& f : referenceFrames) refs.push_back(f->handle());
    if (!refs.empty()) lsdg_depth_update(dm_, refs.data(), (int32_t)refs.size());
  }
  void createKeyFrame(Frame* new_keyframe) { lsdg_depth_create_keyframe(dm_, new_keyframe->handle()); }
  void createKeyFrame(const Frame::SharedPtr& new_keyframe) { createKeyFrame(new_keyframe.get()); }
  void finalizeKeyFrame() { lsdg_depth_finalize_keyframe(dm_); }
  lsdg_depthmap handle() const { return dm_; }

 private:
  lsdg_depthmap dm_ = nullptr;
};

// SlamSystem's front end (trackFrame with blockUntilMapped, gtDepthInit) on
// one thread; no pose graph, loop closure or relocalisation (see
// lsdg_system_track).  Images are 8-bit, row stride `stride` bytes.
class SlamSystem {
 public:
  SlamSystem(Context& ctx, int w, int h, float fx, float fy, float cx, float cy, int32_t precision = LSDG_FP32)
      : w_(w), h_(h) {
    lsdg_camera cam{w, h, fx, fy, cx, cy};
    if (lsdg_system_create(ctx.get(), &cam, precision, nullptr, &s_) != LSDG_OK) s_ = nullptr;
  }
  ~SlamSystem() {
    if (s_) lsdg_system_destroy(s_);
  }
  SlamSystem(const SlamSystem&) = delete;
  SlamSystem& operator=(const SlamSystem&) = delete;

  // gtDepthInit(image, depth, timeStamp, id): depth in metres (<= 0: none);
  // the frame's variance is VAR_GT_INIT_INITIAL, as setDepthFromGroundTruth sets it.
  bool gtDepthInit(const uint8_t* image, int stride, const float* depth, int id) {
    lsdg_params p;
    lsdg_params_default(&p);
    std::vector<float> idepth((size_t)w_ * h_), var((size_t)w_ * h_);
    for (size_t i = 0; i < idepth.size(); ++i) {
      const bool ok = depth[i] > 0.0f && depth[i] == depth[i];
      idepth[i] = ok ? 1.0f / depth[i] : -1.0f;
      var[i] = ok ? p.var_gt_init_initial : -1.0f;
    }
    return lsdg_system_init_gt(s_, id, image, stride, idepth.data(), var.data()) == LSDG_OK;
  }
  // trackFrame(image, frameID, blockUntilMapped = true, timestamp); the
  // per-frame outcome (pose, keyframe decision, tracker statistics) in `out`.
  bool trackFrame(const uint8_t* image, int stride, int frameID, lsdg_system_frame* out = nullptr) {
    lsdg_system_frame f;
    const bool ok = lsdg_system_track(s_, frameID, image, stride, &f) == LSDG_OK;
    if (ok) trackingIsGood_ = f.status == LSDG_SYSTEM_TRACKED;
    if (out) *out = f;
    return ok;
  }
  bool trackingIsGood() const { return trackingIsGood_; }
  lsdg_system handle() const { return s_; }

 private:
  int w_, h_;
  lsdg_system s_ = nullptr;
  bool trackingIsGood_ = true;
};

}  // namespace lsd_slam
