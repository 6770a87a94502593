// C-ABI of lsd-slam's SlamSystem front end (the direct caller of the path:
// lib/App/InputThread.cpp:71 hands it every image, tools/LSD.cpp:102,139
// builds it): trackFrame's tracking, loss test and keyframe selection, then
// the mapping iterations trackFrame(blockUntilMapped) waits for
// (doMappingIteration: updateKeyframe with the queued frames, or
// finishCurrentKeyframe + createNewCurrentKeyframe).  Single-threaded and
// deterministic; the pose graph, loop closures and relocalisation are not
// part of it.
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "capi_common.h"

using namespace lsdg;

namespace {

using FramePtr = std::shared_ptr<lsdg_frame_s>;

FramePtr own(lsdg_frame f) {
  return FramePtr(f, [](lsdg_frame x) { lsdg_frame_destroy(x); });
}

// se3FromSim3: the rotation and translation of a Sim3, scale dropped.
void se3_of(const Pose& p, double* out7) {
  Pose q = p;
  q.s = 1.0;
  pose_to7(q, out7);
}

}  // namespace

struct lsdg_system_s {
  lsdg_ctx ctx = nullptr;
  lsdg_camera cam{};
  int precision = LSDG_FP32;
  lsdg_system_params P{};
  lsdg_depthmap dm = nullptr;
  lsdg_ref ref = nullptr;        // trackingReference, on the current keyframe
  FramePtr kf;                    // currentKeyFrame
  std::vector<FramePtr> keyframes;  // keyFrameGraph->keyframesAll
  std::deque<FramePtr> unmapped;    // unmappedTrackedFrames
  FramePtr latest;                  // latestTrackedFrame
  Pose lastCamToWorld = pose_identity();  // keyFrameGraph->allFramePoses.back()
  bool trackingIsGood = true;
  bool createNewKeyFrame = false;
  bool initialised = false;
  // meanIdepth of the current keyframe for its depth version
  lsdg_frame meanFor = nullptr;
  int meanVersion = -1;
  float meanIdepth = 0.0f;
  std::vector<float> scratch;
};

extern "C" void lsdg_system_params_default(lsdg_system_params* p) {
  if (!p) return;
  p->kf_dist_weight = 4.0f;
  p->kf_usage_weight = 3.0f;
  p->min_num_mapped = 5;
  p->initialization_phase_count = 5;
}

extern "C" lsdg_status lsdg_system_create(lsdg_ctx c, const lsdg_camera* cam, int32_t precision,
                                          const lsdg_system_params* params, lsdg_system* out) {
  if (!c || !cam || !out) return LSDG_EINVAL;
  lsdg_system s = new (std::nothrow) lsdg_system_s();
  if (!s) return LSDG_ENOMEM;
  s->ctx = c;
  s->cam = *cam;
  s->precision = precision;
  if (params) s->P = *params;
  else lsdg_system_params_default(&s->P);
  lsdg_status st = lsdg_depthmap_create(c, cam, &s->dm);
  if (st != LSDG_OK) {
    delete s;
    return st;
  }
  *out = s;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_system_destroy(lsdg_system s) {
  if (!s) return LSDG_EINVAL;
  if (s->ref) lsdg_ref_destroy(s->ref);
  if (s->dm) lsdg_depthmap_destroy(s->dm);
  delete s;  // releases the frames
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_system_init_gt(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride,
                                           const float* idepth, const float* var) {
  if (!s || !image || !idepth || !var) return LSDG_EINVAL;
  if (s->initialised) return fail(s->ctx, LSDG_ESTATE, "system already initialised");
  lsdg_frame f;
  lsdg_status st = lsdg_frame_create_u8(s->ctx, id, &s->cam, image, stride, s->precision, &f);
  if (st != LSDG_OK) return st;
  FramePtr fp = own(f);
  st = lsdg_frame_set_depth(f, idepth, var);
  if (st == LSDG_OK) st = lsdg_depth_initialize_from_gt(s->dm, f);
  if (st != LSDG_OK) return st;
  f->camToWorld = pose_identity();
  s->kf = fp;
  s->keyframes.push_back(fp);
  s->lastCamToWorld = f->camToWorld;
  s->initialised = true;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_system_init_random(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride,
                                               uint64_t seed) {
  if (!s || !image) return LSDG_EINVAL;
  if (s->initialised) return fail(s->ctx, LSDG_ESTATE, "system already initialised");
  lsdg_frame f;
  lsdg_status st = lsdg_frame_create_u8(s->ctx, id, &s->cam, image, stride, s->precision, &f);
  if (st != LSDG_OK) return st;
  FramePtr fp = own(f);
  st = lsdg_depth_initialize_randomly(s->dm, f, seed);
  if (st != LSDG_OK) return st;
  f->camToWorld = pose_identity();
  s->kf = fp;
  s->keyframes.push_back(fp);
  s->lastCamToWorld = f->camToWorld;
  s->initialised = true;
  return LSDG_OK;
}

// Frame::meanIdepth of the keyframe: the float mean, in raster order, of the
// level-0 idepth setDepth wrote (pixels without depth hold -1).
static lsdg_status mean_idepth(lsdg_system s, float* out) {
  lsdg_frame k = s->kf.get();
  if (s->meanFor != k || s->meanVersion != k->depthVersion) {
    const size_t n = (size_t)s->cam.w * s->cam.h;
    s->scratch.resize(n);
    lsdg_status st = lsdg_frame_read(k, LSDG_WHAT_IDEPTH, 0, s->scratch.data(), n * sizeof(float));
    if (st != LSDG_OK) return st;
    float sum = 0.0f;
    int num = 0;
    for (size_t i = 0; i < n; ++i)
      if (s->scratch[i] != -1.0f) {
        sum += s->scratch[i];
        ++num;
      }
    s->meanIdepth = sum / (float)num;
    s->meanFor = k;
    s->meanVersion = k->depthVersion;
  }
  *out = s->meanIdepth;
  return LSDG_OK;
}

// doMappingIteration while it has work (what trackFrame(blockUntilMapped) waits for).
static lsdg_status map_until_idle(lsdg_system s, lsdg_system_frame* info) {
  for (;;) {
    if (s->createNewKeyFrame) {
      // finishCurrentKeyframe, then createNewCurrentKeyframe(latestTrackedFrame)
      lsdg_status st = lsdg_depth_finalize_keyframe(s->dm);
      if (st != LSDG_OK) return st;
      FramePtr nk = s->latest;
      st = lsdg_depth_create_keyframe(s->dm, nk.get());
      if (st != LSDG_OK) return st;
      // camToWorld through the parent, with the rescaled thisToParent
      nk->camToWorld = pose_mul(s->kf->camToWorld, nk->thisToParent);
      // the latest tracked frame (allFramePoses.back()) is now the keyframe
      s->lastCamToWorld = nk->camToWorld;
      s->kf = nk;
      s->keyframes.push_back(nk);
      s->createNewKeyFrame = false;
      if (info && info->id == nk->id) info->new_keyframe = 1;
      continue;
    }
    // updateKeyframe: drop frames tracked on another keyframe, then map with the whole queue
    while (!s->unmapped.empty() && s->unmapped.front()->parent != s->kf.get()) s->unmapped.pop_front();
    if (s->unmapped.empty()) return LSDG_OK;
    std::vector<lsdg_frame> refs;
    for (auto& f : s->unmapped) refs.push_back(f.get());
    lsdg_status st = lsdg_depth_update(s->dm, refs.data(), (int32_t)refs.size());
    if (st != LSDG_OK) return st;
    s->unmapped.pop_front();
  }
}

extern "C" lsdg_status lsdg_system_track(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride,
                                         lsdg_system_frame* out) {
  if (!s || !image || !out) return LSDG_EINVAL;
  if (!s->initialised) return fail(s->ctx, LSDG_ESTATE, "system not initialised");
  std::memset(out, 0, sizeof(*out));
  out->id = id;
  out->closeness_score = -1.0f;
  if (!s->trackingIsGood) {
    out->status = LSDG_SYSTEM_NOT_TRACKED;
    return LSDG_OK;
  }
  lsdg_frame f;
  lsdg_status st = lsdg_frame_create_u8(s->ctx, id, &s->cam, image, stride, s->precision, &f);
  if (st != LSDG_OK) return st;
  FramePtr fp = own(f);
  const bool myCreate = s->createNewKeyFrame;
  // trackingReference->importFrame(currentKeyFrame) when the keyframe changed
  // (a depth change is picked up by the reference itself)
  if (!s->ref || s->ref->kf != s->kf.get()) {
    if (s->ref) lsdg_ref_destroy(s->ref);
    s->ref = nullptr;
    st = lsdg_ref_create(s->ctx, s->kf.get(), &s->ref);
    if (st != LSDG_OK) return st;
  }
  double init[7], pose7[7];
  se3_of(pose_mul(pose_inv(s->kf->camToWorld), s->lastCamToWorld), init);
  st = lsdg_se3_track(s->ctx, s->ref, f, init, pose7, &out->stats);
  if (st != LSDG_OK && st != LSDG_DIVERGED) return st;
  out->keyframe_id = s->kf->id;
  std::memcpy(out->frame_to_keyframe, pose7, sizeof(pose7));
  // keyFrameGraph->keyframesAll: the finished keyframes (finishCurrentKeyframe
  // adds a keyframe when the next one replaces it; SLAM enabled, as by default)
  const int nKF = (int)s->keyframes.size() - 1;
  if (out->stats.diverged || (nKF > s->P.initialization_phase_count && !out->stats.tracking_was_good)) {
    s->trackingIsGood = false;
    out->status = LSDG_SYSTEM_LOST;
    return LSDG_OK;
  }
  // keyFrameGraph->addFrame: camToWorld through the tracking parent
  Pose toKf = pose_from7(pose7);
  f->camToWorld = pose_mul(s->kf->camToWorld, toKf);
  s->lastCamToWorld = f->camToWorld;
  s->latest = fp;
  // keyframe selection
  if (!myCreate && s->kf->numMappedOnThis > s->P.min_num_mapped) {
    float mean;
    st = mean_idepth(s, &mean);
    if (st != LSDG_OK) return st;
    double dist[3];
    for (int i = 0; i < 3; ++i) dist[i] = toKf.t[i] * mean;
    float minVal = std::fmin(0.2f + nKF * 0.8f / s->P.initialization_phase_count, 1.0f);
    if (nKF < s->P.initialization_phase_count) minVal *= 0.7f;
    const float usage = out->stats.point_usage;
    // TrackableKeyFrameSearch::getRefFrameScore
    const float score = (float)(dist[0] * dist[0] + dist[1] * dist[1] + dist[2] * dist[2]) * s->P.kf_dist_weight *
                            s->P.kf_dist_weight +
                        (1.0f - usage) * (1.0f - usage) * s->P.kf_usage_weight * s->P.kf_usage_weight;
    out->closeness_score = score;
    out->keyframe_mean_idepth = mean;
    if (score > minVal) s->createNewKeyFrame = true;
  }
  s->unmapped.push_back(fp);
  st = map_until_idle(s, out);
  if (st != LSDG_OK) return st;
  pose_to8(f->camToWorld, out->cam_to_world);
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_system_keyframes(lsdg_system s, int32_t* ids, double* c2w, int32_t cap, int32_t* count) {
  if (!s || !count || cap < 0) return LSDG_EINVAL;
  const int n = (int)s->keyframes.size();
  for (int i = 0; i < n && i < cap; ++i) {
    if (ids) ids[i] = s->keyframes[i]->id;
    if (c2w) pose_to8(s->keyframes[i]->camToWorld, c2w + 8 * i);
  }
  *count = n;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_system_depth_read(lsdg_system s, lsdg_hypotheses* out) {
  if (!s || !out) return LSDG_EINVAL;
  return lsdg_depth_read(s->dm, out);
}
