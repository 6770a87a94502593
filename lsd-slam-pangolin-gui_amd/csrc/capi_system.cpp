// C-ABI of lsd-slam's SlamSystem front end (the direct caller of the path:
// lib/App/InputThread.cpp:71 hands it every image, tools/LSD.cpp:102,139
// builds it): trackFrame's tracking, loss test and keyframe selection, then
// the mapping iterations trackFrame(blockUntilMapped) waits for
// (doMappingIteration: updateKeyframe with the queued frames, or
// finishCurrentKeyframe + createNewCurrentKeyframe), and after a loss the
// relocaliser (upstream's Relocalizer, fed by trackFrame while tracking is
// lost) over one batched SE3 plan per image.  Single-threaded and
// deterministic; the pose graph and loop closures are not part of it.
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
  std::vector<lsdg_ref> relocRefs;  // the relocaliser's tracking references, one per keyframe
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
  p->relocalise = 1;
  p->reloc_level = 4;
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
  for (lsdg_ref r : s->relocRefs)
    if (r) lsdg_ref_destroy(r);
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

// Frame::meanIdepth of frame k: the float mean, in raster order, of the
// level-0 idepth setDepth wrote (pixels without depth hold -1).
static lsdg_status frame_mean_idepth(lsdg_system s, lsdg_frame k, float* out) {
  const size_t n = (size_t)s->cam.w * s->cam.h;
  std::vector<float> d(n);
  lsdg_status st = lsdg_frame_read(k, LSDG_WHAT_IDEPTH, 0, d.data(), n * sizeof(float));
  if (st != LSDG_OK) return st;
  float sum = 0.0f;
  int num = 0;
  for (size_t i = 0; i < n; ++i)
    if (d[i] != -1.0f) {
      sum += d[i];
      ++num;
    }
  *out = sum / (float)num;
  return LSDG_OK;
}

// meanIdepth of the current keyframe, cached per depth version.
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

// ------------------------------------------------------------ relocaliser
namespace {

// Outcome of one quick check of an image against a keyframe.
struct RelocCheck {
  Pose toKf;        // frameToKeyframe found
  bool good;        // not diverged and good / (good + bad) above the threshold
  float residual;   // lastResidual
};

}  // namespace

// Quick checks of frame f against keyframes kfs[i] from init[i] (frameToKf)
// at P.reloc_level only, all in one batch plan: trackFrameOnPermaref's
// single-level SE3 alignment.  The context's tracked levels are set to that
// level while the plan is built and run.
static lsdg_status reloc_checks(lsdg_system s, lsdg_frame f, const std::vector<int>& kfs,
                                const std::vector<Pose>& init, std::vector<RelocCheck>& out) {
  lsdg_ctx c = s->ctx;
  const int n = (int)kfs.size();
  out.assign(n, RelocCheck{pose_identity(), false, 0.0f});
  if (n == 0) return LSDG_OK;
  if (s->relocRefs.size() < s->keyframes.size()) s->relocRefs.resize(s->keyframes.size(), nullptr);
  std::vector<lsdg_job> jobs(n);
  for (int i = 0; i < n; ++i) {
    lsdg_ref& r = s->relocRefs[kfs[i]];
    if (!r) {
      lsdg_status st = lsdg_ref_create(c, s->keyframes[kfs[i]].get(), &r);
      if (st != LSDG_OK) return st;
    }
    lsdg_job& J = jobs[i];
    std::memset(&J, 0, sizeof(J));
    J.ref = r;
    J.frame = f;
    Pose p = init[i];
    p.s = 1.0;
    pose_to8(p, J.frame_to_ref_init);
    J.kind = LSDG_JOB_SE3;
    J.start_level = s->P.reloc_level;
    J.final_level = s->P.reloc_level;
    J.rebuild_pyramid = 0;
  }
  const lsdg_params saved = c->P;
  lsdg_params q = saved;
  q.se3_max_level = s->P.reloc_level + 1;
  q.se3_min_level = s->P.reloc_level;
  lsdg_status st = lsdg_ctx_set_params(c, &q);
  if (st != LSDG_OK) return st;
  std::vector<lsdg_job_result> res(n);
  lsdg_batch b = nullptr;
  st = lsdg_batch_create(c, jobs.data(), n, &b);
  if (st == LSDG_OK) st = lsdg_batch_run(b, res.data());
  if (b) lsdg_batch_destroy(b);
  lsdg_ctx_set_params(c, &saved);
  if (st != LSDG_OK) return st;
  const double thr = 1.0 - 0.75 * (1.0 - (double)saved.min_goodpergoodbad_pixel);
  for (int i = 0; i < n; ++i) {
    const lsdg_track_stats& t = res[i].stats;
    const int cnt = t.last_good_count + t.last_bad_count;
    const double ratio = cnt > 0 ? (double)t.last_good_count / (double)cnt : 0.0;
    out[i].toKf = pose_from8(res[i].frame_to_ref);
    out[i].toKf.s = 1.0;
    out[i].good = !t.diverged && ratio > thr;
    out[i].residual = t.last_residual;
  }
  return LSDG_OK;
}

// The relocaliser on image f (see lsdg.h): *found = keyframe index or -1,
// *toKf its frameToKeyframe.
static lsdg_status relocalise(lsdg_system s, lsdg_frame f, int* found, Pose* toKf) {
  *found = -1;
  const int nk = (int)s->keyframes.size();
  std::vector<int> all(nk);
  std::vector<Pose> ident(nk, pose_identity());
  for (int i = 0; i < nk; ++i) all[i] = i;
  std::vector<RelocCheck> r;
  lsdg_status st = reloc_checks(s, f, all, ident, r);
  if (st != LSDG_OK) return st;
  int c = -1;
  for (int i = 0; i < nk; ++i)
    if (r[i].good && (c < 0 || r[i].residual < r[c].residual)) c = i;
  if (c < 0) return LSDG_OK;
  // the candidate's chain neighbours, from the camToWorld the candidate implies
  const Pose c2wF = pose_mul(s->keyframes[c]->camToWorld, r[c].toKf);
  std::vector<int> nb;
  std::vector<Pose> nbInit;
  for (int k : {c - 1, c + 1})
    if (k >= 0 && k < nk) {
      nb.push_back(k);
      nbInit.push_back(pose_mul(pose_inv(s->keyframes[k]->camToWorld), c2wF));
    }
  std::vector<RelocCheck> rn;
  st = reloc_checks(s, f, nb, nbInit, rn);
  if (st != LSDG_OK) return st;
  int best = c, good = 0, bad = 0;
  float bestRes = r[c].residual;
  Pose bestPose = r[c].toKf;
  for (size_t i = 0; i < nb.size(); ++i) {
    if (!rn[i].good) {
      ++bad;
      continue;
    }
    ++good;
    if (rn[i].residual < bestRes) {
      best = nb[i];
      bestRes = rn[i].residual;
      bestPose = rn[i].toKf;
    }
  }
  if (good < bad) return LSDG_OK;
  *found = best;
  *toKf = bestPose;
  return LSDG_OK;
}

// loadNewCurrentKeyframe: keyframe k becomes the current keyframe, its stored
// depth the map's hypotheses again.
static lsdg_status load_keyframe(lsdg_system s, int k) {
  FramePtr kf = s->keyframes[k];
  const size_t n = (size_t)s->cam.w * s->cam.h;
  std::vector<float> idepth(n), var(n);
  lsdg_status st = lsdg_frame_read(kf.get(), LSDG_WHAT_IDEPTH, 0, idepth.data(), n * sizeof(float));
  if (st == LSDG_OK) st = lsdg_frame_read(kf.get(), LSDG_WHAT_IDEPTH_VAR, 0, var.data(), n * sizeof(float));
  if (st == LSDG_OK) st = lsdg_depth_set_prior(s->dm, kf.get(), idepth.data(), var.data(), 20);
  if (st != LSDG_OK) return st;
  s->kf = kf;
  s->unmapped.clear();
  s->createNewKeyFrame = false;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_system_track(lsdg_system s, int32_t id, const uint8_t* image, int32_t stride,
                                         lsdg_system_frame* out) {
  if (!s || !image || !out) return LSDG_EINVAL;
  if (!s->initialised) return fail(s->ctx, LSDG_ESTATE, "system not initialised");
  std::memset(out, 0, sizeof(*out));
  out->id = id;
  out->keyframe_id = -1;
  out->closeness_score = -1.0f;
  if (!s->trackingIsGood && !s->P.relocalise) {
    out->status = LSDG_SYSTEM_NOT_TRACKED;
    return LSDG_OK;
  }
  lsdg_frame f;
  lsdg_status st = lsdg_frame_create_u8(s->ctx, id, &s->cam, image, stride, s->precision, &f);
  if (st != LSDG_OK) return st;
  FramePtr fp = own(f);
  bool relocalised = false;
  if (!s->trackingIsGood) {
    int k = -1;
    Pose toKf = pose_identity();
    st = relocalise(s, f, &k, &toKf);
    if (st != LSDG_OK) return st;
    if (k < 0) {
      out->status = LSDG_SYSTEM_NOT_TRACKED;
      return LSDG_OK;
    }
    st = load_keyframe(s, k);
    if (st != LSDG_OK) return st;
    s->lastCamToWorld = pose_mul(s->kf->camToWorld, toKf);
    s->trackingIsGood = true;
    relocalised = true;
  }
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
  if (relocalised) out->status = LSDG_SYSTEM_RELOCALISED;
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

// ------------------------------------------------------ constraint search
extern "C" lsdg_status lsdg_system_find_constraints(lsdg_system s, int32_t index, float closeness_th,
                                                    float strictness, lsdg_constraint_result* results,
                                                    int32_t* candidate_ids, int32_t cap, int32_t* count) {
  if (!s || !count || cap < 0 || (cap > 0 && (!results || !candidate_ids))) return LSDG_EINVAL;
  const int nk = (int)s->keyframes.size();
  if (index < 0 || index >= nk) return fail(s->ctx, LSDG_EINVAL, "keyframe index out of range");
  std::vector<lsdg_keyframe_pose> poses(nk);
  for (int i = 0; i < nk; ++i) {
    lsdg_frame k = s->keyframes[i].get();
    poses[i].id = k->id;
    lsdg_status st = frame_mean_idepth(s, k, &poses[i].mean_idepth);
    if (st != LSDG_OK) return st;
    pose_to8(k->camToWorld, poses[i].cam_to_world);
  }
  // the keyframe's neighbours: its tracking parent and the keyframes tracked on it
  const lsdg_frame self = s->keyframes[index].get();
  std::vector<int> other;
  std::vector<lsdg_keyframe_pose> otherPoses;
  for (int i = 0; i < nk; ++i) {
    const lsdg_frame k = s->keyframes[i].get();
    if (i == index || k == self->parent || k->parent == self) continue;
    other.push_back(i);
    otherPoses.push_back(poses[i]);
  }
  const int m = (int)other.size();
  std::vector<int32_t> idx(std::max(m, 1));
  int32_t nc = 0;
  lsdg_status st = lsdg_keyframe_candidates(&s->cam, &poses[index], otherPoses.data(), m, closeness_th,
                                            s->P.kf_dist_weight, idx.data(), nullptr, nullptr, m, &nc);
  if (st != LSDG_OK) return st;
  *count = nc;
  if (nc == 0) return LSDG_OK;
  if (s->relocRefs.size() < s->keyframes.size()) s->relocRefs.resize(s->keyframes.size(), nullptr);
  auto ref_of = [&](int i, lsdg_ref* r) -> lsdg_status {
    if (!s->relocRefs[i]) {
      lsdg_status e = lsdg_ref_create(s->ctx, s->keyframes[i].get(), &s->relocRefs[i]);
      if (e != LSDG_OK) return e;
    }
    *r = s->relocRefs[i];
    return LSDG_OK;
  };
  std::vector<lsdg_constraint_candidate> cands(nc);
  const Pose frameInv = pose_inv(s->keyframes[index]->camToWorld);
  for (int c = 0; c < nc; ++c) {
    const int i = other[idx[c]];
    lsdg_constraint_candidate& C = cands[c];
    std::memset(&C, 0, sizeof(C));
    if ((st = ref_of(index, &C.frame_ref)) != LSDG_OK || (st = ref_of(i, &C.candidate_ref)) != LSDG_OK) return st;
    pose_to8(pose_mul(frameInv, s->keyframes[i]->camToWorld), C.candidate_to_frame_init);
  }
  std::vector<lsdg_constraint_result> res(nc);
  st = lsdg_constraint_search(s->ctx, cands.data(), nc, strictness, res.data());
  if (st != LSDG_OK) return st;
  for (int c = 0; c < nc && c < cap; ++c) {
    results[c] = res[c];
    candidate_ids[c] = s->keyframes[other[idx[c]]]->id;
  }
  return LSDG_OK;
}
