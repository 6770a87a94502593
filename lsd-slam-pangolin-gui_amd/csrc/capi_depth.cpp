// C-ABI of the depth map: DepthMap::updateKeyframe / createKeyFrame /
// finalizeKeyFrame / initializeFromGTDepth and state access.  The per-pixel
// work runs in depthmap.hip; this file holds the host side the reference keeps
// in DepthMap.cpp: reference-frame bookkeeping (oldest / newest frame,
// referenceFrameByID), Frame::prepareForStereoWith, the mean-idepth rescale
// and the new keyframe's pose.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "capi_common.h"

using namespace lsdg;

// Keyframe data the depth map reads: level-0 gradients (epipolar checks) and
// max-gradients (observe / fill-holes gates).
static constexpr unsigned KF_DATA = fd_bit(FD_GRAD, 0) | fd_bit(FD_MAXG, 0);

static void free_dm(lsdg_depthmap dm) {
  if (dm->d_pool) (void)hipFree(dm->d_pool);
  if (dm->d_refs) (void)hipFree(dm->d_refs);
  if (dm->d_refById) (void)hipFree(dm->d_refById);
  if (dm->h_refs) (void)hipHostFree(dm->h_refs);
  if (dm->h_refById) (void)hipHostFree(dm->h_refById);
  if (dm->h_red) (void)hipHostFree(dm->h_red);
  delete dm;
}

extern "C" lsdg_status lsdg_depthmap_create(lsdg_ctx c, const lsdg_camera* cam, lsdg_depthmap* out) {
  if (!c || !cam || !out) return LSDG_EINVAL;
  if (cam->w < 16 || cam->h < 16 || cam->w > 16384 || cam->h > 16384)
    return fail(c, LSDG_EINVAL, "depth map size out of range");
  if ((cam->w & 1) || (cam->h & 1))  // the keyframes it maps are even-sized (frame_alloc)
    return fail(c, LSDG_EINVAL, "depth map width and height must be even");
  lsdg_depthmap dm = new (std::nothrow) lsdg_depthmap_s();
  if (!dm) return LSDG_ENOMEM;
  dm->ctx = c;
  DepthGeom& g = dm->g;
  g.w = cam->w;
  g.h = cam->h;
  g.fx = cam->fx;
  g.fy = cam->fy;
  g.cx = cam->cx;
  g.cy = cam->cy;
  g.fxi = 1.0f / g.fx;
  g.fyi = 1.0f / g.fy;
  g.cxi = -g.cx / g.fx;
  g.cyi = -g.cy / g.fy;
  const size_t n = (size_t)g.w * g.h;
  const size_t nb = (n + 255) / 256;
  size_t total = 0;
  auto take = [&](size_t bytes) {
    size_t o = total;
    total += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t ob[3][5];
  for (int k = 0; k < 3; ++k) {
    ob[k][0] = take(n);
    ob[k][1] = take(n * 4);
    ob[k][2] = take(n * 4);
    ob[k][3] = take(n * 4);
    ob[k][4] = take(n * 16);
  }
  size_t o_obs = take(n), o_dest = take(n * 4), o_iv = take(n * 8), o_cnt = take((n + 1) * 4),
         o_fill = take(n * 4), o_bucket = take(n * 4), o_blk = take((nb + 1) * 4), o_red = take(2 * 1024 * 8),
         o_rc = take(64), o_work = take(obs_work_slots(g.w, g.h) * 16),
         o_wc = take(OBS_COUNT_BYTES);
  hipError_t e = hipMalloc(&dm->d_pool, total);
  if (e != hipSuccess) {
    free_dm(dm);
    return fail(c, LSDG_ENOMEM, "hipMalloc(depth map)", e);
  }
  char* b = (char*)dm->d_pool;
  for (int k = 0; k < 3; ++k) {
    dm->buf[k].valid = (uint8_t*)(b + ob[k][0]);
    dm->buf[k].bl = (int*)(b + ob[k][1]);
    dm->buf[k].nsid = (float*)(b + ob[k][2]);
    dm->buf[k].val = (int*)(b + ob[k][3]);
    dm->buf[k].dv = (float4*)(b + ob[k][4]);
  }
  dm->d_obs = (uint8_t*)(b + o_obs);
  dm->d_dest = (int*)(b + o_dest);
  dm->d_newIV = (float2*)(b + o_iv);
  dm->d_count = (int*)(b + o_cnt);
  dm->d_fill = (int*)(b + o_fill);
  dm->d_bucket = (int*)(b + o_bucket);
  dm->d_scanblk = (int*)(b + o_blk);
  dm->d_red = (double*)(b + o_red);
  dm->d_red_counter = (unsigned*)(b + o_rc);
  dm->d_work = (int4*)(b + o_work);
  dm->d_work_count = (int*)(b + o_wc);
  if ((e = hipMemsetAsync(dm->d_pool, 0, total, c->stream)) != hipSuccess ||
      (e = hipHostMalloc(&dm->h_red, 64, hipHostMallocMapped)) != hipSuccess ||
      (e = hipHostGetDevicePointer((void**)&dm->d_red_out, dm->h_red, 0)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
    free_dm(dm);
    return fail(c, LSDG_EDEVICE, "depth map setup", e);
  }
  *out = dm;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depthmap_destroy(lsdg_depthmap dm) {
  if (!dm) return LSDG_EINVAL;
  (void)hipStreamSynchronize(dm->ctx->stream);
  free_dm(dm);
  return LSDG_OK;
}

// DepthMap::reset: every hypothesis invalid, no active keyframe.
extern "C" lsdg_status lsdg_depth_reset(lsdg_depthmap dm) {
  if (!dm) return LSDG_EINVAL;
  lsdg_ctx c = dm->ctx;
  size_t n = (size_t)dm->g.w * dm->g.h;
  for (int k = 0; k < 3; ++k) CK(c, hipMemsetAsync(dm->buf[k].valid, 0, n, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  dm->kf = nullptr;
  dm->valid = false;
  return LSDG_OK;
}

extern "C" int32_t lsdg_depth_is_valid(lsdg_depthmap dm) { return (dm && dm->valid && dm->kf) ? 1 : 0; }

static lsdg_status check_frame(lsdg_depthmap dm, lsdg_frame f) {
  if (!f) return LSDG_EINVAL;
  if (f->ctx != dm->ctx) return fail(dm->ctx, LSDG_EINVAL, "frame belongs to another context");
  if (f->w[0] != dm->g.w || f->h[0] != dm->g.h) return fail(dm->ctx, LSDG_EINVAL, "frame / depth map size mismatch");
  return LSDG_OK;
}

// Frame::setDepth of the active keyframe from the current hypotheses, then
// its depth pyramid.
static lsdg_status set_depth(lsdg_depthmap dm, lsdg_frame f) {
  lsdg_ctx c = dm->ctx;
  CK(c, launch_set_depth(dm, f, c->stream));
  return frame_depth_from_level0(f);
}

extern "C" lsdg_status lsdg_depth_initialize_from_gt(lsdg_depthmap dm, lsdg_frame kf) {
  if (!dm) return LSDG_EINVAL;
  lsdg_status st = check_frame(dm, kf);
  if (st != LSDG_OK) return st;
  if (!kf->hasDepth) return fail(dm->ctx, LSDG_ESTATE, "initializeFromGTDepth needs a frame with depth");
  lsdg_ctx c = dm->ctx;
  // kf's level-0 idepth is about to be rewritten by setDepth: use a copy
  size_t n = (size_t)dm->g.w * dm->g.h;
  float* tmp = (float*)dm->d_newIV;
  CK(c, hipMemcpyAsync(tmp, kf->d_idepth[0], n * 4, hipMemcpyDeviceToDevice, c->stream));
  CK(c, launch_hyp_prior(dm, tmp, nullptr, 20, 1, c->P.var_gt_init_initial, c->stream));
  dm->kf = kf;
  dm->kfReactivated = false;
  dm->valid = true;
  st = set_depth(dm, kf);
  if (st != LSDG_OK) return st;
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_initialize_randomly(lsdg_depthmap dm, lsdg_frame kf, uint64_t seed) {
  if (!dm) return LSDG_EINVAL;
  lsdg_status st = check_frame(dm, kf);
  if (st != LSDG_OK) return st;
  lsdg_ctx c = dm->ctx;
  st = frame_require(kf, fd_bit(FD_MAXG, 0));
  if (st != LSDG_OK) return st;
  dm->kf = kf;
  dm->kfReactivated = false;
  dm->valid = true;
  CK(c, launch_random_init(dm, c->P, (unsigned long long)seed, c->stream));
  st = set_depth(dm, kf);
  if (st != LSDG_OK) return st;
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_set_prior(lsdg_depthmap dm, lsdg_frame kf, const float* idepth, const float* var,
                                            int32_t validity) {
  if (!dm || !idepth || !var) return LSDG_EINVAL;
  lsdg_status st = check_frame(dm, kf);
  if (st != LSDG_OK) return st;
  lsdg_ctx c = dm->ctx;
  size_t n = (size_t)dm->g.w * dm->g.h;
  float* d_id = (float*)dm->d_newIV;  // scratch: 2n floats
  float* d_var = d_id + n;
  CK(c, hipMemcpyAsync(d_id, idepth, n * 4, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(d_var, var, n * 4, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_hyp_prior(dm, d_id, d_var, validity, 0, 0.f, c->stream));
  dm->kf = kf;
  dm->kfReactivated = false;
  dm->valid = true;
  st = set_depth(dm, kf);
  if (st != LSDG_OK) return st;
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

// Frame::prepareForStereoWith(keyframe, thisToOther = refToKf, K, 0): FP32
// products cached per reference frame.
static void prepare_stereo(const DepthGeom& g, const Pose& thisToOther, StereoRef& r) {
  Pose o2t = pose_inv(thisToOther);
  double Ro[9], Rt[9];
  quat_R(o2t.q, Ro);
  quat_R(thisToOther.q, Rt);
  const float K[9] = {g.fx, 0.f, g.cx, 0.f, g.fy, g.cy, 0.f, 0.f, 1.f};
  float M[9];
  for (int i = 0; i < 9; ++i) M[i] = (float)(Ro[i] * o2t.s);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.KR[3 * i + j] = K[3 * i] * M[j] + K[3 * i + 1] * M[3 + j] + K[3 * i + 2] * M[6 + j];
  for (int i = 0; i < 3; ++i) r.t_o2t[i] = (float)o2t.t[i];
  for (int i = 0; i < 3; ++i)
    r.Kt[i] = K[3 * i] * r.t_o2t[0] + K[3 * i + 1] * r.t_o2t[1] + K[3 * i + 2] * r.t_o2t[2];
  for (int i = 0; i < 3; ++i) r.t_t2o[i] = (float)thisToOther.t[i];
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i) r.Rrow[3 * k + i] = (float)(Rt[3 * i + k] * thisToOther.s);
}

static lsdg_status ensure_ref_storage(lsdg_depthmap dm, int nrefs, int nbyid) {
  lsdg_ctx c = dm->ctx;
  if (nrefs > dm->refs_cap) {
    if (dm->d_refs) (void)hipFree(dm->d_refs);
    if (dm->h_refs) (void)hipHostFree(dm->h_refs);
    dm->d_refs = nullptr;
    dm->h_refs = nullptr;
    dm->refs_cap = 0;
    CK(c, hipMalloc(&dm->d_refs, sizeof(StereoRef) * nrefs));
    CK(c, hipHostMalloc(&dm->h_refs, sizeof(StereoRef) * nrefs, hipHostMallocDefault));
    dm->refs_cap = nrefs;
  }
  if (nbyid > dm->refById_cap) {
    if (dm->d_refById) (void)hipFree(dm->d_refById);
    if (dm->h_refById) (void)hipHostFree(dm->h_refById);
    dm->d_refById = nullptr;
    dm->h_refById = nullptr;
    dm->refById_cap = 0;
    CK(c, hipMalloc(&dm->d_refById, sizeof(int) * nbyid));
    CK(c, hipHostMalloc(&dm->h_refById, sizeof(int) * nbyid, hipHostMallocDefault));
    dm->refById_cap = nbyid;
  }
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_update(lsdg_depthmap dm, const lsdg_frame* refs, int32_t n) {
  if (!dm || !refs || n < 1) return LSDG_EINVAL;
  lsdg_ctx c = dm->ctx;
  if (!dm->valid || !dm->kf) return fail(c, LSDG_ESTATE, "depth map has no active keyframe");
  for (int i = 0; i < n; ++i) {
    lsdg_status st = check_frame(dm, refs[i]);
    if (st != LSDG_OK) return st;
  }
  lsdg_frame kf = dm->kf;
  // referenceFrameByID: entry k covers frame id offset + k; ids are increasing
  const int offset = refs[0]->id;
  std::vector<int> byId;
  for (int i = 0; i < n; ++i) {
    if (i > 0 && refs[i]->id < refs[i - 1]->id) return fail(c, LSDG_EINVAL, "reference frames must be oldest first");
    if ((long long)refs[i]->id - offset + 1 > (1 << 24)) return fail(c, LSDG_EINVAL, "reference frame ids too sparse");
    while ((int)byId.size() + offset <= refs[i]->id) byId.push_back(i);
  }
  lsdg_status st = ensure_ref_storage(dm, n, (int)byId.size() + 1);
  if (st != LSDG_OK) return st;
  CK(c, hipStreamSynchronize(c->stream));  // staging buffers are free
  for (int i = 0; i < n; ++i) {
    lsdg_frame f = refs[i];
    Pose refToKf = (f->parentId == kf->id) ? f->thisToParent : pose_mul(pose_inv(kf->camToWorld), f->camToWorld);
    StereoRef& r = dm->h_refs[i];
    prepare_stereo(dm->g, refToKf, r);
    r.img = f->d_image[0];
    r.wasGood = (f->parentId == kf->id && f->wasGoodAllocated) ? f->d_wasGood : nullptr;
    r.initialTrackedResidual = f->initialTrackedResidual;
    r.id = f->id;
  }
  for (size_t k = 0; k < byId.size(); ++k) dm->h_refById[k] = byId[k];
  const lsdg_params& P = c->P;
  st = frame_require(kf, KF_DATA);
  if (st != LSDG_OK) return st;
  ProfMark pm = prof_begin(c, LSDG_KERNEL_DEPTH);
  CK(c, launch_observe(dm, P, dm->h_refs, n, dm->h_refById, (int)byId.size(), offset, 0, n - 1,
                       (float)kf->numFramesTrackedOnThis, (float)kf->numMappedOnThis, c->stream));
  CK(c, launch_fill_regularize(dm, P, P.val_sum_min_for_keep, kf, c->stream));
  frame_depth_written(kf);
  // algorithmic bytes of one update (181 B per pixel: observe 70, fill holes 66,
  // regularize 37, setDepth 8)
  prof_end(c, pm, 181.0 * dm->g.w * dm->g.h);
  kf->numMappedOnThis++;
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_create_keyframe(lsdg_depthmap dm, lsdg_frame nkf) {
  if (!dm) return LSDG_EINVAL;
  lsdg_ctx c = dm->ctx;
  lsdg_status st = check_frame(dm, nkf);
  if (st != LSDG_OK) return st;
  if (!dm->valid || !dm->kf) return fail(c, LSDG_ESTATE, "depth map has no active keyframe");
  if (nkf->parentId < 0) return fail(c, LSDG_ESTATE, "new keyframe has no tracking parent");
  const lsdg_params& P = c->P;
  Pose t2p = nkf->thisToParent;
  t2p.s = 1.0;
  Pose oldToNew = pose_inv(t2p);
  const uint8_t* wg = (nkf->parentId == dm->kf->id && nkf->wasGoodAllocated) ? nkf->d_wasGood : nullptr;
  st = frame_require(nkf, KF_DATA);
  if (st != LSDG_OK) return st;
  ProfMark pm = prof_begin(c, LSDG_KERNEL_DEPTH);
  CK(c, launch_propagate(dm, P, nkf, pose_xf(oldToNew, false), wg, c->stream));
  dm->kf = nkf;
  dm->kfReactivated = false;
  CK(c, launch_regularize(dm, P, true, P.val_sum_min_for_keep, c->stream));
  CK(c, launch_fill_regularize(dm, P, P.val_sum_min_for_keep, nullptr, c->stream));
  CK(c, launch_mean_idepth(dm, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  double cnt = dm->h_red[0], sum = dm->h_red[1];
  float rescale = (float)(cnt / sum);
  CK(c, launch_rescale(dm, rescale, c->stream));
  Pose tp = pose_inv(oldToNew);
  tp.s = (double)rescale;
  nkf->thisToParent = tp;
  st = set_depth(dm, nkf);
  if (st != LSDG_OK) return st;
  prof_end(c, pm, 0.0);
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_finalize_keyframe(lsdg_depthmap dm) {
  if (!dm) return LSDG_EINVAL;
  lsdg_ctx c = dm->ctx;
  if (!dm->valid || !dm->kf) return fail(c, LSDG_ESTATE, "depth map has no active keyframe");
  lsdg_status st0 = frame_require(dm->kf, KF_DATA);
  if (st0 != LSDG_OK) return st0;
  CK(c, launch_fill_regularize(dm, c->P, c->P.val_sum_min_for_keep, dm->kf, c->stream));
  frame_depth_written(dm->kf);
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_read(lsdg_depthmap dm, lsdg_hypotheses* o) {
  if (!dm || !o) return LSDG_EINVAL;
  lsdg_ctx c = dm->ctx;
  size_t n = (size_t)dm->g.w * dm->g.h;
  const HypBuf& b = dm->buf[dm->cur];
  if (o->is_valid) CK(c, hipMemcpyAsync(o->is_valid, b.valid, n, hipMemcpyDeviceToHost, c->stream));
  if (o->blacklisted) CK(c, hipMemcpyAsync(o->blacklisted, b.bl, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (o->next_stereo_frame_min_id)
    CK(c, hipMemcpyAsync(o->next_stereo_frame_min_id, b.nsid, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (o->validity_counter) CK(c, hipMemcpyAsync(o->validity_counter, b.val, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (o->observe_code) CK(c, hipMemcpyAsync(o->observe_code, dm->d_obs, n, hipMemcpyDeviceToHost, c->stream));
  if (o->idepth || o->idepth_var || o->idepth_smoothed || o->idepth_var_smoothed) {
    std::vector<float> dv(4 * n);
    CK(c, hipMemcpyAsync(dv.data(), b.dv, n * 16, hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) {
      if (o->idepth) o->idepth[i] = dv[4 * i];
      if (o->idepth_var) o->idepth_var[i] = dv[4 * i + 1];
      if (o->idepth_smoothed) o->idepth_smoothed[i] = dv[4 * i + 2];
      if (o->idepth_var_smoothed) o->idepth_var_smoothed[i] = dv[4 * i + 3];
    }
  }
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_write(lsdg_depthmap dm, const lsdg_hypotheses* o) {
  if (!dm || !o) return LSDG_EINVAL;
  lsdg_ctx c = dm->ctx;
  size_t n = (size_t)dm->g.w * dm->g.h;
  const HypBuf& b = dm->buf[dm->cur];
  if (o->is_valid) CK(c, hipMemcpyAsync(b.valid, o->is_valid, n, hipMemcpyHostToDevice, c->stream));
  if (o->blacklisted) CK(c, hipMemcpyAsync(b.bl, o->blacklisted, n * 4, hipMemcpyHostToDevice, c->stream));
  if (o->next_stereo_frame_min_id)
    CK(c, hipMemcpyAsync(b.nsid, o->next_stereo_frame_min_id, n * 4, hipMemcpyHostToDevice, c->stream));
  if (o->validity_counter) CK(c, hipMemcpyAsync(b.val, o->validity_counter, n * 4, hipMemcpyHostToDevice, c->stream));
  if (o->idepth || o->idepth_var || o->idepth_smoothed || o->idepth_var_smoothed) {
    std::vector<float> dv(4 * n);
    CK(c, hipMemcpyAsync(dv.data(), b.dv, n * 16, hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) {
      if (o->idepth) dv[4 * i] = o->idepth[i];
      if (o->idepth_var) dv[4 * i + 1] = o->idepth_var[i];
      if (o->idepth_smoothed) dv[4 * i + 2] = o->idepth_smoothed[i];
      if (o->idepth_var_smoothed) dv[4 * i + 3] = o->idepth_var_smoothed[i];
    }
    CK(c, hipMemcpyAsync(b.dv, dv.data(), n * 16, hipMemcpyHostToDevice, c->stream));
  }
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_depth_active_keyframe(lsdg_depthmap dm, int32_t* id) {
  if (!dm || !id) return LSDG_EINVAL;
  *id = dm->kf ? dm->kf->id : -1;
  return LSDG_OK;
}
