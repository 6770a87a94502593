// C-ABI of the keyframe export: publishKeyframe's point records and the
// viewer's computeVbo filter (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:42-97,
// lib/Pangolin_IOWrapper/Keyframe.h:66-158), computed on the device.
#include <algorithm>
#include <cstring>

#include "capi_common.h"

using namespace lsdg;

extern "C" void lsdg_vbo_params_default(lsdg_vbo_params* p) {
  if (!p) return;
  p->scaled_th = 1e-3f;
  p->abs_th = 1e-1f;
  p->scale = 1.0f;
  p->min_near_support = 9;
}

// Scratch for n pixels: records | block offsets | vertices.
struct ExportScratch {
  lsdg_point_dense* rec;
  int* blk;
  lsdg_vertex* vtx;
};

static lsdg_status export_scratch(lsdg_ctx c, int w, int h, ExportScratch* out) {
  const size_t n = (size_t)w * h;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t oRec = 0, oBlk = up(n * sizeof(lsdg_point_dense)), oVtx = oBlk + up((vbo_block_count(w, h) + 1) * 4);
  const size_t total = oVtx + up(n * sizeof(lsdg_vertex));
  if (total > c->exportCap) {
    CK(c, hipStreamSynchronize(c->stream));
    if (c->d_export) (void)hipFree(c->d_export);
    c->d_export = nullptr;
    c->exportCap = 0;
    CK(c, hipMalloc(&c->d_export, total));
    c->exportCap = total;
  }
  char* b = (char*)c->d_export;
  out->rec = (lsdg_point_dense*)(b + oRec);
  out->blk = (int*)(b + oBlk);
  out->vtx = (lsdg_vertex*)(b + oVtx);
  return LSDG_OK;
}

static VboArgs vbo_args(int w, int h, float fx, float fy, float cx, float cy, const lsdg_vbo_params& p) {
  VboArgs a;
  a.w = w;
  a.h = h;
  // Keyframe.h:86-89
  a.fxi = 1 / fx;
  a.fyi = 1 / fy;
  a.cxi = -cx / fx;
  a.cyi = -cy / fy;
  a.scaledTH = p.scaled_th;
  a.absTH = p.abs_th;
  a.scale = p.scale;
  a.minNear = p.min_near_support;
  return a;
}

// Records of frame f at `level` into s.rec (device).
static lsdg_status dense_records(lsdg_frame f, int level, const ExportScratch& s) {
  lsdg_ctx c = f->ctx;
  if (!f->hasDepth) return fail(c, LSDG_ESTATE, "keyframe has no depth");
  lsdg_status st = frame_require(f, fd_bit(FD_IMAGE, level));
  if (st != LSDG_OK) return st;
  CK(c, launch_kf_dense(f->d_idepth[level], f->d_ivar[level], f->d_image[level], f->w[level] * f->h[level], s.rec,
                        c->stream));
  return LSDG_OK;
}

// Runs the filter on s.rec and copies the kept vertices to host `out`.
static lsdg_status run_vbo(lsdg_ctx c, const VboArgs& a, const ExportScratch& s, lsdg_vertex* out, int32_t* count) {
  const int nb = (int)vbo_block_count(a.w, a.h);
  int total = 0;
  if (nb > 0) {
    CK(c, launch_vbo(s.rec, a, s.blk, s.vtx, c->stream));
    CK(c, hipMemcpyAsync(&total, s.blk + nb, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
  }
  if (total > 0) {
    CK(c, hipMemcpyAsync(out, s.vtx, sizeof(lsdg_vertex) * (size_t)total, hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
  }
  *count = total;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_keyframe_points(lsdg_frame kf, int32_t level, lsdg_point_dense* out) {
  if (!kf || !out || level < 0 || level >= PYR) return LSDG_EINVAL;
  lsdg_ctx c = kf->ctx;
  ExportScratch s;
  lsdg_status st = export_scratch(c, kf->w[level], kf->h[level], &s);
  if (st == LSDG_OK) st = dense_records(kf, level, s);
  if (st != LSDG_OK) return st;
  CK(c, hipMemcpyAsync(out, s.rec, sizeof(lsdg_point_dense) * (size_t)kf->w[level] * kf->h[level],
                       hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_points_vbo(lsdg_ctx c, const lsdg_point_dense* in, int32_t w, int32_t h, float fx,
                                       float fy, float cx, float cy, const lsdg_vbo_params* p, lsdg_vertex* out,
                                       int32_t* count) {
  if (!c || !in || !p || !out || !count || w < 1 || h < 1 || w > 16384 || h > 16384) return LSDG_EINVAL;
  ExportScratch s;
  lsdg_status st = export_scratch(c, w, h, &s);
  if (st != LSDG_OK) return st;
  CK(c, hipMemcpyAsync(s.rec, in, sizeof(lsdg_point_dense) * (size_t)w * h, hipMemcpyHostToDevice, c->stream));
  return run_vbo(c, vbo_args(w, h, fx, fy, cx, cy, *p), s, out, count);
}

extern "C" lsdg_status lsdg_keyframe_vbo(lsdg_frame kf, int32_t level, const lsdg_vbo_params* p, lsdg_vertex* out,
                                         int32_t* count) {
  if (!kf || !p || !out || !count || level < 0 || level >= PYR) return LSDG_EINVAL;
  lsdg_ctx c = kf->ctx;
  const int w = kf->w[level], h = kf->h[level];
  ExportScratch s;
  lsdg_status st = export_scratch(c, w, h, &s);
  if (st == LSDG_OK) st = dense_records(kf, level, s);
  if (st != LSDG_OK) return st;
  return run_vbo(c, vbo_args(w, h, kf->fx[level], kf->fy[level], kf->cx[level], kf->cy[level], *p), s, out, count);
}
