// C-ABI of the ingest step before the hot path (lib/App/InputThread.cpp:57-71):
// the undistortion table, the fused undistort + convert + pyramid launch, and
// TextOutputIOWrapper's pose.txt line.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "capi_common.h"

using namespace lsdg;

struct lsdg_undistorter_s {
  lsdg_ctx ctx = nullptr;
  int inW = 0, inH = 0, outW = 0, outH = 0;
  float2* d_map = nullptr;   // one source position per output pixel
  uint8_t* d_src = nullptr;  // the distorted input, staged per call
  uint8_t* h_src = nullptr;  // pinned host staging of the input (one copy per call)
};

extern "C" lsdg_status lsdg_undistorter_create(lsdg_ctx c, int32_t inW, int32_t inH, int32_t outW, int32_t outH,
                                               const float* map, lsdg_undistorter* out) {
  if (!c || !map || !out) return LSDG_EINVAL;
  if (inW < 2 || inH < 2 || outW < 1 || outH < 1 || inW > 16384 || inH > 16384 || outW > 16384 || outH > 16384)
    return fail(c, LSDG_EINVAL, "undistorter size out of range (2..16384 input, 1..16384 output)");
  if ((outW & 1) || (outH & 1))  // the output is a frame's level 0 (even-sized)
    return fail(c, LSDG_EINVAL, "undistorted width and height must be even");
  const size_t n = (size_t)outW * outH;
  // every source must keep both bilinear taps inside the input (NaN fails too)
  for (size_t i = 0; i < n; ++i) {
    const float x = map[2 * i], y = map[2 * i + 1];
    if (x < 0.0f) continue;
    if (!(x >= 0.0f && x < (float)(inW - 1) && y >= 0.0f && y < (float)(inH - 1)))
      return fail(c, LSDG_EINVAL, "undistortion table entry outside the input");
  }
  lsdg_undistorter u = new (std::nothrow) lsdg_undistorter_s();
  if (!u) return LSDG_ENOMEM;
  u->ctx = c;
  u->inW = inW;
  u->inH = inH;
  u->outW = outW;
  u->outH = outH;
  hipError_t e = hipMalloc(&u->d_map, n * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc(&u->d_src, (size_t)inW * inH);
  if (e == hipSuccess) e = hipHostMalloc(&u->h_src, (size_t)inW * inH, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemcpyAsync(u->d_map, map, n * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    lsdg_undistorter_destroy(u);
    return fail(c, LSDG_EDEVICE, "undistorter setup", e);
  }
  *out = u;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_undistorter_destroy(lsdg_undistorter u) {
  if (!u) return LSDG_EINVAL;
  (void)hipStreamSynchronize(u->ctx->stream);
  if (u->d_map) (void)hipFree(u->d_map);
  if (u->d_src) (void)hipFree(u->d_src);
  if (u->h_src) (void)hipHostFree(u->h_src);
  delete u;
  return LSDG_OK;
}

// The FOV model's radial map in double precision, each source rounded to float
// once; out-of-range sources become (-1, -1).
extern "C" lsdg_status lsdg_fov_undistort_table(const float in_k[4], float omega, int32_t inW, int32_t inH,
                                                const float out_k[4], int32_t outW, int32_t outH, float* map) {
  if (!in_k || !out_k || !map || inW < 2 || inH < 2 || outW < 1 || outH < 1) return LSDG_EINVAL;
  if (!(out_k[0] != 0.0f && out_k[1] != 0.0f) || !std::isfinite(omega) || omega < 0.0f) return LSDG_EINVAL;
  const double w = omega, t2 = 2.0 * std::tan(w / 2.0);
  for (int y = 0; y < outH; ++y)
    for (int x = 0; x < outW; ++x) {
      const double ix = (x - (double)out_k[2]) / out_k[0], iy = (y - (double)out_k[3]) / out_k[1];
      const double r = std::sqrt(ix * ix + iy * iy);
      const double fac = (r == 0.0 || w == 0.0) ? 1.0 : std::atan(r * t2) / (w * r);
      const float sx = (float)(in_k[0] * fac * ix + in_k[2]), sy = (float)(in_k[1] * fac * iy + in_k[3]);
      const bool ok = sx > 0.0f && sy > 0.0f && sx < (float)(inW - 1) && sy < (float)(inH - 1);
      float* m = map + 2 * ((size_t)x + (size_t)y * outW);
      m[0] = ok ? sx : -1.0f;
      m[1] = ok ? sy : -1.0f;
    }
  return LSDG_OK;
}

// Uploads the distorted image and runs the fused ingest launch into frame f
// (sizes checked by the caller); leaves f with image levels 0-4 valid.
static lsdg_status ingest_into(lsdg_frame f, lsdg_undistorter u, const uint8_t* image, int32_t stride,
                               uint8_t* undistorted) {
  lsdg_ctx c = f->ctx;
  const int w = f->w[0], h = f->h[0];
  // rows into pinned staging on the host, then one copy (a 2D copy from
  // pageable memory goes row by row); the previous call drained the stream
  for (int y = 0; y < u->inH; ++y)
    std::memcpy(u->h_src + (size_t)y * u->inW, image + (size_t)y * stride, (size_t)u->inW);
  hipError_t e = hipMemcpyAsync(u->d_src, u->h_src, (size_t)u->inW * u->inH, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    ProfMark pm = prof_begin(c, LSDG_KERNEL_INGEST);
    // the 8-bit copy is always written: later pyramid rebuilds read it
    e = launch_ingest(f, u->d_src, u->inW, u->d_map, f->d_input_u8, c->stream);
    // table 8 B, one input byte, level 0 (4 B) and the 8-bit copy per output pixel, plus levels 1-4
    prof_end(c, pm, pyramid_bytes(f, true, 0) + (double)w * h * (8.0 + 1.0 + 1.0));
  }
  if (e == hipSuccess && undistorted)
    e = hipMemcpyAsync(undistorted, f->d_input_u8, (size_t)w * h, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LSDG_EDEVICE, "ingest", e);
  f->valid = fd_levels(FD_IMAGE, 0, PYR - 1);  // gradients are built on first use
  f->u8Level0 = true;
  return LSDG_OK;
}

static lsdg_status check_ingest(lsdg_ctx c, int w, int h, lsdg_undistorter u, const uint8_t* image, int32_t stride) {
  if (!u || !image || u->ctx != c) return LSDG_EINVAL;
  if (w != u->outW || h != u->outH) return fail(c, LSDG_EINVAL, "camera size differs from the undistorter's output");
  if (stride < u->inW) return fail(c, LSDG_EINVAL, "input row stride below the input width");
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_create_undistorted(lsdg_ctx c, int32_t id, const lsdg_camera* cam,
                                                     lsdg_undistorter u, const uint8_t* image, int32_t stride,
                                                     int32_t precision, uint8_t* undistorted, lsdg_frame* out) {
  if (!c || !cam || !out) return LSDG_EINVAL;
  lsdg_status st = check_ingest(c, cam->w, cam->h, u, image, stride);
  if (st != LSDG_OK) return st;
  lsdg_frame f;
  st = frame_alloc(c, id, cam, precision, &f);
  if (st != LSDG_OK) return st;
  st = ingest_into(f, u, image, stride, undistorted);
  if (st != LSDG_OK) {
    lsdg_frame_destroy(f);
    return st;
  }
  *out = f;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_ingest(lsdg_frame f, int32_t id, lsdg_undistorter u, const uint8_t* image,
                                         int32_t stride, uint8_t* undistorted) {
  if (!f) return LSDG_EINVAL;
  lsdg_ctx c = f->ctx;
  lsdg_status st = check_ingest(c, f->w[0], f->h[0], u, image, stride);
  if (st != LSDG_OK) return st;
  // every field as in a new frame; the arrays and the camera stay, and the
  // depth version moves on so point clouds built from the old content are
  // rebuilt
  CK(c, hipStreamSynchronize(c->stream));
  lsdg_frame_s fresh{};
  fresh.ctx = c;
  fresh.id = id;
  std::memcpy(fresh.w, f->w, sizeof(f->w));
  std::memcpy(fresh.h, f->h, sizeof(f->h));
  std::memcpy(fresh.fx, f->fx, sizeof(f->fx));
  std::memcpy(fresh.fy, f->fy, sizeof(f->fy));
  std::memcpy(fresh.cx, f->cx, sizeof(f->cx));
  std::memcpy(fresh.cy, f->cy, sizeof(f->cy));
  std::memcpy(fresh.fxi, f->fxi, sizeof(f->fxi));
  std::memcpy(fresh.fyi, f->fyi, sizeof(f->fyi));
  std::memcpy(fresh.cxi, f->cxi, sizeof(f->cxi));
  std::memcpy(fresh.cyi, f->cyi, sizeof(f->cyi));
  fresh.precision = f->precision;
  fresh.d_pool = f->d_pool;
  for (int l = 0; l < PYR; ++l) {
    fresh.d_image[l] = f->d_image[l];
    fresh.d_grad[l] = f->d_grad[l];
    fresh.d_gradh[l] = f->d_gradh[l];
    fresh.d_maxg[l] = f->d_maxg[l];
    fresh.d_idepth[l] = f->d_idepth[l];
    fresh.d_ivar[l] = f->d_ivar[l];
  }
  fresh.d_input_u8 = f->d_input_u8;
  fresh.d_wasGood = f->d_wasGood;
  fresh.depthVersion = f->depthVersion + 1;
  *f = fresh;
  return ingest_into(f, u, image, stride, undistorted);
}

extern "C" lsdg_status lsdg_pose_txt_line(int32_t id, const double cam_to_world[8], const double this_to_parent[8],
                                          char* buf, size_t cap) {
  if (!cam_to_world || !this_to_parent || !buf || cap == 0) return LSDG_EINVAL;
  const double* a = cam_to_world + 4;
  const double* b = this_to_parent + 4;
  // "%g" is the ostream default for a double (precision 6, no floatfield)
  const int n = std::snprintf(buf, cap, "%d,%g,%g,%g,%g,%g,%g\n", (int)id, a[0], a[1], a[2], b[0], b[1], b[2]);
  if (n < 0 || (size_t)n >= cap) {
    buf[0] = '\0';
    return LSDG_EINVAL;
  }
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_pose_txt(lsdg_frame f, char* buf, size_t cap) {
  if (!f) return LSDG_EINVAL;
  double a[8], b[8];
  pose_to8(f->camToWorld, a);
  pose_to8(f->thisToParent, b);
  return lsdg_pose_txt_line(f->id, a, b, buf, cap);
}
