// C-ABI of liblsdg.so: contexts, frames, tracking references and the plain
// (one alignment per call) SE3 / Sim3 trackers.  The LM control flow is the
// reference's (SE3Tracker::trackFrame, Sim3Tracker::trackFrameSim3); each
// evaluation is one fused kernel whose FP64 sums land in host-mapped memory,
// and the 6x6 / 7x7 systems are solved in FP64 on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "capi_common.h"

using namespace lsdg;

// ------------------------------------------------------------ defaults
extern "C" void lsdg_params_default(lsdg_params* p) {
  std::memset(p, 0, sizeof(*p));
  const int its[PYR] = {5, 20, 50, 100, 100};
  for (int l = 0; l < PYR; ++l) {
    p->lambda_initial[l] = 0.0f;
    p->lambda_success_fac[l] = 0.5f;
    p->lambda_fail_fac[l] = 2.0f;
    p->step_size_min[l] = 1e-8f;
    p->convergence_eps[l] = 0.999f;
    p->max_its_per_lvl[l] = its[l];
  }
  p->var_weight = 1.0f;
  p->huber_d = 3.0f;
  p->camera_pixel_noise2 = 4.0f * 4.0f;
  p->se3_min_level = 1;
  p->se3_max_level = 5;
  p->min_goodperall_pixel = 0.04f;
  p->min_goodperall_pixel_absmin = 0.01f;
  p->min_goodpergoodbad_pixel = 0.5f;
  p->max_diff_constant = 40.0f * 40.0f;
  p->max_diff_grad_mult = 0.5f * 0.5f;
  p->use_affine_lightning_estimation = 0;
  p->use_esm_tracking = 1;
  p->min_abs_grad_create = 5.0f;
  p->min_abs_grad_decrease = 5.0f;
  p->min_epl_length_squared = 1.0f;
  p->min_epl_grad_squared = 2.0f * 2.0f;
  p->min_epl_angle_squared = 0.3f * 0.3f;
  p->gradient_sample_dist = 1.0f;
  p->sample_point_to_border = 7.0f;
  p->max_epl_length_crop = 30.0f;
  p->min_epl_length_crop = 3.0f;
  p->max_error_stereo = 1300.0f;
  p->min_distance_error_stereo = 1.5f;
  p->stereo_epl_var_fac = 2.0f;
  p->min_depth = 0.05f;
  p->max_var = 0.5f * 0.5f;
  p->var_random_init_initial = 0.5f * 0.5f * 0.5f;
  p->var_gt_init_initial = 0.01f * 0.01f;
  p->succ_var_inc_fac = 1.01f;
  p->fail_var_inc_fac = 1.1f;
  p->diff_fac_observe = 1.0f;
  p->diff_fac_prop_merge = 1.0f;
  p->diff_fac_smoothing = 1.0f;
  p->depth_smoothing_factor = 1.0f;
  p->division_eps = 1e-10f;
  p->validity_counter_inc = 5;
  p->validity_counter_dec = 5;
  p->validity_counter_initial_observe = 5;
  p->validity_counter_max = 5;
  p->validity_counter_max_variable = 250;
  p->val_sum_min_for_create = 30;
  p->val_sum_min_for_keep = 24;
  p->val_sum_min_for_unblacklist = 100;
  p->min_blacklist = -1;
  p->allow_negative_idepths = 1;
  p->use_subpixel_stereo = 1;
  p->stereo_subpixel_oob_guard = 1;
  p->stereo_max_steps = 1024;
}

static bool params_ok(const lsdg_params* p) {
  if (p->se3_min_level < 0 || p->se3_max_level > PYR || p->se3_min_level >= p->se3_max_level) return false;
  for (int l = 0; l < PYR; ++l)
    if (p->max_its_per_lvl[l] < 0) return false;
  if (p->stereo_max_steps < 1 || p->stereo_max_steps > 1 << 20) return false;
  // epipolar samples reach 2 * gradient_sample_dist past the search interval,
  // which starts sample_point_to_border pixels inside the image
  if (!(p->gradient_sample_dist > 0.f) || !(2.f * p->gradient_sample_dist + 1.f < p->sample_point_to_border)) return false;
  if (!(p->max_epl_length_crop > 0.f) || !(p->min_depth > 0.f)) return false;
  return true;
}

// ------------------------------------------------------------- context
lsdg_status lsdg::fail(lsdg_ctx c, lsdg_status st, const char* what, hipError_t e) {
  if (c) {
    char buf[512];
    if (e != hipSuccess)
      std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
      std::snprintf(buf, sizeof buf, "%s", what);
    c->err = buf;
  }
  return st;
}

extern "C" lsdg_status lsdg_ctx_create(int32_t device, const lsdg_params* params, lsdg_ctx* out) {
  if (!out) return LSDG_EINVAL;
  *out = nullptr;
  lsdg_params P;
  if (params) P = *params;
  else lsdg_params_default(&P);
  if (!params_ok(&P)) return LSDG_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return LSDG_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return LSDG_EDEVICE;
  lsdg_ctx c = new (std::nothrow) lsdg_ctx_s();
  if (!c) return LSDG_ENOMEM;
  c->device = device;
  c->P = P;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  c->partials_blocks = 4096;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->d_partials, sizeof(double) * NS_MAX * c->partials_blocks) != hipSuccess ||
      hipMalloc(&c->d_counter, 64) != hipSuccess || hipMemset(c->d_counter, 0, 64) != hipSuccess ||
      hipHostMalloc(&c->h_result, sizeof(double) * 64, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c->d_result, c->h_result, 0) != hipSuccess ||
      noalloc_scratch_init(device) != hipSuccess) {  // the allocation-free entry point's device array
    lsdg_ctx_destroy(c);
    return LSDG_EDEVICE;
  }
  *out = c;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ctx_destroy(lsdg_ctx c) {
  if (!c) return LSDG_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_partials) (void)hipFree(c->d_partials);
  if (c->d_counter) (void)hipFree(c->d_counter);
  if (c->h_result) (void)hipHostFree(c->h_result);
  if (c->d_export) (void)hipFree(c->d_export);
  if (c->d_adj) (void)hipFree(c->d_adj);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ctx_set_params(lsdg_ctx c, const lsdg_params* p) {
  if (!c || !p) return LSDG_EINVAL;
  if (!params_ok(p)) return fail(c, LSDG_EINVAL, "invalid parameters");
  c->P = *p;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ctx_synchronize(lsdg_ctx c) {
  if (!c) return LSDG_EINVAL;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? LSDG_OK : fail(c, LSDG_EDEVICE, "hipStreamSynchronize", e);
}

extern "C" lsdg_status lsdg_ctx_set_sums(lsdg_ctx c, int32_t mode) {
  if (!c) return LSDG_EINVAL;
  if (mode != LSDG_SUMS_FP64 && mode != LSDG_SUMS_KAHAN) return fail(c, LSDG_EINVAL, "unknown summation mode");
  c->sums = mode;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ctx_get_sums(lsdg_ctx c, int32_t* mode) {
  if (!c || !mode) return LSDG_EINVAL;
  *mode = c->sums;
  return LSDG_OK;
}

extern "C" const char* lsdg_ctx_last_error(lsdg_ctx c) { return c ? c->err.c_str() : "null context"; }

// --------------------------------------------------------------- frames
// Frame::initialize: fx_l = fx_{l-1} * 0.5, cx_l = (cx_{l-1} + 0.5) * 0.5 - 0.5,
// fxi = 1 / fx_l, cxi = -cx_l / fx_l.
static void frame_intrinsics(lsdg_frame f, const lsdg_camera* cam) {
  for (int l = 0; l < PYR; ++l) {
    f->w[l] = l ? f->w[l - 1] / 2 : cam->w;
    f->h[l] = l ? f->h[l - 1] / 2 : cam->h;
    if (l == 0) {
      f->fx[0] = cam->fx; f->fy[0] = cam->fy; f->cx[0] = cam->cx; f->cy[0] = cam->cy;
    } else {
      f->fx[l] = f->fx[l - 1] * 0.5f;
      f->fy[l] = f->fy[l - 1] * 0.5f;
      f->cx[l] = (f->cx[l - 1] + 0.5f) * 0.5f - 0.5f;
      f->cy[l] = (f->cy[l - 1] + 0.5f) * 0.5f - 0.5f;
    }
    f->fxi[l] = 1.0f / f->fx[l];
    f->fyi[l] = 1.0f / f->fy[l];
    f->cxi[l] = -f->cx[l] / f->fx[l];
    f->cyi[l] = -f->cy[l] / f->fy[l];
  }
}

lsdg_status lsdg::frame_alloc(lsdg_ctx c, int32_t id, const lsdg_camera* cam, int32_t precision,
                                lsdg_frame* out) {
  if (!c || !cam || !out) return LSDG_EINVAL;
  if (cam->w < 32 || cam->h < 32 || cam->w > 16384 || cam->h > 16384 || (cam->w >> (PYR - 1)) < 4 ||
      (cam->h >> (PYR - 1)) < 4)
    return fail(c, LSDG_EINVAL, "frame size out of range (32..16384, >= 4 px at the coarsest level)");
  // The level-1 refPixelWasGood mask is indexed with (x >> 1, y >> 1) of
  // level-0 pixels (DepthMap::propagateDepth's tracking gate): an odd width
  // or height would put the last column or row one past the mask.  Upstream
  // requires multiples of 16 (SlamSystem's constructor); even suffices here.
  if ((cam->w & 1) || (cam->h & 1)) return fail(c, LSDG_EINVAL, "frame width and height must be even");
  if (precision < LSDG_FP32 || precision > LSDG_BF16) return fail(c, LSDG_EINVAL, "unknown precision");
  lsdg_frame f = new (std::nothrow) lsdg_frame_s();
  if (!f) return LSDG_ENOMEM;
  f->ctx = c;
  f->id = id;
  f->precision = precision;
  frame_intrinsics(f, cam);
  // one pool: per level image, grad4, [gradh], maxg, idepth, ivar; then u8 staging and the level-1 mask
  size_t off[PYR][6];
  size_t total = 0;
  auto take = [&](size_t bytes) {
    size_t o = total;
    total += (bytes + 255) & ~(size_t)255;
    return o;
  };
  for (int l = 0; l < PYR; ++l) {
    size_t n = (size_t)f->w[l] * f->h[l];
    off[l][0] = take(n * 4);
    off[l][1] = take(n * 16);
    off[l][2] = precision != LSDG_FP32 ? take(n * 8) : 0;
    off[l][3] = take(n * 4);
    off[l][4] = take(n * 4);
    off[l][5] = take(n * 4);
  }
  size_t off_u8 = take((size_t)f->w[0] * f->h[0]);
  size_t off_wg = take((size_t)f->w[1] * f->h[1]);
  hipError_t e = hipMalloc(&f->d_pool, total);
  if (e != hipSuccess) {
    delete f;
    return fail(c, LSDG_ENOMEM, "hipMalloc(frame)", e);
  }
  char* b = (char*)f->d_pool;
  for (int l = 0; l < PYR; ++l) {
    f->d_image[l] = (float*)(b + off[l][0]);
    f->d_grad[l] = (float4*)(b + off[l][1]);
    f->d_gradh[l] = precision != LSDG_FP32 ? (uint2*)(b + off[l][2]) : nullptr;
    f->d_maxg[l] = (float*)(b + off[l][3]);
    f->d_idepth[l] = (float*)(b + off[l][4]);
    f->d_ivar[l] = (float*)(b + off[l][5]);
  }
  f->d_input_u8 = (uint8_t*)(b + off_u8);
  f->d_wasGood = (uint8_t*)(b + off_wg);
  *out = f;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_destroy(lsdg_frame f) {
  if (!f) return LSDG_EINVAL;
  (void)hipStreamSynchronize(f->ctx->stream);
  if (f->d_pool) (void)hipFree(f->d_pool);
  delete f;
  return LSDG_OK;
}

// Algorithmic bytes of a pyramid build: the image stage reads level 0 and
// writes levels 1-4; the gradient stage reads each built level's image and
// writes what `grads` names (16 B float4, 8 B packed, 4 B max-gradient).
double lsdg::pyramid_bytes(lsdg_frame f, bool images, unsigned grads) {
  double b = 0.0;
  if (images) {
    b += (double)f->w[0] * f->h[0] * 4.0;
    for (int l = 1; l < PYR; ++l) b += (double)f->w[l] * f->h[l] * 4.0;
  }
  for (int l = 0; l < PYR; ++l) {
    double n = (double)f->w[l] * f->h[l], w = 0.0;
    if (grads & fd_bit(FD_GRAD, l)) w += 16.0;
    if (grads & fd_bit(FD_GRADH, l)) w += 8.0;
    if (grads & fd_bit(FD_MAXG, l)) w += 4.0;
    if (w > 0.0) b += n * (4.0 + w);
  }
  return b;
}

unsigned lsdg::fd_track(lsdg_frame f, int lo, int hi) {
  return fd_levels(f->precision == LSDG_FP32 ? FD_GRAD : FD_GRADH, lo, hi);
}

lsdg_status lsdg::frame_require(lsdg_frame f, unsigned need) {
  if (!f->d_gradh[0]) need &= ~fd_levels(FD_GRADH, 0, PYR - 1);
  unsigned miss = need & ~f->valid;
  if (!miss) return LSDG_OK;
  lsdg_ctx c = f->ctx;
  const unsigned images = fd_levels(FD_IMAGE, 1, PYR - 1);
  const unsigned derived = fd_levels(FD_GRAD, 0, PYR - 1) | fd_levels(FD_GRADH, 0, PYR - 1) |
                           fd_levels(FD_MAXG, 0, PYR - 1);
  const unsigned level0 = fd_bit(FD_IMAGE, 0) | fd_bit(FD_GRAD, 0) | fd_bit(FD_GRADH, 0) | fd_bit(FD_MAXG, 0);
  // data above level 0 reads the image pyramid, which one launch builds whole
  bool buildImages = (f->valid & images) != images && (miss & ~level0) != 0;
  unsigned grads = miss & derived;
  ProfMark pm = prof_begin(c, LSDG_KERNEL_PYRAMID);
  if (buildImages) CK(c, launch_pyr_images(f, false, c->stream));
  if (grads) CK(c, launch_pyr_grads(f, grads, c->stream));
  prof_end(c, pm, pyramid_bytes(f, buildImages, grads));
  f->valid |= (buildImages ? images : 0u) | grads;
  return LSDG_OK;
}

// Every pyramid level of every kind the frame can hold.
static unsigned fd_all(lsdg_frame f) {
  unsigned m = fd_levels(FD_IMAGE, 0, PYR - 1) | fd_levels(FD_GRAD, 0, PYR - 1) | fd_levels(FD_MAXG, 0, PYR - 1);
  if (f->d_gradh[0]) m |= fd_levels(FD_GRADH, 0, PYR - 1);
  return m;
}

extern "C" lsdg_status lsdg_frame_create(lsdg_ctx c, int32_t id, const lsdg_camera* cam, const float* image,
                                         int32_t precision, lsdg_frame* out) {
  if (!image) return LSDG_EINVAL;
  lsdg_frame f;
  lsdg_status st = frame_alloc(c, id, cam, precision, &f);
  if (st != LSDG_OK) return st;
  size_t n = (size_t)cam->w * cam->h;
  hipError_t e = hipMemcpyAsync(f->d_image[0], image, n * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) f->valid = fd_bit(FD_IMAGE, 0);  // the rest is built on first use
  else st = fail(c, LSDG_EDEVICE, "upload image", e);
  if (st == LSDG_OK) {
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) st = fail(c, LSDG_EDEVICE, "frame pyramid", e);
  }
  if (st != LSDG_OK) {
    lsdg_frame_destroy(f);
    return st;
  }
  *out = f;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_create_u8(lsdg_ctx c, int32_t id, const lsdg_camera* cam, const uint8_t* image,
                                            int32_t stride, int32_t precision, lsdg_frame* out) {
  if (!image || !cam || stride < cam->w) return LSDG_EINVAL;
  lsdg_frame f;
  lsdg_status st = frame_alloc(c, id, cam, precision, &f);
  if (st != LSDG_OK) return st;
  // a 2D copy from pageable memory goes row by row: pack padded rows first
  std::vector<uint8_t> packed;
  const uint8_t* src = image;
  if (stride != cam->w) {
    packed.resize((size_t)cam->w * cam->h);
    for (int y = 0; y < cam->h; ++y)
      std::memcpy(packed.data() + (size_t)y * cam->w, image + (size_t)y * stride, (size_t)cam->w);
    src = packed.data();
  }
  hipError_t e = hipMemcpyAsync(f->d_input_u8, src, (size_t)cam->w * cam->h, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    // the image stage converts level 0 and builds levels 1-4; the rest is built on first use
    ProfMark pm = prof_begin(c, LSDG_KERNEL_PYRAMID);
    e = launch_pyr_images(f, true, c->stream);
    prof_end(c, pm, pyramid_bytes(f, true, 0));
    if (e == hipSuccess) {
      f->valid = fd_levels(FD_IMAGE, 0, PYR - 1);
      f->u8Level0 = true;
    }
    else st = fail(c, LSDG_EDEVICE, "frame pyramid", e);
  } else {
    st = fail(c, LSDG_EDEVICE, "upload image", e);
  }
  if (st == LSDG_OK) {
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) st = fail(c, LSDG_EDEVICE, "frame pyramid", e);
  }
  if (st != LSDG_OK) {
    lsdg_frame_destroy(f);
    return st;
  }
  *out = f;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_build_pyramid(lsdg_frame f) {
  if (!f) return LSDG_EINVAL;
  f->valid &= fd_bit(FD_IMAGE, 0);
  lsdg_status st = frame_require(f, fd_all(f));
  if (st != LSDG_OK) return st;
  CK(f->ctx, hipStreamSynchronize(f->ctx->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_invalidate(lsdg_frame f) {
  if (!f) return LSDG_EINVAL;
  f->valid &= fd_bit(FD_IMAGE, 0);
  return LSDG_OK;
}

lsdg_status lsdg::frame_depth_from_level0(lsdg_frame f) {
  lsdg_ctx c = f->ctx;
  CK(c, launch_pyr_depth(f, c->stream));
  frame_depth_written(f);
  return LSDG_OK;
}

void lsdg::frame_depth_written(lsdg_frame f) {
  f->hasDepth = true;
  f->depthVersion++;
}

extern "C" lsdg_status lsdg_frame_set_depth(lsdg_frame f, const float* idepth, const float* var) {
  if (!f || !idepth || !var) return LSDG_EINVAL;
  lsdg_ctx c = f->ctx;
  size_t n = (size_t)f->w[0] * f->h[0];
  CK(c, hipMemcpyAsync(f->d_idepth[0], idepth, n * 4, hipMemcpyHostToDevice, c->stream));
  CK(c, hipMemcpyAsync(f->d_ivar[0], var, n * 4, hipMemcpyHostToDevice, c->stream));
  CK(c, launch_depth_normalize(f->d_idepth[0], f->d_ivar[0], (int)n, c->stream));
  lsdg_status st = frame_depth_from_level0(f);
  if (st != LSDG_OK) return st;
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

lsdg_status lsdg::frame_ensure_was_good(lsdg_frame f) {
  if (f->wasGoodAllocated) return LSDG_OK;
  CK(f->ctx, hipMemsetAsync(f->d_wasGood, 1, (size_t)f->w[1] * f->h[1], f->ctx->stream));
  f->wasGoodAllocated = true;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_read(lsdg_frame f, int32_t what, int32_t level, void* out, size_t bytes) {
  if (!f || !out || level < 0 || level >= PYR) return LSDG_EINVAL;
  lsdg_ctx c = f->ctx;
  size_t n = (size_t)f->w[level] * f->h[level];
  const void* src = nullptr;
  size_t need = n * 4;
  unsigned req = 0;
  switch (what) {
    case LSDG_WHAT_IMAGE: src = f->d_image[level]; req = fd_bit(FD_IMAGE, level); break;
    case LSDG_WHAT_GRADIENTS: src = f->d_grad[level]; need = n * 16; req = fd_bit(FD_GRAD, level); break;
    case LSDG_WHAT_MAXGRAD: src = f->d_maxg[level]; req = fd_bit(FD_MAXG, level); break;
    case LSDG_WHAT_IDEPTH:
    case LSDG_WHAT_IDEPTH_VAR:
      if (!f->hasDepth) return fail(c, LSDG_ESTATE, "frame has no depth");
      src = what == LSDG_WHAT_IDEPTH ? f->d_idepth[level] : f->d_ivar[level];
      break;
    case LSDG_WHAT_REF_PIXEL_WAS_GOOD:
      need = (size_t)f->w[1] * f->h[1];
      if (bytes < need) return fail(c, LSDG_EINVAL, "output buffer too small");
      if (!f->wasGoodAllocated) {
        std::memset(out, 1, need);
        return LSDG_OK;
      }
      src = f->d_wasGood;
      break;
    default: return fail(c, LSDG_EINVAL, "unknown 'what'");
  }
  if (bytes < need) return fail(c, LSDG_EINVAL, "output buffer too small");
  lsdg_status st = frame_require(f, req);
  if (st != LSDG_OK) return st;
  CK(c, hipMemcpyAsync(out, src, need, hipMemcpyDeviceToHost, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_set_tracking(lsdg_frame f, lsdg_frame parent, const double tp[8], float itr,
                                               const uint8_t* wg) {
  if (!f || !tp) return LSDG_EINVAL;
  f->parent = parent;
  f->parentId = parent ? parent->id : -1;
  f->thisToParent = pose_from8(tp);
  f->initialTrackedResidual = itr;
  if (wg) {
    lsdg_ctx c = f->ctx;
    CK(c, hipMemcpyAsync(f->d_wasGood, wg, (size_t)f->w[1] * f->h[1], hipMemcpyHostToDevice, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
    f->wasGoodAllocated = true;
  }
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_get_tracking(lsdg_frame f, int32_t* pid, double tp[8], float* itr) {
  if (!f) return LSDG_EINVAL;
  if (pid) *pid = f->parentId;
  if (tp) pose_to8(f->thisToParent, tp);
  if (itr) *itr = f->initialTrackedResidual;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_set_cam_to_world(lsdg_frame f, const double c2w[8]) {
  if (!f || !c2w) return LSDG_EINVAL;
  f->camToWorld = pose_from8(c2w);
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_counters(lsdg_frame f, int32_t* tracked, int32_t* mapped) {
  if (!f) return LSDG_EINVAL;
  if (tracked) *tracked = f->numFramesTrackedOnThis;
  if (mapped) *mapped = f->numMappedOnThis;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_frame_set_counters(lsdg_frame f, int32_t tracked, int32_t mapped) {
  if (!f) return LSDG_EINVAL;
  f->numFramesTrackedOnThis = tracked;
  f->numMappedOnThis = mapped;
  return LSDG_OK;
}

// ----------------------------------------------------- tracking reference
extern "C" lsdg_status lsdg_ref_create(lsdg_ctx c, lsdg_frame kf, lsdg_ref* out) {
  if (!c || !kf || !out) return LSDG_EINVAL;
  if (kf->ctx != c) return fail(c, LSDG_EINVAL, "keyframe belongs to another context");
  lsdg_ref r = new (std::nothrow) lsdg_ref_s();
  if (!r) return LSDG_ENOMEM;
  r->ctx = c;
  r->kf = kf;
  *out = r;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ref_destroy(lsdg_ref r) {
  if (!r) return LSDG_EINVAL;
  (void)hipStreamSynchronize(r->ctx->stream);
  for (int l = 0; l < PYR; ++l)
    if (r->d_pool[l]) (void)hipFree(r->d_pool[l]);
  if (r->d_blockcounts) (void)hipFree(r->d_blockcounts);
  delete r;
  return LSDG_OK;
}

lsdg_status lsdg::ref_points(lsdg_ref r, int level) {
  lsdg_ctx c = r->ctx;
  if (level < 0 || level >= PYR) return fail(c, LSDG_EINVAL, "level out of range");
  if (!r->kf->hasDepth) return fail(c, LSDG_ESTATE, "keyframe has no depth");
  if (r->depthVersion != r->kf->depthVersion) {
    for (int l = 0; l < PYR; ++l) r->num[l] = -1;
    r->depthVersion = r->kf->depthVersion;
  }
  if (r->num[level] >= 0) return LSDG_OK;
  lsdg_status st = frame_require(r->kf, fd_bit(FD_IMAGE, level) | fd_bit(FD_GRAD, level));
  if (st != LSDG_OK) return st;
  CK(c, launch_point_cloud(r, level, c->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ref_make_point_cloud(lsdg_ref r, int32_t level) {
  if (!r) return LSDG_EINVAL;
  lsdg_status st = ref_points(r, level);
  if (st != LSDG_OK) return st;
  CK(r->ctx, hipStreamSynchronize(r->ctx->stream));
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ref_num_points(lsdg_ref r, int32_t level, int32_t* n) {
  if (!r || !n) return LSDG_EINVAL;
  lsdg_status st = lsdg_ref_make_point_cloud(r, level);
  if (st != LSDG_OK) return st;
  *n = r->num[level];
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ref_read_points(lsdg_ref r, int32_t level, float* out, int32_t max_points) {
  if (!r || !out) return LSDG_EINVAL;
  lsdg_status st = lsdg_ref_make_point_cloud(r, level);
  if (st != LSDG_OK) return st;
  int n = r->num[level];
  if (max_points < n) return fail(r->ctx, LSDG_EINVAL, "output buffer too small");
  std::vector<float4> p(n);
  std::vector<float> v(n);
  std::vector<float2> g(n);
  std::vector<int> ix(n);
  const PointsLevel& pl = r->lv[level];
  lsdg_ctx c = r->ctx;
  if (n) {
    CK(c, hipMemcpyAsync(p.data(), pl.posI, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipMemcpyAsync(v.data(), pl.var, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipMemcpyAsync(g.data(), pl.grad, n * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipMemcpyAsync(ix.data(), pl.idx, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
  }
  for (int i = 0; i < n; ++i) {
    float* o = out + 8 * (size_t)i;
    o[0] = p[i].x; o[1] = p[i].y; o[2] = p[i].z; o[3] = p[i].w;
    o[4] = v[i]; o[5] = g[i].x; o[6] = g[i].y; o[7] = (float)ix[i];
  }
  return LSDG_OK;
}

// ---------------------------------------------------------------- tracking
Xf lsdg::pose_xf(const Pose& T, bool scaled) {
  Xf x;
  pose_rt_f(T, scaled, x.R, x.t);
  return x;
}

TrackKParams lsdg::track_kparams(const lsdg_params& P, bool sim3) {
  TrackKParams k;
  k.huber = sim3 ? P.huber_d : P.huber_d / 2.0f;
  k.var_weight = P.var_weight;
  k.noise2 = P.camera_pixel_noise2;
  k.max_diff_constant = P.max_diff_constant;
  k.max_diff_grad_mult = P.max_diff_grad_mult;
  k.aff_a = 1.0f;
  k.aff_b = 0.0f;
  k.use_esm = P.use_esm_tracking;
  k.aff_stats = 1;  // the plain evaluation reports them (lsdg_eval)
  k.roll[0] = 1.f; k.roll[1] = 0.f; k.roll[2] = 0.f; k.roll[3] = 1.f;
  return k;
}

void lsdg::esm_roll(const Pose& T, float* r4) { pose_roll_f(T, r4); }

// Algorithmic bytes of one evaluation at level l: per point xyz + I_ref + var
// (20 B) and the isGood byte, plus the level's (I, gx, gy) image; Sim3 adds
// the depth residual's point terms and the frame's idepth / variance.
double lsdg::eval_bytes(int n, lsdg_frame f, int lvl, bool sim3) {
  double P = (double)f->w[lvl] * f->h[lvl];
  double bpx = f->precision == LSDG_FP32 ? 12.0 : 6.0;
  double b = n * 21.0 + P * bpx;
  if (sim3) b += n * 9.0 + P * 8.0;
  return b;
}

// One SE3 evaluation; sums into S (host, NS_SE3 doubles).
static lsdg_status se3_eval_sync(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int lvl, const Pose& T, float aa, float ab,
                                 uint8_t* d_mask, double* S) {
  lsdg_status st = ref_points(r, lvl);
  if (st != LSDG_OK) return st;
  const PointsLevel& pl = r->lv[lvl];
  if (pl.n == 0) {
    std::memset(S, 0, sizeof(double) * NS_SE3);
    return LSDG_OK;
  }
  st = frame_require(f, fd_track(f, lvl, lvl));
  if (st != LSDG_OK) return st;
  TrackKParams kp = track_kparams(c->P, false);
  kp.aff_a = aa;
  kp.aff_b = ab;
  KTimer kt = prof_timer(c);
  CK(c, launch_se3_eval(c, pl, f, lvl, pose_xf(T, false), kp, d_mask, c->d_result, &kt));
  prof_timer_done(c, LSDG_KERNEL_EVAL, kt, eval_bytes(pl.n, f, lvl, false));
  CK(c, hipStreamSynchronize(c->stream));
  std::memcpy(S, c->h_result, sizeof(double) * NS_SE3);
  return LSDG_OK;
}

static void unpack_A6(const double* S, double n, double* A, double* b) {
  int k = 0;
  for (int a = 0; a < 6; ++a)
    for (int cc = a; cc < 6; ++cc) {
      A[6 * a + cc] = S[S3_A + k] / n;
      A[6 * cc + a] = A[6 * a + cc];
      ++k;
    }
  for (int a = 0; a < 6; ++a) b[a] = -S[S3_B + a] / n;
}

// Statistics of the last residual pass; sums and LM comparisons are FP64 (the
// per-point terms are FP32, as in the reference).
struct Se3Out {
  double lastGoodCount = 0, lastBadCount = 0, lastMeanRes = 0, pointUsage = 0;
  float affA = 1, affB = 0, affAIt = 1, affBIt = 0;
};

static void se3_pass_stats(const double* S, int N, Se3Out& o) {
  o.lastGoodCount = S[S3_GOOD];
  o.lastBadCount = S[S3_BAD];
  o.lastMeanRes = S[S3_SIGNED] / S[S3_GOOD];
  o.pointUsage = S[S3_USAGE] / (double)N;
  double sxx = S[S3_SXX], syy = S[S3_SYY], sx = S[S3_SX], sy = S[S3_SY], sw = S[S3_SW];
  o.affAIt = (float)std::sqrt((syy - sy * sy / sw) / (sxx - sx * sx / sw));
  o.affBIt = (float)((sy - (double)o.affAIt * sx) / sw);
}

extern "C" lsdg_status lsdg_se3_track(lsdg_ctx c, lsdg_ref r, lsdg_frame f, const double init[7], double out[7],
                                      lsdg_track_stats* stats) {
  if (!c || !r || !f || !init || !out) return LSDG_EINVAL;
  if (r->ctx != c || f->ctx != c) return fail(c, LSDG_EINVAL, "objects belong to another context");
  if (f->w[0] != r->kf->w[0] || f->h[0] != r->kf->h[0]) return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
  const lsdg_params& P = c->P;
  lsdg_track_stats st;
  std::memset(&st, 0, sizeof st);
  st.tracking_was_good = 1;
  Se3Out o;
  Pose T = pose_inv(pose_from7(init));
  T.s = 1.0;
  double last_residual = 0.0;
  double S[NS_MAX], Sacc[NS_MAX];
  bool diverged = false;
  for (int lvl = P.se3_max_level - 1; lvl >= P.se3_min_level && !diverged; --lvl) {
    lsdg_status s0 = ref_points(r, lvl);
    if (s0 != LSDG_OK) return s0;
    int N = r->lv[lvl].n;
    uint8_t* mask = nullptr;
    if (lvl == P.se3_min_level && lvl == 1) {  // refPixelWasGood is a level-1 mask
      s0 = frame_ensure_was_good(f);
      if (s0 != LSDG_OK) return s0;
      mask = f->d_wasGood;
    }
    const float lim = P.min_goodperall_pixel_absmin * (float)(f->w[lvl] * f->h[lvl]);
    s0 = se3_eval_sync(c, r, f, lvl, T, o.affA, o.affB, mask, S);
    if (s0 != LSDG_OK) return s0;
    se3_pass_stats(S, N, o);
    if ((float)S[S3_WARPED] < lim) {
      diverged = true;
      break;
    }
    if (P.use_affine_lightning_estimation) { o.affA = o.affAIt; o.affB = o.affBIt; }
    double lastErr = S[S3_ERR] / S[S3_WARPED];
    std::memcpy(Sacc, S, sizeof(double) * NS_SE3);
    st.num_residual_calls[lvl]++;
    float lambda = P.lambda_initial[lvl];
    for (int it = 0; it < P.max_its_per_lvl[lvl] && !diverged; ++it) {
      double A[36], b[6], inc[6];
      unpack_A6(Sacc, Sacc[S3_WARPED], A, b);
      st.num_warp_update_calls[lvl]++;
      int incTry = 0;
      while (true) {
        double Al[36];
        std::memcpy(Al, A, sizeof Al);
        for (int a = 0; a < 6; ++a) Al[7 * a] *= 1.0 + (double)lambda;
        ldlt_solve<6>(Al, b, inc);
        incTry++;
        Pose Tn = pose_mul(se3_exp(inc), T);
        s0 = se3_eval_sync(c, r, f, lvl, Tn, o.affA, o.affB, mask, S);
        if (s0 != LSDG_OK) return s0;
        se3_pass_stats(S, N, o);
        if ((float)S[S3_WARPED] < lim) {
          diverged = true;
          break;
        }
        double error = S[S3_ERR] / S[S3_WARPED];
        st.num_residual_calls[lvl]++;
        if (error < lastErr) {
          T = Tn;
          std::memcpy(Sacc, S, sizeof(double) * NS_SE3);
          if (P.use_affine_lightning_estimation) { o.affA = o.affAIt; o.affB = o.affBIt; }
          if (error / lastErr > (double)P.convergence_eps[lvl]) it = P.max_its_per_lvl[lvl];
          last_residual = lastErr = error;
          if (lambda <= 0.2f) lambda = 0.f;
          else lambda *= P.lambda_success_fac[lvl];
          break;
        } else {
          double dd = 0;
          for (int a = 0; a < 6; ++a) dd += inc[a] * inc[a];
          if (!(dd > (double)P.step_size_min[lvl])) {
            it = P.max_its_per_lvl[lvl];
            break;
          }
          if (lambda == 0.f) lambda = 0.2f;
          else lambda *= std::pow(P.lambda_fail_fac[lvl], (float)incTry);
        }
      }
    }
  }
  st.last_good_count = (int)o.lastGoodCount;
  st.last_bad_count = (int)o.lastBadCount;
  st.last_mean_res = (float)o.lastMeanRes;
  st.point_usage = (float)o.pointUsage;
  st.affine_a = o.affA;
  st.affine_b = o.affB;
  if (diverged) {
    st.diverged = 1;
    st.tracking_was_good = 0;
    if (stats) *stats = st;
    pose_to7(pose_identity(), out);
    return LSDG_DIVERGED;
  }
  st.last_residual = (float)last_residual;
  int ml = P.se3_min_level;
  bool good = o.lastGoodCount / (double)(f->w[ml] * f->h[ml]) > (double)P.min_goodperall_pixel &&
              o.lastGoodCount / (o.lastGoodCount + o.lastBadCount) > (double)P.min_goodpergoodbad_pixel;
  st.tracking_was_good = good ? 1 : 0;
  if (good) r->kf->numFramesTrackedOnThis++;
  Pose res = pose_inv(T);
  res.s = 1.0;
  f->initialTrackedResidual = (float)(last_residual / o.pointUsage);
  f->thisToParent = res;
  f->parent = r->kf;
  f->parentId = r->kf->id;
  pose_to7(res, out);
  if (stats) *stats = st;
  return LSDG_OK;
}

static void fill_eval_se3(const double* S, lsdg_eval* o) {
  std::memset(o, 0, sizeof(*o));
  int k = 0;
  for (int a = 0; a < 6; ++a)
    for (int cc = a; cc < 6; ++cc) {
      o->A[7 * a + cc] = S[S3_A + k];
      o->A[7 * cc + a] = S[S3_A + k];
      ++k;
    }
  for (int a = 0; a < 6; ++a) o->b[a] = S[S3_B + a];
  o->error = S[S3_ERR];
  o->sum_res_p = S[S3_ERR];
  o->num_terms_p = S[S3_WARPED];
  o->n_constraints = S[S3_WARPED];
  o->warped_count = (int32_t)S[S3_WARPED];
  o->good_count = (int32_t)S[S3_GOOD];
  o->bad_count = (int32_t)S[S3_BAD];
  o->sum_res_unweighted = S[S3_RESUNW];
  o->sum_signed_res = S[S3_SIGNED];
  o->usage = S[S3_USAGE];
  o->sxx = S[S3_SXX]; o->syy = S[S3_SYY]; o->sx = S[S3_SX]; o->sy = S[S3_SY]; o->sw = S[S3_SW];
}

extern "C" lsdg_status lsdg_se3_evaluate(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int32_t level, const double pose[7],
                                         float aa, float ab, lsdg_eval* out, uint8_t* is_good) {
  if (!c || !r || !f || !pose || !out) return LSDG_EINVAL;
  if (level < 0 || level >= PYR) return fail(c, LSDG_EINVAL, "level out of range");
  if (f->w[0] != r->kf->w[0] || f->h[0] != r->kf->h[0]) return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
  uint8_t* d_mask = nullptr;
  size_t mbytes = (size_t)r->kf->w[1] * r->kf->h[1];
  if (is_good && level == 1) {  // refPixelWasGood is a level-1 mask
    CK(c, hipMalloc(&d_mask, mbytes));
    hipError_t e = hipMemcpyAsync(d_mask, is_good, mbytes, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      (void)hipFree(d_mask);
      return fail(c, LSDG_EDEVICE, "mask upload", e);
    }
  }
  double S[NS_MAX];
  lsdg_status st = se3_eval_sync(c, r, f, level, pose_from7(pose), aa, ab, d_mask, S);
  if (st == LSDG_OK && d_mask) {
    hipError_t e = hipMemcpy(is_good, d_mask, mbytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = fail(c, LSDG_EDEVICE, "mask download", e);
  }
  if (d_mask) (void)hipFree(d_mask);
  if (st != LSDG_OK) return st;
  fill_eval_se3(S, out);
  return LSDG_OK;
}

// ------------------------------------------------------------------ Sim3
static lsdg_status sim3_eval_sync(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int lvl, const Pose& T, float aa, float ab,
                                  double* S) {
  lsdg_status st = ref_points(r, lvl);
  if (st != LSDG_OK) return st;
  const PointsLevel& pl = r->lv[lvl];
  if (pl.n == 0) {
    std::memset(S, 0, sizeof(double) * NS_SIM3);
    return LSDG_OK;
  }
  st = frame_require(f, fd_track(f, lvl, lvl));
  if (st != LSDG_OK) return st;
  TrackKParams kp = track_kparams(c->P, true);
  kp.aff_a = aa;
  kp.aff_b = ab;
  esm_roll(T, kp.roll);
  KTimer kt = prof_timer(c);
  CK(c, launch_sim3_eval(c, pl, f, lvl, pose_xf(T, true), kp, c->d_result, &kt));
  prof_timer_done(c, LSDG_KERNEL_EVAL, kt, eval_bytes(pl.n, f, lvl, true));
  CK(c, hipStreamSynchronize(c->stream));
  std::memcpy(S, c->h_result, sizeof(double) * NS_SIM3);
  return LSDG_OK;
}

struct Sim3Res {
  double sumResD = 0, sumResP = 0;
  int numTermsD = 0, numTermsP = 0;
  double meanD = 0, meanP = 0, mean = 0;
};

static Sim3Res sim3_res(const double* S) {
  Sim3Res r;
  r.sumResD = S[SM_RESD];
  r.sumResP = S[SM_RESP];
  r.numTermsD = (int)S[SM_NTD];
  r.numTermsP = (int)S[SM_WARPED];
  r.mean = (S[SM_RESD] + S[SM_RESP]) / (S[SM_NTD] + S[SM_WARPED]);
  r.meanD = S[SM_RESD] / S[SM_NTD];
  r.meanP = S[SM_RESP] / S[SM_WARPED];
  return r;
}

static void unpack_A7(const double* S, double* A7, double* b7) {
  int k = 0;
  for (int a = 0; a < 7; ++a)
    for (int cc = a; cc < 7; ++cc) {
      A7[7 * a + cc] = S[SM_A + k];
      A7[7 * cc + a] = S[SM_A + k];
      ++k;
    }
  for (int a = 0; a < 7; ++a) b7[a] = S[SM_B + a];
}

extern "C" lsdg_status lsdg_sim3_track(lsdg_ctx c, lsdg_ref r, lsdg_frame f, const double init[8], int32_t startLevel,
                                       int32_t finalLevel, double out[8], lsdg_track_stats* stats) {
  if (!c || !r || !f || !init || !out) return LSDG_EINVAL;
  if (r->ctx != c || f->ctx != c) return fail(c, LSDG_EINVAL, "objects belong to another context");
  if (startLevel < 0 || startLevel >= PYR || finalLevel < 0 || finalLevel > startLevel)
    return fail(c, LSDG_EINVAL, "level range");
  if (f->w[0] != r->kf->w[0] || f->h[0] != r->kf->h[0]) return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
  if (!f->hasDepth) return fail(c, LSDG_ESTATE, "Sim3 tracking needs a frame with depth");
  const lsdg_params& P = c->P;
  lsdg_track_stats st;
  std::memset(&st, 0, sizeof st);
  Pose T = pose_inv(pose_from8(init));
  double S[NS_MAX], Sacc[NS_MAX];
  std::memset(Sacc, 0, sizeof Sacc);
  Sim3Res finalRes;
  bool upToDate = false;
  float affA = 1.f, affB = 0.f;
  double pointUsage = 0.0;
  auto diverge = [&](bool zeroH) {
    st.diverged = 1;
    st.tracking_was_good = 0;
    if (zeroH) std::memset(st.hessian, 0, sizeof st.hessian);
    if (stats) *stats = st;
    Pose I = pose_identity();
    pose_to8(I, out);
    return LSDG_DIVERGED;
  };
  for (int lvl = startLevel; lvl >= finalLevel; --lvl) {
    if (P.max_its_per_lvl[lvl] == 0) continue;
    lsdg_status s0 = ref_points(r, lvl);
    if (s0 != LSDG_OK) return s0;
    int N = r->lv[lvl].n;
    float lim = 0.5f * P.min_goodperall_pixel_absmin * (float)(f->w[lvl] * f->h[lvl]);
    s0 = sim3_eval_sync(c, r, f, lvl, T, affA, affB, S);
    if (s0 != LSDG_OK) return s0;
    pointUsage = S[SM_USAGE] / (double)N;
    if ((float)S[SM_WARPED] < lim || S[SM_WARPED] < 10) return diverge(false);
    Sim3Res lastErr = sim3_res(S);
    std::memcpy(Sacc, S, sizeof(double) * NS_SIM3);
    st.num_residual_calls[lvl]++;
    float lambda = P.lambda_initial[lvl];
    upToDate = false;
    for (int it = 0; it < P.max_its_per_lvl[lvl]; ++it) {
      double A7[49], b7[49];
      unpack_A7(Sacc, A7, b7);
      double n = 2.0 * Sacc[SM_WARPED];
      upToDate = true;
      st.num_warp_update_calls[lvl]++;
      int incTry = 0;
      while (true) {
        double A[49], b[7], inc[7];
        for (int a = 0; a < 7; ++a) {
          for (int cc = 0; cc < 7; ++cc) A[7 * a + cc] = A7[7 * a + cc] / n;
          b[a] = -b7[a] / n;
        }
        for (int a = 0; a < 7; ++a) A[8 * a] *= 1.0 + (double)lambda;
        ldlt_solve<7>(A, b, inc);
        incTry++;
        double absInc = 0;
        for (int a = 0; a < 7; ++a) absInc += inc[a] * inc[a];
        if (!(absInc >= 0 && absInc < 1)) return diverge(true);
        Pose Tn = pose_mul(sim3_exp(inc), T);
        s0 = sim3_eval_sync(c, r, f, lvl, Tn, affA, affB, S);
        if (s0 != LSDG_OK) return s0;
        pointUsage = S[SM_USAGE] / (double)N;
        if ((float)S[SM_WARPED] < lim || S[SM_WARPED] < 10) return diverge(false);
        Sim3Res error = sim3_res(S);
        st.num_residual_calls[lvl]++;
        if (error.mean < lastErr.mean) {
          T = Tn;
          std::memcpy(Sacc, S, sizeof(double) * NS_SIM3);
          upToDate = false;
          if (error.mean / lastErr.mean > (double)P.convergence_eps[lvl]) it = P.max_its_per_lvl[lvl];
          finalRes = lastErr = error;
          if (lambda <= 0.2f) lambda = 0.f;
          else lambda *= P.lambda_success_fac[lvl];
          break;
        } else {
          if (!(absInc > (double)P.step_size_min[lvl])) {
            it = P.max_its_per_lvl[lvl];
            break;
          }
          if (lambda == 0.f) lambda = 0.2f;
          else lambda *= std::pow(P.lambda_fail_fac[lvl], (float)incTry);
        }
      }
    }
  }
  if (!upToDate) {
    lsdg_status s0 = sim3_eval_sync(c, r, f, finalLevel, T, affA, affB, S);
    if (s0 != LSDG_OK) return s0;
    finalRes = sim3_res(S);
    std::memcpy(Sacc, S, sizeof(double) * NS_SIM3);
  }
  {
    double A7[49], b7[7];
    unpack_A7(Sacc, A7, b7);
    std::memcpy(st.hessian, A7, sizeof A7);
  }
  if (T.s <= 0) return diverge(false);
  st.last_residual = (float)finalRes.mean;
  st.last_depth_residual = (float)finalRes.meanD;
  st.last_photometric_residual = (float)finalRes.meanP;
  st.point_usage = (float)pointUsage;
  st.tracking_was_good = 1;
  Pose res = pose_inv(T);
  pose_to8(res, out);
  if (stats) *stats = st;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_sim3_evaluate(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int32_t level, const double pose[8],
                                          lsdg_eval* out) {
  if (!c || !r || !f || !pose || !out) return LSDG_EINVAL;
  if (level < 0 || level >= PYR) return fail(c, LSDG_EINVAL, "level out of range");
  if (!f->hasDepth) return fail(c, LSDG_ESTATE, "Sim3 evaluation needs a frame with depth");
  if (f->w[0] != r->kf->w[0] || f->h[0] != r->kf->h[0]) return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
  double S[NS_MAX];
  lsdg_status st = sim3_eval_sync(c, r, f, level, pose_from8(pose), 1.f, 0.f, S);
  if (st != LSDG_OK) return st;
  std::memset(out, 0, sizeof(*out));
  unpack_A7(S, out->A, out->b);
  out->sum_res_p = S[SM_RESP];
  out->sum_res_d = S[SM_RESD];
  out->error = S[SM_RESP] + S[SM_RESD];
  out->num_terms_p = S[SM_WARPED];
  out->num_terms_d = S[SM_NTD];
  out->n_constraints = 2.0 * S[SM_WARPED];
  out->warped_count = (int32_t)S[SM_WARPED];
  out->usage = S[SM_USAGE];
  out->sxx = S[SM_SXX]; out->syy = S[SM_SYY]; out->sx = S[SM_SX]; out->sy = S[SM_SY]; out->sw = S[SM_SW];
  return LSDG_OK;
}

// --------------------------------------------------------------- adjoint
// Device buffer of per-point inverse-depth gradients, grown on demand.
static lsdg_status adjoint_buffer(lsdg_ctx c, int n, float** out) {
  if (n > c->adjCap) {
    if (c->d_adj) (void)hipFree(c->d_adj);
    c->d_adj = nullptr;
    c->adjCap = 0;
    CK(c, hipMalloc(&c->d_adj, sizeof(float) * (size_t)n));
    c->adjCap = n;
  }
  *out = c->d_adj;
  return LSDG_OK;
}

// Shared body of the SE3 and Sim3 adjoints: point cloud, pyramid level,
// launch, result row in host-mapped memory, per-point gradient copied out.
static lsdg_status energy_grad_run(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int level, const Pose& T, bool sim3,
                                   double* R, float* grad_idepth, int max_points) {
  if (level < 0 || level >= PYR) return fail(c, LSDG_EINVAL, "level out of range");
  if (r->ctx != c || f->ctx != c) return fail(c, LSDG_EINVAL, "objects belong to another context");
  if (f->w[0] != r->kf->w[0] || f->h[0] != r->kf->h[0]) return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
  if (sim3 && !f->hasDepth) return fail(c, LSDG_ESTATE, "the Sim3 adjoint needs a frame with depth");
  lsdg_status st = ref_points(r, level);
  if (st != LSDG_OK) return st;
  const PointsLevel& pl = r->lv[level];
  if (grad_idepth && max_points < pl.n) return fail(c, LSDG_EINVAL, "grad_idepth buffer too small");
  const int ns = sim3 ? NS_ADJ_SIM3 : NS_ADJ;
  std::memset(R, 0, sizeof(double) * ns);
  if (pl.n == 0) return LSDG_OK;
  st = frame_require(f, fd_track(f, level, level));
  if (st != LSDG_OK) return st;
  float* d_g = nullptr;
  st = adjoint_buffer(c, pl.n, &d_g);
  if (st != LSDG_OK) return st;
  TrackKParams kp = track_kparams(c->P, sim3);
  kp.use_esm = 0;  // the exact derivative samples the frame's own gradient
  KTimer kt = prof_timer(c);
  hipError_t e = sim3 ? launch_energy_grad_sim3(c, pl, f, level, pose_xf(T, true), kp, d_g, c->d_result, &kt)
                      : launch_energy_grad(c, pl, f, level, pose_xf(T, false), kp, d_g, c->d_result, &kt);
  if (e != hipSuccess) return fail(c, LSDG_EDEVICE, "energy_grad launch", e);
  prof_timer_done(c, LSDG_KERNEL_ADJOINT, kt, adjoint_bytes(pl.n, f, level, sim3));
  if (grad_idepth) {
    e = hipMemcpyAsync(grad_idepth, d_g, sizeof(float) * pl.n, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return fail(c, LSDG_EDEVICE, "energy_grad download", e);
  }
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LSDG_EDEVICE, "energy_grad", e);
  std::memcpy(R, c->h_result, sizeof(double) * ns);
  return LSDG_OK;
}

// Algorithmic bytes of one adjoint evaluation: the tracker evaluation's
// (every point record and the level's image; Sim3 adds the depth terms) plus
// the 4-byte dE/dd written per point.
double lsdg::adjoint_bytes(int n, lsdg_frame f, int lvl, bool sim3) {
  return eval_bytes(n, f, lvl, sim3) + 4.0 * n;
}

extern "C" lsdg_status lsdg_energy_grad(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int32_t level, const double pose[7],
                                        double* energy, double grad_xi[6], float* grad_idepth, int32_t max_points) {
  if (!c || !r || !f || !pose) return LSDG_EINVAL;
  double R[NS_MAX];
  lsdg_status st = energy_grad_run(c, r, f, level, pose_from7(pose), false, R, grad_idepth, max_points);
  if (st != LSDG_OK) return st;
  if (energy) *energy = R[6];
  if (grad_xi)
    for (int a = 0; a < 6; ++a) grad_xi[a] = R[a];
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_energy_grad_sim3(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int32_t level,
                                             const double pose[8], double* energy, double grad_xi[7],
                                             float* grad_idepth, int32_t max_points) {
  if (!c || !r || !f || !pose) return LSDG_EINVAL;
  double R[NS_MAX];
  lsdg_status st = energy_grad_run(c, r, f, level, pose_from8(pose), true, R, grad_idepth, max_points);
  if (st != LSDG_OK) return st;
  if (energy) *energy = 0.5 * (R[AS_EP] + R[AS_ED]);
  if (grad_xi)
    for (int a = 0; a < 7; ++a) grad_xi[a] = R[AS_G + a];
  return LSDG_OK;
}

// The SE3 adjoint with no device allocation: its reduction rows, counter,
// merged row and per-point staging live in a file-scope device array
// (track.hip), so the point cloud and the frame's pyramid level it reads must
// already exist (building them allocates).
extern "C" lsdg_status lsdg_energy_grad_noalloc(lsdg_ctx c, lsdg_ref r, lsdg_frame f, int32_t level,
                                                const double pose[7], double* energy, double grad_xi[6],
                                                float* grad_idepth, int32_t max_points) {
  if (!c || !r || !f || !pose) return LSDG_EINVAL;
  if (level < 0 || level >= PYR) return fail(c, LSDG_EINVAL, "level out of range");
  if (r->ctx != c || f->ctx != c) return fail(c, LSDG_EINVAL, "objects belong to another context");
  if (f->w[0] != r->kf->w[0] || f->h[0] != r->kf->h[0]) return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
  if (!r->kf->hasDepth || r->depthVersion != r->kf->depthVersion || r->num[level] < 0)
    return fail(c, LSDG_ESTATE, "point cloud of this level not built (lsdg_ref_make_point_cloud first)");
  const unsigned need = fd_track(f, level, level);
  if ((f->valid & need) != need)
    return fail(c, LSDG_ESTATE, "frame pyramid level not built (lsdg_frame_build_pyramid first)");
  const PointsLevel& pl = r->lv[level];
  if (pl.n > NOALLOC_MAX_POINTS)
    return fail(c, LSDG_EINVAL, "more points than the allocation-free adjoint holds (1048576)");
  if (grad_idepth && max_points < pl.n) return fail(c, LSDG_EINVAL, "grad_idepth buffer too small");
  double R[NS_ADJ] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (pl.n > 0) {
    TrackKParams kp = track_kparams(c->P, false);
    kp.use_esm = 0;
    hipError_t e = launch_energy_grad_noalloc(c, pl, f, level, pose_xf(pose_from7(pose), false), kp, R, grad_idepth);
    if (e != hipSuccess) return fail(c, LSDG_EDEVICE, "energy_grad_noalloc", e);
  }
  if (energy) *energy = R[6];
  if (grad_xi)
    for (int a = 0; a < 6; ++a) grad_xi[a] = R[a];
  return LSDG_OK;
}
