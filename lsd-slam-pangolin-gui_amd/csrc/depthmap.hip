// DepthMap (semi-dense inverse-depth filter) on gfx950, one thread per pixel.
//
//   k_observe_*    DepthMap::observeDepthRow -> observeDepthCreate / Update ->
//                  makeAndCheckEPL -> doLineStereo, in place (each pixel only
//                  touches its own hypothesis); writes a per-pixel outcome code.
//                  A gate pass compacts the pixels that search an epipolar
//                  line into a work list; a stereo pass runs one thread per item.
//   k_regularize   regularizeDepthMap(removeOcclusions, validityTH), 5x5 gather
//                  from a 16x16 tile plus its 2-pixel halo in LDS.
//   k_fill_regularize  regularizeDepthMapFillHoles, regularizeDepthMap(false)
//                  and optionally Frame::setDepth in one pass over a 4-pixel
//                  halo; the reference's integral image of validity is
//                  replaced by the equal integer 5x5 box sum.
//   k_set_depth    Frame::setDepth (level 0; the pyramid follows).
//   k_prop_*       propagateDepth as a deterministic scatter: project every
//                  source, count per destination, scan, bucket, then one
//                  thread per destination folds its sources in raster order,
//                  which is the order the reference's sequential loop uses.
//   k_mean_idepth  createKeyFrame's mean inverse depth (FP64, deterministic).
//
// The hypothesis maps rotate through three buffers: `cur` and `oth` hold what
// the reference's current and other maps hold, the third is spare.  A pass
// reads `cur` and writes every pixel of `oth` (copy-through for untouched
// pixels), then the two swap, as the reference's memcpy-then-update passes
// leave them.  The fused fill + regularize pass writes the filled map (the
// reference's other map afterwards: propagateDepth reuses it with only the
// flags wiped, so its stale fields stay visible) and the regularized map
// (the new current one) into the two buffers it does not read.  Every per-pixel float expression keeps
// the CPU restatement's operation order (compiled without FP contraction), so
// masks, codes and values match it bit for bit.
#include "lsdg_internal.h"

namespace lsdg {

typedef __attribute__((address_space(1))) const float GlobalFloat;

__device__ inline float dinterp(const float* __restrict__ m, float x, float y, int w) {
  int ix = (int)x, iy = (int)y;
  float dx = x - (float)ix, dy = y - (float)iy;
  float dxdy = dx * dy;
  GlobalFloat* bp = (GlobalFloat*)(m + ix + iy * w);
  return dxdy * bp[1 + w] + (dy - dxdy) * bp[w] + (dx - dxdy) * bp[1] + (1.0f - dx - dy + dxdy) * bp[0];
}

// getInterpolatedElement42 on (gx, gy, I, 0) gradients.
__device__ inline float2 dinterp42(const float4* __restrict__ g, float x, float y, int w) {
  int ix = (int)x, iy = (int)y;
  float dx = x - (float)ix, dy = y - (float)iy;
  float dxdy = dx * dy;
  const float4* bp = g + ix + iy * w;
  float4 a = bp[1 + w], b = bp[w], c = bp[1], d = bp[0];
  float w11 = dxdy, w01 = dy - dxdy, w10 = dx - dxdy, w00 = 1.0f - dx - dy + dxdy;
  return make_float2(w11 * a.x + w01 * b.x + w10 * c.x + w00 * d.x, w11 * a.y + w01 * b.y + w10 * c.y + w00 * d.y);
}

// The four texels dinterp reads at (x, y), as two 8-byte loads; the index is
// clamped to the image so that positions outside it (never blended) stay in
// bounds, and is unchanged for every position dinterp may read.
// The image pointer reaches the kernel through device memory (a reference
// frame's descriptor), so it is cast to the global address space: generic
// (flat) loads would make every wait drain all outstanding loads.
typedef float Texel2 __attribute__((ext_vector_type(2), aligned(4)));
typedef __attribute__((address_space(1))) const Texel2* GlobalTexel2;
struct Quad {
  Texel2 t, b;  // (I(ix, iy), I(ix + 1, iy)), (I(ix, iy + 1), I(ix + 1, iy + 1))
};
__device__ inline Quad fetch_quad(const float* __restrict__ m, float x, float y, int w, int h) {
  // the linear index, clamped to [0, w (h - 1) - 2]: unchanged for
  // 0 <= ix <= w - 2, 0 <= iy <= h - 2 (every blended position; sizes are
  // at most 16384, so the 24-bit product is exact there), and byte offsets
  // in 32 bits (an image is far below 4 GiB)
  const unsigned lin = (unsigned)(int)x + __umul24((unsigned)(int)y, (unsigned)w);
  const unsigned off = (unsigned)min(max((int)lin, 0), w * (h - 1) - 2);
  const char* base = (const char*)m;
  const unsigned top = off * 4u, bottom = top + (unsigned)w * 4u;
  Quad q;
  q.t = *(GlobalTexel2)(base + top);
  q.b = *(GlobalTexel2)(base + bottom);
  return q;
}
// dinterp's bilinear blend of fetched texels (same expression and order).
__device__ inline float blend_quad(const Quad& q, float x, float y) {
  // truncf(x) == (float)(int)x over the int range
  float dx = x - truncf(x), dy = y - truncf(y);
  float dxdy = dx * dy;
  return dxdy * q.b.y + (dy - dxdy) * q.b.x + (dx - dxdy) * q.t.y + (1.0f - dx - dy + dxdy) * q.t.x;
}

// The quad fetch through a buffer resource over an image that is the same
// for the whole wave.  The hardware's range check returns zeros for byte
// offsets at or past the image's end (negative offsets wrap there too), so
// no clamp is needed: only lead fetches past a line's end, which are never
// blended, leave [0, w (h - 1) - 2].
struct TexBuf {
  __amdgpu_buffer_rsrc_t r;
  unsigned rowBytes;
};
__device__ inline TexBuf tex_buf(const float* img, int w, int h) {
  // word 3 of the descriptor: 32-bit data format, as for any raw access
  return TexBuf{__builtin_amdgcn_make_buffer_rsrc((void*)img, (short)0, w * h * 4, 0x00020000), (unsigned)w * 4u};
}
__device__ inline Quad fetch_quad(const TexBuf& b, float x, float y, int w) {
  const unsigned top = ((unsigned)(int)x + __umul24((unsigned)(int)y, (unsigned)w)) * 4u;
  Quad q;
  q.t = __builtin_bit_cast(Texel2, __builtin_amdgcn_raw_buffer_load_b64(b.r, top, 0, 0));
  q.b = __builtin_bit_cast(Texel2, __builtin_amdgcn_raw_buffer_load_b64(b.r, top + b.rowBytes, 0, 0));
  return q;
}

#ifndef LSDG_STEREO_AHEAD
#define LSDG_STEREO_AHEAD 4
#endif
// Steps a texel fetch runs ahead of its use in the epipolar search.  With
// every fetch pinned to its step, 4 measured best (2: 33.7 us for the
// 640x480 stereo pass, 4: 31.6, 8: 32.3, 12: 36.2): a longer lead keeps more
// cache lines of the wave's 64 lines in flight, and the L1 keeps fewer of the
// lines the next steps read again.
constexpr int STEREO_AHEAD = LSDG_STEREO_AHEAD;
static_assert(STEREO_AHEAD >= 1, "at least one fetch in flight");

__device__ inline float unzero(float v) {
  return v < 0.0f ? (v > -1e-10f ? -1e-10f : v) : (v < 1e-10f ? 1e-10f : v);
}

struct Hyp {
  bool valid;
  int bl;
  float nsid;
  int val;
  float4 dv;
};

__device__ inline Hyp load_hyp(const HypBuf& b, int i) {
  Hyp h;
  h.valid = b.valid[i] != 0;
  h.bl = b.bl[i];
  h.nsid = b.nsid[i];
  h.val = b.val[i];
  h.dv = b.dv[i];
  return h;
}
__device__ inline void store_hyp(const HypBuf& b, int i, const Hyp& h) {
  b.valid[i] = h.valid ? 1 : 0;
  b.bl[i] = h.bl;
  b.nsid[i] = h.nsid;
  b.val[i] = h.val;
  b.dv[i] = h.dv;
}
// DepthMapPixelHypothesis(idepth, var, validity): smoothed values -1.
__device__ inline Hyp make_hyp(float id, float var, int val) {
  Hyp h;
  h.valid = true;
  h.bl = 0;
  h.nsid = 0.0f;
  h.val = val;
  h.dv = make_float4(id, var, -1.0f, -1.0f);
  return h;
}

// ------------------------------------------------------------ stereo
struct ObsCtx {
  DepthGeom g;
  const float* kfImg;
  const float4* kfGrad;
};

// gx, gy: the keyframe image's central differences at (x, y),
// I(x+1, y) - I(x-1, y) and I(x, y+1) - I(x, y-1).
__device__ inline bool make_and_check_epl(const DepthGeom& g, const lsdg_params& P, int x, int y, float gx, float gy,
                                          const StereoRef& ref, float* pepx, float* pepy) {
  float epx = -g.fx * ref.t_t2o[0] + ref.t_t2o[2] * ((float)x - g.cx);
  float epy = -g.fy * ref.t_t2o[1] + ref.t_t2o[2] * ((float)y - g.cy);
  if (isnan(epx + epy)) return false;
  float eplLengthSquared = epx * epx + epy * epy;
  if (eplLengthSquared < P.min_epl_length_squared) return false;
  float eplGradSquared = gx * epx + gy * epy;
  eplGradSquared = eplGradSquared * eplGradSquared / eplLengthSquared;
  if (eplGradSquared < P.min_epl_grad_squared) return false;
  if (eplGradSquared / (gx * gx + gy * gy) < P.min_epl_angle_squared) return false;
  float fac = P.gradient_sample_dist / sqrtf(eplLengthSquared);
  *pepx = epx * fac;
  *pepy = epy * fac;
  return true;
}

// Outcome of the epipolar search: the smallest error (first step on ties)
// with its neighbours' errors and error-vector dot products and its
// position, and the smallest error of the other steps (first step on ties).
// Empty entries hold (INFINITY, INT_MAX): the reference only records errors
// below INFINITY.
struct Scan {
  float best, second;
  int ibest, isecond;
  float errPre, diffPre, errPost, diffPost;
  float bx, by;
};

// Advances the position chain cp_{j+1} = cp_j + inc (float additions, one
// at a time, as the reference's loop does) from step j to step `to`.
__device__ inline void walk(float& x, float& y, int& j, int to, float ix, float iy) {
  while (j + 4 <= to) {
    x += ix;
    x += ix;
    x += ix;
    x += ix;
    y += iy;
    y += iy;
    y += iy;
    y += iy;
    j += 4;
  }
  while (j < to) {
    x += ix;
    y += iy;
    ++j;
  }
}

// Samples V_{m0} .. V_{m0+N-1} of the epipolar search (V_m, m >= 2, is
// sampled at cp_{m-2} + 2 inc; V_{-2} .. V_1 are the initial samples), with
// the walker (x, y) at cp_j, j <= max(m0 - 2, 0), on entry and at
// cp_{m0+N-3} on return (when m0 + N - 3 >= j).  The positions come first and
// every texel fetch is issued before the first blend, so the N samples cost
// one round trip to memory, not N.  With bx set, (*bx, *by) receives
// cp_{m0+N-4}: the position of the step whose newest sample is V_{m0+N-2}.
// Fetches for m <= 1 are clamped to the image and not used.
template <int N>
__device__ inline void window_samples(const float* __restrict__ img, int w, int h, float ix, float iy, int m0,
                                      float i_m2, float i_m1, float i_0, float i_p1, float* out, float& x, float& y,
                                      int& j, float* bx = nullptr, float* by = nullptr) {
  float sx[N], sy[N];
#pragma unroll
  for (int t = 0; t < N; ++t) {
    walk(x, y, j, m0 + t - 2, ix, iy);  // at most one step after the first
    sx[t] = x + 2.0f * ix;
    sy[t] = y + 2.0f * iy;
    if (bx && t == N - 2) {
      *bx = x;
      *by = y;
    }
  }
  Quad q[N];
#pragma unroll
  for (int t = 0; t < N; ++t) q[t] = fetch_quad(img, sx[t], sy[t], w, h);
  // every fetch's value is taken here, so none is deferred into the
  // per-sample selects below (one wait for all N, not one per branch)
#pragma unroll
  for (int t = 0; t < N; ++t) asm volatile("" ::"v"(q[t].t.x), "v"(q[t].t.y), "v"(q[t].b.x), "v"(q[t].b.y));
#pragma unroll
  for (int t = 0; t < N; ++t) {
    const int m = m0 + t;
    out[t] = m >= 2 ? blend_quad(q[t], sx[t], sy[t]) : m == 1 ? i_p1 : m == 0 ? i_0 : m == -1 ? i_m1 : i_m2;
  }
}

// A step's matching error: the sum of the squared differences of its five
// samples (newest first) and the keyframe's, in the reference's order.
__device__ inline float stereo_error(float v_p2, float v_p1, float v, float v_m1, float v_m2, float r_p2, float r_p1,
                                     float r, float r_m1, float r_m2) {
  float e, ee = 0.0f;
  e = v_p2 - r_p2; ee += e * e;
  e = v_p1 - r_p1; ee += e * e;
  e = v - r; ee += e * e;
  e = v_m1 - r_m1; ee += e * e;
  e = v_m2 - r_m2; ee += e * e;
  return ee;
}
__device__ inline float stereo_sq(const float* e) {
  float ee = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) ee += e[i] * e[i];
  return ee;
}
// The dot product of two steps' error vectors, as the reference forms it.
__device__ inline float stereo_dot(const float* a, const float* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4];
}

// doLineStereo: the matching error (>= 0) or a code -1..-4.  With UNI the
// reference image is the same for the whole wave and the search loop fetches
// through a buffer resource.
template <bool UNI>
__device__ float line_stereo(const ObsCtx& o, const lsdg_params& P, float u, float v, float epxn, float epyn,
                             float min_idepth, float prior_idepth, float max_idepth, const StereoRef& ref,
                             float& r_idepth, float& r_var, float& r_epl) {
  const int width = o.g.w, height = o.g.h;
  const float* kfImg = o.kfImg;
  const float* refImg = ref.img;
  float KinvP[3] = {o.g.fxi * u + o.g.cxi, o.g.fyi * v + o.g.cyi, 1.0f};
  float pInf[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) pInf[i] = ref.KR[3 * i] * KinvP[0] + ref.KR[3 * i + 1] * KinvP[1] + ref.KR[3 * i + 2] * KinvP[2];
  float pReal2 = pInf[2] / prior_idepth + ref.Kt[2];
  float rescale = pReal2 * prior_idepth;
  float firstX = u - 2.0f * epxn * rescale, firstY = v - 2.0f * epyn * rescale;
  float lastX = u + 2.0f * epxn * rescale, lastY = v + 2.0f * epyn * rescale;
  if (firstX <= 0.0f || firstX >= (float)(width - 2) || firstY <= 0.0f || firstY >= (float)(height - 2) ||
      lastX <= 0.0f || lastX >= (float)(width - 2) || lastY <= 0.0f || lastY >= (float)(height - 2))
    return -1.0f;
  if (!(rescale > 0.7f && rescale < 1.4f)) return -1.0f;
  float rv_p1 = dinterp(kfImg, u + epxn * rescale, v + epyn * rescale, width);
  float rv_m1 = dinterp(kfImg, u - epxn * rescale, v - epyn * rescale, width);
  float rv = dinterp(kfImg, u, v, width);
  float rv_m2 = dinterp(kfImg, u - 2.0f * epxn * rescale, v - 2.0f * epyn * rescale, width);
  float rv_p2 = dinterp(kfImg, u + 2.0f * epxn * rescale, v + 2.0f * epyn * rescale, width);

  float pClose[3], pFar[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) pClose[i] = pInf[i] + ref.Kt[i] * max_idepth;
  if (pClose[2] < 0.001f) {
    max_idepth = (0.001f - pInf[2]) / ref.Kt[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) pClose[i] = pInf[i] + ref.Kt[i] * max_idepth;
  }
  {
    float zc = pClose[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) pClose[i] = pClose[i] / zc;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) pFar[i] = pInf[i] + ref.Kt[i] * min_idepth;
  if (pFar[2] < 0.001f || max_idepth < min_idepth) return -1.0f;
  {
    float zf = pFar[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) pFar[i] = pFar[i] / zf;
  }
  if (isnan(pFar[0] + pClose[0]) || isnan(pFar[1] + pClose[1])) return -4.0f;

  float incx = pClose[0] - pFar[0];
  float incy = pClose[1] - pFar[1];
  float eplLength = sqrtf(incx * incx + incy * incy);
  if (!(eplLength > 0.0f) || isinf(eplLength)) return -4.0f;
  const float crop = P.max_epl_length_crop;
  if (eplLength > crop) {
    pClose[0] = pFar[0] + incx * crop / eplLength;
    pClose[1] = pFar[1] + incy * crop / eplLength;
  }
  incx *= P.gradient_sample_dist / eplLength;
  incy *= P.gradient_sample_dist / eplLength;
  pFar[0] -= incx;
  pFar[1] -= incy;
  pClose[0] += incx;
  pClose[1] += incy;
  const float mincrop = P.min_epl_length_crop;
  if (eplLength < mincrop) {
    float pad = (mincrop - eplLength) / 2.0f;
    pFar[0] -= incx * pad;
    pFar[1] -= incy * pad;
    pClose[0] += incx * pad;
    pClose[1] += incy * pad;
  }
  const float B = P.sample_point_to_border;
  const float Wb = (float)width - B, Hb = (float)height - B;
  if (pFar[0] <= B || pFar[0] >= Wb || pFar[1] <= B || pFar[1] >= Hb) return -1.0f;
  if (pClose[0] <= B || pClose[0] >= Wb || pClose[1] <= B || pClose[1] >= Hb) {
    if (pClose[0] <= B) {
      float toAdd = (B - pClose[0]) / incx;
      pClose[0] += toAdd * incx;
      pClose[1] += toAdd * incy;
    } else if (pClose[0] >= Wb) {
      float toAdd = (Wb - pClose[0]) / incx;
      pClose[0] += toAdd * incx;
      pClose[1] += toAdd * incy;
    }
    if (pClose[1] <= B) {
      float toAdd = (B - pClose[1]) / incy;
      pClose[0] += toAdd * incx;
      pClose[1] += toAdd * incy;
    } else if (pClose[1] >= Hb) {
      float toAdd = (Hb - pClose[1]) / incy;
      pClose[0] += toAdd * incx;
      pClose[1] += toAdd * incy;
    }
    float fincx = pClose[0] - pFar[0], fincy = pClose[1] - pFar[1];
    float newLen = sqrtf(fincx * fincx + fincy * fincy);
    if (pClose[0] <= B || pClose[0] >= Wb || pClose[1] <= B || pClose[1] >= Hb || newLen < 8.0f) return -1.0f;
  }

  // Epipolar search, in two passes.  The search keeps only what decides the
  // match: each step's error and the first and second minima (the
  // reference's selects, in step order).  The best step's other records (the
  // errors and error-vector dot products of its neighbours, its position) are
  // then recomputed from the samples around it, which the same chain of float
  // additions gives again (walk), so they equal what the reference's loop
  // records when it meets the best.
  const float cp0x = pFar[0], cp0y = pFar[1];
  const float init_m2 = dinterp(refImg, cp0x - 2.0f * incx, cp0y - 2.0f * incy, width);
  const float init_m1 = dinterp(refImg, cp0x - incx, cp0y - incy, width);
  const float init_0 = dinterp(refImg, cp0x, cp0y, width);
  const float init_p1 = dinterp(refImg, cp0x + incx, cp0y + incy, width);
  const int maxSteps = P.stereo_max_steps;
  const bool negx = incx < 0.0f, negy = incy < 0.0f;
  const float closeX = pClose[0], closeY = pClose[1];
  Scan sc;
  sc.best = INFINITY;
  sc.second = INFINITY;
  sc.ibest = INT_MAX;
  sc.isecond = INT_MAX;
  sc.errPre = NAN;
  sc.diffPre = NAN;
  sc.errPost = NAN;
  sc.diffPost = NAN;
  sc.bx = -1.0f;
  sc.by = -1.0f;
  float val_m2 = init_m2, val_m1 = init_m1, val = init_0, val_p1 = init_p1;
  float cpx = cp0x, cpy = cp0y;
  int loop = 0;
  // the step bound, held in a vector register: as a scalar the compiler
  // reloads it from the kernel arguments at every step (a scalar-memory
  // round trip per step)
  int stepLimit;
  asm("v_mov_b32 %0, %1" : "=v"(stepLimit) : "s"(maxSteps));
  {
    float lx = cpx, ly = cpy;
    const TexBuf tb = tex_buf(refImg, width, height);
    const auto ring_fetch = [&](float fx, float fy) {
      if constexpr (UNI) return fetch_quad(tb, fx, fy, width);
      else return fetch_quad(refImg, fx, fy, width, height);
    };
    // The texels of the next STEREO_AHEAD steps' samples are in flight
    // while a step is evaluated: a lead copy of the position chain fetches
    // each sample STEREO_AHEAD steps early.  Lead fetches past the line's
    // end are clamped to the image and never used.
    Quad ring[STEREO_AHEAD];
#pragma unroll
    for (int k = 0; k < STEREO_AHEAD; ++k) {
      ring[k] = ring_fetch(lx + 2.0f * incx, ly + 2.0f * incy);
      lx += incx;
      ly += incy;
    }
    // Lanes leave the loop together (a wave-uniform exit every
    // STEREO_AHEAD steps; finished lanes idle), and every step issues its
    // fetch whether or not the lane is still searching: with per-lane exits
    // the wave's outstanding-load count would differ between paths and the
    // compiler would wait for all fetches at every step.
    bool alive = true;
    for (;;) {
#pragma unroll
      for (int k = 0; k < STEREO_AHEAD; ++k) {
        // the reference's while condition, without short-circuit branches
        const bool go = ((negx == (cpx > closeX)) & (negy == (cpy > closeY))) | (loop == 0);
        const bool act = alive & go & (loop < stepLimit);
        const float val_p2 = blend_quad(ring[k], cpx + 2.0f * incx, cpy + 2.0f * incy);
        ring[k] = ring_fetch(lx + 2.0f * incx, ly + 2.0f * incy);
        lx += incx;
        ly += incy;
        const float ee = stereo_error(val_p2, val_p1, val, val_m1, val_m2, rv_p2, rv_p1, rv, rv_m1, rv_m2);
        const bool nb = act & (ee < sc.best);          // new best
        const bool ns = act & !nb & (ee < sc.second);  // new second
        sc.second = nb ? sc.best : (ns ? ee : sc.second);
        sc.isecond = nb ? sc.ibest : (ns ? loop : sc.isecond);
        sc.best = nb ? ee : sc.best;
        sc.ibest = nb ? loop : sc.ibest;
        alive = act;
        val_m2 = val_m1;
        val_m1 = val;
        val = val_p1;
        val_p1 = val_p2;
        // the position and count of a lane that has left the line are never
        // read again, so they advance regardless
        cpx += incx;
        cpy += incy;
        ++loop;
        // Keeps each step's fetch in its step: the next step's position
        // comes out of a barrier that no memory access crosses.  Without it
        // the compiler either merges a ring slot's loop-carried loaded value
        // into one load of a loop-carried address, issued where the value is
        // used, or gathers the group's fetches after its last step: either
        // way every group of STEREO_AHEAD steps starts with a round trip to
        // memory.
        asm volatile("" : "+v"(cpx), "+v"(cpy)::"memory");
      }
      if (!__any(alive)) break;
    }
  }
  if (sc.best < INFINITY) {
    // Samples V_{b-3} .. V_{b+3} around the best step b (V_m, m >= 2, at
    // cp_{m-2} + 2 inc; V_{-2} .. V_1 the initial samples): the error
    // vectors of steps b - 1, b and b + 1 (step k's window is V_{k-2} ..
    // V_{k+2}), and the positions cp_b and cp_{b+1}.
    const int b = sc.ibest;
    float V[7];
    float px = cp0x, py = cp0y, bpx, bpy;
    int pj = 0;
    window_samples<7>(refImg, width, height, incx, incy, b - 3, init_m2, init_m1, init_0, init_p1, V, px,
                      py, pj, &bpx, &bpy);
    // (the walker is now at cp_{b+1}, and (bpx, bpy) = cp_b)
    const float ePrev[5] = {V[4] - rv_p2, V[3] - rv_p1, V[2] - rv, V[1] - rv_m1, V[0] - rv_m2};
    const float eBest[5] = {V[5] - rv_p2, V[4] - rv_p1, V[3] - rv, V[2] - rv_m1, V[1] - rv_m2};
    const float eNext[5] = {V[6] - rv_p2, V[5] - rv_p1, V[4] - rv, V[3] - rv_m1, V[2] - rv_m2};
    // the step after the best, if the line went on (the search was on it at b)
    const bool post = (b + 1 < maxSteps) & (negx == (px > closeX)) & (negy == (py > closeY));
    sc.errPre = b == 0 ? -1.0f : stereo_sq(ePrev);
    sc.diffPre = b == 0 ? NAN : stereo_dot(eBest, ePrev);
    sc.errPost = post ? stereo_sq(eNext) : -1.0f;
    sc.diffPost = post ? stereo_dot(eNext, eBest) : -1.0f;
    sc.bx = bpx;
    sc.by = bpy;
  }
  float best_err = sc.best;
  const float second_err = sc.second;
  const int loopBest = sc.best < INFINITY ? sc.ibest : -1;
  const int loopSecond = sc.second < INFINITY ? sc.isecond : -1;
  const float best_errPre = sc.errPre, best_DiffPre = sc.diffPre;
  const float best_errPost = sc.errPost, best_DiffPost = sc.diffPost;
  float best_x = sc.bx, best_y = sc.by;
  if (best_err > 4.0f * P.max_error_stereo) return -3.0f;
  if ((float)abs(loopBest - loopSecond) > 1.0f && P.min_distance_error_stereo * best_err > second_err) return -2.0f;

  bool didSubpixel = false;
  if (P.use_subpixel_stereo) {
    float gradPre_pre = -(best_errPre - best_DiffPre);
    float gradPre_this = +(best_err - best_DiffPre);
    float gradPost_this = -(best_err - best_DiffPost);
    float gradPost_post = +(best_errPost - best_DiffPost);
    bool interpPost = false, interpPre = false;
    if (P.stereo_subpixel_oob_guard && (best_errPre < 0.0f || best_errPost < 0.0f)) {
    } else if ((gradPost_this < 0.0f) ^ (gradPre_this < 0.0f)) {
    } else if ((gradPre_pre < 0.0f) ^ (gradPre_this < 0.0f)) {
      if (!((gradPost_post < 0.0f) ^ (gradPost_this < 0.0f))) interpPre = true;
    } else if ((gradPost_post < 0.0f) ^ (gradPost_this < 0.0f)) {
      interpPost = true;
    }
    if (interpPre) {
      float d = gradPre_this / (gradPre_this - gradPre_pre);
      best_x -= d * incx;
      best_y -= d * incy;
      best_err = best_err - 2.0f * d * gradPre_this - (gradPre_pre - gradPre_this) * d * d;
      didSubpixel = true;
    } else if (interpPost) {
      float d = gradPost_this / (gradPost_this - gradPost_post);
      best_x += d * incx;
      best_y += d * incy;
      best_err = best_err + 2.0f * d * gradPost_this + (gradPost_post - gradPost_this) * d * d;
      didSubpixel = true;
    }
  }
  float sampleDist = P.gradient_sample_dist * rescale;
  float gradAlongLine = 0.0f, tmp;
  tmp = rv_p2 - rv_p1; gradAlongLine += tmp * tmp;
  tmp = rv_p1 - rv; gradAlongLine += tmp * tmp;
  tmp = rv - rv_m1; gradAlongLine += tmp * tmp;
  tmp = rv_m1 - rv_m2; gradAlongLine += tmp * tmp;
  gradAlongLine /= sampleDist * sampleDist;
  if (best_err > P.max_error_stereo + sqrtf(gradAlongLine) * 20.0f) return -3.0f;

  float idnew, alpha;
  const float* t = ref.t_o2t;
  if (incx * incx > incy * incy) {
    float oldX = o.g.fxi * best_x + o.g.cxi;
    float nominator = oldX * t[2] - t[0];
    float dot0 = KinvP[0] * ref.Rrow[0] + KinvP[1] * ref.Rrow[1] + KinvP[2] * ref.Rrow[2];
    float dot2 = KinvP[0] * ref.Rrow[6] + KinvP[1] * ref.Rrow[7] + KinvP[2] * ref.Rrow[8];
    idnew = (dot0 - oldX * dot2) / nominator;
    alpha = incx * o.g.fxi * (dot0 * t[2] - dot2 * t[0]) / (nominator * nominator);
  } else {
    float oldY = o.g.fyi * best_y + o.g.cyi;
    float nominator = oldY * t[2] - t[1];
    float dot1 = KinvP[0] * ref.Rrow[3] + KinvP[1] * ref.Rrow[4] + KinvP[2] * ref.Rrow[5];
    float dot2 = KinvP[0] * ref.Rrow[6] + KinvP[1] * ref.Rrow[7] + KinvP[2] * ref.Rrow[8];
    idnew = (dot1 - oldY * dot2) / nominator;
    alpha = incy * o.g.fyi * (dot1 * t[2] - dot2 * t[1]) / (nominator * nominator);
  }
  if (idnew < 0.0f && !P.allow_negative_idepths) return -2.0f;
  float photoDispError = 4.0f * P.camera_pixel_noise2 / (gradAlongLine + P.division_eps);
  float trackingErrorFac = 0.25f * (1.0f + ref.initialTrackedResidual);
  float2 g2 = dinterp42(o.kfGrad, u, v, width);
  float geo = (g2.x * epxn + g2.y * epyn) + P.division_eps;
  geo = trackingErrorFac * trackingErrorFac * (g2.x * g2.x + g2.y * g2.y) / (geo * geo);
  r_var = alpha * alpha * ((didSubpixel ? 0.05f : 0.5f) * sampleDist * sampleDist + geo + photoDispError);
  r_idepth = idnew;
  r_epl = eplLength;
  return best_err;
}

// Reference frames of one update: in the kernel arguments when they fit
// (the common case of a few frames), else in device memory.
constexpr int OBS_INLINE_REFS = 4;
constexpr int OBS_INLINE_BYID = 32;

struct ObsArgs {
  DepthGeom g;
  HypBuf cur;
  uint8_t* obs;
  const float* kfImg;
  const float* kfMaxg;
  const float4* kfGrad;
  const StereoRef* refs;  // device copies when not inline
  const int* refById;
  StereoRef inlRefs[OBS_INLINE_REFS];
  int inlById[OBS_INLINE_BYID];
  int nRefById;
  int refByIdOffset;
  int oldestIdx, newestIdx;
  int kfReactivated;
  int kfId;
  float numTracked, numMapped;
};

template <bool INL>
__device__ inline const StereoRef& obs_ref(const ObsArgs& a, int i) {
  if constexpr (INL) return a.inlRefs[i];
  else return a.refs[i];
}
template <bool INL>
__device__ inline int obs_by_id(const ObsArgs& a, int k) {
  if constexpr (INL) return a.inlById[k];
  else return a.refById[k];
}

__device__ inline bool was_good_gate(const StereoRef& r, int x, int y, int w) {
  typedef __attribute__((address_space(1))) const uint8_t GlobalByte;
  if (r.wasGood && !((GlobalByte*)r.wasGood)[(x >> 1) + (w >> 1) * (y >> 1)]) return false;
  return true;
}

// observeDepthRow in two passes.  The gate pass runs the per-pixel tests of
// observeDepth / observeDepthCreate / observeDepthUpdate up to and including
// makeAndCheckEPL and writes the outcome of every pixel that stops there;
// pixels that go on to doLineStereo are appended to a work list.  The stereo
// pass runs one thread per work item, so waves hold only pixels that search
// an epipolar line, and creates (long lines over the full inverse-depth
// range) are not mixed with updates (short lines around the prior).  Each
// pixel still reads and writes only its own hypothesis, so the item order
// does not matter.
//
// Appends are aggregated per block, and the blocks are spread over OBS_SEGS
// list segments with one counter pair each (gate block b appends to segment
// b % OBS_SEGS): same-address device atomics serialise, so one counter for
// the whole grid would cost more than the search itself.  Segment s holds
// creates from its front and updates from its back; its capacity is the
// pixel count of its blocks (GATE_PX x 256 each), so it cannot overflow.
constexpr int OBS_SEGS = 16;
constexpr int OBS_COUNTER_STRIDE = 32;  // ints between counters: one 128-byte line each

__device__ inline int lane_rank(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// A gate block covers GATE_PX x 16 by 16 pixels: thread (tx, ty), tx < 16,
// takes pixels (16 k + tx, ty), k < GATE_PX, of its block's region, so every
// load instruction reads 16 consecutive pixels of four rows and each thread
// has GATE_PX pixels' loads in flight at once.  The shape also sets the work
// list's order, which the stereo pass feels: a wave's appends are its 16 x 4
// patches in turn, so the items of a patch are consecutive in the list and
// a stereo wave's lanes search the lines of neighbouring pixels.  Measured
// on the 640x480 update: 1, 2, 4 or 8 pixels per thread gave gate + stereo
// of 43.2, 41.4, 42.1 and 44.5 us; rows of 32 or 64 threads instead of 16
// (same pixel count) made the stereo pass 45 us instead of 32.
constexpr int GATE_PX = 2;
constexpr int GATE_TW = 16;                // threads per row of the block
constexpr int GATE_TH = 256 / GATE_TW;     // rows of the block
constexpr int GATE_W = GATE_TW * GATE_PX;  // block region width in pixels

// Pixels outside [3, w-3) x [3, h-3) get LSDG_OBS_NONE.  Appends go to the
// counter set `count`; the other set, `other`, is cleared for the next update
// (the two alternate, so no clearing pass runs between updates).
template <bool INL>
__global__ __launch_bounds__(256) void k_observe_gate(ObsArgs a, lsdg_params P, int4* __restrict__ work,
                                                      int* __restrict__ count, int* __restrict__ other, int segcap) {
  const int w = a.g.w, h = a.g.h;
  const int y = blockIdx.y * GATE_TH + (int)(threadIdx.x / GATE_TW);
  const int x0 = blockIdx.x * GATE_W + (int)(threadIdx.x % GATE_TW);
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2 * OBS_SEGS)
    other[threadIdx.x * OBS_COUNTER_STRIDE] = 0;
  const float* I = a.kfImg;
  // every per-pixel input of the thread's pixels is read up front (interior
  // pixels only, so every neighbour is in bounds): the tests below wait on
  // memory once, and once more for the reference's tracking mask
  bool has[GATE_PX], inner[GATE_PX];
  int bl[GATE_PX];
  float mg[GATE_PX], nsid[GATE_PX], gx[GATE_PX], gy[GATE_PX];
#pragma unroll
  for (int k = 0; k < GATE_PX; ++k) {
    const int x = x0 + GATE_TW * k;
    inner[k] = x >= 3 && x < w - 3 && y >= 3 && y < h - 3;
    const int idx = inner[k] ? x + y * w : w + 1;  // (an interior pixel when not inner: loads stay in bounds)
    has[k] = a.cur.valid[idx] != 0;
    bl[k] = a.cur.bl[idx];
    mg[k] = a.kfMaxg[idx];
    nsid[k] = a.cur.nsid[idx];
    gx[k] = I[idx + 1] - I[idx - 1];
    gy[k] = I[idx + w] - I[idx - w];
  }
  // the tests up to the choice of reference frame, then every pixel's read
  // of that frame's tracking mask together, then makeAndCheckEPL
  int kind[GATE_PX];  // 0 done here, 1 observeDepthCreate search, 2 observeDepthUpdate search
  int ri[GATE_PX];
  uint8_t code[GATE_PX];
  bool go[GATE_PX];
#pragma unroll
  for (int k = 0; k < GATE_PX; ++k) {
    const int x = x0 + GATE_TW * k;
    kind[k] = 0;
    ri[k] = 0;
    code[k] = LSDG_OBS_NONE;
    go[k] = false;
    if (!inner[k]) continue;  // (border pixel: LSDG_OBS_NONE; none outside the map)
    if (has[k] && mg[k] < P.min_abs_grad_decrease) {
      a.cur.valid[x + y * w] = 0;
      code[k] = LSDG_OBS_LOWGRAD_INVALIDATE;
    } else if (mg[k] < P.min_abs_grad_create || bl[k] < P.min_blacklist) {
      code[k] = LSDG_OBS_NONE;
    } else if (!has[k]) {
      ri[k] = a.kfReactivated ? a.newestIdx : a.oldestIdx;
      go[k] = true;
    } else if (!a.kfReactivated) {
      const int q = (int)nsid[k] - a.refByIdOffset;
      if (q >= a.nRefById) {
        code[k] = LSDG_OBS_SKIP_ALREADY_GOOD;
      } else {
        ri[k] = q < 0 ? a.oldestIdx : obs_by_id<INL>(a, q);
        go[k] = true;
      }
    } else {
      ri[k] = a.newestIdx;
      go[k] = true;
    }
  }
  bool good[GATE_PX];
#pragma unroll
  for (int k = 0; k < GATE_PX; ++k) good[k] = !go[k] || was_good_gate(obs_ref<INL>(a, ri[k]), x0 + GATE_TW * k, y, w);
  float epx[GATE_PX], epy[GATE_PX];
#pragma unroll
  for (int k = 0; k < GATE_PX; ++k) {
    const int x = x0 + GATE_TW * k;
    epx[k] = 0.0f;
    epy[k] = 0.0f;
    if (go[k]) {
      if (!good[k]) code[k] = LSDG_OBS_SKIP_NOT_GOOD_TRACKED;
      else if (!make_and_check_epl(a.g, P, x, y, gx[k], gy[k], obs_ref<INL>(a, ri[k]), &epx[k], &epy[k]))
        code[k] = LSDG_OBS_EPL_FAIL;
      else kind[k] = has[k] ? 2 : 1;
    }
    if (kind[k] == 0 && x < w && y < h) a.obs[x + y * w] = code[k];
  }
  // block-wide ranks: per-wave ballots, then one atomic pair per block; an
  // item's slot counts the block's earlier waves, then the wave's earlier
  // pixel columns k, then the lanes before it
  __shared__ int wcount[4][2], base[2];
  const int wave = threadIdx.x >> 6;
  unsigned long long mc[GATE_PX], mu[GATE_PX];
  int tc = 0, tu = 0;
#pragma unroll
  for (int k = 0; k < GATE_PX; ++k) {
    mc[k] = __ballot(kind[k] == 1);
    mu[k] = __ballot(kind[k] == 2);
    tc += __popcll(mc[k]);
    tu += __popcll(mu[k]);
  }
  if (__lane_id() == 0) {
    wcount[wave][0] = tc;
    wcount[wave][1] = tu;
  }
  __syncthreads();
  const int seg = (blockIdx.x + blockIdx.y * gridDim.x) % OBS_SEGS;
  if (threadIdx.x < 2) {
    const int k = threadIdx.x;
    const int tot = wcount[0][k] + wcount[1][k] + wcount[2][k] + wcount[3][k];
    base[k] = tot ? atomicAdd(&count[(2 * seg + k) * OBS_COUNTER_STRIDE], tot) : 0;
  }
  __syncthreads();
  int rc = base[0], ru = base[1];
  for (int v = 0; v < wave; ++v) {
    rc += wcount[v][0];
    ru += wcount[v][1];
  }
  int4* segw = work + (size_t)seg * segcap;
#pragma unroll
  for (int k = 0; k < GATE_PX; ++k) {
    const int x = x0 + GATE_TW * k;
    const int4 item = make_int4(x + y * w, ri[k], __float_as_int(epx[k]), __float_as_int(epy[k]));
    if (kind[k] == 1) segw[rc + lane_rank(mc[k])] = item;
    else if (kind[k] == 2) segw[segcap - 1 - (ru + lane_rank(mu[k]))] = item;
    rc += __popcll(mc[k]);
    ru += __popcll(mu[k]);
  }
}

#ifndef LSDG_STEREO_MIN_BLOCKS
#define LSDG_STEREO_MIN_BLOCKS 3
#endif
// Stereo blocks per CU the register budget is held to (3: 168 VGPRs, no spills).
constexpr int STEREO_MIN_BLOCKS = LSDG_STEREO_MIN_BLOCKS;

// observeDepthCreate item `i` of segment `seg`; creates search the whole
// inverse-depth range, so their lines are long.
template <bool INL>
__device__ inline void observe_create(const ObsArgs& a, const lsdg_params& P, const int4* __restrict__ work,
                                      int segcap, int seg, int i) {
  const int4 item = work[(size_t)seg * segcap + i];
  const int w = a.g.w;
  const int idx = item.x, x = idx % w, y = idx / w;
  const float epx = __int_as_float(item.z), epy = __int_as_float(item.w);
  // every create searches the same frame (item.y holds this index): read
  // as a launch-uniform value its image pointer and products stay scalar
  const StereoRef& ref = obs_ref<INL>(a, a.kfReactivated ? a.newestIdx : a.oldestIdx);
  ObsCtx o{a.g, a.kfImg, a.kfGrad};
  float rid, rvar, repl;
  float err =
      line_stereo<true>(o, P, (float)x, (float)y, epx, epy, 0.0f, 1.0f, 1.0f / P.min_depth, ref, rid, rvar, repl);
  uint8_t code;
  if (err == -3.0f || err == -2.0f) a.cur.bl[idx] = a.cur.bl[idx] - 1;
  if (err < 0.0f) {
    code = (uint8_t)(LSDG_OBS_STEREO_FAIL_BASE - (int)err);
  } else if (rvar > P.max_var) {
    code = LSDG_OBS_CREATE_VAR_TOO_BIG;
  } else {
    rid = unzero(rid);
    store_hyp(a.cur, idx, make_hyp(rid, rvar, P.validity_counter_initial_observe));
    code = LSDG_OBS_CREATED;
  }
  a.obs[idx] = code;
}

// observeDepthUpdate item `i` of segment `seg` (slot segcap - 1 - i), one
// lane (short lines around the prior).
template <bool INL>
__device__ inline void observe_update(const ObsArgs& a, const lsdg_params& P, const int4* __restrict__ work,
                                      int segcap, int seg, int i) {
  const int4 item = work[(size_t)seg * segcap + segcap - 1 - i];
  const int w = a.g.w;
  const int idx = item.x, x = idx % w, y = idx / w;
  const float epx = __int_as_float(item.z), epy = __int_as_float(item.w);
  const StereoRef& ref = obs_ref<INL>(a, item.y);
  ObsCtx o{a.g, a.kfImg, a.kfGrad};
  uint8_t code;
  // observeDepthUpdate
  Hyp t = load_hyp(a.cur, idx);
  float sv = sqrtf(t.dv.w);
  float min_id = t.dv.z - sv * P.stereo_epl_var_fac;
  float max_id = t.dv.z + sv * P.stereo_epl_var_fac;
  if (min_id < 0.0f) min_id = 0.0f;
  if (max_id > 1.0f / P.min_depth) max_id = 1.0f / P.min_depth;
  float rid = 0.0f, rvar = 0.0f, repl = 0.0f;
  float err = line_stereo<false>(o, P, (float)x, (float)y, epx, epy, min_id, t.dv.z, max_id, ref, rid, rvar, repl);
  float diff = rid - t.dv.z;
  if (err == -1.0f) {
    code = LSDG_OBS_STEREO_FAIL_BASE + 1;
  } else if (err == -2.0f) {
    t.val -= P.validity_counter_dec;
    if (t.val < 0) t.val = 0;
    t.nsid = 0.0f;
    t.dv.y *= P.fail_var_inc_fac;
    if (t.dv.y > P.max_var) {
      t.valid = false;
      t.bl--;
    }
    store_hyp(a.cur, idx, t);
    code = LSDG_OBS_STEREO_FAIL_BASE + 2;
  } else if (err == -3.0f) {
    code = LSDG_OBS_STEREO_FAIL_BASE + 3;
  } else if (err == -4.0f) {
    code = LSDG_OBS_STEREO_FAIL_BASE + 4;
  } else if (P.diff_fac_observe * diff * diff > rvar + t.dv.w) {
    t.dv.y *= P.fail_var_inc_fac;
    if (t.dv.y > P.max_var) t.valid = false;
    store_hyp(a.cur, idx, t);
    code = LSDG_OBS_INCONSISTENT;
  } else {
    const float mg = a.kfMaxg[idx];
    float id_var = t.dv.y * P.succ_var_inc_fac;
    float wgt = rvar / (rvar + id_var);
    float new_id = (1.0f - wgt) * rid + wgt * t.dv.x;
    t.dv.x = unzero(new_id);
    id_var = id_var * wgt;
    if (id_var < t.dv.y) t.dv.y = id_var;
    t.val += P.validity_counter_inc;
    float cap_v = (float)P.validity_counter_max + mg * (float)P.validity_counter_max_variable / 255.0f;
    if ((float)t.val > cap_v) t.val = (int)cap_v;
    if (repl < P.min_epl_length_crop) {
      float inc = a.numTracked / (a.numMapped + 5.0f);
      if (inc < 3.0f) inc = 3.0f;
      inc += (float)(((int)(repl * 10000.0f)) % 2);
      if ((double)repl < 0.5 * (double)P.min_epl_length_crop) inc *= 3.0f;
      t.nsid = (float)ref.id + inc;
    }
    store_hyp(a.cur, idx, t);
    code = LSDG_OBS_UPDATED;
  }
  a.obs[idx] = code;
}

// Both searches in one persistent launch.  The work is cut into wave-sized
// chunks, numbered creates first, then updates.  The creates of all segments,
// in segment order, form one list cut into chunks of 64; an update chunk is
// 64 consecutive updates of one segment.  One lane searches one line.
//
// The grid is what the device holds at once, so every wave starts at launch.
// Blocks are dealt to the CUs in turn (block b to the same CU as block
// b mod (simds / 4)), so the waves fall into rounds of `simds` waves, one per
// SIMD.  A create wave (lines of about 100 steps) outlasts an update wave
// (about 30) and the pass ends with its busiest SIMD: round 0 takes the
// first create chunks, round 1 the rest of the creates and then updates, and
// round 2 updates only on the CUs whose round-1 block took none of the
// creates.  Later chunks go to every wave in turn.
template <bool INL>
__global__ __launch_bounds__(256, STEREO_MIN_BLOCKS) void k_observe_stereo(ObsArgs a, lsdg_params P, const int4* __restrict__ work,
                                                        const int* __restrict__ count, int segcap, int simds) {
  const int lane = (int)(threadIdx.x & 63);
  const int v = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)), V = (int)(gridDim.x * 4);
  // The counts, one per lane (one vector load for the whole table): lane
  // l < OBS_SEGS holds segment l's creates and lane OBS_SEGS + l its update
  // chunks, each with its inclusive prefix over its half of the table.
  int nItems = 0, pre = 0;
  if (lane < 2 * OBS_SEGS) {
    const int s = lane % OBS_SEGS, k = lane / OBS_SEGS;
    nItems = count[(2 * s + k) * OBS_COUNTER_STRIDE];
    pre = k == 0 ? nItems : (nItems + 63) / 64;
  }
#pragma unroll
  for (int d = 1; d < OBS_SEGS; d <<= 1) {
    const int o = __shfl_up(pre, d, 64);
    if (lane % OBS_SEGS >= d) pre += o;
  }
  // the create prefixes as wave-uniform values
  int cpre[OBS_SEGS];
#pragma unroll
  for (int k = 0; k < OBS_SEGS; ++k) cpre[k] = __builtin_amdgcn_readlane(pre, k);
  const int nCreate = cpre[OBS_SEGS - 1];
  const int nUpd = __builtin_amdgcn_readlane(pre, 2 * OBS_SEGS - 1);
  const int nCreateChunks = (nCreate + 63) / 64;
  const int nChunk = nCreateChunks + nUpd;
  // the round-2 waves left idle: those of the CUs with creates in round 1
  const int skip = V >= 3 * simds ? min(simds, (max(nCreateChunks - simds, 0) + 3) & ~3) : 0;
  const int r = v / simds, rs = v - r * simds;
  int c = r < 2 || skip == 0 ? v : r == 2 ? (rs < skip ? -1 : v - skip) : v - skip;
  for (int k = 0; c < nChunk; ++k, c = V - skip + (k - 1) * V + v) {
    if (c < 0) continue;
    if (c < nCreateChunks) {
      const int vi = 64 * c + lane;
      if (vi < nCreate) {
        // its segment: the number of segments whose prefix is at or below
        // it; its index there: vi less the largest such prefix
        int seg = 0, base = 0;
#pragma unroll
        for (int t = 0; t < OBS_SEGS - 1; ++t) {
          const bool below = cpre[t] <= vi;
          seg += below ? 1 : 0;
          base = below ? cpre[t] : base;
        }
        observe_create<INL>(a, P, work, segcap, seg, vi - base);
      }
    } else {
      const int u = c - nCreateChunks;
      // the update table lane of chunk u: the first whose prefix exceeds u
      // (prefixes do not decrease, so the lanes at or below u are the ones
      // before it)
      const int seg = (int)__popcll(__ballot(lane >= OBS_SEGS && lane < 2 * OBS_SEGS && pre <= u));
      const int q = u - (seg > 0 ? __shfl(pre, OBS_SEGS + seg - 1, 64) : 0);
      const int n = __shfl(nItems, OBS_SEGS + seg, 64);
      const int i = q * 64 + lane;
      if (i < n) observe_update<INL>(a, P, work, segcap, seg, i);
    }
  }
}

// ---------------------------------------------------------- fill holes
// The tile kernels below issue all of their global loads before their first
// LDS store (one round trip to memory per block) and read their 5x5
// neighbourhoods without branches.  They are bound by VALU issue (a wave64
// instruction takes four cycles on a 16-lane SIMD), so the divisions a 5x5
// gather repeats per centre pixel are done once per tile element instead:
// idepth / var and 1 / var for the fill's inverse-variance mean, and
// 1 / (var + distFac) for the six distinct squared distances of a 5x5 window
// for the regularization (each the same IEEE single operation the per-centre
// loop would do).  The tiles hold only what the neighbour reads use; a
// thread reads its own pixel's full hypothesis from global memory.
constexpr int TILE_THREADS = 256;

// squared distances dx^2 + dy^2 of a 5x5 window: 0, 1, 2, 4, 5, 8
constexpr int DIST_SLOTS = 6;
__host__ __device__ constexpr int dist_slot(int d2) { return d2 < 3 ? d2 : (d2 == 4 ? 3 : (d2 == 5 ? 4 : 5)); }
__host__ __device__ constexpr int slot_dist(int k) { return k < 3 ? k : (k == 3 ? 4 : (k == 4 ? 5 : 8)); }

__device__ inline float reg_dist_var(const lsdg_params& P) {
  return 0.075f * 0.075f * P.depth_smoothing_factor * P.depth_smoothing_factor;
}

// Neighbours of a regularization: validity flag, validity counter (as the
// float the reference's sum adds), idepth / variance, and per squared
// distance 1 / (var + d2 * regDistVar).  Positions outside the image are
// invalid.
template <int N>
struct RegTile {
  uint8_t valid[N][N];
  float val[N][N];
  float2 xy[N][N];
  float iv[DIST_SLOTS][N][N];
  __device__ void put(int ty, int tx, bool v, int c, float2 d, float regDistVar) {
    valid[ty][tx] = v ? 1 : 0;
    val[ty][tx] = (float)c;
    xy[ty][tx] = d;
#pragma unroll
    for (int k = 0; k < DIST_SLOTS; ++k) iv[k][ty][tx] = 1.0f / (d.y + (float)slot_dist(k) * regDistVar);
  }
};

// Neighbours of a hole fill: validity flag, validity counter, idepth /
// variance, and (idepth / var, 1 / var).
template <int N>
struct FillTile {
  uint8_t valid[N][N];
  int val[N][N];
  float2 xy[N][N];
  float2 qr[N][N];
};

// One pixel's hypothesis fields, loaded from global memory.
struct PixelLoad {
  uint8_t valid;
  int bl;
  float nsid;
  int val;
  float4 dv;
};
__device__ inline PixelLoad load_pixel(const HypBuf& b, int i) {
  return PixelLoad{b.valid[i], b.bl[i], b.nsid[i], b.val[i], b.dv[i]};
}
__device__ inline Hyp as_hyp(const PixelLoad& p) {
  Hyp d;
  d.valid = p.valid != 0;
  d.bl = p.bl;
  d.nsid = p.nsid;
  d.val = p.val;
  d.dv = p.dv;
  return d;
}

// Tile position k of an N x N tile with top-left corner (x0, y0): whether it
// lies in the image, and the pixel index to load (pixel 0, in bounds, if not).
template <int N>
__device__ inline bool tile_pos(int k, int x0, int y0, int w, int h, int& ty, int& tx, int& j) {
  ty = k / N;
  tx = k - ty * N;
  const int x = x0 + tx, y = y0 + ty;
  const bool ok = k < N * N && x >= 0 && x < w && y >= 0 && y < h;
  j = ok ? x + y * w : 0;
  return ok;
}

// regularizeDepthMapFillHoles at pixel (x, y) (tile position (cx, cy)) with
// validity flag `valid`, blacklist counter bl and keyframe gradient mg:
// whether a hypothesis is created there, and its idepth.
template <int N>
__device__ __forceinline__ bool fill_creates(const FillTile<N>& t, int cx, int cy, int x, int y, int w, int h,
                                             bool valid, int bl, float mg, const lsdg_params& P, float& obs) {
  if (!(x >= 3 && x < w - 2 && y >= 3 && y < h - 2 && !valid && !(mg < P.min_abs_grad_decrease))) return false;
  // 5x5 box sum of the validity of valid pixels (the reference reads it off
  // an integral image; the integer sum is the same)
  int val = 0;
#pragma unroll
  for (int yy = -2; yy <= 2; ++yy)
#pragma unroll
    for (int xx = -2; xx <= 2; ++xx) {
      const int c = t.val[cy + yy][cx + xx];
      val += t.valid[cy + yy][cx + xx] ? c : 0;
    }
  if (!((bl >= P.min_blacklist && val > P.val_sum_min_for_create) || val > P.val_sum_min_for_unblacklist))
    return false;
  // inverse-variance mean over the valid neighbours, in raster order
  float sumId = 0.0f, sumIv = 0.0f;
#pragma unroll
  for (int yy = -2; yy <= 2; ++yy)
#pragma unroll
    for (int xx = -2; xx <= 2; ++xx) {
      const bool ok = t.valid[cy + yy][cx + xx] != 0;
      const float2 e = t.qr[cy + yy][cx + xx];
      const float a = sumId + e.x, b = sumIv + e.y;
      sumId = ok ? a : sumId;
      sumIv = ok ? b : sumIv;
    }
  obs = unzero(sumId / sumIv);
  return true;
}

// regularizeDepthMapRow for pixel (x, y) with hypothesis d; the 5x5
// neighbourhood is read from tile t around (cx, cy).  Every neighbour is
// evaluated and the sums select what the reference's loop would add, in its
// order (dx outer, dy inner).
template <int N>
__device__ __forceinline__ Hyp regularize_pixel(const RegTile<N>& t, int cx, int cy, int x, int y, int w, int h, Hyp d,
                                                const lsdg_params& P, int removeOcc, int validityTH) {
  if (x >= 2 && x < w - 2 && y >= 2 && y < h - 2 && d.valid) {
    float sum = 0.0f, val_sum = 0.0f, sumIvar = 0.0f;
    int nOcc = 0, nNot = 0;
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx)
#pragma unroll
      for (int dy = -2; dy <= 2; ++dy) {
        const bool ok = t.valid[cy + dy][cx + dx] != 0;
        const float2 sv = t.xy[cy + dy][cx + dx];
        const float diff = sv.x - d.dv.x;
        const bool rej = P.diff_fac_smoothing * diff * diff > sv.y + d.dv.y;
        const bool take = ok && !rej;
        nOcc += (ok && rej && sv.x > d.dv.x) ? 1 : 0;
        nNot += take ? 1 : 0;
        const float ivar = t.iv[dist_slot(dx * dx + dy * dy)][cy + dy][cx + dx];
        const float vs = val_sum + t.val[cy + dy][cx + dx], s = sum + sv.x * ivar, si = sumIvar + ivar;
        val_sum = take ? vs : val_sum;
        sum = take ? s : sum;
        sumIvar = take ? si : sumIvar;
      }
    if (val_sum < (float)validityTH) {
      d.valid = false;
      d.bl--;
    } else if (removeOcc && nOcc > nNot) {
      d.valid = false;
    } else {
      sum = sum / sumIvar;
      sum = unzero(sum);
      d.dv.z = sum;
      d.dv.w = 1.0f / sumIvar;
    }
  }
  return d;
}

// Frame::setDepth for one pixel: the smoothed idepth / variance of a valid
// hypothesis not behind the camera, else -1 for both; returns what it wrote.
__device__ inline float2 set_depth_pixel(const Hyp& d, int i, float* __restrict__ idepth, float* __restrict__ var) {
  float2 r = make_float2(-1.0f, -1.0f);
  if (d.valid && (double)d.dv.z >= -0.05) r = make_float2(unzero(d.dv.z), unzero(d.dv.w));
  idepth[i] = r.x;
  var[i] = r.y;
  return r;
}

// ---------------------------------------------------------- regularize
// regularizeDepthMap(removeOcclusions, validityTH) on a 16x16 block and its
// 2-pixel halo.
__global__ __launch_bounds__(TILE_THREADS) void k_regularize(DepthGeom g, HypBuf in, HypBuf out, lsdg_params P,
                                                             int removeOcc, int validityTH) {
  __shared__ RegTile<20> t;
  constexpr int ITER = (20 * 20 + TILE_THREADS - 1) / TILE_THREADS;
  const int w = g.w, h = g.h;
  const int bx = blockIdx.x * 16, by = blockIdx.y * 16;
  const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
  const int x = bx + lx, y = by + ly;
  const bool inside = x < w && y < h;
  const int idx = inside ? x + y * w : 0;
  const PixelLoad self = load_pixel(in, idx);
  uint8_t v[ITER];
  int c[ITER];
  float4 q[ITER];
  bool ok[ITER];
#pragma unroll
  for (int i = 0; i < ITER; ++i) {
    int ty, tx, j;
    ok[i] = tile_pos<20>(threadIdx.x + i * TILE_THREADS, bx - 2, by - 2, w, h, ty, tx, j);
    v[i] = in.valid[j];
    c[i] = in.val[j];
    q[i] = in.dv[j];
  }
  const float rdv = reg_dist_var(P);
#pragma unroll
  for (int i = 0; i < ITER; ++i) {
    const int k = threadIdx.x + i * TILE_THREADS;
    if (k >= 20 * 20) break;
    const int ty = k / 20, tx = k - ty * 20;
    t.put(ty, tx, ok[i] && v[i], ok[i] ? c[i] : 0, ok[i] ? make_float2(q[i].x, q[i].y) : make_float2(0.f, 1.f), rdv);
  }
  __syncthreads();
  if (!inside) return;
  store_hyp(out, idx, regularize_pixel(t, lx + 2, ly + 2, x, y, w, h, as_hyp(self), P, removeOcc, validityTH));
}

// Fill holes, then regularize (no occlusion removal), in one pass.  A block
// fills its pixels and their 2-pixel halo from a 4-pixel-halo input tile (a
// halo pixel's fill depends only on the input, so every block that fills it
// gets the same result) and regularizes its pixels from the filled tile.  It
// stores the filled map (`filled`: the reference's other map afterwards) and
// the regularized one (`out`), and, when kfIdepth is set, Frame::setDepth's
// level-0 idepth / variance of the regularized map.
// The tile a block of a gx x gy grid takes so that each XCD holds one band
// of consecutive tiles in raster order.  Blocks are dealt to the XCDs in
// turn (block b to XCD b % 8), and a tile kernel's halo is its neighbours'
// pixels: with the tiles in raster order the neighbours sit on other XCDs
// and every halo is fetched again into another L2.
__device__ inline int2 xcd_tile(int gx, int gy) {
  constexpr int XCDS = 8;
  const int n = gx * gy, b = (int)(blockIdx.x + blockIdx.y * gx);
  const int q = n / XCDS, r = n % XCDS, x = b % XCDS, slot = b / XCDS;
  const int t = x * q + min(x, r) + slot;  // XCD x takes q + (x < r) tiles
  return make_int2(t % gx, t / gx);
}

static_assert(PYR == 5, "a 16x16 tile holds whole 2^(PYR-1) blocks of level-0 pixels");
__global__ __launch_bounds__(TILE_THREADS) void k_fill_regularize(DepthGeom g, HypBuf in, HypBuf filled, HypBuf out,
                                                                  const float* __restrict__ maxg, lsdg_params P,
                                                                  int validityTH, PyrDepthArgs kf) {
  __shared__ FillTile<24> a;  // input, pixels and a 4-pixel halo
  __shared__ RegTile<20> f;   // filled, pixels and a 2-pixel halo
  constexpr int AITER = (24 * 24 + TILE_THREADS - 1) / TILE_THREADS;
  constexpr int FITER = (20 * 20 + TILE_THREADS - 1) / TILE_THREADS;
  const int w = g.w, h = g.h;
  const int2 tile = xcd_tile(gridDim.x, gridDim.y);
  const int bx = tile.x * 16, by = tile.y * 16;
  const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
  const int x = bx + lx, y = by + ly;
  const bool inside = x < w && y < h;
  const int idx = inside ? x + y * w : 0;
  // every global load first: this pixel, the input tile, and the blacklist
  // counter and keyframe gradient of the filled-tile positions
  const PixelLoad self = load_pixel(in, idx);
  uint8_t av[AITER];
  int ac[AITER];
  float4 aq[AITER];
  bool aok[AITER];
#pragma unroll
  for (int i = 0; i < AITER; ++i) {
    int ty, tx, j;
    aok[i] = tile_pos<24>(threadIdx.x + i * TILE_THREADS, bx - 4, by - 4, w, h, ty, tx, j);
    av[i] = in.valid[j];
    ac[i] = in.val[j];
    aq[i] = in.dv[j];
  }
  int fb[FITER];
  float fm[FITER];
#pragma unroll
  for (int i = 0; i < FITER; ++i) {
    int ty, tx, j;
    tile_pos<20>(threadIdx.x + i * TILE_THREADS, bx - 2, by - 2, w, h, ty, tx, j);
    fb[i] = in.bl[j];
    fm[i] = maxg[j];
  }
#pragma unroll
  for (int i = 0; i < AITER; ++i) {
    const int k = threadIdx.x + i * TILE_THREADS;
    if (k >= 24 * 24) break;
    const int ty = k / 24, tx = k - ty * 24;
    const float2 d = aok[i] ? make_float2(aq[i].x, aq[i].y) : make_float2(0.f, 1.f);
    a.valid[ty][tx] = aok[i] ? av[i] : 0;
    a.val[ty][tx] = aok[i] ? ac[i] : 0;
    a.xy[ty][tx] = d;
    a.qr[ty][tx] = make_float2(d.x / d.y, 1.0f / d.y);
  }
  __syncthreads();
  const float rdv = reg_dist_var(P);
#pragma unroll
  for (int i = 0; i < FITER; ++i) {
    const int k = threadIdx.x + i * TILE_THREADS;
    if (k >= 20 * 20) break;
    int ty, tx, j;
    const bool ok = tile_pos<20>(k, bx - 2, by - 2, w, h, ty, tx, j);
    bool v = ok && a.valid[ty + 2][tx + 2];
    int c = ok ? a.val[ty + 2][tx + 2] : 0;
    float2 e = a.xy[ty + 2][tx + 2];
    float obs = 0.0f;
    if (ok && fill_creates(a, tx + 2, ty + 2, bx - 2 + tx, by - 2 + ty, w, h, v, fb[i], fm[i], P, obs)) {
      v = true;
      c = 0;
      e = make_float2(obs, P.var_random_init_initial);
    }
    f.put(ty, tx, v, c, e, rdv);
  }
  __syncthreads();
  float2 d0 = make_float2(-1.0f, -1.0f);
  if (inside) {
    const int cx = lx + 2, cy = ly + 2;
    // the filled hypothesis: DepthMapPixelHypothesis(obs, VAR_RANDOM_INIT_INITIAL, 0) where created
    Hyp d = as_hyp(self);
    if (f.valid[cy][cx] && !d.valid) {
      const float2 e = f.xy[cy][cx];
      d = make_hyp(e.x, e.y, 0);
    }
    store_hyp(filled, idx, d);
    const Hyp r = regularize_pixel(f, cx, cy, x, y, w, h, d, P, 0, validityTH);
    store_hyp(out, idx, r);
    if (kf.id[0]) d0 = set_depth_pixel(r, idx, kf.id[0], kf.var[0]);
  }
  if (!kf.id[0]) return;  // (uniform)
  // The keyframe's depth pyramid: each level-l pixel fuses the 2^l x 2^l
  // level-0 pixels below it, which lie in this tile, so levels 1-4 follow
  // from the tile's setDepth values in shared memory (level by level, in
  // place: a level's reads end at a barrier before its writes).
  __shared__ float pid[16][17], pvar[16][17];
  pid[ly][lx] = d0.x;
  pvar[ly][lx] = d0.y;
  __syncthreads();
#pragma unroll
  for (int l = 1, n = 8; l < PYR; ++l, n >>= 1) {
    const int tx = threadIdx.x % n, ty = threadIdx.x / n;
    float oi = -1.0f, ov = -1.0f;
    if ((int)threadIdx.x < n * n) {
      const int px = tile.x * n + tx, py = tile.y * n + ty;
      if (px < kf.w[l] && py < kf.h[l]) {
        const float ia[4] = {pid[2 * ty][2 * tx], pid[2 * ty][2 * tx + 1], pid[2 * ty + 1][2 * tx],
                             pid[2 * ty + 1][2 * tx + 1]};
        const float va[4] = {pvar[2 * ty][2 * tx], pvar[2 * ty][2 * tx + 1], pvar[2 * ty + 1][2 * tx],
                             pvar[2 * ty + 1][2 * tx + 1]};
        fuse4(ia, va, oi, ov);
        kf.id[l][px + py * kf.w[l]] = oi;
        kf.var[l][px + py * kf.w[l]] = ov;
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < n * n) {
      pid[ty][tx] = oi;
      pvar[ty][tx] = ov;
    }
    __syncthreads();
  }
}

// -------------------------------------------------------- Frame::setDepth
__global__ void k_set_depth(HypBuf cur, int n, float* __restrict__ idepth, float* __restrict__ var) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  set_depth_pixel(load_hyp(cur, i), i, idepth, var);
}

// ------------------------------------------------------------ propagate
struct PropArgs {
  DepthGeom g;
  HypBuf src;
  Xf x;                      // old -> new keyframe
  const uint8_t* wasGood;    // level-1 mask of the new keyframe, or null
  const float* kfImg;        // old keyframe image
  const float* newMaxg;
  const float* newImg;
  int* dest;
  float2* newIV;
  int* count;
};

__global__ __launch_bounds__(256) void k_prop_project(PropArgs a, lsdg_params P) {
  const int W = a.g.w, H = a.g.h;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  a.dest[i] = -1;
  if (!a.src.valid[i]) return;
  int y = i / W, x = i - y * W;
  float4 s = a.src.dv[i];
  float k0 = (float)x * a.g.fxi + a.g.cxi, k1 = (float)y * a.g.fyi + a.g.cyi, k2 = 1.0f;
  float pn[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    pn[r] = (a.x.R[3 * r] * k0 + a.x.R[3 * r + 1] * k1 + a.x.R[3 * r + 2] * k2) / s.z + a.x.t[r];
  float new_id = 1.0f / pn[2];
  float u = pn[0] * new_id * a.g.fx + a.g.cx;
  float v = pn[1] * new_id * a.g.fy + a.g.cy;
  if (!(u > 2.1f && v > 2.1f && u < (float)W - 3.1f && v < (float)H - 3.1f)) return;
  int nidx = (int)(u + 0.5f) + ((int)(v + 0.5f)) * W;
  float destAbsGrad = a.newMaxg[nidx];
  if (a.wasGood) {
    if (!a.wasGood[(x >> 1) + (W >> 1) * (y >> 1)] || destAbsGrad < P.min_abs_grad_decrease) return;
  } else {
    float sc = a.kfImg[i];
    float dc = dinterp(a.newImg, u, v, W);
    float res = dc - sc;
    if (res * res / (P.max_diff_constant + P.max_diff_grad_mult * destAbsGrad * destAbsGrad) > 1.0f ||
        destAbsGrad < P.min_abs_grad_decrease)
      return;
  }
  float r4 = new_id / s.z;
  r4 *= r4;
  r4 *= r4;
  float new_var = r4 * s.y;
  a.dest[i] = nidx;
  a.newIV[i] = make_float2(new_id, new_var);
  atomicAdd(&a.count[nidx], 1);
}

__global__ __launch_bounds__(256) void k_prop_bucket(const int* __restrict__ dest, const int* __restrict__ offsets,
                                                     int* __restrict__ fill, int* __restrict__ bucket, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int d = dest[i];
  if (d < 0) return;
  int pos = atomicAdd(&fill[d], 1);
  bucket[offsets[d] + pos] = i;
}

// One thread per destination: sources in ascending (raster) order, the
// reference's occlusion / create / merge rules.  Empty destinations only get
// the wipe (isValid = false, blacklisted = 0).
__global__ __launch_bounds__(256) void k_prop_fold(HypBuf src, HypBuf out, const float2* __restrict__ newIV,
                                                   const int* __restrict__ offsets, const int* __restrict__ bucket,
                                                   int n, lsdg_params P) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  int b0 = offsets[d], b1 = offsets[d + 1];
  if (b0 == b1) {
    out.valid[d] = 0;
    out.bl[d] = 0;
    return;
  }
  Hyp t;
  t.valid = false;
  t.bl = 0;
  t.nsid = out.nsid[d];
  t.val = out.val[d];
  t.dv = out.dv[d];
  int last = -1;
  for (int k = b0; k < b1; ++k) {
    // next source in raster order
    int si = 0x7fffffff;
    for (int j = b0; j < b1; ++j) {
      int c = bucket[j];
      if (c > last && c < si) si = c;
    }
    last = si;
    float2 iv = newIV[si];
    float new_id = iv.x, new_var = iv.y;
    int sval = src.val[si];
    if (t.valid) {
      float diff = t.dv.x - new_id;
      if (P.diff_fac_prop_merge * diff * diff > new_var + t.dv.y) {
        if (new_id < t.dv.x) continue;
        t.valid = false;
      }
    }
    if (!t.valid) {
      t = make_hyp(new_id, new_var, min(P.validity_counter_max_variable, sval));
    } else {
      float wgt = new_var / (t.dv.y + new_var);
      float merged = wgt * t.dv.x + (1.0f - wgt) * new_id;
      int mv = sval + t.val;
      int cap = P.validity_counter_max + P.validity_counter_max_variable;
      if (mv > cap) mv = cap;
      t = make_hyp(merged, 1.0f / (1.0f / t.dv.y + 1.0f / new_var), mv);
    }
  }
  store_hyp(out, d, t);
}

// Exclusive scan of count[0..n) into offsets (in place), offsets[n] = total.
// Three kernels: per-block sums, a one-block scan of block sums, per-block scan.
__global__ __launch_bounds__(256) void k_scan_blocksum(const int* __restrict__ a, int n, int* __restrict__ blk) {
  __shared__ int sh[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  sh[threadIdx.x] = i < n ? a[i] : 0;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(1024) void k_scan_top(int* blk, int nb) {
  __shared__ int sh[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < nb ? blk[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) blk[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[nb] = carry;
}

__global__ __launch_bounds__(256) void k_scan_apply(int* a, int n, const int* __restrict__ blk) {
  __shared__ int sh[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  int v = i < n ? a[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) a[i] = blk[blockIdx.x] + sh[threadIdx.x] - v;
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) a[n] = blk[gridDim.x];
}

// ---------------------------------------------------- mean inverse depth
// Count and FP64 sum of idepth_smoothed over valid pixels; the last block
// merges the per-block rows in block order and writes (count, sum).
__global__ __launch_bounds__(256) void k_mean_idepth(HypBuf cur, int n, double* __restrict__ partials,
                                                     unsigned* counter, double* __restrict__ out) {
  __shared__ double shc[256], shs[256];
  __shared__ bool amLast;
  double c = 0.0, s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    if (cur.valid[i]) {
      c += 1.0;
      s += (double)cur.dv[i].z;
    }
  shc[threadIdx.x] = c;
  shs[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      shc[threadIdx.x] += shc[threadIdx.x + k];
      shs[threadIdx.x] += shs[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = shc[0];
    partials[2 * blockIdx.x + 1] = shs[0];
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!amLast) return;
  __threadfence();
  merge_rows<2, 256>(partials, gridDim.x, out);
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_rescale(HypBuf cur, int n, float f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !cur.valid[i]) return;
  float f2 = f * f;
  float4 d = cur.dv[i];
  d.x *= f;
  d.z *= f;
  d.y *= f2;
  d.w *= f2;
  cur.dv[i] = d;
}

// Hypotheses from (idepth, var) arrays: valid where var > 0 and idepth is
// finite (gt != 0: also idepth > 0, as initializeFromGTDepth).
__global__ void k_hyp_prior(HypBuf cur, int n, const float* __restrict__ id, const float* __restrict__ var,
                            int validity, int gt, float gtVar) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = id[i], s = gt ? gtVar : var[i];
  bool ok = gt ? (!isnan(v) && v > 0.0f) : (isfinite(v) && s > 0.0f);
  Hyp h;
  if (ok) {
    h = make_hyp(v, s, validity);
    h.dv.z = v;
    h.dv.w = s;
  } else {
    h.valid = false;
    h.bl = 0;
    h.nsid = gt ? cur.nsid[i] : 0.0f;
    h.val = gt ? cur.val[i] : 0;
    h.dv = gt ? cur.dv[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  store_hyp(cur, i, h);
}

// DepthMap::initializeRandomly: interior pixels whose keyframe max-gradient
// exceeds MIN_ABS_GRAD_CREATE get DepthMapPixelHypothesis(r, r, VAR_RANDOM_INIT_INITIAL,
// VAR_RANDOM_INIT_INITIAL, 20) with r = 0.5 + (draw % 100001) / 100000, the
// others lose their hypothesis (flag and blacklist cleared, the rest kept);
// the border is untouched.  The draw is random_init_draw in place of the
// reference's rand().
__global__ void k_random_init(HypBuf cur, int w, int h, const float* __restrict__ maxg, float minGrad, float var,
                              unsigned long long seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int y = i / w, x = i - y * w;
  if (x < 1 || x >= w - 1 || y < 1 || y >= h - 1) return;
  if (maxg[i] > minGrad) {
    const float id = 0.5f + 1.0f * ((float)(random_init_draw(seed, i) % 100001) / 100000.0f);
    Hyp t = make_hyp(id, var, 20);
    t.dv.z = id;
    t.dv.w = var;
    store_hyp(cur, i, t);
  } else {
    cur.valid[i] = 0;
    cur.bl[i] = 0;
  }
}

// ------------------------------------------------------------ launchers
static inline dim3 grid16(int w, int h) { return dim3((w + 15) / 16, (h + 15) / 16); }

static_assert(OBS_COUNT_BYTES >= 2 * 2 * OBS_SEGS * OBS_COUNTER_STRIDE * sizeof(int), "observe counter sets");

size_t obs_work_slots(int w, int h) {
  const size_t bx = (size_t)(w + GATE_W - 1) / GATE_W, by = (size_t)(h + GATE_TH - 1) / GATE_TH;
  return ((bx * by + OBS_SEGS - 1) / OBS_SEGS) * OBS_SEGS * 256 * GATE_PX;
}

hipError_t launch_observe(lsdg_depthmap dm, const lsdg_params& P, const StereoRef* h_refs, int nRefs,
                          const int* h_refById, int nRefById, int refByIdOffset, int oldestIdx, int newestIdx,
                          float numTracked, float numMapped, hipStream_t s) {
  ObsArgs a{};
  a.g = dm->g;
  a.cur = dm->buf[dm->cur];
  a.obs = dm->d_obs;
  lsdg_frame kf = dm->kf;
  a.kfImg = kf->d_image[0];
  a.kfMaxg = kf->d_maxg[0];
  a.kfGrad = kf->d_grad[0];
  a.nRefById = nRefById;
  a.refByIdOffset = refByIdOffset;
  a.oldestIdx = oldestIdx;
  a.newestIdx = newestIdx;
  a.kfReactivated = dm->kfReactivated ? 1 : 0;
  a.kfId = kf->id;
  a.numTracked = numTracked;
  a.numMapped = numMapped;
  const bool inl = nRefs <= OBS_INLINE_REFS && nRefById <= OBS_INLINE_BYID;
  hipError_t e = hipSuccess;
  if (inl) {
    for (int i = 0; i < nRefs; ++i) a.inlRefs[i] = h_refs[i];
    for (int k = 0; k < nRefById; ++k) a.inlById[k] = h_refById[k];
  } else {
    // h_refs / h_refById are pinned staging buffers the caller keeps until the stream drains
    e = hipMemcpyAsync(dm->d_refs, h_refs, sizeof(StereoRef) * nRefs, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dm->d_refById, h_refById, sizeof(int) * nRefById, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    a.refs = dm->d_refs;
    a.refById = dm->d_refById;
  }
  const dim3 gg((unsigned)((dm->g.w + GATE_W - 1) / GATE_W), (unsigned)((dm->g.h + GATE_TH - 1) / GATE_TH));
  // a segment takes every OBS_SEGS-th gate block: one slot per pixel
  const int segBlocks = (int)((gg.x * gg.y + OBS_SEGS - 1) / OBS_SEGS);
  const int segcap = segBlocks * 256 * GATE_PX;
  if ((size_t)segcap * OBS_SEGS > obs_work_slots(dm->g.w, dm->g.h)) return hipErrorInvalidValue;
  const int set = 2 * OBS_SEGS * OBS_COUNTER_STRIDE;
  int* count = dm->d_work_count + dm->obsParity * set;
  int* other = dm->d_work_count + (1 - dm->obsParity) * set;
  dm->obsParity ^= 1;
  // stereo grid: the blocks the device holds at once, at most one wave per
  // chunk the lists can hold
  int& resident = dm->stereoBlocks[inl ? 0 : 1];
  if (resident == 0) {
    int perCU = 0;
    e = inl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_observe_stereo<true>, 256, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_observe_stereo<false>, 256, 0);
    if (e != hipSuccess) return e;
    resident = perCU > 0 ? perCU * dm->ctx->num_cus : 1;
  }
  const long long maxChunks = (long long)OBS_SEGS * 2 * ((segcap + 63) / 64);
  const int simds = 4 * dm->ctx->num_cus;  // 4 SIMDs per CU
  const dim3 gs((unsigned)std::min<long long>(resident, (maxChunks + 3) / 4));
  if (inl) {
    k_observe_gate<true><<<gg, dim3(256), 0, s>>>(a, P, dm->d_work, count, other, segcap);
    k_observe_stereo<true><<<gs, dim3(256), 0, s>>>(a, P, dm->d_work, count, segcap, simds);
  } else {
    k_observe_gate<false><<<gg, dim3(256), 0, s>>>(a, P, dm->d_work, count, other, segcap);
    k_observe_stereo<false><<<gs, dim3(256), 0, s>>>(a, P, dm->d_work, count, segcap, simds);
  }
  return hipGetLastError();
}

hipError_t launch_regularize(lsdg_depthmap dm, const lsdg_params& P, bool removeOcclusions, int th, hipStream_t s) {
  k_regularize<<<grid16(dm->g.w, dm->g.h), dim3(TILE_THREADS), 0, s>>>(dm->g, dm->buf[dm->cur], dm->buf[dm->oth], P,
                                                                       removeOcclusions ? 1 : 0, th);
  std::swap(dm->cur, dm->oth);
  return hipGetLastError();
}

hipError_t launch_fill_regularize(lsdg_depthmap dm, const lsdg_params& P, int th, lsdg_frame setDepth,
                                  hipStream_t s) {
  const int in = dm->cur, fill = dm->oth, reg = 3 - dm->cur - dm->oth;
  PyrDepthArgs kf{};
  if (setDepth) {
    for (int l = 0; l < PYR; ++l) {
      kf.id[l] = setDepth->d_idepth[l];
      kf.var[l] = setDepth->d_ivar[l];
      kf.w[l] = setDepth->w[l];
      kf.h[l] = setDepth->h[l];
    }
  }
  k_fill_regularize<<<grid16(dm->g.w, dm->g.h), dim3(TILE_THREADS), 0, s>>>(
      dm->g, dm->buf[in], dm->buf[fill], dm->buf[reg], dm->kf->d_maxg[0], P, th, kf);
  dm->oth = fill;
  dm->cur = reg;
  return hipGetLastError();
}

hipError_t launch_set_depth(lsdg_depthmap dm, lsdg_frame f, hipStream_t s) {
  int n = dm->g.w * dm->g.h;
  k_set_depth<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(dm->buf[dm->cur], n, f->d_idepth[0], f->d_ivar[0]);
  return hipGetLastError();
}

hipError_t launch_propagate(lsdg_depthmap dm, const lsdg_params& P, lsdg_frame nkf, const Xf& oldToNew,
                            const uint8_t* d_wasGood, hipStream_t s) {
  const int n = dm->g.w * dm->g.h;
  const int nb = (n + 255) / 256;
  hipError_t e;
  if ((e = hipMemsetAsync(dm->d_count, 0, sizeof(int) * (n + 1), s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(dm->d_fill, 0, sizeof(int) * n, s)) != hipSuccess) return e;
  PropArgs a;
  a.g = dm->g;
  a.src = dm->buf[dm->cur];
  a.x = oldToNew;
  a.wasGood = d_wasGood;
  a.kfImg = dm->kf->d_image[0];
  a.newMaxg = nkf->d_maxg[0];
  a.newImg = nkf->d_image[0];
  a.dest = dm->d_dest;
  a.newIV = dm->d_newIV;
  a.count = dm->d_count;
  k_prop_project<<<dim3(nb), dim3(256), 0, s>>>(a, P);
  k_scan_blocksum<<<dim3(nb), dim3(256), 0, s>>>(dm->d_count, n, dm->d_scanblk);
  k_scan_top<<<dim3(1), dim3(1024), 0, s>>>(dm->d_scanblk, nb);
  k_scan_apply<<<dim3(nb), dim3(256), 0, s>>>(dm->d_count, n, dm->d_scanblk);
  k_prop_bucket<<<dim3(nb), dim3(256), 0, s>>>(dm->d_dest, dm->d_count, dm->d_fill, dm->d_bucket, n);
  k_prop_fold<<<dim3(nb), dim3(256), 0, s>>>(dm->buf[dm->cur], dm->buf[dm->oth], dm->d_newIV, dm->d_count,
                                             dm->d_bucket, n, P);
  std::swap(dm->cur, dm->oth);
  return hipGetLastError();
}

hipError_t launch_mean_idepth(lsdg_depthmap dm, hipStream_t s) {
  int n = dm->g.w * dm->g.h;
  int nb = (n + 255 * 8) / (256 * 8);
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  k_mean_idepth<<<dim3(nb), dim3(256), 0, s>>>(dm->buf[dm->cur], n, dm->d_red, dm->d_red_counter, dm->d_red_out);
  return hipGetLastError();
}

hipError_t launch_rescale(lsdg_depthmap dm, float f, hipStream_t s) {
  int n = dm->g.w * dm->g.h;
  k_rescale<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(dm->buf[dm->cur], n, f);
  return hipGetLastError();
}

hipError_t launch_hyp_prior(lsdg_depthmap dm, const float* d_id, const float* d_var, int validity, int gt,
                            float gtVar, hipStream_t s) {
  int n = dm->g.w * dm->g.h;
  k_hyp_prior<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(dm->buf[dm->cur], n, d_id, d_var, validity, gt, gtVar);
  return hipGetLastError();
}

hipError_t launch_random_init(lsdg_depthmap dm, const lsdg_params& P, unsigned long long seed, hipStream_t s) {
  const int n = dm->g.w * dm->g.h;
  k_random_init<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(dm->buf[dm->cur], dm->g.w, dm->g.h, dm->kf->d_maxg[0],
                                                            P.min_abs_grad_create, P.var_random_init_initial, seed);
  return hipGetLastError();
}

}  // namespace lsdg
