// Per-point tracker math shared by the plain (one alignment per launch) and
// batched (many alignments per launch) evaluation kernels: bilinear taps of the
// (I, gx, gy) pyramid in FP32 / FP16 / BF16, the Jacobian row, the per-thread
// accumulators, and one function per point for SE3, Sim3 and the adjoint.
// The float expressions keep the reference's operation order (SE3Tracker::
// calcResidualAndBuffers / calcWeightsAndResidual / calculateWarpUpdate and the
// Sim3Tracker counterparts) so masks and counts match the CPU restatement.
#pragma once
#include <type_traits>

#include "lsdg_internal.h"

namespace lsdg {

// Waits until this wave's earlier memory operations have completed.  Stores
// made as agent-scope atomics (sc1) are coherent at agent scope once complete,
// so a wait orders them before a later counter update with no write-back of
// the XCD's L2 (which an agent-scope release adds for ordinary stores).  The
// signal fences keep the compiler from moving memory operations across it.
__device__ inline void wait_own_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

constexpr int EVAL_BS = 256;
constexpr int EVAL_WAVES = EVAL_BS / 64;
constexpr int EVAL_U = 2;  // points per thread with loads in flight together (plain kernels)

struct Taps {
  float gx, gy, I;
};

__device__ inline float bf16_to_f32(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

template <int PREC>
__device__ inline Taps load_tap(const void* img, int i) {
  if constexpr (PREC == LSDG_FP32) {
    float4 g = ((const float4*)img)[i];
    return {g.x, g.y, g.z};
  } else {
    uint2 g = ((const uint2*)img)[i];
    if constexpr (PREC == LSDG_FP16) {
      return {__half2float(__ushort_as_half((unsigned short)(g.x & 0xFFFFu))),
              __half2float(__ushort_as_half((unsigned short)(g.x >> 16))),
              __half2float(__ushort_as_half((unsigned short)(g.y & 0xFFFFu)))};
    } else {
      return {bf16_to_f32((unsigned short)(g.x & 0xFFFFu)), bf16_to_f32((unsigned short)(g.x >> 16)),
              bf16_to_f32((unsigned short)(g.y & 0xFFFFu))};
    }
  }
}

// getInterpolatedElement43: weights dxdy, dy-dxdy, dx-dxdy, 1-dx-dy+dxdy.
template <int PREC>
__device__ inline Taps interp3(const void* img, float x, float y, int w) {
  int ix = (int)x, iy = (int)y;
  float dx = x - (float)ix, dy = y - (float)iy;
  float dxdy = dx * dy;
  int b = ix + iy * w;
  Taps a11 = load_tap<PREC>(img, b + 1 + w), a01 = load_tap<PREC>(img, b + w);
  Taps a10 = load_tap<PREC>(img, b + 1), a00 = load_tap<PREC>(img, b);
  float w11 = dxdy, w01 = dy - dxdy, w10 = dx - dxdy, w00 = 1.0f - dx - dy + dxdy;
  Taps o;
  o.gx = w11 * a11.gx + w01 * a01.gx + w10 * a10.gx + w00 * a00.gx;
  o.gy = w11 * a11.gy + w01 * a01.gy + w10 * a10.gy + w00 * a00.gy;
  o.I = w11 * a11.I + w01 * a01.I + w10 * a10.I + w00 * a00.I;
  return o;
}

// The four raw taps of one bilinear sample, fetched ahead of their use so a
// thread can keep several points' loads in flight.
template <int PREC>
struct TapRaw {
  using T = typename std::conditional<PREC == LSDG_FP32, float4, uint2>::type;
  T a11, a01, a10, a00;
  float dx, dy;
};

template <int PREC>
__device__ inline Taps decode_tap(const typename TapRaw<PREC>::T& g) {
  if constexpr (PREC == LSDG_FP32) {
    return {g.x, g.y, g.z};
  } else if constexpr (PREC == LSDG_FP16) {
    return {__half2float(__ushort_as_half((unsigned short)(g.x & 0xFFFFu))),
            __half2float(__ushort_as_half((unsigned short)(g.x >> 16))),
            __half2float(__ushort_as_half((unsigned short)(g.y & 0xFFFFu)))};
  } else {
    return {bf16_to_f32((unsigned short)(g.x & 0xFFFFu)), bf16_to_f32((unsigned short)(g.x >> 16)),
            bf16_to_f32((unsigned short)(g.y & 0xFFFFu))};
  }
}

template <int PREC>
__device__ inline void fetch_taps(const void* img, float x, float y, int w, TapRaw<PREC>& t) {
  using T = typename TapRaw<PREC>::T;
  const T* im = (const T*)img;
  int ix = (int)x, iy = (int)y;
  t.dx = x - (float)ix;
  t.dy = y - (float)iy;
  int b = ix + iy * w;
  t.a11 = im[b + 1 + w];
  t.a01 = im[b + w];
  t.a10 = im[b + 1];
  t.a00 = im[b];
}

// interp3's weighted sum of fetched taps (same operations, same order).
template <int PREC>
__device__ inline Taps blend_taps(const TapRaw<PREC>& t) {
  float dx = t.dx, dy = t.dy;
  float dxdy = dx * dy;
  Taps a11 = decode_tap<PREC>(t.a11), a01 = decode_tap<PREC>(t.a01);
  Taps a10 = decode_tap<PREC>(t.a10), a00 = decode_tap<PREC>(t.a00);
  float w11 = dxdy, w01 = dy - dxdy, w10 = dx - dxdy, w00 = 1.0f - dx - dy + dxdy;
  Taps o;
  o.gx = w11 * a11.gx + w01 * a01.gx + w10 * a10.gx + w00 * a00.gx;
  o.gy = w11 * a11.gy + w01 * a01.gy + w10 * a10.gy + w00 * a00.gy;
  o.I = w11 * a11.I + w01 * a01.I + w10 * a10.I + w00 * a00.I;
  return o;
}

// Division, reciprocal and square root of the per-point math.  FP32 pyramids
// use IEEE correctly rounded operations (the CPU restatement's results, bit
// for bit); FP16 / BF16 pyramids use the hardware reciprocal and square-root
// approximations (about 1 ulp, far below the taps' 11- / 8-bit quantisation),
// which cut the evaluation's VALU work by about a third.
template <int PREC>
__device__ inline float fdiv(float a, float b) {
  if constexpr (PREC == LSDG_FP32) return a / b;
  else return a * __builtin_amdgcn_rcpf(b);
}
template <int PREC>
__device__ inline float frcp(float b) {
  if constexpr (PREC == LSDG_FP32) return 1.0f / b;
  else return __builtin_amdgcn_rcpf(b);
}
template <int PREC>
__device__ inline float fsqrt(float x) {
  if constexpr (PREC == LSDG_FP32) return sqrtf(x);
  else return __builtin_amdgcn_sqrtf(x);
}

// Projection of one reference point into the frame (the warp every tracker
// evaluation starts with).
struct Proj {
  float Wx, Wy, Wz, u, v;
  bool in;
};

template <int PREC>
__device__ inline Proj project(const float4& p, const Xf& x, const LevelGeom& g) {
  const float wlim = (float)(g.w - 2), hlim = (float)(g.h - 2);
  Proj o;
  o.Wx = x.R[0] * p.x + x.R[1] * p.y + x.R[2] * p.z + x.t[0];
  o.Wy = x.R[3] * p.x + x.R[4] * p.y + x.R[5] * p.z + x.t[1];
  o.Wz = x.R[6] * p.x + x.R[7] * p.y + x.R[8] * p.z + x.t[2];
  o.u = fdiv<PREC>(o.Wx, o.Wz) * g.fx + g.cx;
  o.v = fdiv<PREC>(o.Wy, o.Wz) * g.fy + g.cy;
  o.in = o.u > 1.0f && o.v > 1.0f && o.u < wlim && o.v < hlim;
  return o;
}

// -v = dr/dxi of the photometric residual (SE3Tracker::calculateWarpUpdate).
template <int PREC>
__device__ inline void jac6(float px, float py, float pz, float gx, float gy, float* v) {
  float z = frcp<PREC>(pz);
  float z_sqr = frcp<PREC>(pz * pz);
  v[0] = z * gx;
  v[1] = z * gy;
  v[2] = (-px * z_sqr) * gx + (-py * z_sqr) * gy;
  v[3] = (-px * py * z_sqr) * gx + (-(1.0f + py * py * z_sqr)) * gy;
  v[4] = (1.0f + px * px * z_sqr) * gx + (px * py * z_sqr) * gy;
  v[5] = (-py * z) * gx + (px * z) * gy;
}

// Per-thread accumulators of the per-point FP32 terms.
//  ACC_F64:   each FP32 term is widened and summed in FP64 (the parity mode:
//             the same terms as the CPU restatement, summed to ~1e-16).
//  ACC_KAHAN: compensated FP32.  Each thread keeps a (sum, correction) pair
//             per value and adds a term with Kahan's update (4 FP32
//             operations); the wave and block trees merge pairs with TwoSum
//             (Knuth), so a block's total is its FP32 pair widened to FP64
//             once, and only the cross-block merge runs in FP64.  The error
//             of a block total is ~2 * 2^-24 * sum |term|, independent of the
//             point count.
constexpr int ACC_F64 = 0;
constexpr int ACC_KAHAN = 1;

// lo + up after swapping lo's upper half-wave (offset 32) or odd rows
// (offset 16) with up's lower half-wave / even rows, 32 bits at a time
// (v_permlane32_swap / v_permlane16_swap).
__device__ inline double swap_sum(double lo, double up, int offset) {
  const unsigned long long l = (unsigned long long)__double_as_longlong(lo);
  const unsigned long long u = (unsigned long long)__double_as_longlong(up);
  unsigned l0 = (unsigned)l, l1 = (unsigned)(l >> 32), u0 = (unsigned)u, u1 = (unsigned)(u >> 32);
  if (offset == 32) {
    auto r0 = __builtin_amdgcn_permlane32_swap(l0, u0, false, false);
    auto r1 = __builtin_amdgcn_permlane32_swap(l1, u1, false, false);
    l0 = r0[0]; u0 = r0[1]; l1 = r1[0]; u1 = r1[1];
  } else {
    auto r0 = __builtin_amdgcn_permlane16_swap(l0, u0, false, false);
    auto r1 = __builtin_amdgcn_permlane16_swap(l1, u1, false, false);
    l0 = r0[0]; u0 = r0[1]; l1 = r1[0]; u1 = r1[1];
  }
  const double nl = __longlong_as_double((long long)(((unsigned long long)l1 << 32) | l0));
  const double nu = __longlong_as_double((long long)(((unsigned long long)u1 << 32) | u0));
  return nl + nu;
}

// Reduce-scatter over the wave (NS <= 64): afterwards lane l holds the wave
// total of value l (0 for l >= NS).  Halving butterfly: at offset o each lane
// keeps one half of its values and receives the other half from lane ^ o.
// Offsets 32 and 16 use the permlane swaps (the sum of the swapped pair is
// already the kept half: no lane selects); offsets 8..1 shuffle.
template <int NS>
__device__ inline double wave_reduce_scatter_f64(const double* v) {
  static_assert(NS <= 64, "reduce-scatter holds one value per lane");
  const int lane = threadIdx.x & 63;
  double a[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) a[i] = i < NS ? swap_sum(v[i], (i + 32 < NS) ? v[i + 32] : 0.0, 32) : 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = swap_sum(a[i], a[i + 16], 16);
#pragma unroll
  for (int h = 8; h >= 1; h >>= 1) {
    const bool hi = (lane & h) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      double send = hi ? a[i] : a[i + h];
      double keep = hi ? a[i + h] : a[i];
      a[i] = keep + __shfl_xor(send, h);
    }
  }
  return a[0];
}

template <int MODE, int NS>
struct Accum;

template <int NS>
struct Accum<ACC_F64, NS> {
  double v[NS];
  __device__ inline void zero() {
#pragma unroll
    for (int k = 0; k < NS; ++k) v[k] = 0.0;
  }
  __device__ inline void add(int k, float x) { v[k] += (double)x; }
  __device__ inline void wave_reduce() {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double a = v[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
      v[k] = a;
    }
  }
  __device__ inline double value(int k) const { return v[k]; }
  __device__ inline double wave_slot() const { return wave_reduce_scatter_f64<NS>(v); }
  __device__ static inline double merge_waves(const double* cells, int stride, int waves) {
    double t = 0.0;
    for (int w = 0; w < waves; ++w) t += cells[(size_t)w * stride];
    return t;
  }
};

// Kahan's compensated update of the pair (s, c), whose value is s + c.
__device__ inline void kahan_add(float& s, float& c, float x) {
  const float y = x + c;
  const float t = s + y;
  c = y - (t - s);
  s = t;
}

// (s, c) += (s2, c2): the sums' rounding error is recovered exactly with
// TwoSum and folded into the correction.
__device__ inline void pair_merge(float& s, float& c, float s2, float c2) {
  const float t = s + s2;
  const float bp = t - s;
  const float e = (s - (t - bp)) + (s2 - bp);
  c = (c + c2) + e;
  s = t;
}

// A pair packed into one 64-bit LDS cell (low word sum, high word correction).
__device__ inline double pack_pair(float s, float c) {
  return __longlong_as_double((long long)(((unsigned long long)__float_as_uint(c) << 32) | __float_as_uint(s)));
}
__device__ inline void unpack_pair(double cell, float& s, float& c) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(cell);
  s = __uint_as_float((unsigned)b);
  c = __uint_as_float((unsigned)(b >> 32));
}

// The reduce-scatter of wave_reduce_scatter_f64 over (sum, correction) pairs:
// afterwards lane l holds the wave's pair of value l ((0, 0) for l >= NS).
template <int NS>
__device__ inline void wave_reduce_scatter_pair(const float* vs, const float* vc, float& rs, float& rc) {
  static_assert(NS <= 64, "reduce-scatter holds one value per lane");
  const int lane = threadIdx.x & 63;
  float s[32], c[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    float ls = i < NS ? vs[i] : 0.f, lc = i < NS ? vc[i] : 0.f;
    float us = i + 32 < NS ? vs[i + 32] : 0.f, uc = i + 32 < NS ? vc[i + 32] : 0.f;
    auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(ls), __float_as_uint(us), false, false);
    auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(lc), __float_as_uint(uc), false, false);
    s[i] = __uint_as_float(a[0]);
    c[i] = __uint_as_float(b[0]);
    pair_merge(s[i], c[i], __uint_as_float(a[1]), __uint_as_float(b[1]));
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(s[i]), __float_as_uint(s[i + 16]), false, false);
    auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(c[i]), __float_as_uint(c[i + 16]), false, false);
    s[i] = __uint_as_float(a[0]);
    c[i] = __uint_as_float(b[0]);
    pair_merge(s[i], c[i], __uint_as_float(a[1]), __uint_as_float(b[1]));
  }
#pragma unroll
  for (int h = 8; h >= 1; h >>= 1) {
    const bool hi = (lane & h) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      const float ss = hi ? s[i] : s[i + h], sc = hi ? c[i] : c[i + h];
      float ks = hi ? s[i + h] : s[i], kc = hi ? c[i + h] : c[i];
      pair_merge(ks, kc, __shfl_xor(ss, h), __shfl_xor(sc, h));
      s[i] = ks;
      c[i] = kc;
    }
  }
  rs = s[0];
  rc = c[0];
}

template <int NS>
struct Accum<ACC_KAHAN, NS> {
  float s[NS], c[NS];
  __device__ inline void zero() {
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = c[k] = 0.f;
  }
  __device__ inline void add(int k, float x) { kahan_add(s[k], c[k], x); }
  // The LDS cell lane l of each wave writes for value l.
  __device__ inline double wave_slot() const {
    float rs, rc;
    wave_reduce_scatter_pair<NS>(s, c, rs, rc);
    return pack_pair(rs, rc);
  }
  // Block total of one value from its wave cells (cells[w * stride], wave
  // order), widened to FP64.
  __device__ static inline double merge_waves(const double* cells, int stride, int waves) {
    float ts, tc;
    unpack_pair(cells[0], ts, tc);
    for (int w = 1; w < waves; ++w) {
      float ws, wc;
      unpack_pair(cells[(size_t)w * stride], ws, wc);
      pair_merge(ts, tc, ws, wc);
    }
    return (double)ts + (double)tc;
  }
};

// Index of (i, j), i <= j, in the row-major upper triangle of a 7x7 matrix.
__host__ __device__ constexpr int tri7(int i, int j) { return i * 7 - i * (i - 1) / 2 + (j - i); }

// One reference point of an SE3 evaluation: warp, bounds, taps, residual,
// isGood, affine statistics, weights, Jacobian and normal-equation terms.
// Rest of one SE3 point once its projection and taps are in hand.
template <int PREC, class ACC>
__device__ inline void se3_finish(const float4& p, float pvar, int pidx, const Proj& w, const TapRaw<PREC>& t,
                                  const Xf& x, const LevelGeom& g, const TrackKParams& kp,
                                  uint8_t* __restrict__ isGood, ACC& acc) {
  const float Wx = w.Wx, Wy = w.Wy, Wz = w.Wz;
  if (!w.in) {
    if (isGood) isGood[pidx] = 0;
    return;
  }
  Taps ri = blend_taps<PREC>(t);
  float c1 = kp.aff_a * p.w + kp.aff_b;
  float c2 = ri.I;
  float r = c1 - c2;
  if (kp.aff_stats) {  // affine-brightness statistics (read only when that estimation is on)
    float wgt = fabsf(r) < 5.0f ? 1.0f : fdiv<PREC>(5.0f, fabsf(r));
    acc.add(S3_SXX, c1 * c1 * wgt);
    acc.add(S3_SYY, c2 * c2 * wgt);
    acc.add(S3_SX, c1 * wgt);
    acc.add(S3_SY, c2 * wgt);
    acc.add(S3_SW, wgt);
  }
  bool good = fdiv<PREC>(r * r, kp.max_diff_constant + kp.max_diff_grad_mult * (ri.gx * ri.gx + ri.gy * ri.gy)) < 1.0f;
  if (isGood) isGood[pidx] = good ? 1 : 0;
  float gx = g.fx * ri.gx, gy = g.fy * ri.gy;
  float d = frcp<PREC>(p.z);
  if (good) {
    acc.add(S3_RESUNW, r * r);
    acc.add(S3_SIGNED, r);
    acc.add(S3_GOOD, 1.0f);
  } else {
    acc.add(S3_BAD, 1.0f);
  }
  float dc = fdiv<PREC>(p.z, Wz);
  acc.add(S3_USAGE, dc < 1.0f ? dc : 1.0f);
  acc.add(S3_WARPED, 1.0f);
  // calcWeightsAndResidual
  float s = kp.var_weight * pvar;
  float g0 = fdiv<PREC>(x.t[0] * Wz - x.t[2] * Wx, Wz * Wz * d);
  float g1 = fdiv<PREC>(x.t[1] * Wz - x.t[2] * Wy, Wz * Wz * d);
  float drpdd = gx * g0 + gy * g1;
  float w_p = frcp<PREC>(kp.noise2 + s * drpdd * drpdd);
  float wr = fabsf(r * fsqrt<PREC>(w_p));
  float wh = fabsf(wr < kp.huber ? 1.0f : fdiv<PREC>(kp.huber, wr));
  float W = wh * w_p;
  // calculateWarpUpdate
  float jv[6];
  jac6<PREC>(Wx, Wy, Wz, gx, gy, jv);
  int k = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int c = a; c < 6; ++c) acc.add(S3_A + (k++), jv[a] * jv[c] * W);
  float rW = r * W;
#pragma unroll
  for (int a = 0; a < 6; ++a) acc.add(S3_B + a, -(jv[a] * rW));
  acc.add(S3_ERR, r * r * W);
}

// SE3 evaluation of points i, i + stride, ... < end of one thread, U points
// at a time: the U points' loads are issued, then their projections' taps,
// then the U points are finished in index order (the per-thread summation
// order of a one-point-at-a-time loop).
template <int PREC, int U, class ACC>
__device__ inline void se3_points(const PointsLevel& pts, int i, int stride, int end, const void* __restrict__ img,
                                  const LevelGeom& g, const Xf& x, const TrackKParams& kp,
                                  uint8_t* __restrict__ isGood, ACC& acc) {
  for (; i < end; i += U * stride) {
    float4 p[U];
    float pv[U];
    int pi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = i + u * stride;
      if (k < end) {
        p[u] = pts.posI[k];
        pv[u] = pts.var[k];
        pi[u] = isGood ? pts.idx[k] : 0;
      }
    }
    Proj w[U];
    TapRaw<PREC> t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i + u * stride < end) {
        w[u] = project<PREC>(p[u], x, g);
        if (w[u].in) fetch_taps<PREC>(img, w[u].u, w[u].v, g.w, t[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i + u * stride < end) se3_finish<PREC>(p[u], pv[u], pi[u], w[u], t[u], x, g, kp, isGood, acc);
  }
}

// Rest of one Sim3 point (photometric + depth residual) once its
// projection, taps and the frame's depth at the rounded pixel are in hand.
template <int PREC, class ACC>
__device__ inline void sim3_finish(const float4& p, float pvar, float2 rg, const Proj& w, const TapRaw<PREC>& t,
                                   float fidR, float varF, const Xf& x, const LevelGeom& g, const TrackKParams& kp,
                                   ACC& acc) {
  if (!w.in) return;
  const float Wx = w.Wx, Wy = w.Wy, Wz = w.Wz;
  Taps ri = blend_taps<PREC>(t);
  float gx, gy;
  if (kp.use_esm) {
    float rgx = kp.roll[0] * rg.x + kp.roll[1] * rg.y;
    float rgy = kp.roll[2] * rg.x + kp.roll[3] * rg.y;
    gx = g.fx * 0.5f * (ri.gx + rgx);
    gy = g.fy * 0.5f * (ri.gy + rgy);
  } else {
    gx = g.fx * ri.gx;
    gy = g.fy * ri.gy;
  }
  float c1 = kp.aff_a * p.w + kp.aff_b;
  float c2 = ri.I;
  float rp = c1 - c2;
  if (kp.aff_stats) {
    float wgt = fabsf(rp) < 2.0f ? 1.0f : fdiv<PREC>(2.0f, fabsf(rp));
    acc.add(SM_SXX, c1 * c1 * wgt);
    acc.add(SM_SYY, c2 * c2 * wgt);
    acc.add(SM_SX, c1 * wgt);
    acc.add(SM_SY, c2 * wgt);
    acc.add(SM_SW, wgt);
  }
  float refId = frcp<PREC>(Wz);
  float d = frcp<PREC>(p.z);
  float rd, vard;
  if (varF > 0.0f) {
    rd = refId - fidR;
    vard = varF;
  } else {
    rd = -1.0f;
    vard = -1.0f;
  }
  float dc = fdiv<PREC>(p.z, Wz);
  acc.add(SM_USAGE, dc < 1.0f ? dc : 1.0f);
  acc.add(SM_WARPED, 1.0f);
  // calcSim3WeightsAndResidual
  float s = kp.var_weight * pvar;
  float sv = kp.var_weight * vard;
  float g0 = fdiv<PREC>(x.t[0] * Wz - x.t[2] * Wx, Wz * Wz * d);
  float g1 = fdiv<PREC>(x.t[1] * Wz - x.t[2] * Wy, Wz * Wz * d);
  float g2 = fdiv<PREC>(Wz - x.t[2], Wz * Wz * d);
  float drpdd = gx * g0 + gy * g1;
  float w_p = frcp<PREC>(kp.noise2 + s * drpdd * drpdd);
  float w_d = frcp<PREC>(sv + g2 * g2 * s);
  float wrd = fabsf(rd * fsqrt<PREC>(w_d));
  float wrp = fabsf(rp * fsqrt<PREC>(w_p));
  float wabs = sv > 0.0f ? wrd + wrp : wrp;
  float wh = fabsf(wabs < kp.huber ? 1.0f : fdiv<PREC>(kp.huber, wabs));
  if (sv > 0.0f) {
    acc.add(SM_RESD, wh * w_d * rd * rd);
    acc.add(SM_NTD, 1.0f);
  }
  acc.add(SM_RESP, wh * w_p * rp * rp);
  float wp = wh * w_p;
  float wd = sv > 0.0f ? wh * w_d : 0.0f;
  // calcSim3LGS: LGS6 over (v, w) and LGS4 over (v3, w1, w2, sigma), merged
  // into the 7x7 upper triangle at rows / columns {2, 3, 4, 6}.
  float jv[6];
  jac6<PREC>(Wx, Wy, Wz, gx, gy, jv);
  float z = frcp<PREC>(Wz);
  float z_sqr = frcp<PREC>(Wz * Wz);
  float v4[4] = {z_sqr, z_sqr * Wy, -z_sqr * Wx, z};
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int c = a; c < 6; ++c) acc.add(SM_A + tri7(a, c), jv[a] * jv[c] * wp);
  float rpW = rp * wp;
#pragma unroll
  for (int a = 0; a < 6; ++a) acc.add(SM_B + a, -(jv[a] * rpW));
  if (wd != 0.0f) {
    const int remap[4] = {2, 3, 4, 6};
    float rdW = rd * wd;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
      for (int c = a; c < 4; ++c) acc.add(SM_A + tri7(remap[a], remap[c]), v4[a] * v4[c] * wd);
      acc.add(SM_B + remap[a], -(v4[a] * rdW));
    }
  }
}

// Sim3 evaluation of points i, i + stride, ... < end of one thread, U points
// at a time (loads first, then taps and the frame depth at the rounded
// projection, then the points in index order).
template <int PREC, int U, class ACC>
__device__ inline void sim3_points(const PointsLevel& pts, int i, int stride, int end, const void* __restrict__ img,
                                   const float* __restrict__ fid, const float* __restrict__ fvar, const LevelGeom& g,
                                   const Xf& x, const TrackKParams& kp, ACC& acc) {
  for (; i < end; i += U * stride) {
    float4 p[U];
    float pv[U];
    float2 rg[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = i + u * stride;
      if (k < end) {
        p[u] = pts.posI[k];
        pv[u] = pts.var[k];
        rg[u] = kp.use_esm ? pts.grad[k] : make_float2(0.f, 0.f);
      }
    }
    Proj w[U];
    TapRaw<PREC> t[U];
    float fd[U], fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i + u * stride < end) {
        w[u] = project<PREC>(p[u], x, g);
        if (w[u].in) {
          fetch_taps<PREC>(img, w[u].u, w[u].v, g.w, t[u]);
          int ir = (int)(w[u].u + 0.5f) + g.w * (int)(w[u].v + 0.5f);
          fv[u] = fvar[ir];
          fd[u] = fid[ir];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i + u * stride < end) sim3_finish<PREC>(p[u], pv[u], rg[u], w[u], t[u], fd[u], fv[u], x, g, kp, acc);
  }
}

// One reference point of the adjoint: dE/dxi into acc[0..6), E into acc[6],
// count into acc[7], dE/d(idepth) into grad_d[i].
template <int PREC, class ACC>
__device__ inline void adjoint_point(const PointsLevel& pts, int i, const void* __restrict__ img, const LevelGeom& g,
                                     const Xf& x, const TrackKParams& kp, float* __restrict__ grad_d, ACC& acc) {
  const float wlim = (float)(g.w - 2), hlim = (float)(g.h - 2);
  float4 p = pts.posI[i];
  float Wx = x.R[0] * p.x + x.R[1] * p.y + x.R[2] * p.z + x.t[0];
  float Wy = x.R[3] * p.x + x.R[4] * p.y + x.R[5] * p.z + x.t[1];
  float Wz = x.R[6] * p.x + x.R[7] * p.y + x.R[8] * p.z + x.t[2];
  float u = (Wx / Wz) * g.fx + g.cx;
  float v = (Wy / Wz) * g.fy + g.cy;
  if (!(u > 1.0f && v > 1.0f && u < wlim && v < hlim)) {
    grad_d[i] = 0.0f;
    return;
  }
  Taps ri = interp3<PREC>(img, u, v, g.w);
  float r = p.w - ri.I;
  float gx = g.fx * ri.gx, gy = g.fy * ri.gy;
  float d = 1.0f / p.z;
  float s = kp.var_weight * pts.var[i];
  float g0 = (x.t[0] * Wz - x.t[2] * Wx) / (Wz * Wz * d);
  float g1 = (x.t[1] * Wz - x.t[2] * Wy) / (Wz * Wz * d);
  float drpdd = gx * g0 + gy * g1;
  float w_p = 1.0f / (kp.noise2 + s * drpdd * drpdd);
  float e = fabsf(r * sqrtf(w_p));
  float k = kp.huber;
  float wh = e < k ? 1.0f : k / e;
  acc.add(6, e < k ? 0.5f * e * e : k * e - 0.5f * k * k);
  acc.add(7, 1.0f);
  float Wr = wh * w_p * r;
  float jv[6];
  jac6<PREC>(Wx, Wy, Wz, gx, gy, jv);
#pragma unroll
  for (int a = 0; a < 6; ++a) acc.add(a, -(Wr * jv[a]));
  grad_d[i] = -Wr * drpdd;
}

// One reference point of the Sim3 adjoint.  E is the Sim3 tracker's
// reweighted least-squares energy with its weights frozen at x:
//   E = 1/2 sum_i wh_i (w_p,i r_p,i^2 + w_d,i r_d,i^2),
// wh the joint Huber weight, w_p / w_d the photometric / depth variance
// weights, r_d = 1/z' - idepth_frame(nearest pixel); the frame's idepth
// sample, the weights and the in-image set are constants of E.  Its twist
// gradient is Sim3Tracker::calcSim3LGS's b (unnormalised) with the frame's own
// gradient in place of the ESM average (kp.use_esm must be 0).  The
// expressions are sim3_finish's, in the same order.
//   dE/dd_i = -wh w_p r_p drpdd + wh w_d r_d g2   (g2 = d(1/z')/dd).
template <int PREC, class ACC>
__device__ inline void adjoint_point_sim3(const PointsLevel& pts, int i, const void* __restrict__ img,
                                          const float* __restrict__ fid, const float* __restrict__ fvar,
                                          const LevelGeom& g, const Xf& x, const TrackKParams& kp,
                                          float* __restrict__ grad_d, ACC& acc) {
  const float4 p = pts.posI[i];
  const Proj w = project<PREC>(p, x, g);
  if (!w.in) {
    grad_d[i] = 0.0f;
    return;
  }
  const float Wx = w.Wx, Wy = w.Wy, Wz = w.Wz;
  Taps ri = interp3<PREC>(img, w.u, w.v, g.w);
  const int ir = (int)(w.u + 0.5f) + g.w * (int)(w.v + 0.5f);
  const float varF = fvar[ir], fidR = fid[ir];
  float gx = g.fx * ri.gx, gy = g.fy * ri.gy;
  float rp = (kp.aff_a * p.w + kp.aff_b) - ri.I;
  float refId = frcp<PREC>(Wz);
  float d = frcp<PREC>(p.z);
  float rd, vard;
  if (varF > 0.0f) {
    rd = refId - fidR;
    vard = varF;
  } else {
    rd = -1.0f;
    vard = -1.0f;
  }
  float s = kp.var_weight * pts.var[i];
  float sv = kp.var_weight * vard;
  float g0 = fdiv<PREC>(x.t[0] * Wz - x.t[2] * Wx, Wz * Wz * d);
  float g1 = fdiv<PREC>(x.t[1] * Wz - x.t[2] * Wy, Wz * Wz * d);
  float g2 = fdiv<PREC>(Wz - x.t[2], Wz * Wz * d);
  float drpdd = gx * g0 + gy * g1;
  float w_p = frcp<PREC>(kp.noise2 + s * drpdd * drpdd);
  float w_d = frcp<PREC>(sv + g2 * g2 * s);
  float wrd = fabsf(rd * fsqrt<PREC>(w_d));
  float wrp = fabsf(rp * fsqrt<PREC>(w_p));
  float wabs = sv > 0.0f ? wrd + wrp : wrp;
  float wh = fabsf(wabs < kp.huber ? 1.0f : fdiv<PREC>(kp.huber, wabs));
  float wp = wh * w_p;
  float wd = sv > 0.0f ? wh * w_d : 0.0f;
  acc.add(AS_EP, wp * rp * rp);
  if (sv > 0.0f) acc.add(AS_ED, wh * w_d * rd * rd);
  acc.add(AS_COUNT, 1.0f);
  float jv[6];
  jac6<PREC>(Wx, Wy, Wz, gx, gy, jv);
  float rpW = rp * wp;
#pragma unroll
  for (int a = 0; a < 6; ++a) acc.add(AS_G + a, -(jv[a] * rpW));
  float gd = -rpW * drpdd;
  if (wd != 0.0f) {
    float z = frcp<PREC>(Wz);
    float z_sqr = frcp<PREC>(Wz * Wz);
    float v4[4] = {z_sqr, z_sqr * Wy, -z_sqr * Wx, z};
    const int remap[4] = {2, 3, 4, 6};
    float rdW = rd * wd;
#pragma unroll
    for (int a = 0; a < 4; ++a) acc.add(AS_G + remap[a], -(v4[a] * rdW));
    gd = gd + rdW * g2;
  }
  grad_d[i] = gd;
}

}  // namespace lsdg
