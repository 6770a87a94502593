// oracle/lsd_oracle.hpp — TEST INFRASTRUCTURE (the oracle).  Not product code:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it, and only as the checker or the timed CPU baseline.
//
// CPU restatement of LSD-SLAM's per-pixel front end, templated on the pixel
// arithmetic type S: S = float is the "restated reference CPU path" (the
// upstream 3-pass tracker structure, sequential FP32 sums, row-chunked
// threads for the depth map), S = double is the FP64 accuracy oracle.
//
// The arithmetic lives in the un-vendored fips dependency `lsd-slam` @ branch
// `unstable` (fips.yml:2-4), absent from the reference tree and from this
// machine.  Every function below restates the published method (Engel et al.
// ICCV 2013 / ECCV 2014) and the upstream lsd_slam_core structure; each cites
// the upstream symbol it follows and, where the wrapper holds evidence, the
// wrapper file:line.  The operation order of every per-pixel float expression
// is fixed here, and the HIP kernels follow the same order so that masks and
// per-pixel values match bit for bit.
//
// Conventions for quantities the upstream core leaves uninitialised: pyramid
// rows the upstream loops never write (gradient rows 0 and h-1, max-gradient
// rows 0 and h-1, the temporary max-gradient buffer outside its written range)
// are zero here.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "../include/lsdg.h"
#include "lie.hpp"

namespace orc {

// SplitMix64 (Steele, Lea, Flood 2014): output number i + 1 from state
// `seed`, top 31 bits: the draw DepthMap::initializeRandomly takes from
// rand() upstream.
inline int random_draw(uint64_t seed, int i) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (int)(z >> 33);
}
inline uint64_t splitmix64_output(uint64_t seed, int i) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr int PYR = LSDG_PYRAMID_LEVELS;

void params_default(lsdg_params* p);

// getInterpolatedElement (upstream util/globalFuncs.h): bilinear with weights
// dxdy, dy-dxdy, dx-dxdy, 1-dx-dy+dxdy, summed in that order.
template <class S>
inline S interp(const S* mat, S x, S y, int w) {
  int ix = (int)x, iy = (int)y;
  S dx = x - (S)ix, dy = y - (S)iy;
  S dxdy = dx * dy;
  const S* bp = mat + ix + iy * w;
  return dxdy * bp[1 + w] + (dy - dxdy) * bp[w] + (dx - dxdy) * bp[1] + ((S)1 - dx - dy + dxdy) * bp[0];
}
// getInterpolatedElement43/42 on 4-lane gradient images (gx, gy, I, 0).
template <class S>
inline void interp4(const S* g4, S x, S y, int w, int lanes, S* out) {
  int ix = (int)x, iy = (int)y;
  S dx = x - (S)ix, dy = y - (S)iy;
  S dxdy = dx * dy;
  const S* bp = g4 + 4 * (ix + iy * w);
  S w11 = dxdy, w01 = dy - dxdy, w10 = dx - dxdy, w00 = (S)1 - dx - dy + dxdy;
  for (int c = 0; c < lanes; ++c)
    out[c] = w11 * bp[4 * (1 + w) + c] + w01 * bp[4 * w + c] + w10 * bp[4 + c] + w00 * bp[c];
}

template <class S>
inline S unzero(S v) {
  return v < 0 ? (v > (S)-1e-10 ? (S)-1e-10 : v) : (v < (S)1e-10 ? (S)1e-10 : v);
}

// ----------------------------------------------------------------- Frame
// Upstream DataStructures/Frame: pyramids image(l), gradients(l) (gx,gy,I,pad),
// maxGradients(l), idepth(l)/idepthVar(l); pose fields thisToParent_raw,
// trackingParent, initialTrackedResidual; refPixelWasGood (level-1 mask,
// default all true).  Wrapper evidence: lib/GUI.h:24,
// PangolinOutputIOWrapper.cpp:50-79, TextOutputIOWrapper.cpp:107-114.
template <class S>
struct Frame {
  int id = 0;
  int w[PYR], h[PYR];
  S fx[PYR], fy[PYR], cx[PYR], cy[PYR], fxi[PYR], fyi[PYR], cxi[PYR], cyi[PYR];
  std::vector<S> image[PYR], grad[PYR], maxg[PYR], idepth[PYR], ivar[PYR];
  bool hasDepth = false;
  float numMappablePixels = 0;

  // pose
  int parentId = -1;
  Frame* parent = nullptr;
  Sim3 thisToParent;
  S initialTrackedResidual = 0;
  std::vector<uint8_t> wasGood;  // level 1
  bool wasGoodAllocated = false;
  Sim3 camToWorld;
  int numFramesTrackedOnThis = 0, numMappedOnThis = 0;

  // prepareForStereoWith products
  S K_otherToThis_R[9], otherToThis_t[3], K_otherToThis_t[3], thisToOther_t[3], K_thisToOther_t[3];
  S R_row[3][3];  // otherToThis_R_row0..2
  double distSquared = 0;

  // Frame::initialize: level intrinsics fx_l = fx/2^l,
  // cx_l = (cx_{l-1} + 0.5) * 0.5 - 0.5, inverse fxi = 1/fx_l, cxi = -cx_l/fx_l.
  void init(int id_, int w0, int h0, float fx0, float fy0, float cx0, float cy0) {
    id = id_;
    for (int l = 0; l < PYR; ++l) {
      w[l] = l ? w[l - 1] / 2 : w0;
      h[l] = l ? h[l - 1] / 2 : h0;
      if (l == 0) {
        fx[0] = fx0; fy[0] = fy0; cx[0] = cx0; cy[0] = cy0;
      } else {
        fx[l] = fx[l - 1] * (S)0.5;
        fy[l] = fy[l - 1] * (S)0.5;
        cx[l] = (cx[l - 1] + (S)0.5) * (S)0.5 - (S)0.5;
        cy[l] = (cy[l - 1] + (S)0.5) * (S)0.5 - (S)0.5;
      }
      fxi[l] = (S)1 / fx[l];
      fyi[l] = (S)1 / fy[l];
      cxi[l] = -cx[l] / fx[l];
      cyi[l] = -cy[l] / fy[l];
    }
  }

  // Frame::buildImage: 2x2 box, ((TL + TR) + BL) + BR, times 0.25.  Rows/cols
  // beyond 2*(w_l, h_l) of an odd-sized parent are ignored.
  void buildImage(int l) {
    int sw = w[l - 1];
    const S* s = image[l - 1].data();
    image[l].assign((size_t)w[l] * h[l], (S)0);
    for (int y = 0; y < h[l]; ++y)
      for (int x = 0; x < w[l]; ++x) {
        const S* p = s + 2 * x + 2 * y * sw;
        image[l][x + y * w[l]] = (p[0] + p[1] + p[sw] + p[sw + 1]) * (S)0.25;
      }
  }
  // Frame::buildGradients on the linear range [w, w(h-1)): central half
  // differences; columns 0 and w-1 read the neighbouring row.
  void buildGradients(int l) {
    int W = w[l], H = h[l];
    grad[l].assign((size_t)4 * W * H, (S)0);
    const S* I = image[l].data();
    for (int idx = W; idx < W * (H - 1); ++idx) {
      grad[l][4 * idx + 0] = (S)0.5 * (I[idx + 1] - I[idx - 1]);
      grad[l][4 * idx + 1] = (S)0.5 * (I[idx + W] - I[idx - W]);
      grad[l][4 * idx + 2] = I[idx];
    }
  }
  // Frame::buildMaxGradients: |grad| on interior rows, vertical 3-max into a
  // temporary, horizontal 3-max back; level 0 counts mappable pixels.
  void buildMaxGradients(int l, float minAbsGradCreate) {
    int W = w[l], H = h[l];
    maxg[l].assign((size_t)W * H, (S)0);
    std::vector<S> tmp((size_t)W * H, (S)0);
    S* m = maxg[l].data();
    const S* g = grad[l].data();
    for (int idx = W; idx < W * (H - 1); ++idx) {
      S dx = g[4 * idx], dy = g[4 * idx + 1];
      m[idx] = std::sqrt(dx * dx + dy * dy);
    }
    for (int idx = W + 1; idx < W * (H - 1) - 1; ++idx) {
      S g1 = m[idx - W], g2 = m[idx];
      if (g1 < g2) g1 = g2;
      S g3 = m[idx + W];
      tmp[idx] = (g1 < g3) ? g3 : g1;
    }
    float num = 0;
    for (int idx = W + 1; idx < W * (H - 1) - 1; ++idx) {
      S g1 = tmp[idx - 1], g2 = tmp[idx];
      if (g1 < g2) g1 = g2;
      S g3 = tmp[idx + 1];
      S v = (g1 < g3) ? g3 : g1;
      m[idx] = v;
      if (v >= (S)minAbsGradCreate) num++;
    }
    if (l == 0) numMappablePixels = num;
  }
  void setImage(const float* img, const lsdg_params& P) {
    image[0].assign(img, img + (size_t)w[0] * h[0]);
    for (int l = 1; l < PYR; ++l) buildImage(l);
    for (int l = 0; l < PYR; ++l) {
      buildGradients(l);
      buildMaxGradients(l, P.min_abs_grad_create);
    }
  }
  // Frame::buildIDepthAndIDepthVar: inverse-variance fusion of the valid
  // (var > 0) children; idepth = 1 / (sum ivar / sum ivar*idepth),
  // var = n / sum ivar; none valid -> -1, -1.
  void buildDepthLevel(int l) {
    int W = w[l], H = h[l], sw = w[l - 1];
    idepth[l].assign((size_t)W * H, (S)-1);
    ivar[l].assign((size_t)W * H, (S)-1);
    const S* ids = idepth[l - 1].data();
    const S* vs = ivar[l - 1].data();
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        int idx = 2 * (x + y * sw);
        S isum = 0, ivsum = 0;
        int num = 0;
        const int off[4] = {0, 1, sw, sw + 1};
        for (int k = 0; k < 4; ++k) {
          S var = vs[idx + off[k]];
          if (var > 0) {
            S iv = (S)1 / var;
            ivsum += iv;
            isum += iv * ids[idx + off[k]];
            num++;
          }
        }
        if (num > 0) {
          S depth = ivsum / isum;
          idepth[l][x + y * W] = (S)1 / depth;
          ivar[l][x + y * W] = (S)num / ivsum;
        }
      }
  }
  void setDepth0(const S* id, const S* var) {
    size_t n = (size_t)w[0] * h[0];
    idepth[0].assign(id, id + n);
    ivar[0].assign(var, var + n);
    for (int l = 1; l < PYR; ++l) buildDepthLevel(l);
    hasDepth = true;
  }
  uint8_t* refPixelWasGood() {
    if (!wasGoodAllocated) {
      wasGood.assign((size_t)w[1] * h[1], 1);
      wasGoodAllocated = true;
    }
    return wasGood.data();
  }
  const uint8_t* refPixelWasGoodNoCreate() const { return wasGoodAllocated ? wasGood.data() : nullptr; }

  // Frame::prepareForStereoWith(other, thisToOther, K, 0): this = reference
  // frame, other = keyframe.
  void prepareForStereoWith(const Sim3& thisToOther, const S* K /*3x3*/) {
    Sim3 o2t = thisToOther.inverse();
    double Ro[9], Rt[9];
    o2t.R(Ro);
    thisToOther.R(Rt);
    S M[9];
    for (int i = 0; i < 9; ++i) M[i] = (S)(Ro[i] * o2t.s);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        K_otherToThis_R[3 * i + j] = K[3 * i] * M[j] + K[3 * i + 1] * M[3 + j] + K[3 * i + 2] * M[6 + j];
    for (int i = 0; i < 3; ++i) otherToThis_t[i] = (S)o2t.t[i];
    for (int i = 0; i < 3; ++i)
      K_otherToThis_t[i] = K[3 * i] * otherToThis_t[0] + K[3 * i + 1] * otherToThis_t[1] + K[3 * i + 2] * otherToThis_t[2];
    for (int i = 0; i < 3; ++i) thisToOther_t[i] = (S)thisToOther.t[i];
    for (int i = 0; i < 3; ++i)
      K_thisToOther_t[i] = K[3 * i] * thisToOther_t[0] + K[3 * i + 1] * thisToOther_t[1] + K[3 * i + 2] * thisToOther_t[2];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) R_row[k][i] = (S)(Rt[3 * i + k] * thisToOther.s);
    distSquared = o2t.t[0] * o2t.t[0] + o2t.t[1] * o2t.t[1] + o2t.t[2] * o2t.t[2];
  }
};

// ---------------------------------------------------- TrackingReference
// Upstream Tracking/TrackingReference::makePointCloud(level): for x in [1,w-1)
// (outer), y in [1,h-1) (inner), skip var <= 0 or idepth == 0; emit
// p = (1/idepth) * (fxi x + cxi, fyi y + cyi, 1), (gx, gy), (I, var), idx.
template <class S>
struct TrackingReference {
  Frame<S>* kf = nullptr;
  int num[PYR] = {-1, -1, -1, -1, -1};
  std::vector<S> pos[PYR], gradd[PYR], colVar[PYR];
  std::vector<int> idx[PYR];
  void makePointCloud(int l) {
    if (num[l] >= 0) return;
    Frame<S>& F = *kf;
    int W = F.w[l], H = F.h[l];
    const S* id = F.idepth[l].data();
    const S* iv = F.ivar[l].data();
    const S* col = F.image[l].data();
    const S* g = F.grad[l].data();
    pos[l].clear(); gradd[l].clear(); colVar[l].clear(); idx[l].clear();
    for (int x = 1; x < W - 1; ++x)
      for (int y = 1; y < H - 1; ++y) {
        int i = x + y * W;
        if (iv[i] <= 0 || id[i] == 0) continue;
        S inv = (S)1 / id[i];
        pos[l].push_back(inv * (F.fxi[l] * (S)x + F.cxi[l]));
        pos[l].push_back(inv * (F.fyi[l] * (S)y + F.cyi[l]));
        pos[l].push_back(inv * (S)1);
        gradd[l].push_back(g[4 * i]);
        gradd[l].push_back(g[4 * i + 1]);
        colVar[l].push_back(col[i]);
        colVar[l].push_back(iv[i]);
        idx[l].push_back(i);
      }
    num[l] = (int)idx[l].size();
  }
};

// float 3x3 (row-major) from a double pose, and translation.
template <class S>
inline void pose_RT(const Sim3& T, S* R, S* t, bool scaled) {
  double M[9];
  T.R(M);
  for (int i = 0; i < 9; ++i) R[i] = (S)(scaled ? M[i] * T.s : M[i]);
  for (int i = 0; i < 3; ++i) t[i] = (S)T.t[i];
}

// Sums of one fused evaluation (lsdg_eval layout, accumulated in the accumulator type).
template <class S>
struct EvalSums {
  S A[49] = {0}, b[7] = {0};
  S err = 0, sumResP = 0, sumResD = 0;
  int numTermsP = 0, numTermsD = 0, n = 0;
  int warped = 0, good = 0, bad = 0;
  S sumResUnw = 0, sumSigned = 0, usage = 0, sxx = 0, syy = 0, sx = 0, sy = 0, sw = 0;
};

// ------------------------------------------------------------ SE3Tracker
// Upstream Tracking/SE3Tracker: calcResidualAndBuffers, calcWeightsAndResidual,
// calculateWarpUpdate (LGS6), trackFrame.  Reached through
// SlamSystem::nextImage (lib/App/InputThread.cpp:71).
// S: per-point arithmetic; Acc: accumulation of the per-point terms and the
// LM error comparisons (float = upstream FP32 sums, double = FP64 accumulation).
template <class S, class Acc = S>
struct SE3Tracker {
  lsdg_params P;
  std::vector<S> bx, by, bz, bdx, bdy, bres, bd, bvar, bw;
  int warped = 0;
  Acc pointUsage = 0, lastGoodCount = 0, lastBadCount = 0, lastMeanRes = 0, lastResidual = 0;
  S aff_a = 1, aff_b = 0, aff_a_it = 1, aff_b_it = 0;
  bool diverged = false, trackingWasGood = true;
  int numResidualCalls[PYR] = {0}, numWarpCalls[PYR] = {0};
  EvalSums<Acc> last;  // statistics of the last residual pass

  // calcResidualAndBuffers: warp, bounds, bilinear (gx, gy, I), residual,
  // isGood, affine statistics, usage, compacted buffers.
  Acc calcResidualAndBuffers(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& refToFrame, int lvl,
                           uint8_t* isGoodOut) {
    int W = frame.w[lvl], H = frame.h[lvl];
    S fx = frame.fx[lvl], fy = frame.fy[lvl], cx = frame.cx[lvl], cy = frame.cy[lvl];
    S R[9], t[3];
    pose_RT(refToFrame, R, t, false);
    const S* grad = frame.grad[lvl].data();
    int N = ref.num[lvl];
    size_t cap = (size_t)N;
    bx.resize(cap); by.resize(cap); bz.resize(cap); bdx.resize(cap); bdy.resize(cap);
    bres.resize(cap); bd.resize(cap); bvar.resize(cap); bw.resize(cap);
    int idx = 0, goodCount = 0, badCount = 0;
    Acc sumResUnw = 0, sumSigned = 0, sxx = 0, syy = 0, sx = 0, sy = 0, sw = 0, usage = 0;
    for (int i = 0; i < N; ++i) {
      const S* p = &ref.pos[lvl][3 * i];
      const S* cv = &ref.colVar[lvl][2 * i];
      S Wx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
      S Wy = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
      S Wz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
      S u = (Wx / Wz) * fx + cx;
      S v = (Wy / Wz) * fy + cy;
      if (!(u > 1 && v > 1 && u < (S)(W - 2) && v < (S)(H - 2))) {
        if (isGoodOut) isGoodOut[ref.idx[lvl][i]] = 0;
        continue;
      }
      S ri[3];
      interp4(grad, u, v, W, 3, ri);
      S c1 = aff_a * cv[0] + aff_b;
      S c2 = ri[2];
      S r = c1 - c2;
      S wgt = std::fabs(r) < (S)5 ? (S)1 : (S)5 / std::fabs(r);
      sxx += (Acc)(c1 * c1 * wgt);
      syy += (Acc)(c2 * c2 * wgt);
      sx += (Acc)(c1 * wgt);
      sy += (Acc)(c2 * wgt);
      sw += (Acc)wgt;
      bool isGood = r * r / ((S)P.max_diff_constant + (S)P.max_diff_grad_mult * (ri[0] * ri[0] + ri[1] * ri[1])) < (S)1;
      if (isGoodOut) isGoodOut[ref.idx[lvl][i]] = isGood ? 1 : 0;
      bx[idx] = Wx; by[idx] = Wy; bz[idx] = Wz;
      bdx[idx] = fx * ri[0];
      bdy[idx] = fy * ri[1];
      bres[idx] = r;
      bd[idx] = (S)1 / p[2];
      bvar[idx] = cv[1];
      idx++;
      if (isGood) {
        sumResUnw += (Acc)(r * r);
        sumSigned += (Acc)r;
        goodCount++;
      } else {
        badCount++;
      }
      S depthChange = p[2] / Wz;
      usage += (Acc)(depthChange < 1 ? depthChange : (S)1);
    }
    warped = idx;
    pointUsage = usage / (Acc)N;
    lastGoodCount = (Acc)goodCount;
    lastBadCount = (Acc)badCount;
    lastMeanRes = sumSigned / (Acc)goodCount;
    aff_a_it = (S)std::sqrt((syy - sy * sy / sw) / (sxx - sx * sx / sw));
    aff_b_it = (S)((sy - (Acc)aff_a_it * sx) / sw);
    last.warped = idx; last.good = goodCount; last.bad = badCount;
    last.sumResUnw = sumResUnw; last.sumSigned = sumSigned; last.usage = usage;
    last.sxx = sxx; last.syy = syy; last.sx = sx; last.sy = sy; last.sw = sw;
    return sumResUnw / (Acc)goodCount;
  }

  // calcWeightsAndResidual: w_p = 1/(cameraPixelNoise2 + s drpdd^2), Huber at
  // huber_d/2, E = sum wh w_p r^2 / M.
  Acc calcWeightsAndResidual(const Sim3& refToFrame) {
    S tx = (S)refToFrame.t[0], ty = (S)refToFrame.t[1], tz = (S)refToFrame.t[2];
    Acc sumRes = 0;
    S hh = (S)P.huber_d / (S)2;
    for (int i = 0; i < warped; ++i) {
      S px = bx[i], py = by[i], pz = bz[i], d = bd[i], rp = bres[i], gx = bdx[i], gy = bdy[i];
      S s = (S)P.var_weight * bvar[i];
      S g0 = (tx * pz - tz * px) / (pz * pz * d);
      S g1 = (ty * pz - tz * py) / (pz * pz * d);
      S drpdd = gx * g0 + gy * g1;
      S w_p = (S)1 / ((S)P.camera_pixel_noise2 + s * drpdd * drpdd);
      S weighted_rp = std::fabs(rp * std::sqrt(w_p));
      S wh = std::fabs(weighted_rp < hh ? (S)1 : hh / weighted_rp);
      sumRes += (Acc)(wh * w_p * rp * rp);
      bw[i] = wh * w_p;
    }
    last.err = sumRes;
    return sumRes / (Acc)warped;
  }

  // v = -dr/dxi (6 terms), as the reference's calculateWarpUpdate.
  static void jac6(S px, S py, S pz, S gx, S gy, S* v) {
    S z = (S)1 / pz;
    S z_sqr = (S)1 / (pz * pz);
    v[0] = z * gx;
    v[1] = z * gy;
    v[2] = (-px * z_sqr) * gx + (-py * z_sqr) * gy;
    v[3] = (-px * py * z_sqr) * gx + (-((S)1 + py * py * z_sqr)) * gy;
    v[4] = ((S)1 + px * px * z_sqr) * gx + (px * py * z_sqr) * gy;
    v[5] = (-py * z) * gx + (px * z) * gy;
  }

  // calculateWarpUpdate: A = sum W v v^T, b = -sum W r v, err = sum W r^2, n = M
  // (LGS6 update; not divided here).
  void calculateWarpUpdate(EvalSums<Acc>& ls) {
    std::fill(ls.A, ls.A + 49, (Acc)0);
    std::fill(ls.b, ls.b + 7, (Acc)0);
    Acc e = 0;
    int n = 0;
    for (int i = 0; i < warped; ++i) {
      S v[6];
      jac6(bx[i], by[i], bz[i], bdx[i], bdy[i], v);
      S r = bres[i], w = bw[i];
      for (int a = 0; a < 6; ++a)
        for (int c = a; c < 6; ++c) ls.A[7 * a + c] += (Acc)(v[a] * v[c] * w);
      for (int a = 0; a < 6; ++a) ls.b[a] -= (Acc)(v[a] * (r * w));
      e += (Acc)(r * r * w);
      n++;
    }
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < a; ++c) ls.A[7 * a + c] = ls.A[7 * c + a];
    ls.err = e;
    ls.n = n;
  }

  // SE3Tracker::trackFrame (S7 of the behavioural spec): levels
  // se3_max_level-1 .. se3_min_level, LM with lambda schedule, divergence on
  // M < absmin * w_l * h_l, good/bad counts from the last residual pass.
  Sim3 trackFrame(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& frameToRefInit) {
    diverged = false;
    trackingWasGood = true;
    aff_a = 1; aff_b = 0;
    Sim3 refToFrame = frameToRefInit.inverse();
    refToFrame.s = 1.0;
    Acc last_residual = 0;
    for (int l = 0; l < PYR; ++l) numResidualCalls[l] = numWarpCalls[l] = 0;
    for (int lvl = P.se3_max_level - 1; lvl >= P.se3_min_level; --lvl) {
      ref.makePointCloud(lvl);
      uint8_t* goodOut = (lvl == P.se3_min_level && lvl == 1) ? frame.refPixelWasGood() : nullptr;
      calcResidualAndBuffers(ref, frame, refToFrame, lvl, goodOut);
      if ((S)warped < (S)P.min_goodperall_pixel_absmin * (S)(frame.w[lvl] * frame.h[lvl])) {
        diverged = true;
        trackingWasGood = false;
        return Sim3();
      }
      if (P.use_affine_lightning_estimation) { aff_a = aff_a_it; aff_b = aff_b_it; }
      Acc lastErr = calcWeightsAndResidual(refToFrame);
      numResidualCalls[lvl]++;
      float lambda = P.lambda_initial[lvl];
      EvalSums<Acc> ls;
      for (int it = 0; it < P.max_its_per_lvl[lvl]; ++it) {
        calculateWarpUpdate(ls);
        numWarpCalls[lvl]++;
        int incTry = 0;
        while (true) {
          double A[36], bb[6], inc[6];
          for (int a = 0; a < 6; ++a) {
            for (int c = 0; c < 6; ++c) A[6 * a + c] = (double)ls.A[7 * a + c] / (double)ls.n;
            bb[a] = -(double)ls.b[a] / (double)ls.n;
          }
          for (int a = 0; a < 6; ++a) A[7 * a] *= 1.0 + (double)lambda;
          ldlt_solve<6>(A, bb, inc);
          incTry++;
          Sim3 newT = se3_exp(inc) * refToFrame;
          calcResidualAndBuffers(ref, frame, newT, lvl, goodOut);
          if ((S)warped < (S)P.min_goodperall_pixel_absmin * (S)(frame.w[lvl] * frame.h[lvl])) {
            diverged = true;
            trackingWasGood = false;
            return Sim3();
          }
          Acc error = calcWeightsAndResidual(newT);
          numResidualCalls[lvl]++;
          if (error < lastErr) {
            refToFrame = newT;
            if (P.use_affine_lightning_estimation) { aff_a = aff_a_it; aff_b = aff_b_it; }
            if (error / lastErr > (Acc)P.convergence_eps[lvl]) it = P.max_its_per_lvl[lvl];
            last_residual = lastErr = error;
            if (lambda <= 0.2f) lambda = 0;
            else lambda *= P.lambda_success_fac[lvl];
            break;
          } else {
            double dd = 0;
            for (int a = 0; a < 6; ++a) dd += inc[a] * inc[a];
            if (!(dd > (double)P.step_size_min[lvl])) {
              it = P.max_its_per_lvl[lvl];
              break;
            }
            if (lambda == 0) lambda = 0.2f;
            else lambda *= std::pow(P.lambda_fail_fac[lvl], (float)incTry);
          }
        }
      }
    }
    lastResidual = last_residual;
    int ml = P.se3_min_level;
    trackingWasGood = !diverged &&
                      lastGoodCount / (Acc)(frame.w[ml] * frame.h[ml]) > (Acc)P.min_goodperall_pixel &&
                      lastGoodCount / (lastGoodCount + lastBadCount) > (Acc)P.min_goodpergoodbad_pixel;
    if (trackingWasGood) ref.kf->numFramesTrackedOnThis++;
    frame.initialTrackedResidual = (S)(lastResidual / pointUsage);
    frame.thisToParent = refToFrame.inverse();
    frame.thisToParent.s = 1.0;
    frame.parentId = ref.kf->id;
    frame.parent = ref.kf;
    return refToFrame.inverse();
  }

  // One fused evaluation (for parity of the per-evaluation sums).
  EvalSums<Acc> evaluate(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& refToFrame, int lvl,
                         uint8_t* isGoodOut, S a, S b) {
    aff_a = a; aff_b = b;
    ref.makePointCloud(lvl);
    calcResidualAndBuffers(ref, frame, refToFrame, lvl, lvl == 1 ? isGoodOut : nullptr);
    calcWeightsAndResidual(refToFrame);
    EvalSums<Acc> ls = last;
    calculateWarpUpdate(ls);
    ls.warped = last.warped; ls.good = last.good; ls.bad = last.bad;
    ls.sumResUnw = last.sumResUnw; ls.sumSigned = last.sumSigned; ls.usage = last.usage;
    ls.sxx = last.sxx; ls.syy = last.syy; ls.sx = last.sx; ls.sy = last.sy; ls.sw = last.sw;
    ls.sumResP = ls.err;
    ls.numTermsP = warped;
    return ls;
  }
};

// ---------------------------------------------------------- Sim3Tracker
// Upstream Tracking/Sim3Tracker: calcSim3Buffers, calcSim3WeightsAndResidual,
// calcSim3LGS (LGS6 + LGS4 -> LGS7 at {2,3,4,6}), trackFrameSim3.
template <class S, class Acc = S>
struct Sim3Tracker {
  lsdg_params P;
  std::vector<S> bx, by, bz, bdx, bdy, bres, bresd, bd, bvar, bvard, bwp, bwd;
  int warped = 0;
  Acc pointUsage = 0;
  S aff_a = 1, aff_b = 0, aff_a_it = 1, aff_b_it = 0;
  bool diverged = false;
  Acc lastResidual = 0, lastDepthResidual = 0, lastPhotometricResidual = 0;
  double lastHessian[49] = {0};
  int numResidualCalls[PYR] = {0}, numWarpCalls[PYR] = {0};
  EvalSums<Acc> last;

  struct Res {
    Acc sumResD = 0, sumResP = 0;
    int numTermsD = 0, numTermsP = 0;
    Acc meanD = 0, meanP = 0, mean = 0;
  };

  // In-plane roll of R (ESM gradient rotation): rollMat = Q^T R with Q the
  // shortest rotation taking (0,0,-1) to R (0,0,-1).
  static void rollOf(const Sim3& T, S* r4) {
    double R[9];
    T.R(R);
    double f[3] = {0, 0, -1};
    double g[3] = {-R[2], -R[5], -R[8]};
    double c = f[0] * g[0] + f[1] * g[1] + f[2] * g[2];
    double ax[3] = {f[1] * g[2] - f[2] * g[1], f[2] * g[0] - f[0] * g[2], f[0] * g[1] - f[1] * g[0]};
    Quat q;
    if (c < -1.0 + 1e-12) {
      q = {1, 0, 0, 0};  // half turn about x (any axis orthogonal to f)
    } else {
      double s = std::sqrt((1.0 + c) * 2.0);
      q = {ax[0] / s, ax[1] / s, ax[2] / s, 0.5 * s};
    }
    double Q[9];
    qmat(qnormalize(q), Q);
    double M[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M[3 * i + j] = Q[i] * R[j] + Q[3 + i] * R[3 + j] + Q[6 + i] * R[6 + j];
    r4[0] = (S)M[0]; r4[1] = (S)M[1]; r4[2] = (S)M[3]; r4[3] = (S)M[4];
  }

  void calcSim3Buffers(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& refToFrame, int lvl) {
    int W = frame.w[lvl], H = frame.h[lvl];
    S fx = frame.fx[lvl], fy = frame.fy[lvl], cx = frame.cx[lvl], cy = frame.cy[lvl];
    S R[9], t[3];
    pose_RT(refToFrame, R, t, true);
    S roll[4];
    rollOf(refToFrame, roll);
    const S* grad = frame.grad[lvl].data();
    const S* fid = frame.idepth[lvl].data();
    const S* fiv = frame.ivar[lvl].data();
    int N = ref.num[lvl];
    size_t cap = (size_t)N;
    bx.resize(cap); by.resize(cap); bz.resize(cap); bdx.resize(cap); bdy.resize(cap); bres.resize(cap);
    bresd.resize(cap); bd.resize(cap); bvar.resize(cap); bvard.resize(cap); bwp.resize(cap); bwd.resize(cap);
    Acc sxx = 0, syy = 0, sx = 0, sy = 0, sw = 0, usage = 0;
    int idx = 0;
    for (int i = 0; i < N; ++i) {
      const S* p = &ref.pos[lvl][3 * i];
      const S* cv = &ref.colVar[lvl][2 * i];
      const S* rg = &ref.gradd[lvl][2 * i];
      S Wx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
      S Wy = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
      S Wz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
      S u = (Wx / Wz) * fx + cx;
      S v = (Wy / Wz) * fy + cy;
      if (!(u > 1 && v > 1 && u < (S)(W - 2) && v < (S)(H - 2))) continue;
      bx[idx] = Wx; by[idx] = Wy; bz[idx] = Wz;
      S ri[3];
      interp4(grad, u, v, W, 3, ri);
      if (P.use_esm_tracking) {
        S rgx = roll[0] * rg[0] + roll[1] * rg[1];
        S rgy = roll[2] * rg[0] + roll[3] * rg[1];
        bdx[idx] = fx * (S)0.5 * (ri[0] + rgx);
        bdy[idx] = fy * (S)0.5 * (ri[1] + rgy);
      } else {
        bdx[idx] = fx * ri[0];
        bdy[idx] = fy * ri[1];
      }
      S c1 = aff_a * cv[0] + aff_b;
      S c2 = ri[2];
      S r = c1 - c2;
      S wgt = std::fabs(r) < (S)2 ? (S)1 : (S)2 / std::fabs(r);
      sxx += (Acc)(c1 * c1 * wgt);
      syy += (Acc)(c2 * c2 * wgt);
      sx += (Acc)(c1 * wgt);
      sy += (Acc)(c2 * wgt);
      sw += (Acc)wgt;
      bres[idx] = r;
      bvar[idx] = cv[1];
      int ir = (int)(u + (S)0.5) + W * (int)(v + (S)0.5);
      S varF = fiv[ir];
      S refId = (S)1 / Wz;
      bd[idx] = (S)1 / p[2];
      if (varF > 0) {
        bresd[idx] = refId - fid[ir];
        bvard[idx] = varF;
      } else {
        bresd[idx] = -1;
        bvard[idx] = -1;
      }
      S depthChange = p[2] / Wz;
      usage += (Acc)(depthChange < 1 ? depthChange : (S)1);
      idx++;
    }
    warped = idx;
    pointUsage = usage / (Acc)N;
    aff_a_it = (S)std::sqrt((syy - sy * sy / sw) / (sxx - sx * sx / sw));
    aff_b_it = (S)((sy - (Acc)aff_a_it * sx) / sw);
    last.warped = idx; last.usage = usage;
    last.sxx = sxx; last.syy = syy; last.sx = sx; last.sy = sy; last.sw = sw;
  }

  Res calcSim3WeightsAndResidual(const Sim3& refToFrame) {
    S tx = (S)refToFrame.t[0], ty = (S)refToFrame.t[1], tz = (S)refToFrame.t[2];
    Res rs;
    S hd = (S)P.huber_d;
    for (int i = 0; i < warped; ++i) {
      S px = bx[i], py = by[i], pz = bz[i], d = bd[i], rp = bres[i], rd = bresd[i];
      S gx = bdx[i], gy = bdy[i];
      S s = (S)P.var_weight * bvar[i];
      S sv = (S)P.var_weight * bvard[i];
      S g0 = (tx * pz - tz * px) / (pz * pz * d);
      S g1 = (ty * pz - tz * py) / (pz * pz * d);
      S g2 = (pz - tz) / (pz * pz * d);
      S drpdd = gx * g0 + gy * g1;
      S w_p = (S)1 / ((S)P.camera_pixel_noise2 + s * drpdd * drpdd);
      S w_d = (S)1 / (sv + g2 * g2 * s);
      S wrd = std::fabs(rd * std::sqrt(w_d));
      S wrp = std::fabs(rp * std::sqrt(w_p));
      S wabs = sv > 0 ? wrd + wrp : wrp;
      S wh = std::fabs(wabs < hd ? (S)1 : hd / wabs);
      if (sv > 0) {
        rs.sumResD += (Acc)(wh * w_d * rd * rd);
        rs.numTermsD++;
      }
      rs.sumResP += (Acc)(wh * w_p * rp * rp);
      rs.numTermsP++;
      bwp[i] = wh * w_p;
      bwd[i] = sv > 0 ? wh * w_d : (S)0;
    }
    rs.mean = (rs.sumResD + rs.sumResP) / (Acc)(rs.numTermsD + rs.numTermsP);
    rs.meanD = rs.sumResD / (Acc)rs.numTermsD;
    rs.meanP = rs.sumResP / (Acc)rs.numTermsP;
    last.sumResD = rs.sumResD; last.sumResP = rs.sumResP;
    last.numTermsD = rs.numTermsD; last.numTermsP = rs.numTermsP;
    return rs;
  }

  // calcSim3LGS: photometric LGS6 and depth LGS4 (v4 = (z^2, z^2 py, -z^2 px, z)
  // over (v3, w1, w2, sigma)), not divided; merged into 7x7 at {2,3,4,6}.
  void calcSim3LGS(EvalSums<Acc>& ls) {
    Acc A6[36] = {0}, b6[6] = {0}, A4[16] = {0}, b4[4] = {0}, e6 = 0, e4 = 0;
    int n6 = 0, n4 = 0;
    for (int i = 0; i < warped; ++i) {
      S px = bx[i], py = by[i], pz = bz[i];
      S wp = bwp[i], wd = bwd[i], rp = bres[i], rd = bresd[i];
      S v[6];
      SE3Tracker<S>::jac6(px, py, pz, bdx[i], bdy[i], v);
      S z = (S)1 / pz;
      S z_sqr = (S)1 / (pz * pz);
      S v4[4] = {z_sqr, z_sqr * py, -z_sqr * px, z};
      for (int a = 0; a < 6; ++a)
        for (int c = a; c < 6; ++c) A6[6 * a + c] += (Acc)(v[a] * v[c] * wp);
      for (int a = 0; a < 6; ++a) b6[a] -= (Acc)(v[a] * (rp * wp));
      e6 += (Acc)(rp * rp * wp);
      n6++;
      for (int a = 0; a < 4; ++a)
        for (int c = a; c < 4; ++c) A4[4 * a + c] += (Acc)(v4[a] * v4[c] * wd);
      for (int a = 0; a < 4; ++a) b4[a] -= (Acc)(v4[a] * (rd * wd));
      e4 += (Acc)(rd * rd * wd);
      n4++;
    }
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < a; ++c) A6[6 * a + c] = A6[6 * c + a];
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < a; ++c) A4[4 * a + c] = A4[4 * c + a];
    std::fill(ls.A, ls.A + 49, (Acc)0);
    std::fill(ls.b, ls.b + 7, (Acc)0);
    for (int a = 0; a < 6; ++a) {
      for (int c = 0; c < 6; ++c) ls.A[7 * a + c] = A6[6 * a + c];
      ls.b[a] = b6[a];
    }
    const int remap[4] = {2, 3, 4, 6};
    for (int a = 0; a < 4; ++a) {
      ls.b[remap[a]] += b4[a];
      for (int c = 0; c < 4; ++c) ls.A[7 * remap[a] + remap[c]] += A4[4 * a + c];
    }
    ls.err = e6 + e4;
    ls.n = n6 + n4;
  }

  Sim3 trackFrameSim3(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& frameToRefInit, int startLevel,
                      int finalLevel) {
    diverged = false;
    aff_a = 1; aff_b = 0;
    Sim3 refToFrame = frameToRefInit.inverse();
    EvalSums<Acc> ls7;
    Res finalResidual;
    bool upToDate = false;
    for (int l = 0; l < PYR; ++l) numResidualCalls[l] = numWarpCalls[l] = 0;
    std::fill(lastHessian, lastHessian + 49, 0.0);
    for (int lvl = startLevel; lvl >= finalLevel; --lvl) {
      if (P.max_its_per_lvl[lvl] == 0) continue;
      ref.makePointCloud(lvl);
      calcSim3Buffers(ref, frame, refToFrame, lvl);
      S lim = (S)0.5 * (S)P.min_goodperall_pixel_absmin * (S)(frame.w[lvl] * frame.h[lvl]);
      if ((S)warped < lim || warped < 10) {
        diverged = true;
        return Sim3();
      }
      Res lastErr = calcSim3WeightsAndResidual(refToFrame);
      numResidualCalls[lvl]++;
      if (P.use_affine_lightning_estimation) { aff_a = aff_a_it; aff_b = aff_b_it; }
      float lambda = P.lambda_initial[lvl];
      upToDate = false;
      for (int it = 0; it < P.max_its_per_lvl[lvl]; ++it) {
        calcSim3LGS(ls7);
        upToDate = true;
        numWarpCalls[lvl]++;
        int incTry = 0;
        while (true) {
          double A[49], bb[7], inc[7];
          for (int a = 0; a < 7; ++a) {
            for (int c = 0; c < 7; ++c) A[7 * a + c] = (double)ls7.A[7 * a + c] / (double)ls7.n;
            bb[a] = -(double)ls7.b[a] / (double)ls7.n;
          }
          for (int a = 0; a < 7; ++a) A[8 * a] *= 1.0 + (double)lambda;
          ldlt_solve<7>(A, bb, inc);
          incTry++;
          double absInc = 0;
          for (int a = 0; a < 7; ++a) absInc += inc[a] * inc[a];
          if (!(absInc >= 0 && absInc < 1)) {
            diverged = true;
            std::fill(lastHessian, lastHessian + 49, 0.0);
            return Sim3();
          }
          Sim3 newT = sim3_exp(inc) * refToFrame;
          calcSim3Buffers(ref, frame, newT, lvl);
          if ((S)warped < lim || warped < 10) {
            diverged = true;
            return Sim3();
          }
          Res error = calcSim3WeightsAndResidual(newT);
          numResidualCalls[lvl]++;
          if (error.mean < lastErr.mean) {
            refToFrame = newT;
            upToDate = false;
            if (P.use_affine_lightning_estimation) { aff_a = aff_a_it; aff_b = aff_b_it; }
            if (error.mean / lastErr.mean > (Acc)P.convergence_eps[lvl]) it = P.max_its_per_lvl[lvl];
            finalResidual = lastErr = error;
            if (lambda <= 0.2f) lambda = 0;
            else lambda *= P.lambda_success_fac[lvl];
            break;
          } else {
            if (!(absInc > (double)P.step_size_min[lvl])) {
              it = P.max_its_per_lvl[lvl];
              break;
            }
            if (lambda == 0) lambda = 0.2f;
            else lambda *= std::pow(P.lambda_fail_fac[lvl], (float)incTry);
          }
        }
      }
    }
    if (!upToDate) {
      ref.makePointCloud(finalLevel);
      calcSim3Buffers(ref, frame, refToFrame, finalLevel);
      finalResidual = calcSim3WeightsAndResidual(refToFrame);
      calcSim3LGS(ls7);
    }
    for (int i = 0; i < 49; ++i) lastHessian[i] = (double)ls7.A[i];
    if (refToFrame.s <= 0) {
      diverged = true;
      return Sim3();
    }
    lastResidual = finalResidual.mean;
    lastDepthResidual = finalResidual.meanD;
    lastPhotometricResidual = finalResidual.meanP;
    return refToFrame.inverse();
  }

  EvalSums<Acc> evaluate(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& refToFrame, int lvl) {
    aff_a = 1; aff_b = 0;
    ref.makePointCloud(lvl);
    calcSim3Buffers(ref, frame, refToFrame, lvl);
    calcSim3WeightsAndResidual(refToFrame);
    EvalSums<Acc> ls;
    calcSim3LGS(ls);
    ls.warped = last.warped; ls.usage = last.usage;
    ls.sxx = last.sxx; ls.syy = last.syy; ls.sx = last.sx; ls.sy = last.sy; ls.sw = last.sw;
    ls.sumResP = last.sumResP; ls.sumResD = last.sumResD;
    ls.numTermsP = last.numTermsP; ls.numTermsD = last.numTermsD;
    ls.err = last.sumResP + last.sumResD;
    return ls;
  }
};

// ------------------------------------------------------------- Adjoint
// Gradient of E = sum_i rho(sqrt(w_p,i) r_i) (Huber at huber_d/2, w_p and the
// in-image set frozen at refToFrame): dE/dxi = -sum W r v (= LGS6 b),
// dE/dd_i = -W_i r_i drpdd_i.  energy = sum_i rho(.) with
// rho(e) = e^2/2 (|e| <= k), k|e| - k^2/2 otherwise.
template <class S, class Acc = S>
struct Adjoint {
  lsdg_params P;
  // grad_d has one entry per reference point (0 outside the image).
  void eval(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& refToFrame, int lvl, double* energy,
            double* gxi, S* grad_d) {
    ref.makePointCloud(lvl);
    int W = frame.w[lvl], H = frame.h[lvl];
    S fx = frame.fx[lvl], fy = frame.fy[lvl], cx = frame.cx[lvl], cy = frame.cy[lvl];
    S R[9], t[3];
    pose_RT(refToFrame, R, t, false);
    const S* grad = frame.grad[lvl].data();
    int N = ref.num[lvl];
    S k = (S)P.huber_d / (S)2;
    Acc E = 0, G[6] = {0};
    for (int i = 0; i < N; ++i) {
      const S* p = &ref.pos[lvl][3 * i];
      const S* cv = &ref.colVar[lvl][2 * i];
      grad_d[i] = 0;
      S Wx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
      S Wy = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
      S Wz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
      S u = (Wx / Wz) * fx + cx;
      S v = (Wy / Wz) * fy + cy;
      if (!(u > 1 && v > 1 && u < (S)(W - 2) && v < (S)(H - 2))) continue;
      S ri[3];
      interp4(grad, u, v, W, 3, ri);
      S r = cv[0] - ri[2];
      S gx = fx * ri[0], gy = fy * ri[1];
      S d = (S)1 / p[2];
      S s = (S)P.var_weight * cv[1];
      S g0 = (t[0] * Wz - t[2] * Wx) / (Wz * Wz * d);
      S g1 = (t[1] * Wz - t[2] * Wy) / (Wz * Wz * d);
      S drpdd = gx * g0 + gy * g1;
      S w_p = (S)1 / ((S)P.camera_pixel_noise2 + s * drpdd * drpdd);
      S e = std::fabs(r * std::sqrt(w_p));
      S wh = e < k ? (S)1 : k / e;
      E += (Acc)(e < k ? (S)0.5 * e * e : k * e - (S)0.5 * k * k);
      S Wr = wh * w_p * r;
      S vv[6];
      SE3Tracker<S>::jac6(Wx, Wy, Wz, gx, gy, vv);
      for (int a = 0; a < 6; ++a) G[a] -= (Acc)(Wr * vv[a]);
      grad_d[i] = -Wr * drpdd;
    }
    *energy = (double)E;
    for (int a = 0; a < 6; ++a) gxi[a] = (double)G[a];
  }
};

// Sim3 adjoint: gradient of the Sim3 tracker's reweighted least-squares
// energy with its weights frozen at refToFrame,
//   E = 1/2 sum_i wh_i (w_p,i rp_i^2 + w_d,i rd_i^2),
// with the per-point terms of calcSim3Buffers / calcSim3WeightsAndResidual
// above (no ESM: the frame's own gradient) and the LGS6 + LGS4 -> LGS7 right
// hand side of calcSim3LGS: dE/dxi = -sum (wp rp v + wd rd v4 at {2,3,4,6}),
// dE/dd_i = -wp rp drpdd + wd rd g2.  out: gxi[7], E.
template <class S, class Acc = S>
struct AdjointSim3 {
  lsdg_params P;
  void eval(TrackingReference<S>& ref, Frame<S>& frame, const Sim3& refToFrame, int lvl, double* energy,
            double* gxi, S* grad_d) {
    ref.makePointCloud(lvl);
    int W = frame.w[lvl], H = frame.h[lvl];
    S fx = frame.fx[lvl], fy = frame.fy[lvl], cx = frame.cx[lvl], cy = frame.cy[lvl];
    S R[9], t[3];
    pose_RT(refToFrame, R, t, true);
    const S* grad = frame.grad[lvl].data();
    const S* fid = frame.idepth[lvl].data();
    const S* fiv = frame.ivar[lvl].data();
    S hd = (S)P.huber_d;
    int N = ref.num[lvl];
    Acc Ep = 0, Ed = 0, G[7] = {0};
    const int remap[4] = {2, 3, 4, 6};
    for (int i = 0; i < N; ++i) {
      const S* p = &ref.pos[lvl][3 * i];
      const S* cv = &ref.colVar[lvl][2 * i];
      grad_d[i] = 0;
      S Wx = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + t[0];
      S Wy = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + t[1];
      S Wz = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + t[2];
      S u = (Wx / Wz) * fx + cx;
      S v = (Wy / Wz) * fy + cy;
      if (!(u > 1 && v > 1 && u < (S)(W - 2) && v < (S)(H - 2))) continue;
      S ri[3];
      interp4(grad, u, v, W, 3, ri);
      S gx = fx * ri[0], gy = fy * ri[1];
      S rp = cv[0] - ri[2];
      int ir = (int)(u + (S)0.5) + W * (int)(v + (S)0.5);
      S varF = fiv[ir];
      S rd, vard;
      if (varF > 0) {
        rd = (S)1 / Wz - fid[ir];
        vard = varF;
      } else {
        rd = -1;
        vard = -1;
      }
      S d = (S)1 / p[2];
      S s = (S)P.var_weight * cv[1];
      S sv = (S)P.var_weight * vard;
      S g0 = (t[0] * Wz - t[2] * Wx) / (Wz * Wz * d);
      S g1 = (t[1] * Wz - t[2] * Wy) / (Wz * Wz * d);
      S g2 = (Wz - t[2]) / (Wz * Wz * d);
      S drpdd = gx * g0 + gy * g1;
      S w_p = (S)1 / ((S)P.camera_pixel_noise2 + s * drpdd * drpdd);
      S w_d = (S)1 / (sv + g2 * g2 * s);
      S wrd = std::fabs(rd * std::sqrt(w_d));
      S wrp = std::fabs(rp * std::sqrt(w_p));
      S wabs = sv > 0 ? wrd + wrp : wrp;
      S wh = std::fabs(wabs < hd ? (S)1 : hd / wabs);
      S wp = wh * w_p;
      S wd = sv > 0 ? wh * w_d : (S)0;
      Ep += (Acc)(wp * rp * rp);
      if (sv > 0) Ed += (Acc)(wh * w_d * rd * rd);
      S vv[6];
      SE3Tracker<S>::jac6(Wx, Wy, Wz, gx, gy, vv);
      S rpW = rp * wp;
      for (int a = 0; a < 6; ++a) G[a] -= (Acc)(vv[a] * rpW);
      S gd = -rpW * drpdd;
      if (wd != 0) {
        S z = (S)1 / Wz;
        S z_sqr = (S)1 / (Wz * Wz);
        S v4[4] = {z_sqr, z_sqr * Wy, -z_sqr * Wx, z};
        S rdW = rd * wd;
        for (int a = 0; a < 4; ++a) G[remap[a]] -= (Acc)(v4[a] * rdW);
        gd = gd + rdW * g2;
      }
      grad_d[i] = gd;
    }
    *energy = 0.5 * ((double)Ep + (double)Ed);
    for (int a = 0; a < 7; ++a) gxi[a] = (double)G[a];
  }
};

// ------------------------------------------------------------ DepthMap
// Upstream DepthEstimation/DepthMap + DepthMapPixelHypothesis.  Output
// evidence: the keyframe idepth / idepthVar arrays the wrapper reads
// (PangolinOutputIOWrapper.cpp:75-89).
template <class S>
struct Hyp {
  bool isValid = false;
  int blacklisted = 0;
  float nextStereoFrameMinID = 0;
  int validity_counter = 0;
  S idepth = 0, idepth_var = 0, idepth_smoothed = 0, idepth_var_smoothed = 0;
  static Hyp make(S id, S var, int val) {
    Hyp h;
    h.isValid = true; h.blacklisted = 0; h.nextStereoFrameMinID = 0; h.validity_counter = val;
    h.idepth = id; h.idepth_var = var; h.idepth_smoothed = -1; h.idepth_var_smoothed = -1;
    return h;
  }
  static Hyp make(S id, S ids, S var, S vars, int val) {
    Hyp h = make(id, var, val);
    h.idepth_smoothed = ids; h.idepth_var_smoothed = vars;
    return h;
  }
};

template <class S>
struct DepthMap {
  lsdg_params P;
  int width = 0, height = 0;
  S fx, fy, cx, cy, fxi, fyi, cxi, cyi, K[9];
  std::vector<Hyp<S>> cur, oth;
  std::vector<uint8_t> obsCode;
  Frame<S>* kf = nullptr;
  bool kfReactivated = false;
  Frame<S>* oldestRef = nullptr;
  Frame<S>* newestRef = nullptr;
  std::vector<Frame<S>*> refById;
  int refByIdOffset = 0;
  int threads = 1;
  bool accDouble = false;  // FP64 accumulation of the rescale mean

  void init(int w, int h, float fx0, float fy0, float cx0, float cy0) {
    width = w; height = h;
    fx = fx0; fy = fy0; cx = cx0; cy = cy0;
    fxi = (S)1 / fx; fyi = (S)1 / fy; cxi = -cx / fx; cyi = -cy / fy;
    S k[9] = {fx, 0, cx, 0, fy, cy, 0, 0, 1};
    std::copy(k, k + 9, K);
    cur.assign((size_t)w * h, Hyp<S>());
    oth.assign((size_t)w * h, Hyp<S>());
    obsCode.assign((size_t)w * h, 0);
  }
  void reset() {
    for (auto& h : cur) h.isValid = false;
    for (auto& h : oth) h.isValid = false;
    kf = nullptr;
  }

  // IndexThreadReduce: rows [y0, y1) in chunks of `step`, spread over threads.
  void reduce(int y0, int y1, int step, const std::function<void(int, int)>& fn) {
    if (threads <= 1) {
      for (int y = y0; y < y1; y += step) fn(y, std::min(y + step, y1));
      return;
    }
    std::vector<std::thread> pool;
    int nchunks = (y1 - y0 + step - 1) / step;
    for (int t = 0; t < threads; ++t)
      pool.emplace_back([&, t] {
        for (int c = t; c < nchunks; c += threads) {
          int a = y0 + c * step;
          fn(a, std::min(a + step, y1));
        }
      });
    for (auto& th : pool) th.join();
  }

  // makeAndCheckEPL
  bool makeAndCheckEPL(int x, int y, const Frame<S>* ref, S* pepx, S* pepy) {
    int idx = x + y * width;
    S epx = -fx * ref->thisToOther_t[0] + ref->thisToOther_t[2] * ((S)x - cx);
    S epy = -fy * ref->thisToOther_t[1] + ref->thisToOther_t[2] * ((S)y - cy);
    if (std::isnan(epx + epy)) return false;
    S eplLengthSquared = epx * epx + epy * epy;
    if (eplLengthSquared < (S)P.min_epl_length_squared) return false;
    const S* I = kf->image[0].data();
    S gx = I[idx + 1] - I[idx - 1];
    S gy = I[idx + width] - I[idx - width];
    S eplGradSquared = gx * epx + gy * epy;
    eplGradSquared = eplGradSquared * eplGradSquared / eplLengthSquared;
    if (eplGradSquared < (S)P.min_epl_grad_squared) return false;
    if (eplGradSquared / (gx * gx + gy * gy) < (S)P.min_epl_angle_squared) return false;
    S fac = (S)P.gradient_sample_dist / std::sqrt(eplLengthSquared);
    *pepx = epx * fac;
    *pepy = epy * fac;
    return true;
  }

  // doLineStereo: returns the matching error or a code -1..-4.
  S doLineStereo(S u, S v, S epxn, S epyn, S min_idepth, S prior_idepth, S max_idepth, const Frame<S>* ref,
                 const S* refImg, S& r_idepth, S& r_var, S& r_epl) {
    const S* kfImg = kf->image[0].data();
    S KinvP[3] = {fxi * u + cxi, fyi * v + cyi, (S)1};
    const S* M = ref->K_otherToThis_R;
    S pInf[3];
    for (int i = 0; i < 3; ++i) pInf[i] = M[3 * i] * KinvP[0] + M[3 * i + 1] * KinvP[1] + M[3 * i + 2] * KinvP[2];
    const S* Kt = ref->K_otherToThis_t;
    S pReal2 = pInf[2] / prior_idepth + Kt[2];
    S rescale = pReal2 * prior_idepth;
    S firstX = u - (S)2 * epxn * rescale, firstY = v - (S)2 * epyn * rescale;
    S lastX = u + (S)2 * epxn * rescale, lastY = v + (S)2 * epyn * rescale;
    if (firstX <= 0 || firstX >= (S)(width - 2) || firstY <= 0 || firstY >= (S)(height - 2) || lastX <= 0 ||
        lastX >= (S)(width - 2) || lastY <= 0 || lastY >= (S)(height - 2))
      return -1;
    if (!(rescale > (S)0.7f && rescale < (S)1.4f)) return -1;
    S rv_p1 = interp(kfImg, u + epxn * rescale, v + epyn * rescale, width);
    S rv_m1 = interp(kfImg, u - epxn * rescale, v - epyn * rescale, width);
    S rv = interp(kfImg, u, v, width);
    S rv_m2 = interp(kfImg, u - (S)2 * epxn * rescale, v - (S)2 * epyn * rescale, width);
    S rv_p2 = interp(kfImg, u + (S)2 * epxn * rescale, v + (S)2 * epyn * rescale, width);

    S pClose[3], pFar[3];
    for (int i = 0; i < 3; ++i) pClose[i] = pInf[i] + Kt[i] * max_idepth;
    if (pClose[2] < (S)0.001f) {
      max_idepth = ((S)0.001f - pInf[2]) / Kt[2];
      for (int i = 0; i < 3; ++i) pClose[i] = pInf[i] + Kt[i] * max_idepth;
    }
    {
      S zc = pClose[2];
      for (int i = 0; i < 3; ++i) pClose[i] = pClose[i] / zc;
    }
    for (int i = 0; i < 3; ++i) pFar[i] = pInf[i] + Kt[i] * min_idepth;
    if (pFar[2] < (S)0.001f || max_idepth < min_idepth) return -1;
    {
      S zf = pFar[2];
      for (int i = 0; i < 3; ++i) pFar[i] = pFar[i] / zf;
    }
    // the reference tests only the x components; the y components are tested
    // too so that no NaN coordinate reaches the bilinear sampler
    if (std::isnan(pFar[0] + pClose[0]) || std::isnan(pFar[1] + pClose[1])) return -4;

    S incx = pClose[0] - pFar[0];
    S incy = pClose[1] - pFar[1];
    S eplLength = std::sqrt(incx * incx + incy * incy);
    if (!(eplLength > 0) || std::isinf(eplLength)) return -4;
    S crop = (S)P.max_epl_length_crop;
    if (eplLength > crop) {
      pClose[0] = pFar[0] + incx * crop / eplLength;
      pClose[1] = pFar[1] + incy * crop / eplLength;
    }
    incx *= (S)P.gradient_sample_dist / eplLength;
    incy *= (S)P.gradient_sample_dist / eplLength;
    pFar[0] -= incx; pFar[1] -= incy;
    pClose[0] += incx; pClose[1] += incy;
    S mincrop = (S)P.min_epl_length_crop;
    if (eplLength < mincrop) {
      S pad = (mincrop - eplLength) / (S)2;
      pFar[0] -= incx * pad; pFar[1] -= incy * pad;
      pClose[0] += incx * pad; pClose[1] += incy * pad;
    }
    S B = (S)P.sample_point_to_border;
    S Wb = (S)width - B, Hb = (S)height - B;
    if (pFar[0] <= B || pFar[0] >= Wb || pFar[1] <= B || pFar[1] >= Hb) return -1;
    if (pClose[0] <= B || pClose[0] >= Wb || pClose[1] <= B || pClose[1] >= Hb) {
      if (pClose[0] <= B) {
        S toAdd = (B - pClose[0]) / incx;
        pClose[0] += toAdd * incx; pClose[1] += toAdd * incy;
      } else if (pClose[0] >= Wb) {
        S toAdd = (Wb - pClose[0]) / incx;
        pClose[0] += toAdd * incx; pClose[1] += toAdd * incy;
      }
      if (pClose[1] <= B) {
        S toAdd = (B - pClose[1]) / incy;
        pClose[0] += toAdd * incx; pClose[1] += toAdd * incy;
      } else if (pClose[1] >= Hb) {
        S toAdd = (Hb - pClose[1]) / incy;
        pClose[0] += toAdd * incx; pClose[1] += toAdd * incy;
      }
      S fincx = pClose[0] - pFar[0], fincy = pClose[1] - pFar[1];
      S newLen = std::sqrt(fincx * fincx + fincy * fincy);
      if (pClose[0] <= B || pClose[0] >= Wb || pClose[1] <= B || pClose[1] >= Hb || newLen < (S)8.0f) return -1;
    }

    S cpx = pFar[0], cpy = pFar[1];
    S val_m2 = interp(refImg, cpx - (S)2.0f * incx, cpy - (S)2.0f * incy, width);
    S val_m1 = interp(refImg, cpx - incx, cpy - incy, width);
    S val = interp(refImg, cpx, cpy, width);
    S val_p1 = interp(refImg, cpx + incx, cpy + incy, width);
    S val_p2;
    int loop = 0;
    S best_x = -1, best_y = -1;
    S best_err = (S)INFINITY, second_err = (S)INFINITY;
    S best_errPre = NAN, best_errPost = NAN, best_DiffPre = NAN, best_DiffPost = NAN;
    bool bestWasLast = false;
    S eeLast = -1;
    S e1A = NAN, e1B = NAN, e2A = NAN, e2B = NAN, e3A = NAN, e3B = NAN, e4A = NAN, e4B = NAN, e5A = NAN, e5B = NAN;
    int loopBest = -1, loopSecond = -1;
    const int maxSteps = P.stereo_max_steps;
    while ((((incx < 0) == (cpx > pClose[0]) && (incy < 0) == (cpy > pClose[1])) || loop == 0) && loop < maxSteps) {
      val_p2 = interp(refImg, cpx + (S)2 * incx, cpy + (S)2 * incy, width);
      S ee = 0;
      if (loop % 2 == 0) {
        e1A = val_p2 - rv_p2; ee += e1A * e1A;
        e2A = val_p1 - rv_p1; ee += e2A * e2A;
        e3A = val - rv; ee += e3A * e3A;
        e4A = val_m1 - rv_m1; ee += e4A * e4A;
        e5A = val_m2 - rv_m2; ee += e5A * e5A;
      } else {
        e1B = val_p2 - rv_p2; ee += e1B * e1B;
        e2B = val_p1 - rv_p1; ee += e2B * e2B;
        e3B = val - rv; ee += e3B * e3B;
        e4B = val_m1 - rv_m1; ee += e4B * e4B;
        e5B = val_m2 - rv_m2; ee += e5B * e5B;
      }
      if (ee < best_err) {
        second_err = best_err;
        loopSecond = loopBest;
        best_err = ee;
        loopBest = loop;
        best_errPre = eeLast;
        best_DiffPre = e1A * e1B + e2A * e2B + e3A * e3B + e4A * e4B + e5A * e5B;
        best_errPost = -1;
        best_DiffPost = -1;
        best_x = cpx;
        best_y = cpy;
        bestWasLast = true;
      } else {
        if (bestWasLast) {
          best_errPost = ee;
          best_DiffPost = e1A * e1B + e2A * e2B + e3A * e3B + e4A * e4B + e5A * e5B;
          bestWasLast = false;
        }
        if (ee < second_err) {
          second_err = ee;
          loopSecond = loop;
        }
      }
      eeLast = ee;
      val_m2 = val_m1; val_m1 = val; val = val_p1; val_p1 = val_p2;
      cpx += incx;
      cpy += incy;
      loop++;
    }
    if (best_err > (S)4.0f * (S)P.max_error_stereo) return -3;
    if ((S)std::abs(loopBest - loopSecond) > (S)1.0f && (S)P.min_distance_error_stereo * best_err > second_err)
      return -2;

    bool didSubpixel = false;
    if (P.use_subpixel_stereo) {
      S gradPre_pre = -(best_errPre - best_DiffPre);
      S gradPre_this = +(best_err - best_DiffPre);
      S gradPost_this = -(best_err - best_DiffPost);
      S gradPost_post = +(best_errPost - best_DiffPost);
      bool interpPost = false, interpPre = false;
      if (P.stereo_subpixel_oob_guard && (best_errPre < 0 || best_errPost < 0)) {
      } else if ((gradPost_this < 0) ^ (gradPre_this < 0)) {
      } else if ((gradPre_pre < 0) ^ (gradPre_this < 0)) {
        if (!((gradPost_post < 0) ^ (gradPost_this < 0))) interpPre = true;
      } else if ((gradPost_post < 0) ^ (gradPost_this < 0)) {
        interpPost = true;
      }
      if (interpPre) {
        S d = gradPre_this / (gradPre_this - gradPre_pre);
        best_x -= d * incx;
        best_y -= d * incy;
        best_err = best_err - (S)2 * d * gradPre_this - (gradPre_pre - gradPre_this) * d * d;
        didSubpixel = true;
      } else if (interpPost) {
        S d = gradPost_this / (gradPost_this - gradPost_post);
        best_x += d * incx;
        best_y += d * incy;
        best_err = best_err + (S)2 * d * gradPost_this + (gradPost_post - gradPost_this) * d * d;
        didSubpixel = true;
      }
    }
    S sampleDist = (S)P.gradient_sample_dist * rescale;
    S gradAlongLine = 0, tmp;
    tmp = rv_p2 - rv_p1; gradAlongLine += tmp * tmp;
    tmp = rv_p1 - rv; gradAlongLine += tmp * tmp;
    tmp = rv - rv_m1; gradAlongLine += tmp * tmp;
    tmp = rv_m1 - rv_m2; gradAlongLine += tmp * tmp;
    gradAlongLine /= sampleDist * sampleDist;
    if (best_err > (S)P.max_error_stereo + std::sqrt(gradAlongLine) * (S)20) return -3;

    S idnew, alpha;
    const S* t = ref->otherToThis_t;
    if (incx * incx > incy * incy) {
      S oldX = fxi * best_x + cxi;
      S nominator = oldX * t[2] - t[0];
      S dot0 = KinvP[0] * ref->R_row[0][0] + KinvP[1] * ref->R_row[0][1] + KinvP[2] * ref->R_row[0][2];
      S dot2 = KinvP[0] * ref->R_row[2][0] + KinvP[1] * ref->R_row[2][1] + KinvP[2] * ref->R_row[2][2];
      idnew = (dot0 - oldX * dot2) / nominator;
      alpha = incx * fxi * (dot0 * t[2] - dot2 * t[0]) / (nominator * nominator);
    } else {
      S oldY = fyi * best_y + cyi;
      S nominator = oldY * t[2] - t[1];
      S dot1 = KinvP[0] * ref->R_row[1][0] + KinvP[1] * ref->R_row[1][1] + KinvP[2] * ref->R_row[1][2];
      S dot2 = KinvP[0] * ref->R_row[2][0] + KinvP[1] * ref->R_row[2][1] + KinvP[2] * ref->R_row[2][2];
      idnew = (dot1 - oldY * dot2) / nominator;
      alpha = incy * fyi * (dot1 * t[2] - dot2 * t[1]) / (nominator * nominator);
    }
    if (idnew < 0 && !P.allow_negative_idepths) return -2;
    S photoDispError = (S)4.0f * (S)P.camera_pixel_noise2 / (gradAlongLine + (S)P.division_eps);
    S trackingErrorFac = (S)0.25f * ((S)1.0f + ref->initialTrackedResidual);
    S g2[2];
    interp4(kf->grad[0].data(), u, v, width, 2, g2);
    S geo = (g2[0] * epxn + g2[1] * epyn) + (S)P.division_eps;
    geo = trackingErrorFac * trackingErrorFac * (g2[0] * g2[0] + g2[1] * g2[1]) / (geo * geo);
    r_var = alpha * alpha * ((didSubpixel ? (S)0.05f : (S)0.5f) * sampleDist * sampleDist + geo + photoDispError);
    r_idepth = idnew;
    r_epl = eplLength;
    return best_err;
  }

  bool wasGoodGate(const Frame<S>* refFrame, int x, int y) {
    if (refFrame->parentId == kf->id) {
      const uint8_t* wg = refFrame->refPixelWasGoodNoCreate();
      int ml = 1;
      if (wg && !wg[(x >> ml) + (width >> ml) * (y >> ml)]) return false;
    }
    return true;
  }

  uint8_t observeDepthCreate(int x, int y, int idx) {
    Hyp<S>* target = &cur[idx];
    Frame<S>* refFrame = kfReactivated ? newestRef : oldestRef;
    if (!wasGoodGate(refFrame, x, y)) return LSDG_OBS_SKIP_NOT_GOOD_TRACKED;
    S epx, epy;
    if (!makeAndCheckEPL(x, y, refFrame, &epx, &epy)) return LSDG_OBS_EPL_FAIL;
    S rid, rvar, repl;
    S err = doLineStereo((S)x, (S)y, epx, epy, (S)0, (S)1, (S)1 / (S)P.min_depth, refFrame, refFrame->image[0].data(),
                         rid, rvar, repl);
    if (err == -3 || err == -2) target->blacklisted--;
    if (err < 0) return (uint8_t)(LSDG_OBS_STEREO_FAIL_BASE - (int)err);
    if (rvar > (S)P.max_var) return LSDG_OBS_CREATE_VAR_TOO_BIG;
    rid = unzero(rid);
    *target = Hyp<S>::make(rid, rvar, P.validity_counter_initial_observe);
    return LSDG_OBS_CREATED;
  }

  uint8_t observeDepthUpdate(int x, int y, int idx, const S* maxg) {
    Hyp<S>* target = &cur[idx];
    Frame<S>* refFrame;
    if (!kfReactivated) {
      int k = (int)target->nextStereoFrameMinID - refByIdOffset;
      if (k >= (int)refById.size()) return LSDG_OBS_SKIP_ALREADY_GOOD;
      refFrame = k < 0 ? oldestRef : refById[k];
    } else {
      refFrame = newestRef;
    }
    if (!wasGoodGate(refFrame, x, y)) return LSDG_OBS_SKIP_NOT_GOOD_TRACKED;
    S epx, epy;
    if (!makeAndCheckEPL(x, y, refFrame, &epx, &epy)) return LSDG_OBS_EPL_FAIL;
    S sv = std::sqrt(target->idepth_var_smoothed);
    S min_id = target->idepth_smoothed - sv * (S)P.stereo_epl_var_fac;
    S max_id = target->idepth_smoothed + sv * (S)P.stereo_epl_var_fac;
    if (min_id < 0) min_id = 0;
    if (max_id > (S)1 / (S)P.min_depth) max_id = (S)1 / (S)P.min_depth;
    S rid, rvar, repl;
    S err = doLineStereo((S)x, (S)y, epx, epy, min_id, target->idepth_smoothed, max_id, refFrame,
                         refFrame->image[0].data(), rid, rvar, repl);
    S diff = rid - target->idepth_smoothed;
    if (err == -1) return LSDG_OBS_STEREO_FAIL_BASE + 1;
    if (err == -2) {
      target->validity_counter -= P.validity_counter_dec;
      if (target->validity_counter < 0) target->validity_counter = 0;
      target->nextStereoFrameMinID = 0;
      target->idepth_var *= (S)P.fail_var_inc_fac;
      if (target->idepth_var > (S)P.max_var) {
        target->isValid = false;
        target->blacklisted--;
      }
      return LSDG_OBS_STEREO_FAIL_BASE + 2;
    }
    if (err == -3) return LSDG_OBS_STEREO_FAIL_BASE + 3;
    if (err == -4) return LSDG_OBS_STEREO_FAIL_BASE + 4;
    if ((S)P.diff_fac_observe * diff * diff > rvar + target->idepth_var_smoothed) {
      target->idepth_var *= (S)P.fail_var_inc_fac;
      if (target->idepth_var > (S)P.max_var) target->isValid = false;
      return LSDG_OBS_INCONSISTENT;
    }
    S id_var = target->idepth_var * (S)P.succ_var_inc_fac;
    S w = rvar / (rvar + id_var);
    S new_id = ((S)1 - w) * rid + w * target->idepth;
    target->idepth = unzero(new_id);
    id_var = id_var * w;
    if (id_var < target->idepth_var) target->idepth_var = id_var;
    target->validity_counter += P.validity_counter_inc;
    S absGrad = maxg[idx];
    S cap = (S)P.validity_counter_max + absGrad * (S)P.validity_counter_max_variable / (S)255.0f;
    if ((S)target->validity_counter > cap) target->validity_counter = (int)cap;
    if (repl < (S)P.min_epl_length_crop) {
      float inc = (float)kf->numFramesTrackedOnThis / (float)(kf->numMappedOnThis + 5);
      if (inc < 3) inc = 3;
      inc += (float)(((int)((float)repl * 10000.0f)) % 2);
      if ((double)repl < 0.5 * (double)P.min_epl_length_crop) inc *= 3;
      target->nextStereoFrameMinID = (float)refFrame->id + inc;
    }
    return LSDG_OBS_UPDATED;
  }

  void observeDepthRow(int y0, int y1) {
    const S* maxg = kf->maxg[0].data();
    for (int y = y0; y < y1; ++y)
      for (int x = 3; x < width - 3; ++x) {
        int idx = x + y * width;
        Hyp<S>* target = &cur[idx];
        bool has = target->isValid;
        if (has && maxg[idx] < (S)P.min_abs_grad_decrease) {
          target->isValid = false;
          obsCode[idx] = LSDG_OBS_LOWGRAD_INVALIDATE;
          continue;
        }
        if (maxg[idx] < (S)P.min_abs_grad_create || target->blacklisted < P.min_blacklist) continue;
        obsCode[idx] = has ? observeDepthUpdate(x, y, idx, maxg) : observeDepthCreate(x, y, idx);
      }
  }
  void observeDepth() {
    std::fill(obsCode.begin(), obsCode.end(), 0);
    reduce(3, height - 3, 10, [this](int a, int b) { observeDepthRow(a, b); });
  }

  // regularizeDepthMapFillHoles + buildRegIntegralBuffer.
  void fillHoles() {
    int W = width, H = height;
    std::vector<int> integral((size_t)W * H);
    for (int y = 0; y < H; ++y) {
      int s = 0;
      for (int x = 0; x < W; ++x) {
        const Hyp<S>& h = cur[x + y * W];
        if (h.isValid) s += h.validity_counter;
        integral[x + y * W] = s;
      }
    }
    for (int idx = W; idx < W * H; ++idx) integral[idx] += integral[idx - W];
    oth = cur;
    const S* maxg = kf->maxg[0].data();
    reduce(3, H - 2, 10, [&](int y0, int y1) {
      for (int y = y0; y < y1; ++y)
        for (int x = 3; x < W - 2; ++x) {
          int idx = x + y * W;
          const Hyp<S>& dest = oth[idx];
          if (dest.isValid) continue;
          if (maxg[idx] < (S)P.min_abs_grad_decrease) continue;
          const int* io = integral.data() + idx;
          int val = io[2 + 2 * W] - io[2 - 3 * W] - io[-3 + 2 * W] + io[-3 - 3 * W];
          if ((dest.blacklisted >= P.min_blacklist && val > P.val_sum_min_for_create) ||
              val > P.val_sum_min_for_unblacklist) {
            S sumId = 0, sumIv = 0;
            for (int yy = y - 2; yy <= y + 2; ++yy)
              for (int xx = x - 2; xx <= x + 2; ++xx) {
                const Hyp<S>& s = oth[xx + yy * W];
                if (!s.isValid) continue;
                sumId += s.idepth / s.idepth_var;
                sumIv += (S)1 / s.idepth_var;
              }
            S obs = sumId / sumIv;
            obs = unzero(obs);
            cur[idx] = Hyp<S>::make(obs, (S)P.var_random_init_initial, 0);
          }
        }
    });
  }

  // regularizeDepthMap(removeOcclusions, validityTH).
  void regularize(bool removeOcclusions, int validityTH) {
    oth = cur;
    int W = width;
    S regDistVar = (S)0.075f * (S)0.075f * (S)P.depth_smoothing_factor * (S)P.depth_smoothing_factor;
    reduce(2, height - 2, 10, [&](int y0, int y1) {
      for (int y = y0; y < y1; ++y)
        for (int x = 2; x < W - 2; ++x) {
          Hyp<S>* dest = &cur[x + y * W];
          const Hyp<S>* dr = &oth[x + y * W];
          if (!dr->isValid) continue;
          S sum = 0, val_sum = 0, sumIvar = 0;
          int nOcc = 0, nNot = 0;
          for (int dx = -2; dx <= 2; ++dx)
            for (int dy = -2; dy <= 2; ++dy) {
              const Hyp<S>* s = dr + dx + dy * W;
              if (!s->isValid) continue;
              S diff = s->idepth - dr->idepth;
              if ((S)P.diff_fac_smoothing * diff * diff > s->idepth_var + dr->idepth_var) {
                if (removeOcclusions && s->idepth > dr->idepth) nOcc++;
                continue;
              }
              val_sum += (S)s->validity_counter;
              if (removeOcclusions) nNot++;
              S distFac = (S)(dx * dx + dy * dy) * regDistVar;
              S ivar = (S)1.0f / (s->idepth_var + distFac);
              sum += s->idepth * ivar;
              sumIvar += ivar;
            }
          if (val_sum < (S)validityTH) {
            dest->isValid = false;
            dest->blacklisted--;
            continue;
          }
          if (removeOcclusions && nOcc > nNot) {
            dest->isValid = false;
            continue;
          }
          sum = sum / sumIvar;
          sum = unzero(sum);
          dest->idepth_smoothed = sum;
          dest->idepth_var_smoothed = (S)1.0f / sumIvar;
        }
    });
  }

  // propagateDepth(new_keyframe): raster-order forward warp into the other
  // buffer with occlusion / create / merge rules, then swap.
  void propagateDepth(Frame<S>* nkf) {
    for (auto& h : oth) { h.isValid = false; h.blacklisted = 0; }
    Sim3 oldToNew = nkf->thisToParent;
    oldToNew.s = 1.0;
    oldToNew = oldToNew.inverse();
    S R[9], t[3];
    pose_RT(oldToNew, R, t, false);
    const uint8_t* twg = (nkf->parentId == kf->id) ? nkf->refPixelWasGoodNoCreate() : nullptr;
    const S* kfImg = kf->image[0].data();
    const S* newMaxG = nkf->maxg[0].data();
    const S* newImg = nkf->image[0].data();
    int W = width, H = height;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const Hyp<S>* src = &cur[x + y * W];
        if (!src->isValid) continue;
        S k0 = (S)x * fxi + cxi, k1 = (S)y * fyi + cyi, k2 = 1;
        S pn[3];
        for (int i = 0; i < 3; ++i)
          pn[i] = (R[3 * i] * k0 + R[3 * i + 1] * k1 + R[3 * i + 2] * k2) / src->idepth_smoothed + t[i];
        S new_id = (S)1.0f / pn[2];
        S u = pn[0] * new_id * fx + cx;
        S v = pn[1] * new_id * fy + cy;
        if (!(u > (S)2.1f && v > (S)2.1f && u < (S)W - (S)3.1f && v < (S)H - (S)3.1f)) continue;
        int nidx = (int)(u + (S)0.5f) + ((int)(v + (S)0.5f)) * W;
        S destAbsGrad = newMaxG[nidx];
        if (twg) {
          if (!twg[(x >> 1) + (W >> 1) * (y >> 1)] || destAbsGrad < (S)P.min_abs_grad_decrease) continue;
        } else {
          S sc = kfImg[x + y * W];
          S dc = interp(newImg, u, v, W);
          S res = dc - sc;
          if (res * res / ((S)P.max_diff_constant + (S)P.max_diff_grad_mult * destAbsGrad * destAbsGrad) > (S)1.0f ||
              destAbsGrad < (S)P.min_abs_grad_decrease)
            continue;
        }
        Hyp<S>* tb = &oth[nidx];
        S r4 = new_id / src->idepth_smoothed;
        r4 *= r4;
        r4 *= r4;
        S new_var = r4 * src->idepth_var;
        if (tb->isValid) {
          S diff = tb->idepth - new_id;
          if ((S)P.diff_fac_prop_merge * diff * diff > new_var + tb->idepth_var) {
            if (new_id < tb->idepth) continue;
            tb->isValid = false;
          }
        }
        if (!tb->isValid) {
          *tb = Hyp<S>::make(new_id, new_var, std::min(P.validity_counter_max_variable, src->validity_counter));
        } else {
          S w = new_var / (tb->idepth_var + new_var);
          S merged = w * tb->idepth + ((S)1.0f - w) * new_id;
          int mv = src->validity_counter + tb->validity_counter;
          int cap = P.validity_counter_max + P.validity_counter_max_variable;
          if (mv > cap) mv = cap;
          *tb = Hyp<S>::make(merged, (S)1.0f / ((S)1.0f / tb->idepth_var + (S)1.0f / new_var), mv);
        }
      }
    std::swap(cur, oth);
  }

  // Frame::setDepth: valid && smoothed >= -0.05 -> (smoothed, var_smoothed),
  // else (-1, -1); then the depth pyramid.
  void setDepth(Frame<S>* f) {
    size_t n = (size_t)width * height;
    std::vector<S> id(n), var(n);
    for (size_t i = 0; i < n; ++i) {
      const Hyp<S>& h = cur[i];
      if (h.isValid && (double)h.idepth_smoothed >= -0.05) {
        id[i] = unzero(h.idepth_smoothed);
        var[i] = unzero(h.idepth_var_smoothed);
      } else {
        id[i] = -1;
        var[i] = -1;
      }
    }
    f->setDepth0(id.data(), var.data());
  }

  void initializeFromGT(Frame<S>* f) {
    kf = f;
    kfReactivated = false;
    const S* id = f->idepth[0].data();
    for (size_t i = 0; i < cur.size(); ++i) {
      S v = id[i];
      if (!std::isnan(v) && v > 0)
        cur[i] = Hyp<S>::make(v, v, (S)P.var_gt_init_initial, (S)P.var_gt_init_initial, 20);
      else {
        cur[i].isValid = false;
        cur[i].blacklisted = 0;
      }
    }
    setDepth(f);
  }
  // DepthMap::initializeRandomly, with random_draw (SplitMix64) in place of rand().
  void initializeRandomly(Frame<S>* f, uint64_t seed) {
    kf = f;
    kfReactivated = false;
    const S* maxg = f->maxg[0].data();
    for (int y = 1; y < height - 1; ++y)
      for (int x = 1; x < width - 1; ++x) {
        const int i = x + y * width;
        if (maxg[i] > (S)P.min_abs_grad_create) {
          const float idf = 0.5f + 1.0f * ((float)(random_draw(seed, i) % 100001) / 100000.0f);
          const S id = (S)idf;
          cur[i] = Hyp<S>::make(id, id, (S)P.var_random_init_initial, (S)P.var_random_init_initial, 20);
        } else {
          cur[i].isValid = false;
          cur[i].blacklisted = 0;
        }
      }
    setDepth(f);
  }
  void setPrior(Frame<S>* f, const float* id, const float* var, int validity) {
    kf = f;
    kfReactivated = false;
    for (size_t i = 0; i < cur.size(); ++i) {
      S v = (S)id[i], s = (S)var[i];
      if (std::isfinite(v) && s > 0)
        cur[i] = Hyp<S>::make(v, v, s, s, validity);
      else {
        cur[i] = Hyp<S>();
      }
    }
    setDepth(f);
  }

  void updateKeyframe(std::vector<Frame<S>*>& refs) {
    oldestRef = refs.front();
    newestRef = refs.back();
    refById.clear();
    refByIdOffset = oldestRef->id;
    for (Frame<S>* f : refs) {
      Sim3 refToKf;
      if (f->parentId == kf->id) refToKf = f->thisToParent;
      else refToKf = kf->camToWorld.inverse() * f->camToWorld;
      f->prepareForStereoWith(refToKf, K);
      while ((int)refById.size() + refByIdOffset <= f->id) refById.push_back(f);
    }
    observeDepth();
    fillHoles();
    regularize(false, P.val_sum_min_for_keep);
    setDepth(kf);
    kf->numMappedOnThis++;
  }

  void createKeyFrame(Frame<S>* nkf) {
    Sim3 oldToNew = nkf->thisToParent;
    oldToNew.s = 1.0;
    oldToNew = oldToNew.inverse();
    propagateDepth(nkf);
    kf = nkf;
    kfReactivated = false;
    regularize(true, P.val_sum_min_for_keep);
    fillHoles();
    regularize(false, P.val_sum_min_for_keep);
    // mean-idepth rescale: a sequential FP32 sum as the reference, or an FP64
    // sum (FP64 oracle, and the FP32 restatement with FP64 accumulation)
    S rescale;
    if (sizeof(S) == 8 || accDouble) {
      double s = 0, n = 0;
      for (auto& h : cur)
        if (h.isValid) { s += (double)h.idepth_smoothed; n += 1; }
      rescale = (S)(n / s);
    } else {
      float sumIdepth = 0, numIdepth = 0;
      for (auto& h : cur) {
        if (!h.isValid) continue;
        sumIdepth += (float)h.idepth_smoothed;
        numIdepth++;
      }
      rescale = (S)(numIdepth / sumIdepth);
    }
    S rescale2 = rescale * rescale;
    for (auto& h : cur) {
      if (!h.isValid) continue;
      h.idepth *= rescale;
      h.idepth_smoothed *= rescale;
      h.idepth_var *= rescale2;
      h.idepth_var_smoothed *= rescale2;
    }
    Sim3 tp = oldToNew.inverse();
    tp.s = (double)rescale;
    kf->thisToParent = tp;
    setDepth(kf);
  }

  void finalizeKeyFrame() {
    fillHoles();
    regularize(false, P.val_sum_min_for_keep);
    setDepth(kf);
  }
};

}  // namespace orc
