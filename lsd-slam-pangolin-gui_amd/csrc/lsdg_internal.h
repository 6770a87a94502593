// Internal object layouts of liblsdg.so: contexts, frames (device pyramids),
// tracking references (device point clouds), depth maps (device hypothesis
// buffers) and the kernel launchers the C-ABI layer calls.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lsdg.h"
#include "lsdg_lie.h"

namespace lsdg {

constexpr int PYR = LSDG_PYRAMID_LEVELS;

// Number of FP64 sums one fused tracker evaluation produces.
//  SE3 layout: [0,21) A upper triangle (row-major), [21,27) b, 27 err,
//  28 sumResUnweighted(good), 29 sumSignedRes(good), 30 usage,
//  31..35 sxx syy sx sy sw, 36 warped, 37 good, 38 bad.
//  Sim3 layout: [0,28) A upper triangle of 7x7, [28,35) b, 35 sumResP,
//  36 sumResD, 37 numTermsD, 38 warped, 39 usage, 40..44 sxx syy sx sy sw,
//  45 errP (= sumResP, sum W_p r_p^2 from the LGS pass), 46 errD.
constexpr int NS_SE3 = 40;
constexpr int NS_SIM3 = 48;
constexpr int NS_MAX = 48;

// Pyramid data of a frame, one validity bit per (kind, level): the
// reference's Frame::require() flags (IMAGE, GRADIENTS, MAX_GRADIENTS), built
// on first use.  FD_GRADH is the packed FP16/BF16 (gx, gy, I) copy that
// reduced-precision frames track on.
enum FrameData : unsigned { FD_IMAGE = 0, FD_GRAD = PYR, FD_GRADH = 2 * PYR, FD_MAXG = 3 * PYR };
constexpr unsigned fd_bit(unsigned kind, int l) { return 1u << (kind + (unsigned)l); }
constexpr unsigned fd_levels(unsigned kind, int lo, int hi) {
  return lo > hi ? 0u : (fd_bit(kind, lo) | fd_levels(kind, lo + 1, hi));
}

enum Se3Slot { S3_A = 0, S3_B = 21, S3_ERR = 27, S3_RESUNW = 28, S3_SIGNED = 29, S3_USAGE = 30, S3_SXX = 31,
               S3_SYY = 32, S3_SX = 33, S3_SY = 34, S3_SW = 35, S3_WARPED = 36, S3_GOOD = 37, S3_BAD = 38 };
enum Sim3Slot { SM_A = 0, SM_B = 28, SM_RESP = 35, SM_RESD = 36, SM_NTD = 37, SM_WARPED = 38, SM_USAGE = 39,
                SM_SXX = 40, SM_SYY = 41, SM_SX = 42, SM_SY = 43, SM_SW = 44 };
// Sums of the adjoint evaluations.  SE3: dE/dxi (6), E, count.  Sim3: dE/dxi
// (7), the photometric and depth parts of 2E, count.
constexpr int NS_ADJ = 8;
enum AdjSim3Slot { AS_G = 0, AS_EP = 7, AS_ED = 8, AS_COUNT = 9, NS_ADJ_SIM3 = 10 };

struct LevelGeom {
  int w, h;
  float fx, fy, cx, cy;
};

// Rigid (or scaled) transform in FP32 for per-point work.
struct Xf {
  float R[9];
  float t[3];
};

struct TrackKParams {
  float huber;        // SE3: huber_d / 2 ; Sim3: huber_d
  float var_weight;
  float noise2;       // cameraPixelNoise2
  float max_diff_constant;
  float max_diff_grad_mult;
  float aff_a, aff_b;
  int use_esm;
  int aff_stats;      // accumulate the affine-brightness sums (sxx, syy, sx, sy, sw)
  float roll[4];      // Sim3 ESM roll (xRoll0, xRoll1, yRoll0, yRoll1)
};

// Depth hypotheses (DepthMapPixelHypothesis) as structure of arrays:
// dv = (idepth, idepth_var, idepth_smoothed, idepth_var_smoothed).
struct HypBuf {
  uint8_t* valid;
  int* bl;       // blacklisted
  float* nsid;   // nextStereoFrameMinID
  int* val;      // validity_counter
  float4* dv;
};

// Frame::prepareForStereoWith products of one reference frame, FP32.
struct StereoRef {
  const float* img;
  const uint8_t* wasGood;  // refPixelWasGood gate, null when not tracked on the keyframe
  float KR[9];             // K_otherToThis_R
  float Kt[3];             // K_otherToThis_t
  float t_o2t[3];          // otherToThis_t
  float t_t2o[3];          // thisToOther_t
  float Rrow[9];           // otherToThis_R_row0..2
  float initialTrackedResidual;
  int id;
};

struct DepthGeom {
  int w, h;
  float fx, fy, cx, cy, fxi, fyi, cxi, cyi;
};

// Device tables of a batched pyramid rebuild.
// A frame's inverse-depth pyramid (levels 0 .. PYR - 1, each half the one
// above, rounded down).
struct PyrDepthArgs {
  float* id[PYR];
  float* var[PYR];
  int w[PYR], h[PYR];
};

// Frame::buildIDepthAndIDepthVar for one pixel: inverse-variance fusion of
// its 4 children (order: (0,0), (1,0), (0,1), (1,1)); (-1, -1) when none has
// a positive variance.
__device__ inline void fuse4(const float* ida, const float* va, float& oid, float& ov) {
  float isum = 0.f, ivsum = 0.f;
  int num = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float v = va[k];
    if (v > 0.f) {
      float iv = 1.0f / v;
      ivsum += iv;
      isum += iv * ida[k];
      num++;
    }
  }
  if (num > 0) {
    float depth = ivsum / isum;
    oid = 1.0f / depth;
    ov = (float)num / ivsum;
  } else {
    oid = -1.f;
    ov = -1.f;
  }
}

// The merge by the last block of a launch: out[k] = the sum over b of
// rows[b * NS + k], in block order (k < NS; BS threads).  The rows are staged
// through LDS a chunk at a time, every thread issuing its agent-scope atomic
// loads back to back (the rows were written with agent-scope stores; no L2
// invalidation needed), so a chunk costs one memory round trip rather than
// one per row; NS lanes then add their column in block order, as a
// sequential loop over the rows would.
constexpr int MERGE_CHUNK = 2048;  // doubles staged per round (16 KB of LDS)
template <int NS, int BS>
__device__ inline void merge_rows(const double* __restrict__ rows, unsigned nrows, double* __restrict__ out) {
  static_assert(MERGE_CHUNK % BS == 0 && NS <= BS, "merge layout");
  constexpr int LOADS = MERGE_CHUNK / BS;
  __shared__ double stage[MERGE_CHUNK];
  const unsigned total = nrows * NS;
  double s = 0.0;
  for (unsigned base = 0; base < total; base += MERGE_CHUNK) {
    const unsigned m = min((unsigned)MERGE_CHUNK, total - base);
    unsigned long long v[LOADS];
#pragma unroll
    for (int u = 0; u < LOADS; ++u) {
      const unsigned i = threadIdx.x + u * BS;
      v[u] = i < m ? __hip_atomic_load((const unsigned long long*)&rows[base + i], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT)
                   : 0ull;
    }
#pragma unroll
    for (int u = 0; u < LOADS; ++u) {
      const unsigned i = threadIdx.x + u * BS;
      if (i < m) stage[i] = __longlong_as_double((long long)v[u]);
    }
    __syncthreads();
    if (threadIdx.x < NS) {
      // this column's entries of the chunk, added in row order; eight LDS
      // reads are in flight ahead of the additions
      unsigned i = (threadIdx.x + NS - base % NS) % NS;
      for (; i + 7 * NS < m; i += 8 * NS) {
        double r[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) r[u] = stage[i + u * NS];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += r[u];
      }
      for (; i < m; i += NS) s += stage[i];
    }
    __syncthreads();
  }
  if (threadIdx.x < NS) out[threadIdx.x] = s;
}

struct PyrBatch {
  const void* img_args;
  const void* grad_args;
  const int* img_off;
  const int* img_tx;
  const int* grad_off;
  int img_blocks, grad_blocks, nf;
};

// Start / stop events of one timed kernel launch (both null: untimed).
struct KTimer {
  hipEvent_t a = nullptr, b = nullptr;
};

// Launches kernel f on stream s.  A timed launch goes through
// hipExtLaunchKernel, whose events carry the kernel's own start and end
// timestamps (the figures rocprofv3 reports), not the stream's around it.
template <typename... KArgs, typename... Args>
inline hipError_t launch_k(void (*f)(KArgs...), dim3 grid, dim3 block, hipStream_t s, const KTimer* t,
                           Args... args) {
  if (t && t->a) hipExtLaunchKernelGGL(f, grid, block, 0, s, t->a, t->b, 0, static_cast<KArgs>(args)...);
  else f<<<grid, block, 0, s>>>(static_cast<KArgs>(args)...);
  return hipGetLastError();
}

struct PointsLevel {
  const float4* posI;  // x, y, z, I
  const float* var;
  const float2* grad;  // keyframe gx, gy (Sim3 ESM)
  const int* idx;      // pixel index in the keyframe level
  int n;
};

}  // namespace lsdg

// ------------------------------------------------------------ C-ABI objects
namespace lsdg {
// Keyframe::computeVbo inputs: geometry, camera inverse (fxi = 1 / fx,
// cxi = -cx / fx, ...) and the filter constants.
struct VboArgs {
  int w, h;
  float fxi, fyi, cxi, cyi;
  float scaledTH, absTH, scale;
  int minNear;
};
}  // namespace lsdg

struct lsdg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  lsdg_params P;
  int sums = LSDG_SUMS_FP64;  // lsdg_ctx_set_sums
  std::string err;
  // reduction scratch for the plain (one-alignment-at-a-time) path
  double* d_partials = nullptr;
  int partials_blocks = 0;
  unsigned* d_counter = nullptr;
  double* h_result = nullptr;  // pinned, host-mapped
  double* d_result = nullptr;  // device view of h_result
  int num_cus = 256;
  // per-kernel timing (lsdg_ctx_profile)
  struct ProfRec {
    int kind;
    hipEvent_t a, b;
    double bytes;
  };
  bool prof_on = false;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_events;
  lsdg_kernel_stats prof[LSDG_KERNEL_KINDS];
  // per-point inverse-depth gradients of the adjoint, grown on demand
  float* d_adj = nullptr;
  int adjCap = 0;
  // keyframe export scratch (records, block offsets, vertices), grown on demand
  void* d_export = nullptr;
  size_t exportCap = 0;
};

struct lsdg_frame_s {
  lsdg_ctx ctx = nullptr;
  int id = 0;
  int w[lsdg::PYR], h[lsdg::PYR];
  float fx[lsdg::PYR], fy[lsdg::PYR], cx[lsdg::PYR], cy[lsdg::PYR];
  float fxi[lsdg::PYR], fyi[lsdg::PYR], cxi[lsdg::PYR], cyi[lsdg::PYR];
  int precision = LSDG_FP32;
  // one device allocation holds every pyramid
  void* d_pool = nullptr;
  float* d_image[lsdg::PYR];
  float4* d_grad[lsdg::PYR];
  uint2* d_gradh[lsdg::PYR];  // packed FP16/BF16 (gx, gy, I, 0); null for FP32 frames
  float* d_maxg[lsdg::PYR];
  float* d_idepth[lsdg::PYR];
  float* d_ivar[lsdg::PYR];
  uint8_t* d_input_u8 = nullptr;  // staging for u8 uploads (w*h)
  bool u8Level0 = false;          // d_input_u8 holds level 0 (every level-0 writer keeps it so)
  uint8_t* d_wasGood = nullptr;   // refPixelWasGood (level 1)
  unsigned valid = 0;             // built pyramid data (fd_bit flags)
  bool wasGoodAllocated = false;
  bool hasDepth = false;
  // pose fields
  int parentId = -1;
  lsdg_frame parent = nullptr;
  lsdg::Pose thisToParent = lsdg::pose_identity();
  float initialTrackedResidual = 0.f;
  lsdg::Pose camToWorld = lsdg::pose_identity();
  int numFramesTrackedOnThis = 0, numMappedOnThis = 0;
  // bumped when depth changes (invalidates reference point clouds)
  int depthVersion = 0;
  // stereo preparation (prepareForStereoWith), FP32 values as the reference caches them
  float K_otherToThis_R[9], otherToThis_t[3], K_otherToThis_t[3], thisToOther_t[3], R_row[3][3];
};

struct lsdg_ref_s {
  lsdg_ctx ctx = nullptr;
  lsdg_frame kf = nullptr;
  int depthVersion = -1;
  int num[lsdg::PYR] = {-1, -1, -1, -1, -1};
  int cap[lsdg::PYR] = {0, 0, 0, 0, 0};
  void* d_pool[lsdg::PYR] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  lsdg::PointsLevel lv[lsdg::PYR];
  int* d_blockcounts = nullptr;
  int blockcounts_cap = 0;
};

struct lsdg_depthmap_s {
  lsdg_ctx ctx = nullptr;
  lsdg::DepthGeom g;
  void* d_pool = nullptr;
  // hypothesis maps: buf[cur] and buf[oth] are the reference's current and
  // other maps; the third is the spare the fused fill + regularize pass writes
  lsdg::HypBuf buf[3];
  int cur = 0, oth = 1;
  uint8_t* d_obs = nullptr;
  lsdg_frame kf = nullptr;
  bool kfReactivated = false;
  bool valid = false;
  // observe inputs
  lsdg::StereoRef* d_refs = nullptr;
  int refs_cap = 0;
  int* d_refById = nullptr;
  int refById_cap = 0;
  int4* d_work = nullptr;    // epipolar-search work items, one slot per pixel
  int* d_work_count = nullptr;  // two sets of per-segment (create, update) item counters
  int obsParity = 0;            // counter set of the next observe pass
  int stereoBlocks[2] = {0, 0};  // resident blocks of the stereo pass (refs inline, in memory), 0 = not yet queried
  lsdg::StereoRef* h_refs = nullptr;  // pinned staging
  int* h_refById = nullptr;
  // propagation scratch
  int* d_dest = nullptr;
  float2* d_newIV = nullptr;
  int* d_count = nullptr;   // P + 1
  int* d_fill = nullptr;    // P
  int* d_bucket = nullptr;  // P
  int* d_scanblk = nullptr;
  // rescale reduction
  double* d_red = nullptr;  // partial rows
  unsigned* d_red_counter = nullptr;
  double* h_red = nullptr;  // pinned, host-mapped (count, sum)
  double* d_red_out = nullptr;
};

// ---------------------------------------------------------- launchers
namespace lsdg {

// pyramid.hip
// image pyramid of f: with fromU8, level 0 converted from the 8-bit staging
// and levels 1-4; otherwise levels 1-4 from level 0 (read from the 8-bit
// copy when it still holds level 0)
hipError_t launch_pyr_images(lsdg_frame f, bool fromU8, hipStream_t s);
// undistorts the 8-bit image d_src (row length inW) through the table d_map
// (one source position per level-0 pixel) into f's level 0, builds levels
// 1-4, and writes the 8-bit undistorted image to d_undist unless it is null
hipError_t launch_ingest(lsdg_frame f, const uint8_t* d_src, int inW, const float2* d_map, uint8_t* d_undist,
                         hipStream_t s);
// Gradients / packed copies / max-gradients of the levels named in `need`
// (FD_GRAD, FD_GRADH, FD_MAXG bits); other levels launch no blocks.
hipError_t launch_pyr_grads(lsdg_frame f, unsigned need, hipStream_t s);
hipError_t launch_pyr_depth(lsdg_frame f, hipStream_t s);
hipError_t launch_depth_normalize(float* idepth, float* ivar, int n, hipStream_t s);
hipError_t launch_point_cloud(lsdg_ref r, int level, hipStream_t s);  // fills r->lv[level], r->num[level]
size_t pyr_batch_bytes(int nf);
void pyr_batch_tables(lsdg_frame const* fs, const unsigned* need, int nf, void* h, void* d, PyrBatch* out);
hipError_t launch_pyr_batch(const PyrBatch& b, hipStream_t s);

// keyframe.hip
hipError_t launch_kf_dense(const float* id, const float* var, const float* img, int n, lsdg_point_dense* out,
                           hipStream_t s);
size_t vbo_block_count(int w, int h);
// blk: vbo_block_count + 1 ints; out: capacity (w - 2) * (h - 2)
hipError_t launch_vbo(const lsdg_point_dense* in, const VboArgs& a, int* blk, lsdg_vertex* out, hipStream_t s);

// track.hip
hipError_t launch_se3_eval(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                           const TrackKParams& kp, uint8_t* d_isgood, double* d_out, const KTimer* t = nullptr);
hipError_t launch_sim3_eval(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                            const TrackKParams& kp, double* d_out, const KTimer* t = nullptr);
hipError_t launch_energy_grad(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                              const TrackKParams& kp, float* d_grad_d, double* d_out, const KTimer* t = nullptr);
// lsdg_energy_grad_noalloc: limits of its file-scope device array, the
// once-per-device lookup of that array's address (lsdg_ctx_create calls it
// with the device current) and the launch: clear the counter, run
// k_energy_grad over the array, copy the merged row (NS_ADJ doubles) and, if
// h_grad is set, the per-point dE/dd to host memory, synchronise.
constexpr int NOALLOC_MAX_POINTS = 1 << 20;
constexpr int NOALLOC_MAX_BLOCKS = 1024;
hipError_t noalloc_scratch_init(int device);
hipError_t launch_energy_grad_noalloc(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                                      const TrackKParams& kp, double* h_row, float* h_grad);
// Sim3 adjoint (kp.use_esm must be 0): d_out receives dE/dxi[7], then the
// photometric and depth parts of 2E and the in-image count.
hipError_t launch_energy_grad_sim3(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                                   const TrackKParams& kp, float* d_grad_d, double* d_out,
                                   const KTimer* t = nullptr);

// depthmap.hip
// Observe work list: slots (int4 each) for a w x h depth map, and the bytes of
// its two alternating sets of segment counters.
constexpr size_t OBS_COUNT_BYTES = 8192;
size_t obs_work_slots(int w, int h);
// h_refs / h_refById: host arrays (pinned: copied asynchronously when they do
// not fit in the kernel arguments).
hipError_t launch_observe(lsdg_depthmap dm, const lsdg_params& P, const StereoRef* h_refs, int nRefs,
                          const int* h_refById, int nRefById, int refByIdOffset, int oldestIdx, int newestIdx,
                          float numTracked, float numMapped, hipStream_t s);
hipError_t launch_regularize(lsdg_depthmap dm, const lsdg_params& P, bool removeOcclusions, int th,
                             hipStream_t s);  // swaps buffers
// regularizeDepthMapFillHoles then regularizeDepthMap(false, th) in one pass,
// leaving the current and other maps as the two passes would; with setDepth
// non-null, also launch_set_depth(setDepth) and the frame's inverse-depth
// levels 1-4 (what launch_pyr_depth would write after it)
hipError_t launch_fill_regularize(lsdg_depthmap dm, const lsdg_params& P, int th, lsdg_frame setDepth,
                                  hipStream_t s);
hipError_t launch_set_depth(lsdg_depthmap dm, lsdg_frame f, hipStream_t s);
// DepthMap::initializeRandomly on dm->kf's level-0 max-gradients
hipError_t launch_random_init(lsdg_depthmap dm, const lsdg_params& P, unsigned long long seed, hipStream_t s);
hipError_t launch_propagate(lsdg_depthmap dm, const lsdg_params& P, lsdg_frame nkf, const Xf& oldToNew,
                            const uint8_t* d_wasGood, hipStream_t s);  // swaps buffers
hipError_t launch_mean_idepth(lsdg_depthmap dm, hipStream_t s);  // h_red[0] = count, h_red[1] = sum
hipError_t launch_rescale(lsdg_depthmap dm, float f, hipStream_t s);
hipError_t launch_hyp_prior(lsdg_depthmap dm, const float* d_id, const float* d_var, int validity, int gt,
                            float gtVar, hipStream_t s);

}  // namespace lsdg
