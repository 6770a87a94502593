// Fused tracker evaluation on gfx950: one pass over the reference points per
// LM evaluation.
//
// The reference runs three passes per evaluation with compacted buffers in
// between (SE3Tracker::calcResidualAndBuffers -> calcWeightsAndResidual ->
// calculateWarpUpdate / LGS6; Sim3Tracker::calcSim3Buffers ->
// calcSim3WeightsAndResidual -> calcSim3LGS).  Every quantity those passes
// need is per point and the pose is known up front, so one kernel computes
// warp, bilinear sample, residual, isGood, depth-variance weight, Huber weight,
// the Jacobian row and its normal-equation terms, and reduces them:
//   per thread  FP64 accumulation of the FP32 per-point terms over a
//               grid-stride slice of points (or, with LSDG_SUMS_KAHAN,
//               Kahan-compensated FP32 pairs),
//   per wave    butterfly (__shfl_xor; pairs: TwoSum reduce-scatter),
//   per block   sum of the wave results in LDS -> one FP64 partial row,
//   last block  FP64 sum of all partial rows in block order (deterministic),
//               written straight to the caller's (possibly host-mapped) buffer.
// The per-point float expressions keep the reference's operation order so the
// isGood mask and counts match the CPU restatement bit for bit.
#include <mutex>

#include "track_device.h"

namespace lsdg {

// Block-wide reduction of the per-thread accumulators into one FP64 partial
// row, then the last block to finish merges all rows in block order.
template <int MODE, int NS>
__device__ inline void reduce_and_merge(Accum<MODE, NS>& acc, double* __restrict__ partials, unsigned* counter,
                                        double* __restrict__ out) {
  __shared__ double sh[EVAL_WAVES][NS];
  __shared__ bool amLast;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (MODE == ACC_F64) {
    acc.wave_reduce();
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < NS; ++k) sh[wave][k] = acc.value(k);
    }
  } else {
    const double cell = acc.wave_slot();
    if (lane < NS) sh[wave][lane] = cell;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const double s = Accum<MODE, NS>::merge_waves(&sh[0][threadIdx.x], NS, EVAL_WAVES);
    __hip_atomic_store((unsigned long long*)&partials[(size_t)blockIdx.x * NS + threadIdx.x],
                       (unsigned long long)__double_as_longlong(s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  wait_own_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(counter, 1u);
    amLast = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!amLast) return;
  merge_rows<NS, EVAL_BS>(partials, gridDim.x, out);
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide reduction of the per-thread accumulators into this block's FP64
// row (the merge is a second launch, k_merge_rows: no completion counter, so
// no same-address atomics, which serialise once the grid has more than a
// few hundred blocks).
template <int MODE, int NS>
__device__ inline void reduce_to_row(Accum<MODE, NS>& acc, double* __restrict__ rows) {
  __shared__ double sh[EVAL_WAVES][NS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (MODE == ACC_F64) {
    acc.wave_reduce();
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < NS; ++k) sh[wave][k] = acc.value(k);
    }
  } else {
    const double cell = acc.wave_slot();
    if (lane < NS) sh[wave][lane] = cell;
  }
  __syncthreads();
  if (threadIdx.x < NS)
    rows[(size_t)blockIdx.x * NS + threadIdx.x] = Accum<MODE, NS>::merge_waves(&sh[0][threadIdx.x], NS, EVAL_WAVES);
}

// out[k] = the sum of rows[b * NS + k] over the nrows blocks, in a fixed
// order: EVAL_BS / NS lanes per column each add the rows b = lane, lane + L,
// ... in order, then the lanes' sums are added pairwise in a fixed tree.  The
// result does not depend on timing (the adjoint's sums are held to the FP64
// oracle by tolerance, not to a particular order), and the column sums run in
// parallel instead of one dependent addition per row.
template <int NS>
__global__ __launch_bounds__(EVAL_BS) void k_merge_rows(const double* __restrict__ rows, unsigned nrows,
                                                        double* __restrict__ out) {
  constexpr int L = EVAL_BS / NS;  // lanes per column
  __shared__ double part[NS][L];
  const int col = threadIdx.x % NS, lane = threadIdx.x / NS;
  if (lane < L) {
    double s = 0.0;
    unsigned b = lane;
    for (; b + 15 * L < nrows; b += 16 * L) {  // 16 loads in flight, added in row order
      double r[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) r[u] = rows[(size_t)(b + u * L) * NS + col];
#pragma unroll
      for (int u = 0; u < 16; ++u) s += r[u];
    }
    for (; b < nrows; b += L) s += rows[(size_t)b * NS + col];
    part[col][lane] = s;
  }
  __syncthreads();
  // pairwise tree over the lanes (a lane without a partner passes its sum up)
  for (int w = 1; w < L; w <<= 1) {
    if (lane < L && lane % (2 * w) == 0 && lane + w < L) part[col][lane] += part[col][lane + w];
    __syncthreads();
  }
  if (threadIdx.x < NS) out[threadIdx.x] = part[threadIdx.x][0];
}

// ------------------------------------------------------------------- SE3
template <int PREC, int ACC>
__global__ __launch_bounds__(EVAL_BS) void k_se3_eval(PointsLevel pts, const void* __restrict__ img, LevelGeom g,
                                                      Xf x, TrackKParams kp, uint8_t* __restrict__ isGood,
                                                      double* __restrict__ partials, unsigned* counter,
                                                      double* __restrict__ out) {
  Accum<ACC, NS_SE3> acc;
  acc.zero();
  se3_points<PREC, EVAL_U>(pts, blockIdx.x * EVAL_BS + threadIdx.x, gridDim.x * EVAL_BS, pts.n, img, g, x, kp, isGood,
                           acc);
  reduce_and_merge(acc, partials, counter, out);
}

// ------------------------------------------------------------------ Sim3
template <int PREC, int ACC>
__global__ __launch_bounds__(EVAL_BS) void k_sim3_eval(PointsLevel pts, const void* __restrict__ img,
                                                       const float* __restrict__ fid, const float* __restrict__ fvar,
                                                       LevelGeom g, Xf x, TrackKParams kp,
                                                       double* __restrict__ partials, unsigned* counter,
                                                       double* __restrict__ out) {
  Accum<ACC, NS_SIM3> acc;
  acc.zero();
  sim3_points<PREC, EVAL_U>(pts, blockIdx.x * EVAL_BS + threadIdx.x, gridDim.x * EVAL_BS, pts.n, img, fid, fvar, g, x,
                            kp, acc);
  reduce_and_merge(acc, partials, counter, out);
}

// --------------------------------------------------------------- adjoint
// dE/dxi (6) and dE/d(idepth_i) of E = sum rho(sqrt(w_p) r) with w_p and the
// in-image set frozen at x; rho Huber at kp.huber.  out: G[6], E, count.
template <int PREC, int ACC>
__global__ __launch_bounds__(EVAL_BS) void k_energy_grad(PointsLevel pts, const void* __restrict__ img, LevelGeom g,
                                                         Xf x, TrackKParams kp, float* __restrict__ grad_d,
                                                         double* __restrict__ rows) {
  Accum<ACC, NS_ADJ> acc;
  acc.zero();
  for (int i = blockIdx.x * EVAL_BS + threadIdx.x; i < pts.n; i += gridDim.x * EVAL_BS)
    adjoint_point<PREC>(pts, i, img, g, x, kp, grad_d, acc);
  reduce_to_row(acc, rows);
}

// Sim3 adjoint: dE/dxi (7), 2E split into its photometric and depth parts,
// count (AdjSim3Slot), dE/d(idepth_i) into grad_d.
template <int PREC, int ACC>
__global__ __launch_bounds__(EVAL_BS) void k_energy_grad_sim3(PointsLevel pts, const void* __restrict__ img,
                                                              const float* __restrict__ fid,
                                                              const float* __restrict__ fvar, LevelGeom g, Xf x,
                                                              TrackKParams kp, float* __restrict__ grad_d,
                                                              double* __restrict__ rows) {
  Accum<ACC, NS_ADJ_SIM3> acc;
  acc.zero();
  for (int i = blockIdx.x * EVAL_BS + threadIdx.x; i < pts.n; i += gridDim.x * EVAL_BS)
    adjoint_point_sim3<PREC>(pts, i, img, fid, fvar, g, x, kp, grad_d, acc);
  reduce_to_row(acc, rows);
}

// ------------------------------------------------------------ launchers
// Plain-path evaluation grid: one point per thread (the shortest dependency
// chain per thread; 1.74k frames/s on the 64 headline instances against 1.61k
// at four points per thread), at most EVAL_MAX_BLOCKS blocks: every block
// adds once to the launch's completion counter, and same-address atomics
// serialise, so larger levels take more points per thread instead.
constexpr int EVAL_MAX_BLOCKS = 256;
static inline int eval_blocks(lsdg_ctx c, int n) {
  int nb = (n + EVAL_BS - 1) / EVAL_BS;
  if (nb < 1) nb = 1;
  if (nb > EVAL_MAX_BLOCKS) nb = EVAL_MAX_BLOCKS;
  if (nb > c->partials_blocks) nb = c->partials_blocks;
  return nb;
}

static inline const void* level_image(lsdg_frame f, int level) {
  return f->precision == LSDG_FP32 ? (const void*)f->d_grad[level] : (const void*)f->d_gradh[level];
}

template <int ACC>
static auto kern_se3_eval(int prec) {
  return prec == LSDG_FP16   ? k_se3_eval<LSDG_FP16, ACC>
         : prec == LSDG_BF16 ? k_se3_eval<LSDG_BF16, ACC>
                             : k_se3_eval<LSDG_FP32, ACC>;
}
template <int ACC>
static auto kern_sim3_eval(int prec) {
  return prec == LSDG_FP16   ? k_sim3_eval<LSDG_FP16, ACC>
         : prec == LSDG_BF16 ? k_sim3_eval<LSDG_BF16, ACC>
                             : k_sim3_eval<LSDG_FP32, ACC>;
}

hipError_t launch_se3_eval(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                           const TrackKParams& kp, uint8_t* d_isgood, double* d_out, const KTimer* t) {
  LevelGeom g{f->w[level], f->h[level], f->fx[level], f->fy[level], f->cx[level], f->cy[level]};
  const dim3 grid(eval_blocks(c, pts.n)), block(EVAL_BS);
  const void* img = level_image(f, level);
  auto k = c->sums == LSDG_SUMS_KAHAN ? kern_se3_eval<ACC_KAHAN>(f->precision) : kern_se3_eval<ACC_F64>(f->precision);
  return launch_k(k, grid, block, c->stream, t, pts, img, g, x, kp, d_isgood, c->d_partials, c->d_counter, d_out);
}

hipError_t launch_sim3_eval(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                            const TrackKParams& kp, double* d_out, const KTimer* t) {
  LevelGeom g{f->w[level], f->h[level], f->fx[level], f->fy[level], f->cx[level], f->cy[level]};
  const dim3 grid(eval_blocks(c, pts.n)), block(EVAL_BS);
  const void* img = level_image(f, level);
  auto k = c->sums == LSDG_SUMS_KAHAN ? kern_sim3_eval<ACC_KAHAN>(f->precision) : kern_sim3_eval<ACC_F64>(f->precision);
  return launch_k(k, grid, block, c->stream, t, pts, img, (const float*)f->d_idepth[level],
                  (const float*)f->d_ivar[level], g, x, kp, c->d_partials, c->d_counter, d_out);
}

// Adjoint grid: two points per thread (one, two and four measure within 10%
// for the kernel; fewer blocks leave fewer rows to merge), at most
// NOALLOC_MAX_BLOCKS blocks: lsdg_energy_grad and its allocation-free variant
// cut the points alike, so their sums agree bit for bit.
static inline int adjoint_blocks(int n) {
  static_assert(NOALLOC_MAX_BLOCKS <= 4096, "the context's partial rows hold the adjoint's blocks");
  int nb = (n + EVAL_BS * 2 - 1) / (EVAL_BS * 2);
  if (nb < 1) nb = 1;
  if (nb > NOALLOC_MAX_BLOCKS) nb = NOALLOC_MAX_BLOCKS;
  return nb;
}

// The adjoint kernel, then the merge of its rows; a timer spans both (the
// first launch's start to the second's end).
template <typename... KArgs, typename... Args>
static hipError_t launch_adjoint(void (*k)(KArgs...), void (*merge)(const double*, unsigned, double*), int nb,
                                 double* rows, double* out, hipStream_t s, const KTimer* t, Args... args) {
  if (t && t->a) hipExtLaunchKernelGGL(k, dim3(nb), dim3(EVAL_BS), 0, s, t->a, nullptr, 0, static_cast<KArgs>(args)...);
  else k<<<dim3(nb), dim3(EVAL_BS), 0, s>>>(static_cast<KArgs>(args)...);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (t && t->a) hipExtLaunchKernelGGL(merge, dim3(1), dim3(EVAL_BS), 0, s, nullptr, t->b, 0, (const double*)rows,
                                       (unsigned)nb, out);
  else merge<<<dim3(1), dim3(EVAL_BS), 0, s>>>(rows, (unsigned)nb, out);
  return hipGetLastError();
}

hipError_t launch_energy_grad(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                              const TrackKParams& kp, float* d_grad_d, double* d_out, const KTimer* t) {
  LevelGeom g{f->w[level], f->h[level], f->fx[level], f->fy[level], f->cx[level], f->cy[level]};
  const int nb = adjoint_blocks(pts.n);
  const void* img = level_image(f, level);
  auto k = f->precision == LSDG_FP16 ? k_energy_grad<LSDG_FP16, ACC_F64>
           : f->precision == LSDG_BF16 ? k_energy_grad<LSDG_BF16, ACC_F64>
                                       : k_energy_grad<LSDG_FP32, ACC_F64>;
  return launch_adjoint(k, k_merge_rows<NS_ADJ>, nb, c->d_partials, d_out, c->stream, t, pts, img, g, x, kp, d_grad_d,
                        c->d_partials);
}

hipError_t launch_energy_grad_sim3(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                                   const TrackKParams& kp, float* d_grad_d, double* d_out, const KTimer* t) {
  LevelGeom g{f->w[level], f->h[level], f->fx[level], f->fy[level], f->cx[level], f->cy[level]};
  const int nb = adjoint_blocks(pts.n);
  const void* img = level_image(f, level);
  auto k = f->precision == LSDG_FP16 ? k_energy_grad_sim3<LSDG_FP16, ACC_F64>
           : f->precision == LSDG_BF16 ? k_energy_grad_sim3<LSDG_BF16, ACC_F64>
                                       : k_energy_grad_sim3<LSDG_FP32, ACC_F64>;
  return launch_adjoint(k, k_merge_rows<NS_ADJ_SIM3>, nb, c->d_partials, d_out, c->stream, t, pts, img,
                        (const float*)f->d_idepth[level], (const float*)f->d_ivar[level], g, x, kp, d_grad_d,
                        c->d_partials);
}

// ------------------------------------------------ allocation-free adjoint
// lsdg_energy_grad_noalloc's device memory: one file-scope array per GPU,
// loaded with the code object and held until the process ends.  Layout:
//   [0, PARTIALS)              FP64 partial rows, NOALLOC_MAX_BLOCKS x NS_ADJ
//   [RESULT, +256)             the merged row (NS_ADJ doubles)
//   [GRAD, +4 * max points)    per-point dE/dd staging
// Every byte a call reads is written earlier in the same call (rows by the
// adjoint kernel, the merged row by the merge), so nothing needs clearing.
// Host code uses the address hipGetSymbolAddress returns for the current
// device, never the array's name (which is a host-side stand-in).  Every
// call on a device shares that device's array: calls that overlap on one
// GPU race.
namespace {
constexpr size_t NOALLOC_PARTIALS = (size_t)NOALLOC_MAX_BLOCKS * NS_ADJ * sizeof(double);
constexpr size_t NOALLOC_RESULT = NOALLOC_PARTIALS;
constexpr size_t NOALLOC_GRAD = NOALLOC_RESULT + 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_GRAD + (size_t)NOALLOC_MAX_POINTS * sizeof(float);
static_assert(NOALLOC_PARTIALS % 256 == 0, "sections stay 256-byte aligned");
static_assert(NS_ADJ * sizeof(double) <= 256, "the merged row fits its section");

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = (unsigned char*)ptr;
  return hipSuccess;
}

hipError_t launch_energy_grad_noalloc(lsdg_ctx c, const PointsLevel& pts, lsdg_frame f, int level, const Xf& x,
                                      const TrackKParams& kp, double* h_row, float* h_grad) {
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) base = g_scratch_ptr[c->device];
  }
  if (!base) return hipErrorNotInitialized;
  if (pts.n > NOALLOC_MAX_POINTS) return hipErrorInvalidValue;
  double* partials = (double*)base;
  double* row = (double*)(base + NOALLOC_RESULT);
  float* grad = (float*)(base + NOALLOC_GRAD);
  LevelGeom g{f->w[level], f->h[level], f->fx[level], f->fy[level], f->cx[level], f->cy[level]};
  const int nb = adjoint_blocks(pts.n);
  auto k = f->precision == LSDG_FP16 ? k_energy_grad<LSDG_FP16, ACC_F64>
           : f->precision == LSDG_BF16 ? k_energy_grad<LSDG_BF16, ACC_F64>
                                       : k_energy_grad<LSDG_FP32, ACC_F64>;
  hipError_t e = launch_adjoint(k, k_merge_rows<NS_ADJ>, nb, partials, row, c->stream, nullptr, pts,
                                level_image(f, level), g, x, kp, grad, partials);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_row, row, sizeof(double) * NS_ADJ, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && h_grad)
    e = hipMemcpyAsync(h_grad, grad, sizeof(float) * pts.n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

}  // namespace lsdg
