// Batched tracking: many frame-vs-keyframe alignments of uneven size, with no
// host round trip between LM iterations.  Three schedules share the LM code:
// the cooperative kernel (k_batch_coop, further down: several blocks per job,
// every level in one launch) and the two below.
//
// Whole-job kernel (k_batch_job).  While a job's level is small, one block
// runs the job's LM loop by itself: evaluation over the level's points,
// deterministic block reduction, LM step in lane 0, repeat.  No launch and no
// inter-block communication per iteration.
//
// Multi-block steps (k_batch_eval + k_batch_solve).  For larger levels each
// step is two launches.  Each job's points are cut into tiles of
// TILE = 256 * ppt points; k_batch_eval's persistent grid walks the
// concatenated tile list of all active jobs (the per-job tile counts are
// double-buffered per step so every block sees the same list) and writes each
// tile's FP64 partial sums to a fixed slot.  k_batch_solve then merges each
// job's slots in a fixed order (deterministic) and runs the job's LM step in
// one lane: the reference's accept / reject / lambda / convergence logic
// (SE3Tracker::trackFrame, Sim3Tracker::trackFrameSim3), an FP64 LDL^T solve
// and exp(d) * T, and writes the pose the job's next evaluation uses.  The
// kernel boundary between the two orders the partial sums, so no block waits
// on or fences for another.  A hipGraph replays K such steps.
#include "batch_internal.h"
#include "track_device.h"

namespace lsdg {

constexpr int BATCH_BS = 256;
constexpr int BATCH_WAVES = BATCH_BS / 64;
#ifndef LSDG_BATCH_U
#define LSDG_BATCH_U 2
#endif
constexpr int BATCH_U = LSDG_BATCH_U;  // points per thread with loads in flight together
constexpr int BATCH_SOLVE_BS = 256;

// f^n by repeated multiplication (exact for the default factor 2).
__host__ __device__ inline float pow_int(float f, int n) {
  float r = 1.0f;
  for (int i = 0; i < n; ++i) r *= f;
  return r;
}

// ------------------------------------------------------------ LM steps
__device__ inline void set_eval_pose(JobState& s, const Pose& T, bool sim3) {
  pose_rt_f(T, sim3, s.xf.R, s.xf.t);
  if (sim3) pose_roll_f(T, s.roll);
}

// The LM step runs on the 64 lanes of the block's first wave: its scalar
// control flow is uniform over the wave (every lane computes the same values
// from the same LDS state and writes the same values back), and the
// independent divisions and trigonometric calls of the solve are spread over
// the lanes.  `sc` is LM_SCRATCH doubles of LDS.
constexpr int LM_SCRATCH = 64;

__device__ inline int lm_lane() { return (int)(threadIdx.x & 63); }

// Orders the wave's LDS writes before its later LDS reads.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// remaining may be null: blocks that repeat another block's LM step (the
// cooperative kernel) leave the count to that block.
__device__ inline void finish_job(JobState& s, int status, int* ntNext, int* remaining) {
  s.status = status;
  *ntNext = 0;
  if (remaining && lm_lane() == 0) atomicSub(remaining, 1);
}

// se3_exp (lsdg_lie.h) with sin and cos of theta / 2 (lane 0) and theta
// (lane 1) evaluated at once, one sincos per lane.
__device__ inline Pose se3_exp_wave(const double* __restrict__ xi, double* __restrict__ sc) {
  const double* w = xi + 3;
  const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double th = sqrt(th2);
  if (th >= 1e-10) {
    const int lane = lm_lane();
    if (lane < 2) {
      double sv, cv;
      sincos(lane == 0 ? 0.5 * th : th, &sv, &cv);
      sc[lane] = sv;
      sc[2 + lane] = cv;
    }
    wave_sync();
  }
  Pose r;
  double k, c;
  if (th < 1e-10) {
    k = 0.5 - th2 / 48.0;
    c = 1.0 - th2 / 8.0;
  } else {
    k = sc[0] / th;
    c = sc[2];
  }
  r.q[0] = k * w[0]; r.q[1] = k * w[1]; r.q[2] = k * w[2]; r.q[3] = c;
  quat_normalize(r.q);
  double bb, cc;
  if (th < 1e-5) {
    bb = 0.5 - th2 / 24.0;
    cc = 1.0 / 6.0 - th2 / 120.0;
  } else {
    bb = (1.0 - sc[3]) / th2;
    cc = (th - sc[1]) / (th2 * th);
  }
  apply_abc(1.0, bb, cc, w, xi, r.t);
  r.s = 1.0;
  return r;
}

// Solve (A diag*(1+lambda)) inc = b from the stored LGS and propose Tn = exp(inc) T.
// Returns false when a Sim3 step diverges (|inc|^2 not in [0, 1)).
template <int KIND>
__device__ inline bool solve_and_propose(JobState& __restrict__ s, double* __restrict__ sc) {
  constexpr int N = KIND == LSDG_JOB_SIM3 ? 7 : 6;
  constexpr int NA = N * (N + 1) / 2;
  double A[N * N], b[N], inc[N];
  const double n = s.lgsN;
  // the normalisation's NA + N divisions, one per lane
  {
    const int lane = lm_lane();
    if (lane < NA) sc[lane] = s.lgsA[lane] / n;
    else if (lane < NA + N) sc[lane] = -s.lgsB[lane - NA] / n;
    wave_sync();
  }
  int k = 0;
  for (int a = 0; a < N; ++a)
    for (int c = a; c < N; ++c) {
      double v = sc[k++];
      A[N * a + c] = v;
      A[N * c + a] = v;
    }
  for (int a = 0; a < N; ++a) b[a] = sc[NA + a];
  wave_sync();
  for (int a = 0; a < N; ++a) A[(N + 1) * a] *= 1.0 + (double)s.lambda;
  ldlt_solve<N>(A, b, inc);
  s.incTry++;
  double dd = 0;
  for (int a = 0; a < N; ++a) {
    s.inc[a] = inc[a];
    dd += inc[a] * inc[a];
  }
  s.incSq = dd;
  if (KIND == LSDG_JOB_SIM3) {
    if (!(dd >= 0 && dd < 1)) return false;
    s.Tn = pose_mul(sim3_exp(inc), s.T);
  } else {
    s.Tn = pose_mul(se3_exp_wave(inc, sc), s.T);
  }
  s.phase = 1;
  set_eval_pose(s, s.Tn, KIND == LSDG_JOB_SIM3);
  return true;
}

// Keeps the evaluation's unnormalised LGS (SE3: 6x6 upper triangle from the
// 7-stride sums; Sim3: 7x7) as the accepted pose's normal equations.
template <int KIND>
__device__ inline void store_lgs(JobState& __restrict__ s, const double* __restrict__ S) {
  if (KIND == LSDG_JOB_SIM3) {
    for (int k = 0; k < 28; ++k) s.lgsA[k] = S[SM_A + k];
    for (int a = 0; a < 7; ++a) s.lgsB[a] = S[SM_B + a];
    s.lgsN = 2.0 * S[SM_WARPED];
  } else {
    for (int k = 0; k < 21; ++k) s.lgsA[k] = S[S3_A + k];
    for (int a = 0; a < 6; ++a) s.lgsB[a] = S[S3_B + a];
    s.lgsN = S[S3_WARPED];
  }
}

template <int KIND>
__device__ inline void begin_level(JobState& s, const JobDesc& jd, const lsdg_params& P, int lvl, int* ntNext,
                                   int* remaining) {
  if (KIND == LSDG_JOB_SIM3) {
    while (lvl >= jd.finalLevel && P.max_its_per_lvl[lvl] == 0) lvl--;
  }
  if (lvl < jd.finalLevel) {
    if (KIND == LSDG_JOB_SIM3 && !s.upToDate) {
      // a final evaluation at the accepted pose for the information matrix
      s.level = jd.finalLevel;
      s.phase = 2;
      set_eval_pose(s, s.T, true);
      *ntNext = jd.lv[s.level].ntiles;
      return;
    }
    if (KIND == LSDG_JOB_SIM3 && s.T.s <= 0) {
      finish_job(s, 2, ntNext, remaining);
      return;
    }
    finish_job(s, 1, ntNext, remaining);
    return;
  }
  s.level = lvl;
  s.phase = 0;
  set_eval_pose(s, s.T, KIND == LSDG_JOB_SIM3);
  *ntNext = jd.lv[lvl].ntiles;
}

// One LM step of an SE3 job, after the evaluation whose sums are S.
__device__ void lm_step_se3(JobState& __restrict__ s, const JobDesc& __restrict__ jd, const double* __restrict__ S,
                            const lsdg_params& P, int* ntNext, int* remaining, double* __restrict__ sc) {
  const int lvl = s.level;
  const JobLevel& L = jd.lv[lvl];
  s.evals++;
  s.evalsAt[lvl]++;
  s.good = S[S3_GOOD];
  s.bad = S[S3_BAD];
  s.meanRes = S[S3_SIGNED] / S[S3_GOOD];
  s.usage = S[S3_USAGE] / (double)L.pts.n;
  // the affine iterate is read only when the affine brightness is estimated
  if (P.use_affine_lightning_estimation) {
    const double sxx = S[S3_SXX], syy = S[S3_SYY], sx = S[S3_SX], sy = S[S3_SY], sw = S[S3_SW];
    s.affAIt = (float)sqrt((syy - sy * sy / sw) / (sxx - sx * sx / sw));
    s.affBIt = (float)((sy - (double)s.affAIt * sx) / sw);
  }
  const float lim = P.min_goodperall_pixel_absmin * (float)(L.g.w * L.g.h);
  if ((float)S[S3_WARPED] < lim) {
    finish_job(s, 2, ntNext, remaining);
    return;
  }
  const double err = S[S3_ERR] / S[S3_WARPED];
  s.numRes[lvl]++;
  if (s.phase == 0) {
    if (P.use_affine_lightning_estimation) { s.affA = s.affAIt; s.affB = s.affBIt; }
    s.lastErr = err;
    s.lambda = P.lambda_initial[lvl];
    s.iteration = 0;
    if (P.max_its_per_lvl[lvl] <= 0) {
      begin_level<LSDG_JOB_SE3>(s, jd, P, lvl - 1, ntNext, remaining);
      return;
    }
    store_lgs<LSDG_JOB_SE3>(s, S);
    s.numWarp[lvl]++;
    s.incTry = 0;
    solve_and_propose<LSDG_JOB_SE3>(s, sc);
    *ntNext = L.ntiles;
    return;
  }
  if (err < s.lastErr) {
    s.T = s.Tn;
    if (P.use_affine_lightning_estimation) { s.affA = s.affAIt; s.affB = s.affBIt; }
    bool conv = err / s.lastErr > (double)P.convergence_eps[lvl];
    s.lastResidual = s.lastErr = err;
    s.lambda = s.lambda <= 0.2f ? 0.0f : s.lambda * P.lambda_success_fac[lvl];
    s.iteration++;
    if (conv || s.iteration >= P.max_its_per_lvl[lvl]) {
      begin_level<LSDG_JOB_SE3>(s, jd, P, lvl - 1, ntNext, remaining);
      return;
    }
    store_lgs<LSDG_JOB_SE3>(s, S);
    s.numWarp[lvl]++;
    s.incTry = 0;
    solve_and_propose<LSDG_JOB_SE3>(s, sc);
    *ntNext = L.ntiles;
    return;
  }
  if (!(s.incSq > (double)P.step_size_min[lvl])) {
    begin_level<LSDG_JOB_SE3>(s, jd, P, lvl - 1, ntNext, remaining);
    return;
  }
  s.lambda = s.lambda == 0.0f ? 0.2f : s.lambda * pow_int(P.lambda_fail_fac[lvl], s.incTry);
  solve_and_propose<LSDG_JOB_SE3>(s, sc);
  *ntNext = L.ntiles;
}

__device__ inline void sim3_res(const double* S, double* r) {
  // r = {sumResD, sumResP, numTermsD, numTermsP, meanD, meanP, mean}
  r[0] = S[SM_RESD];
  r[1] = S[SM_RESP];
  r[2] = S[SM_NTD];
  r[3] = S[SM_WARPED];
  r[6] = (S[SM_RESD] + S[SM_RESP]) / (S[SM_NTD] + S[SM_WARPED]);
  r[4] = S[SM_RESD] / S[SM_NTD];
  r[5] = S[SM_RESP] / S[SM_WARPED];
}

// One LM step of a Sim3 job.
__device__ void lm_step_sim3(JobState& __restrict__ s, const JobDesc& __restrict__ jd, const double* __restrict__ S,
                             const lsdg_params& P, int* ntNext, int* remaining, double* __restrict__ sc) {
  const int lvl = s.level;
  const JobLevel& L = jd.lv[lvl];
  s.evals++;
  s.evalsAt[lvl]++;
  s.usage = S[SM_USAGE] / (double)L.pts.n;
  if (s.phase == 2) {
    // final evaluation for lastSim3Hessian
    sim3_res(S, s.finalRes);
    store_lgs<LSDG_JOB_SIM3>(s, S);
    s.upToDate = 1;
    finish_job(s, s.T.s <= 0 ? 2 : 1, ntNext, remaining);
    return;
  }
  const float lim = 0.5f * P.min_goodperall_pixel_absmin * (float)(L.g.w * L.g.h);
  if ((float)S[SM_WARPED] < lim || S[SM_WARPED] < 10) {
    finish_job(s, 2, ntNext, remaining);
    return;
  }
  double res[7];
  sim3_res(S, res);
  s.numRes[lvl]++;
  if (s.phase == 0) {
    for (int k = 0; k < 7; ++k) s.lastRes[k] = res[k];
    s.lambda = P.lambda_initial[lvl];
    s.upToDate = 0;
    s.iteration = 0;
    store_lgs<LSDG_JOB_SIM3>(s, S);
    s.upToDate = 1;
    s.numWarp[lvl]++;
    s.incTry = 0;
    if (!solve_and_propose<LSDG_JOB_SIM3>(s, sc)) {
      for (int k = 0; k < 28; ++k) s.lgsA[k] = 0.0;
      finish_job(s, 2, ntNext, remaining);
      return;
    }
    *ntNext = L.ntiles;
    return;
  }
  if (res[6] < s.lastRes[6]) {
    s.T = s.Tn;
    s.upToDate = 0;
    bool conv = res[6] / s.lastRes[6] > (double)P.convergence_eps[lvl];
    for (int k = 0; k < 7; ++k) s.finalRes[k] = s.lastRes[k] = res[k];
    s.lambda = s.lambda <= 0.2f ? 0.0f : s.lambda * P.lambda_success_fac[lvl];
    s.iteration++;
    if (conv || s.iteration >= P.max_its_per_lvl[lvl]) {
      begin_level<LSDG_JOB_SIM3>(s, jd, P, lvl - 1, ntNext, remaining);
      return;
    }
    store_lgs<LSDG_JOB_SIM3>(s, S);
    s.upToDate = 1;
    s.numWarp[lvl]++;
    s.incTry = 0;
  } else {
    if (!(s.incSq > (double)P.step_size_min[lvl])) {
      begin_level<LSDG_JOB_SIM3>(s, jd, P, lvl - 1, ntNext, remaining);
      return;
    }
    s.lambda = s.lambda == 0.0f ? 0.2f : s.lambda * pow_int(P.lambda_fail_fac[lvl], s.incTry);
  }
  if (!solve_and_propose<LSDG_JOB_SIM3>(s, sc)) {
    for (int k = 0; k < 28; ++k) s.lgsA[k] = 0.0;
    finish_job(s, 2, ntNext, remaining);
    return;
  }
  *ntNext = L.ntiles;
}

// ------------------------------------------------------------- kernels
// Start state of every job: T = (frameToRef_init)^-1, first level, counters
// zero.  One block, so the remaining-jobs counter can be reset first.
__global__ __launch_bounds__(256) void k_batch_init(BatchArgs a) {
  __shared__ bool clearRows;
  if (threadIdx.x == 0) {
    *a.remaining = 0;
    // the run's epoch tags the cooperative row exchange; its low 16 bits
    // never take the value 0, which the rows are cleared to on the first run
    // and whenever the low bits wrap, so no word left by an earlier run can
    // carry a tag of this one
    unsigned e = *a.epoch + 1;
    if ((e & 0xFFFFu) == 0) ++e;
    *a.epoch = e;
    clearRows = (e & 0xFFFFu) == 1;
  }
  __syncthreads();
  if (clearRows && a.coopRows)
    for (size_t i = threadIdx.x; i < a.coopWords; i += blockDim.x) a.coopRows[i] = 0ull;
  for (int j = threadIdx.x; j < a.njobs; j += blockDim.x) {
    const JobDesc& jd = a.jobs[j];
    JobState s = {};
    s.T = pose_inv(jd.init);
    if (jd.kind == LSDG_JOB_SE3) s.T.s = 1.0;
    s.affA = 1.0f;
    s.affB = 0.0f;
    int nt = 0, dummy = 0;
    if (jd.kind == LSDG_JOB_SIM3) begin_level<LSDG_JOB_SIM3>(s, jd, a.P, jd.startLevel, &nt, &dummy);
    else begin_level<LSDG_JOB_SE3>(s, jd, a.P, jd.startLevel, &nt, &dummy);
    a.st[j] = s;
    a.ntiles[j] = nt;  // parity 0
    a.ntiles[a.njobs + j] = 0;
    a.counters[j] = 0u;
    if (nt > 0) atomicAdd(a.remaining, 1);
  }
}

// Tile evaluation of one LM step: a persistent grid walks the concatenated
// tile list of this step's jobs; each tile's FP64 partial row goes to a fixed
// slot.  No cross-block communication: the merge runs in k_batch_solve.
template <int KIND, int PREC, int ACC>
__global__ __launch_bounds__(BATCH_BS) void k_batch_eval(BatchArgs a, int parity) {
  constexpr int NS = KIND == LSDG_JOB_SIM3 ? NS_SIM3 : NS_SE3;
  __shared__ int pref[MAX_BATCH_JOBS + 1];
  __shared__ int scan_tmp[BATCH_BS];
  __shared__ double sh[BATCH_WAVES][NS];
  const int tid = threadIdx.x;
  const int nj = a.njobs;
  const int* nt = a.ntiles + parity * nj;
  // exclusive prefix of this step's tile counts (chunked block scan)
  {
    const int per = (nj + BATCH_BS - 1) / BATCH_BS;
    int base = tid * per, local = 0;
    for (int k = 0; k < per; ++k)
      if (base + k < nj) local += nt[base + k];
    scan_tmp[tid] = local;
    __syncthreads();
    for (int off = 1; off < BATCH_BS; off <<= 1) {
      int v = tid >= off ? scan_tmp[tid - off] : 0;
      __syncthreads();
      scan_tmp[tid] += v;
      __syncthreads();
    }
    int run = scan_tmp[tid] - local;
    for (int k = 0; k < per; ++k)
      if (base + k < nj) {
        pref[base + k] = run;
        run += nt[base + k];
      }
    if (tid == BATCH_BS - 1) pref[nj] = scan_tmp[BATCH_BS - 1];
    __syncthreads();
  }
  const int total = pref[nj];
  const int lane = tid & 63, wave = tid >> 6;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (pref[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int j = lo;
    const int tile = t - pref[j];
    const JobDesc& jd = a.jobs[j];
    const JobState* st = a.st + j;
    const int lvl = st->level;
    const JobLevel& L = jd.lv[lvl];
    const Xf x = st->xf;
    TrackKParams kp = a.kp;
    kp.aff_a = st->affA;
    kp.aff_b = st->affB;
    Accum<ACC, NS> acc;
    acc.zero();
    const int tileSize = BATCH_BS * a.ppt;
    const int i0 = tile * tileSize;
    const int i1 = min(i0 + tileSize, L.pts.n);
    if (KIND == LSDG_JOB_SIM3) {
      kp.roll[0] = st->roll[0]; kp.roll[1] = st->roll[1]; kp.roll[2] = st->roll[2]; kp.roll[3] = st->roll[3];
      sim3_points<PREC, BATCH_U>(L.pts, i0 + tid, BATCH_BS, i1, L.img, L.fid, L.fvar, L.g, x, kp, acc);
    } else {
      uint8_t* mask = lvl == 1 ? jd.isGood : nullptr;
      se3_points<PREC, BATCH_U>(L.pts, i0 + tid, BATCH_BS, i1, L.img, L.g, x, kp, mask, acc);
    }
    const double wsum = acc.wave_slot();
    if (lane < NS) sh[wave][lane] = wsum;
    __syncthreads();
    if (tid < NS)
      a.partials[((size_t)jd.slotBase + tile) * NS + tid] = Accum<ACC, NS>::merge_waves(&sh[0][tid], NS, BATCH_WAVES);
    __syncthreads();
  }
}

// Copies a trivially copyable object between global memory and LDS with the
// whole block (32-bit words).
template <class T>
__device__ inline void block_copy(T* dst, const T* src, int nthreads) {
  static_assert(sizeof(T) % 4 == 0, "copied as 32-bit words");
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (int w = threadIdx.x; w < (int)(sizeof(T) / 4); w += nthreads) d[w] = s[w];
}

// Merge and LM step of one LM iteration: one block per job.  Wave w sums the
// job's tile rows w, w + 4, ... (row order within a wave), the four wave sums
// are added in wave order, and lane 0 runs the LM step on the job's
// descriptor and control block staged in LDS (their loads are issued
// alongside the row loads).  Also zeroes the next step's tile count of jobs
// that are already finished.
template <int KIND>
__global__ __launch_bounds__(BATCH_SOLVE_BS) void k_batch_solve(BatchArgs a, int parity) {
  constexpr int NS = KIND == LSDG_JOB_SIM3 ? NS_SIM3 : NS_SE3;
  constexpr int WAVES = BATCH_SOLVE_BS / 64;
  __shared__ double part[WAVES][NS];
  __shared__ double S[NS];
  __shared__ JobState sst;
  __shared__ JobDesc sjd;
  __shared__ int next;
  __shared__ double lmsc[LM_SCRATCH];
  const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nj = a.njobs;
  const int n = a.ntiles[parity * nj + j];
  const int slotBase = a.jobs[j].slotBase;  // loaded alongside n
  int* ntNext = a.ntiles + (1 - parity) * nj + j;
  if (n == 0) {
    if (tid == 0) *ntNext = 0;
    return;
  }
  block_copy(&sst, a.st + j, BATCH_SOLVE_BS);
  block_copy(&sjd, a.jobs + j, BATCH_SOLVE_BS);
  if (lane < NS) {
    const double* rows = a.partials + (size_t)slotBase * NS + lane;
    double v = 0.0;
    int k = wave;
    for (; k + 7 * WAVES < n; k += 8 * WAVES) {
      double r[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) r[u] = rows[(size_t)(k + u * WAVES) * NS];
#pragma unroll
      for (int u = 0; u < 8; ++u) v += r[u];
    }
    for (; k < n; k += WAVES) v += rows[(size_t)k * NS];
    part[wave][lane] = v;
  }
  __syncthreads();
  if (tid < NS) {
    double v = part[0][tid];
    for (int w = 1; w < WAVES; ++w) v += part[w][tid];
    S[tid] = v;
  }
  __syncthreads();
  if (tid < 64) {
    int nx = 0;
    if (KIND == LSDG_JOB_SIM3) lm_step_sim3(sst, sjd, S, a.P, &nx, a.remaining, lmsc);
    else lm_step_se3(sst, sjd, S, a.P, &nx, a.remaining, lmsc);
    if (tid == 0) next = nx;
  }
  __syncthreads();
  block_copy(a.st + j, &sst, BATCH_SOLVE_BS);
  if (tid == 0) *ntNext = next;
}


// Whole LM of one job in one block, for as long as the job's current level
// has at most a.jobMaxPoints points: the level's evaluation over the block,
// the block's deterministic reduction (wave reduce-scatter, then waves in
// order) and the LM step in lane 0, repeated with no launch and no
// inter-block communication in between.  The job's descriptor and control
// block live in LDS meanwhile.  On exit the job is finished or stands at a
// larger level, whose tile count goes to the parity-0 slot for the
// multi-block steps.
template <int KIND, int PREC, int ACC>
__global__ __launch_bounds__(BATCH_JOB_BS) void k_batch_job(BatchArgs a) {
  constexpr int NS = KIND == LSDG_JOB_SIM3 ? NS_SIM3 : NS_SE3;
  constexpr int WAVES = BATCH_JOB_BS / 64;
  __shared__ JobState s;
  __shared__ JobDesc jd;
  __shared__ double sh[WAVES][NS];
  __shared__ double S[NS];
  __shared__ double lmsc[LM_SCRATCH];
  const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  block_copy(&s, a.st + j, BATCH_JOB_BS);
  block_copy(&jd, a.jobs + j, BATCH_JOB_BS);
  __syncthreads();
  for (int it = 0; it < BATCH_JOB_MAX_EVALS; ++it) {
    // s and jd are read after a barrier: the exit test is uniform over the block
    if (s.status != 0) break;
    const int lvl = s.level;
    const JobLevel& L = jd.lv[lvl];
    if (L.pts.n > a.jobMaxPoints) break;
    const Xf x = s.xf;
    TrackKParams kp = a.kp;
    kp.aff_a = s.affA;
    kp.aff_b = s.affB;
    Accum<ACC, NS> acc;
    acc.zero();
    const int n = L.pts.n;
    if (KIND == LSDG_JOB_SIM3) {
      kp.roll[0] = s.roll[0]; kp.roll[1] = s.roll[1]; kp.roll[2] = s.roll[2]; kp.roll[3] = s.roll[3];
      sim3_points<PREC, BATCH_U>(L.pts, tid, BATCH_JOB_BS, n, L.img, L.fid, L.fvar, L.g, x, kp, acc);
    } else {
      uint8_t* mask = lvl == 1 ? jd.isGood : nullptr;
      se3_points<PREC, BATCH_U>(L.pts, tid, BATCH_JOB_BS, n, L.img, L.g, x, kp, mask, acc);
    }
    const double wsum = acc.wave_slot();
    if (lane < NS) sh[wave][lane] = wsum;
    __syncthreads();
    if (tid < NS) S[tid] = Accum<ACC, NS>::merge_waves(&sh[0][tid], NS, WAVES);
    __syncthreads();
    if (tid < 64) {
      int next = 0;
      if (KIND == LSDG_JOB_SIM3) lm_step_sim3(s, jd, S, a.P, &next, a.remaining, lmsc);
      else lm_step_se3(s, jd, S, a.P, &next, a.remaining, lmsc);
    }
    __syncthreads();
  }
  block_copy(a.st + j, &s, BATCH_JOB_BS);
  if (tid == 0) a.ntiles[j] = s.status == 0 ? jd.lv[s.level].ntiles : 0;
}

// Cooperative whole-LM kernel: K blocks per job (blockIdx.x = job * K + part)
// run the job's whole LM loop together, every level included.  A level of at
// most soloMax points is evaluated whole by every block of the job.  A larger
// level is cut into K contiguous point ranges: each block reduces its range
// to an FP64 row, publishes it as tagged words (device-scope stores), reads
// the K rows until every word carries this iteration's tag, and sums them in
// part order.  In both cases every block
// of the job then holds the same sums and runs the same LM step on its own
// LDS copy of the control block, so no result is broadcast; part 0 alone
// counts finished jobs and writes the control block back.  Rows are double
// buffered by iteration parity: a block can be at most one iteration ahead of
// the slowest block of its job.  The wait is bounded (a.spinMax polls): if the
// K blocks of a job cannot all be resident (the grid is sized from the
// occupancy, so this does not happen on an idle device), the blocks stop and
// set *remaining negative, which the host reports as an error.  The block
// size BS (256 or 512 threads) is one of the shapes the plan's tuner picks.
#ifndef LSDG_SPIN_SLEEP
#define LSDG_SPIN_SLEEP 2
#endif

template <int KIND, int PREC, int ACC, int COOP_BS>
__global__ __launch_bounds__(COOP_BS) void k_batch_coop(BatchArgs a) {
  constexpr int NS = KIND == LSDG_JOB_SIM3 ? NS_SIM3 : NS_SE3;
  constexpr int WAVES = COOP_BS / 64;
  __shared__ JobState s;
  __shared__ JobDesc jd;
  __shared__ double sh[WAVES][NS];
  __shared__ double S[NS];
  __shared__ double lmsc[LM_SCRATCH];
  __shared__ int arrived;
  const int K = a.coopK;
  // Blocks are dispatched round-robin over the device's XCDs.  With
  // a.xcds > 1, block b belongs to group b % xcds, and the K blocks of job j
  // all fall in group j % xcds: a job's points and taps stay in one XCD's L2.
  int j, part;
  if (a.xcds > 1) {
    const int slot = blockIdx.x / a.xcds;
    j = (slot / K) * a.xcds + (int)(blockIdx.x % a.xcds);
    part = slot % K;
  } else {
    j = blockIdx.x / K;
    part = blockIdx.x % K;
  }
  if (j >= a.njobs) return;  // padding of the last group (the whole block leaves)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int* remaining = part == 0 ? a.remaining : nullptr;
  unsigned long long* rows = a.coopRows + (size_t)j * 2 * K * NS * 2;
  const unsigned epochTag = (*a.epoch & 0xFFFFu) << 16;
  block_copy(&s, a.st + j, COOP_BS);
  block_copy(&jd, a.jobs + j, COOP_BS);
  __syncthreads();
  unsigned gen = 0;  // synchronised iterations of this job so far
  for (int it = 0; it < BATCH_JOB_MAX_EVALS; ++it) {
    // s is read after a barrier: the exit test is uniform over the block
    if (s.status != 0) break;
    const int lvl = s.level;
    const JobLevel& L = jd.lv[lvl];
    const int n = L.pts.n;
    const bool split = K > 1 && n > a.soloMax;
    const int i0 = split ? (int)((long long)n * part / K) : 0;
    const int i1 = split ? (int)((long long)n * (part + 1) / K) : n;
    const Xf x = s.xf;
    TrackKParams kp = a.kp;
    kp.aff_a = s.affA;
    kp.aff_b = s.affB;
    Accum<ACC, NS> acc;
    acc.zero();
    if (KIND == LSDG_JOB_SIM3) {
      kp.roll[0] = s.roll[0]; kp.roll[1] = s.roll[1]; kp.roll[2] = s.roll[2]; kp.roll[3] = s.roll[3];
      sim3_points<PREC, BATCH_U>(L.pts, i0 + tid, COOP_BS, i1, L.img, L.fid, L.fvar, L.g, x, kp, acc);
    } else {
      uint8_t* mask = lvl == 1 ? jd.isGood : nullptr;
      se3_points<PREC, BATCH_U>(L.pts, i0 + tid, COOP_BS, i1, L.img, L.g, x, kp, mask, acc);
    }
    const double wsum = acc.wave_slot();
    if (lane < NS) sh[wave][lane] = wsum;
    __syncthreads();
    if (!split) {
      if (tid < NS) S[tid] = Accum<ACC, NS>::merge_waves(&sh[0][tid], NS, WAVES);
    } else {
      // Tagged row exchange: each value is published as two 64-bit words
      // (one 32-bit half of the FP64 value, one 32-bit tag: the run's epoch
      // and this iteration's generation), so every word tells by itself
      // whether it belongs to this iteration.  Wave 0 stores its part's row
      // and then reads all K rows until every word carries the tag: one
      // round of loads after the last part's stores land, with no arrival
      // counter and no store-completion wait in between.
      unsigned long long* buf = rows + (size_t)(gen & 1) * K * NS * 2;
      const unsigned tag = epochTag | (gen & 0xFFFFu);
      if (tid < 64) {
        if (tid < NS) {
          const double v = Accum<ACC, NS>::merge_waves(&sh[0][tid], NS, WAVES);
          const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
          unsigned long long* w = buf + ((size_t)part * NS + tid) * 2;
          const unsigned long long t64 = (unsigned long long)tag << 32;
          __hip_atomic_store(w, t64 | (bits & 0xFFFFFFFFull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(w + 1, t64 | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // at most spinMax polls after the first look; the sum is taken in
        // part order once every word of every row carries this tag
        bool ok = false;
        double v = 0.0;
        for (unsigned spins = 0;; ++spins) {
          bool mine = true;
          double t = 0.0;
          if (tid < NS) {
            for (int q = 0; q < K; ++q) {
              const unsigned long long* w = buf + ((size_t)q * NS + tid) * 2;
              const unsigned long long lo = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              const unsigned long long hi = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              mine = mine && (unsigned)(lo >> 32) == tag && (unsigned)(hi >> 32) == tag;
              t += __longlong_as_double((long long)((hi << 32) | (lo & 0xFFFFFFFFull)));
            }
          }
          if (__all(mine)) {
            ok = true;
            v = t;
            break;
          }
          if (spins >= a.spinMax) break;
          __builtin_amdgcn_s_sleep(LSDG_SPIN_SLEEP);
        }
        if (tid < NS) S[tid] = v;
        if (tid == 0) arrived = ok;
      }
      __syncthreads();
      if (!arrived) {
        if (tid == 0) __hip_atomic_store(a.remaining, -(1 << 20), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      ++gen;
    }
    __syncthreads();
    if (tid < 64) {
      int next = 0;
      if (KIND == LSDG_JOB_SIM3) lm_step_sim3(s, jd, S, a.P, &next, remaining, lmsc);
      else lm_step_se3(s, jd, S, a.P, &next, remaining, lmsc);
    }
    __syncthreads();
  }
  if (part == 0) {
    block_copy(a.st + j, &s, COOP_BS);
    if (tid == 0) a.ntiles[j] = s.status == 0 ? jd.lv[s.level].ntiles : 0;
  }
}

// ------------------------------------------------------------ launchers
template <int KIND, int PREC>
static void (*job_kernel(int acc))(BatchArgs) {
  return acc == BATCH_ACC_KAHAN ? k_batch_job<KIND, PREC, ACC_KAHAN> : k_batch_job<KIND, PREC, ACC_F64>;
}

hipError_t launch_batch_job(const BatchArgs& a, int kind, int prec, int acc, hipStream_t s, const KTimer* t) {
  auto k = kind == LSDG_JOB_SIM3 ? (prec == LSDG_FP16   ? job_kernel<LSDG_JOB_SIM3, LSDG_FP16>(acc)
                                    : prec == LSDG_BF16 ? job_kernel<LSDG_JOB_SIM3, LSDG_BF16>(acc)
                                                        : job_kernel<LSDG_JOB_SIM3, LSDG_FP32>(acc))
                                 : (prec == LSDG_FP16   ? job_kernel<LSDG_JOB_SE3, LSDG_FP16>(acc)
                                    : prec == LSDG_BF16 ? job_kernel<LSDG_JOB_SE3, LSDG_BF16>(acc)
                                                        : job_kernel<LSDG_JOB_SE3, LSDG_FP32>(acc));
  return launch_k(k, dim3(a.njobs), dim3(BATCH_JOB_BS), s, t, a);
}

template <int ACC, int BS>
static void (*coop_kernel_bs(int kind, int prec))(BatchArgs) {
  if (kind == LSDG_JOB_SIM3)
    return prec == LSDG_FP16 ? k_batch_coop<LSDG_JOB_SIM3, LSDG_FP16, ACC, BS>
           : prec == LSDG_BF16 ? k_batch_coop<LSDG_JOB_SIM3, LSDG_BF16, ACC, BS>
                               : k_batch_coop<LSDG_JOB_SIM3, LSDG_FP32, ACC, BS>;
  return prec == LSDG_FP16 ? k_batch_coop<LSDG_JOB_SE3, LSDG_FP16, ACC, BS>
         : prec == LSDG_BF16 ? k_batch_coop<LSDG_JOB_SE3, LSDG_BF16, ACC, BS>
                             : k_batch_coop<LSDG_JOB_SE3, LSDG_FP32, ACC, BS>;
}

static void (*coop_kernel(int kind, int prec, int acc, int bs))(BatchArgs) {
  if (acc == BATCH_ACC_KAHAN)
    return bs == 256 ? coop_kernel_bs<ACC_KAHAN, 256>(kind, prec) : coop_kernel_bs<ACC_KAHAN, 512>(kind, prec);
  return bs == 256 ? coop_kernel_bs<ACC_F64, 256>(kind, prec) : coop_kernel_bs<ACC_F64, 512>(kind, prec);
}

bool batch_coop_block_size_ok(int bs) { return bs == 256 || bs == 512; }

int batch_coop_grid(int njobs, int K, int xcds) {
  if (xcds <= 1) return njobs * K;
  return ((njobs + xcds - 1) / xcds) * xcds * K;
}

hipError_t batch_coop_capacity(int kind, int prec, int acc, int bs, int* blocks) {
  *blocks = 0;
  if (!batch_coop_block_size_ok(bs)) return hipErrorInvalidValue;
  int dev = 0, perCU = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, coop_kernel(kind, prec, acc, bs), bs, 0);
  *blocks = e == hipSuccess ? perCU * cus : 0;
  return e;
}

hipError_t launch_batch_coop(const BatchArgs& a, int kind, int prec, int acc, int bs, hipStream_t s,
                             const KTimer* t) {
  if (!batch_coop_block_size_ok(bs)) return hipErrorInvalidValue;
  return launch_k(coop_kernel(kind, prec, acc, bs), dim3(batch_coop_grid(a.njobs, a.coopK, a.xcds)), dim3(bs), s,
                  t, a);
}

hipError_t launch_batch_init(const BatchArgs& a, hipStream_t s) {
  k_batch_init<<<dim3(1), dim3(256), 0, s>>>(a);
  return hipGetLastError();
}

template <int KIND, int PREC>
static void (*eval_kernel(int acc))(BatchArgs, int) {
  return acc == BATCH_ACC_KAHAN ? k_batch_eval<KIND, PREC, ACC_KAHAN> : k_batch_eval<KIND, PREC, ACC_F64>;
}

hipError_t launch_batch_eval(const BatchArgs& a, int kind, int prec, int acc, int parity, int grid, hipStream_t s,
                             const KTimer* t) {
  auto k = kind == LSDG_JOB_SIM3 ? (prec == LSDG_FP16   ? eval_kernel<LSDG_JOB_SIM3, LSDG_FP16>(acc)
                                    : prec == LSDG_BF16 ? eval_kernel<LSDG_JOB_SIM3, LSDG_BF16>(acc)
                                                        : eval_kernel<LSDG_JOB_SIM3, LSDG_FP32>(acc))
                                 : (prec == LSDG_FP16   ? eval_kernel<LSDG_JOB_SE3, LSDG_FP16>(acc)
                                    : prec == LSDG_BF16 ? eval_kernel<LSDG_JOB_SE3, LSDG_BF16>(acc)
                                                        : eval_kernel<LSDG_JOB_SE3, LSDG_FP32>(acc));
  return launch_k(k, dim3(grid), dim3(BATCH_BS), s, t, a, parity);
}

hipError_t launch_batch_solve(const BatchArgs& a, int kind, int parity, hipStream_t s, const KTimer* t) {
  auto k = kind == LSDG_JOB_SIM3 ? k_batch_solve<LSDG_JOB_SIM3> : k_batch_solve<LSDG_JOB_SE3>;
  return launch_k(k, dim3(a.njobs), dim3(BATCH_SOLVE_BS), s, t, a, parity);
}

hipError_t launch_batch_step(const BatchArgs& a, int kind, int prec, int acc, int parity, int grid, hipStream_t s) {
  hipError_t e = launch_batch_eval(a, kind, prec, acc, parity, grid, s);
  if (e != hipSuccess) return e;
  return launch_batch_solve(a, kind, parity, s);
}

}  // namespace lsdg
