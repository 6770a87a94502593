// C-ABI of batched tracking: lsdg_batch_create / run / destroy / track.
//
// A plan holds the device job table, one LM control block per job, the
// per-tile partial-sum rows and a hipGraph of
//     [batched pyramid rebuild] -> init -> k_batch_coop (every level, coopK
//     blocks per job) -> copy of the remaining-jobs counter to host memory
// when the device holds at least two blocks per job at once, else
//     [batched pyramid rebuild] -> init -> k_batch_job (levels with at most
//     jobMaxPoints points, one block per job) -> K x (k_batch_eval,
//     k_batch_solve) for larger levels -> copy of the remaining-jobs counter
//     to host memory.
// lsdg_batch_run replays it (plus K-step continuation graphs while jobs are
// still running), sizes K to what the workload needs, and copies the control
// blocks back into lsdg_job_result.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "batch_internal.h"
#include "capi_common.h"

using namespace lsdg;

struct lsdg_batch_s {
  lsdg_ctx ctx = nullptr;
  int n = 0;
  int kind = LSDG_JOB_SE3;
  int prec = LSDG_FP32;
  int acc = BATCH_ACC_F64;
  int ppt = 4;
  int grid = 1024;
  // Independent chains: the jobs are split into `chains` groups whose LM
  // steps run as parallel graph branches, so one group's latency-bound solve
  // overlaps another group's evaluation.
  int chains = 1;
  std::vector<hipStream_t> side;  // streams of chains 1..chains-1 during capture
  std::vector<hipEvent_t> forkJoin;
  int steps = 16;  // multi-block LM steps in the first graph (even, may be 0)
  int contSteps = 16;  // LM steps per continuation graph (even, > 0)
  // Levels of up to this many points run in k_batch_job (12 points per
  // thread of its 512-thread block); larger ones in the multi-block steps.
  int jobMaxPoints = 12 * BATCH_JOB_BS;
  // Cooperative schedule: coopK blocks per job run every level's LM loop in
  // one launch (0: the whole-job kernel and multi-block steps instead).
  int coopK = 0;
  int coopBS = 512;  // threads per cooperative block (256 or 512)
  int xcds = 1;      // dispatch groups a job's cooperative blocks are kept in (1: job order)
  int soloMax = 2048;
  unsigned spinMax = COOP_SPIN_MAX;
  unsigned long long* d_coop = nullptr;  // tagged row-exchange words of the cooperative kernel
  unsigned* d_epoch = nullptr;           // per-chain run counters
  std::vector<lsdg_job> jobs;
  std::vector<int> depthVersion;
  std::vector<JobDesc> h_jobs;
  JobDesc* d_jobs = nullptr;
  JobState* d_st = nullptr;
  int* d_ntiles = nullptr;
  unsigned* d_counters = nullptr;
  int* d_remaining = nullptr;
  int* h_remaining = nullptr;  // pinned
  double* d_partials = nullptr;
  size_t rows = 0;
  // every distinct frame of the plan and the pyramid data its jobs track on
  // (fd_bit flags), re-established at each run: a frame may have been given
  // new content (lsdg_frame_ingest) since the plan was built
  std::vector<lsdg_frame> frames;
  std::vector<unsigned> needs;
  std::vector<char> rebuilt;  // frames[k] is rebuilt by the graph itself
  // pyramid rebuild
  std::vector<lsdg_frame> pyrFrames;
  std::vector<unsigned> pyrNeed;  // data the rebuild produces per frame (fd_bit flags)
  void* d_pyr = nullptr;
  void* h_pyr = nullptr;
  PyrBatch pyr{};
  hipGraphExec_t first = nullptr, cont = nullptr;
  hipGraph_t g_first = nullptr, g_cont = nullptr;
  JobState* h_st = nullptr;  // pinned: every graph ends by copying the control blocks here
};

static void batch_free(lsdg_batch b) {
  if (b->first) (void)hipGraphExecDestroy(b->first);
  if (b->cont) (void)hipGraphExecDestroy(b->cont);
  if (b->g_first) (void)hipGraphDestroy(b->g_first);
  if (b->g_cont) (void)hipGraphDestroy(b->g_cont);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->d_st) (void)hipFree(b->d_st);
  if (b->d_ntiles) (void)hipFree(b->d_ntiles);
  if (b->d_counters) (void)hipFree(b->d_counters);
  for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : b->forkJoin) (void)hipEventDestroy(ev);
  if (b->d_remaining) (void)hipFree(b->d_remaining);
  if (b->h_remaining) (void)hipHostFree(b->h_remaining);
  if (b->h_st) (void)hipHostFree(b->h_st);
  if (b->d_partials) (void)hipFree(b->d_partials);
  if (b->d_coop) (void)hipFree(b->d_coop);
  if (b->d_epoch) (void)hipFree(b->d_epoch);
  if (b->d_pyr) (void)hipFree(b->d_pyr);
  if (b->h_pyr) (void)hipHostFree(b->h_pyr);
  delete b;
}

// Per-job device description: point clouds of every tracked level (built if
// needed), frame images, geometry, tile counts and partial-row slots.
static lsdg_status build_job_table(lsdg_batch b) {
  lsdg_ctx c = b->ctx;
  const int tile = 256 * b->ppt;
  b->h_jobs.assign(b->n, JobDesc{});
  size_t slot = 0;
  for (int j = 0; j < b->n; ++j) {
    const lsdg_job& J = b->jobs[j];
    lsdg_ref r = J.ref;
    lsdg_frame f = J.frame;
    JobDesc& d = b->h_jobs[j];
    d.kind = J.kind;
    d.startLevel = J.start_level;
    d.finalLevel = J.final_level;
    d.init = J.kind == LSDG_JOB_SIM3 ? pose_from8(J.frame_to_ref_init) : pose_from7(J.frame_to_ref_init);
    d.isGood = J.kind == LSDG_JOB_SE3 ? f->d_wasGood : nullptr;
    int maxTiles = 1;
    for (int l = 0; l < PYR; ++l) {
      JobLevel& L = d.lv[l];
      L.g = LevelGeom{f->w[l], f->h[l], f->fx[l], f->fy[l], f->cx[l], f->cy[l]};
      L.img = f->precision == LSDG_FP32 ? (const void*)f->d_grad[l] : (const void*)f->d_gradh[l];
      L.fid = f->d_idepth[l];
      L.fvar = f->d_ivar[l];
      if (l < d.finalLevel || l > d.startLevel) {
        L.pts = PointsLevel{nullptr, nullptr, nullptr, nullptr, 0};
        L.ntiles = 1;
        continue;
      }
      lsdg_status st = ref_points(r, l);
      if (st != LSDG_OK) return st;
      L.pts = r->lv[l];
      L.ntiles = std::max(1, (L.pts.n + tile - 1) / tile);
      maxTiles = std::max(maxTiles, L.ntiles);
    }
    d.slotBase = (int)slot;
    slot += (size_t)maxTiles;
    b->depthVersion[j] = r->kf->depthVersion;
  }
  const int NS = b->kind == LSDG_JOB_SIM3 ? NS_SIM3 : NS_SE3;
  if (slot > b->rows) {
    if (b->d_partials) (void)hipFree(b->d_partials);
    b->d_partials = nullptr;
    CK(c, hipMalloc(&b->d_partials, sizeof(double) * NS * slot));
    b->rows = slot;
  }
  CK(c, hipMemcpyAsync(b->d_jobs, b->h_jobs.data(), sizeof(JobDesc) * b->n, hipMemcpyHostToDevice, c->stream));
  CK(c, hipStreamSynchronize(c->stream));
  return LSDG_OK;
}

// First job of chain c (chain c holds jobs [chain_begin(c), chain_begin(c + 1))).
static int chain_begin(lsdg_batch b, int c) { return (int)((long long)c * b->n / b->chains); }

// Kernel arguments of chain c: its jobs, control blocks, tile counts
// ([2][jobs of the chain]) and remaining-jobs counter.
static BatchArgs batch_args(lsdg_batch b, int c) {
  const int j0 = chain_begin(b, c), j1 = chain_begin(b, c + 1);
  BatchArgs a;
  a.jobs = b->d_jobs + j0;
  a.st = b->d_st + j0;
  a.njobs = j1 - j0;
  a.ntiles = b->d_ntiles + 2 * j0;
  a.counters = b->d_counters + j0;
  a.partials = b->d_partials;
  a.remaining = b->d_remaining + c;
  a.ppt = b->ppt;
  a.jobMaxPoints = b->jobMaxPoints;
  const size_t rowWords = (size_t)2 * std::max(b->coopK, 1) * (b->kind == LSDG_JOB_SIM3 ? NS_SIM3 : NS_SE3) * 2;
  a.coopRows = b->d_coop ? b->d_coop + (size_t)j0 * rowWords : nullptr;
  a.coopWords = b->d_coop ? (size_t)(j1 - j0) * rowWords : 0;
  a.epoch = b->d_epoch + c;
  a.coopK = b->coopK;
  a.soloMax = b->soloMax;
  a.spinMax = b->spinMax;
  a.P = b->ctx->P;
  a.kp = track_kparams(b->ctx->P, b->kind == LSDG_JOB_SIM3);
  // the batched LM reads the affine sums only when it estimates the affine
  // brightness; the results report the estimate, not the sums
  a.kp.aff_stats = b->ctx->P.use_affine_lightning_estimation;
  a.xcds = b->coopK > 0 ? b->xcds : 1;
  return a;
}

// Captures one graph: optional pyramid rebuild; then per chain, as parallel
// branches: init and whole-job kernel (first graph only), `steps` LM steps
// from parity 0 and the chain's remaining-jobs counter to host memory.
static lsdg_status capture(lsdg_batch b, bool first, hipGraph_t* g, hipGraphExec_t* exec) {
  lsdg_ctx c = b->ctx;
  CK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (first && b->pyr.nf > 0) e = launch_pyr_batch(b->pyr, c->stream);
  // fork: every side stream starts after the main stream's work so far
  if (e == hipSuccess && b->chains > 1) e = hipEventRecord(b->forkJoin[0], c->stream);
  for (int k = 1; k < b->chains && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->side[k - 1], b->forkJoin[0], 0);
  const int steps = first ? b->steps : b->contSteps;
  for (int k = 0; k < b->chains && e == hipSuccess; ++k) {
    hipStream_t st = k == 0 ? c->stream : b->side[k - 1];
    BatchArgs a = batch_args(b, k);
    if (first) {
      e = launch_batch_init(a, st);
      if (e == hipSuccess && b->coopK > 0) e = launch_batch_coop(a, b->kind, b->prec, b->acc, b->coopBS, st);
      else if (e == hipSuccess && b->jobMaxPoints > 0) e = launch_batch_job(a, b->kind, b->prec, b->acc, st);
    }
    for (int s = 0; s < steps && e == hipSuccess; ++s)
      e = launch_batch_step(a, b->kind, b->prec, b->acc, s & 1, b->grid, st);
    if (e == hipSuccess) e = hipMemcpyAsync(b->h_remaining + k, b->d_remaining + k, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->h_st + chain_begin(b, k), b->d_st + chain_begin(b, k), sizeof(JobState) * a.njobs,
                         hipMemcpyDeviceToHost, st);
  }
  // join
  for (int k = 1; k < b->chains && e == hipSuccess; ++k) {
    e = hipEventRecord(b->forkJoin[k], b->side[k - 1]);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, b->forkJoin[k], 0);
  }
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(c, LSDG_EDEVICE, "batch graph capture", e != hipSuccess ? e : e2);
  }
  hipError_t e3 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  if (e3 != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(c, LSDG_EDEVICE, "hipGraphInstantiate", e3);
  }
  *g = graph;
  return LSDG_OK;
}

static lsdg_status batch_graphs(lsdg_batch b) {
  if (b->first) (void)hipGraphExecDestroy(b->first);
  if (b->cont) (void)hipGraphExecDestroy(b->cont);
  if (b->g_first) (void)hipGraphDestroy(b->g_first);
  if (b->g_cont) (void)hipGraphDestroy(b->g_cont);
  b->first = b->cont = nullptr;
  b->g_first = b->g_cont = nullptr;
  lsdg_status st = capture(b, true, &b->g_first, &b->first);
  if (st != LSDG_OK) return st;
  return capture(b, false, &b->g_cont, &b->cont);
}

static lsdg_status validate_jobs(lsdg_ctx c, const lsdg_job* jobs, int32_t n) {
  if (!c || !jobs || n < 1) return LSDG_EINVAL;
  if (n > MAX_BATCH_JOBS) return fail(c, LSDG_EINVAL, "too many jobs in one batch");
  const int kind = jobs[0].kind;
  const int prec = jobs[0].frame ? jobs[0].frame->precision : LSDG_FP32;
  for (int j = 0; j < n; ++j) {
    const lsdg_job& J = jobs[j];
    if (!J.ref || !J.frame) return fail(c, LSDG_EINVAL, "job without reference or frame");
    if (J.ref->ctx != c || J.frame->ctx != c) return fail(c, LSDG_EINVAL, "job objects belong to another context");
    if (J.kind != kind) return fail(c, LSDG_EINVAL, "a batch holds one kind of job (SE3 or Sim3)");
    if (J.frame->precision != prec) return fail(c, LSDG_EINVAL, "a batch holds frames of one precision");
    if (J.frame->w[0] != J.ref->kf->w[0] || J.frame->h[0] != J.ref->kf->h[0])
      return fail(c, LSDG_EINVAL, "frame / keyframe size mismatch");
    if (!J.ref->kf->hasDepth) return fail(c, LSDG_ESTATE, "keyframe has no depth");
    if (kind == LSDG_JOB_SIM3) {
      if (!J.frame->hasDepth) return fail(c, LSDG_ESTATE, "Sim3 job needs a frame with depth");
      if (J.start_level < 0 || J.start_level >= PYR || J.final_level < 0 || J.final_level > J.start_level)
        return fail(c, LSDG_EINVAL, "level range");
    } else if (kind != LSDG_JOB_SE3) {
      return fail(c, LSDG_EINVAL, "unknown job kind");
    }
  }
  return LSDG_OK;
}

// The default shape: the cooperative kernel with 512-thread blocks, as many
// blocks per job as the device holds at once (at most 16), levels above 2048
// points split; when the device cannot hold two blocks per job, the
// whole-job kernel for levels up to 6144 points and multi-block steps above,
// in two chains from 16 jobs on.
lsdg_batch_config lsdg::batch_default_shape(int n) {
  lsdg_batch_config s{};
  s.schedule = LSDG_SCHEDULE_COOP;
  s.coop_block = 512;
  s.coop_k = 16;
  s.solo_max = 2048;
  s.job_max_points = 12 * BATCH_JOB_BS;
  s.chains = n >= 16 ? 2 : 1;
  s.points_per_thread = 4;
  s.xcd_groups = 1;
  return s;
}

// Environment knobs that pin a schedule (tests and measurements); each
// overrides the shape it names.
static void env_shape(lsdg_batch_config& s, int n) {
  if (const char* e = std::getenv("LSDG_BATCH_COOP")) s.schedule = std::atoi(e) ? LSDG_SCHEDULE_COOP : LSDG_SCHEDULE_STEPS;
  if (const char* e = std::getenv("LSDG_BATCH_COOP_K")) s.coop_k = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("LSDG_BATCH_COOP_BS")) s.coop_block = std::atoi(e);
  if (const char* e = std::getenv("LSDG_BATCH_SOLO")) s.solo_max = std::max(0, std::atoi(e));
  if (const char* e = std::getenv("LSDG_BATCH_PPT")) s.points_per_thread = std::max(1, std::min(64, std::atoi(e)));
  if (const char* e = std::getenv("LSDG_BATCH_JOB_POINTS")) s.job_max_points = std::max(0, std::atoi(e));
  if (const char* e = std::getenv("LSDG_BATCH_CHAINS")) s.chains = std::max(1, std::min(8, std::min(n, std::atoi(e))));
  if (const char* e = std::getenv("LSDG_BATCH_XCDS")) s.xcd_groups = std::max(1, std::min(64, std::atoi(e)));
}

lsdg_status lsdg::batch_create_shaped(lsdg_ctx c, const lsdg_job* jobs, int32_t n, const lsdg_batch_config& shape,
                                      lsdg_batch* out) {
  if (!out) return LSDG_EINVAL;
  lsdg_status vst = validate_jobs(c, jobs, n);
  if (vst != LSDG_OK) return vst;
  if (shape.schedule == LSDG_SCHEDULE_COOP && !batch_coop_block_size_ok(shape.coop_block))
    return fail(c, LSDG_EINVAL, "cooperative block size must be 256 or 512");
  const lsdg_params& P = c->P;
  const int kind = jobs[0].kind;
  const int prec = jobs[0].frame->precision;
  lsdg_batch b = new (std::nothrow) lsdg_batch_s();
  if (!b) return LSDG_ENOMEM;
  b->ctx = c;
  b->n = n;
  b->kind = kind;
  b->prec = prec;
  // The context's summation mode at creation (lsdg_ctx_set_sums): FP64 sums
  // take the plain path's LM decisions; compensated FP32 sums differ from
  // them by ~1e-7 relative.
  b->acc = c->sums == LSDG_SUMS_KAHAN ? BATCH_ACC_KAHAN : BATCH_ACC_F64;
  if (const char* e = std::getenv("LSDG_BATCH_STEPS")) b->steps = std::max(0, std::min(512, std::atoi(e) & ~1));
  // polls of the cooperative row-exchange wait before a job's blocks give up
  // (LSDG_EDEVICE); a small value makes the give-up path run on purpose
  if (const char* e = std::getenv("LSDG_BATCH_SPIN_MAX")) b->spinMax = (unsigned)std::max(0, std::atoi(e));
  b->ppt = std::max(1, std::min(64, shape.points_per_thread));
  b->jobMaxPoints = std::max(0, shape.job_max_points);
  b->chains = std::max(1, std::min(8, std::min(n, shape.chains)));
  b->soloMax = std::max(0, shape.solo_max);
  // The cooperative schedule needs the K blocks of every job resident
  // together: K is capped by the blocks the device holds at once; below two
  // blocks per job the plan falls back to the whole-job kernel and steps.
  if (shape.schedule == LSDG_SCHEDULE_COOP) {
    int cap = 0;
    if (batch_coop_capacity(kind, prec, b->acc, shape.coop_block, &cap) == hipSuccess) {
      // with dispatch groups the jobs are padded to a multiple of the groups
      const int xcds = std::max(1, std::min(64, shape.xcd_groups));
      const int nPad = xcds > 1 ? ((n + xcds - 1) / xcds) * xcds : n;
      const int K = std::min(std::max(1, shape.coop_k), std::min(16, cap / nPad));
      if (K >= 2) {
        b->coopK = K;
        b->coopBS = shape.coop_block;
        b->xcds = xcds;
        b->chains = 1;
        b->jobMaxPoints = 1 << 30;
      }
    }
  }
  b->grid = 1024 / b->chains;  // evaluation blocks per chain
  if (const char* e = std::getenv("LSDG_BATCH_GRID")) b->grid = std::max(1, std::min(65536, std::atoi(e)));
  b->jobs.assign(jobs, jobs + n);
  for (auto& J : b->jobs) {
    if (kind == LSDG_JOB_SE3) {
      J.start_level = P.se3_max_level - 1;
      J.final_level = P.se3_min_level;
    }
  }
  b->depthVersion.assign(n, -1);
  lsdg_status st = LSDG_OK;
  hipError_t e;
  if ((e = hipMalloc(&b->d_jobs, sizeof(JobDesc) * n)) != hipSuccess ||
      (e = hipMalloc(&b->d_st, sizeof(JobState) * n)) != hipSuccess ||
      (e = hipMalloc(&b->d_ntiles, sizeof(int) * 2 * n)) != hipSuccess ||
      (e = hipMalloc(&b->d_counters, sizeof(unsigned) * n)) != hipSuccess ||
      (e = hipMalloc(&b->d_remaining, 64)) != hipSuccess ||
      (b->coopK > 0 && (e = hipMalloc(&b->d_coop, sizeof(unsigned long long) * 2 * b->coopK * NS_SIM3 * 2 * (size_t)n)) !=
                           hipSuccess) ||
      (e = hipMalloc(&b->d_epoch, 64)) != hipSuccess || (e = hipMemset(b->d_epoch, 0, 64)) != hipSuccess ||
      (e = hipHostMalloc(&b->h_remaining, 64, hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc(&b->h_st, sizeof(JobState) * n, hipHostMallocDefault)) != hipSuccess) {
    batch_free(b);
    return fail(c, LSDG_ENOMEM, "batch allocation", e);
  }
  for (int k = 0; k < b->chains; ++k) {
    hipEvent_t ev = nullptr;
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
      batch_free(b);
      return fail(c, LSDG_EDEVICE, "batch events", e);
    }
    b->forkJoin.push_back(ev);
    if (k == 0) continue;
    hipStream_t st = nullptr;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) {
      batch_free(b);
      return fail(c, LSDG_EDEVICE, "batch streams", e);
    }
    b->side.push_back(st);
  }
  // pyramid data each frame is tracked on (union over its jobs), built now so
  // the frame holds it; frames the batch rebuilds get it again on every run
  std::vector<lsdg_frame>& frames = b->frames;
  std::vector<unsigned>& needs = b->needs;
  for (auto& J : b->jobs) {
    unsigned m = fd_levels(FD_IMAGE, 0, PYR - 1) | fd_track(J.frame, J.final_level, J.start_level);
    auto it = std::find(frames.begin(), frames.end(), J.frame);
    if (it == frames.end()) {
      frames.push_back(J.frame);
      needs.push_back(m);
    } else {
      needs[it - frames.begin()] |= m;
    }
    if (J.rebuild_pyramid && std::find(b->pyrFrames.begin(), b->pyrFrames.end(), J.frame) == b->pyrFrames.end())
      b->pyrFrames.push_back(J.frame);
  }
  for (size_t k = 0; k < frames.size(); ++k)
    if ((st = frame_require(frames[k], needs[k])) != LSDG_OK) {
      batch_free(b);
      return st;
    }
  for (lsdg_frame f : b->pyrFrames) b->pyrNeed.push_back(needs[std::find(frames.begin(), frames.end(), f) - frames.begin()]);
  for (lsdg_frame f : frames)
    b->rebuilt.push_back(std::find(b->pyrFrames.begin(), b->pyrFrames.end(), f) != b->pyrFrames.end());
  if (!b->pyrFrames.empty()) {
    size_t bytes = pyr_batch_bytes((int)b->pyrFrames.size());
    if ((e = hipMalloc(&b->d_pyr, bytes)) != hipSuccess || (e = hipHostMalloc(&b->h_pyr, bytes)) != hipSuccess) {
      batch_free(b);
      return fail(c, LSDG_ENOMEM, "batch pyramid tables", e);
    }
    pyr_batch_tables(b->pyrFrames.data(), b->pyrNeed.data(), (int)b->pyrFrames.size(), b->h_pyr, b->d_pyr, &b->pyr);
    if ((e = hipMemcpy(b->d_pyr, b->h_pyr, bytes, hipMemcpyHostToDevice)) != hipSuccess) {
      batch_free(b);
      return fail(c, LSDG_EDEVICE, "batch pyramid tables", e);
    }
  }
  for (auto& J : b->jobs)
    if (kind == LSDG_JOB_SE3 && (st = frame_ensure_was_good(J.frame)) != LSDG_OK) {
      batch_free(b);
      return st;
    }
  if ((st = build_job_table(b)) != LSDG_OK) {
    batch_free(b);
    return st;
  }
  if (!std::getenv("LSDG_BATCH_STEPS")) {
    // no multi-block steps in the first graph when every level runs in k_batch_job
    int maxPts = 0;
    for (const JobDesc& d : b->h_jobs)
      for (int l = 0; l < PYR; ++l) maxPts = std::max(maxPts, d.lv[l].pts.n);
    if (maxPts <= b->jobMaxPoints) b->steps = 0;
  }
  if ((st = batch_graphs(b)) != LSDG_OK) {
    batch_free(b);
    return st;
  }
  *out = b;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_batch_create(lsdg_ctx c, const lsdg_job* jobs, int32_t n, lsdg_batch* out) {
  if (!c || !jobs || !out || n < 1) return LSDG_EINVAL;
  lsdg_status st = validate_jobs(c, jobs, n);
  if (st != LSDG_OK) return st;
  // the tuned shape of this class of batch when the table has one
  lsdg_batch_config shape = batch_default_shape(n);
  char key[LSDG_TUNE_KEY_BYTES];
  if (lsdg_batch_tune_key(c, jobs, n, key) == LSDG_OK) tune_table_lookup(key, &shape);
  env_shape(shape, n);
  return batch_create_shaped(c, jobs, n, shape, out);
}

extern "C" lsdg_status lsdg_batch_create_with_config(lsdg_ctx c, const lsdg_job* jobs, int32_t n,
                                                     const lsdg_batch_config* config, lsdg_batch* out) {
  if (!c || !jobs || !config || !out || n < 1) return LSDG_EINVAL;
  if (config->schedule != LSDG_SCHEDULE_COOP && config->schedule != LSDG_SCHEDULE_STEPS)
    return fail(c, LSDG_EINVAL, "unknown schedule");
  return batch_create_shaped(c, jobs, n, *config, out);
}

extern "C" lsdg_status lsdg_batch_get_config(lsdg_batch b, lsdg_batch_config* out) {
  if (!b || !out) return LSDG_EINVAL;
  std::memset(out, 0, sizeof(*out));
  out->schedule = b->coopK > 0 ? LSDG_SCHEDULE_COOP : LSDG_SCHEDULE_STEPS;
  out->coop_block = b->coopBS;
  out->coop_k = b->coopK;
  out->solo_max = b->soloMax;
  out->job_max_points = b->jobMaxPoints;
  out->chains = b->chains;
  out->points_per_thread = b->ppt;
  out->xcd_groups = b->xcds;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_batch_destroy(lsdg_batch b) {
  if (!b) return LSDG_EINVAL;
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_free(b);
  return LSDG_OK;
}

static void fill_result(lsdg_batch b, int j, const JobState& s, lsdg_job_result* r) {
  const lsdg_job& J = b->jobs[j];
  std::memset(r, 0, sizeof(*r));
  lsdg_track_stats& st = r->stats;
  for (int l = 0; l < PYR; ++l) {
    st.num_residual_calls[l] = s.numRes[l];
    st.num_warp_update_calls[l] = s.numWarp[l];
  }
  st.point_usage = (float)s.usage;
  st.affine_a = s.affA;
  st.affine_b = s.affB;
  Pose I = pose_identity();
  if (s.status == 2) {
    st.diverged = 1;
    st.tracking_was_good = 0;
    st.last_good_count = (int)s.good;
    st.last_bad_count = (int)s.bad;
    pose_to8(I, r->frame_to_ref);
    r->status = LSDG_DIVERGED;
    if (b->kind == LSDG_JOB_SIM3) std::memset(st.hessian, 0, sizeof st.hessian);
    return;
  }
  Pose res = pose_inv(s.T);
  if (b->kind == LSDG_JOB_SE3) {
    res.s = 1.0;
    st.last_good_count = (int)s.good;
    st.last_bad_count = (int)s.bad;
    st.last_mean_res = (float)s.meanRes;
    st.last_residual = (float)s.lastResidual;
    lsdg_frame f = J.frame;
    const int ml = b->ctx->P.se3_min_level;
    bool good = s.good / (double)(f->w[ml] * f->h[ml]) > (double)b->ctx->P.min_goodperall_pixel &&
                s.good / (s.good + s.bad) > (double)b->ctx->P.min_goodpergoodbad_pixel;
    st.tracking_was_good = good ? 1 : 0;
    // the frame fields SE3Tracker::trackFrame writes
    if (good) J.ref->kf->numFramesTrackedOnThis++;
    f->initialTrackedResidual = (float)(s.lastResidual / s.usage);
    f->thisToParent = res;
    f->parent = J.ref->kf;
    f->parentId = J.ref->kf->id;
  } else {
    st.last_residual = (float)s.finalRes[6];
    st.last_depth_residual = (float)s.finalRes[4];
    st.last_photometric_residual = (float)s.finalRes[5];
    st.tracking_was_good = 1;
    int k = 0;
    for (int a = 0; a < 7; ++a)
      for (int cc = a; cc < 7; ++cc) {
        st.hessian[7 * a + cc] = s.lgsA[k];
        st.hessian[7 * cc + a] = s.lgsA[k];
        ++k;
      }
  }
  pose_to8(res, r->frame_to_ref);
  r->status = LSDG_OK;
}

// The same work as the graphs, launched in stream order with an event pair
// around every LM-step kernel (and the pyramid rebuild) for lsdg_ctx_profile.
static lsdg_status batch_run_profiled(lsdg_batch b) {
  lsdg_ctx c = b->ctx;
  if (b->pyr.nf > 0) {
    double pb = 0.0;
    for (size_t k = 0; k < b->pyrFrames.size(); ++k) pb += pyramid_bytes(b->pyrFrames[k], true, b->pyrNeed[k]);
    ProfMark pm = prof_begin(c, LSDG_KERNEL_PYRAMID);
    CK(c, launch_pyr_batch(b->pyr, c->stream));
    prof_end(c, pm, pb);
  }
  // chains one after the other on the context stream
  for (int k = 0; k < b->chains; ++k) {
    BatchArgs a = batch_args(b, k);
    CK(c, launch_batch_init(a, c->stream));
    if (b->coopK > 0) {
      KTimer kt = prof_timer(c);
      CK(c, launch_batch_coop(a, b->kind, b->prec, b->acc, b->coopBS, c->stream, &kt));
      prof_timer_done(c, LSDG_KERNEL_BATCH_COOP, kt, 0.0);
    } else if (b->jobMaxPoints > 0) {
      KTimer kt = prof_timer(c);
      CK(c, launch_batch_job(a, b->kind, b->prec, b->acc, c->stream, &kt));
      prof_timer_done(c, LSDG_KERNEL_BATCH_JOB, kt, 0.0);
    }
    bool done = false;
    for (int round = 0; round <= 64 && !done; ++round) {
      const int steps = round == 0 ? b->steps : b->contSteps;
      for (int s = 0; s < steps; ++s) {
        KTimer ke = prof_timer(c);
        CK(c, launch_batch_eval(a, b->kind, b->prec, b->acc, s & 1, b->grid, c->stream, &ke));
        prof_timer_done(c, LSDG_KERNEL_BATCH_EVAL, ke, 0.0);
        KTimer ks = prof_timer(c);
        CK(c, launch_batch_solve(a, b->kind, s & 1, c->stream, &ks));
        prof_timer_done(c, LSDG_KERNEL_BATCH_SOLVE, ks, 0.0);
      }
      CK(c, hipMemcpyAsync(b->h_remaining + k, b->d_remaining + k, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      CK(c, hipStreamSynchronize(c->stream));
      if (b->h_remaining[k] < 0) return fail(c, LSDG_EDEVICE, "batched LM: cooperative blocks not co-resident");
      done = b->h_remaining[k] <= 0;
    }
    if (!done) return fail(c, LSDG_ESTATE, "batched LM did not finish");
  }
  return prof_resolve(c);
}

static int remaining_jobs(lsdg_batch b) {
  int r = 0;
  for (int k = 0; k < b->chains; ++k) r += b->h_remaining[k];
  return r;
}

// Brings every frame of the plan back to what the plan reads: a frame given
// new content since the last run (lsdg_frame_ingest keeps only its image
// levels, drops its depth and its refPixelWasGood mask) gets its derived
// levels rebuilt unless the graph rebuilds it, and its mask reset; a Sim3
// job's frame must still carry depth.  Frames whose data is in place cost
// nothing here.
static lsdg_status batch_refresh_frames(lsdg_batch b) {
  lsdg_ctx c = b->ctx;
  for (size_t k = 0; k < b->frames.size(); ++k) {
    lsdg_frame f = b->frames[k];
    if (b->kind == LSDG_JOB_SIM3 && !f->hasDepth)
      return fail(c, LSDG_ESTATE, "a Sim3 job's frame has no depth (new content since the plan was built?)");
    if (!b->rebuilt[k]) {
      lsdg_status st = frame_require(f, b->needs[k]);
      if (st != LSDG_OK) return st;
    }
    if (b->kind == LSDG_JOB_SE3) {
      lsdg_status st = frame_ensure_was_good(f);
      if (st != LSDG_OK) return st;
    }
  }
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_batch_run(lsdg_batch b, lsdg_job_result* results) {
  if (!b) return LSDG_EINVAL;
  lsdg_ctx c = b->ctx;
  {
    lsdg_status st = batch_refresh_frames(b);
    if (st != LSDG_OK) return st;
  }
  // keyframes whose depth changed since the plan was built need new point clouds
  bool stale = false;
  for (int j = 0; j < b->n; ++j) stale |= b->jobs[j].ref->kf->depthVersion != b->depthVersion[j];
  if (stale) {
    lsdg_status st = build_job_table(b);
    if (st == LSDG_OK) st = batch_graphs(b);
    if (st != LSDG_OK) return st;
  }
  if (c->prof_on) {
    lsdg_status st = batch_run_profiled(b);
    if (st != LSDG_OK) return st;
  } else {
    CK(c, hipGraphLaunch(b->first, c->stream));
    CK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < b->chains; ++k)
      if (b->h_remaining[k] < 0) return fail(c, LSDG_EDEVICE, "batched LM: cooperative blocks not co-resident");
    int launches = 0;
    while (remaining_jobs(b) > 0) {
      if (++launches > 64) return fail(c, LSDG_ESTATE, "batched LM did not finish within 64 continuation graphs");
      CK(c, hipGraphLaunch(b->cont, c->stream));
      CK(c, hipStreamSynchronize(c->stream));
    }
    // Size the first graph's multi-block steps to what this run needed (the
    // most evaluations one job ran at levels above jobMaxPoints), when it was
    // too short or more than a few steps too long.
    // (the graphs copied the control blocks to h_st)
    int need = 0;
    for (int j = 0; j < b->n; ++j) {
      int e = 0;
      for (int l = 0; l < PYR; ++l)
        if (b->h_jobs[j].lv[l].pts.n > b->jobMaxPoints) e += b->h_st[j].evalsAt[l];
      need = std::max(need, e);
    }
    const int steps = std::min(128, (need + 1) & ~1);
    if ((launches > 0 && steps > b->steps) || steps + 4 < b->steps) {
      b->steps = steps;
      lsdg_status st = batch_graphs(b);
      if (st != LSDG_OK) return st;
    }
  }
  // the graph rebuilt these frames' levels
  for (size_t k = 0; k < b->pyrFrames.size(); ++k) b->pyrFrames[k]->valid |= b->pyrNeed[k];
  if (results || c->prof_on) {
    if (c->prof_on) CK(c, hipMemcpy(b->h_st, b->d_st, sizeof(JobState) * b->n, hipMemcpyDeviceToHost));
    if (results)
      for (int j = 0; j < b->n; ++j) fill_result(b, j, b->h_st[j], &results[j]);
    if (c->prof_on) {
      // algorithmic bytes of every evaluation, split by the kernel that ran it
      double bytesSplit = 0.0, bytesJob = 0.0;
      for (int j = 0; j < b->n; ++j)
        for (int l = 0; l < PYR; ++l) {
          const int np = b->h_jobs[j].lv[l].pts.n;
          const double eb = b->h_st[j].evalsAt[l] * eval_bytes(np, b->jobs[j].frame, l, b->kind == LSDG_JOB_SIM3);
          if (np > b->jobMaxPoints) bytesSplit += eb;
          else bytesJob += eb;
        }
      if (b->coopK > 0) {
        c->prof[LSDG_KERNEL_BATCH_COOP].bytes += bytesSplit + bytesJob;
      } else {
        c->prof[LSDG_KERNEL_BATCH_EVAL].bytes += bytesSplit;
        c->prof[LSDG_KERNEL_BATCH_JOB].bytes += bytesJob;
      }
    }
  }
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_batch_track(lsdg_ctx c, const lsdg_job* jobs, int32_t n, lsdg_job_result* results) {
  lsdg_batch b = nullptr;
  lsdg_status st = lsdg_batch_create(c, jobs, n, &b);
  if (st != LSDG_OK) return st;
  st = lsdg_batch_run(b, results);
  lsdg_batch_destroy(b);
  return st;
}
