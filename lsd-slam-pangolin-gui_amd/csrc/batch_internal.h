// Device-side layout of a batched tracking plan (jobs, per-job LM control
// blocks) shared by batch.hip and capi_batch.cpp.
#pragma once
#include "../../include/lsdg_tune.h"
#include "lsdg_internal.h"

namespace lsdg {

constexpr int MAX_BATCH_JOBS = 2048;
constexpr int BATCH_ACC_F64 = 0;    // LSDG_SUMS_FP64
constexpr int BATCH_ACC_KAHAN = 1;  // LSDG_SUMS_KAHAN
// Whole-job kernel: block size and the bound on the evaluations one job may
// run inside it (the LM loop ends long before; the bound makes every block's
// exit unconditional).
constexpr int BATCH_JOB_BS = 512;
constexpr int BATCH_JOB_MAX_EVALS = 1 << 16;

// One pyramid level of one job: the reference's points, the frame's
// (I, gx, gy) image (FP32 float4 or packed FP16/BF16) and, for Sim3, the
// frame's inverse depth and variance.
struct JobLevel {
  PointsLevel pts;
  const void* img;
  const float* fid;
  const float* fvar;
  LevelGeom g;
  int ntiles;  // ceil(n / tile), at least 1
};

struct JobDesc {
  JobLevel lv[PYR];
  uint8_t* isGood;  // the frame's level-1 refPixelWasGood (SE3), written at level-1 evaluations
  Pose init;        // frameToRef_init
  int startLevel, finalLevel;
  int slotBase;     // first partial-sum row of this job
  int kind;
};

// LM control block of one job (SE3Tracker / Sim3Tracker locals, on the device).
struct JobState {
  Pose T, Tn;        // accepted referenceToFrame, trial
  Xf xf;             // pose of the next evaluation, FP32
  float roll[4];     // its ESM roll (Sim3)
  double inc[7];
  double incSq;
  double lgsA[28], lgsB[7], lgsN;  // unnormalised LGS of the accepted pose
  double lastErr, lastResidual;    // SE3
  double lastRes[7], finalRes[7];  // Sim3 (sumResD, sumResP, numTermsD, numTermsP, meanD, meanP, mean)
  double good, bad, meanRes, usage;
  float lambda, affA, affB, affAIt, affBIt;
  int status;        // 0 running, 1 done, 2 diverged
  int level, phase, iteration, incTry, upToDate, evals;
  int numRes[PYR], numWarp[PYR];
  int evalsAt[PYR];  // evaluations run per level (every step the job was active)
};

struct BatchArgs {
  const JobDesc* jobs;
  JobState* st;
  int njobs;
  int* ntiles;        // [2][njobs] tile counts of this / next step
  unsigned* counters; // [njobs] finished tiles of this step
  double* partials;   // rows of NS doubles
  int* remaining;     // jobs still running
  int ppt;            // points per thread per tile
  int jobMaxPoints;   // levels with at most this many points run in k_batch_job
  unsigned long long* coopRows;  // k_batch_coop: [njobs][2][coopK][NS][2] tagged words
  size_t coopWords;              // words of coopRows (0 without the cooperative schedule)
  unsigned* epoch;               // run counter of this chain (k_batch_init increments it)
  int coopK;          // k_batch_coop: blocks per job
  int soloMax;        // k_batch_coop: levels of at most this many points are not split
  unsigned spinMax;   // k_batch_coop: polls of the row-exchange wait before the blocks give up
  int xcds;           // k_batch_coop: > 1: block b belongs to dispatch group b % xcds, which holds
                      // every block of jobs j with j % xcds == b % xcds (grid padded to a multiple)
  int affineStats;    // accumulate the affine-brightness sums (only read when estimation is on)
  lsdg_params P;
  TrackKParams kp;
};

hipError_t launch_batch_init(const BatchArgs& a, hipStream_t s);
// Cooperative whole-LM kernel (a.coopK blocks of `bs` threads per job, all
// levels; bs is 256 or 512) and the number of its blocks the device holds at
// once.
constexpr unsigned COOP_SPIN_MAX = 1u << 22;
bool batch_coop_block_size_ok(int bs);
// Blocks of a cooperative launch: njobs * K, or with xcds > 1 the jobs
// padded to a multiple of xcds (the padding blocks exit at once).
int batch_coop_grid(int njobs, int K, int xcds);
hipError_t launch_batch_coop(const BatchArgs& a, int kind, int prec, int acc, int bs, hipStream_t s,
                             const KTimer* t = nullptr);
hipError_t batch_coop_capacity(int kind, int prec, int acc, int bs, int* blocks);
// Launches with an optional kernel timer (see launch_k).
hipError_t launch_batch_job(const BatchArgs& a, int kind, int prec, int acc, hipStream_t s,
                            const KTimer* t = nullptr);
hipError_t launch_batch_eval(const BatchArgs& a, int kind, int prec, int acc, int parity, int grid, hipStream_t s,
                             const KTimer* t = nullptr);
hipError_t launch_batch_solve(const BatchArgs& a, int kind, int parity, hipStream_t s, const KTimer* t = nullptr);
hipError_t launch_batch_step(const BatchArgs& a, int kind, int prec, int acc, int parity, int grid, hipStream_t s);

// Plan shapes (capi_batch.cpp) and the tuning table (capi_tune.cpp).
lsdg_batch_config batch_default_shape(int n);
lsdg_status batch_create_shaped(lsdg_ctx c, const lsdg_job* jobs, int32_t n, const lsdg_batch_config& shape,
                                lsdg_batch* out);
// The table's entry for key into *out; false when there is none.
bool tune_table_lookup(const char* key, lsdg_batch_config* out);

}  // namespace lsdg
