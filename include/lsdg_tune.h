/*
 * lsdg_tune.h — shapes of a batched-tracking plan and their tuner.
 *
 * A batch plan (lsdg_batch_create, include/lsdg.h) runs many frame-vs-keyframe
 * alignments in one persistent launch.  How the work is cut is the plan's
 * shape: the schedule (the cooperative whole-LM kernel, or the whole-job
 * kernel plus multi-block LM steps), the cooperative block size and blocks
 * per job, the largest pyramid level a job's blocks evaluate whole rather
 * than split, and for the step schedule the whole-job threshold, the number
 * of parallel chains and the points per thread of a reduction tile.  Every
 * shape computes the same FP32 per-point terms and sums them in the context's
 * summation mode (lsdg_ctx_set_sums); shapes differ only in the order of the
 * sum (~1e-16 relative under LSDG_SUMS_FP64), which leaves the LM path
 * unchanged: the tests hold every shape to 1e-9 twist of the plain path with
 * the same evaluation counts.  Only the time differs.
 *
 * lsdg_batch_tune searches the shapes for a set of jobs and records the best
 * in a process-wide table keyed by the batch's class (job kind, pyramid
 * precision, resolution class, job count class); lsdg_batch_create takes the
 * recorded shape for its class when there is one.  The table is saved to and
 * loaded from a JSON file, so a deployment tunes once per device type.
 *
 * No counterpart in the reference, whose trackers run one alignment at a time
 * on the CPU (SE3Tracker::trackFrame, Sim3Tracker::trackFrameSim3, reached
 * through SlamSystem::nextImage, lib/App/InputThread.cpp:71).
 */
#ifndef LSDG_TUNE_H
#define LSDG_TUNE_H

#include "lsdg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LSDG_SCHEDULE_COOP 0   /* cooperative whole-LM kernel, coop_k blocks of coop_block threads per job */
#define LSDG_SCHEDULE_STEPS 1  /* whole-job kernel up to job_max_points, then multi-block LM steps */

typedef struct lsdg_batch_config {
  int32_t schedule;           /* LSDG_SCHEDULE_* */
  int32_t coop_block;         /* cooperative block size: 256 or 512 threads */
  int32_t coop_k;             /* blocks per job (capped by what the device holds at once, at most 16) */
  int32_t solo_max;           /* cooperative: levels of at most this many points are evaluated whole */
  int32_t job_max_points;     /* steps: levels of at most this many points run in the whole-job kernel */
  int32_t chains;             /* steps: groups of jobs stepping as parallel graph branches */
  int32_t points_per_thread;  /* steps: reduction tile = 256 x points_per_thread points */
  int32_t xcd_groups;         /* cooperative: 1 = blocks in job order; G > 1 = the K blocks of a job
                                 placed in one of G round-robin dispatch groups (the device's
                                 XCDs), so a job's points and taps stay in one L2 */
} lsdg_batch_config;

#define LSDG_TUNE_KEY_BYTES 64

/* The class a batch belongs to, as the tuning table's key, e.g.
 * "se3/fp32/p19/single/n64": job kind, pyramid precision (with "-kahan" under
 * LSDG_SUMS_KAHAN), ceil(log2) of the largest level-0 pixel count, whether the
 * batch mixes resolutions, and ceil(log2) of the job count. */
lsdg_status lsdg_batch_tune_key(lsdg_ctx ctx, const lsdg_job* jobs, int32_t n, char key[LSDG_TUNE_KEY_BYTES]);

/* A plan with an explicit shape (a cooperative schedule falls back to the
 * step schedule when the device cannot hold two blocks per job at once). */
lsdg_status lsdg_batch_create_with_config(lsdg_ctx ctx, const lsdg_job* jobs, int32_t n,
                                          const lsdg_batch_config* config, lsdg_batch* out);
/* The shape a plan took (after the fallback and device caps). */
lsdg_status lsdg_batch_get_config(lsdg_batch b, lsdg_batch_config* out);

/* Times the default shape and a coordinate search over the shapes (each
 * candidate: one warm-up run, then `runs` timed runs, median), records the
 * fastest under the batch's key and returns it with both median run times.
 * The frames' content is left as the runs leave it (tracked poses and
 * refPixelWasGood masks written). */
lsdg_status lsdg_batch_tune(lsdg_ctx ctx, const lsdg_job* jobs, int32_t n, int32_t runs, lsdg_batch_config* best,
                            double* default_ms, double* best_ms);

/* The process-wide tuning table. */
lsdg_status lsdg_tune_table_put(const char* key, const lsdg_batch_config* config);
/* LSDG_ESTATE when the key has no entry. */
lsdg_status lsdg_tune_table_get(const char* key, lsdg_batch_config* out);
lsdg_status lsdg_tune_table_clear(void);
int32_t lsdg_tune_table_size(void);
/* JSON: {"lsdg_tune": 1, "entries": [{"key": "...", "schedule": 0, ...}, ...]};
 * load merges into the table (entries of the file win). */
lsdg_status lsdg_tune_table_save(const char* path);
lsdg_status lsdg_tune_table_load(const char* path);

#ifdef __cplusplus
}
#endif

#endif /* LSDG_TUNE_H */
