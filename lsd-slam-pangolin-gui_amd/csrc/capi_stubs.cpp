// Entry points whose kernels are not built yet: they report LSDG_ESTATE.
#include "capi_common.h"
using namespace lsdg;
extern "C" {
lsdg_status lsdg_depthmap_create(lsdg_ctx c, const lsdg_camera*, lsdg_depthmap*) { return fail(c, LSDG_ESTATE, "depth map not available"); }
lsdg_status lsdg_depthmap_destroy(lsdg_depthmap) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_reset(lsdg_depthmap) { return LSDG_ESTATE; }
int32_t lsdg_depth_is_valid(lsdg_depthmap) { return 0; }
lsdg_status lsdg_depth_initialize_from_gt(lsdg_depthmap, lsdg_frame) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_set_prior(lsdg_depthmap, lsdg_frame, const float*, const float*, int32_t) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_update(lsdg_depthmap, const lsdg_frame*, int32_t) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_create_keyframe(lsdg_depthmap, lsdg_frame) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_finalize_keyframe(lsdg_depthmap) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_read(lsdg_depthmap, lsdg_hypotheses*) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_write(lsdg_depthmap, const lsdg_hypotheses*) { return LSDG_ESTATE; }
lsdg_status lsdg_depth_active_keyframe(lsdg_depthmap, int32_t*) { return LSDG_ESTATE; }
lsdg_status lsdg_batch_create(lsdg_ctx c, const lsdg_job*, int32_t, lsdg_batch*) { return fail(c, LSDG_ESTATE, "batch not available"); }
lsdg_status lsdg_batch_run(lsdg_batch, lsdg_job_result*) { return LSDG_ESTATE; }
lsdg_status lsdg_batch_destroy(lsdg_batch) { return LSDG_ESTATE; }
lsdg_status lsdg_batch_track(lsdg_ctx c, const lsdg_job*, int32_t, lsdg_job_result*) { return fail(c, LSDG_ESTATE, "batch not available"); }
}
