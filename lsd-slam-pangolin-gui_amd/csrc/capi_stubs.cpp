// Entry points whose kernels are not built yet: they report LSDG_ESTATE.
#include "capi_common.h"
using namespace lsdg;
extern "C" {
lsdg_status lsdg_batch_create(lsdg_ctx c, const lsdg_job*, int32_t, lsdg_batch*) { return fail(c, LSDG_ESTATE, "batch not available"); }
lsdg_status lsdg_batch_run(lsdg_batch, lsdg_job_result*) { return LSDG_ESTATE; }
lsdg_status lsdg_batch_destroy(lsdg_batch) { return LSDG_ESTATE; }
lsdg_status lsdg_batch_track(lsdg_ctx c, const lsdg_job*, int32_t, lsdg_job_result*) { return fail(c, LSDG_ESTATE, "batch not available"); }
}
