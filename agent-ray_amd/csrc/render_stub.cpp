// Entry points whose kernels land in later milestones.
#include "ar_internal.h"
using namespace ar;
extern "C" {
ar_status ar_server_create(const ar_scene*, const ar_server_config*, ar_server**) {
  return fail(AR_ERR_UNSUPPORTED_TYPE, "serving: not built yet");
}
ar_status ar_server_destroy(ar_server*) { return AR_OK; }
ar_status ar_job_submit(ar_server*, const ar_job_desc*, uint64_t*) { return fail(AR_ERR_UNSUPPORTED_TYPE, "x"); }
ar_status ar_job_poll(ar_server*, uint64_t, int32_t*) { return fail(AR_ERR_UNSUPPORTED_TYPE, "x"); }
ar_status ar_job_wait(ar_server*, uint64_t, float*, ar_job_stats*) { return fail(AR_ERR_UNSUPPORTED_TYPE, "x"); }
ar_status ar_job_cancel(ar_server*, uint64_t) { return fail(AR_ERR_UNSUPPORTED_TYPE, "x"); }
double ar_server_clock(const ar_server*) { return 0.0; }
}
