// placeholder entry points of the batched / serving API (replaced by sp_server.cpp)
#include <hip/hip_runtime.h>
#include "sp_host.h"
extern "C" {
int sp_run_batch(sp_ctx*, uint32_t, const sp_request*, const sp_particles*, sp_result*) { return sp::set_error(SP_ERR_UNSUPPORTED, "batch path not built"); }
int sp_submit(sp_ctx*, const sp_request*, const sp_particles*, uint64_t*) { return sp::set_error(SP_ERR_UNSUPPORTED, "server not built"); }
int sp_wait(sp_ctx*, uint64_t, int) { return sp::set_error(SP_ERR_UNSUPPORTED, "server not built"); }
int sp_result_get(sp_ctx*, uint64_t, sp_result*) { return sp::set_error(SP_ERR_UNSUPPORTED, "server not built"); }
int sp_cancel(sp_ctx*, uint64_t) { return sp::set_error(SP_ERR_UNSUPPORTED, "server not built"); }
int sp_server_stats(sp_ctx*, double*) { return sp::set_error(SP_ERR_UNSUPPORTED, "server not built"); }
int sp_autotune(sp_ctx*, const sp_request*, const sp_particles*, sp_tuning*, double*) { return sp::set_error(SP_ERR_UNSUPPORTED, "autotuner not built"); }
}
