// hwt_internal.h -- entry points shared between the library's translation
// units but not part of the public C ABI.
#pragma once

#include "../../include/hwtree.h"

extern "C" {
// The plan's pipeline over arrays holding up to dopts->n options, of which
// the first *dev_count (device memory) are priced, enqueued on `stream` for
// capture into a caller's graph: no plan events are recorded and no timing is
// kept.  The plan's last stream becomes `stream`.
int hwt_internal_plan_enqueue_counted(hwt_plan *p, const hwt_options_f64 *dopts, const int32_t *dev_count,
                                      const hwt_yield_f64 *dyield, double *dout, void *stream);
// A number unique to the plan in this process (graphs that capture the plan's
// kernels key on it, since a freed plan's address can come back).
uint64_t hwt_internal_plan_serial(const hwt_plan *p);
// A caller's graph holding the plan's kernels was launched on `stream`: the
// next hwt_plan_sync reads that pass's status.
void hwt_internal_plan_replayed_on(hwt_plan *p, void *stream);
// Records msg as the calling thread's hwt_last_error() and returns code.
int hwt_internal_fail(int code, const char *msg);
}
