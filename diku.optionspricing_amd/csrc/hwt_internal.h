// hwt_internal.h -- entry points shared between the library's translation
// units but not part of the public C ABI.
#pragma once

#include "../../include/hwtree.h"

extern "C" {
// hwt_plan_execute over arrays holding up to dopts->n options, of which the
// first *dev_count (device memory) are priced; launched directly on `stream`.
int hwt_plan_execute_counted(hwt_plan *p, const hwt_options_f64 *dopts, const int32_t *dev_count,
                             const hwt_yield_f64 *dyield, double *dout, void *stream);
// Records msg as the calling thread's hwt_last_error() and returns code.
int hwt_internal_fail(int code, const char *msg);
}
