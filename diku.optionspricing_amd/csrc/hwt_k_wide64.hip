// hwt_k_wide64.hip -- the FP64 wide-tree kernel, and mixed mode's FP64 re-pricing
// kernel (same per-option routine).
#include "hwt_pricers.h"

namespace hwt {

template __global__ void k_price_wide<double>(const OptConst *__restrict__, const double *__restrict__, const int32_t *__restrict__,
                                           int32_t *__restrict__, double *__restrict__, double *__restrict__,
                                           int32_t *__restrict__);

// Mixed mode's second pass: the options the FP32 kernels handed back (state
// out of FP32's range, F32_SAFE_STATE) priced with FP64 states, any width, from
// their own queue.  Every wave reads the count and leaves at once when the
// list is empty, which is the usual case.
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2))) void k_price_redo(
    const OptConst *__restrict__ cst, const double *__restrict__ edge, const int32_t *__restrict__ redo,
    int32_t *__restrict__ tally, double *__restrict__ out, double *__restrict__ scratch)
{
    const int nredo = uni(tally[T_NREDO]);
    price_tasks<double, MAXK, false, true>(cst, edge, redo, tally, T_QUEUE_X, 0, nredo, out, scratch, nullptr);
    stamp_wave_end(tally);
}

} // namespace hwt
