// hwt_k_fused64.hip -- the single persistent kernel for every tree class, FP64.
#include "hwt_pricers.h"

namespace hwt {

template __global__ void k_price_fused<double>(const OptConst *__restrict__, const double *__restrict__, const int32_t *__restrict__,
                                           int32_t *__restrict__, double *__restrict__, double *__restrict__,
                                           int32_t *__restrict__);

} // namespace hwt
