// hwt_k_fused32.hip -- the single persistent kernel for every tree class, FP32
// tree states (mixed mode).
#include "hwt_pricers.h"

namespace hwt {

template __global__ void k_price_fused<float>(const OptConst *__restrict__, const double *__restrict__, const int32_t *__restrict__,
                                           int32_t *__restrict__, double *__restrict__, double *__restrict__,
                                           int32_t *__restrict__);

} // namespace hwt
