// hwt_k_narrow.hip -- the narrow-tree kernel (trees up to 128 nodes wide), FP64
// and FP32 tree states.
#include "hwt_pricers.h"

namespace hwt {

template __global__ void k_price_narrow<double>(const OptConst *__restrict__, const double *__restrict__, const int32_t *__restrict__,
                                             int32_t *__restrict__, double *__restrict__, int32_t *__restrict__);
template __global__ void k_price_narrow<float>(const OptConst *__restrict__, const double *__restrict__, const int32_t *__restrict__,
                                             int32_t *__restrict__, double *__restrict__, int32_t *__restrict__);

} // namespace hwt
