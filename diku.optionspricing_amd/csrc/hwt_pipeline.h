// hwt_pipeline.h -- what the host side of the pricer and the pricing-kernel
// translation units share: the device views of the batch, the task classes
// and bucket layout, the counter slots, and the pricing kernels' signatures.
// The kernels are defined in hwt_pricers.h and instantiated one per
// translation unit (hwt_k_*.hip), so the library's slowest code compiles in
// parallel.
#pragma once

#include "../../include/hwtree.h"
#include "hwt_flat.h"
#include "hwt_rows.h"
#include "hwt_wave.h"

namespace hwt {

struct DevOpts {
    int32_t n;  // options, or the capacity bound when `count` is set
    const double *strike, *maturity, *length;
    const uint16_t *term_unit, *term_steps;
    const double *a, *sigma;
    const int8_t *type;
    const int32_t *count;  // device-side option count (<= n), or null
    // per-option tree constants given by the caller instead of the fields
    // above (which are then null), or null
    const hwt_option_constants_f64 *consts;
    // with consts: the constants are a float caller's, widened; the exercise
    // step is then (int)(t / dt) in float arithmetic
    int32_t consts_f32;
};

struct DevYield {
    int32_t n;
    const double *rate;
    const int32_t *day;
};

// Task classes, in task-list order:
//   wide  (w > 64*KSMALL)           k_price_wide, one option per wave, LPT queue
//   mid   (ROW_MAX_WIDTH < w <= 64*KSMALL)  k_price_narrow, one option per wave
//   row   (w <= ROW_MAX_WIDTH)      k_price_narrow, four options per wave
//   flat  (w <= FLAT_MAX_WIDTH)     k_price_narrow, 64 / w options per wave, one
//                                   node per lane (hwt_flat.h)
// Bucket ranges: wide [B_WIDE, NBUCKET) and mid [B_MID, B_WIDE), log-spaced
// by the sort key (work unless a width/height order is asked for); row
// [B_ROW, B_MID), by (nodes per lane, height) so the four rows of a wave share
// their step count; flat [0, B_ROW), by (width, height).  The scan walks each
// class from its largest key down.
constexpr int NBUCKET = 2048;
constexpr int B_ROW = 640, B_MID = 1280, B_WIDE = 1536;
constexpr int ROW_HBUCKETS = 160;   // height buckets per row part (nodes per lane 1..4)
constexpr int FLAT_HBUCKETS = 64;   // height buckets per flat part (one width each)
constexpr int MAX_PADS = 128;       // idle task slots padding row and flat parts, at most
constexpr int KSMALL = 2;  // trees up to 64*KSMALL nodes wide go to the narrow-tree kernel
#ifndef WIDE_WINDOW_STEP
#define WIDE_WINDOW_STEP 4  // nodes per lane a wide tree's window grows or shrinks by
#endif
// k_setup's class switches
constexpr int CLASS_ROWS = 1, CLASS_FLAT = 2;
enum Tally {
    T_QUEUE = 0,    // next task of the wide-tree kernel (tasks [0, T_NLARGE))
    T_QUEUE_S = 1,  // (unused: the narrow-tree kernel strides statically)
    T_NTASK = 2,
    T_NLARGE = 3,
    T_INVALID = 4,
    T_UNSUP = 5,
    T_WORK = 6,     // int64
    T_SEQWORK = 8,  // int64
    T_WORKW = 10,   // int64, wide trees only
    T_NMID = 12,    // end of the mid-class tasks
    T_NPAD = 13,    // idle slots padding the row class
    T_NCOUNT = 14,  // end of the execution counts (slots 16-23 are k_setup's accumulators)
    T_QUEUE_R = 15, // next chunk of row groups (fused kernel)
    // accumulated by k_setup's blocks, then moved into the slots above and
    // cleared by k_order's writer block
    A_WORK = 16,    // int64
    A_SEQWORK = 18, // int64
    A_WORKW = 20,   // int64
    A_INVALID = 22,
    A_UNSUP = 23,
    T_NREDO = 24,   // mixed mode: options handed to the FP64 re-pricing (host-visible)
    T_QUEUE_X = 25, // next task of the FP64 re-pricing kernel
    T_NROW = 26,    // end of the row-class tasks (start of the flat class)
    T_QUEUE_F = 27, // next chunk of flat groups (fused kernel)
    T_NFLATG = 28,  // flat groups in all
    // first task and first group of each flat part (one width each, widest
    // first, as the scan lays them out), then the end of the last part
    T_FLAT_START = 32,  // FLAT_PARTS + 1 entries
    // int64, device wall clock (constant rate) of the last execution at
    // k_setup's start (its first block)
    T_T0 = 44,
    T_FLAT_GROUP = 48,  // FLAT_PARTS + 1 entries
    // int64: what the windowed kernels' stencil steps executed (StepCount):
    // operations, unit-form node slots, plain-form node slots
    T_XFLOP = 64,
    T_XUNIT = 66,
    T_XPLAIN = 68,
    // int64 x N_END: the latest end of a pricing wavefront, wavefront w in
    // slot w % N_END (one address for all of them serialised 4,096 atomics
    // into 37 us); cleared by k_order's writer block, the maximum taken at sync
    T_END = 72,
    T_HOST = T_END + 2 * 256,  // [0, T_HOST) is copied to the host by hwt_plan_sync
    // the bucket histogram (k_setup's atomics), cleared by the first pricing
    // kernel of each execution
    T_HIST = T_HOST,
    T_ALLOC = T_HIST + NBUCKET
};
constexpr int N_END = (T_HOST - T_END) / 2;

// The bucket histogram (in the tally) back to zero for the next execution:
// its last readers, k_order's blocks, are done once a pricing kernel runs.
__device__ __forceinline__ void clear_bucket_counts(int32_t *tally)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NBUCKET; i += gridDim.x * blockDim.x)
        tally[T_HIST + i] = 0;
}

// The execution's latest pricing-wavefront end so far, for
// hwt_stats.total_ms: one 64-bit atomic per wavefront, as it exits, into the
// wavefront's end slot.
__device__ __forceinline__ void stamp_wave_end(int32_t *tally)
{
    const int w = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (threadIdx.x % WAVE == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(tally + T_END) + w % N_END,
                  (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

// Pricing kernels (hwt_pricers.h).  All are persistent: one single-wave
// workgroup per resident slot, tasks from the ordered task list.
template <class T>
__global__ void k_price_wide(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                             const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                             double *__restrict__ out, double *__restrict__ scratch, int32_t *__restrict__ redo);
template <class T>
__global__ void k_price_narrow(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                               const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                               double *__restrict__ out, int32_t *__restrict__ redo);
template <class T>
__global__ void k_price_fused(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                              const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                              double *__restrict__ out, double *__restrict__ scratch, int32_t *__restrict__ redo);
__global__ void k_price_redo(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                             const int32_t *__restrict__ redo, int32_t *__restrict__ tally,
                             double *__restrict__ out, double *__restrict__ scratch);

} // namespace hwt
