/*
 * hwtree.h -- C ABI of the MI355X-native Hull-White one-factor trinomial-tree
 * pricer (European options on zero-coupon bonds).
 *
 * This is the drop-in boundary for the reference's price-vector entry points:
 *
 *   void seq::computeOptions(const Options&, const Yield&, std::vector<real>&)
 *        -- reference seq/Seq.hpp:181-190
 *   void cuda::option::KernelRunBase::run(const Options&, const Yield&,
 *        std::vector<real>&, int blockSize, SortType, bool isTest)
 *        -- reference cuda-option/Kernel.cuh:303-314
 *   void cuda::multi::KernelRunBase::run(...same...)
 *        -- reference cuda-multi/Kernel.cuh:374-403
 *
 * Inputs are the reference's structure-of-arrays trinom::Options
 * (common/Options.hpp:48-58, same fields, order and element types) and
 * trinom::Yield (common/Yield.hpp:14-18).  Prices come back in input order.
 * Nothing here throws; every call returns 0 or a negative HWT_E* code.
 * The caller owns every buffer.  A plan is bound to one device and is not
 * safe to share between host threads.
 */
#ifndef HWTREE_H
#define HWTREE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HWT_ABI_VERSION 2

/* Return codes. */
#define HWT_OK            0
#define HWT_EINVAL       -1  /* bad argument: null pointer, n < 0, bad type, a <= 0, ... */
#define HWT_EUNSUPPORTED -2  /* valid option the selected kernel form cannot price (e.g. width) */
#define HWT_ENOMEM       -3  /* device or host allocation failed, or capacity exceeded */
#define HWT_EHIP         -4  /* HIP runtime error */

/* Kernel forms (the reference's two CUDA runners plus an inspector). */
#define HWT_FORM_AUTO 0  /* the fastest form for the batch: currently the wave form */
#define HWT_FORM_LANE 1  /* form 1: one lane per option, columns interleaved in HBM
                            (replaces cuda-option, reference cuda-option/Kernel.cuh:65-216) */
#define HWT_FORM_WAVE 2  /* form 2: columns on chip, many options per workgroup
                            (replaces cuda-multi, reference cuda-multi/Kernel.cuh:60-292) */

/* Precision modes. */
#define HWT_PREC_FP64  0 /* all tree arithmetic in FP64 */
#define HWT_PREC_MIXED 1 /* Q propagation and backward induction in FP32; alpha
                            calibration sums, logs and yield interpolation in FP64 */

/* Processing-order hint, the reference's SortType (common/Options.hpp:15-22).
 * Results are always returned in input order. */
#define HWT_SORT_NONE        '-'
#define HWT_SORT_WIDTH_ASC   'w'
#define HWT_SORT_WIDTH_DESC  'W'
#define HWT_SORT_HEIGHT_ASC  'h'
#define HWT_SORT_HEIGHT_DESC 'H'
#define HWT_SORT_WORK_DESC   'L'  /* longest-processing-time first (default) */

/* trinom::Options, FP64 (common/Options.hpp:48-58).  Pointers are host
 * pointers for hwt_price_f64 and device pointers for hwt_plan_execute. */
typedef struct {
    int32_t n;
    const double   *strike;      /* StrikePrices   */
    const double   *maturity;    /* Maturities: bond maturity in years, sets the tree height */
    const double   *length;      /* Lengths: option expiry in years, sets the exercise step */
    const uint16_t *term_unit;   /* TermUnits (days per term unit, 365 = yearly) */
    const uint16_t *term_steps;  /* TermStepCounts (steps per term unit) */
    const double   *a;           /* ReversionRates */
    const double   *sigma;       /* Volatilities   */
    const int8_t   *type;        /* Types: CALL = 0, PUT = 1 */
} hwt_options_f64;

/* trinom::Yield (common/Yield.hpp:14-18): continuously compounded zero rates
 * at day offsets, ascending. */
typedef struct {
    int32_t n;
    const double  *rate;         /* Prices    */
    const int32_t *day;          /* TimeSteps */
} hwt_yield_f64;

typedef struct {
    int32_t form;        /* HWT_FORM_*                                  (0 = auto) */
    int32_t precision;   /* HWT_PREC_*                                  (0 = FP64) */
    int32_t block_size;  /* form 1: lanes per workgroup (64..1024); <= 0 picks the
                            default.  Form 2 runs one wavefront per workgroup. */
    int32_t sort;        /* HWT_SORT_*: processing order of the wide and mid-width
                            trees (0 = work-descending) */
    int32_t device;      /* HIP device ordinal for hwt_price_f64; < 0 = current */
    int32_t flags;       /* HWT_FLAG_* */
    int32_t wide_per_cu;   /* form 2: resident workgroups per CU of the persistent
                              wide-tree kernel; 0 = as many as fit */
    int32_t narrow_per_cu; /* form 2: the same for the narrow-tree kernel */
    int32_t lane_max_width;/* form 1: the widest tree the plan prices (its tree columns in
                              HBM are sized for it; wider options are unsupported);
                              0 = hwt_max_width.  The one-shot calls set it from the batch. */
} hwt_config;

/* Launch the plan's kernels one by one instead of replaying its HIP graph;
 * hwt_stats.kernel_ms then times the pricing kernel alone. */
#define HWT_FLAG_NO_GRAPH 1
/* Form 2: price trees up to 64 nodes wide one per wavefront instead of four
 * per wavefront (16-lane rows). */
#define HWT_FLAG_NO_ROWS  2
/* Form 2: run the narrow-tree kernel beside the wide-tree kernel on a second
 * stream instead of after it on the same stream. */
#define HWT_FLAG_CONCURRENT 4
/* Form 2: one persistent kernel for every tree: its wavefronts take the
 * wide trees first, then claim the narrow ones in chunks as they run out, so
 * narrow trees fill the wide trees' tail.  Overrides HWT_FLAG_CONCURRENT. */
#define HWT_FLAG_FUSED 8
/* Form 2: price trees up to 21 nodes wide in 16-lane rows (or one per
 * wavefront with HWT_FLAG_NO_ROWS) instead of side by side across a whole
 * wavefront, one node per lane (the flattened form). */
#define HWT_FLAG_NO_FLAT 16
/* Launch the plan's kernels one by one, like HWT_FLAG_NO_GRAPH but without its
 * timing events, instead of replaying the plan's HIP graph.  Back to back on a
 * large batch the direct launches cost less per execution than a graph replay
 * (0.273 against 0.278 ms on the 60k skewed batch); on small batches the
 * host's launch calls cost more.  The autotuner tries both.  HWT_FLAG_NO_GRAPH
 * takes precedence. */
#define HWT_FLAG_DIRECT 32

typedef struct {
    double  kernel_ms;        /* pricing kernels, first start to last end (HWT_FLAG_NO_GRAPH),
                                 else = total_ms */
    double  total_ms;         /* whole plan execution: constants, ordering, pricing;
                                 excludes H2D/D2H copies.  Device wall clock, stamped
                                 by the kernels: the set-up kernel's first block at
                                 its start to the last pricing wavefront's end (the
                                 first kernel's launch is not included); 0 for an
                                 empty batch. */
    double  wide_ms;          /* wide-tree pricing kernel alone (HWT_FLAG_NO_GRAPH), else 0 */
    double  narrow_ms;        /* narrow-tree pricing kernel alone (HWT_FLAG_NO_GRAPH), else 0 */
    int64_t device_memory;    /* bytes of device memory held by the plan */
    int64_t node_updates;     /* stencil node updates performed (active tree region) */
    int64_t wide_node_updates;/* ... of which by the wide-tree kernel */
    int64_t seq_node_updates; /* node updates the reference's seq would perform (both passes) */
    int32_t tasks;            /* work items scheduled */
    int32_t wide_tasks;       /* ... of which wide trees */
    int32_t unsupported;      /* options that could not be priced (output NaN) */
    int32_t fp64_repriced;    /* HWT_PREC_MIXED: options whose FP32 tree state left FP32's
                                 range (long, weakly mean-reverting bonds) and were priced
                                 again with FP64 states in the same execution */
    int64_t wide_step_ops;    /* arithmetic operations the wide-tree kernel's stencil steps
                                 executed, padding slots included: per lane and step 4K+1 in
                                 the unit-diagonal form (K nodes per lane), 5K in the plain
                                 form (FP64 FLOP in FP64 mode).  Counted by hwt_plan_sync
                                 from the window schedule the kernel walks (a small kernel
                                 after the execution, run only when stats are asked for);
                                 0 for form 1 */
    int64_t wide_unit_slots;  /* node slots it stepped in the unit-diagonal form (64 K per step) */
    int64_t wide_plain_slots; /* ... in the plain form */
} hwt_stats;

/* Fills *cfg with defaults (auto form, FP64, tuned sizes, work-descending order). */
void hwt_config_init(hwt_config *cfg);

/* One-shot pricing of host arrays; replaces seq::computeOptions and
 * KernelRunBase::run.  out_prices (host, n doubles) receives prices in input
 * order.  cfg and stats may be NULL. */
int hwt_price_f64(const hwt_options_f64 *opts, const hwt_yield_f64 *yield,
                  double *out_prices, const hwt_config *cfg, hwt_stats *stats);

/* Reusable plan over device-resident inputs (no allocation per execution;
 * the pipeline is captured into a HIP graph).  capacity = max options per
 * execution, yield_capacity = max yield knots. */
typedef struct hwt_plan hwt_plan;
int  hwt_plan_create(hwt_plan **plan, int32_t capacity, int32_t yield_capacity,
                     const hwt_config *cfg);
/* Enqueue pricing of device arrays on `stream` (a hipStream_t, NULL = the
 * plan's own stream).  Asynchronous: call hwt_plan_sync for the status. */
int  hwt_plan_execute(hwt_plan *plan, const hwt_options_f64 *dev_opts,
                      const hwt_yield_f64 *dev_yield, double *dev_out, void *stream);
/* Wait for the last execution; returns its status (HWT_EUNSUPPORTED if any
 * option could not be priced) and fills stats if non-NULL. */
int  hwt_plan_sync(hwt_plan *plan, hwt_stats *stats);
void hwt_plan_destroy(hwt_plan *plan);

/* trinom::Options and trinom::Yield as a default-built reference caller holds
 * them: `real` is float unless USE_DOUBLE (common/Real.hpp:17-21). */
typedef struct {
    int32_t n;
    const float    *strike;
    const float    *maturity;
    const float    *length;
    const uint16_t *term_unit;
    const uint16_t *term_steps;
    const float    *a;
    const float    *sigma;
    const int8_t   *type;
} hwt_options_f32;

typedef struct {
    int32_t n;
    const float   *rate;
    const int32_t *day;
} hwt_yield_f32;

/* One-shot pricing for FP32 callers; replaces seq::computeOptions and the
 * KernelRunBase::run overloads of the reference's default FP32 build
 * (seq/Seq.hpp:181-190, cuda-option/Kernel.cuh:303-314,
 * cuda-multi/Kernel.cuh:374-403, all with real = float).  Each option's tree
 * is the one that build derives: its constants (height, dt, dr, M, jmax) as
 * common/OptionConstants.hpp:29-53 computes them in float, and the exercise
 * step (int)(t / dt) in float (seq/Seq.hpp:125), which for some step counts
 * is one step earlier than in FP64 (15 steps a year at a 3-year expiry: 44,
 * not 45).  That tree is priced as hwt_price_f64 prices (cfg->precision:
 * HWT_PREC_FP64 by default, or HWT_PREC_MIXED), on the curve's rates widened
 * exactly, and each price is rounded once to float. */
int hwt_price_f32(const hwt_options_f32 *opts, const hwt_yield_f32 *yield, float *out_prices,
                  const hwt_config *cfg, hwt_stats *stats);

/* trinom::OptionConstants (common/OptionConstants.hpp:14-53): an option's
 * tree constants as the reference derives them from an Options row.  Pricing
 * from constants replaces seq::computeSingleOption(const OptionConstants&,
 * const Yield&) (seq/Seq.hpp:24): the exercise step is (int)(t / dt) as in
 * seq/Seq.hpp:125, the tree is width = 2*jmax+1 nodes wide and n steps high,
 * and M, dr and dt are taken as given. */
typedef struct {
    double   t;          /* option expiry in years (the Length field) */
    double   dt;         /* years per step */
    double   dr;         /* rate spacing of the tree */
    double   X;          /* strike */
    double   M;          /* e^{-a dt} - 1 */
    int32_t  jmax;
    int32_t  n;          /* height in steps */
    int32_t  width;      /* must equal 2*jmax + 1 */
    uint16_t term_unit;  /* days per term unit (yield lookups) */
    int8_t   type;       /* CALL = 0, PUT = 1 */
} hwt_option_constants_f64;

/* Prices n options given by their constants (host arrays) into out_prices
 * (host, input order).  cfg and stats may be NULL. */
int hwt_price_constants_f64(const hwt_option_constants_f64 *consts, int32_t n, const hwt_yield_f64 *yield,
                            double *out_prices, const hwt_config *cfg, hwt_stats *stats);

/* The same for constants a default-built (real = float) caller derived: the
 * exercise step is (int)(t / dt) evaluated in float, as the FP32 reference
 * evaluates it; the tree is then priced as hwt_price_f32 prices. */
typedef struct {
    float    t, dt, dr, X, M;
    int32_t  jmax, n, width;
    uint16_t term_unit;
    int8_t   type;
} hwt_option_constants_f32;

int hwt_price_constants_f32(const hwt_option_constants_f32 *consts, int32_t n, const hwt_yield_f32 *yield,
                            float *out_prices, const hwt_config *cfg, hwt_stats *stats);

/* A batch held in device memory across pricing calls: the upload of
 * cuda::CudaOptions (cuda/CudaDomain.cuh:88-135, constructed once, then passed
 * to KernelRunBase::run(CudaOptions&, ...), cuda-option/Kernel.cuh:316-325).
 * hwt_batch_create copies the host arrays to the current device;
 * hwt_batch_price prices the resident batch into host_out (input order), with
 * no re-upload: the batch keeps the plan of its last call and reuses it, and
 * its captured graph, while the settings stay the same (a change of settings
 * builds a new plan).  hwt_batch_device_memory counts the kept plan too.
 * Calls on one batch must not overlap (from several host threads): they
 * share its device output array and its plan. */
typedef struct hwt_batch hwt_batch;
int hwt_batch_create(hwt_batch **batch, const hwt_options_f64 *host_opts, const hwt_yield_f64 *host_yield);
/* The same for a real = float caller's arrays: the batch holds each option's
 * tree as the FP32 build derives it (see hwt_price_f32). */
int hwt_batch_create_f32(hwt_batch **batch, const hwt_options_f32 *host_opts, const hwt_yield_f32 *host_yield);
int hwt_batch_price(hwt_batch *batch, double *host_out, const hwt_config *cfg, hwt_stats *stats);
/* Bytes of device memory the batch holds (inputs and output). */
int64_t hwt_batch_device_memory(const hwt_batch *batch);
void hwt_batch_destroy(hwt_batch *batch);

/* Pricing with no allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  A variant of
 * hwt_plan_execute with the default configuration (form 2, FP64): device
 * arrays in, prices in input order into dev_out, enqueued on `stream` (a
 * hipStream_t, NULL = the default stream) without waiting.  Its kernels'
 * buffers live in one static device array per GPU (about 37 MB, held from the
 * library's load until the process ends) sized for at most
 * HWT_NOALLOC_MAX_OPTIONS options; larger batches return HWT_ENOMEM.  Nothing
 * is allocated and no memset is issued.  Invalid or too-wide options come
 * back NaN; no per-option status is reported.  Calls that overlap on one GPU
 * (concurrent streams or host threads) share the array and race. */
#define HWT_NOALLOC_MAX_OPTIONS 65536
int hwt_price_device_noalloc(const hwt_options_f64 *dev_opts, const hwt_yield_f64 *dev_yield, double *dev_out,
                             void *stream);

/* Per-step alpha calibration of each option (alphas[0..n] of the reference's
 * forward pass, seq/Seq.hpp:56-106), host arrays.  alpha_offsets (n_opts+1
 * int64, host) is filled first when alphas == NULL so the caller can size the
 * buffer; alphas (host, alpha_offsets[n_opts] doubles) then receives the
 * sequences back to back. */
int hwt_alphas_f64(const hwt_options_f64 *opts, const hwt_yield_f64 *yield,
                   int64_t *alpha_offsets, double *alphas);

/* Project autotuner (no counterpart in the reference, whose runners take a
 * block size and sort order from the command line, cuda-option/Kernel.cuh:303
 * and cuda-multi/Kernel.cuh:374).  Prices the device batch repeatedly with
 * candidate settings of form 2 -- processing order of the wide trees (work,
 * height or width), narrow trees four or one per wavefront, the two pricing
 * kernels in sequence, side by side or fused, graph replays or direct
 * launches (HWT_FLAG_DIRECT), and resident workgroups per CU of each -- and
 * then form 1 instead of form 2 for FP64 batches of at most 262,144 options
 * (an inspector choosing the kernel form per batch), by greedy coordinate
 * descent from *cfg (form 0 or 2), timing each candidate with HIP events
 * around bursts of four executions queued back to back (median of seven
 * bursts after five warm-ups; a change must win by 1%), and writes the
 * fastest back into *cfg.
 * The choice is cached per batch signature (precision, option count and the
 * tree-class counts and total work, each to a power of two) for the life of
 * the process, and appended to the file named by the environment variable
 * HWT_TUNE_CACHE when it is set (read at the first call).  dev_out receives
 * the prices of the last run. */
typedef struct {
    double  ms;          /* pipeline time of the chosen configuration */
    double  default_ms;  /* ... of the configuration passed in */
    double  tune_ms;     /* wall time spent tuning (0 on a cache hit) */
    int32_t candidates;  /* configurations timed */
    int32_t cached;      /* 1 if the choice came from the cache */
} hwt_tune_report;

int hwt_autotune(const hwt_options_f64 *dev_opts, const hwt_yield_f64 *dev_yield, double *dev_out,
                 hwt_config *cfg, hwt_tune_report *report);

/* Largest tree width (2*jmax+1) each form supports. */
int32_t hwt_max_width(int32_t form);

const char *hwt_strerror(int code);
/* Detail of the last error on the calling thread ("" if none). */
const char *hwt_last_error(void);
int32_t hwt_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HWTREE_H */
