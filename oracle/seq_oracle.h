/*
 * seq_oracle.h -- CPU restatement of the reference's sequential FP64 pricer.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C ABI, the C++ drop-in headers) may include, link or call this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker / CPU baseline.
 *
 * Every function follows the reference operation-for-operation in IEEE
 * double (compiled with -ffp-contract=off), so that its output is bit-equal
 * to the reference seq binary built with REAL=64 (see oracle/Makefile, which
 * builds that binary into oracle/_ref/ when the reference sources exist).
 *
 * Parity pins (all in tests/test_oracle_*.py):
 *   orc_constants        -- reference width table (data-generator), closed forms
 *   orc_yield_at_year    -- hand cases: before/on/between/after knots
 *   orc_prob_*           -- rows sum to 1, Hull-White moment matching
 *   orc_price_single     -- golden files data/out/<set>.out (FP64 seq output),
 *                           thesis book KATs, ZCB identity, put-call parity,
 *                           the compiled reference binary
 *   orc_alphas           -- sum_j Q_{m+1}[j] e^{-j dr dt} = P((m+2) dt) identity
 *   orc_option_constants_f32, orc_price_single_f32step
 *                        -- the reference's FP32 golden files data/out32 and
 *                           its FP32 seq binary (oracle/_ref/Seq32), at the
 *                           reference's FP32 comparison (test/Compare.hpp)
 */
#ifndef HWTREE_SEQ_ORACLE_H
#define HWTREE_SEQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* common/OptionConstants.hpp:14-27 (field set), :29-53 (computation). */
typedef struct {
    double t;      /* option expiry (Lengths[i]) in years */
    double dt;     /* years per step */
    double dr;
    double X;
    double M;
    int32_t jmax;
    int32_t n;     /* tree height in steps (Maturities[i] drives it) */
    int32_t width;
    uint16_t termUnit;
    int8_t type;   /* CALL = 0, PUT = 1 (common/Options.hpp:24-28) */
} orc_constants;

/* common/OptionConstants.hpp:29-53 */
void orc_option_constants(double strike, double maturity, double length,
                          uint16_t termUnit, uint16_t termStepCount,
                          double a, double sigma, int8_t type,
                          orc_constants *out);

/* The same constants as the reference's default build (real = float)
 * derives them: float arithmetic with the C library's double exp, sqrt and
 * ceil, stored widened. */
void orc_option_constants_f32(float strike, float maturity, float length,
                              uint16_t termUnit, uint16_t termStepCount,
                              float a, float sigma, int8_t type,
                              orc_constants *out);

/* common/Domain.hpp:11-46 */
double orc_yield_at_year(double t, int termUnit, const double *prices,
                         const int32_t *timeSteps, int size);

/* common/Domain.hpp:50-96; which: 1 = up, 2 = mid, 3 = down;
 * branch: 'A' standard, 'B' bottom (j == -jmax), 'C' top (j == jmax). */
double orc_prob(char branch, int which, int j, double M);

/* common/Domain.hpp:98-104 */
double orc_compute_alpha(double aggregatedQs, int i, double dt, int termUnit,
                         const double *prices, const int32_t *timeSteps, int size);

/* common/Domain.hpp:147-169 */
double orc_call_value(int isMaturity, double X, int8_t type, double res);

/* seq/Seq.hpp:24-179.  alphas_out (nullable) receives alphas[0..n]. */
double orc_price_single(const orc_constants *c, const double *yprices,
                        const int32_t *ydays, int ysize, double *alphas_out);

/* seq/Seq.hpp:24-179 evaluated in double on the tree of float-build
 * constants, with the exercise step (int)(t / dt) taken in float. */
double orc_price_single_f32step(const orc_constants *c, const double *yprices,
                                const int32_t *ydays, int ysize);

/* seq/Seq.hpp:181-190 (OpenMP parallel-for over options when threads > 1). */
void orc_price_batch(int n, const double *strike, const double *maturity,
                     const double *length, const uint16_t *termUnit,
                     const uint16_t *termSteps, const double *a,
                     const double *sigma, const int8_t *type,
                     const double *yprices, const int32_t *ydays, int ysize,
                     double *out, int threads);

/* Node-update counts of seq's two passes for one option:
 * forward sum_{i=1..n} (2 min(i,jmax)+1), backward sum_{i=0..n-1} (2 min(i,jmax)+1). */
void orc_node_updates(const orc_constants *c, int64_t *fwd, int64_t *bwd);

#ifdef __cplusplus
}
#endif

#endif
