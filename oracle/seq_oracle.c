/*
 * seq_oracle.c -- CPU restatement of the reference's sequential FP64 pricer.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); never linked
 * into the product.  Each function cites the reference lines it follows and
 * keeps their IEEE double operation order, so that with -ffp-contract=off the
 * results are bit-equal to the reference seq binary built with REAL=64.
 */
#include "seq_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* common/Real.hpp:23-32 constants, as doubles (REAL=64 build). */
#define R_ONE 1.0
#define R_TWO 2.0
#define R_HALF 0.5
#define R_THREE 3.0
#define R_SIX 6.0
#define R_SEVEN 7.0
#define R_YEAR 365.0
#define R_MINUS184 (-0.184)

/* common/OptionConstants.hpp:29-53 */
void orc_option_constants(double strike, double maturity, double length,
                          uint16_t termUnit, uint16_t termStepCount,
                          double a, double sigma, int8_t type,
                          orc_constants *c)
{
    c->termUnit = termUnit;
    const double T = maturity;
    const int termUnitsInYearCount = (int)ceil(R_YEAR / termUnit);
    /* uint16 * int -> int, then * double, truncated on assignment to int */
    c->n = (int32_t)((int)termStepCount * termUnitsInYearCount * T);
    c->t = length;
    c->dt = termUnitsInYearCount / (double)termStepCount;
    c->type = type;
    c->X = strike;
    const double V = sigma * sigma * (R_ONE - exp(-R_TWO * a * c->dt)) / (R_TWO * a);
    c->dr = sqrt(R_THREE * V);
    c->M = exp(-a * c->dt) - R_ONE;
    c->jmax = (int32_t)(R_MINUS184 / c->M) + 1;
    c->width = 2 * c->jmax + 1;
}

/* common/OptionConstants.hpp:29-53 in the reference's default build, where
 * real = float (common/Real.hpp:17-21): float arithmetic, except that the
 * unqualified exp, sqrt and ceil there resolve to the C library's double
 * functions (<cmath> declares only those in the global namespace), whose
 * results are rounded to float where they are stored in the float fields. */
void orc_option_constants_f32(float strike, float maturity, float length,
                              uint16_t termUnit, uint16_t termStepCount,
                              float a, float sigma, int8_t type,
                              orc_constants *c)
{
    c->termUnit = termUnit;
    const float T = maturity;
    const int termUnitsInYearCount = (int)ceil((double)((float)R_YEAR / (float)termUnit));
    /* uint16 * int -> int, then * float, truncated on assignment to int */
    c->n = (int32_t)((float)((int)termStepCount * termUnitsInYearCount) * T);
    c->t = length;
    const float dt = (float)termUnitsInYearCount / (float)termStepCount;
    c->dt = dt;
    c->type = type;
    c->X = strike;
    /* sigma*sigma (float) * (one - exp(...)) (double) / (two*a) (float) */
    const double V = (double)(sigma * sigma) * (1.0 - exp((double)(-2.0f * a * dt))) / (double)(2.0f * a);
    const float dr = (float)sqrt(3.0 * V);
    const float M = (float)(exp((double)(-a * dt)) - 1.0);
    c->dr = dr;
    c->M = M;
    c->jmax = (int32_t)(-0.184f / M) + 1;
    c->width = 2 * c->jmax + 1;
}

/* common/Domain.hpp:11-46 */
double orc_yield_at_year(double t, int termUnit, const double *prices,
                         const int32_t *timeSteps, int size)
{
    const int tDays = (int)round(t * termUnit);
    int first = 0, second = 0;
    for (int i = 0; i < size; ++i) {
        const int ts = timeSteps[i];
        if (ts >= tDays) {
            second = i;
            break;
        }
        first = i;
    }
    if (first == second)
        return prices[0];
    if (first == size - 1)
        return prices[size - 1];
    const int t1 = timeSteps[first], t2 = timeSteps[second];
    const double p1 = prices[first], p2 = prices[second];
    const double coefficient = (tDays - t1) / (double)(t2 - t1);
    return p1 + coefficient * (p2 - p1);
}

/* common/Domain.hpp:50-96 (Hull-White 1996 exhibits 1A/1B/1C). */
double orc_prob(char branch, int which, int j, double M)
{
    const double jjMM = (double)(j * j) * M * M;
    const double jM = (double)j * M;
    switch (branch) {
    case 'A':
        if (which == 1) return R_ONE / R_SIX + (jjMM + jM) * R_HALF;
        if (which == 2) return R_TWO / R_THREE - jjMM;
        if (which == 3) return R_ONE / R_SIX + (jjMM - jM) * R_HALF;
        break;
    case 'B':
        if (which == 1) return R_ONE / R_SIX + (jjMM - jM) * R_HALF;
        if (which == 2) return -R_ONE / R_THREE - jjMM + (R_TWO * j) * M;
        if (which == 3) return R_SEVEN / R_SIX + (jjMM - (R_THREE * j) * M) * R_HALF;
        break;
    case 'C':
        if (which == 1) return R_SEVEN / R_SIX + (jjMM + (R_THREE * j) * M) * R_HALF;
        if (which == 2) return -R_ONE / R_THREE - jjMM - (R_TWO * j) * M;
        if (which == 3) return R_ONE / R_SIX + (jjMM + jM) * R_HALF;
        break;
    }
    return 0.0;
}

/* common/Domain.hpp:98-104 */
double orc_compute_alpha(double aggregatedQs, int i, double dt, int termUnit,
                         const double *prices, const int32_t *timeSteps, int size)
{
    const double ti = (i + 2) * dt;
    const double R = orc_yield_at_year(ti, termUnit, prices, timeSteps, size);
    const double P = exp(-R * ti);
    return log(aggregatedQs / P) / dt;
}

/* common/Domain.hpp:147-169 (std::max(a, b) == (a < b) ? b : a). */
double orc_call_value(int isMaturity, double X, int8_t type, double res)
{
    double ret = res;
    if (isMaturity) {
        if (type == 1) {            /* PUT */
            const double v = X - res;
            ret = (v < 0.0) ? 0.0 : v;
        } else if (type == 0) {     /* CALL */
            const double v = res - X;
            ret = (v < 0.0) ? 0.0 : v;
        }
    }
    if (ret == INFINITY)
        return DBL_MAX;
    else if (ret == -INFINITY)
        return DBL_MIN;
    return ret;
}

typedef struct {
    double rate, pu, pm, pd;
} jvalue;

static double price_single(const orc_constants *c, const double *yprices, const int32_t *ydays, int ysize,
                           double *alphas_out, int f32_step);

/* seq/Seq.hpp:24-179 */
double orc_price_single(const orc_constants *c, const double *yprices,
                        const int32_t *ydays, int ysize, double *alphas_out)
{
    return price_single(c, yprices, ydays, ysize, alphas_out, 0);
}

/* seq/Seq.hpp:24-179 on the tree of a real = float build's constants
 * (orc_option_constants_f32): the exercise test i == (int)(t / dt) of
 * seq/Seq.hpp:125 is then evaluated in float; everything else in double. */
double orc_price_single_f32step(const orc_constants *c, const double *yprices,
                                const int32_t *ydays, int ysize)
{
    return price_single(c, yprices, ydays, ysize, NULL, 1);
}

static double price_single(const orc_constants *c, const double *yprices, const int32_t *ydays, int ysize,
                           double *alphas_out, int f32_step)
{
    const int width = c->width, jmax = c->jmax, n = c->n;
    const int jmin = -jmax;
    jvalue *jv = (jvalue *)malloc(sizeof(jvalue) * (size_t)width);

    /* seq/Seq.hpp:27-50: per-j rates and probabilities */
    jv[0].rate = jmin * c->dr;
    jv[0].pu = orc_prob('B', 1, jmin, c->M);
    jv[0].pm = orc_prob('B', 2, jmin, c->M);
    jv[0].pd = orc_prob('B', 3, jmin, c->M);
    jv[width - 1].rate = jmax * c->dr;
    jv[width - 1].pu = orc_prob('C', 1, jmax, c->M);
    jv[width - 1].pm = orc_prob('C', 2, jmax, c->M);
    jv[width - 1].pd = orc_prob('C', 3, jmax, c->M);
    for (int i = 1; i < width - 1; ++i) {
        const int j = i + jmin;
        jv[i].rate = j * c->dr;
        jv[i].pu = orc_prob('A', 1, j, c->M);
        jv[i].pm = orc_prob('A', 2, j, c->M);
        jv[i].pd = orc_prob('A', 3, j, c->M);
    }

    double *Qs = (double *)calloc((size_t)width, sizeof(double));
    double *QsCopy = (double *)calloc((size_t)width, sizeof(double));
    double *alphas = (double *)calloc((size_t)n + 1, sizeof(double));
    Qs[jmax] = R_ONE;
    alphas[0] = orc_yield_at_year(c->dt, c->termUnit, yprices, ydays, ysize);

    /* seq/Seq.hpp:60-113: forward induction (scatter form) */
    for (int i = 0; i < n; ++i) {
        const int jhigh = i < jmax ? i : jmax;
        const double alpha = alphas[i];
        for (int j = -jhigh; j <= jhigh; ++j) {
            const int jind = j - jmin;
            const jvalue jval = jv[jind];
            const double qexp = Qs[jind] * exp(-(alpha + jval.rate) * c->dt);
            if (j == jmin) {
                QsCopy[jind + 2] += jval.pu * qexp;
                QsCopy[jind + 1] += jval.pm * qexp;
                QsCopy[jind] += jval.pd * qexp;
            } else if (j == jmax) {
                QsCopy[jind] += jval.pu * qexp;
                QsCopy[jind - 1] += jval.pm * qexp;
                QsCopy[jind - 2] += jval.pd * qexp;
            } else {
                QsCopy[jind + 1] += jval.pu * qexp;
                QsCopy[jind] += jval.pm * qexp;
                QsCopy[jind - 1] += jval.pd * qexp;
            }
        }
        const int jhigh1 = (i + 1) < jmax ? (i + 1) : jmax;
        double alpha_val = 0;
        for (int j = -jhigh1; j <= jhigh1; ++j) {
            const int jind = j - jmin;
            alpha_val += QsCopy[jind] * exp(-jv[jind].rate * c->dt);
        }
        alphas[i + 1] = orc_compute_alpha(alpha_val, i, c->dt, c->termUnit, yprices, ydays, ysize);
        double *QsT = Qs;
        Qs = QsCopy;
        QsCopy = QsT;
        memset(QsCopy, 0, sizeof(double) * (size_t)width);
    }

    /* seq/Seq.hpp:115-171: backward induction */
    double *call = Qs, *callCopy = QsCopy;
    for (int k = 0; k < width; ++k)
        call[k] = 100.0;
    for (int i = n - 1; i >= 0; --i) {
        const int jhigh = i < jmax ? i : jmax;
        const double alpha = alphas[i];
        const int isMaturity = i == (f32_step ? (int)((float)c->t / (float)c->dt) : (int)(c->t / c->dt));
        for (int j = -jhigh; j <= jhigh; ++j) {
            const int jind = j - jmin;
            const jvalue jval = jv[jind];
            const double callExp = exp(-(alpha + jval.rate) * c->dt);
            double res;
            if (j == jmax)
                res = (jval.pu * call[jind] + jval.pm * call[jind - 1] + jval.pd * call[jind - 2]) * callExp;
            else if (j == jmin)
                res = (jval.pu * call[jind + 2] + jval.pm * call[jind + 1] + jval.pd * call[jind]) * callExp;
            else
                res = (jval.pu * call[jind + 1] + jval.pm * call[jind] + jval.pd * call[jind - 1]) * callExp;
            callCopy[jind] = orc_call_value(isMaturity, c->X, c->type, res);
        }
        double *callT = call;
        call = callCopy;
        callCopy = callT;
        memset(callCopy, 0, sizeof(double) * (size_t)width);
    }

    const double result = call[jmax];
    if (alphas_out)
        memcpy(alphas_out, alphas, sizeof(double) * ((size_t)n + 1));
    free(jv);
    free(alphas);
    free(Qs);
    free(QsCopy);
    return result;
}

/* seq/Seq.hpp:181-190 */
void orc_price_batch(int n, const double *strike, const double *maturity,
                     const double *length, const uint16_t *termUnit,
                     const uint16_t *termSteps, const double *a,
                     const double *sigma, const int8_t *type,
                     const double *yprices, const int32_t *ydays, int ysize,
                     double *out, int threads)
{
#ifdef _OPENMP
    if (threads < 1)
        threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int i = 0; i < n; ++i) {
        orc_constants c;
        orc_option_constants(strike[i], maturity[i], length[i], termUnit[i], termSteps[i],
                             a[i], sigma[i], type[i], &c);
        out[i] = orc_price_single(&c, yprices, ydays, ysize, NULL);
    }
    (void)threads;
}

void orc_node_updates(const orc_constants *c, int64_t *fwd, int64_t *bwd)
{
    int64_t f = 0, b = 0;
    for (int i = 1; i <= c->n; ++i)
        f += 2 * (int64_t)(i < c->jmax ? i : c->jmax) + 1;
    for (int i = 0; i < c->n; ++i)
        b += 2 * (int64_t)(i < c->jmax ? i : c->jmax) + 1;
    *fwd = f;
    *bwd = b;
}
