/*
 * hwtsurr.h -- C ABI of the learned surrogate in front of the trinomial-tree
 * pricer: an ensemble of small MLPs trained on prices the FP64 tree itself
 * produces, with the FP64 tree as fallback.
 *
 * The reference has no counterpart (it prices every option on the tree,
 * seq/Seq.hpp:181-190); this is an addition beside the drop-in entry points of
 * hwtree.h, and hwt_surr_price has the same contract as hwt_plan_execute:
 * device SoA options in, prices in input order out.
 *
 * Network (per member): 8 inputs -> 4 hidden layers of 256 (SiLU) -> 1 output,
 * 199,937 parameters.  The three 256x256 layers run on BF16 MFMA with FP32
 * accumulation; the input layer, the output layer and every master weight
 * stay FP32.  The output is the residual of the tree price over the
 * closed-form Hull-White bond-option price on the tree's own dates, in units
 * of HWT_SURR_RESIDUAL_SCALE.
 *
 * Inputs: (X-80)/20, (T-15.5)/14.5, (S-T-5.5)/4.5, (steps-24)/12,
 * (a-0.105)/0.095, (sigma-0.0125)/0.0075, sin 2 pi u, cos 2 pi u, where T and
 * S are the exercise and bond dates on the tree grid and u locates the
 * strike on the tree's node lattice at the exercise step.
 *
 * An option is priced by the tree instead when it lies outside the training
 * box (term unit 365; 12 or 36 steps/yr; strike [60,100]; expiry [1,30] y;
 * bond maturity expiry + [1,10] y; a [0.01,0.2]; sigma [0.005,0.02]) or when
 * the ensemble's spread (max - min of the members' prices) exceeds the
 * configured threshold.
 *
 * Return codes are hwtree.h's HWT_* codes.
 */
#ifndef HWTSURR_H
#define HWTSURR_H

#include "hwtree.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HWT_SURR_INPUTS   8
#define HWT_SURR_HIDDEN   256
#define HWT_SURR_MAX_MEMBERS 4
#define HWT_SURR_NPARAMS  199937     /* per member */
#define HWT_SURR_RESIDUAL_SCALE 0.01 /* network output unit, price per 100 notional */

/* Offsets of each tensor in a member's FP32 parameter vector (row-major,
 * W[out][in]). */
#define HWT_SURR_W1 0        /* 256 x 8   */
#define HWT_SURR_B1 2048     /* 256       */
#define HWT_SURR_W2 2304     /* 256 x 256 */
#define HWT_SURR_B2 67840
#define HWT_SURR_W3 68096
#define HWT_SURR_B3 133632
#define HWT_SURR_W4 133888
#define HWT_SURR_B4 199424
#define HWT_SURR_W5 199680   /* 1 x 256   */
#define HWT_SURR_B5 199936   /* 1         */

typedef struct {
    int32_t members;          /* ensemble size, 1..4 (default 4) */
    int32_t batch;            /* training rows per step, multiple of 128 (default 8192) */
    float   lr_max;           /* peak Adam step size (default 2e-3) */
    float   lr_min;           /* final step size of the cosine decay (default 2e-5) */
    float   warmup_frac;      /* share of steps with linear warm-up (default 0.02) */
    float   beta1, beta2, eps;/* Adam (defaults 0.9, 0.999, 1e-8) */
    float   spread_threshold; /* fallback when max-min of member prices exceeds it (default 1e-3) */
    int32_t reserved0;
    uint64_t seed;            /* initial weights and data order */
    int32_t reserved[6];
} hwt_surr_config;

typedef struct {
    double  ms;               /* wall time of the call (HIP events) */
    double  loss;             /* mean squared error of the last epoch (normalised units) */
    int64_t steps;            /* optimiser steps taken so far */
    int32_t fallback;         /* options re-priced by the tree (hwt_surr_price) */
    int32_t out_of_box;       /* ... of which outside the training box */
    double  surrogate_ms;     /* box split + features + ensemble inference + gate (hwt_surr_price;
                                 read from the device's wall clock at the pass's marks, since the
                                 pass runs as one graph) */
    double  tree_ms;          /* fallback tree pricing (hwt_surr_price; the same way) */
    int32_t reserved[4];
} hwt_surr_stats;

typedef struct hwt_surrogate hwt_surrogate;

void hwt_surr_config_init(hwt_surr_config *cfg);

/* train_capacity: most training samples; infer_capacity: most options per
 * hwt_surr_price call; yield_capacity: most yield knots. */
int hwt_surr_create(hwt_surrogate **s, int32_t train_capacity, int32_t infer_capacity,
                    int32_t yield_capacity, const hwt_surr_config *cfg);
void hwt_surr_destroy(hwt_surrogate *s);

/* Training set: device options (all inside the box), the yield curve and their
 * FP64 tree prices (device).  Computes the inputs and residual targets. */
int hwt_surr_set_training_data(hwt_surrogate *s, const hwt_options_f64 *dev_opts,
                               const hwt_yield_f64 *dev_yield, const double *dev_prices);
/* Runs `epochs` epochs of Adam over the training set; the learning-rate
 * schedule spans total_epochs (>= epochs already run + epochs). */
int hwt_surr_train(hwt_surrogate *s, int32_t epochs, int32_t total_epochs, hwt_surr_stats *stats);

/* Ensemble prediction without fallback, for inspection: per option the
 * ensemble-mean price, the spread (max - min over members) and whether the
 * option is inside the box.  All device pointers. */
int hwt_surr_predict(hwt_surrogate *s, const hwt_options_f64 *dev_opts, const hwt_yield_f64 *dev_yield,
                     double *dev_mean, double *dev_spread, int8_t *dev_in_box);

/* Surrogate with fallback: prices in input order into dev_out.  A first
 * kernel splits the batch: options outside the box go to the fallback list,
 * the ensemble runs on the in-box options only, and those whose spread
 * exceeds the threshold join the list; the list is then compacted and priced
 * by the FP64 tree in one pass, in the same stream.  Counts stay on the
 * device (no host round trip).  `tree` is a plan of capacity >=
 * infer_capacity created with HWT_FLAG_NO_GRAPH. */
int hwt_surr_price(hwt_surrogate *s, hwt_plan *tree, const hwt_options_f64 *dev_opts,
                   const hwt_yield_f64 *dev_yield, double *dev_out, hwt_surr_stats *stats);

/* Parameters (FP32 masters), host arrays of members x HWT_SURR_NPARAMS. */
int hwt_surr_get_params(hwt_surrogate *s, float *host_params);
int hwt_surr_set_params(hwt_surrogate *s, const float *host_params);

/* Test hooks.  Forward and backward of one member on training rows (host
 * indices, n a multiple of 128 and <= batch): network outputs (normalised),
 * the mean loss, and the gradient of the loss (host, HWT_SURR_NPARAMS), with
 * no parameter update.  Any output pointer may be NULL. */
int hwt_surr_eval_rows(hwt_surrogate *s, int32_t member, const int32_t *host_rows, int32_t n,
                       float *host_out, float *host_loss, float *host_grad);
/* Inputs and residual targets of training rows (host: n x 8 and n). */
int hwt_surr_get_inputs(hwt_surrogate *s, const int32_t *host_rows, int32_t n, float *host_x,
                        float *host_target);
/* One Adam update of every member with the given gradients (host, members x
 * HWT_SURR_NPARAMS) at step size lr. */
int hwt_surr_adam_step(hwt_surrogate *s, const float *host_grad, float lr);

/* Checkpoint: parameters, Adam moments, step count and data-order RNG state. */
int hwt_surr_save(hwt_surrogate *s, const char *path);
int hwt_surr_load(hwt_surrogate *s, const char *path);

#ifdef __cplusplus
}
#endif

#endif /* HWTSURR_H */
