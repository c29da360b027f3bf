// hwt_pricers.h -- the persistent pricing kernels of form 2 and the per-option
// pricing routines they run (duality sums over the two on-chip tree states).
// Included only by the hwt_k_*.hip translation units, each of which
// instantiates one kernel.
#pragma once

#include "hwt_pipeline.h"

namespace hwt {

// Payoff of the option on a bond worth B at the exercise step, clamped to the
// largest finite value of the propagation type (common/Domain.hpp:163-168).
template <class T>
__device__ __forceinline__ double payoff(const OptConst &cv, double B)
{
    double v = (cv.type == 1) ? cv.X - B : B - cv.X;  // PUT : CALL
    v = (v < 0.0) ? 0.0 : v;
    return (v > RealMax<T>::value) ? RealMax<T>::value : v;
}

// Price from the two states at the exercise step: y (forward, Arrow-Debreu)
// and b (backward, bond value per unit face), both untransformed:
//   S1 = sum y, S2 = sum y b, scale = 100 S1 Pn / (S2 Pe),
//   price = Pe / S1 * sum y payoff(scale b).
// Sums are FP64 (compensated in mixed mode).

template <class T, int KMAX>
__device__ __forceinline__ double price_option(const OptConst &cv, const double *edge, int lane,
                                               typename Lds<T>::type *lds, lds_f64 *etab, T *fscratch)
{
    const TreeShape t = uni(TreeShape{cv.w, cv.jmax, cv.M, cv.dr, cv.dt, cv.rho, cv.rhop});
    const int n = uni(cv.n), e = uni(cv.e);
    load_edge_table(edge, lane, etab);
    int flo, fhi, blo, bhi;
    const bool unit = uni((int)cv.unit) != 0;
    // wide trees change window four nodes per lane at a time (a 60k batch of
    // wide and narrow trees: 283 us with single steps, 252 us with 2, 243 us
    // with 4); the mid class (up to 128 nodes wide) keeps single steps
    constexpr int G = KMAX > KSMALL ? WIDE_WINDOW_STEP : 1;
    forward_pass<T, KMAX, G>(t, e, lane, lds, etab, fscratch, flo, fhi, unit);
    backward_pass<T, KMAX, G>(t, n, e, lane, lds, etab, blo, bhi, unit);

    const int ep = min(e, t.jmax);
    const int klo = t.jmax - ep, khi = t.jmax + ep;
    auto fy = [&](int k) { return (k >= flo && k < fhi) ? (double)fscratch[k] : 0.0; };
    auto fb = [&](int k) { return (k >= blo && k < bhi) ? (double)lds[k] : 0.0; };
    auto yk = [&](int k) {
        double v = fy(k);
        if (k == 1) v -= t.rho * fy(0);
        if (k == t.w - 2) v -= t.rhop * fy(t.w - 1);
        return v;
    };
    auto bk = [&](int k) {
        double v = fb(k);
        if (k == 0) v += t.rho * fb(1);
        if (k == t.w - 1) v += t.rhop * fb(t.w - 2);
        return v;
    };
    typename SumOf<T>::type a1, a2, a3;
    bool big = false;
    for (int k = klo + lane; k <= khi; k += WAVE) {
        const double yv = yk(k), bv = bk(k);
        a1.add(yv);
        a2.add(yv * bv);
        big |= !(fabs(bv) <= F32_SAFE_STATE);
    }
    const double s1 = a1.wave_total(), s2 = a2.wave_total();
    if constexpr (sizeof(T) == sizeof(float)) {
        // FP32 state out of range: NaN asks for the FP64 re-pricing
        if (__ballot(big) != 0 || !isfinite(s1) || !isfinite(s2)) {
            __syncthreads();
            return NAN;
        }
    }
    const double scale = 100.0 * s1 * cv.Pn / (s2 * cv.Pe);
    for (int k = klo + lane; k <= khi; k += WAVE) a3.add(yk(k) * payoff<T>(cv, scale * bk(k)));
    const double s3 = a3.wave_total();
    __syncthreads();
    return cv.Pe / s1 * s3;
}

// Narrow trees (width <= 64*K): the whole tree in one full window, both passes
// in registers, and the duality sums formed straight from the registers.
template <class T, int K>
__device__ __forceinline__ double price_narrow(const OptConst &cv, const double *edge, int lane, lds_f64 *etab)
{
    const TreeShape t = uni(TreeShape{cv.w, cv.jmax, cv.M, cv.dr, cv.dt, cv.rho, cv.rhop});
    const int n = uni(cv.n), e = uni(cv.e);
    load_edge_table(edge, lane, etab);
    T cl[K], cm[K], cr[K], zf[K], zb[K];
    window_coefs<T, K, true>(t, 0, lane, etab, cl, cm, cr);
    init_window<T, K, true>(zf, t, 0, lane);
    stencil_steps<T, K>(zf, cl, cm, cr, e);
    window_coefs<T, K, false>(t, 0, lane, etab, cl, cm, cr);
    init_window<T, K, false>(zb, t, 0, lane);
    stencil_steps<T, K>(zb, cl, cm, cr, n - e);

    // undo the edge similarity: y = T^{-1} zf, b = T^T zb
    const double fl = dpp_from_lower_lane(zf[K - 1]), fr = dpp_from_upper_lane(zf[0]);
    const double bl = dpp_from_lower_lane(zb[K - 1]), br = dpp_from_upper_lane(zb[0]);
    double yv[K], bv[K];
    typename SumOf<T>::type a1, a2, a3;
    bool big = false;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = lane * K + s;
        const double fL = s > 0 ? (double)zf[s - 1] : fl, fR = s + 1 < K ? (double)zf[s + 1] : fr;
        const double bL = s > 0 ? (double)zb[s - 1] : bl, bR = s + 1 < K ? (double)zb[s + 1] : br;
        double y = zf[s], b = zb[s];
        if (k == 1) y -= t.rho * fL;
        if (k == t.w - 2) y -= t.rhop * fR;
        if (k == 0) b += t.rho * bR;
        if (k == t.w - 1) b += t.rhop * bL;
        if (k >= t.w) y = b = 0.0;
        yv[s] = y;
        bv[s] = b;
        a1.add(y);
        a2.add(y * b);
        big |= !(fabs(b) <= F32_SAFE_STATE);
    }
    const double s1 = a1.wave_total(), s2 = a2.wave_total();
    if constexpr (sizeof(T) == sizeof(float)) {
        if (__ballot(big) != 0 || !isfinite(s1) || !isfinite(s2)) return NAN;  // FP64 re-pricing
    }
    const double scale = 100.0 * s1 * cv.Pn / (s2 * cv.Pe);
#pragma unroll
    for (int s = 0; s < K; ++s) a3.add(yv[s] * payoff<T>(cv, scale * bv[s]));
    const double s3 = a3.wave_total();
    return cv.Pe / s1 * s3;
}

// Persistent pricing kernels: one single-wave workgroup per resident slot;
// the waves work through tasks [first, last) of the task list.
//   DYNAMIC (wide trees): each wave takes task first + blockIdx.x, then claims
//     further tasks from the queue counter, so the long trees at the head of
//     the list balance themselves.  Lane 0 stores the finished price and
//     claims the next task in one divergent region; every lane then reads the
//     claim back through a convergent broadcast, so all 64 lanes always leave
//     the loop together.
//   static (narrow trees): wave b takes tasks first + b, first + b + grid, ...
//     Narrow tasks are many and short, and one counter bumped once per task
//     by every wave serialises in L2; the list is in decreasing work order, so
//     a fixed stride spreads the work evenly.
// A finished price into out[o], or -- a NaN from an FP32 pricer, whose state
// left FP32's range -- option o onto the list the FP64 kernel re-prices.
// Called by one lane.
template <class T>
__device__ __forceinline__ void store_price(double p, int o, double *__restrict__ out, int32_t *__restrict__ tally,
                                            int32_t *__restrict__ redo)
{
    if (sizeof(T) == sizeof(float) && p != p) redo[atomicAdd(&tally[T_NREDO], 1)] = o;
    else out[o] = p;
}

template <class T, int KMAX, bool NARROW, bool DYNAMIC>
__device__ __forceinline__ void price_tasks(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                                            const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                                            int queue, int first, int last, double *__restrict__ out,
                                            double *__restrict__ scratch, int32_t *__restrict__ redo)
{
    __shared__ T lds[NARROW ? 1 : WAVE * KMAX];
    __shared__ double etab[EDGE_TAB];
    const int lane = threadIdx.x;
    T *fscratch = reinterpret_cast<T *>(scratch + (size_t)blockIdx.x * (WAVE * KMAX));
    const int stride = uni((int)gridDim.x);
    int t = uni((int)blockIdx.x) + first;
    while (t < last) {
        const int o = uni(order[t]);
        const OptConst c = cst[o];
        const double *eo = edge + (size_t)o * EDGE_TAB;
        double p;
        if (NARROW) {
            const int K = (uni(c.w) + WAVE - 1) / WAVE;
            p = (K == 1) ? price_narrow<T, 1>(c, eo, lane, (lds_f64 *)etab)
                         : price_narrow<T, (KMAX >= 2 ? 2 : 1)>(c, eo, lane, (lds_f64 *)etab);
        } else {
            p = price_option<T, KMAX>(c, eo, lane, (typename Lds<T>::type *)lds, (lds_f64 *)etab, fscratch);
        }
        if (DYNAMIC) {
            int claim = 0;
            if (lane == 0) {
                store_price<T>(p, o, out, tally, redo);
                claim = atomicAdd(&tally[queue], 1);
            }
            t = uni(__shfl(claim, 0, WAVE)) + first + stride;
        } else {
            if (lane == 0) store_price<T>(p, o, out, tally, redo);
            t += stride;
        }
    }
}

template <class T>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2))) void k_price_wide(
    const OptConst *__restrict__ cst, const double *__restrict__ edge, const int32_t *__restrict__ order,
    int32_t *__restrict__ tally, double *__restrict__ out, double *__restrict__ scratch, int32_t *__restrict__ redo)
{
    clear_bucket_counts(tally);
    const int nlarge = uni(tally[T_NLARGE]);
    price_tasks<T, MAXK, false, true>(cst, edge, order, tally, T_QUEUE, 0, nlarge, out, scratch, redo);
    stamp_wave_end(tally);
}


// Row-class tasks [first, last) in groups of four consecutive tasks, one
// group per wave at a time (static stride); the last group may hold fewer
// options, and its idle rows repeat the group's first option without
// writing a price.
//   DYNAMIC: groups are claimed ROW_CHUNK at a time from tally[T_QUEUE_R]
//   (the fused kernel, whose wavefronts reach the row class at different
//   times); otherwise wave b takes groups b, b + grid, ...
constexpr int ROW_CHUNK = 4;
template <class T, bool DYNAMIC>
__device__ __forceinline__ void price_row_groups(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                                                 const int32_t *__restrict__ order, int first, int last,
                                                 double *__restrict__ out, int32_t *__restrict__ tally,
                                                 int32_t *__restrict__ redo)
{
    __shared__ double rowbuf[2 * ROWS * ROW_MAX_WIDTH];
    const int lane = threadIdx.x, row = lane / ROW, l = lane % ROW;
    lds_f64 *ybuf = (lds_f64 *)rowbuf + row * ROW_MAX_WIDTH;
    lds_f64 *bbuf = (lds_f64 *)rowbuf + (ROWS + row) * ROW_MAX_WIDTH;
    const int ngroups = (last - first + ROWS - 1) / ROWS;
    int g, gend;
    auto claim = [&]() {
        int c = 0;
        if (lane == 0) c = atomicAdd(&tally[T_QUEUE_R], 1);
        c = uni(__shfl(c, 0, WAVE));
        g = c * ROW_CHUNK;
        gend = min(g + ROW_CHUNK, ngroups);
    };
    if (DYNAMIC) claim();
    else g = uni((int)blockIdx.x);
    while (g < ngroups) {
        const int base = first + g * ROWS;
        const int mine = base + row < last ? order[base + row] : -1;
        const bool active = mine >= 0;
        const int o = active ? mine : order[base];  // a group's first slot is never idle
        const OptConst c = cst[o];
        RowOpt r;
        r.t = TreeShape{c.w, c.jmax, c.M, c.dr, c.dt, c.rho, c.rhop};
        r.n = c.n;
        r.e = c.e;
        r.type = c.type;
        r.X = c.X;
        r.Pe = c.Pe;
        r.Pn = c.Pn;
        r.edge = edge + (size_t)o * EDGE_TAB;
        const int e_row = active ? c.e : 0, n_row = active ? c.n : 0;
        int ev[8];
        int K = 1;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            ev[q] = __builtin_amdgcn_readlane(e_row, q * ROW);
            ev[ROWS + q] = __builtin_amdgcn_readlane(n_row, q * ROW);
            K = max(K, (__builtin_amdgcn_readlane(c.w, q * ROW) + ROW - 1) / ROW);
        }
        double p;
        switch (K) {
        case 1: p = price_rows<T, 1>(r, l, e_row, n_row, ev, ybuf, bbuf); break;
        case 2: p = price_rows<T, 2>(r, l, e_row, n_row, ev, ybuf, bbuf); break;
        case 3: p = price_rows<T, 3>(r, l, e_row, n_row, ev, ybuf, bbuf); break;
        default: p = price_rows<T, 4>(r, l, e_row, n_row, ev, ybuf, bbuf); break;
        }
        if (active && l == 0) store_price<T>(p, o, out, tally, redo);
        if (DYNAMIC) {
            if (++g >= gend) claim();
        } else {
            g += uni((int)gridDim.x);
        }
    }
}

// Flat-class groups, one per wave at a time: group G of the flat class is
// part j's group G - T_FLAT_GROUP[j] (parts hold one width each, widest
// first), 64 / w consecutive tasks from T_FLAT_START[j].  A part's padding
// slots come last, so a group's first task is always a real one; lanes of an
// idle tree price nothing.
//   DYNAMIC: groups are claimed ROW_CHUNK at a time from tally[T_QUEUE_F]
//   (the fused kernel); otherwise wave b takes groups b, b + grid, ...
template <class T, bool DYNAMIC>
__device__ __forceinline__ void price_flat_groups(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                                                  const int32_t *__restrict__ order, double *__restrict__ out,
                                                  int32_t *__restrict__ tally, int32_t *__restrict__ redo)
{
    __shared__ double flatbuf[2 * WAVE + 4 * flat_per_wave(3)];
    const int lane = threadIdx.x;
    lds_f64 *ybuf = (lds_f64 *)flatbuf, *bbuf = ybuf + WAVE, *sbuf = bbuf + WAVE;
    const int ngroups = uni(tally[T_NFLATG]);
    int G, gend;
    auto claim = [&]() {
        int c = 0;
        if (lane == 0) c = atomicAdd(&tally[T_QUEUE_F], 1);
        c = uni(__shfl(c, 0, WAVE));
        G = c * ROW_CHUNK;
        gend = min(G + ROW_CHUNK, ngroups);
    };
    if (DYNAMIC) claim();
    else G = uni((int)blockIdx.x);
    while (G < ngroups) {
        int j = 0;
        while (j + 1 < FLAT_PARTS && uni(tally[T_FLAT_GROUP + j + 1]) <= G) ++j;
        const int w = flat_width(FLAT_PARTS - 1 - j), g = flat_per_wave(w);
        const int start = uni(tally[T_FLAT_START + j]) + (G - uni(tally[T_FLAT_GROUP + j])) * g;
        const int q = lane / w, k = lane - q * w;
        const int mine = q < g ? order[start + q] : -1;
        const bool active = mine >= 0;
        const int o = active ? mine : order[start];
        const OptConst c = cst[o];
        RowOpt r;
        r.t = TreeShape{c.w, c.jmax, c.M, c.dr, c.dt, c.rho, c.rhop};
        r.n = c.n;
        r.e = c.e;
        r.type = c.type;
        r.X = c.X;
        r.Pe = c.Pe;
        r.Pn = c.Pn;
        r.edge = edge + (size_t)o * EDGE_TAB;
        const double p = price_flat<T>(r, k, active, g, w, ybuf, bbuf, sbuf);
        if (lane < g) {
            const int oq = order[start + lane];
            if (oq >= 0) store_price<T>(p, oq, out, tally, redo);
        }
        if (DYNAMIC) {
            if (++G >= gend) claim();
        } else {
            G += uni((int)gridDim.x);
        }
    }
}

// Four waves per SIMD (at most 128 VGPRs): on a 1M batch of mostly tall
// narrow trees the kernel runs 5% faster than at the three its 149 VGPRs
// would otherwise allow.
template <class T>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(4))) void k_price_narrow(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                                                       const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                                                       double *__restrict__ out, int32_t *__restrict__ redo)
{
    const int nlarge = uni(tally[T_NLARGE]), nmid = uni(tally[T_NMID]), nrow = uni(tally[T_NROW]);
    price_tasks<T, KSMALL, true, false>(cst, edge, order, tally, T_QUEUE_S, nlarge, nmid, out, nullptr, redo);
    price_row_groups<T, false>(cst, edge, order, nmid, nrow, out, tally, redo);
    price_flat_groups<T, false>(cst, edge, order, out, tally, redo);
    stamp_wave_end(tally);
}

// Every class in one persistent kernel: the wide trees from the LPT queue,
// then the mid class and the row groups from their own queues, so waves that
// run out of wide trees early price the narrow ones while the last wide
// trees finish.
template <class T>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2))) void k_price_fused(
    const OptConst *__restrict__ cst, const double *__restrict__ edge, const int32_t *__restrict__ order,
    int32_t *__restrict__ tally, double *__restrict__ out, double *__restrict__ scratch, int32_t *__restrict__ redo)
{
    clear_bucket_counts(tally);
    const int nlarge = uni(tally[T_NLARGE]), nmid = uni(tally[T_NMID]), nrow = uni(tally[T_NROW]);
    price_tasks<T, MAXK, false, true>(cst, edge, order, tally, T_QUEUE, 0, nlarge, out, scratch, redo);
    price_tasks<T, KSMALL, true, true>(cst, edge, order, tally, T_QUEUE_S, nlarge, nmid, out, nullptr, redo);
    price_row_groups<T, true>(cst, edge, order, nmid, nrow, out, tally, redo);
    price_flat_groups<T, true>(cst, edge, order, out, tally, redo);
    stamp_wave_end(tally);
}

} // namespace hwt
