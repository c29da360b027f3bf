// hwt_lane.h -- form 1 device code: one lane prices one option.
//
// Each lane owns its option's two tree columns (forward state y, backward
// state b) in global memory, interleaved across the 64 lanes of its wave:
// node k of column c of lane l lives at  base_wave + (c*W + k)*64 + l, so the
// nodes a wave touches in one sweep position form one contiguous 512-byte row
// and every load and store is coalesced.  A step sweeps the active region with
// a three-value sliding window (previous, current, next) in registers and
// recomputes the stencil coefficients on the fly (closed forms, node discounts
// advanced by one multiplication per node, edge nodes from the option's edge
// table).  The same duality as the wave form gives the price from the two
// columns at the exercise step.
#pragma once

#include "hwt_wave.h"

namespace hwt {

constexpr int LANE_MAX_WIDTH = WAVE_MAX_WIDTH;

struct LaneColumn {
    double *p;  // node 0 of this lane's column
    __device__ double &operator[](int k) const { return p[(size_t)k * WAVE]; }
};

// Sweep one step over [lo, hi] (new region), reading the old state (nonzero
// only on [olo, ohi]) in place.  FWD: forward operator G, else G^T.
template <bool FWD>
__device__ __forceinline__ void lane_sweep(const TreeShape &t, const double *etab, LaneColumn col, int lo, int hi,
                                           int olo, int ohi)
{
    const double q = exp(-t.dr * t.dt);
    double Em = node_disc(lo - 1, t.jmax, t.dr, t.dt);  // E_{k-1}
    double E0 = Em * q, Ep = E0 * q;
    // old value at lo-1: inside the old region when the backward light cone
    // narrows, zero when it lies outside the old region or the tree
    double prev = (lo - 1 >= olo && lo - 1 <= ohi && lo >= 1) ? col[lo - 1] : 0.0;
    double cur = (lo >= olo && lo <= ohi) ? col[lo] : 0.0;
    for (int k = lo; k <= hi; ++k) {
        const double nxt = (k + 1 >= olo && k + 1 <= ohi) ? col[k + 1] : 0.0;
        const int j = k - t.jmax;
        double cl, cm, cr;
        if (k < 3 || k > t.w - 4) {
            const double *e = etab + (FWD ? 0 : 18) + 3 * edge_slot_of(k, t.w);
            cl = e[0];
            cm = e[1];
            cr = e[2];
        } else if (FWD) {
            cl = pu_std(j - 1, t.M) * Em;
            cm = pm_std(j, t.M) * E0;
            cr = pd_std(j + 1, t.M) * Ep;
        } else {
            cl = pd_std(j, t.M) * E0;
            cm = pm_std(j, t.M) * E0;
            cr = pu_std(j, t.M) * E0;
        }
        col[k] = fma(cr, nxt, fma(cl, prev, cm * cur));
        prev = cur;
        cur = nxt;
        Em = E0;
        E0 = Ep;
        Ep *= q;
    }
}

// Prices one option; columns at Y (forward) and B (backward).
__device__ __forceinline__ double lane_price(const OptConst &c, const double *etab, LaneColumn Y, LaneColumn B)
{
    const TreeShape t{c.w, c.jmax, c.M, c.dr, c.dt, c.rho, c.rhop};
    const int J = c.jmax, n = c.n, e = c.e;
    // forward: delta at the root, e steps
    Y[J] = 1.0;
    for (int i = 0; i < e; ++i) {
        const int ho = min(i, J), hn = min(i + 1, J);
        lane_sweep<true>(t, etab, Y, J - hn, J + hn, J - ho, J + ho);
    }
    // backward: face value at step n over the nodes that reach the exercise
    // region, then n - e steps, each one node narrower per side
    const int ep = min(e, J);
    auto need = [&](int s) { return min(J, ep + (s - e)); };
    {
        const int h = need(n);
        for (int k = J - h; k <= J + h; ++k) {
            double v = 100.0;
            if (k == 0) v -= t.rho * 100.0;
            if (k == t.w - 1) v -= t.rhop * 100.0;
            B[k] = v;
        }
    }
    for (int s = n; s > e; --s) {
        const int ho = need(s), hn = need(s - 1);
        lane_sweep<false>(t, etab, B, J - hn, J + hn, J - ho, J + ho);
    }
    // combine at the exercise step
    const int lo = J - ep, hi = J + ep;
    auto yk = [&](int k) {
        double v = Y[k];
        if (k == 1) v -= t.rho * (0 >= lo ? Y[0] : 0.0);
        if (k == t.w - 2) v -= t.rhop * (t.w - 1 <= hi ? Y[t.w - 1] : 0.0);
        return v;
    };
    auto bk = [&](int k) {
        double v = B[k];
        if (k == 0) v += t.rho * B[1];
        if (k == t.w - 1) v += t.rhop * B[t.w - 2];
        return v;
    };
    double s1 = 0.0, s2 = 0.0;
    for (int k = lo; k <= hi; ++k) {
        const double yv = yk(k);
        s1 += yv;
        s2 += yv * bk(k);
    }
    const double scale = 100.0 * s1 * c.Pn / (s2 * c.Pe);
    double s3 = 0.0;
    for (int k = lo; k <= hi; ++k) {
        const double Bv = scale * bk(k);
        double v = (c.type == 1) ? c.X - Bv : Bv - c.X;  // PUT : CALL
        v = (v < 0.0) ? 0.0 : v;
        if (v == INFINITY) v = DBL_MAX;
        s3 += yk(k) * v;
    }
    return c.Pe / s1 * s3;
}

} // namespace hwt
