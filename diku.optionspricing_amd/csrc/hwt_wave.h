// hwt_wave.h -- form 2 device code: one wavefront prices one option with the
// tree column resident in VGPRs.
//
// Layout.  A "window" of 64*K consecutive tree nodes starting at node `base`
// is blocked over the 64 lanes: lane l holds nodes base + l*K + s, s < K, in
// registers z[s].  One propagation step is the three-point stencil
//     z'[k] = cl[k] z[k-1] + cm[k] z[k] + cr[k] z[k+1]
// with per-node coefficients held in registers; the two values that cross a
// lane boundary move with DPP wave shifts (wave_shr:1 / wave_shl:1), so a step
// costs 3K FP64 FMA-class ops plus 4 DPP moves per lane and touches no memory.
//
// Phases.  The active part of the tree grows by one node per side per forward
// step until it reaches the edges, and the part of the backward pass that can
// still reach the exercise nodes shrinks by one node per side per step.  The
// window therefore comes in sizes K = 1, 2, ... (centred, half width 32K) and
// a final "full" window (base 0, K = ceil(w/64)); the wave moves between sizes
// through an LDS re-layout, so the work per step tracks the active width.
#pragma once

#include "hwt_math.h"

namespace hwt {

constexpr int WAVE = 64;
constexpr int MAXK = 21;                 // widths up to 64 * 21 = 1344 nodes
constexpr int WAVE_MAX_WIDTH = WAVE * MAXK;

// ------------------------------------------------------------------- DPP
__device__ __forceinline__ double dpp_from_lower_lane(double x)
{
    // lane l receives lane l-1; lane 0 receives 0 (bound_ctrl)
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double dpp_from_upper_lane(double x)
{
    // lane l receives lane l+1; lane 63 receives 0 (bound_ctrl)
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// ------------------------------------------------------------- stencil
// One step from src into dst (distinct register arrays).
template <int K>
__device__ __forceinline__ void stencil_step(const double (&src)[K], double (&dst)[K], const double (&cl)[K],
                                             const double (&cm)[K], const double (&cr)[K])
{
    const double left = dpp_from_lower_lane(src[K - 1]);
    const double right = dpp_from_upper_lane(src[0]);
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const double pv = (s > 0) ? src[s - 1] : left;
        const double nx = (s + 1 < K) ? src[s + 1] : right;
        double v = cm[s] * src[s];
        v = fma(cl[s], pv, v);
        dst[s] = fma(cr[s], nx, v);
    }
}

// nsteps steps in place; two steps per iteration ping-pong between z and a
// second register array so no value has to be copied.
template <int K>
__device__ __forceinline__ void stencil_steps(double (&z)[K], const double (&cl)[K], const double (&cm)[K],
                                              const double (&cr)[K], int nsteps)
{
    double t[K];
    for (int it = 0; it + 1 < nsteps; it += 2) {
        stencil_step<K>(z, t, cl, cm, cr);
        stencil_step<K>(t, z, cl, cm, cr);
    }
    if (nsteps & 1) {
        stencil_step<K>(z, t, cl, cm, cr);
#pragma unroll
        for (int s = 0; s < K; ++s) z[s] = t[s];
    }
}

typedef __attribute__((address_space(3))) double lds_f64;

// Scalar (wave-uniform) copies of values the caller guarantees are uniform.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v)
{
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                            __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
__device__ __forceinline__ TreeShape uni(const TreeShape &t)
{
    return TreeShape{uni(t.w), uni(t.jmax), uni(t.M), uni(t.dr), uni(t.dt), uni(t.rho), uni(t.rhop)};
}

// Edge coefficient table: for each pass (0 forward, 1 backward) the
// three-point coefficients of nodes 0, 1, 2, w-3, w-2, w-1, i.e. every node
// where the edge-corrected operator differs from the closed forms.
constexpr int EDGE_TAB = 2 * 6 * 3;

__device__ __forceinline__ int edge_slot(int k, int w) { return k < 3 ? k : 3 + (k - (w - 3)); }

// Fills the table; lanes 0..35 each compute one entry.
__device__ __forceinline__ void fill_edge_table(const TreeShape &t, int lane, lds_f64 *tab)
{
    if (lane < EDGE_TAB) {
        const int pass = lane / 18, slot = (lane % 18) / 3, which = lane % 3;
        const int k = slot < 3 ? slot : t.w - 3 + (slot - 3);
        double v;
        if (pass == 0) v = op_G(t, k, k - 1 + which);             // forward: G[k][k-1..k+1]
        else v = op_G(t, k - 1 + which, k);                        // backward: G^T[k][k-1..k+1]
        tab[lane] = v;
    }
    __syncthreads();
}

// One coefficient of the three-point stencil (which: 0 left, 1 mid, 2 right)
// for every node of the window [base, base + 64K), written to buf[0, 64K) in
// node order.  Lanes take nodes base+lane, base+lane+64, ...; the node
// discounts E_k = e^{-(k-jmax) dr dt} advance by e^{-64 dr dt} per node.
// Interior nodes use the closed forms; nodes next to the edges read the
// precomputed table; padding beyond the top edge is zero.
template <bool FWD>
__device__ __forceinline__ void stage_coef(const TreeShape &t, int base, int K, int lane, int which, lds_f64 *buf,
                                           const lds_f64 *etab)
{
    const int d = FWD ? which - 1 : 0;  // forward coefficients use the source node's discount
    const double q64 = exp(-64.0 * t.dr * t.dt);
    double E = node_disc(base + lane + d, t.jmax, t.dr, t.dt);
    const int nwin = WAVE * K;
#pragma unroll 1
    for (int m = lane; m < nwin; m += WAVE) {
        const int k = base + m;
        const int j = k - t.jmax;
        double v;
        if (k >= t.w) {
            v = 0.0;
        } else if (k < 3 || k > t.w - 4) {
            v = etab[(FWD ? 0 : 18) + 3 * edge_slot(k, t.w) + which];
        } else if (FWD) {
            v = (which == 0 ? pu_std(j - 1, t.M) : which == 1 ? pm_std(j, t.M) : pd_std(j + 1, t.M)) * E;
        } else {
            v = (which == 0 ? pd_std(j, t.M) : which == 1 ? pm_std(j, t.M) : pu_std(j, t.M)) * E;
        }
        buf[m] = v;
        E *= q64;
    }
    __syncthreads();
}

template <int K>
__device__ __forceinline__ void fetch_coef(int lane, const lds_f64 *buf, double (&c)[K])
{
#pragma unroll
    for (int s = 0; s < K; ++s) c[s] = buf[lane * K + s];
    __syncthreads();
}

template <int K, bool FWD>
__device__ __forceinline__ void window_coefs(const TreeShape &t, int base, int lane, lds_f64 *buf, const lds_f64 *etab,
                                             double (&cl)[K], double (&cm)[K], double (&cr)[K])
{
    stage_coef<FWD>(t, base, K, lane, 0, buf, etab);
    fetch_coef<K>(lane, buf, cl);
    stage_coef<FWD>(t, base, K, lane, 1, buf, etab);
    fetch_coef<K>(lane, buf, cm);
    stage_coef<FWD>(t, base, K, lane, 2, buf, etab);
    fetch_coef<K>(lane, buf, cr);
}

// Initial values: forward delta at the root node; backward T^{-T} * 100.
template <int K, bool FWD>
__device__ __forceinline__ void init_window(double (&z)[K], const TreeShape &t, int base, int lane)
{
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = base + lane * K + s;
        double v;
        if (FWD) {
            v = (k == t.jmax) ? 1.0 : 0.0;
        } else {
            v = (k < t.w) ? 100.0 : 0.0;
            if (k == 0) v -= t.rho * 100.0;
            if (k == t.w - 1) v -= t.rhop * 100.0;
        }
        z[s] = v;
    }
}

// One phase: bring the window in (initial values, or the previous window's
// nodes [in_lo, in_hi) from LDS, zeros elsewhere), run nsteps stencil steps,
// and leave the window's tree nodes in LDS for the next phase.
template <int K, bool FWD>
__device__ __noinline__ void phase_fn(TreeShape tv, int basev, int nstepsv, int in_lov, int in_hiv, int lane,
                                      lds_f64 *lds, const lds_f64 *etab)
{
    const TreeShape t = uni(tv);
    const int base = uni(basev), nsteps = uni(nstepsv), in_lo = uni(in_lov), in_hi = uni(in_hiv);
    double z[K];
    if (in_lo >= in_hi) {
        init_window<K, FWD>(z, t, base, lane);
    } else {
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int k = base + lane * K + s;
            z[s] = (k >= in_lo && k < in_hi) ? lds[k] : 0.0;
        }
    }
    __syncthreads();
    if (nsteps > 0) {
        double cl[K], cm[K], cr[K];
        window_coefs<K, FWD>(t, base, lane, lds, etab, cl, cm, cr);
        stencil_steps<K>(z, cl, cm, cr, nsteps);
    }
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = base + lane * K + s;
        if (k < t.w) lds[k] = z[s];
    }
    __syncthreads();
}

#define HWT_KCASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) \
    X(14) X(15) X(16) X(17) X(18) X(19) X(20) X(21)

template <bool FWD>
__device__ __forceinline__ void dispatch_phase(int K, const TreeShape &t, int base, int nsteps, int in_lo, int in_hi,
                                               int lane, lds_f64 *lds, const lds_f64 *etab)
{
    switch (K) {
#define HWT_C(KK) case KK: phase_fn<KK, FWD>(t, base, nsteps, in_lo, in_hi, lane, lds, etab); break;
        HWT_KCASES(HWT_C)
#undef HWT_C
    default: break;
    }
}

struct Window {
    int K, base;
};

// Smallest centred window of half width 32K that can hold half-width D, or
// the full window when none fits inside the tree.
__device__ __forceinline__ Window window_for(int D, const TreeShape &t)
{
    const int Kg = D / 32 + 1;
    if (32 * Kg <= t.jmax) return Window{Kg, t.jmax - 32 * Kg};
    return Window{(t.w + WAVE - 1) / WAVE, 0};
}

// Forward pass: state 0 (delta at the root) -> state e.  The state is left in
// lds[lo, hi); tree nodes outside that range are zero.
__device__ __forceinline__ void forward_pass(const TreeShape &t, int e, int lane, lds_f64 *lds, const lds_f64 *etab,
                                             int &lo, int &hi)
{
    Window win = window_for(0, t);
    int in_lo = 0, in_hi = 0;  // empty: initial values
    int i = 0;
    for (;;) {
        const bool full = 32 * win.K > t.jmax;  // growth windows satisfy 32K <= jmax
        const int iend = full ? e : min(e, 32 * win.K - 1);
        dispatch_phase<true>(win.K, t, win.base, iend - i, in_lo, in_hi, lane, lds, etab);
        i = iend;
        in_lo = win.base;
        in_hi = min(win.base + WAVE * win.K, t.w);
        if (i >= e) break;
        win = window_for(i + 1, t);
    }
    lo = in_lo;
    hi = in_hi;
}

// Backward pass: state n (bond face value 100) -> state e, keeping only the
// nodes that can still reach the exercise nodes |j| <= min(e, jmax).
__device__ __forceinline__ void backward_pass(const TreeShape &t, int n, int e, int lane, lds_f64 *lds,
                                              const lds_f64 *etab, int &lo, int &hi)
{
    const int ep = min(e, t.jmax);
    const int Kgmax = t.jmax / 32;
    int s = n;
    Window win = window_for(min(t.jmax, ep + (s - e)), t);
    int in_lo = 0, in_hi = 0;
    while (s > e) {
        const bool full = 32 * win.K > t.jmax;
        const int klo = full ? Kgmax : win.K - 1;
        const int send = max(e, e + 32 * klo - 1 - ep);
        dispatch_phase<false>(win.K, t, win.base, s - send, in_lo, in_hi, lane, lds, etab);
        s = send;
        in_lo = win.base;
        in_hi = min(win.base + WAVE * win.K, t.w);
        if (s > e) win = window_for(min(t.jmax, ep + (s - e)), t);
    }
    lo = in_lo;
    hi = in_hi;
}

} // namespace hwt
