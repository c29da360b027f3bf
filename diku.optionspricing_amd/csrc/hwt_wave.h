// hwt_wave.h -- form 2 device code: one wavefront prices one option with the
// tree column resident in VGPRs.
//
// Layout.  A "window" of 64*K consecutive tree nodes starting at node `base`
// is blocked over the 64 lanes: lane l holds nodes base + l*K + s, s < K, in
// registers z[s].  One propagation step is the three-point stencil
//     z'[k] = cl[k] z[k-1] + cm[k] z[k] + cr[k] z[k+1]
// with per-node coefficients held in registers; the two values that cross a
// lane boundary move with DPP wave shifts (wave_shr:1 / wave_shl:1), so a step
// costs 3K FP64 ops (mul, fma, fma) plus 4 DPP moves per lane and touches no
// memory.
//
// Phases.  The active part of the tree grows by one node per side per forward
// step until it reaches the edges, and the part of the backward pass that can
// still reach the exercise nodes shrinks by one node per side per step.  The
// window therefore comes in sizes K = 1, 2, ... (centred, half width 32K) and
// a final "full" window (base 0, K = ceil(w/64)); the wave moves between sizes
// through an LDS re-layout, so the work per step tracks the active width.
#pragma once

#include "hwt_math.h"

#include <cfloat>

namespace hwt {

constexpr int WAVE = 64;
constexpr int MAXK = 21;                 // widths up to 64 * 21 = 1344 nodes
constexpr int WAVE_MAX_WIDTH = WAVE * MAXK;

typedef __attribute__((address_space(3))) double lds_f64;

// ------------------------------------------------------------------- DPP
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// lane l receives lane l-1; lane 0 receives 0
__device__ __forceinline__ double dpp_from_lower_lane(double x) { return dpp_mov<0x138>(x); }
// lane l receives lane l+1; lane 63 receives 0
__device__ __forceinline__ double dpp_from_upper_lane(double x) { return dpp_mov<0x130>(x); }

__device__ __forceinline__ double readlane_f64(double x, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                            __builtin_amdgcn_readlane(__double2loint(x), lane));
}

// Sum over the 64 lanes, the same in every lane: DPP butterflies inside each
// 16-lane row (quad_perm [1,0,3,2], quad_perm [2,3,0,1], half-row mirror, row
// mirror), then the four row sums are read out and added in a fixed order.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Sum over each 16-lane row, the same in every lane of the row: the first
// four butterflies of wave_sum.
__device__ __forceinline__ double row_sum(double v)
{
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    return v;
}

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

// ------------------------------------------------------------ edge table
// For each pass (0 forward, 1 backward) the three-point coefficients of nodes
// 0, 1, 2, w-3, w-2, w-1 -- every node where the edge-corrected operator
// differs from the closed forms -- 36 entries.  Computed once per option by the
// setup kernel and handed to the pricing wave through LDS.
constexpr int EDGE_TAB = 2 * 6 * 3;

HWT_HD int edge_node(int slot, int w) { return slot < 3 ? slot : w - 6 + slot; }
// Table slot of edge node k (k < 3 or k > w-4).
HWT_HD int edge_slot_of(int k, int w) { return k < 3 ? k : 3 + (k - (w - 3)); }

// Node discounts of nodes 0..3 and w-4..w-1 (all the edge coefficients need).
struct EdgeDisc {
    double E[8];
    int w;
    HWT_HD double operator()(int k) const { return k < 4 ? E[k] : E[4 + (k - (w - 4))]; }
};

// General form: every entry from the definition G = T F T^{-1}.
HWT_HD void edge_table_general(const TreeShape &t, double *tab)
{
    EdgeDisc d;
    d.w = t.w;
    for (int i = 0; i < 8; ++i) {
        const int k = i < 4 ? i : t.w - 8 + i;
        d.E[i] = (k >= 0 && k < t.w) ? node_disc(k, t.jmax, t.dr, t.dt) : 0.0;
    }
    for (int slot = 0; slot < 6; ++slot) {
        const int k = edge_node(slot, t.w);
        for (int which = 0; which < 3; ++which) {
            tab[3 * slot + which] = op_G(t, d, k, k - 1 + which);       // forward: G[k][k-1..k+1]
            tab[18 + 3 * slot + which] = op_G(t, d, k - 1 + which, k);  // backward: G^T[k][k-1..k+1]
        }
    }
}

// Node discounts of the four bottom and four top nodes (E0..E3 of nodes 0..3,
// Ea..Ed of nodes w-1..w-4) from four exponentials instead of eight: with
// h = dr dt and X = jmax h, E0 = e^X and E_{k+1} = E_k e^{-h} going up,
// Ea = e^{-X} and E_{k-1} = E_k e^{h} going down (the pricers advance their
// interior discounts the same way, one multiplication per node).
struct EdgeDiscounts {
    double E0, E1, E2, E3, Ea, Eb, Ec, Ed;
};

HWT_HD EdgeDiscounts edge_discounts(const TreeShape &t)
{
    const double h = t.dr * t.dt, X = t.jmax * h;
    const double q = exp(-h), qi = exp(h);
    EdgeDiscounts e;
    e.E0 = exp(X);
    e.E1 = e.E0 * q;
    e.E2 = e.E1 * q;
    e.E3 = e.E2 * q;
    e.Ea = exp(-X);
    e.Eb = e.Ea * qi;
    e.Ec = e.Eb * qi;
    e.Ed = e.Ec * qi;
    return e;
}

// Edge-elimination factors from those discounts (as edge_factors, trees of
// at least 5 nodes): rho = F[2][0] / F[2][1], rhop = F[w-3][w-1] / F[w-3][w-2].
HWT_HD void edge_factors(TreeShape &t, const EdgeDiscounts &e)
{
    const int J = t.jmax;
    t.rho = (pu_bot(-J, t.M) * e.E0) / (pu_std(1 - J, t.M) * e.E1);
    t.rhop = (pd_top(J, t.M) * e.Ea) / (pd_std(J - 1, t.M) * e.Eb);
}

// Closed form for trees at least 7 nodes wide, where the two edge
// corrections touch disjoint rows and columns.  With F[r][s] = p(s->r) E_s,
// bottom rho = F20/F21, top rhop = F[c][a]/F[c][b] (a = w-1, b = w-2, c = w-3):
//   G00 = F00 - rho F01,  G10 = F10 - rho F11 + rho G00,  G11 = F11 + rho F01,
//   G[a][a] = Faa - rhop Fab,  G[b][a] = Fba - rhop Fbb + rhop G[a][a],
//   G[b][b] = Fbb + rhop Fab,  every other entry equals F.
HWT_HD void edge_table_closed(const TreeShape &t, const EdgeDiscounts &e, double *tab)
{
    const double M = t.M, rho = t.rho, rhop = t.rhop;
    const int J = t.jmax, b = t.w - 2, c = t.w - 3, d = t.w - 4;
    const double E0 = e.E0, E1 = e.E1, E2 = e.E2, E3 = e.E3;
    const double Ea = e.Ea, Eb = e.Eb, Ec = e.Ec, Ed = e.Ed;
    // bottom: node 0 branches (j=-J) to 0,1,2; nodes 1,2,3 are standard
    const double F00 = pd_bot(-J, M) * E0, F10 = pm_bot(-J, M) * E0;
    const double F01 = pd_std(1 - J, M) * E1, F11 = pm_std(1 - J, M) * E1, F21 = pu_std(1 - J, M) * E1;
    const double F12 = pd_std(2 - J, M) * E2, F22 = pm_std(2 - J, M) * E2, F32 = pu_std(2 - J, M) * E2;
    const double F23 = pd_std(3 - J, M) * E3;
    const double G00 = F00 - rho * F01;
    const double G10 = F10 - rho * F11 + rho * G00;
    const double G11 = F11 + rho * F01;
    // top: node a branches (j=J) to a,b,c; nodes b,c,d are standard
    const double Faa = pu_top(J, M) * Ea, Fba = pm_top(J, M) * Ea;
    const double Fab = pu_std(b - J, M) * Eb, Fbb = pm_std(b - J, M) * Eb, Fcb = pd_std(b - J, M) * Eb;
    const double Fbc = pu_std(c - J, M) * Ec, Fcc = pm_std(c - J, M) * Ec, Fdc = pd_std(c - J, M) * Ec;
    const double Fcd = pu_std(d - J, M) * Ed;
    const double Gaa = Faa - rhop * Fab;
    const double Gba = Fba - rhop * Fbb + rhop * Gaa;
    const double Gbb = Fbb + rhop * Fab;
    const double fwd[18] = {0.0, G00, F01,  G10, G11, F12,  F21, F22, F23,
                            Fcd, Fcc, Fcb,  Fbc, Gbb, Gba,  Fab, Gaa, 0.0};
    const double bwd[18] = {0.0, G00, G10,  F01, G11, F21,  F12, F22, F32,
                            Fdc, Fcc, Fbc,  Fcb, Gbb, Fab,  Gba, Gaa, 0.0};
    for (int i = 0; i < 18; ++i) {
        tab[i] = fwd[i];
        tab[18 + i] = bwd[i];
    }
}

HWT_HD void edge_table(const TreeShape &t, double *tab)
{
    if (t.w < 7) edge_table_general(t, tab);
    else edge_table_closed(t, edge_discounts(t), tab);
}

__device__ __forceinline__ void load_edge_table(const double *src, int lane, lds_f64 *tab)
{
    if (lane < EDGE_TAB) tab[lane] = src[lane];
    __syncthreads();
}

// ------------------------------------------------------------- stencil
// Propagation type T: double (FP64 mode) or float (mixed mode, where the two
// tree passes run in FP32 and every discount factor and sum stays FP64).
template <class T> struct Lds;
template <> struct Lds<double> { typedef __attribute__((address_space(3))) double type; };
template <> struct Lds<float> { typedef __attribute__((address_space(3))) float type; };

__device__ __forceinline__ float dpp_from_lower_lane(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_from_upper_lane(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, true));
}

// Mixed mode's range guard.  The undiscounted backward (bond) state of a long,
// weakly mean-reverting bond grows without bound at the low-rate edge of the
// tree: 1e40 to 1e80 in FP64 for 100-year bonds at a ~ 0.01, past FP32's
// 3.4e38.  An overflowed node turns its light cone into inf/NaN, so the FP32
// pricers report an option as NaN when a duality sum is not finite or a bond
// value at an exercise node exceeds this bound, and the option is re-priced by
// the FP64 kernel.  Below the bound a forward value lost to FP32 underflow
// (< 1.2e-38) changes S2 by less than 1e-9 per node.
constexpr double F32_SAFE_STATE = 0x1p96;

// Largest finite value of the propagation type: a payoff that overflows is
// clamped to it, as the reference does in either precision
// (common/Domain.hpp:163-168).
template <class T> struct RealMax;
template <> struct RealMax<double> { static constexpr double value = DBL_MAX; };
template <> struct RealMax<float> { static constexpr double value = FLT_MAX; };

// One step from src into dst (distinct register arrays).
template <class T, int K>
__device__ __forceinline__ void stencil_step(const T (&src)[K], T (&dst)[K], const T (&cl)[K], const T (&cm)[K],
                                             const T (&cr)[K])
{
    const T left = dpp_from_lower_lane(src[K - 1]);
    const T right = dpp_from_upper_lane(src[0]);
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const T pv = (s > 0) ? src[s - 1] : left;
        const T nx = (s + 1 < K) ? src[s + 1] : right;
        T v = cm[s] * src[s];
        v = fma(cl[s], pv, v);
        v = fma(cr[s], nx, v);
        dst[s] = v;
    }
}

// nsteps steps in place; two steps per iteration ping-pong between z and a
// second register array so no value has to be copied.
template <class T, int K>
__device__ __forceinline__ void stencil_steps(T (&z)[K], const T (&cl)[K], const T (&cm)[K], const T (&cr)[K],
                                              int nsteps)
{
    T t[K];
    for (int it = 0; it + 1 < nsteps; it += 2) {
        stencil_step<T, K>(z, t, cl, cm, cr);
        stencil_step<T, K>(t, z, cl, cm, cr);
    }
    if (nsteps & 1) {
        stencil_step<T, K>(z, t, cl, cm, cr);
#pragma unroll
        for (int s = 0; s < K; ++s) z[s] = t[s];
    }
}

// Coefficients of this lane's K nodes for the forward operator G (FWD) or the
// backward operator G^T, computed in FP64 and stored as T.  Interior nodes use
// the closed forms with the node discounts E_k = e^{-(k-jmax) dr dt} advanced
// by one multiplication per node; padding beyond the top edge is zero; the (at
// most six) edge nodes inside the window are then overwritten from the table,
// with a wave-uniform switch on their slot so every register index stays static.
template <class T, int K, bool FWD>
__device__ __forceinline__ void window_coefs(const TreeShape &t, int base, int lane, const lds_f64 *etab,
                                             T (&cl)[K], T (&cm)[K], T (&cr)[K])
{
    const int k0 = base + lane * K;
    const double q = exp(-t.dr * t.dt);
    if (FWD) {
        // outgoing weights pu*E (to k+1), pm*E (to k), pd*E (to k-1) of the
        // source nodes k0-1 .. k0+K, rolled into the targets' three slots
        double E = node_disc(k0 - 1, t.jmax, t.dr, t.dt);
#pragma unroll
        for (int m = 0; m < K + 2; ++m) {
            const int k = k0 - 1 + m;
            const int j = k - t.jmax;
            const double Ek = (k >= 0 && k < t.w) ? E : 0.0;
            if (m < K) cl[m < K ? m : 0] = (T)(pu_std(j, t.M) * Ek);
            if (m >= 1 && m <= K) cm[(m >= 1 && m <= K) ? m - 1 : 0] = (T)(pm_std(j, t.M) * Ek);
            if (m >= 2) cr[m >= 2 ? m - 2 : 0] = (T)(pd_std(j, t.M) * Ek);
            E *= q;
        }
#pragma unroll
        for (int s = 0; s < K; ++s)
            if (k0 + s >= t.w) cl[s] = cm[s] = cr[s] = (T)0;
    } else {
        double E = node_disc(k0, t.jmax, t.dr, t.dt);
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int j = k0 + s - t.jmax;
            const double Ek = (k0 + s < t.w) ? E : 0.0;
            cl[s] = (T)(pd_std(j, t.M) * Ek);
            cm[s] = (T)(pm_std(j, t.M) * Ek);
            cr[s] = (T)(pu_std(j, t.M) * Ek);
            E *= q;
        }
    }
    const lds_f64 *tab = etab + (FWD ? 0 : 18);
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int k = edge_node(e, t.w);
        const int off = k - base;
        if (k >= 0 && k < t.w && off >= 0 && off < WAVE * K) {
            const int L = off / K, S = off - L * K;  // wave-uniform
            const T a = (T)tab[3 * e], b = (T)tab[3 * e + 1], c = (T)tab[3 * e + 2];
            if (lane == L) {
                switch (S) {
#define HWT_SLOT(SS)                                       \
    case SS:                                               \
        if (SS < K) {                                      \
            cl[SS < K ? SS : 0] = a;                       \
            cm[SS < K ? SS : 0] = b;                       \
            cr[SS < K ? SS : 0] = c;                       \
        }                                                  \
        break;
                    HWT_SLOT(0) HWT_SLOT(1) HWT_SLOT(2) HWT_SLOT(3) HWT_SLOT(4) HWT_SLOT(5) HWT_SLOT(6)
                    HWT_SLOT(7) HWT_SLOT(8) HWT_SLOT(9) HWT_SLOT(10) HWT_SLOT(11) HWT_SLOT(12) HWT_SLOT(13)
                    HWT_SLOT(14) HWT_SLOT(15) HWT_SLOT(16) HWT_SLOT(17) HWT_SLOT(18) HWT_SLOT(19) HWT_SLOT(20)
#undef HWT_SLOT
                default: break;
                }
            }
        }
    }
}

// Initial values: forward delta at the root node; backward T^{-T} * 100.
template <class T, int K, bool FWD>
__device__ __forceinline__ void init_window(T (&z)[K], const TreeShape &t, int base, int lane)
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
        z[s] = (T)v;
    }
}

// ------------------------------------------------- unit-diagonal stencil
// The diagonal similarity u = D^{-1} z, D = diag(d), with d = 1 on each lane's
// first node and d_k = G[k][k-1] d_{k-1} along the lane, makes every in-lane
// sub-diagonal entry 1, so a node update costs two FMAs instead of three ops:
//     u'_k = u_{k-1} + cm_k u_k + cr'_k u_{k+1}           (s >= 1)
//     u'_k = cl'_k u_{k-1} + cm_k u_k + cr'_k u_{k+1}     (s = 0)
// with cr'_k = cr_k d_{k+1} / d_k and cl'_k = cl_k d_{k-1} (d_{k-1} is the
// lower lane's last scale).  Padding nodes (k >= w) take d_k = d_{k-1}: their
// values become copies of real values, which real nodes read only through
// zero coefficients.  The setup kernel checks per option that every scale
// stays in range (unit_form_ok); the plain stencil prices the rest.
//
// The coefficients are formed one node at a time and transformed on the
// way (the closed forms, advanced incrementally; in windows that reach an
// edge, edge nodes are read from the LDS table at a per-lane index and
// padding nodes get zero), so no plain coefficient array is ever live beside the
// transformed ones: the set-up needs no more registers than the steps, and
// the form fits every window size at two waves per SIMD.  Undoing D after the
// steps needs the scales again: each lane parks them in its own node slots of
// the LDS state buffer, which the steps leave untouched.  In mixed mode the
// coefficients and scales are formed in FP64 and rounded to FP32, and the
// state and the steps are FP32.

// Transformed coefficients of this lane's nodes, and z -> D^{-1} z: the chain
// d_s = c_s d_{s-1} (c_s = cl_s on tree nodes, 1 on padding) is built going
// up with cr'_{s-1} = cr_{s-1} c_s, then walked down with 1/d_{s-1} = c_s / d_s.
template <class T, int K, bool FWD>
__device__ __forceinline__ void unit_in(const TreeShape &t, int base, int lane, const lds_f64 *etab, T (&z)[K],
                                        T &cl0, T (&um)[K], T (&ur)[K], typename Lds<T>::type *lds)
{
    const lds_f64 *tab = etab + (FWD ? 0 : 18);
    const int k0 = base + lane * K;
    // a window strictly inside the tree (every growth and shrink window) has
    // no edge or padding node: its coefficients are the closed forms alone
    const bool inner = base >= 4 && base + WAVE * K <= t.w - 4;
    const double q = exp(-t.dr * t.dt);
    auto inside = [&](int k) { return k >= 0 && k < t.w; };
    double E = node_disc(k0 - 1, t.jmax, t.dr, t.dt);  // discount of node k0-1+m at step m
    double Eprev = inside(k0 - 1) ? E : 0.0;
    E *= q;
    double Ek = inside(k0) ? E : 0.0;
    // the standard probabilities from a_k = (k - jmax) M, advanced by one
    // addition per node: pu, pm, pd = 1/6 + (a^2 + a)/2, 2/3 - a^2,
    // 1/6 + (a^2 - a)/2 (forward: the outgoing weights of the sources k-1, k,
    // k+1 of target k, rolled along)
    double a = (double)(k0 - 1 - t.jmax) * t.M;
    double a2 = a * a, pu_prev = 0.0, pm_cur = 0.0;
    if (FWD) {
        pu_prev = fma(0.5, a2 + a, 1.0 / 6.0) * Eprev;
        a += t.M;
        a2 = a * a;
        pm_cur = (2.0 / 3.0 - a2) * Ek;
    }
    double c[K];  // the chain factors, kept for the walk down
    double cl, cm, cr, cr_prev = 0.0, d = 1.0, cl0d = 0.0;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = k0 + s;
        E *= q;
        const double Enext = inside(k + 1) ? E : 0.0;
        if (FWD) {
            const double pu_cur = fma(0.5, a2 + a, 1.0 / 6.0) * Ek;
            a += t.M;
            a2 = a * a;
            cl = pu_prev;
            cm = pm_cur;
            cr = fma(0.5, a2 - a, 1.0 / 6.0) * Enext;
            pu_prev = pu_cur;
            pm_cur = (2.0 / 3.0 - a2) * Enext;
        } else {
            a += t.M;
            a2 = a * a;
            cl = fma(0.5, a2 - a, 1.0 / 6.0) * Ek;
            cm = (2.0 / 3.0 - a2) * Ek;
            cr = fma(0.5, a2 + a, 1.0 / 6.0) * Ek;
        }
        if (!inner) {  // wave-uniform: only windows that reach an edge
            if (k >= t.w) cl = cm = cr = 0.0;
            if (k < 3 || (k >= t.w - 3 && k < t.w)) {
                const int e = 3 * edge_slot_of(k, t.w);
                cl = tab[e];
                cm = tab[e + 1];
                cr = tab[e + 2];
            }
        }
        um[s] = (T)cm;
        if (s == 0) {
            cl0d = cl;
        } else {
            c[s] = (k < t.w) ? cl : 1.0;
            d *= c[s];
            ur[s - 1] = (T)(cr_prev * c[s]);
            lds[k] = (T)d;
        }
        cr_prev = cr;
        Eprev = Ek;
        Ek = Enext;
    }
    double inv = 1.0 / d;
    ur[K - 1] = (T)(cr_prev * inv);
    cl0 = (T)(cl0d * dpp_from_lower_lane(d));
#pragma unroll
    for (int s = K - 1; s > 0; --s) {
        z[s] = (T)((double)z[s] * inv);
        inv *= c[s];
    }
}

// z -> D z with the scales unit_in parked in LDS.
template <class T, int K>
__device__ __forceinline__ void unit_out(int base, int lane, T (&z)[K], const typename Lds<T>::type *lds)
{
    const int k0 = base + lane * K;
#pragma unroll
    for (int s = 1; s < K; ++s) z[s] *= lds[k0 + s];
}

// Largest window the unit form is used for: every window (21 nodes per
// lane), in both precisions.  In FP32, D^{-1} (scales down to 1e-30, checked
// per option by the setup kernel) can lift the undiscounted backward state of
// a long, weakly mean-reverting bond past FP32's range; such an option's sums
// or exercise-node values leave the range guard's bound, and it is re-priced
// by the FP64 kernel like any other out-of-range option.
template <class T> struct UnitMaxK;
template <> struct UnitMaxK<double> { static constexpr int value = MAXK; };
template <> struct UnitMaxK<float> { static constexpr int value = MAXK; };

// a*b + c into a fresh register.  Written as the three-address v_fma_f64 /
// v_fma_f32: left to itself the compiler accumulates into c's register
// (v_fmac) and then spends a move per node restoring the loop-carried
// registers (in FP32 as many moves as the unit form saves multiplies).  The
// result only feeds an ordinary FMA, which is what the DPP shifts read.
__device__ __forceinline__ double fma3(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
__device__ __forceinline__ float fma3(float a, float b, float c)
{
    float d;
    asm("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

template <class T, int K>
__device__ __forceinline__ void unit_step(const T (&src)[K], T (&dst)[K], T cl0, const T (&cm)[K], const T (&cr)[K])
{
    const T left = dpp_from_lower_lane(src[K - 1]);
    const T right = dpp_from_upper_lane(src[0]);
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const T nx = (s + 1 < K) ? src[s + 1] : right;
        T v;
        if (s == 0) {
            v = cm[0] * src[0];
            v = fma(cl0, left, v);
        } else {
            v = fma3(cm[s], src[s], src[s - 1]);
        }
        dst[s] = fma(cr[s], nx, v);
    }
}

template <class T, int K>
__device__ __forceinline__ void unit_steps(T (&z)[K], T cl0, const T (&cm)[K], const T (&cr)[K], int nsteps)
{
    T t[K];
    for (int it = 0; it + 1 < nsteps; it += 2) {
        unit_step<T, K>(z, t, cl0, cm, cr);
        unit_step<T, K>(t, z, cl0, cm, cr);
    }
    if (nsteps & 1) {
        unit_step<T, K>(z, t, cl0, cm, cr);
#pragma unroll
        for (int s = 0; s < K; ++s) z[s] = t[s];
    }
}

// Range the unit form's scales must stay in: D^{-1} z and the transformed
// coefficients then stay finite for any state below about 1e208.
template <class T> struct UnitRange;
template <> struct UnitRange<double> { static constexpr double dmin = 1e-100; };
template <> struct UnitRange<float> { static constexpr double dmin = 1e-30; };

// Whether every scale d of every window of this option lies in
// [dmin, 1/dmin]: a scale is a product of at most K-1 sub-diagonal entries
// (K = ceil(w/64)), each interior entry is a standard probability
// (between 1/6 - jmax|M|/2 and 1/6 + (jmax^2 M^2 + jmax|M|)/2) times a node
// discount in [e^{-X}, e^{X}], X = jmax dr dt, and the edge entries are the
// table's sub-diagonals of nodes 1, 2, w-3, w-2, w-1 (node 0 always starts a
// lane).

// x^f for a small count f >= 0, by repeated squaring.
HWT_HD double pow_count(double x, int f)
{
    double r = 1.0;
    for (; f > 0; f >>= 1, x *= x)
        if (f & 1) r *= x;
    return r;
}

// eX = e^{X} and emX = e^{-X} are the discounts of nodes 0 and w-1.  The range
// test lo^f >= dmin, hi^f <= 1/dmin is taken on powers rather than logarithms:
// an underflow or overflow on the way only makes it fail.
HWT_HD bool unit_form_ok(const TreeShape &t, const double *tab, double dmin, double eX, double emX)
{
    const double jm = t.jmax * fabs(t.M);
    double lo = (1.0 / 6.0 - 0.5 * jm) * emX, hi = (1.0 / 6.0 + 0.5 * (jm * jm + jm)) * eX;
    for (int pass = 0; pass < 2; ++pass)
        for (int slot = 1; slot < 6; ++slot) {
            const double a = fabs(tab[18 * pass + 3 * slot]);
            lo = a < lo ? a : lo;
            hi = a > hi ? a : hi;
        }
    if (!(lo > 0.0) || !(hi < 1e300)) return false;
    const int f = (t.w + WAVE - 1) / WAVE - 1;  // most factors in one lane
    return pow_count(lo < 1.0 ? lo : 1.0, f) >= dmin && pow_count(hi > 1.0 ? hi : 1.0, f) <= 1.0 / dmin;
}

HWT_HD bool unit_form_ok(const TreeShape &t, const double *tab, double dmin)
{
    const double X = t.jmax * t.dr * t.dt;
    return unit_form_ok(t, tab, dmin, exp(X), exp(-X));
}

// One phase: bring the window in (initial values, or the previous window's
// nodes [in_lo, in_hi) from LDS, zeros elsewhere), run nsteps stencil steps,
// and leave the window's tree nodes in LDS, or in `out` when given.
template <class T, int K, bool FWD>
__device__ __forceinline__ void phase_fn(const TreeShape &t, int base, int nsteps, int in_lo, int in_hi, int lane,
                                         typename Lds<T>::type *lds, const lds_f64 *etab, T *out, bool unit)
{
    T z[K];
    if (in_lo >= in_hi) {
        init_window<T, K, FWD>(z, t, base, lane);
    } else {
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int k = base + lane * K + s;
            z[s] = (k >= in_lo && k < in_hi) ? lds[k] : (T)0;
        }
    }
    __syncthreads();
    if (nsteps > 0) {
        bool done = false;
        if constexpr (K > 1 && K <= UnitMaxK<T>::value) {
            if (unit) {
                T cl0, um[K], ur[K];
                unit_in<T, K, FWD>(t, base, lane, etab, z, cl0, um, ur, lds);
                unit_steps<T, K>(z, cl0, um, ur, nsteps);
                unit_out<T, K>(base, lane, z, lds);
                done = true;
            }
        }
        if (!done) {
            T cl[K], cm[K], cr[K];
            window_coefs<T, K, FWD>(t, base, lane, etab, cl, cm, cr);
            stencil_steps<T, K>(z, cl, cm, cr, nsteps);
        }
    }
    if (out) {
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int k = base + lane * K + s;
            if (k < t.w) out[k] = z[s];
        }
    } else {
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int k = base + lane * K + s;
            if (k < t.w) lds[k] = z[s];
        }
    }
    __syncthreads();
}

#define HWT_KCASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) \
    X(14) X(15) X(16) X(17) X(18) X(19) X(20) X(21)

template <class T, int KMAX, bool FWD>
__device__ __forceinline__ void dispatch_phase(int K, const TreeShape &t, int base, int nsteps, int in_lo, int in_hi,
                                               int lane, typename Lds<T>::type *lds, const lds_f64 *etab, T *out,
                                               bool unit)
{
    switch (K) {
#define HWT_C(KK)                                                                                        \
    case KK:                                                                                             \
        if (KK <= KMAX)                                                                                  \
            phase_fn<T, (KK <= KMAX ? KK : 1), FWD>(t, base, nsteps, in_lo, in_hi, lane, lds, etab, out, unit); \
        break;
        HWT_KCASES(HWT_C)
#undef HWT_C
    default: break;
    }
}

struct Window {
    int K, base;
};

// Smallest centred window of half width 32K, K a multiple of G, that can hold
// half-width D, or the full window when none fits inside the tree.  Windows
// grow (forward) and shrink (backward) G nodes per lane at a time: a larger G
// steps a few padding nodes more but changes window -- and pays the
// coefficient set-up of a phase -- G times less often.
template <int G>
__device__ __forceinline__ Window window_for(int D, const TreeShape &t)
{
    const int Kg = (D / 32 / G + 1) * G;
    if (32 * Kg <= t.jmax) return Window{Kg, t.jmax - 32 * Kg};
    return Window{(t.w + WAVE - 1) / WAVE, 0};
}

// Window schedules of the two passes: the phases (window, steps) in order.
// The pricing passes below run a phase at each; the step counter at sync
// time (k_count_steps) walks the same schedules to count what they executed.
// visit(win, nsteps, in_lo, in_hi, last): in_lo/in_hi bound the previous
// window's tree nodes (empty for the first phase); last marks the forward
// pass's final phase.
template <int G, class Visit>
__device__ __forceinline__ void forward_schedule(const TreeShape &t, int e, int &lo, int &hi, Visit &&visit)
{
    Window win = window_for<G>(0, t);
    int in_lo = 0, in_hi = 0;  // empty: initial values
    int i = 0;
    for (;;) {
        const bool full = 32 * win.K > t.jmax;  // growth windows satisfy 32K <= jmax
        const int iend = full ? e : min(e, 32 * win.K - 1);
        const bool last = iend >= e;
        visit(win, iend - i, in_lo, in_hi, last);
        i = iend;
        in_lo = win.base;
        in_hi = min(win.base + WAVE * win.K, t.w);
        if (last) break;
        win = window_for<G>(i + 1, t);
    }
    lo = in_lo;
    hi = in_hi;
}

// The backward pass keeps only the nodes that can still reach the exercise
// nodes |j| <= min(e, jmax).
template <int G, class Visit>
__device__ __forceinline__ void backward_schedule(const TreeShape &t, int n, int e, int &lo, int &hi, Visit &&visit)
{
    const int ep = min(e, t.jmax);
    const int Kgmax = t.jmax / 32 / G * G;  // the largest window below the full one (0: none)
    int s = n;
    Window win = window_for<G>(min(t.jmax, ep + (s - e)), t);
    int in_lo = 0, in_hi = 0;
    while (s > e) {
        const bool full = 32 * win.K > t.jmax;
        const int klo = full ? Kgmax : win.K - G;  // the next window down
        const int send = max(e, e + 32 * klo - 1 - ep);
        visit(win, s - send, in_lo, in_hi, false);
        s = send;
        in_lo = win.base;
        in_hi = min(win.base + WAVE * win.K, t.w);
        if (s > e) win = window_for<G>(min(t.jmax, ep + (s - e)), t);
    }
    lo = in_lo;
    hi = in_hi;
}

// Whether a phase of window size K runs the unit-diagonal form.
template <class T>
__device__ __forceinline__ bool unit_phase(int K, bool unit)
{
    return unit && K > 1 && K <= UnitMaxK<T>::value;
}

// Forward pass: state 0 (delta at the root) -> state e.  The state ends up in
// out[lo, hi) (global scratch); tree nodes outside that range are zero.
template <class T, int KMAX, int G>
__device__ __forceinline__ void forward_pass(const TreeShape &t, int e, int lane, typename Lds<T>::type *lds,
                                             const lds_f64 *etab, T *out, int &lo, int &hi, bool unit)
{
    forward_schedule<G>(t, e, lo, hi, [&](const Window &win, int nsteps, int in_lo, int in_hi, bool last) {
        dispatch_phase<T, KMAX, true>(win.K, t, win.base, nsteps, in_lo, in_hi, lane, lds, etab,
                                      last ? out : nullptr, unit);
    });
}

// Backward pass: state n (bond face value 100) -> state e.  The state ends up
// in lds[lo, hi).
template <class T, int KMAX, int G>
__device__ __forceinline__ void backward_pass(const TreeShape &t, int n, int e, int lane, typename Lds<T>::type *lds,
                                              const lds_f64 *etab, int &lo, int &hi, bool unit)
{
    backward_schedule<G>(t, n, e, lo, hi, [&](const Window &win, int nsteps, int in_lo, int in_hi, bool) {
        dispatch_phase<T, KMAX, false>(win.K, t, win.base, nsteps, in_lo, in_hi, lane, lds, etab, (T *)nullptr,
                                       unit);
    });
}

// FP64 accumulators for the duality sums: plain in FP64 mode; compensated
// (Neumaier) in mixed mode, where the FP32 tree states are summed in FP64.
struct PlainSum {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double total() const { return s; }
    __device__ __forceinline__ double wave_total() const { return wave_sum(s); }
    __device__ __forceinline__ double row_total() const { return row_sum(s); }
};

struct KahanSum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x)
    {
        const double t = s + x;
        c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double total() const { return s + c; }
    __device__ __forceinline__ double wave_total() const { return wave_sum(s) + wave_sum(c); }
    __device__ __forceinline__ double row_total() const { return row_sum(s) + row_sum(c); }
};

template <class T> struct SumOf { typedef PlainSum type; };
template <> struct SumOf<float> { typedef KahanSum type; };

} // namespace hwt
