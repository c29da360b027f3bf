// hwt_rows.h -- narrow trees (width <= 64) packed four to a wavefront.
//
// Each 16-lane row of the wave prices one option: lane l of the row holds
// nodes l*K + s, s < K (K <= 4), and values cross lane boundaries with DPP
// row shifts (row_shr:1 / row_shl:1), which never leave the row.  One
// instruction stream therefore advances four trees at once.
//
// Each row runs its option's two passes as one sequence of n steps: forward
// for steps [0, e), then -- at its own exercise step -- it parks the forward
// state, loads the backward start values and the backward coefficients, and
// runs the remaining n - e steps backward; a row whose n steps are done gets
// identity coefficients (mid 1, sides 0), which leave its state unchanged to
// the bit.  The four rows differ only in register contents, so the stencil
// code stays uniform; the task list groups options by (K, n) so that the rows
// of a wave finish together.  The sums of the duality identity are formed
// per row over the nodes' LDS slots in a fixed order, so a price is the same
// bit for bit whichever options share its wave.
#pragma once

#include "hwt_wave.h"

namespace hwt {

constexpr int ROW = 16;
constexpr int ROWS = WAVE / ROW;
constexpr int ROW_MAXK = 4;
constexpr int ROW_MAX_WIDTH = ROW * ROW_MAXK;

__device__ __forceinline__ double row_from_lower(double x) { return dpp_mov<0x111>(x); }  // row_shr:1
__device__ __forceinline__ double row_from_upper(double x) { return dpp_mov<0x101>(x); }  // row_shl:1
__device__ __forceinline__ float row_from_lower(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_from_upper(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x101, 0xF, 0xF, true));
}

// The option of this lane's row (values differ between rows).
struct RowOpt {
    TreeShape t;
    int32_t n, e, type;
    double X, Pe, Pn;
    const double *edge;  // the option's 36-entry edge table
};

// Coefficients of this lane's K nodes (target node k = l*K + s): forward
// operator G (FWD) or its transpose, closed forms in the interior, the edge
// table on nodes 0, 1, 2, w-3, w-2, w-1, zero on padding nodes k >= w.
template <class T, int K, bool FWD>
__device__ __forceinline__ void row_coefs(const RowOpt &r, int l, T (&cl)[K], T (&cm)[K], T (&cr)[K])
{
    const TreeShape &t = r.t;
    const int k0 = l * K;
    const double q = exp(-t.dr * t.dt);
    double E = node_disc(k0 - 1, t.jmax, t.dr, t.dt);  // E_{k0-1}
    double Es[K + 2];
#pragma unroll
    for (int m = 0; m < K + 2; ++m) {
        const int k = k0 - 1 + m;
        Es[m] = (k >= 0 && k < t.w) ? E : 0.0;
        E *= q;
    }
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = k0 + s;
        double a, b, c;
        if (FWD) {  // sources k-1 (up), k (mid), k+1 (down)
            a = pu_std(k - 1 - t.jmax, t.M) * Es[s];
            b = pm_std(k - t.jmax, t.M) * Es[s + 1];
            c = pd_std(k + 1 - t.jmax, t.M) * Es[s + 2];
        } else {    // node k's own branches, discounted at k
            a = pd_std(k - t.jmax, t.M) * Es[s + 1];
            b = pm_std(k - t.jmax, t.M) * Es[s + 1];
            c = pu_std(k - t.jmax, t.M) * Es[s + 1];
        }
        if (k < t.w && (k < 3 || k > t.w - 4)) {
            const double *tab = r.edge + (FWD ? 0 : 18) + 3 * edge_slot_of(k, t.w);
            a = tab[0];
            b = tab[1];
            c = tab[2];
        }
        if (k >= t.w) a = b = c = 0.0;
        cl[s] = (T)a;
        cm[s] = (T)b;
        cr[s] = (T)c;
    }
}

// Start values: forward delta at the root, backward T^{-T} * 100.
template <class T, int K, bool FWD>
__device__ __forceinline__ void row_init(T (&z)[K], const RowOpt &r, int l)
{
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = l * K + s;
        double v;
        if (FWD) {
            v = (k == r.t.jmax) ? 1.0 : 0.0;
        } else {
            v = (k < r.t.w) ? 100.0 : 0.0;
            if (k == 0) v -= r.t.rho * 100.0;
            if (k == r.t.w - 1) v -= r.t.rhop * 100.0;
        }
        z[s] = (T)v;
    }
}

template <class T, int K>
__device__ __forceinline__ void row_step(const T (&src)[K], T (&dst)[K], const T (&cl)[K], const T (&cm)[K],
                                         const T (&cr)[K])
{
    const T left = row_from_lower(src[K - 1]);
    const T right = row_from_upper(src[0]);
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

template <class T, int K>
__device__ __forceinline__ void row_steps(T (&z)[K], const T (&cl)[K], const T (&cm)[K], const T (&cr)[K],
                                          int nsteps)
{
    T t[K];
    for (int it = 0; it + 1 < nsteps; it += 2) {
        row_step<T, K>(z, t, cl, cm, cr);
        row_step<T, K>(t, z, cl, cm, cr);
    }
    if (nsteps & 1) {
        row_step<T, K>(z, t, cl, cm, cr);
#pragma unroll
        for (int s = 0; s < K; ++s) z[s] = t[s];
    }
}

// Prices the four rows' options; returns each row's price in every lane of
// the row.  Called by all 64 lanes together (it has workgroup barriers).
// ev[0..3] are the rows' exercise steps and ev[4..7] their heights
// (wave-uniform copies); an idle row carries e = n = 0.  ybuf/bbuf are this
// row's 64 node slots of LDS (zeroed and filled here).
template <class T, int K>
__device__ __forceinline__ double price_rows(const RowOpt &r, int l, int e_row, int n_row, const int (&ev)[8],
                                             lds_f64 *ybuf, lds_f64 *bbuf)
{
    T cl[K], cm[K], cr[K], bl[K], bm[K], br[K], z[K], zf[K];
    row_coefs<T, K, true>(r, l, cl, cm, cr);
    row_coefs<T, K, false>(r, l, bl, bm, br);
    row_init<T, K, true>(z, r, l);
#pragma unroll
    for (int s = 0; s < K; ++s) zf[s] = z[s];
    int nmax = 0;
#pragma unroll
    for (int i = 4; i < 8; ++i) nmax = max(nmax, ev[i]);
    int i = 0;
    for (;;) {
        // events at step i: a row reaching its exercise step turns backward,
        // a row reaching its height freezes
        if (i == e_row) {
#pragma unroll
            for (int s = 0; s < K; ++s) {
                zf[s] = z[s];
                cl[s] = bl[s];
                cm[s] = bm[s];
                cr[s] = br[s];
            }
            row_init<T, K, false>(z, r, l);
        }
        if (i == n_row) {
#pragma unroll
            for (int s = 0; s < K; ++s) {
                cl[s] = (T)0;
                cm[s] = (T)1;
                cr[s] = (T)0;
            }
        }
        if (i >= nmax) break;
        int next = nmax;  // the next event step after i (wave-uniform)
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (ev[k] > i && ev[k] < next) next = ev[k];
        row_steps<T, K>(z, cl, cm, cr, next - i);
        i = next;
    }

    // undo the edge similarity (y = T^{-1} zf, b = T^T z); the sums are then
    // formed over node slots of the row's LDS buffer in a fixed order (lane l
    // adds nodes 4l .. 4l+3, then the row butterfly), so a price does not
    // depend on the K its wave happened to run with
    const TreeShape &t = r.t;
    const double fl = row_from_lower(zf[K - 1]), fr = row_from_upper(zf[0]);
    const double blo = row_from_lower(z[K - 1]), bhi = row_from_upper(z[0]);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        ybuf[4 * l + s] = 0.0;
        bbuf[4 * l + s] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = l * K + s;
        const double fL = s > 0 ? (double)zf[s - 1] : fl, fR = s + 1 < K ? (double)zf[s + 1] : fr;
        const double bL = s > 0 ? (double)z[s - 1] : blo, bR = s + 1 < K ? (double)z[s + 1] : bhi;
        double yk = zf[s], bk = z[s];
        if (k == 1) yk -= t.rho * fL;
        if (k == t.w - 2) yk -= t.rhop * fR;
        if (k == 0) bk += t.rho * bR;
        if (k == t.w - 1) bk += t.rhop * bL;
        if (k < t.w) {
            ybuf[k] = yk;
            bbuf[k] = bk;
        }
    }
    __syncthreads();
    double yv[4], bv[4];
    typename SumOf<T>::type a1, a2, a3;
    bool big = false;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        yv[s] = ybuf[4 * l + s];
        bv[s] = bbuf[4 * l + s];
        a1.add(yv[s]);
        a2.add(yv[s] * bv[s]);
        big |= !(fabs(bv[s]) <= F32_SAFE_STATE);
    }
    __syncthreads();  // the buffer is reused by the wave's next group
    const double s1 = a1.row_total(), s2 = a2.row_total();
    if constexpr (sizeof(T) == sizeof(float)) {
        // FP32 state of this row's option out of range: NaN asks for the
        // FP64 re-pricing (the row's 16 bits of the ballot)
        const int row = __lane_id() / ROW;
        if (((__ballot(big) >> (ROW * row)) & 0xFFFFull) != 0 || !isfinite(s1) || !isfinite(s2)) return NAN;
    }
    const double scale = 100.0 * s1 * r.Pn / (s2 * r.Pe);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        double v = (r.type == 1) ? r.X - scale * bv[s] : scale * bv[s] - r.X;  // PUT : CALL
        v = (v < 0.0) ? 0.0 : v;
        v = (v > RealMax<T>::value) ? RealMax<T>::value : v;
        a3.add(yv[s] * v);
    }
    const double s3 = a3.row_total();
    return r.Pe / s1 * s3;
}

} // namespace hwt
