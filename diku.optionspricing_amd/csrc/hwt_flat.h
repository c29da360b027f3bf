// hwt_flat.h -- tiny trees fully flattened across a wavefront.
//
// A tree at most FLAT_MAX_WIDTH nodes wide leaves most of a 16-lane row idle
// (a 7-node tree uses 7 of 16 lanes).  Here a wave holds floor(64 / w) trees
// of one width w side by side, one node per lane: tree q owns lanes
// [q w, (q+1) w), its node k on lane q w + k.  A step is one three-point
// stencil over all 64 lanes with DPP whole-wave shifts (wave_shr:1 /
// wave_shl:1): a value that crosses into a neighbouring tree meets a zero
// coefficient (node 0 has no lower branch, node w-1 no upper one), so the
// trees never mix.  This is the wave-level form of the reference's fully
// flattened version, where all nodes of all options form one segmented array
// (futhark/trinom-flat.fut:259-385).
//
// Each lane runs its own tree's sequence of n steps: forward to its exercise
// step e, then -- at its own e -- it parks the forward state, switches to the
// backward coefficients and start values, and at its own n freezes with
// identity coefficients; the wave stops at each step where some tree changes
// phase (a wave-wide minimum over the lanes' next events).  The duality sums
// are then taken per tree, node by node in a fixed order, by the lane of the
// tree's index, so a price does not depend on which trees share the wave.
#pragma once

#include "hwt_rows.h"

#include <climits>

namespace hwt {

constexpr int FLAT_MAX_WIDTH = 21;
constexpr int FLAT_PARTS = (FLAT_MAX_WIDTH - 3) / 2 + 1;  // odd widths 3..21

constexpr HWT_HD int flat_part(int w) { return (w - 3) / 2; }
constexpr HWT_HD int flat_width(int part) { return 3 + 2 * part; }
constexpr HWT_HD int flat_per_wave(int w) { return WAVE / w; }

// Minimum over the 64 lanes, the same in every lane (all lanes active).
__device__ __forceinline__ int wave_min_i32(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0xB1, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x4E, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x141, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x140, 0xF, 0xF, false));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

template <class T>
__device__ __forceinline__ void flat_steps(T &z, T cl, T cm, T cr, int nsteps)
{
    for (int it = 0; it < nsteps; ++it) {
        const T left = dpp_from_lower_lane(z);
        const T right = dpp_from_upper_lane(z);
        T v = cm * z;
        v = fma(cl, left, v);
        z = fma(cr, right, v);
    }
}

// Prices the trees of one flat group.  Called by all 64 lanes together (it
// has workgroup barriers).  `active` is false on lanes past the group's trees
// and on the lanes of an idle (padding) tree, whose e and n are then 0.
// ybuf/bbuf are 64 LDS slots each, sbuf 4 per tree (strike, P(e dt), P(n dt),
// type).  Returns, in lane q (q < trees), the price
// of tree q; NaN when its FP32 state (mixed mode) left FP32's range.
template <class T>
__device__ __forceinline__ double price_flat(const RowOpt &r, int k, bool active, int trees, int w,
                                             lds_f64 *ybuf, lds_f64 *bbuf, lds_f64 *sbuf)
{
    const int lane = threadIdx.x;
    const int e_l = active ? r.e : 0, n_l = active ? r.n : 0;
    T cl = 0, cm = 0, cr = 0, bl = 0, bm = 0, br = 0, z = 0, zf = 0;
    if (active) {
        T a[1], b[1], c[1];
        row_coefs<T, 1, true>(r, k, a, b, c);
        cl = a[0];
        cm = b[0];
        cr = c[0];
        row_coefs<T, 1, false>(r, k, a, b, c);
        bl = a[0];
        bm = b[0];
        br = c[0];
        row_init<T, 1, true>(a, r, k);
        z = a[0];
    }
    int nmax = n_l;
    for (int off = 32; off >= 1; off >>= 1) nmax = max(nmax, __shfl_xor(nmax, off, WAVE));
    int i = 0;
    for (;;) {
        // events at step i: a tree reaching its exercise step turns
        // backward, a tree reaching its height freezes
        if (i == e_l && active) {
            zf = z;
            cl = bl;
            cm = bm;
            cr = br;
            T a[1];
            row_init<T, 1, false>(a, r, k);
            z = a[0];
        }
        if (i == n_l) {
            cl = (T)0;
            cm = (T)1;
            cr = (T)0;
        }
        if (i >= nmax) break;
        const int next = wave_min_i32(e_l > i ? e_l : (n_l > i ? n_l : INT_MAX));
        flat_steps<T>(z, cl, cm, cr, next - i);
        i = next;
    }

    // undo the edge similarity (y = T^{-1} zf, b = T^T z), then the sums per
    // tree over its node slots of LDS, node by node
    const TreeShape &t = r.t;
    const double fL = dpp_from_lower_lane((double)zf), fR = dpp_from_upper_lane((double)zf);
    const double bL = dpp_from_lower_lane((double)z), bR = dpp_from_upper_lane((double)z);
    double yk = zf, bk = z;
    if (k == 1) yk -= t.rho * fL;
    if (k == t.w - 2) yk -= t.rhop * fR;
    if (k == 0) bk += t.rho * bR;
    if (k == t.w - 1) bk += t.rhop * bL;
    ybuf[lane] = active ? yk : 0.0;
    bbuf[lane] = active ? bk : 0.0;
    // the tree's own constants, for the lane that sums it
    if (active && k == 0) {
        sbuf[4 * (lane / w)] = r.X;
        sbuf[4 * (lane / w) + 1] = r.Pe;
        sbuf[4 * (lane / w) + 2] = r.Pn;
        sbuf[4 * (lane / w) + 3] = r.type;
    }
    __syncthreads();
    double price = 0.0;
    if (lane < trees) {
        const int base = lane * w;
        typename SumOf<T>::type a1, a2, a3;
        bool big = false;
        for (int j = 0; j < w; ++j) {
            const double yv = ybuf[base + j], bv = bbuf[base + j];
            a1.add(yv);
            a2.add(yv * bv);
            big |= !(fabs(bv) <= F32_SAFE_STATE);
        }
        const double s1 = a1.total(), s2 = a2.total();
        const double X = sbuf[4 * lane], Pe = sbuf[4 * lane + 1], Pn = sbuf[4 * lane + 2];
        const int type = (int)sbuf[4 * lane + 3];
        const double scale = 100.0 * s1 * Pn / (s2 * Pe);
        for (int j = 0; j < w; ++j) {
            const double yv = ybuf[base + j], Bv = scale * bbuf[base + j];
            double v = (type == 1) ? X - Bv : Bv - X;  // PUT : CALL
            v = (v < 0.0) ? 0.0 : v;
            v = (v > RealMax<T>::value) ? RealMax<T>::value : v;
            a3.add(yv * v);
        }
        const double s3 = a3.total();
        price = Pe / s1 * s3;
        if (sizeof(T) == sizeof(float) && (big || !isfinite(s1) || !isfinite(s2))) price = NAN;
    }
    __syncthreads();  // the buffers are reused by the wave's next group
    return price;
}

} // namespace hwt
