// hwt_math.h -- Hull-White one-factor trinomial-tree math shared by the
// device kernels: per-option constants, yield-curve interpolation, branching
// probabilities and the edge-corrected tridiagonal propagation operator.
//
// Formulas follow the reference's domain layer so integer tree quantities
// (height n, jmax, exercise step) come out identical:
//   constants          common/OptionConstants.hpp:29-53
//   yield interpolation common/Domain.hpp:11-46
//   probabilities      common/Domain.hpp:50-96 (Hull-White 1996 exhibits 1A/1B/1C)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HWT_HD __host__ __device__ __forceinline__

namespace hwt {

// Option status codes written by the setup kernel.
enum : int32_t { ST_OK = 0, ST_UNSUPPORTED = 1, ST_INVALID = 2 };

// Per-option constants, computed once per execution on the device.
struct OptConst {
    double dt, dr, M, X;   // years per step, rate spacing, e^{-a dt}-1, strike
    double Pe, Pn;         // discount bonds P(e dt), P(n dt) from the yield curve
    double rho, rhop;      // edge-elimination factors (bottom, top)
    int32_t jmax, n, e, w; // half width, height, exercise step (-1: none), width
    int16_t type, unit;    // CALL 0 / PUT 1; 1 if the unit-diagonal stencil form is safe
    int32_t status;        // ST_*
    int64_t work;          // stencil node updates this option needs
};
static_assert(sizeof(OptConst) == 96, "OptConst layout");

// ---------------------------------------------------------------- yield curve
// Linear interpolation in days with flat extrapolation (common/Domain.hpp:11-46).
HWT_HD double yield_at(double t, int termUnit, const double *rates, const int32_t *days, int size)
{
    const int tDays = (int)round(t * termUnit);
    int first = 0, second = 0;
    for (int i = 0; i < size; ++i) {
        if (days[i] >= tDays) {
            second = i;
            break;
        }
        first = i;
    }
    if (first == second) return rates[0];
    if (first == size - 1) return rates[size - 1];
    const int t1 = days[first], t2 = days[second];
    const double p1 = rates[first], p2 = rates[second];
    const double coef = (tDays - t1) / (double)(t2 - t1);
    return p1 + coef * (p2 - p1);
}

// The same value by binary search, for curves whose days never decrease: the
// linear scan above stops at the first knot with day >= tDays, which is the
// lower bound; with no such knot it returns the last rate, and with the first
// knot already past tDays the first rate (so one knot gives rates[0] either way).
HWT_HD double yield_at_sorted(double t, int termUnit, const double *rates, const int32_t *days, int size)
{
    const int tDays = (int)round(t * termUnit);
    int lo = 0, hi = size;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (days[mid] >= tDays) hi = mid;
        else lo = mid + 1;
    }
    if (lo == 0) return rates[0];
    if (lo == size) return rates[size - 1];
    const int t1 = days[lo - 1], t2 = days[lo];
    const double p1 = rates[lo - 1], p2 = rates[lo];
    const double coef = (tDays - t1) / (double)(t2 - t1);
    return p1 + coef * (p2 - p1);
}

// ------------------------------------------------------------- probabilities
// Standard branching (-jmax < j < jmax): up, mid, down.
HWT_HD double pu_std(int j, double M) { return 1.0 / 6.0 + ((double)(j * j) * M * M + (double)j * M) * 0.5; }
HWT_HD double pm_std(int j, double M) { return 2.0 / 3.0 - (double)(j * j) * M * M; }
HWT_HD double pd_std(int j, double M) { return 1.0 / 6.0 + ((double)(j * j) * M * M - (double)j * M) * 0.5; }
// Bottom edge (j == -jmax) branches to j+2, j+1, j.
HWT_HD double pu_bot(int j, double M) { return 1.0 / 6.0 + ((double)(j * j) * M * M - (double)j * M) * 0.5; }
HWT_HD double pm_bot(int j, double M) { return -1.0 / 3.0 - (double)(j * j) * M * M + (2.0 * j) * M; }
HWT_HD double pd_bot(int j, double M) { return 7.0 / 6.0 + ((double)(j * j) * M * M - (3.0 * j) * M) * 0.5; }
// Top edge (j == jmax) branches to j, j-1, j-2.
HWT_HD double pu_top(int j, double M) { return 7.0 / 6.0 + ((double)(j * j) * M * M + (3.0 * j) * M) * 0.5; }
HWT_HD double pm_top(int j, double M) { return -1.0 / 3.0 - (double)(j * j) * M * M - (2.0 * j) * M; }
HWT_HD double pd_top(int j, double M) { return 1.0 / 6.0 + ((double)(j * j) * M * M + (double)j * M) * 0.5; }

// Discount of node k over one step, without the alpha shift: e^{-(j dr) dt}.
HWT_HD double node_disc(int k, int jmax, double dr, double dt) { return exp(-((double)(k - jmax) * dr) * dt); }

// ---------------------------------------------------------------- operators
// Node indices k = j + jmax in [0, w).  The forward operator in "undiscounted
// Arrow-Debreu" variables y_i (Q_i = e^{-A_{i-1}} y_i, A the running sum of
// alpha*dt) is  y_{i+1} = F y_i  with  F[r][s] = p(s -> r) * E_s.  The bond
// backward operator is exactly its transpose.  F is tridiagonal except for the
// two edge branchings F[2][0] (bottom -> j+2) and F[w-3][w-1] (top -> j-2),
// which the similarity  G = T F T^{-1},  T = I + rho e1 e0^T + rhop e_{w-2} e_{w-1}^T
// (rho = F[2][0]/F[2][1], rhop = F[w-3][w-1]/F[w-3][w-2]) eliminates, so both
// passes become pure three-point stencils.

struct TreeShape {
    int32_t w, jmax;
    double M, dr, dt, rho, rhop;
};

// p(s -> r) for 0 <= r, s < w.
HWT_HD double branch_prob(const TreeShape &t, int r, int s)
{
    if (s == 0) {
        const int j = -t.jmax;
        if (r == 0) return pd_bot(j, t.M);
        if (r == 1) return pm_bot(j, t.M);
        if (r == 2) return pu_bot(j, t.M);
        return 0.0;
    }
    if (s == t.w - 1) {
        const int j = t.jmax;
        if (r == t.w - 1) return pu_top(j, t.M);
        if (r == t.w - 2) return pm_top(j, t.M);
        if (r == t.w - 3) return pd_top(j, t.M);
        return 0.0;
    }
    const int j = s - t.jmax;
    if (r == s + 1) return pu_std(j, t.M);
    if (r == s) return pm_std(j, t.M);
    if (r == s - 1) return pd_std(j, t.M);
    return 0.0;
}

// Node discounts E_k come from `disc(k)`, so callers can supply a table.
struct ExactDisc {
    const TreeShape &t;
    HWT_HD double operator()(int k) const { return node_disc(k, t.jmax, t.dr, t.dt); }
};

template <class Disc>
HWT_HD double op_F(const TreeShape &t, const Disc &disc, int r, int s)
{
    if (r < 0 || s < 0 || r >= t.w || s >= t.w) return 0.0;
    const double p = branch_prob(t, r, s);
    return p == 0.0 ? 0.0 : p * disc(s);
}

template <class Disc>
HWT_HD double op_FTinv(const TreeShape &t, const Disc &disc, int r, int s)
{
    double v = op_F(t, disc, r, s);
    if (s == 0) v -= t.rho * op_F(t, disc, r, 1);
    if (s == t.w - 1) v -= t.rhop * op_F(t, disc, r, t.w - 2);
    return v;
}

// G[r][s] of the tridiagonal operator.
template <class Disc>
HWT_HD double op_G(const TreeShape &t, const Disc &disc, int r, int s)
{
    if (r < 0 || s < 0 || r >= t.w || s >= t.w) return 0.0;
    double v = op_FTinv(t, disc, r, s);
    if (r == 1) v += t.rho * op_FTinv(t, disc, 0, s);
    if (r == t.w - 2) v += t.rhop * op_FTinv(t, disc, t.w - 1, s);
    return v;
}

HWT_HD double op_G(const TreeShape &t, int r, int s) { return op_G(t, ExactDisc{t}, r, s); }

HWT_HD void edge_factors(TreeShape &t)
{
    const ExactDisc d{t};
    t.rho = 0.0;
    t.rhop = 0.0;
    t.rho = op_F(t, d, 2, 0) / op_F(t, d, 2, 1);
    t.rhop = op_F(t, d, t.w - 3, t.w - 1) / op_F(t, d, t.w - 3, t.w - 2);
}

// ------------------------------------------------------------- work counts
// S(m) = sum_{u=1..m} min(u, J)
HWT_HD int64_t sum_min(int64_t m, int64_t J)
{
    if (m <= 0) return 0;
    if (m <= J) return m * (m + 1) / 2;
    return J * (J + 1) / 2 + (m - J) * J;
}

// Node updates of this pricer: forward 0 -> e over the active region, backward
// n -> e over the light cone of the exercise nodes.
HWT_HD int64_t dual_node_updates(int32_t n, int32_t e, int32_t jmax)
{
    if (n <= 0 || e < 0 || e >= n) return 0;
    const int64_t ep = e < jmax ? e : jmax;
    const int64_t fwd = 2 * sum_min(e, jmax) + e;
    const int64_t bwd = 2 * (sum_min(ep + (n - 1 - e), jmax) - sum_min(ep - 1, jmax)) + (n - e);
    return fwd + bwd;
}

// Node updates of the reference's seq (forward i=1..n plus backward i=0..n-1).
HWT_HD int64_t seq_node_updates(int32_t n, int32_t jmax)
{
    if (n <= 0) return 0;
    return (2 * sum_min(n, jmax) + n) + (2 * sum_min(n - 1, jmax) + n);
}

// ------------------------------------------------------------- constants
// common/OptionConstants.hpp:29-53, plus validation the reference lacks
// (it divides by zero or overflows int on these inputs).  Returns ST_*.
HWT_HD int32_t option_constants(double strike, double T, double len, uint16_t termUnit,
                                uint16_t termSteps, double a, double sigma, int8_t type,
                                int32_t maxWidth, OptConst &c)
{
    c.X = strike;
    c.type = type;
    c.unit = 0;
    c.Pe = c.Pn = c.rho = c.rhop = 0.0;
    c.work = 0;
    c.n = c.jmax = c.w = 0;
    c.e = -1;
    if ((type != 0 && type != 1) || termUnit == 0 || termSteps == 0 || !(a > 0.0) ||
        !(sigma >= 0.0) || !(T >= 0.0) || !(strike == strike))
        return ST_INVALID;
    const int tuy = (int)ceil(365.0 / termUnit);
    const double nReal = (int)termSteps * tuy * T;
    if (!(nReal < 2147483647.0)) return ST_INVALID;
    c.n = (int32_t)nReal;
    c.dt = tuy / (double)termSteps;
    const double V = sigma * sigma * (1.0 - exp(-2.0 * a * c.dt)) / (2.0 * a);
    c.dr = sqrt(3.0 * V);
    c.M = exp(-a * c.dt) - 1.0;
    const double q = -0.184 / c.M;
    if (!(q < 1.0e9)) return ST_INVALID;
    c.jmax = (int32_t)q + 1;
    c.w = 2 * c.jmax + 1;
    // isMaturity step, seq/Seq.hpp:125: i == (int)(t / dt) for i in [0, n-1]
    const double qe = len / c.dt;
    c.e = (qe > -1.0 && qe < (double)c.n) ? (int32_t)qe : -1;
    if (c.w > maxWidth) return ST_UNSUPPORTED;
    return ST_OK;
}

// Constants given directly (the reference's OptionConstants, as
// seq::computeSingleOption receives them, seq/Seq.hpp:24): the exercise step
// as seq/Seq.hpp:125 derives it, plus the validation the other path does.
// With f32, t and dt are floats widened to double and the step quotient is
// taken in float, as a real = float caller takes it: the double quotient of
// two floats rounds to the float quotient exactly (53 >= 2 * 24 + 2 bits).
HWT_HD int32_t constants_given(double t, double dt, double dr, double X, double M, int32_t jmax, int32_t n,
                               int32_t width, uint16_t termUnit, int8_t type, int32_t maxWidth, bool f32,
                               OptConst &c)
{
    c.X = X;
    c.type = type;
    c.unit = 0;
    c.Pe = c.Pn = c.rho = c.rhop = 0.0;
    c.work = 0;
    c.n = c.jmax = c.w = 0;
    c.e = -1;
    if ((type != 0 && type != 1) || termUnit == 0 || !(dt > 0.0) || !(dr >= 0.0) || !(M > -1.0 && M < 0.0) ||
        jmax < 1 || width != 2 * jmax + 1 || n < 0 || !(X == X) || !(dr < 1e300) || !(t == t))
        return ST_INVALID;
    c.n = n;
    c.dt = dt;
    c.dr = dr;
    c.M = M;
    c.jmax = jmax;
    c.w = width;
    const double qe = f32 ? (double)(float)(t / dt) : t / dt;
    c.e = (qe > -1.0 && qe < (double)c.n) ? (int32_t)qe : -1;
    if (c.w > maxWidth) return ST_UNSUPPORTED;
    return ST_OK;
}

} // namespace hwt
