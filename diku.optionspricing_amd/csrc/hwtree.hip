// hwtree.hip -- kernels and C ABI of the MI355X-native HW1F trinomial pricer.
//
// Pipeline of one execution (captured once into a HIP graph per plan):
//   k_setup    one thread per option: tree constants (the reference's
//              OptionConstants), P(e dt) and P(n dt) from the yield curve, edge
//              factors, work estimate, bucket and rank inside the bucket;
//              trivial options are priced right here
//   k_order    bucket offsets, largest work first (longest-processing-time
//              order; each block scans the histogram itself), and the option
//              indices into the task list; it clears the queue heads and the
//              setup accumulators for the next execution, and the first
//              pricing kernel clears the histogram, so no clearing
//              kernel or memset node runs per execution (k_reset only runs for
//              an empty batch, which has no k_setup)
//   k_price_wide / k_price_narrow  persistent kernels, launched side by side:
//              each wavefront pulls options off its part of the task list and
//              prices them with the on-chip tree column (hwt_wave.h); trees up to
//              128 nodes wide go to the narrow kernel, which needs far fewer
//              registers and so runs at higher occupancy
//
// Pricing identity.  With Q_e the Arrow-Debreu prices at the exercise step e
// and B_e the bond values there, the reference's backward induction gives
// price = sum_j Q_e[j] payoff(B_e[j]) exactly (Arrow-Debreu duality, same tree,
// same discount factors).  Because alpha only shifts every node of a step by
// the same discount, Q_e = (P(e dt)/S1) y_e and B_e = 100 S1 P(n dt)/(S2 P(e dt)) b_e
// where y_e is the undiscounted forward state after e steps, b_e the
// undiscounted backward state after n-e steps, S1 = sum y_e and S2 = y_e . b_e.
// So each option costs n stencil steps (e forward + n-e backward) with no
// per-step reductions, against 2n steps plus n reductions in seq/Seq.hpp.
#include "hwt_lane.h"
#include "hwt_pipeline.h"
#include "../../include/hwtree.h"
#include "hwt_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <chrono>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

using namespace hwt;

namespace {

__device__ __forceinline__ int log_bucket(int64_t key, int per_octave, int nb)
{
    int b = (int)(per_octave * log2((double)key + 1.0));
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

// Row class: 160 height buckets per nodes-per-lane count K (1..4): exact
// heights below 96, then 16 buckets per octave.
__device__ __forceinline__ int row_bucket(int w, int n)
{
    const int K = (w + ROW - 1) / ROW;
    const int h = n < 96 ? n : min(ROW_HBUCKETS - 1, 96 + (int)(16.0 * log2(n / 96.0)));
    return B_ROW + (K - 1) * ROW_HBUCKETS + h;
}

// Flat class: 64 height buckets per width: exact heights below 32, then 8
// buckets per octave.
__device__ __forceinline__ int flat_bucket(int w, int n)
{
    const int h = n < 32 ? n : min(FLAT_HBUCKETS - 1, 32 + (int)(8.0 * log2(n / 32.0)));
    return flat_part(w) * FLAT_HBUCKETS + h;
}

// The part of the row or flat class bucket b belongs to: 0..ROW_MAXK-1 the
// row parts (nodes per lane), then the flat parts (one per width); and the
// number of tasks a wave takes from that part at once.
__device__ __forceinline__ int part_of(int b) { return b >= B_ROW ? (b - B_ROW) / ROW_HBUCKETS : ROW_MAXK + b / FLAT_HBUCKETS; }
__device__ __forceinline__ int part_group(int part)
{
    return part < ROW_MAXK ? ROWS : flat_per_wave(flat_width(part - ROW_MAXK));
}
constexpr int NPARTS = ROW_MAXK + FLAT_PARTS;

// Every counter and tally to zero: the empty batch's whole pipeline.
__global__ void k_reset(int32_t *tally)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < T_ALLOC; i += blockDim.x * gridDim.x) tally[i] = 0;
}

// One thread per option; launched with SETUP_BLOCK threads per block.  The
// histogram and the work tallies are aggregated per block in LDS first, so the
// global atomics number a few per block instead of one per option (single
// counters hit by every option serialise in L2).
constexpr int SETUP_BLOCK = 256;
constexpr int SETUP_YMAX = 2048;  // knots staged in LDS by k_setup (24 KB)
constexpr int SCAN_PER = NBUCKET / SETUP_BLOCK;
static_assert(NBUCKET % SETUP_BLOCK == 0, "scan positions per thread");
static_assert(T_FLAT_START + FLAT_PARTS + 1 <= T_T0 && T_T0 + 2 <= T_FLAT_GROUP &&
                  T_FLAT_GROUP + FLAT_PARTS + 1 <= T_XFLOP && T_XPLAIN + 2 <= T_END && N_END <= SETUP_BLOCK,
              "tally layout");

// Bucket at scan position p: the wide class, then the mid class, each largest
// key first unless an ascending order was asked for, then the row class by
// decreasing (K, height), then the flat class by decreasing (width, height).
__device__ __forceinline__ int scan_bucket(int p, int asc)
{
    constexpr int NW = NBUCKET - B_WIDE, NM = B_WIDE - B_MID;
    if (p < NW) return asc ? B_WIDE + p : NBUCKET - 1 - p;
    if (p < NW + NM) return asc ? B_MID + (p - NW) : B_WIDE - 1 - (p - NW);
    return NBUCKET - 1 - p;
}

// Run by every block of k_order (SETUP_BLOCK threads), after k_setup: the
// exclusive offsets of the buckets in scan order, into soffs (LDS).  The
// writer block also puts the pads into the task list, the class boundaries
// and counts into the host-visible tally, and clears the queue heads and the
// setup accumulators for the next execution; the histogram, which every
// block reads, is cleared by the first pricing kernel.  Each part of the
// row class (one K) is padded to a multiple of four tasks and each part of the
// flat class (one width w) to a multiple of 64 / w, with idle slots (order
// -1), so every group a wave takes shares one K or one width and an option's
// arithmetic never depends on which options share its wave.  The flat parts'
// first tasks and first group numbers go to the tally for the flat kernel.
__device__ __forceinline__ void scan_offsets(const int32_t *__restrict__ hist, int32_t *soffs,
                                             int32_t *__restrict__ order, int32_t *__restrict__ tally, int32_t asc,
                                             int32_t classes, bool writer)
{
    const int tid = threadIdx.x;
    constexpr int NW = NBUCKET - B_WIDE, NM = B_WIDE - B_MID, NR = B_MID - B_ROW;
    __shared__ int32_t ptot[NPARTS];  // options per part
    __shared__ int32_t ppad[NPARTS];  // idle slots that round each part up to whole groups
    __shared__ int32_t wtot[SETUP_BLOCK / WAVE];
    __shared__ int32_t nrow_end;
    int idx[SCAN_PER], cnt[SCAN_PER];
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        idx[q] = scan_bucket(tid * SCAN_PER + q, asc);
        cnt[q] = hist[idx[q]];
    }
    if (tid < NPARTS) ptot[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q)
        if (idx[q] < B_MID && cnt[q]) atomicAdd(&ptot[part_of(idx[q])], cnt[q]);
    __syncthreads();
    // one thread per part works out its pad (a division by the part's group)
    if (tid < NPARTS) {
        const int g = part_group(tid);
        ppad[tid] = classes ? (g - ptot[tid] % g) % g : 0;
    }
    __syncthreads();
    // the lowest bucket of each part comes last in the scan and takes the pad
    int v[SCAN_PER], pad[SCAN_PER], sum = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        const int b = idx[q];
        const bool lowest = b < B_MID && (b >= B_ROW ? (b - B_ROW) % ROW_HBUCKETS == 0 : b % FLAT_HBUCKETS == 0);
        pad[q] = lowest ? ppad[part_of(b)] : 0;
        v[q] = cnt[q] + pad[q];
        sum += v[q];
    }
    // exclusive scan of the per-thread sums: inside each wave, then across
    // the block's waves
    const int lane = tid % WAVE, wv = tid / WAVE;
    int incl = sum;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const int t = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += t;
    }
    if (lane == WAVE - 1) wtot[wv] = incl;
    __syncthreads();
    int run = incl - sum;
    for (int k = 0; k < wv; ++k) run += wtot[k];
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        const int p = tid * SCAN_PER + q;
        soffs[idx[q]] = run;
        run += v[q];
        if (writer) {
            for (int j = 0; j < pad[q]; ++j) order[run - pad[q] + j] = -1;
            if (p == NW - 1) tally[T_NLARGE] = run;
            if (p == NW + NM - 1) tally[T_NMID] = run;
            if (p == NW + NM + NR - 1) {
                tally[T_NROW] = run;
                nrow_end = run;
            }
            if (p == NBUCKET - 1) tally[T_NTASK] = run;
        }
    }
    __syncthreads();  // orders nrow_end above before thread 0 reads it
    if (writer && tid < N_END) reinterpret_cast<unsigned long long *>(tally + T_END)[tid] = 0;
    if (writer && tid == 0) {
        // flat parts in task order, widest first: first task and first group
        int start = nrow_end, group = 0, npad = 0;
#pragma unroll
        for (int k = 0; k < NPARTS; ++k) npad += ppad[k];
#pragma unroll
        for (int j = 0; j < FLAT_PARTS; ++j) {
            const int fp = FLAT_PARTS - 1 - j;
            const int size = ptot[ROW_MAXK + fp] + ppad[ROW_MAXK + fp];
            tally[T_FLAT_START + j] = start;
            tally[T_FLAT_GROUP + j] = group;
            start += size;
            group += size / flat_per_wave(flat_width(fp));
        }
        tally[T_FLAT_START + FLAT_PARTS] = start;
        tally[T_FLAT_GROUP + FLAT_PARTS] = group;
        tally[T_NFLATG] = group;
        tally[T_NPAD] = npad;
        tally[T_QUEUE] = 0;
        tally[T_QUEUE_S] = 0;
        tally[T_QUEUE_R] = 0;
        tally[T_QUEUE_F] = 0;
        int64_t *t64 = reinterpret_cast<int64_t *>(tally);
        t64[T_WORK / 2] = t64[A_WORK / 2];
        t64[T_SEQWORK / 2] = t64[A_SEQWORK / 2];
        t64[T_WORKW / 2] = t64[A_WORKW / 2];
        tally[T_INVALID] = tally[A_INVALID];
        tally[T_UNSUP] = tally[A_UNSUP];
        tally[T_NREDO] = 0;
        tally[T_QUEUE_X] = 0;
        for (int k = T_XFLOP; k < T_XPLAIN + 2; ++k) tally[k] = 0;
        t64[A_WORK / 2] = 0;
        t64[A_SEQWORK / 2] = 0;
        t64[A_WORKW / 2] = 0;
        tally[A_INVALID] = 0;
        tally[A_UNSUP] = 0;
    }
}

// Per-wave LDS staging of k_setup's outputs.  A lane's record (its 288-byte
// edge table, its 96-byte constants) is written to LDS and the wave copies the
// records out as contiguous 1 KB rows instead of one 16-byte piece per lane
// and store (25.7 -> 22.7 us on the 60k skewed batch).  The kernel is bound
// by writing its 24.6 MB, not by its arithmetic.
constexpr int REC_TAB = EDGE_TAB / 2;                 // 16-byte units per edge table
constexpr int REC_CST = (int)(sizeof(OptConst) / 16); // ... per constants record
constexpr int STAGE_LANES = WAVE / 2;                 // tables staged per round
constexpr int STAGE_UNITS = STAGE_LANES * REC_TAB;    // per wave: 9 KB
static_assert(sizeof(OptConst) % 16 == 0 && WAVE * REC_CST <= STAGE_UNITS, "constants fit one round");

// Orders this wave's LDS accesses on either side (the staging slice is the
// wave's own, so no other wave takes part).
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Sum of a count over the wave, in every lane: exact through the FP64 DPP
// reduction while the total stays below 2^53.
__device__ __forceinline__ unsigned long long wave_total_count(unsigned long long v)
{
    return (unsigned long long)wave_sum((double)v);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Copies `units` (at most MAXU) 16-byte units from the slice to dst, 64
// lanes at a time, every LDS read issued before the first store.  (Stores
// marked non-temporal measured slower overall: k_setup gains 5 us, the
// pricing kernels, reading the records back past L2, lose 6.)
template <int MAXU>
__device__ __forceinline__ void copy_out(u32x4 *__restrict__ dst, const u32x4 *slice, int units, int lane)
{
    constexpr int R = (MAXU + WAVE - 1) / WAVE;
    u32x4 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (lane + r * WAVE < units) v[r] = slice[lane + r * WAVE];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (lane + r * WAVE < units) dst[lane + r * WAVE] = v[r];
}

__global__ __launch_bounds__(SETUP_BLOCK) void k_setup(DevOpts o, DevYield y, int32_t sort, int32_t maxWidth,
                                                       double unit_dmin, int32_t classes,
                                                       OptConst *__restrict__ cst, double *__restrict__ edge,
                                                       int2 *__restrict__ slot, int32_t *__restrict__ hist,
                                                       int32_t *__restrict__ tally, double *__restrict__ out)
{
    __shared__ int32_t lh[NBUCKET];  // the block's histogram, then its first rank in each bucket
    __shared__ unsigned long long lsum[3];  // work, seq work, wide work
    __shared__ int32_t lbad[2];             // invalid, unsupported
    __shared__ double yr[SETUP_YMAX];       // the yield curve, staged once per block
    __shared__ int32_t yd[SETUP_YMAX];
    __shared__ u32x4 stage[SETUP_BLOCK / WAVE][STAGE_UNITS];
    const int tid = threadIdx.x;
    const int lane = tid % WAVE;
    const int i = tid + blockIdx.x * SETUP_BLOCK;
    const int nopt = o.count ? min(*o.count, o.n) : o.n;
    const bool mine = i < nopt;
    if (i == 0)
        reinterpret_cast<unsigned long long *>(tally)[T_T0 / 2] = __builtin_amdgcn_s_memrealtime();
    // the option's fields are requested before the curve is staged, so the
    // two loads overlap
    double strike = 0.0, mat = 0.0, len = 0.0, a = 0.0, sigma = 0.0;
    uint16_t tu = 0, ts = 0;
    int8_t type = 0;
    if (mine && !o.consts) {
        strike = o.strike[i];
        mat = o.maturity[i];
        len = o.length[i];
        tu = o.term_unit[i];
        ts = o.term_steps[i];
        a = o.a[i];
        sigma = o.sigma[i];
        type = o.type[i];
    }
    for (int k = tid; k < NBUCKET; k += SETUP_BLOCK) lh[k] = 0;
    if (tid < 3) lsum[tid] = 0ull;
    if (tid < 2) lbad[tid] = 0;
    // Curves that fit are read from LDS, by binary search when their days
    // never decrease (every real curve); anything else keeps the reference's
    // linear scan over global memory.  Both give the same value.
    const bool staged = y.n <= SETUP_YMAX;
    int unsorted = 0;
    if (staged) {
        for (int k = tid; k < y.n; k += SETUP_BLOCK) {
            yr[k] = y.rate[k];
            yd[k] = y.day[k];
            if (k + 1 < y.n && y.day[k] > y.day[k + 1]) unsorted = 1;
        }
    }
    // the barrier runs unconditionally: it also orders the LDS clearing above
    // before the atomics below, for curves too long to stage as well
    const int any_unsorted = __syncthreads_or(unsorted);
    const bool sorted = staged && !any_unsorted;
    auto yield = [&](double t, int tu) {
        return sorted ? yield_at_sorted(t, tu, yr, yd, y.n) : yield_at(t, tu, y.rate, y.day, y.n);
    };
    OptConst c{};
    double tab[EDGE_TAB];  // the edge table; zeros for options that have none
#pragma unroll
    for (int k = 0; k < EDGE_TAB; ++k) tab[k] = 0.0;
    unsigned long long work = 0, seq_work = 0, wide_work = 0;
    bool invalid = false, unsupported = false;
    int b = -1, r = 0;  // bucket, rank among the block's options in it
    if (mine) {
        int32_t st;
        if (o.consts) {
            const hwt_option_constants_f64 k = o.consts[i];
            tu = k.term_unit;
            st = constants_given(k.t, k.dt, k.dr, k.X, k.M, k.jmax, k.n, k.width, k.term_unit, k.type, maxWidth,
                                 o.consts_f32 != 0, c);
        } else {
            st = option_constants(strike, mat, len, tu, ts, a, sigma, type, maxWidth, c);
        }
        c.status = st;
        if (st == ST_INVALID) {
            out[i] = NAN;
            invalid = true;
        } else if (st == ST_UNSUPPORTED) {
            out[i] = NAN;
            unsupported = true;
        } else if (c.n == 0) {
            out[i] = 100.0;  // no steps: the face value (seq/Seq.hpp:119,171)
        } else {
            const double tn = c.n * c.dt;
            c.Pn = exp(-yield(tn, tu) * tn);
            if (c.e < 0) {
                out[i] = 100.0 * c.Pn;  // exercise outside the tree: the zero-coupon bond
            } else {
                const double te = c.e * c.dt;
                c.Pe = exp(-yield(te, tu) * te);
                TreeShape t{c.w, c.jmax, c.M, c.dr, c.dt, 0.0, 0.0};
                // trees of 7 or more nodes (all but the tiniest) take their
                // edge discounts from four exponentials and the table from
                // its closed form
                if (t.w >= 7) {
                    const EdgeDiscounts ed = edge_discounts(t);
                    edge_factors(t, ed);
                    edge_table_closed(t, ed, tab);
                    c.unit = unit_form_ok(t, tab, unit_dmin, ed.E0, ed.Ea) ? 1 : 0;
                } else {
                    edge_factors(t);
                    double gtab[EDGE_TAB];
                    edge_table_general(t, gtab);
#pragma unroll
                    for (int k = 0; k < EDGE_TAB; ++k) tab[k] = gtab[k];
                    c.unit = unit_form_ok(t, tab, unit_dmin) ? 1 : 0;
                }
                c.rho = t.rho;
                c.rhop = t.rhop;
                c.work = dual_node_updates(c.n, c.e, c.jmax);
                work = (unsigned long long)c.work;
                seq_work = (unsigned long long)seq_node_updates(c.n, c.jmax);
                int64_t key;
                switch (sort) {
                case 'W': case 'w': key = c.w; break;
                case 'H': case 'h': key = c.n; break;
                case '-': key = 0; break;
                default: key = c.work; break;
                }
                const bool wide = (c.w + WAVE - 1) / WAVE > KSMALL;
                if (wide) b = B_WIDE + log_bucket(key, 8, NBUCKET - B_WIDE);
                else if ((classes & CLASS_FLAT) && c.w <= FLAT_MAX_WIDTH) b = flat_bucket(c.w, c.n);
                else if (c.w > ROW_MAX_WIDTH || !(classes & CLASS_ROWS)) b = B_MID + log_bucket(key, 4, B_WIDE - B_MID);
                else b = row_bucket(c.w, c.n);
                r = atomicAdd(&lh[b], 1);
                if (wide) wide_work = work;
            }
        }
    }
    // the block's tallies: one LDS atomic per wave and counter (every lane
    // adding to one LDS address took 4.5 us of the kernel)
    work = wave_total_count(work);
    seq_work = wave_total_count(seq_work);
    wide_work = wave_total_count(wide_work);
    const int n_invalid = __popcll(__ballot(invalid));
    const int n_unsup = __popcll(__ballot(unsupported));
    if (lane == 0) {
        if (work) atomicAdd(&lsum[0], work);
        if (seq_work) atomicAdd(&lsum[1], seq_work);
        if (wide_work) atomicAdd(&lsum[2], wide_work);
        if (n_invalid) atomicAdd(&lbad[0], n_invalid);
        if (n_unsup) atomicAdd(&lbad[1], n_unsup);
    }
    // The block's slots in each bucket: the global histogram's old value is
    // the block's first rank there.  Requested after the first round of the
    // copy-out below and read after the last, so the atomics' round trip
    // overlaps the rest of it; requested before the first round, they held
    // its stores back (k_setup 26.5 -> 24.8 us on the 60k skewed batch).
    int first_rank[SCAN_PER];
    // the wave's records out through its LDS slice: edge tables in two rounds
    // of 32 lanes, then the constants
    const int first = i - lane;  // the wave's first option
    const int count = max(0, min(WAVE, nopt - first));
    u32x4 *slice = stage[tid / WAVE];
    for (int half = 0; half < 2; ++half) {
        if (half == 1) {
            __syncthreads();  // every lane's LDS histogram count is in
#pragma unroll
            for (int q = 0; q < SCAN_PER; ++q) {
                const int k = tid + q * SETUP_BLOCK;
                first_rank[q] = lh[k] ? atomicAdd(&hist[k], lh[k]) : 0;
            }
        }
        if (lane / STAGE_LANES == half) {
            u32x4 u[REC_TAB];
            __builtin_memcpy(u, tab, sizeof(u));
            u32x4 *r = slice + (lane % STAGE_LANES) * REC_TAB;
#pragma unroll
            for (int k = 0; k < REC_TAB; ++k) r[k] = u[k];
        }
        wave_lds_order();
        const int recs = max(0, min(STAGE_LANES, count - half * STAGE_LANES));
        copy_out<STAGE_UNITS>(reinterpret_cast<u32x4 *>(edge) + (size_t)(first + half * STAGE_LANES) * REC_TAB,
                              slice, recs * REC_TAB, lane);
        wave_lds_order();
    }
    {
        u32x4 u[REC_CST];
        __builtin_memcpy(u, &c, sizeof(u));
        u32x4 *r = slice + lane * REC_CST;
#pragma unroll
        for (int k = 0; k < REC_CST; ++k) r[k] = u[k];
        wave_lds_order();
        copy_out<WAVE * REC_CST>(reinterpret_cast<u32x4 *>(cst) + (size_t)first * REC_CST, slice, count * REC_CST,
                                 lane);
    }
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) lh[tid + q * SETUP_BLOCK] = first_rank[q];
    __syncthreads();
    if (mine) slot[i] = make_int2(b, b >= 0 ? lh[b] + r : 0);
    if (tid == 0) {
        if (lsum[0]) atomicAdd((unsigned long long *)&tally[A_WORK], lsum[0]);
        if (lsum[1]) atomicAdd((unsigned long long *)&tally[A_SEQWORK], lsum[1]);
        if (lsum[2]) atomicAdd((unsigned long long *)&tally[A_WORKW], lsum[2]);
        if (lbad[0]) atomicAdd(&tally[A_INVALID], lbad[0]);
        if (lbad[1]) atomicAdd(&tally[A_UNSUP], lbad[1]);
    }
}

// After k_setup: the bucket offsets (every block scans the 8 KB histogram
// itself, scan_offsets; block 0 also writes the pads and the tally), then
// each option's index into the task list at its bucket's offset plus the
// rank k_setup gave it there.  Order inside a bucket is the order in which
// k_setup's blocks reached the histogram (options are independent).  One
// kernel instead of a scan kernel and a scatter kernel saves a kernel
// boundary per execution.
__global__ __launch_bounds__(SETUP_BLOCK) void k_order(int32_t n, const int32_t *__restrict__ count,
                                                       const int2 *__restrict__ slot,
                                                       const int32_t *__restrict__ hist, int32_t *__restrict__ order,
                                                       int32_t *__restrict__ tally, int32_t asc, int32_t classes)
{
    __shared__ int32_t soffs[NBUCKET];
    const int tid = threadIdx.x;
    const int i = tid + blockIdx.x * SETUP_BLOCK;
    if (count) n = min(*count, n);
    const int2 s = (i < n) ? slot[i] : make_int2(-1, 0);  // requested first, so it arrives during the scan
    scan_offsets(hist, soffs, order, tally, asc, classes, blockIdx.x == 0);
    if (s.x >= 0) order[soffs[s.x] + s.y] = i;
}

// What the wide-tree kernel's stencil steps executed in the last execution,
// padding slots included, from the window schedule it walks
// (forward_schedule / backward_schedule): operations (per lane and step 4K+1
// in the unit-diagonal form, 5K in the plain one) and node slots (64 K per
// step) per form.  One thread per wide task; run at sync time, never in the
// timed pipeline.
__global__ __launch_bounds__(256) void k_count_steps(const OptConst *__restrict__ cst,
                                                     const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                                                     int32_t mixed)
{
    __shared__ unsigned long long part[3];
    if (threadIdx.x < 3) part[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tally[T_NLARGE]) {
        const OptConst c = cst[order[i]];
        const TreeShape t{c.w, c.jmax, c.M, c.dr, c.dt, c.rho, c.rhop};
        unsigned long long ops = 0, slots[2] = {0, 0};
        auto count = [&](const Window &win, int nsteps, int, int, bool) {
            const bool u = mixed ? unit_phase<float>(win.K, c.unit != 0) : unit_phase<double>(win.K, c.unit != 0);
            const unsigned long long nn = (unsigned long long)nsteps * WAVE;
            ops += nn * (u ? 4 * win.K + 1 : 5 * win.K);
            slots[u ? 0 : 1] += nn * win.K;
        };
        int lo, hi;
        forward_schedule<WIDE_WINDOW_STEP>(t, c.e, lo, hi, count);
        backward_schedule<WIDE_WINDOW_STEP>(t, c.n, c.e, lo, hi, count);
        atomicAdd(&part[0], ops);
        atomicAdd(&part[1], slots[0]);
        atomicAdd(&part[2], slots[1]);
    }
    __syncthreads();
    if (threadIdx.x < 3 && part[threadIdx.x])
        atomicAdd(reinterpret_cast<unsigned long long *>(tally + T_XFLOP) + threadIdx.x, part[threadIdx.x]);
}

// Form 1: one lane per option over every scheduled task (tasks ordered by
// decreasing work, so the lanes of a wave carry similar trees).
__global__ void k_price_lane(const OptConst *__restrict__ cst, const double *__restrict__ edge,
                             const int32_t *__restrict__ order, int32_t *__restrict__ tally,
                             double *__restrict__ out, double *__restrict__ cols, int32_t colw)
{
    clear_bucket_counts(tally);
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int o = gid < tally[T_NTASK] ? order[gid] : -1;  // -1: past the end or an idle pad slot
    if (o >= 0) {
        const OptConst c = cst[o];
        double *base = cols + (size_t)(gid / WAVE) * (2 * (size_t)colw * WAVE) + (gid % WAVE);
        out[o] = lane_price(c, edge + (size_t)o * EDGE_TAB, LaneColumn{base},
                            LaneColumn{base + (size_t)colw * WAVE});
    }
    stamp_wave_end(tally);
}

// ------------------------------------------------------------- alphas
// Per-step calibrated alphas of one option (the reference's forward pass,
// seq/Seq.hpp:56-106): alpha_i dt = A_i - A_{i-1}, A_i = ln(sum_k E_k y_i[k]) + R(t_{i+1}) t_{i+1}.
template <int K>
__device__ __noinline__ void alphas_full(const OptConst c, const DevYield y, int tu, int lane, double *dst,
                                         const lds_f64 *etab)
{
    const TreeShape t = uni(TreeShape{c.w, c.jmax, c.M, c.dr, c.dt, c.rho, c.rhop});
    double z[K], cl[K], cm[K], cr[K], wE[K];
    window_coefs<double, K, true>(t, 0, lane, etab, cl, cm, cr);
    init_window<double, K, true>(z, t, 0, lane);
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int k = lane * K + s;
        double v = (k < c.w) ? node_disc(k, c.jmax, c.dr, c.dt) : 0.0;
        if (k == 0) v -= c.rho * node_disc(1, c.jmax, c.dr, c.dt);
        if (k == c.w - 1) v -= c.rhop * node_disc(c.w - 2, c.jmax, c.dr, c.dt);
        wE[s] = v;
    }
    double Aprev = 0.0;
    for (int i = 0; i <= c.n; ++i) {
        if (i > 0) stencil_steps<double, K>(z, cl, cm, cr, 1);
        double part = 0.0;
#pragma unroll
        for (int s = 0; s < K; ++s) part += wE[s] * z[s];
        const double sigma = wave_sum(part);
        const double ti = (i + 1) * c.dt;
        const double A = log(sigma) + yield_at(ti, tu, y.rate, y.day, y.n) * ti;
        if (lane == 0) dst[i] = (A - Aprev) / c.dt;
        Aprev = A;
    }
}

__global__ __launch_bounds__(WAVE) void k_alphas(DevOpts o, DevYield y, const int64_t *__restrict__ offsets,
                                                 int32_t *__restrict__ tally, double *__restrict__ alphas)
{
    const int lane = threadIdx.x;
    const int i = blockIdx.x;
    OptConst c;
    const int st = option_constants(o.strike[i], o.maturity[i], o.length[i], o.term_unit[i], o.term_steps[i],
                                    o.a[i], o.sigma[i], o.type[i], WAVE_MAX_WIDTH, c);
    if (st != ST_OK) {
        if (lane == 0) atomicAdd(&tally[st == ST_INVALID ? T_INVALID : T_UNSUP], 1);
        return;
    }
    TreeShape t{c.w, c.jmax, c.M, c.dr, c.dt, 0.0, 0.0};
    edge_factors(t);
    c.rho = t.rho;
    c.rhop = t.rhop;
    __shared__ double etab[EDGE_TAB];
    if (lane == 0) {
        double tab[EDGE_TAB];
        edge_table(t, tab);
        for (int k = 0; k < EDGE_TAB; ++k) etab[k] = tab[k];
    }
    __syncthreads();
    double *dst = alphas + offsets[i];
    const int K = (c.w + WAVE - 1) / WAVE;
    switch (K) {
#define HWT_C(KK) case KK: alphas_full<KK>(c, y, o.term_unit[i], lane, dst, (const lds_f64 *)etab); break;
        HWT_KCASES(HWT_C)
#undef HWT_C
    default: break;
    }
}

// ------------------------------------------------------------- host side
thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HWT_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t err_ = (call);                                                              \
        if (err_ != hipSuccess)                                                                \
            return fail(HWT_EHIP, std::string(#call) + ": " + hipGetErrorString(err_));        \
    } while (0)

bool valid_opts_ptrs(const hwt_options_f64 *o)
{
    return o && o->n >= 0 &&
           (o->n == 0 || (o->strike && o->maturity && o->length && o->term_unit && o->term_steps && o->a &&
                          o->sigma && o->type));
}

bool valid_yield_ptrs(const hwt_yield_f64 *y) { return y && y->n > 0 && y->rate && y->day; }

DevOpts dev_opts(const hwt_options_f64 &o)
{
    return DevOpts{o.n, o.strike, o.maturity, o.length, o.term_unit, o.term_steps, o.a, o.sigma, o.type, nullptr};
}

} // namespace

struct hwt_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t capacity = 0, ycap = 0;
    hwt_config cfg{};
    int grid_wide = 0, grid_narrow = 0, grid_fused = 0;
    hipStream_t side = nullptr;       // narrow-tree kernel runs here, concurrently
    hipEvent_t fork = nullptr, join = nullptr;
    double *scratch_wide = nullptr;
    int32_t *redo = nullptr;  // mixed mode: options re-priced in FP64
    int grid_redo = 0;
    double *cols = nullptr;  // form 1 tree columns
    int lane_block = 256;
    int lane_width = LANE_MAX_WIDTH;  // form 1: column length, the widest tree priced
    // device workspace, one allocation
    void *ws = nullptr;
    size_t ws_bytes = 0;
    OptConst *cst = nullptr;
    double *edge = nullptr;
    int2 *slot = nullptr;  // per option: bucket, rank inside it
    int32_t *order = nullptr, *hist = nullptr, *tally = nullptr;
    int32_t *h_tally = nullptr;  // pinned
    hipEvent_t evk0 = nullptr, evk1 = nullptr;
    double wall_khz = 0.0;  // rate of the device wall clock the kernels stamp
    hipEvent_t evw0 = nullptr, evw1 = nullptr, evn0 = nullptr, evn1 = nullptr;
    // captured graph and the arguments it was captured with
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    DevOpts g_opts{};
    DevYield g_yield{};
    double *g_out = nullptr;
    hipStream_t g_stream = nullptr;
    hipStream_t last_stream = nullptr;
    bool timed_kernel = false;  // evk0/evk1 were recorded by the last execution
    bool timed_split = false;   // so were evw0/evw1 and evn0/evn1
    int32_t last_n = 0;         // options bound of the last execution
    bool steps_counted = true;  // k_count_steps has run for the last execution
    uint64_t serial = 0;        // unique per plan in the process (graphs captured over it key on it)
};

static std::atomic<uint64_t> g_plan_serial{0};

static int launch_pipeline(hwt_plan *p, const DevOpts &o, const DevYield &y, double *out, hipStream_t s, bool time_kernel)
{
    const int n = o.n;
    // Report only this launch's errors: an error left by an unrelated earlier
    // runtime call would otherwise surface in the check below.
    (void)hipGetLastError();
    p->timed_kernel = time_kernel && n > 0;
    p->timed_split = p->timed_kernel && p->cfg.form != HWT_FORM_LANE;
    const int sort = p->cfg.sort ? p->cfg.sort : 'L';
    const int asc = (sort == 'w' || sort == 'h') ? 1 : 0;
    // an empty batch launches no k_setup and no k_order: the tallies it
    // reports (all zero) come from a clearing kernel instead
    if (n == 0) k_reset<<<1, 1024, 0, s>>>(p->tally);
    if (n > 0) {
        const int blocks = (n + SETUP_BLOCK - 1) / SETUP_BLOCK;
        const double unit_dmin = p->cfg.precision == HWT_PREC_MIXED ? UnitRange<float>::dmin : UnitRange<double>::dmin;
        const int classes = p->cfg.form == HWT_FORM_LANE
                                ? 0
                                : ((p->cfg.flags & HWT_FLAG_NO_ROWS) ? 0 : CLASS_ROWS) |
                                      ((p->cfg.flags & HWT_FLAG_NO_FLAT) ? 0 : CLASS_FLAT);
        // the lane form prices tasks one per lane: no row or flat groups, no pads
        // form 1 prices trees up to its column length; wider ones are unsupported
        const int max_width = p->cfg.form == HWT_FORM_LANE ? p->lane_width : WAVE_MAX_WIDTH;
        k_setup<<<blocks, SETUP_BLOCK, 0, s>>>(o, y, sort, max_width, unit_dmin, classes, p->cst, p->edge,
                                               p->slot, p->hist, p->tally, out);
        k_order<<<blocks, SETUP_BLOCK, 0, s>>>(n, o.count, p->slot, p->hist, p->order, p->tally, asc, classes);
        if (time_kernel) HWT_HIP(hipEventRecord(p->evk0, s));
        if (p->cfg.form == HWT_FORM_LANE) {
            const int lb = p->lane_block;
            k_price_lane<<<(n + lb - 1) / lb, lb, 0, s>>>(p->cst, p->edge, p->order, p->tally, out, p->cols,
                                                          p->lane_width);
            if (time_kernel) HWT_HIP(hipEventRecord(p->evk1, s));
            HWT_HIP(hipGetLastError());
            return HWT_OK;
        }
        const bool mixed = p->cfg.precision == HWT_PREC_MIXED;
        if (p->cfg.flags & HWT_FLAG_FUSED) {
            if (time_kernel) HWT_HIP(hipEventRecord(p->evw0, s));
            if (mixed)
                k_price_fused<float><<<p->grid_fused, WAVE, 0, s>>>(p->cst, p->edge, p->order, p->tally, out,
                                                                    p->scratch_wide, p->redo);
            else
                k_price_fused<double><<<p->grid_fused, WAVE, 0, s>>>(p->cst, p->edge, p->order, p->tally, out,
                                                                     p->scratch_wide, p->redo);
            if (time_kernel) {
                HWT_HIP(hipEventRecord(p->evw1, s));
                HWT_HIP(hipEventRecord(p->evn0, s));  // no separate narrow kernel: narrow_ms ~ 0
                HWT_HIP(hipEventRecord(p->evn1, s));
            }
        } else {
            // the narrow-tree kernel after the wide-tree one on `s`, or beside it
            // on the side stream (HWT_FLAG_CONCURRENT)
            const bool serial = (p->cfg.flags & HWT_FLAG_CONCURRENT) == 0;
            hipStream_t ns = serial ? s : p->side;
            if (!serial) {
                HWT_HIP(hipEventRecord(p->fork, s));
                HWT_HIP(hipStreamWaitEvent(p->side, p->fork, 0));
            }
            if (time_kernel) HWT_HIP(hipEventRecord(p->evw0, s));
            if (mixed)
                k_price_wide<float><<<p->grid_wide, WAVE, 0, s>>>(p->cst, p->edge, p->order, p->tally, out,
                                                                  p->scratch_wide, p->redo);
            else
                k_price_wide<double><<<p->grid_wide, WAVE, 0, s>>>(p->cst, p->edge, p->order, p->tally, out,
                                                                   p->scratch_wide, p->redo);
            if (time_kernel) HWT_HIP(hipEventRecord(p->evw1, s));
            if (time_kernel) HWT_HIP(hipEventRecord(p->evn0, ns));
            if (mixed)
                k_price_narrow<float><<<p->grid_narrow, WAVE, 0, ns>>>(p->cst, p->edge, p->order, p->tally, out,
                                                                       p->redo);
            else
                k_price_narrow<double><<<p->grid_narrow, WAVE, 0, ns>>>(p->cst, p->edge, p->order, p->tally, out,
                                                                        p->redo);
            if (time_kernel) HWT_HIP(hipEventRecord(p->evn1, ns));
            if (!serial) {
                HWT_HIP(hipEventRecord(p->join, p->side));
                HWT_HIP(hipStreamWaitEvent(s, p->join, 0));
            }
        }
        // mixed mode: FP64 re-pricing of the options whose FP32 state left its
        // range, after every FP32 kernel has handed its options over
        if (mixed)
            k_price_redo<<<p->grid_redo, WAVE, 0, s>>>(p->cst, p->edge, p->redo, p->tally, out, p->scratch_wide);
        if (time_kernel) HWT_HIP(hipEventRecord(p->evk1, s));
    }
    HWT_HIP(hipGetLastError());
    return HWT_OK;
}

extern "C" {

void hwt_config_init(hwt_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->form = HWT_FORM_AUTO;
    cfg->precision = HWT_PREC_FP64;
    cfg->sort = HWT_SORT_WORK_DESC;
    cfg->device = -1;
}

int32_t hwt_abi_version(void) { return HWT_ABI_VERSION; }

int32_t hwt_max_width(int32_t form)
{
    (void)form;
    return WAVE_MAX_WIDTH;
}

const char *hwt_strerror(int code)
{
    switch (code) {
    case HWT_OK: return "success";
    case HWT_EINVAL: return "invalid argument";
    case HWT_EUNSUPPORTED: return "unsupported option for the selected kernel form";
    case HWT_ENOMEM: return "out of memory or capacity exceeded";
    case HWT_EHIP: return "HIP runtime error";
    default: return "unknown error";
    }
}

const char *hwt_last_error(void) { return g_last_error.c_str(); }

int hwt_plan_create(hwt_plan **out, int32_t capacity, int32_t yield_capacity, const hwt_config *cfg)
{
    g_last_error.clear();
    if (!out || capacity < 0 || yield_capacity < 1) return fail(HWT_EINVAL, "hwt_plan_create: bad arguments");
    *out = nullptr;
    hwt_plan *p = new (std::nothrow) hwt_plan();
    if (!p) return fail(HWT_ENOMEM, "hwt_plan_create: host allocation");
    p->serial = ++g_plan_serial;
    if (cfg) p->cfg = *cfg;
    else hwt_config_init(&p->cfg);
    if (p->cfg.form < HWT_FORM_AUTO || p->cfg.form > HWT_FORM_WAVE ||
        (p->cfg.precision != HWT_PREC_FP64 && p->cfg.precision != HWT_PREC_MIXED) || p->cfg.wide_per_cu < 0 ||
        p->cfg.narrow_per_cu < 0) {
        delete p;
        return fail(HWT_EINVAL, "hwt_plan_create: unknown form or precision");
    }
    if (p->cfg.form == HWT_FORM_LANE && p->cfg.precision == HWT_PREC_MIXED) {
        delete p;
        return fail(HWT_EUNSUPPORTED, "hwt_plan_create: the lane-per-option form runs in FP64 only");
    }
    p->capacity = capacity;
    p->ycap = yield_capacity;
    int rc = HWT_OK;
    auto bail = [&](int code, const std::string &msg) {
        hwt_plan_destroy(p);
        return fail(code, msg);
    };
    if (hipGetDevice(&p->device) != hipSuccess) return bail(HWT_EHIP, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess)
        return bail(HWT_EHIP, "hipStreamCreate failed");
    if (hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->join, hipEventDisableTiming) != hipSuccess)
        return bail(HWT_EHIP, "hipEventCreate failed");
    // Persistent grids: as many single-wave workgroups as can be resident.
    int cus = 0, per_cu_w = 0, per_cu_n = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess || cus < 1)
        cus = 256;
    const bool mixed = p->cfg.precision == HWT_PREC_MIXED;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu_w, mixed ? (const void *)k_price_wide<float> : (const void *)k_price_wide<double>, WAVE, 0) !=
            hipSuccess ||
        per_cu_w < 1)
        per_cu_w = 8;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu_n, mixed ? (const void *)k_price_narrow<float> : (const void *)k_price_narrow<double>, WAVE, 0) !=
            hipSuccess ||
        per_cu_n < 1)
        per_cu_n = 8;
    if (p->cfg.wide_per_cu > 0) per_cu_w = std::min(per_cu_w, p->cfg.wide_per_cu);
    if (p->cfg.narrow_per_cu > 0) per_cu_n = std::min(per_cu_n, p->cfg.narrow_per_cu);
    p->grid_wide = per_cu_w * cus;
    p->grid_narrow = per_cu_n * cus;
    // the fused kernel shares the wide kernel's scratch slots, so its grid
    // never exceeds the wide kernel's
    int per_cu_f = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu_f, mixed ? (const void *)k_price_fused<float> : (const void *)k_price_fused<double>, WAVE, 0) !=
            hipSuccess ||
        per_cu_f < 1)
        per_cu_f = per_cu_w;
    p->grid_fused = std::min(per_cu_f, per_cu_w) * cus;
    // mixed mode's FP64 re-pricing kernel shares the wide kernel's scratch
    // slots as well
    int per_cu_x = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_x, (const void *)k_price_redo, WAVE, 0) != hipSuccess ||
        per_cu_x < 1)
        per_cu_x = 1;
    p->grid_redo = std::min(per_cu_x, per_cu_w) * cus;
    if (const char *v = getenv("HWT_VERBOSE"); v && *v && *v != '0')
        fprintf(stderr, "hwtree plan: device %d, %d CUs, wide kernel %d blocks (%d/CU), narrow kernel %d blocks (%d/CU)\n",
                p->device, cus, p->grid_wide, per_cu_w, p->grid_narrow, per_cu_n);
    const size_t cap = (size_t)std::max(capacity, 1);
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_cst = carve(cap * sizeof(OptConst));
    const size_t o_edge = carve(cap * EDGE_TAB * sizeof(double));
    const size_t o_slot = carve(cap * sizeof(int2));
    const size_t o_order = carve((cap + MAX_PADS) * sizeof(int32_t));  // + row and flat pads
    const size_t o_tally = carve(T_ALLOC * sizeof(int32_t));  // with the histogram
    const bool lane = p->cfg.form == HWT_FORM_LANE;
    // the wide-tree kernels' parked forward states (form 2 only)
    const size_t o_scw = carve(lane ? 0 : (size_t)p->grid_wide * WAVE * MAXK * sizeof(double));
    const size_t o_redo = carve(p->cfg.precision == HWT_PREC_MIXED ? cap * sizeof(int32_t) : 0);
    if (lane) {
        const int bs = p->cfg.block_size;
        p->lane_block = (bs >= WAVE && bs <= 1024 && bs % WAVE == 0) ? bs : 256;
        const int lw = p->cfg.lane_max_width;
        p->lane_width = (lw > 0 && lw < LANE_MAX_WIDTH) ? lw : LANE_MAX_WIDTH;
    }
    const size_t lane_waves = lane ? (cap + WAVE - 1) / WAVE : 0;
    const size_t o_cols = carve(lane_waves * 2 * (size_t)p->lane_width * WAVE * sizeof(double));
    p->ws_bytes = off;
    if (hipMalloc(&p->ws, p->ws_bytes) != hipSuccess) return bail(HWT_ENOMEM, "hipMalloc workspace failed");
    char *base = (char *)p->ws;
    p->cst = (OptConst *)(base + o_cst);
    p->edge = (double *)(base + o_edge);
    p->slot = (int2 *)(base + o_slot);
    p->order = (int32_t *)(base + o_order);
    p->tally = (int32_t *)(base + o_tally);
    p->hist = p->tally + T_HIST;
    p->scratch_wide = (double *)(base + o_scw);
    p->redo = p->cfg.precision == HWT_PREC_MIXED ? (int32_t *)(base + o_redo) : nullptr;
    p->cols = lane ? (double *)(base + o_cols) : nullptr;
    // counters start clear; from then on every execution leaves them clear
    if (hipMemsetAsync(p->tally, 0, T_ALLOC * sizeof(int32_t), p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
        return bail(HWT_EHIP, "clearing the plan's counters failed");
    if (hipHostMalloc((void **)&p->h_tally, T_HOST * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return bail(HWT_ENOMEM, "hipHostMalloc failed");
    std::memset(p->h_tally, 0, T_HOST * sizeof(int32_t));  // a plan never executed syncs clean
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, p->device) != hipSuccess || khz <= 0)
        return bail(HWT_EHIP, "the device wall clock rate is unavailable");
    p->wall_khz = khz;
    if (hipEventCreate(&p->evk0) != hipSuccess || hipEventCreate(&p->evk1) != hipSuccess ||
        hipEventCreate(&p->evw0) != hipSuccess || hipEventCreate(&p->evw1) != hipSuccess ||
        hipEventCreate(&p->evn0) != hipSuccess || hipEventCreate(&p->evn1) != hipSuccess)
        return bail(HWT_EHIP, "hipEventCreate failed");
    *out = p;
    (void)rc;
    return HWT_OK;
}

void hwt_plan_destroy(hwt_plan *p)
{
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->evk0) (void)hipEventDestroy(p->evk0);
    if (p->evk1) (void)hipEventDestroy(p->evk1);
    for (hipEvent_t e : {p->evw0, p->evw1, p->evn0, p->evn1})
        if (e) (void)hipEventDestroy(e);
    if (p->fork) (void)hipEventDestroy(p->fork);
    if (p->join) (void)hipEventDestroy(p->join);
    if (p->side) (void)hipStreamDestroy(p->side);
    if (p->h_tally) (void)hipHostFree(p->h_tally);
    if (p->ws) (void)hipFree(p->ws);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

} // extern "C"

// Enqueues one execution over device views (graph replay unless
// HWT_FLAG_NO_GRAPH; the graph is captured again when the views or the
// stream change).
static int plan_execute_views(hwt_plan *p, const DevOpts &o, const DevYield &y, double *dout, hipStream_t s)
{
    p->last_stream = s;
    p->last_n = o.n;
    p->steps_counted = false;
    // no event records around the execution: the kernels stamp its start and
    // end (T_T0, T_T1), which costs nothing between back-to-back executions;
    // two event records per execution measured 9 us per graph replay
    if (p->cfg.flags & (HWT_FLAG_NO_GRAPH | HWT_FLAG_DIRECT)) {
        const int rc = launch_pipeline(p, o, y, dout, s, (p->cfg.flags & HWT_FLAG_NO_GRAPH) != 0);
        if (rc) return rc;
    } else {
        const bool same = p->exec && p->g_stream == s && p->g_out == dout &&
                          std::memcmp(&p->g_opts, &o, sizeof(o)) == 0 && std::memcmp(&p->g_yield, &y, sizeof(y)) == 0;
        if (!same) {
            if (p->exec) (void)hipGraphExecDestroy(p->exec);
            if (p->graph) (void)hipGraphDestroy(p->graph);
            p->exec = nullptr;
            p->graph = nullptr;
            HWT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = launch_pipeline(p, o, y, dout, s, false);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return fail(HWT_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
            p->graph = g;
            HWT_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
            p->g_opts = o;
            p->g_yield = y;
            p->g_out = dout;
            p->g_stream = s;
        }
        HWT_HIP(hipGraphLaunch(p->exec, s));
        p->timed_kernel = p->timed_split = false;
    }
    return HWT_OK;
}

extern "C" {

int hwt_plan_execute(hwt_plan *p, const hwt_options_f64 *dopts, const hwt_yield_f64 *dyield, double *dout, void *stream)
{
    g_last_error.clear();
    if (!p || !valid_opts_ptrs(dopts) || !valid_yield_ptrs(dyield) || (dopts->n > 0 && !dout))
        return fail(HWT_EINVAL, "hwt_plan_execute: null or negative-size argument");
    if (dopts->n > p->capacity) return fail(HWT_ENOMEM, "hwt_plan_execute: more options than the plan capacity");
    if (dyield->n > p->ycap) return fail(HWT_ENOMEM, "hwt_plan_execute: yield curve longer than the plan capacity");
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    return plan_execute_views(p, dev_opts(*dopts), DevYield{dyield->n, dyield->rate, dyield->day}, dout, s);
}

int hwt_internal_fail(int code, const char *msg) { return fail(code, msg ? msg : ""); }

// The pipeline over arrays holding up to dopts->n options, of which the
// first *dev_count (device memory) are priced, enqueued on `stream` for
// capture into a caller's graph (the surrogate's fallback list).
int hwt_internal_plan_enqueue_counted(hwt_plan *p, const hwt_options_f64 *dopts, const int32_t *dev_count,
                                      const hwt_yield_f64 *dyield, double *dout, void *stream)
{
    if (!p || !dev_count || !valid_opts_ptrs(dopts) || !valid_yield_ptrs(dyield) || (dopts->n > 0 && !dout))
        return fail(HWT_EINVAL, "plan enqueue: null or negative-size argument");
    if (dopts->n > p->capacity) return fail(HWT_ENOMEM, "plan enqueue: bound above the plan capacity");
    if (dyield->n > p->ycap) return fail(HWT_ENOMEM, "plan enqueue: yield curve longer than the plan capacity");
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    DevOpts o = dev_opts(*dopts);
    o.count = dev_count;
    p->last_stream = s;
    p->last_n = o.n;
    p->steps_counted = true;  // nothing to count: the caller's graph has no stats
    return launch_pipeline(p, o, DevYield{dyield->n, dyield->rate, dyield->day}, dout, s, false);
}

uint64_t hwt_internal_plan_serial(const hwt_plan *p) { return p ? p->serial : 0; }

void hwt_internal_plan_replayed_on(hwt_plan *p, void *stream)
{
    if (!p) return;
    p->last_stream = stream ? (hipStream_t)stream : p->stream;
    p->steps_counted = true;
}

int hwt_plan_sync(hwt_plan *p, hwt_stats *stats)
{
    if (!p) return fail(HWT_EINVAL, "hwt_plan_sync: null plan");
    hipStream_t s = p->last_stream ? p->last_stream : p->stream;
    // with stats: what the windowed kernel's steps executed in the last
    // execution, counted once from its schedule (form 2 only)
    if (stats && !p->steps_counted && p->last_n > 0 && p->cfg.form != HWT_FORM_LANE) {
        (void)hipGetLastError();
        k_count_steps<<<(p->last_n + 255) / 256, 256, 0, s>>>(p->cst, p->order, p->tally,
                                                             p->cfg.precision == HWT_PREC_MIXED ? 1 : 0);
        HWT_HIP(hipGetLastError());
        p->steps_counted = true;
    }
    HWT_HIP(hipStreamSynchronize(s));
    // the tallies are fetched here, once, rather than by a copy queued behind
    // every execution: back-to-back executions then run without a transfer
    // between them
    HWT_HIP(hipMemcpyAsync(p->h_tally, p->tally, T_HOST * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HWT_HIP(hipStreamSynchronize(s));
    const int32_t *t = p->h_tally;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.f;
        const unsigned long long *t64u = reinterpret_cast<const unsigned long long *>(t);
        const unsigned long long t0 = t64u[T_T0 / 2];
        unsigned long long t1 = 0;
        for (int k = 0; k < N_END; ++k) t1 = std::max(t1, t64u[T_END / 2 + k]);
        stats->total_ms = (t0 && t1 > t0) ? (double)(t1 - t0) / p->wall_khz : 0.0;
        stats->kernel_ms = stats->total_ms;
        if (p->timed_kernel) {
            if (hipEventElapsedTime(&ms, p->evk0, p->evk1) == hipSuccess) stats->kernel_ms = ms;
        }
        if (p->timed_split) {
            if (hipEventElapsedTime(&ms, p->evw0, p->evw1) == hipSuccess) stats->wide_ms = ms;
            if (hipEventElapsedTime(&ms, p->evn0, p->evn1) == hipSuccess) stats->narrow_ms = ms;
        }
        stats->device_memory = (int64_t)p->ws_bytes;
        int64_t w, sw, ww;
        std::memcpy(&w, t + T_WORK, sizeof(w));
        std::memcpy(&sw, t + T_SEQWORK, sizeof(sw));
        std::memcpy(&ww, t + T_WORKW, sizeof(ww));
        stats->node_updates = w;
        stats->wide_node_updates = ww;
        stats->seq_node_updates = sw;
        stats->tasks = t[T_NTASK] - t[T_NPAD];
        stats->wide_tasks = t[T_NLARGE];
        stats->unsupported = t[T_UNSUP];
        stats->fp64_repriced = t[T_NREDO];
        std::memcpy(&stats->wide_step_ops, t + T_XFLOP, sizeof(int64_t));
        std::memcpy(&stats->wide_unit_slots, t + T_XUNIT, sizeof(int64_t));
        std::memcpy(&stats->wide_plain_slots, t + T_XPLAIN, sizeof(int64_t));
    }
    if (t[T_INVALID] > 0) return fail(HWT_EINVAL, std::to_string(t[T_INVALID]) + " option(s) with invalid parameters (price NaN)");
    if (t[T_UNSUP] > 0)
        return fail(HWT_EUNSUPPORTED, std::to_string(t[T_UNSUP]) + " option(s) wider than " +
                                          std::to_string(p->cfg.form == HWT_FORM_LANE ? p->lane_width : WAVE_MAX_WIDTH) +
                                          " nodes (price NaN)");
    return HWT_OK;
}

// Device copies of host SoA arrays, one allocation.
struct DeviceBatch {
    void *mem = nullptr;
    size_t bytes = 0;
    hwt_options_f64 o{};
    hwt_yield_f64 y{};
    double *out = nullptr;
    ~DeviceBatch()
    {
        if (mem) (void)hipFree(mem);
    }
};

static int upload(const hwt_options_f64 *h, const hwt_yield_f64 *hy, DeviceBatch &d, bool need_out)
{
    const size_t n = (size_t)h->n, ny = (size_t)hy->n;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_strike = carve(n * 8), o_mat = carve(n * 8), o_len = carve(n * 8), o_tu = carve(n * 2),
                 o_ts = carve(n * 2), o_a = carve(n * 8), o_sig = carve(n * 8), o_type = carve(n),
                 o_rate = carve(ny * 8), o_day = carve(ny * 4), o_out = carve(need_out ? n * 8 : 0);
    if (hipMalloc(&d.mem, std::max<size_t>(off, 256)) != hipSuccess) return fail(HWT_ENOMEM, "hipMalloc batch failed");
    d.bytes = std::max<size_t>(off, 256);
    char *b = (char *)d.mem;
    auto put = [&](size_t at, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(b + at, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    if (put(o_strike, h->strike, n * 8) || put(o_mat, h->maturity, n * 8) || put(o_len, h->length, n * 8) ||
        put(o_tu, h->term_unit, n * 2) || put(o_ts, h->term_steps, n * 2) || put(o_a, h->a, n * 8) ||
        put(o_sig, h->sigma, n * 8) || put(o_type, h->type, n) || put(o_rate, hy->rate, ny * 8) ||
        put(o_day, hy->day, ny * 4))
        return fail(HWT_EHIP, "hipMemcpy H2D failed");
    d.o = hwt_options_f64{h->n, (double *)(b + o_strike), (double *)(b + o_mat), (double *)(b + o_len),
                          (uint16_t *)(b + o_tu), (uint16_t *)(b + o_ts), (double *)(b + o_a),
                          (double *)(b + o_sig), (int8_t *)(b + o_type)};
    d.y = hwt_yield_f64{hy->n, (double *)(b + o_rate), (int32_t *)(b + o_day)};
    d.out = (double *)(b + o_out);
    return HWT_OK;
}

struct DeviceGuard {
    int prev = -1;
    bool set = false;
    int enter(int dev)
    {
        if (dev < 0) return HWT_OK;
        if (hipGetDevice(&prev) != hipSuccess) return fail(HWT_EHIP, "hipGetDevice failed");
        if (hipSetDevice(dev) != hipSuccess) return fail(HWT_EINVAL, "hipSetDevice: bad device ordinal");
        set = true;
        return HWT_OK;
    }
    ~DeviceGuard()
    {
        if (set) (void)hipSetDevice(prev);
    }
};

} // extern "C"

// Widest tree (2*jmax+1) among the valid options of host arrays, from the
// same constant formulas as the device (0 when none is valid).
static int host_max_width(const hwt_options_f64 *o)
{
    int w = 0;
    for (int i = 0; i < o->n; ++i) {
        OptConst c;
        if (option_constants(o->strike[i], o->maturity[i], o->length[i], o->term_unit[i], o->term_steps[i], o->a[i],
                             o->sigma[i], o->type[i], 1 << 30, c) == ST_OK)
            w = std::max(w, (int)c.w);
    }
    return w;
}

// One execution of a fresh plan over device views, results copied to host.
// One execution of plan p over device views, synchronised, and the prices
// copied back to host_out.
static int price_views(hwt_plan *p, const DevOpts &o, const DevYield &y, double *dev_out, double *host_out,
                       hwt_stats *stats)
{
    const int rc = plan_execute_views(p, o, y, dev_out, p->stream);
    int rs = rc ? rc : hwt_plan_sync(p, stats);
    const std::string keep = g_last_error;
    if (rc == HWT_OK) {
        if (hipMemcpy(host_out, dev_out, (size_t)o.n * 8, hipMemcpyDeviceToHost) != hipSuccess)
            rs = fail(HWT_EHIP, "hipMemcpy D2H failed");
        else
            g_last_error = keep;
    }
    return rs;
}

// A batch resident in device memory: raw options (DeviceBatch), or per-option
// tree constants (the constants entry points and FP32 callers), with the
// curve and an output array.
struct hwt_batch {
    int device = 0;
    int max_width = 0;  // widest valid tree, for form 1's column buffers
    DeviceBatch d;      // raw options
    void *cmem = nullptr;  // constants, curve and output, one allocation
    DevOpts o{};
    DevYield y{};
    double *out = nullptr;
    size_t bytes = 0;
    // the plan of the last hwt_batch_price call, kept while the settings stay
    // the same, so repeated pricing allocates nothing and replays its graph
    hwt_plan *plan = nullptr;
    hwt_config plan_cfg{};
    bool once = false;  // a one-shot call's batch: its single execution runs without a graph
    ~hwt_batch()
    {
        if (plan) hwt_plan_destroy(plan);
        if (cmem) (void)hipFree(cmem);
    }
};

// A batch of constants: k[0..n) and the curve copied to the current device.
static int batch_of_constants(hwt_batch *b, const hwt_option_constants_f64 *k, int32_t n, const hwt_yield_f64 *yield,
                              bool f32)
{
    const size_t ny = (size_t)yield->n;
    auto part = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    const size_t cb = part((size_t)n * sizeof(*k)), rb = part(ny * 8), db = part(ny * 4);
    const size_t bytes = cb + rb + db + part((size_t)n * 8);
    if (hipMalloc(&b->cmem, bytes) != hipSuccess) return fail(HWT_ENOMEM, "hipMalloc constants batch failed");
    b->bytes = bytes;
    char *base = (char *)b->cmem;
    auto *dc = (hwt_option_constants_f64 *)base;
    auto *dr = (double *)(base + cb);
    auto *dd = (int32_t *)(base + cb + rb);
    b->out = (double *)(base + cb + rb + db);
    if ((n > 0 && hipMemcpy(dc, k, (size_t)n * sizeof(*k), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(dr, yield->rate, ny * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dd, yield->day, ny * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail(HWT_EHIP, "hipMemcpy H2D failed");
    b->o = DevOpts{};
    b->o.n = n;
    b->o.consts = dc;
    b->o.consts_f32 = f32 ? 1 : 0;
    b->y = DevYield{yield->n, dr, dd};
    int w = 3;
    for (int i = 0; i < n; ++i) w = std::max(w, std::min(k[i].width, LANE_MAX_WIDTH + 1));
    b->max_width = w;
    return HWT_OK;
}

static int new_batch(hwt_batch **out, hwt_batch *&b)
{
    *out = nullptr;
    b = new (std::nothrow) hwt_batch();
    if (!b) return fail(HWT_ENOMEM, "hwt_batch: host allocation");
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        b = nullptr;
        return fail(HWT_EHIP, "hipGetDevice failed");
    }
    return HWT_OK;
}

// The tree constants of an option as the reference's default build (real =
// float) derives them, common/OptionConstants.hpp:29-53: float arithmetic,
// except that the unqualified exp, sqrt and ceil there are the C library's
// double functions (<cmath> declares only those in the global namespace),
// their results rounded to float where the float fields store them.
// Invalid parameters get width 0, which the constants path rejects.
static hwt_option_constants_f64 constants_f32(float strike, float T, float len, uint16_t tu, uint16_t steps, float a,
                                              float sigma, int8_t type)
{
    hwt_option_constants_f64 k{};
    k.t = len;
    k.X = strike;
    k.term_unit = tu;
    k.type = type;
    if ((type != 0 && type != 1) || tu == 0 || steps == 0 || !(a > 0.0f) || !(sigma >= 0.0f) || !(T >= 0.0f) ||
        !(strike == strike))
        return k;
    const int tuy = (int)std::ceil((double)(365.0f / (float)tu));
    const float nf = (float)((int)steps * tuy) * T;  // int * int, then * float
    if (!(nf < 2147483647.0f)) return k;
    const float dt = (float)tuy / (float)steps;
    const double V = (double)(sigma * sigma) * (1.0 - std::exp((double)(-2.0f * a * dt))) / (double)(2.0f * a);
    const float dr = (float)std::sqrt(3.0 * V);
    const float M = (float)(std::exp((double)(-a * dt)) - 1.0);
    const float q = -0.184f / M;
    if (!(q < 1.0e9f)) return k;
    k.n = (int32_t)nf;
    k.dt = dt;
    k.dr = dr;
    k.M = M;
    k.jmax = (int32_t)q + 1;
    k.width = 2 * k.jmax + 1;
    return k;
}

static bool valid_opts_f32(const hwt_options_f32 *o)
{
    return o && o->n >= 0 &&
           (o->n == 0 || (o->strike && o->maturity && o->length && o->term_unit && o->term_steps && o->a &&
                          o->sigma && o->type));
}

static bool valid_yield_f32(const hwt_yield_f32 *y) { return y && y->n > 0 && y->rate && y->day; }

// The FP32 build's constants of every option of host float arrays, and the
// curve's rates widened (float -> double is exact).
static int widen_f32(const hwt_options_f32 *opts, const hwt_yield_f32 *yield, std::vector<hwt_option_constants_f64> &k,
                     std::vector<double> &rates)
{
    try {
        k.resize((size_t)opts->n);
        rates.assign(yield->rate, yield->rate + yield->n);
    } catch (const std::bad_alloc &) {
        return fail(HWT_ENOMEM, "host allocation");
    }
    for (int i = 0; i < opts->n; ++i)
        k[i] = constants_f32(opts->strike[i], opts->maturity[i], opts->length[i], opts->term_unit[i],
                             opts->term_steps[i], opts->a[i], opts->sigma[i], opts->type[i]);
    return HWT_OK;
}

extern "C" {

int hwt_batch_create(hwt_batch **out, const hwt_options_f64 *opts, const hwt_yield_f64 *yield)
{
    g_last_error.clear();
    if (!out || !valid_opts_ptrs(opts) || !valid_yield_ptrs(yield))
        return fail(HWT_EINVAL, "hwt_batch_create: null or negative-size argument");
    hwt_batch *b = nullptr;
    int rc = new_batch(out, b);
    if (rc) return rc;
    rc = upload(opts, yield, b->d, true);
    if (rc) {
        delete b;
        return rc;
    }
    b->bytes = b->d.bytes;
    b->o = dev_opts(b->d.o);
    b->y = DevYield{b->d.y.n, b->d.y.rate, b->d.y.day};
    b->out = b->d.out;
    b->max_width = host_max_width(opts);
    *out = b;
    return HWT_OK;
}

int hwt_batch_create_f32(hwt_batch **out, const hwt_options_f32 *opts, const hwt_yield_f32 *yield)
{
    g_last_error.clear();
    if (!out || !valid_opts_f32(opts) || !valid_yield_f32(yield))
        return fail(HWT_EINVAL, "hwt_batch_create_f32: null or negative-size argument");
    std::vector<hwt_option_constants_f64> k;
    std::vector<double> rates;
    int rc = widen_f32(opts, yield, k, rates);
    if (rc) return rc;
    hwt_batch *b = nullptr;
    rc = new_batch(out, b);
    if (rc) return rc;
    const hwt_yield_f64 y{yield->n, rates.data(), yield->day};
    rc = batch_of_constants(b, k.data(), opts->n, &y, true);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return HWT_OK;
}

int64_t hwt_batch_device_memory(const hwt_batch *b)
{
    return b ? (int64_t)(b->bytes + (b->plan ? b->plan->ws_bytes : 0)) : 0;
}

void hwt_batch_destroy(hwt_batch *b)
{
    if (!b) return;
    int prev = -1;
    const bool swap = hipGetDevice(&prev) == hipSuccess && prev != b->device && hipSetDevice(b->device) == hipSuccess;
    delete b;
    if (swap) (void)hipSetDevice(prev);
}

int hwt_batch_price(hwt_batch *b, double *host_out, const hwt_config *cfg, hwt_stats *stats)
{
    g_last_error.clear();
    if (!b || (b->o.n > 0 && !host_out)) return fail(HWT_EINVAL, "hwt_batch_price: null argument");
    hwt_config c;
    if (cfg) c = *cfg;
    else hwt_config_init(&c);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (b->o.n == 0) return HWT_OK;
    DeviceGuard g;
    int rc = g.enter(b->device);
    if (rc) return rc;
    if (c.lane_max_width <= 0) c.lane_max_width = std::max(b->max_width, 3);
    if (b->once) c.flags |= HWT_FLAG_NO_GRAPH;  // one execution: a graph would not be replayed
    if (!b->plan || std::memcmp(&b->plan_cfg, &c, sizeof(c)) != 0) {
        if (b->plan) hwt_plan_destroy(b->plan);
        b->plan = nullptr;
        rc = hwt_plan_create(&b->plan, b->o.n, b->y.n, &c);
        if (rc) return rc;
        b->plan_cfg = c;
    }
    return price_views(b->plan, b->o, b->y, b->out, host_out, stats);
}

int hwt_price_f64(const hwt_options_f64 *opts, const hwt_yield_f64 *yield, double *out_prices, const hwt_config *cfg,
                  hwt_stats *stats)
{
    g_last_error.clear();
    if (!valid_opts_ptrs(opts) || !valid_yield_ptrs(yield) || (opts->n > 0 && !out_prices))
        return fail(HWT_EINVAL, "hwt_price_f64: null or negative-size argument");
    hwt_config c;
    if (cfg) c = *cfg;
    else hwt_config_init(&c);
    DeviceGuard g;
    int rc = g.enter(c.device);
    if (rc) return rc;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (opts->n == 0) return HWT_OK;
    hwt_batch *b = nullptr;
    rc = hwt_batch_create(&b, opts, yield);
    if (rc) return rc;
    b->once = true;
    rc = hwt_batch_price(b, out_prices, &c, stats);
    hwt_batch_destroy(b);
    return rc;
}

} // extern "C"

// One-shot pricing of a batch of constants (host arrays) into host_out.
static int price_constants(const hwt_option_constants_f64 *consts, int32_t n, const hwt_yield_f64 *yield,
                           double *host_out, const hwt_config *cfg, hwt_stats *stats, bool f32)
{
    hwt_config c;
    if (cfg) c = *cfg;
    else hwt_config_init(&c);
    DeviceGuard g;
    int rc = g.enter(c.device);
    if (rc) return rc;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return HWT_OK;
    hwt_batch *b = nullptr;
    rc = new_batch(&b, b);
    if (rc) return rc;
    b->once = true;
    rc = batch_of_constants(b, consts, n, yield, f32);
    if (rc == HWT_OK) rc = hwt_batch_price(b, host_out, &c, stats);
    hwt_batch_destroy(b);
    return rc;
}

extern "C" {

int hwt_price_f32(const hwt_options_f32 *opts, const hwt_yield_f32 *yield, float *out_prices, const hwt_config *cfg,
                  hwt_stats *stats)
{
    g_last_error.clear();
    if (!valid_opts_f32(opts) || !valid_yield_f32(yield) || (opts->n > 0 && !out_prices))
        return fail(HWT_EINVAL, "hwt_price_f32: null or negative-size argument");
    std::vector<hwt_option_constants_f64> k;
    std::vector<double> rates, prices;
    int rc = widen_f32(opts, yield, k, rates);
    if (rc) return rc;
    try {
        prices.resize((size_t)opts->n);
    } catch (const std::bad_alloc &) {
        return fail(HWT_ENOMEM, "hwt_price_f32: host allocation");
    }
    const hwt_yield_f64 y{yield->n, rates.data(), yield->day};
    rc = price_constants(k.data(), opts->n, &y, prices.data(), cfg, stats, true);
    if (rc == HWT_OK || rc == HWT_EUNSUPPORTED || rc == HWT_EINVAL)  // NaN rows are part of the output
        for (int i = 0; i < opts->n; ++i) out_prices[i] = (float)prices[i];
    return rc;
}

int hwt_price_constants_f64(const hwt_option_constants_f64 *consts, int32_t n, const hwt_yield_f64 *yield,
                            double *out_prices, const hwt_config *cfg, hwt_stats *stats)
{
    g_last_error.clear();
    if (n < 0 || (n > 0 && (!consts || !out_prices)) || !valid_yield_ptrs(yield))
        return fail(HWT_EINVAL, "hwt_price_constants_f64: null or negative-size argument");
    return price_constants(consts, n, yield, out_prices, cfg, stats, false);
}

int hwt_price_constants_f32(const hwt_option_constants_f32 *consts, int32_t n, const hwt_yield_f32 *yield,
                            float *out_prices, const hwt_config *cfg, hwt_stats *stats)
{
    g_last_error.clear();
    if (n < 0 || (n > 0 && (!consts || !out_prices)) || !valid_yield_f32(yield))
        return fail(HWT_EINVAL, "hwt_price_constants_f32: null or negative-size argument");
    std::vector<hwt_option_constants_f64> wide((size_t)n);
    for (int i = 0; i < n; ++i) {  // float -> double is exact
        const hwt_option_constants_f32 &k = consts[i];
        wide[i] = hwt_option_constants_f64{k.t, k.dt, k.dr, k.X, k.M, k.jmax, k.n, k.width, k.term_unit, k.type};
    }
    const std::vector<double> rates(yield->rate, yield->rate + yield->n);
    const hwt_yield_f64 y{yield->n, rates.data(), yield->day};
    std::vector<double> prices((size_t)n);
    const int rc = price_constants(wide.data(), n, &y, prices.data(), cfg, stats, true);
    if (rc == HWT_OK || rc == HWT_EUNSUPPORTED || rc == HWT_EINVAL)
        for (int i = 0; i < n; ++i) out_prices[i] = (float)prices[i];
    return rc;
}

} // extern "C"

// ------------------------------------------------- allocation-free pricing
// hwt_price_device_noalloc keeps every buffer its kernels need beyond the
// caller's inputs and output in this one static device array, which each
// device holds from the module's load until the process ends.  Layout (bytes,
// each part 256-aligned): OptConst[N], edge tables[N][36], (bucket, rank)[N],
// order[N + row pads], counters with the histogram, and the wide-tree kernel's parked forward states (one
// 64 x MAXK double slot per resident wave, at most NOALLOC_WIDE_SLOTS).
constexpr int NOALLOC_MAX_OPTIONS = HWT_NOALLOC_MAX_OPTIONS;
constexpr int NOALLOC_WIDE_SLOTS = 1024;
constexpr size_t noalloc_part(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
constexpr size_t NOALLOC_O_CST = 0;
constexpr size_t NOALLOC_O_EDGE = NOALLOC_O_CST + noalloc_part((size_t)NOALLOC_MAX_OPTIONS * sizeof(OptConst));
constexpr size_t NOALLOC_O_SLOT = NOALLOC_O_EDGE + noalloc_part((size_t)NOALLOC_MAX_OPTIONS * EDGE_TAB * 8);
constexpr size_t NOALLOC_O_ORDER = NOALLOC_O_SLOT + noalloc_part((size_t)NOALLOC_MAX_OPTIONS * sizeof(int2));
constexpr size_t NOALLOC_O_TALLY = NOALLOC_O_ORDER + noalloc_part(((size_t)NOALLOC_MAX_OPTIONS + MAX_PADS) * 4);
constexpr size_t NOALLOC_O_WIDE = NOALLOC_O_TALLY + noalloc_part(T_ALLOC * 4);
constexpr size_t NOALLOC_BYTES = NOALLOC_O_WIDE + (size_t)NOALLOC_WIDE_SLOTS * WAVE * MAXK * 8;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

// The array's address on each device, taken once per device with that device
// current (in host code `scratch` names a host-side stand-in, not device
// memory), and the persistent grids, which need occupancy queries.
struct NoallocDevice {
    unsigned char *ptr = nullptr;
    int grid_wide = 0, grid_narrow = 0;
};
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[NOALLOC_MAX_DEVICES];

int noalloc_device(int dev, NoallocDevice &out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(HWT_EINVAL, "hwt_price_device_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoallocDevice &d = g_noalloc[dev];
    if (!d.ptr) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(HWT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        int cus = 0, per_w = 0, per_n = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_w, (const void *)k_price_wide<double>, WAVE, 0) !=
                hipSuccess ||
            per_w < 1)
            per_w = 2;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_n, (const void *)k_price_narrow<double>, WAVE, 0) !=
                hipSuccess ||
            per_n < 1)
            per_n = 4;
        d.grid_wide = std::min(per_w * cus, NOALLOC_WIDE_SLOTS);
        d.grid_narrow = per_n * cus;
        d.ptr = (unsigned char *)ptr;
    }
    out = d;
    return HWT_OK;
}

} // namespace

extern "C" {

int hwt_price_device_noalloc(const hwt_options_f64 *dopts, const hwt_yield_f64 *dyield, double *dout, void *stream)
{
    g_last_error.clear();
    if (!valid_opts_ptrs(dopts) || !valid_yield_ptrs(dyield) || (dopts->n > 0 && !dout))
        return fail(HWT_EINVAL, "hwt_price_device_noalloc: null or negative-size argument");
    if (dopts->n > NOALLOC_MAX_OPTIONS)
        return fail(HWT_ENOMEM, "hwt_price_device_noalloc: " + std::to_string(dopts->n) +
                                    " options exceed the fixed scratch capacity of " +
                                    std::to_string(NOALLOC_MAX_OPTIONS) + "; use a plan");
    int dev = 0;
    HWT_HIP(hipGetDevice(&dev));
    NoallocDevice d;
    const int rc = noalloc_device(dev, d);
    if (rc) return rc;
    // a plan-shaped view of the scratch: default settings (form 2, FP64,
    // work-descending order, kernels in sequence), nothing allocated
    hwt_plan v;
    hwt_config_init(&v.cfg);
    v.device = dev;
    v.grid_wide = d.grid_wide;
    v.grid_narrow = d.grid_narrow;
    v.cst = (OptConst *)(d.ptr + NOALLOC_O_CST);
    v.edge = (double *)(d.ptr + NOALLOC_O_EDGE);
    v.slot = (int2 *)(d.ptr + NOALLOC_O_SLOT);
    v.order = (int32_t *)(d.ptr + NOALLOC_O_ORDER);
    v.tally = (int32_t *)(d.ptr + NOALLOC_O_TALLY);
    v.hist = v.tally + T_HIST;
    v.scratch_wide = (double *)(d.ptr + NOALLOC_O_WIDE);
    hipStream_t s = (hipStream_t)stream;
    // the array keeps the last call's contents: counters start from zero by a
    // fill kernel (no memset node, so a caller may capture the call in a graph)
    (void)hipGetLastError();
    k_reset<<<1, 1024, 0, s>>>(v.tally);
    HWT_HIP(hipGetLastError());
    return launch_pipeline(&v, dev_opts(*dopts), DevYield{dyield->n, dyield->rate, dyield->day}, dout, s, false);
}

int hwt_alphas_f64(const hwt_options_f64 *opts, const hwt_yield_f64 *yield, int64_t *offsets, double *alphas)
{
    g_last_error.clear();
    if (!valid_opts_ptrs(opts) || !valid_yield_ptrs(yield) || !offsets)
        return fail(HWT_EINVAL, "hwt_alphas_f64: null or negative-size argument");
    // Heights on the host, with the same constant formulas as the device.
    offsets[0] = 0;
    for (int i = 0; i < opts->n; ++i) {
        OptConst c;
        const int st = option_constants(opts->strike[i], opts->maturity[i], opts->length[i], opts->term_unit[i],
                                        opts->term_steps[i], opts->a[i], opts->sigma[i], opts->type[i],
                                        WAVE_MAX_WIDTH, c);
        if (st == ST_INVALID) return fail(HWT_EINVAL, "hwt_alphas_f64: option " + std::to_string(i) + " invalid");
        if (st == ST_UNSUPPORTED)
            return fail(HWT_EUNSUPPORTED, "hwt_alphas_f64: option " + std::to_string(i) + " too wide");
        offsets[i + 1] = offsets[i] + (int64_t)c.n + 1;
    }
    if (!alphas || opts->n == 0) return HWT_OK;
    DeviceBatch d;
    int rc = upload(opts, yield, d, false);
    if (rc) return rc;
    void *mem = nullptr;
    const size_t total = (size_t)offsets[opts->n];
    const size_t bytes = total * 8 + (size_t)(opts->n + 1) * 8 + T_NCOUNT * 4;
    if (hipMalloc(&mem, bytes) != hipSuccess) return fail(HWT_ENOMEM, "hipMalloc alphas failed");
    double *dal = (double *)mem;
    int64_t *doff = (int64_t *)(dal + total);
    int32_t *tally = (int32_t *)(doff + opts->n + 1);
    rc = HWT_OK;
    if (hipMemcpy(doff, offsets, (size_t)(opts->n + 1) * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(tally, 0, T_NCOUNT * 4) != hipSuccess)
        rc = fail(HWT_EHIP, "hwt_alphas_f64: setup copies failed");
    if (!rc) {
        k_alphas<<<opts->n, WAVE>>>(dev_opts(d.o), DevYield{d.y.n, d.y.rate, d.y.day}, doff, tally, dal);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail(HWT_EHIP, "hwt_alphas_f64: kernel failed");
        else if (hipMemcpy(alphas, dal, total * 8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(HWT_EHIP, "hwt_alphas_f64: D2H failed");
    }
    (void)hipFree(mem);
    return rc;
}

} // extern "C"

// ------------------------------------------------------------- autotuner
namespace {

std::mutex g_tune_mu;
std::map<std::string, hwt_config> g_tune_cache;
bool g_tune_loaded = false;
// the flags the tuner chooses (the caller's others are kept as given)
constexpr int TUNED_FLAGS = HWT_FLAG_NO_ROWS | HWT_FLAG_CONCURRENT | HWT_FLAG_FUSED | HWT_FLAG_NO_FLAT | HWT_FLAG_DIRECT;

int ilog2(int64_t v)
{
    int b = 0;
    while (v > 1) {
        v >>= 1;
        ++b;
    }
    return b;
}

std::string tune_signature(const hwt_config &c, int n, const hwt_stats &st)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::ostringstream os;
    os << "d" << dev << ".p" << c.precision << ".n" << ilog2(n) << ".t" << ilog2(st.tasks) << ".w"
       << ilog2(st.wide_tasks) << ".u" << ilog2(st.node_updates);
    return os.str();
}

void tune_load_file()
{
    const char *path = getenv("HWT_TUNE_CACHE");
    if (!path || !*path) return;
    std::ifstream f(path);
    std::string key;
    hwt_config c;
    hwt_config_init(&c);
    while (f >> key >> c.form >> c.sort >> c.flags >> c.wide_per_cu >> c.narrow_per_cu) g_tune_cache[key] = c;
}

void tune_store_file(const std::string &key, const hwt_config &c)
{
    const char *path = getenv("HWT_TUNE_CACHE");
    if (!path || !*path) return;
    std::ofstream f(path, std::ios::app);
    const int kept = c.flags & TUNED_FLAGS;
    f << key << ' ' << c.form << ' ' << c.sort << ' ' << kept << ' '
      << c.wide_per_cu << ' ' << c.narrow_per_cu << '\n';
}

// Pipeline time per execution, in the mode the caller's plans will run in
// (graph replays unless HWT_FLAG_NO_GRAPH or HWT_FLAG_DIRECT; the first
// warm-up captures the graph): five warm-ups, each synchronised, then seven
// bursts of four executions queued back to back, timed by HIP events around
// each burst; the median burst's time per execution.  Back-to-back periods
// include what a launch costs between executions, which the kernels' own
// stamps (hwt_stats.total_ms) leave out.  warm_ms extends the warm-up to at
// least that much wall time: the first timing of a tuning run would otherwise
// catch the clocks still ramping and read several percent slow.  A candidate
// whose second execution (the first one after the graph capture) already
// takes longer than cutoff_ms is out: that time is returned at once.
constexpr int TUNE_WARM = 5, TUNE_TIMED = 7, TUNE_BURST = 4;
int tune_time(const hwt_options_f64 *o, const hwt_yield_f64 *y, double *out, hwt_config c, double &ms,
              hwt_stats *first, double warm_ms = 0.0, double cutoff_ms = 0.0)
{
    hwt_plan *p = nullptr;
    int rc = hwt_plan_create(&p, o->n, y->n, &c);
    if (rc) return rc;
    auto run = [&](hwt_stats &st) {
        int r = hwt_plan_execute(p, o, y, out, nullptr);
        if (r == HWT_OK) r = hwt_plan_sync(p, &st);
        return r == HWT_EUNSUPPORTED ? HWT_OK : r;  // unsupported options are NaN in every candidate alike
    };
    const auto w0 = std::chrono::steady_clock::now();
    for (int i = 0; rc == HWT_OK; ++i) {
        hwt_stats st{};
        rc = run(st);
        if (i == 0 && first) *first = st;
        if (rc == HWT_OK && i == 1 && cutoff_ms > 0.0 && st.total_ms > cutoff_ms) {
            hwt_plan_destroy(p);
            ms = st.total_ms;
            return HWT_OK;
        }
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        if (i + 1 >= TUNE_WARM && el >= warm_ms) break;
    }
    double t[TUNE_TIMED];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == HWT_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(HWT_EHIP, "hipEventCreate failed");
    for (int i = 0; i < TUNE_TIMED && rc == HWT_OK; ++i) {
        float el = 0.f;
        if (hipEventRecord(e0, p->stream) != hipSuccess) rc = fail(HWT_EHIP, "hipEventRecord failed");
        for (int r = 0; r < TUNE_BURST && rc == HWT_OK; ++r) rc = hwt_plan_execute(p, o, y, out, nullptr);
        if (rc == HWT_OK && (hipEventRecord(e1, p->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&el, e0, e1) != hipSuccess))
            rc = fail(HWT_EHIP, "timing a burst of executions failed");
        t[i] = el / TUNE_BURST;
    }
    if (rc == HWT_OK) {
        hwt_stats st{};
        rc = hwt_plan_sync(p, &st);
        if (rc == HWT_EUNSUPPORTED) rc = HWT_OK;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    hwt_plan_destroy(p);
    if (rc) return rc;
    std::sort(t, t + TUNE_TIMED);
    ms = t[TUNE_TIMED / 2];
    return HWT_OK;
}

} // namespace

extern "C" int hwt_autotune(const hwt_options_f64 *dev_opts, const hwt_yield_f64 *dev_yield, double *dev_out,
                            hwt_config *cfg, hwt_tune_report *report)
{
    g_last_error.clear();
    if (!cfg || !valid_opts_ptrs(dev_opts) || !valid_yield_ptrs(dev_yield) || (dev_opts->n > 0 && !dev_out))
        return fail(HWT_EINVAL, "hwt_autotune: null or negative-size argument");
    const auto w0 = std::chrono::steady_clock::now();
    hwt_tune_report rep{};
    if (cfg->form == HWT_FORM_LANE || dev_opts->n == 0) {  // the caller fixed form 1: nothing to choose
        if (report) *report = rep;
        return HWT_OK;
    }
    const hwt_config base = *cfg;
    hwt_stats st{};
    double base_ms = 0.0;
    int rc = tune_time(dev_opts, dev_yield, dev_out, base, base_ms, &st, 30.0);
    if (rc) return rc;
    rep.default_ms = base_ms;
    const std::string key = tune_signature(base, dev_opts->n, st);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (!g_tune_loaded) {
            tune_load_file();
            g_tune_loaded = true;
        }
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            hwt_config c = base;
            c.form = it->second.form;
            c.sort = it->second.sort;
            c.flags = (base.flags & ~TUNED_FLAGS) | (it->second.flags & TUNED_FLAGS);
            c.wide_per_cu = it->second.wide_per_cu;
            c.narrow_per_cu = it->second.narrow_per_cu;
            *cfg = c;
            rep.cached = 1;
        }
    }
    if (rep.cached) {
        rc = tune_time(dev_opts, dev_yield, dev_out, *cfg, rep.ms, nullptr, 30.0);
        if (rc) return rc;
        rep.candidates = 1;
        if (report) *report = rep;
        return HWT_OK;
    }
    // greedy coordinate descent; a change must win by 1% to be kept (burst
    // medians repeat within about 0.2%)
    hwt_config best = base;
    double best_ms = base_ms;
    int cand = 1;
    auto consider = [&](hwt_config c) -> int {
        double ms = 0.0;
        const int r = tune_time(dev_opts, dev_yield, dev_out, c, ms, nullptr, 0.0, 1.5 * best_ms);
        if (r) return r;
        ++cand;
        if (ms < 0.99 * best_ms) {
            best = c;
            best_ms = ms;
        }
        return HWT_OK;
    };
    for (int so : {(int)HWT_SORT_WORK_DESC, (int)HWT_SORT_HEIGHT_DESC, (int)HWT_SORT_WIDTH_DESC}) {
        if (so == (best.sort ? best.sort : HWT_SORT_WORK_DESC)) continue;
        hwt_config c = best;
        c.sort = so;
        if ((rc = consider(c))) return rc;
    }
    for (int bit : {HWT_FLAG_NO_ROWS, HWT_FLAG_NO_FLAT, HWT_FLAG_CONCURRENT, HWT_FLAG_FUSED, HWT_FLAG_DIRECT}) {
        hwt_config c = best;
        c.flags ^= bit;
        if ((rc = consider(c))) return rc;
    }
    for (int v : {4, 2}) {
        if (v == best.wide_per_cu) continue;
        hwt_config c = best;
        c.wide_per_cu = v;
        if ((rc = consider(c))) return rc;
    }
    for (int v : {8, 4}) {
        if (v == best.narrow_per_cu) continue;
        hwt_config c = best;
        c.narrow_per_cu = v;
        if ((rc = consider(c))) return rc;
    }
    // the inspector's last question: form 1 (one lane per option, columns in
    // HBM) instead of form 2, for FP64 batches small enough that its column
    // buffers (2.75 MB per 64 options) stay modest
    if (base.precision == HWT_PREC_FP64 && dev_opts->n <= (1 << 18)) {
        hwt_config c = best;
        c.form = HWT_FORM_LANE;
        if ((rc = consider(c))) return rc;
    }
    // a winner other than the starting configuration must win again when the
    // two are timed back to back, so that drift over the search (clocks,
    // temperature) cannot pick it
    if (std::memcmp(&best, &base, sizeof(best)) != 0) {
        double again_base = 0.0, again_best = 0.0;
        if ((rc = tune_time(dev_opts, dev_yield, dev_out, base, again_base, nullptr))) return rc;
        if ((rc = tune_time(dev_opts, dev_yield, dev_out, best, again_best, nullptr))) return rc;
        cand += 2;
        rep.default_ms = again_base;  // the fair pair is what the report compares
        if (!(again_best < 0.99 * again_base)) {
            best = base;
            best_ms = again_base;
        } else {
            best_ms = again_best;
        }
    }
    *cfg = best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = best;
        tune_store_file(key, best);
    }
    rep.ms = best_ms;
    rep.candidates = cand;
    rep.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    if (report) *report = rep;
    return HWT_OK;
}


static_assert(ROW_MAXK * (ROWS - 1) + [] {
    int pads = 0;
    for (int p = 0; p < FLAT_PARTS; ++p) pads += flat_per_wave(flat_width(p)) - 1;
    return pads;
}() <= MAX_PADS, "task-list room for the row and flat padding");
