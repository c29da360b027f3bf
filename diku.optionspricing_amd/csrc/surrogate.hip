// surrogate.hip -- the learned surrogate in front of the tree pricer: an
// ensemble of MLPs (8 -> 256 x 4 -> 1, SiLU) trained on FP64 tree prices,
// with hand-written BF16 MFMA kernels for the three 256x256 layers (forward,
// backward, weight gradients) and FP32 kernels for the input layer, the
// output layer and Adam.  Inference is one fused kernel per batch tile that
// runs every member, forms the ensemble mean and spread, and queues the
// options the surrogate must not price for the FP64 tree, which then runs in
// the same stream on the compacted list.
//
// Matrix-core layout (v_mfma_f32_32x32x16_bf16, one wavefront of 64 lanes):
// lane l holds A[l%32][8(l/32) .. 8(l/32)+7] and B[8(l/32) .. +7][l%32]; the
// 16 accumulators of lane l are D[8(v/4) + 4(l/32) + v%4][l%32], v = 0..15.
// Every GEMM here is C = A B^T with both operands stored K-contiguous, so each
// lane's operand is one 16-byte load.
//
// Buffers per member (batch rows b, neurons n, inputs k):
//   z_l     FP32 [b][n]   pre-activations (for SiLU')
//   a_l     BF16 [b][n] and a_l^T [n][b]   (operands of the next forward
//                          GEMM and of this layer's weight-gradient GEMM)
//   dz_l    BF16 [b][n] and dz_l^T [n][b]  (operands of the input-gradient
//                          and weight-gradient GEMMs)
//   W_l     BF16 [n][k] and W_l^T [k][n], refreshed by Adam from FP32 masters
#include "hwt_math.h"
#include "hwt_internal.h"
#include "../../include/hwtsurr.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <sstream>
#include <string>
#include <vector>

using namespace hwt;

namespace {

constexpr int NIN = HWT_SURR_INPUTS;
constexpr int HID = HWT_SURR_HIDDEN;
constexpr int NPAR = HWT_SURR_NPARAMS;
constexpr int PSTRIDE = 200192;          // per-member stride of parameter-shaped vectors
constexpr int LAYER_W = HID * HID;        // elements of one 256x256 weight matrix
constexpr int GEMM_ROWS = 128;            // rows per GEMM workgroup (4 waves x 32)
constexpr int DW_SPLIT = 8;               // split of the batch in the weight-gradient GEMMs
constexpr int HEAD_ROWS = 32;             // rows per head / input-layer workgroup
constexpr int INF_ROWS = 128;             // options per inference workgroup
constexpr int ACT_LD = HID + 8;           // LDS row pitch (bf16) of the inference activations
constexpr double RES = HWT_SURR_RESIDUAL_SCALE;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int woff(int l) { return l == 2 ? HWT_SURR_W2 : (l == 3 ? HWT_SURR_W3 : HWT_SURR_W4); }
constexpr int boff(int l) { return l == 2 ? HWT_SURR_B2 : (l == 3 ? HWT_SURR_B3 : HWT_SURR_B4); }

// BF16 with round-to-nearest-even (the hardware conversion, v_cvt_pk_bf16_f32).
__device__ inline uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// SiLU on the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32).
__device__ inline float sigmoidf(float z)
{
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ inline float silu(float z) { return z * sigmoidf(z); }
__device__ inline float dsilu(float z)
{
    const float s = sigmoidf(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ inline bf16x8 load_frag(const uint16_t *p)
{
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ inline float wave_sum_f(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct SOpts {
    int32_t n;
    const double *strike, *maturity, *length;
    const uint16_t *tu, *steps;
    const double *a, *sigma;
    const int8_t *type;
};
struct SYield {
    int32_t n;
    const double *rate;
    const int32_t *day;
};

SOpts sopts(const hwt_options_f64 &o)
{
    return SOpts{o.n, o.strike, o.maturity, o.length, o.term_unit, o.term_steps, o.a, o.sigma, o.type};
}

// ------------------------------------------------------------------ inputs
// Network inputs and the closed-form baseline of option i.  Returns false
// when the option lies outside the training box (or its tree has no exercise
// step inside it), in which case only the tree may price it.
//
// Baseline: the Hull-White zero-coupon-bond option formula (Jamshidian 1989)
// on the tree's own dates T = e dt and S = n dt and the tree's discount bonds
// P(T), P(S) (the same interpolation of the same curve), so it satisfies the
// same put-call parity as the tree and the residual is independent of the
// option type:
//   sigma_p = sigma/a (1 - e^{-a(S-T)}) sqrt((1 - e^{-2aT}) / (2a))
//   h = ln(100 P(S) / (X P(T))) / sigma_p + sigma_p / 2
//   call = 100 P(S) N(h) - X P(T) N(h - sigma_p),  put = call - 100 P(S) + X P(T).
// Lattice phase: the tree's bond values at the exercise step are close to
// exp(c0 - B~ x_j) with x_j = j dr, so the strike sits at node coordinate
// u = (c0 - ln X) / (B~ dr); the tree price is piecewise linear in X between
// nodes, and sin/cos(2 pi u) let the network place that pattern.
// B~ = dt (1 - q^m) / (1 - q), q = e^{-a dt}, m = n - e (discrete bond
// sensitivity); c0 = ln(100 P(S)/P(T)) minus the continuous-time convexity
// terms B sigma^2 (1-e^{-aT})^2 / (2a^2) + B^2 sigma^2 (1 - e^{-2aT}) / (4a).
// Whether option i's parameters lie inside the training box.
__device__ bool surr_param_box(const SOpts &o, int i)
{
    const double X = o.strike[i], Tb = o.maturity[i], L = o.length[i], a = o.a[i], sg = o.sigma[i];
    const int tu = o.tu[i], st = o.steps[i], ty = o.type[i];
    return tu == 365 && (st == 12 || st == 36) && X >= 60.0 && X <= 100.0 && L >= 1.0 && L <= 30.0 &&
           Tb - L >= 1.0 && Tb - L <= 10.0 && a >= 0.01 && a <= 0.2 && sg >= 0.005 && sg <= 0.02 &&
           (ty == 0 || ty == 1);
}

// ... and its tree has the exercise step inside (always so for box parameters:
// expiry >= 1 y puts e at 12 or more, bond >= expiry + 1 y puts n above e);
// c receives the tree constants.
__device__ bool surr_in_box(const SOpts &o, int i, OptConst &c)
{
    if (!surr_param_box(o, i)) return false;
    return option_constants(o.strike[i], o.maturity[i], o.length[i], o.tu[i], o.steps[i], o.a[i], o.sigma[i],
                            o.type[i], 1 << 30, c) == ST_OK &&
           c.e >= 1 && c.n > c.e;
}

template <class YieldFn>
__device__ bool surr_inputs(const SOpts &o, int i, float x[NIN], double &base, const YieldFn &yield)
{
    const double X = o.strike[i], a = o.a[i], sg = o.sigma[i];
    const int tu = o.tu[i], st = o.steps[i], ty = o.type[i];
    OptConst c;
    if (!surr_in_box(o, i, c)) return false;
    const double T = c.e * c.dt, S = c.n * c.dt;
    const double PT = exp(-yield(T, tu) * T);
    const double PS = exp(-yield(S, tu) * S);
    const double Bc = (1.0 - exp(-a * (S - T))) / a;
    const double sp = sg * Bc * sqrt((1.0 - exp(-2.0 * a * T)) / (2.0 * a));
    const double h = log(100.0 * PS / (X * PT)) / sp + 0.5 * sp;
    const double r2 = 0.70710678118654752440;
    // N(h) = erfc(-h/sqrt2)/2; the put uses the upper tails directly
    const double call = 100.0 * PS * 0.5 * erfc(-h * r2) - X * PT * 0.5 * erfc(-(h - sp) * r2);
    const double put = X * PT * 0.5 * erfc((h - sp) * r2) - 100.0 * PS * 0.5 * erfc(h * r2);
    base = (ty == 1) ? put : call;
    const double q = exp(-a * c.dt);
    const double Bd = c.dt * (1.0 - pow(q, (double)(c.n - c.e))) / (1.0 - q);
    const double e1 = 1.0 - exp(-a * T);
    const double c0 = log(100.0 * PS / PT) - Bc * sg * sg * e1 * e1 / (2.0 * a * a) -
                      Bc * Bc * sg * sg * (1.0 - exp(-2.0 * a * T)) / (4.0 * a);
    const double u = (c0 - log(X)) / (Bd * c.dr);
    const double ph = 2.0 * (u - floor(u));
    x[0] = (float)((X - 80.0) / 20.0);
    x[1] = (float)((T - 15.5) / 14.5);
    x[2] = (float)((S - T - 5.5) / 4.5);
    x[3] = (float)((st - 24.0) / 12.0);
    x[4] = (float)((a - 0.105) / 0.095);
    x[5] = (float)((sg - 0.0125) / 0.0075);
    x[6] = (float)sinpi(ph);
    x[7] = (float)cospi(ph);
    return true;
}

__global__ void k_train_inputs(SOpts o, SYield y, const double *__restrict__ price, float *__restrict__ X,
                               float *__restrict__ target, int32_t *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= o.n) return;
    float x[NIN];
    double base = 0.0;
    const bool ok = surr_inputs(o, i, x, base,
                                [&](double t, int tu) { return yield_at(t, tu, y.rate, y.day, y.n); });
#pragma unroll
    for (int k = 0; k < NIN; ++k) X[(size_t)i * NIN + k] = ok ? x[k] : 0.0f;
    target[i] = ok ? (float)((price[i] - base) / RES) : 0.0f;
    if (!ok) atomicAdd(bad, 1);
}

// ---------------------------------------------------------------- training
// Gathers `count` rows of every member into the batch buffers (member pitch
// `pitch` rows); each member has its own order.
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ target, const int32_t *__restrict__ perm,
                         int64_t perm_stride, int64_t offset, int count, int pitch, float *__restrict__ xb,
                         float *__restrict__ tb, int m0)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = m0 + blockIdx.y;
    if (r >= count) return;
    const int row = perm[(size_t)blockIdx.y * perm_stride + offset + r];
#pragma unroll
    for (int k = 0; k < NIN; ++k) xb[((size_t)m * pitch + r) * NIN + k] = X[(size_t)row * NIN + k];
    tb[(size_t)m * pitch + r] = target[row];
}

struct Net {
    // per member, strides in elements
    float *P;          // FP32 masters [member][PSTRIDE]
    uint16_t *Wb;      // BF16 [member][3][256*256] (layers 2..4), row-major [n][k]
    uint16_t *WbT;     // BF16 transposes [member][3][256*256], [k][n]
};

struct Acts {
    int B;                 // rows per member
    float *xb, *tb;        // [m][B][8], [m][B]
    float *z;              // [m][4][B][256]  z1..z4
    uint16_t *a;           // [m][3][B][256]  a1..a3 (bf16)
    uint16_t *aT;          // [m][3][256][B]
    float *a4;             // [m][B][256]     FP32 last hidden activation
    uint16_t *dz;          // [m][3][B][256]  dz2..dz4 (bf16)
    uint16_t *dzT;         // [m][3][256][B]
    float *dz1T;           // [m][256][B]
    float *y;              // [m][B] network outputs
    float *gw;             // [m][3][DW_SPLIT][256*256] weight-gradient partials (layers 2..4)
    float *gs;             // [m][PSTRIDE] small gradients (W1, b1..b4), dense
    float *hp;             // [m][B/HEAD_ROWS][257] head partials (dw5, db5)
    double *loss;          // [m] running sum of squared errors
    float *g;              // [m][PSTRIDE] dense gradient
};

__host__ __device__ inline size_t zoff(const Acts &s, int m, int l) { return (((size_t)m * 4 + (l - 1)) * s.B) * HID; }
__host__ __device__ inline size_t aoff(const Acts &s, int m, int l) { return (((size_t)m * 3 + (l - 1)) * s.B) * HID; }
__host__ __device__ inline size_t dzoff(const Acts &s, int m, int l) { return (((size_t)m * 3 + (l - 2)) * s.B) * HID; }

// Input layer (FP32): z1 = W1 x + b1, a1 = silu(z1).  Block = 256 neurons x
// HEAD_ROWS rows.
__global__ __launch_bounds__(HID) void k_layer1(Net net, Acts s, int m0)
{
    const int m = m0 + blockIdx.y;
    const int n = threadIdx.x;
    const int r0 = blockIdx.x * HEAD_ROWS;
    __shared__ float xs[HEAD_ROWS * NIN];
    if (n < HEAD_ROWS * NIN) xs[n] = s.xb[((size_t)m * s.B + r0) * NIN + n];
    __syncthreads();
    const float *P = net.P + (size_t)m * PSTRIDE;
    float w[NIN];
#pragma unroll
    for (int k = 0; k < NIN; ++k) w[k] = P[HWT_SURR_W1 + n * NIN + k];
    const float b = P[HWT_SURR_B1 + n];
    float *z = s.z + zoff(s, m, 1);
    uint16_t *a = s.a + aoff(s, m, 1);
    uint16_t *aT = s.aT + aoff(s, m, 1);
    uint32_t packed[HEAD_ROWS / 2];
#pragma unroll
    for (int r = 0; r < HEAD_ROWS; ++r) {
        float acc = b;
#pragma unroll
        for (int k = 0; k < NIN; ++k) acc = fmaf(w[k], xs[r * NIN + k], acc);
        const size_t row = (size_t)(r0 + r);
        z[row * HID + n] = acc;
        const uint16_t h = to_bf16(silu(acc));
        a[row * HID + n] = h;
        if (r & 1) packed[r / 2] |= (uint32_t)h << 16;
        else packed[r / 2] = h;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(aT + (size_t)n * s.B + r0);
#pragma unroll
    for (int q = 0; q < HEAD_ROWS / 8; ++q)
        dst[q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
}

enum GemmMode { G_FWD = 0, G_FWD_LAST = 1, G_DX = 2, G_DX_FIRST = 3, G_DW = 4 };

struct GemmArgs {
    const uint16_t *A;   // [M][K] (member stride sA)
    const uint16_t *Bm;  // [N][K] (member stride sB)
    size_t sA, sB;
    int lda, ldb, K;     // K: reduction length handled by one workgroup
    int kz;              // split index stride: workgroup z = member * nsplit + split
    int nsplit;
    int m0;              // first member
    Net net;
    Acts s;
    int layer;           // forward: output layer l; DX: layer whose input gradient is formed; DW: layer
};

// C = A B^T on BF16 MFMA, FP32 accumulation.  Workgroup: 4 waves, 128 x 64
// outputs (each wave 32 rows x two 32-column blocks).  The epilogue depends
// on the mode (see the per-mode comments below).
template <int MODE>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mz = blockIdx.z;
    const int m = g.m0 + (MODE == G_DW ? mz / g.nsplit : mz);
    const int split = MODE == G_DW ? mz % g.nsplit : 0;
    const int row0 = blockIdx.x * GEMM_ROWS + wave * 32;
    const int col0 = blockIdx.y * 64;
    const int kbase = split * g.K;
    const uint16_t *A = g.A + (size_t)m * g.sA + (size_t)(row0 + (lane & 31)) * g.lda + kbase + 8 * (lane >> 5);
    const uint16_t *B0 = g.Bm + (size_t)m * g.sB + (size_t)(col0 + (lane & 31)) * g.ldb + kbase + 8 * (lane >> 5);
    const uint16_t *B1 = B0 + (size_t)32 * g.ldb;
    f32x16 acc0 = {}, acc1 = {};
    for (int k = 0; k < g.K; k += 16) {
        const bf16x8 a = load_frag(A + k);
        const bf16x8 b0 = load_frag(B0 + k);
        const bf16x8 b1 = load_frag(B1 + k);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc1, 0, 0, 0);
    }
    const Acts &s = g.s;
    const int l = g.layer;
    const float *P = g.net.P + (size_t)m * PSTRIDE;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const f32x16 &acc = cb ? acc1 : acc0;
        const int col = col0 + cb * 32 + (lane & 31);
#pragma unroll
        for (int grp = 0; grp < 4; ++grp) {
            // this lane's rows rbase .. rbase+3 of column col
            const int rbase = row0 + 8 * grp + 4 * (lane >> 5);
            float ov[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = rbase + q;
                const float v = acc[4 * grp + q];
                if (MODE == G_FWD || MODE == G_FWD_LAST) {
                    // z_l = a_{l-1} W_l^T + b_l; a_l = silu(z_l)
                    const float zz = v + P[boff(l) + col];
                    s.z[zoff(s, m, l) + (size_t)row * HID + col] = zz;
                    ov[q] = silu(zz);
                    if (MODE == G_FWD) s.a[aoff(s, m, l) + (size_t)row * HID + col] = to_bf16(ov[q]);
                    else s.a4[(size_t)m * s.B * HID + (size_t)row * HID + col] = ov[q];
                } else if (MODE == G_DX || MODE == G_DX_FIRST) {
                    // d a_{l-1} = dz_l W_l; dz_{l-1} = d a_{l-1} * silu'(z_{l-1})
                    ov[q] = v * dsilu(s.z[zoff(s, m, l - 1) + (size_t)row * HID + col]);
                    if (MODE == G_DX) s.dz[dzoff(s, m, l - 1) + (size_t)row * HID + col] = to_bf16(ov[q]);
                } else {
                    // dW_l[n][k] partial over this split of the batch
                    s.gw[(((size_t)m * 3 + (l - 2)) * DW_SPLIT + split) * LAYER_W + (size_t)row * HID + col] = v;
                    ov[q] = 0.0f;
                }
            }
            if (MODE == G_FWD || MODE == G_DX) {
                uint16_t *T = (MODE == G_FWD) ? s.aT + aoff(s, m, l) : s.dzT + dzoff(s, m, l - 1);
                const uint2 pk = make_uint2((uint32_t)to_bf16(ov[0]) | ((uint32_t)to_bf16(ov[1]) << 16),
                                            (uint32_t)to_bf16(ov[2]) | ((uint32_t)to_bf16(ov[3]) << 16));
                *reinterpret_cast<uint2 *>(T + (size_t)col * s.B + rbase) = pk;
            } else if (MODE == G_DX_FIRST) {
                *reinterpret_cast<float4 *>(s.dz1T + (size_t)m * HID * s.B + (size_t)col * s.B + rbase) =
                    make_float4(ov[0], ov[1], ov[2], ov[3]);
            }
        }
    }
}

// Output layer (FP32), loss and the start of the backward pass:
//   y = w5 . a4 + b5,  loss = mean (y - t)^2,  dy = 2 (y - t) / B,
//   dz4 = dy w5 * silu'(z4) (BF16, both layouts),
//   per-block partial sums of dw5 = sum dy a4 and db5 = sum dy.
__global__ __launch_bounds__(HID) void k_head(Net net, Acts s, int m0, int Bn)
{
    const int m = m0 + blockIdx.y;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r0 = blockIdx.x * HEAD_ROWS;
    const float *P = net.P + (size_t)m * PSTRIDE;
    __shared__ float dys[HEAD_ROWS];
    __shared__ float lossp[4];
    const float *a4 = s.a4 + (size_t)m * s.B * HID;
    float lsum = 0.0f;
    for (int r = wave; r < HEAD_ROWS; r += 4) {
        const size_t row = r0 + r;
        float p = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) p = fmaf(P[HWT_SURR_W5 + lane + 64 * q], a4[row * HID + lane + 64 * q], p);
        const float yv = wave_sum_f(p) + P[HWT_SURR_B5];
        const float err = yv - s.tb[(size_t)m * s.B + row];
        if (lane == 0) {
            s.y[(size_t)m * s.B + row] = yv;
            dys[r] = 2.0f * err / (float)Bn;
        }
        lsum += err * err;
    }
    if (lane == 0) lossp[wave] = lsum;
    __syncthreads();
    const int n = t;
    const float w5 = P[HWT_SURR_W5 + n];
    const float *z4 = s.z + zoff(s, m, 4);
    uint16_t *dz = s.dz + dzoff(s, m, 4);
    float dw = 0.0f, db = 0.0f;
    uint32_t packed[HEAD_ROWS / 2];
#pragma unroll
    for (int r = 0; r < HEAD_ROWS; ++r) {
        const size_t row = r0 + r;
        const float dy = dys[r];
        dw = fmaf(dy, a4[row * HID + n], dw);
        db += dy;
        const uint16_t h = to_bf16(dy * w5 * dsilu(z4[row * HID + n]));
        dz[row * HID + n] = h;
        if (r & 1) packed[r / 2] |= (uint32_t)h << 16;
        else packed[r / 2] = h;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(s.dzT + dzoff(s, m, 4) + (size_t)n * s.B + r0);
#pragma unroll
    for (int q = 0; q < HEAD_ROWS / 8; ++q)
        dst[q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
    float *hp = s.hp + ((size_t)m * (s.B / HEAD_ROWS) + blockIdx.x) * (HID + 1);
    hp[n] = dw;
    if (n == 0) {
        hp[HID] = db;
        atomicAdd(&s.loss[m], (double)(lossp[0] + lossp[1] + lossp[2] + lossp[3]));
    }
}

// Gradients of the input layer and of the biases, one workgroup per neuron:
//   dW1[n][i] = sum_b dz1[b][n] x[b][i], db1[n] = sum_b dz1[b][n],
//   db_l[n] = sum_b dz_l[b][n] (l = 2..4).
__global__ __launch_bounds__(256) void k_small_grads(Acts s, int m0, int Bn)
{
    const int m = m0 + blockIdx.y;
    const int n = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.0f;
    const float *dz1 = s.dz1T + (size_t)m * HID * s.B + (size_t)n * s.B;
    const float *xb = s.xb + (size_t)m * s.B * NIN;
    for (int r = t; r < Bn; r += 256) {
        const float d = dz1[r];
#pragma unroll
        for (int k = 0; k < NIN; ++k) acc[k] = fmaf(d, xb[(size_t)r * NIN + k], acc[k]);
        acc[8] += d;
#pragma unroll
        for (int l = 2; l <= 4; ++l) acc[7 + l] += from_bf16(s.dzT[dzoff(s, m, l) + (size_t)n * s.B + r]);
    }
    __shared__ float red[4][12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const float v = wave_sum_f(acc[k]);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (t < 12) {
        const float v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        float *gs = s.gs + (size_t)m * PSTRIDE;
        if (t < NIN) gs[HWT_SURR_W1 + n * NIN + t] = v;
        else if (t == 8) gs[HWT_SURR_B1 + n] = v;
        else gs[boff(t - 7) + n] = v;
    }
}

// Dense gradient of every parameter from the partial sums.
__global__ void k_grad_reduce(Acts s, int m0, int nhead)
{
    const int m = m0 + blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    float g;
    if (p >= HWT_SURR_W2 && p < HWT_SURR_B4 && !(p >= HWT_SURR_B2 && p < HWT_SURR_W3) &&
        !(p >= HWT_SURR_B3 && p < HWT_SURR_W4)) {
        const int l = p < HWT_SURR_B2 ? 2 : (p < HWT_SURR_B3 ? 3 : 4);
        const int idx = p - woff(l);
        const float *src = s.gw + ((size_t)m * 3 + (l - 2)) * DW_SPLIT * LAYER_W + idx;
        g = 0.0f;
#pragma unroll
        for (int q = 0; q < DW_SPLIT; ++q) g += src[(size_t)q * LAYER_W];
    } else if (p >= HWT_SURR_W5) {
        const int idx = p - HWT_SURR_W5;  // 0..255 = w5, 256 = b5
        const float *src = s.hp + (size_t)m * (s.B / HEAD_ROWS) * (HID + 1) + idx;
        g = 0.0f;
        for (int q = 0; q < nhead; ++q) g += src[(size_t)q * (HID + 1)];
    } else {
        g = s.gs[(size_t)m * PSTRIDE + p];
    }
    s.g[(size_t)m * PSTRIDE + p] = g;
}

// Adam on the FP32 masters; refreshes the BF16 copies (and transposes) of
// the 256x256 layers.  c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t).
__global__ void k_adam(Net net, const float *__restrict__ grad, float *__restrict__ m1, float *__restrict__ m2,
                       float lr, float beta1, float beta2, float eps, float c1, float c2)
{
    const int m = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    const size_t i = (size_t)m * PSTRIDE + p;
    const float g = grad[i];
    const float mm = beta1 * m1[i] + (1.0f - beta1) * g;
    const float vv = beta2 * m2[i] + (1.0f - beta2) * g * g;
    m1[i] = mm;
    m2[i] = vv;
    const float w = net.P[i] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    net.P[i] = w;
    int l = 0;
    if (p >= HWT_SURR_W2 && p < HWT_SURR_B2) l = 2;
    else if (p >= HWT_SURR_W3 && p < HWT_SURR_B3) l = 3;
    else if (p >= HWT_SURR_W4 && p < HWT_SURR_B4) l = 4;
    if (l) {
        const int idx = p - woff(l), nn = idx / HID, kk = idx % HID;
        const size_t base = ((size_t)m * 3 + (l - 2)) * LAYER_W;
        const uint16_t h = to_bf16(w);
        net.Wb[base + idx] = h;
        net.WbT[base + (size_t)kk * HID + nn] = h;
    }
}

// BF16 copies from the masters (after initialisation or a load).
__global__ void k_refresh_bf16(Net net, int members)
{
    const int m = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= members || idx >= 3 * LAYER_W) return;
    const int l = 2 + idx / LAYER_W, e = idx % LAYER_W, nn = e / HID, kk = e % HID;
    const uint16_t h = to_bf16(net.P[(size_t)m * PSTRIDE + woff(l) + e]);
    const size_t base = ((size_t)m * 3 + (l - 2)) * LAYER_W;
    net.Wb[base + e] = h;
    net.WbT[base + (size_t)kk * HID + nn] = h;
}

// --------------------------------------------------------------- inference
// One workgroup (4 waves) prices INF_ROWS = 128 options with every member.
// LDS holds the inputs and the BF16 activation tile act[row][neuron].  The
// hidden layers run as D = W A^T on the matrix cores: the weights are the
// A operand (loaded from L2, K-contiguous rows of W), the activations the B
// operand (LDS rows), so D[neuron][row] lands with four consecutive neurons
// of one row in each lane's accumulator quad and the epilogue writes them as
// one 8-byte store.  Wave w owns neurons [64w, 64w + 64) for all 128 rows,
// so each weight is read once per tile; a layer reads the whole tile, a
// barrier, then the epilogue overwrites it in place.  Layer 1 runs the same
// way on the FP32 matrix instructions (K = 8 in four 32x32x2 steps); every
// accumulator starts at its bias.  The output layer reduces per row across
// lanes and waves in a fixed order.  The yield curve is staged in LDS for the
// inputs.
constexpr int INF_YMAX = 256;  // keeps two 128-row tiles per CU in LDS

struct InferArgs {
    SOpts o;
    SYield y;
    Net net;
    int members;
    float threshold;
    double *out;       // prices (accepted options) or null
    double *mean;      // ensemble-mean price (predict) or null
    double *spread;    // member spread (predict) or null
    int8_t *in_box;    // (predict) or null
    int32_t *fb_idx;   // fallback list or null
    int32_t *fb_count; // [0] total, [1] out of box
    const int32_t *rows;      // option of each row (null: row i is option i)
    const int32_t *row_count; // number of listed rows (with `rows`)
};

// Two floats to a packed BF16 pair (a in the low half), round to nearest
// even, as one v_cvt_pk_bf16_f32 (the value-by-value form costs a convert,
// a shift and an or per pair).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t pack_bf16x2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// Accumulators of a wave's 2 x 4 blocks D[neuron][row] seeded with the
// layer's bias: lane l's value v belongs to neuron 32 mb + 8 (v/4) + 4 (l/32)
// + v%4 of the wave's 64, whatever the row block.
__device__ inline void bias_init(f32x16 (&acc)[2][4], const float *bias, int n0, int lane)
{
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        f32x16 b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 bv = *reinterpret_cast<const float4 *>(bias + n0 + 32 * mb + 8 * q + 4 * (lane >> 5));
            b[4 * q] = bv.x;
            b[4 * q + 1] = bv.y;
            b[4 * q + 2] = bv.z;
            b[4 * q + 3] = bv.w;
        }
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) acc[mb][nb] = b;
    }
}

// SiLU of the pre-activations, rounded to BF16, into the activation tile:
// the lane's four consecutive neurons of one row as one 8-byte store.
__device__ inline void silu_to_act(const f32x16 (&acc)[2][4], uint16_t *act, int n0, int lane)
{
    const int mrow = lane & 31;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = n0 + 32 * mb + 8 * q + 4 * (lane >> 5);
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const f32x16 &a = acc[mb][nb];
                const uint2 pk = make_uint2(pack_bf16x2(silu(a[4 * q]), silu(a[4 * q + 1])),
                                            pack_bf16x2(silu(a[4 * q + 2]), silu(a[4 * q + 3])));
                *reinterpret_cast<uint2 *>(act + (32 * nb + mrow) * ACT_LD + col) = pk;
            }
        }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_infer(InferArgs g)
{
    static_assert(INF_ROWS == 128, "the layer-1 and head mappings assume 128 rows per tile");
    __shared__ __attribute__((aligned(16))) uint16_t act[INF_ROWS * ACT_LD];
    __shared__ float xs[INF_ROWS * NIN];
    __shared__ double bases[INF_ROWS];
    __shared__ float hp[4 * INF_ROWS];
    __shared__ float ysum[INF_ROWS], ymin[INF_ROWS], ymax[INF_ROWS];
    __shared__ int32_t okf[INF_ROWS];
    __shared__ int32_t fb_local[INF_ROWS];
    __shared__ int32_t fb_n, fb_out, fb_base;
    __shared__ double yr[INF_YMAX];
    __shared__ int32_t yd[INF_YMAX];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r0 = blockIdx.x * INF_ROWS;
    const int nrows = g.rows ? *g.row_count : g.o.n;
    if (r0 >= nrows) return;  // the whole tile is past the listed rows

    // curves that fit are read from LDS, by binary search when their days
    // never decrease; both lookups give the reference's value
    const bool staged = g.y.n <= INF_YMAX;
    int unsorted = 0;
    if (staged) {
        for (int k = t; k < g.y.n; k += 256) {
            yr[k] = g.y.rate[k];
            yd[k] = g.y.day[k];
            if (k + 1 < g.y.n && g.y.day[k] > g.y.day[k + 1]) unsorted = 1;
        }
    }
    // unconditional barrier: it also orders the LDS writes below after every
    // thread's curve staging, whatever the curve's length
    const int any_unsorted = __syncthreads_or(unsorted);
    const bool sorted = staged && !any_unsorted;
    if (t < INF_ROWS) {
        const int i = r0 + t < nrows ? (g.rows ? g.rows[r0 + t] : r0 + t) : -1;
        float x[NIN] = {};
        double base = 0.0;
        const bool ok = i >= 0 && surr_inputs(g.o, i, x, base, [&](double tt, int tu) {
            return sorted ? yield_at_sorted(tt, tu, yr, yd, g.y.n) : yield_at(tt, tu, g.y.rate, g.y.day, g.y.n);
        });
#pragma unroll
        for (int k = 0; k < NIN; ++k) xs[t * NIN + k] = x[k];
        bases[t] = base;
        okf[t] = ok;
        ysum[t] = 0.0f;
        ymin[t] = INFINITY;
        ymax[t] = -INFINITY;
    }
    if (t == 0) fb_n = fb_out = 0;
    __syncthreads();

    const int mrow = lane & 31, khalf = 8 * (lane >> 5);
    const int n0 = wave * 64;  // this wave's output neurons in the hidden layers
    for (int m = 0; m < g.members; ++m) {
        const float *P = g.net.P + (size_t)m * PSTRIDE;
        // layer 1 (FP32 MFMA 32x32x2, K = 8 in four steps): D[neuron][row] =
        // b1 + W1 x^T, with the lane's A value W1[n][k] and B value x[row][k]
        {
            f32x16 acc[2][4];
            bias_init(acc, P + HWT_SURR_B1, n0, lane);
#pragma unroll
            for (int k = 0; k < NIN; k += 2) {
                const int kk = k + (lane >> 5);
                const float a0 = P[HWT_SURR_W1 + (n0 + mrow) * NIN + kk];
                const float a1 = P[HWT_SURR_W1 + (n0 + 32 + mrow) * NIN + kk];
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const float bx = xs[(32 * nb + mrow) * NIN + kk];
                    acc[0][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bx, acc[0][nb], 0, 0, 0);
                    acc[1][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bx, acc[1][nb], 0, 0, 0);
                }
            }
            silu_to_act(acc, act, n0, lane);
        }
        __syncthreads();
        for (int l = 2; l <= 4; ++l) {
            const uint16_t *W = g.net.Wb + ((size_t)m * 3 + (l - 2)) * LAYER_W;
            const uint16_t *wa0 = W + (size_t)(n0 + mrow) * HID + khalf;
            const uint16_t *wa1 = wa0 + (size_t)32 * HID;
            const uint16_t *ab = act + mrow * ACT_LD + khalf;
            f32x16 acc[2][4];
            bias_init(acc, P + boff(l), n0, lane);
#pragma unroll 2
            for (int k = 0; k < HID; k += 16) {
                const bf16x8 a0 = load_frag(wa0 + k), a1 = load_frag(wa1 + k);
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const bf16x8 b =
                        __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(ab + nb * 32 * ACT_LD + k));
                    acc[0][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0][nb], 0, 0, 0);
                    acc[1][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1][nb], 0, 0, 0);
                }
            }
            __syncthreads();  // every wave has read this layer's input tile
            if (l < 4) {
                silu_to_act(acc, act, n0, lane);
                __syncthreads();
            } else {
                // output layer in FP32 from the accumulators: per row, the
                // lane's 32 neurons, then the lane pair holding the same row,
                // then the four waves
                float p[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int col = n0 + 32 * mb + 8 * q + 4 * (lane >> 5);
                        const float4 w5 = *reinterpret_cast<const float4 *>(P + HWT_SURR_W5 + col);
#pragma unroll
                        for (int nb = 0; nb < 4; ++nb) {
                            const f32x16 &a = acc[mb][nb];
                            p[nb] = fmaf(w5.x, silu(a[4 * q]), p[nb]);
                            p[nb] = fmaf(w5.y, silu(a[4 * q + 1]), p[nb]);
                            p[nb] = fmaf(w5.z, silu(a[4 * q + 2]), p[nb]);
                            p[nb] = fmaf(w5.w, silu(a[4 * q + 3]), p[nb]);
                        }
                    }
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const float v = p[nb] + __shfl_xor(p[nb], 32, 64);
                    if (lane < 32) hp[wave * INF_ROWS + 32 * nb + lane] = v;
                }
                __syncthreads();
                if (t < INF_ROWS) {
                    const float yv = ((hp[t] + hp[INF_ROWS + t]) + (hp[2 * INF_ROWS + t] + hp[3 * INF_ROWS + t])) +
                                     P[HWT_SURR_B5];
                    ysum[t] += yv;
                    ymin[t] = fminf(ymin[t], yv);
                    ymax[t] = fmaxf(ymax[t], yv);
                }
            }
        }
    }
    if (t < INF_ROWS) {
        const int i = r0 + t < nrows ? (g.rows ? g.rows[r0 + t] : r0 + t) : -1;
        if (i >= 0) {
            const double mean = bases[t] + RES * (double)(ysum[t] / (float)g.members);
            const double spread = RES * (double)(ymax[t] - ymin[t]);
            const bool ok = okf[t] != 0;
            if (g.mean) g.mean[i] = ok ? mean : NAN;
            if (g.spread) g.spread[i] = ok ? spread : NAN;
            if (g.in_box) g.in_box[i] = (int8_t)ok;
            // a non-finite member output makes the mean non-finite while
            // fminf/fmaxf would drop it from the spread: never accepted
            const bool accept = ok && isfinite(mean) && spread <= (double)g.threshold;
            if (g.out && accept) g.out[i] = mean;
            if (g.fb_idx && !accept) {
                fb_local[atomicAdd(&fb_n, 1)] = i;
                if (!ok) atomicAdd(&fb_out, 1);
            }
        }
    }
    if (g.fb_idx) {
        __syncthreads();
        if (t == 0 && fb_n > 0) {
            fb_base = atomicAdd(&g.fb_count[0], fb_n);
            atomicAdd(&g.fb_count[1], fb_out);
        }
        __syncthreads();
        if (t < fb_n) g.fb_idx[fb_base + t] = fb_local[t];
    }
}

// Splits a batch before inference: options outside the training box go to
// the fallback list (fb_idx, fb_count[0] and [1]), the others to the row list
// of the inference kernel (rows, row_count).  A block takes BOX_CHUNK options
// and stages both lists in LDS, so the global counters see one atomic per
// list per block (same-address atomics serialise in L2: one per 256 options
// cost more than the whole test).  Order inside the lists is arbitrary.
constexpr int BOX_CHUNK = 4096;

__global__ __launch_bounds__(256) void k_box(SOpts o, int32_t *__restrict__ fb_idx, int32_t *__restrict__ fb_count,
                                             int32_t *__restrict__ rows, int32_t *__restrict__ row_count)
{
    __shared__ int32_t nin, nout, bin, bout;
    __shared__ int32_t lin[BOX_CHUNK], lout[BOX_CHUNK];
    const int t = threadIdx.x, i0 = blockIdx.x * BOX_CHUNK;
    if (t == 0) nin = nout = 0;
    __syncthreads();
    for (int i = i0 + t; i < min(i0 + BOX_CHUNK, o.n); i += 256) {
        // the parameter test alone; a listed option that still fails the tree
        // test is sent to the fallback list by the inference kernel
        if (surr_param_box(o, i)) lin[atomicAdd(&nin, 1)] = i;
        else lout[atomicAdd(&nout, 1)] = i;
    }
    __syncthreads();
    if (t == 0) {
        bin = nin ? atomicAdd(row_count, nin) : 0;
        bout = nout ? atomicAdd(&fb_count[0], nout) : 0;
        if (nout) atomicAdd(&fb_count[1], nout);
    }
    __syncthreads();
    for (int k = t; k < nin; k += 256) rows[bin + k] = lin[k];
    for (int k = t; k < nout; k += 256) fb_idx[bout + k] = lout[k];
}

// Fallback list -> compact option arrays, and prices back to input order.
struct Compact {
    double *strike, *maturity, *length, *a, *sigma;
    uint16_t *tu, *steps;
    int8_t *type;
};

__global__ void k_fb_gather(SOpts o, const int32_t *__restrict__ idx, const int32_t *__restrict__ count, Compact c)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count[0]) return;
    const int i = idx[j];
    c.strike[j] = o.strike[i];
    c.maturity[j] = o.maturity[i];
    c.length[j] = o.length[i];
    c.a[j] = o.a[i];
    c.sigma[j] = o.sigma[i];
    c.tu[j] = o.tu[i];
    c.steps[j] = o.steps[i];
    c.type[j] = o.type[i];
}

__global__ void k_fb_scatter(const int32_t *__restrict__ idx, const int32_t *__restrict__ count,
                             const double *__restrict__ tmp, double *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count[0]) return;
    out[idx[j]] = tmp[j];
}

// The device's constant-rate wall clock into slot i (a pass's timing marks,
// which events cannot give inside a graph).
__global__ void k_stamp(uint64_t *t, int i) { t[i] = __builtin_amdgcn_s_memrealtime(); }

__global__ void k_zero_i32(int32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_f64(double *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

int fail(int code, const std::string &msg) { return hwt_internal_fail(code, msg.c_str()); }

#define SURR_HIP(call)                                                                         \
    do {                                                                                       \
        hipError_t err_ = (call);                                                              \
        if (err_ != hipSuccess)                                                                \
            return fail(HWT_EHIP, std::string(#call) + ": " + hipGetErrorString(err_));        \
    } while (0)

bool valid_opts(const hwt_options_f64 *o)
{
    return o && o->n >= 0 &&
           (o->n == 0 || (o->strike && o->maturity && o->length && o->term_unit && o->term_steps && o->a &&
                          o->sigma && o->type));
}
bool valid_yield(const hwt_yield_f64 *y) { return y && y->n > 0 && y->rate && y->day; }

} // namespace

struct hwt_surrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    hwt_surr_config cfg{};
    int members = 4, B = 8192;
    int32_t train_cap = 0, infer_cap = 0, ycap = 0;
    int32_t ntrain = 0;
    int64_t step = 0;
    int epochs_done = 0;
    std::mt19937_64 rng;
    // device
    void *mem = nullptr;
    size_t bytes = 0;
    Net net{};
    Acts acts{};
    float *m1 = nullptr, *m2 = nullptr;
    float *X = nullptr, *target = nullptr;
    int32_t *perm = nullptr;  // [members][train_cap]
    int32_t *rows = nullptr;  // test hook rows [B]
    int32_t *bad = nullptr;
    int32_t *fb_idx = nullptr;   // fallback list: out-of-box options, then uncertain in-box ones
    int32_t *inf_rows = nullptr; // in-box options, the rows of the inference kernel
    int32_t *fb_count = nullptr; // [0] fallback total, [1] of which out of box, [4] inference rows;
                                 // [8..13] three u64 wall-clock stamps of a pricing pass
    double *fb_tmp = nullptr;
    Compact comp{};
    int32_t *h_count = nullptr;  // pinned copy of fb_count [16]
    int wall_khz = 100000;       // rate of the device's constant wall clock
    double *h_loss = nullptr;    // pinned [members]
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    // the pricing pass captured as one graph, and the arguments it was
    // captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct PassKey {
        hwt_options_f64 o;
        hwt_yield_f64 y;
        double *out;
        const hwt_plan *tree;
        uint64_t tree_serial;
        float threshold;
        int32_t members;
    } key{};
};

namespace {

int gemm_launch(hwt_surrogate *s, int mode, int l, int m0, int mc, int Bn)
{
    GemmArgs g{};
    g.net = s->net;
    g.s = s->acts;
    g.m0 = m0;
    g.layer = l;
    g.nsplit = 1;
    const Acts &a = s->acts;
    const size_t actm = (size_t)3 * a.B * HID;  // member stride of a/aT/dz/dzT
    dim3 grid;
    switch (mode) {
    case G_FWD:
    case G_FWD_LAST:
        // A = a_{l-1} [B][256], B = W_l [256][256]
        g.A = a.a + aoff(a, 0, l - 1);
        g.sA = actm;
        g.lda = HID;
        g.Bm = s->net.Wb + (size_t)(l - 2) * LAYER_W;
        g.sB = (size_t)3 * LAYER_W;
        g.ldb = HID;
        g.K = HID;
        grid = dim3(Bn / GEMM_ROWS, HID / 64, mc);
        break;
    case G_DX:
    case G_DX_FIRST:
        // A = dz_l [B][256], B = W_l^T [256][256]
        g.A = a.dz + dzoff(a, 0, l);
        g.sA = actm;
        g.lda = HID;
        g.Bm = s->net.WbT + (size_t)(l - 2) * LAYER_W;
        g.sB = (size_t)3 * LAYER_W;
        g.ldb = HID;
        g.K = HID;
        grid = dim3(Bn / GEMM_ROWS, HID / 64, mc);
        break;
    default:
        // A = dz_l^T [256][B], B = a_{l-1}^T [256][B], K = B split DW_SPLIT ways
        g.A = a.dzT + dzoff(a, 0, l);
        g.sA = actm;
        g.lda = a.B;
        g.Bm = a.aT + aoff(a, 0, l - 1);
        g.sB = actm;
        g.ldb = a.B;
        g.K = Bn / DW_SPLIT;
        g.nsplit = DW_SPLIT;
        grid = dim3(HID / GEMM_ROWS, HID / 64, mc * DW_SPLIT);
        break;
    }
    switch (mode) {
    case G_FWD: k_gemm<G_FWD><<<grid, 256, 0, s->stream>>>(g); break;
    case G_FWD_LAST: k_gemm<G_FWD_LAST><<<grid, 256, 0, s->stream>>>(g); break;
    case G_DX: k_gemm<G_DX><<<grid, 256, 0, s->stream>>>(g); break;
    case G_DX_FIRST: k_gemm<G_DX_FIRST><<<grid, 256, 0, s->stream>>>(g); break;
    default: k_gemm<G_DW><<<grid, 256, 0, s->stream>>>(g); break;
    }
    return HWT_OK;
}

// Forward + backward of members [m0, m0+mc) on the gathered rows; leaves the
// dense gradient in acts.g.
int forward_backward(hwt_surrogate *s, int m0, int mc, int Bn)
{
    hipStream_t st = s->stream;
    k_layer1<<<dim3(Bn / HEAD_ROWS, mc), HID, 0, st>>>(s->net, s->acts, m0);
    gemm_launch(s, G_FWD, 2, m0, mc, Bn);
    gemm_launch(s, G_FWD, 3, m0, mc, Bn);
    gemm_launch(s, G_FWD_LAST, 4, m0, mc, Bn);
    k_head<<<dim3(Bn / HEAD_ROWS, mc), HID, 0, st>>>(s->net, s->acts, m0, Bn);
    gemm_launch(s, G_DX, 4, m0, mc, Bn);
    gemm_launch(s, G_DX, 3, m0, mc, Bn);
    gemm_launch(s, G_DX_FIRST, 2, m0, mc, Bn);
    for (int l = 2; l <= 4; ++l) gemm_launch(s, G_DW, l, m0, mc, Bn);
    k_small_grads<<<dim3(HID, mc), 256, 0, st>>>(s->acts, m0, Bn);
    k_grad_reduce<<<dim3((NPAR + 255) / 256, mc), 256, 0, st>>>(s->acts, m0, Bn / HEAD_ROWS);
    SURR_HIP(hipGetLastError());
    return HWT_OK;
}

float lr_at(const hwt_surrogate *s, int64_t step, int64_t total)
{
    const hwt_surr_config &c = s->cfg;
    const double warm = std::max<double>(1.0, c.warmup_frac * (double)total);
    if (step < warm) return (float)(c.lr_max * (step + 1) / warm);
    const double f = std::min(1.0, (step - warm) / std::max(1.0, (double)total - warm));
    return (float)(c.lr_min + 0.5 * (c.lr_max - c.lr_min) * (1.0 + std::cos(M_PI * f)));
}

int adam_launch(hwt_surrogate *s, float lr)
{
    const hwt_surr_config &c = s->cfg;
    const double t = (double)(s->step + 1);
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)c.beta1, t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)c.beta2, t)));
    k_adam<<<dim3((NPAR + 255) / 256, s->members), 256, 0, s->stream>>>(s->net, s->acts.g, s->m1, s->m2, lr, c.beta1,
                                                                         c.beta2, c.eps, c1, c2);
    SURR_HIP(hipGetLastError());
    ++s->step;
    return HWT_OK;
}

void init_params(hwt_surrogate *s, std::vector<float> &P)
{
    // PyTorch-style uniform init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    P.assign((size_t)s->members * PSTRIDE, 0.0f);
    for (int m = 0; m < s->members; ++m) {
        std::mt19937_64 g(s->cfg.seed * 1000003ull + 7919ull * (uint64_t)m + 1ull);
        auto fill = [&](int off, int count, int fan_in) {
            std::uniform_real_distribution<float> U(-1.0f / std::sqrt((float)fan_in), 1.0f / std::sqrt((float)fan_in));
            for (int i = 0; i < count; ++i) P[(size_t)m * PSTRIDE + off + i] = U(g);
        };
        fill(HWT_SURR_W1, HID * NIN, NIN);
        fill(HWT_SURR_B1, HID, NIN);
        for (int l = 2; l <= 4; ++l) {
            fill(woff(l), LAYER_W, HID);
            fill(boff(l), HID, HID);
        }
        fill(HWT_SURR_W5, HID, HID);
        fill(HWT_SURR_B5, 1, HID);
    }
}

struct DeviceGuardS {
    int prev = -1;
    explicit DeviceGuardS(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuardS()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

} // namespace

extern "C" {

void hwt_surr_config_init(hwt_surr_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->members = 4;
    c->batch = 8192;
    c->lr_max = 2e-3f;
    c->lr_min = 2e-5f;
    c->warmup_frac = 0.02f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1e-8f;
    c->spread_threshold = 1e-3f;
    c->seed = 1;
}

void hwt_surr_destroy(hwt_surrogate *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->mem) (void)hipFree(s->mem);
    if (s->h_count) (void)hipHostFree(s->h_count);
    if (s->h_loss) (void)hipHostFree(s->h_loss);
    for (hipEvent_t e : {s->e0, s->e1, s->e2})
        if (e) (void)hipEventDestroy(e);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int hwt_surr_create(hwt_surrogate **out, int32_t train_capacity, int32_t infer_capacity, int32_t yield_capacity,
                    const hwt_surr_config *cfg)
{
    if (!out || train_capacity < 0 || infer_capacity < 0 || yield_capacity < 1)
        return fail(HWT_EINVAL, "hwt_surr_create: bad arguments");
    *out = nullptr;
    hwt_surrogate *s = new (std::nothrow) hwt_surrogate();
    if (!s) return fail(HWT_ENOMEM, "hwt_surr_create: host allocation");
    if (cfg) s->cfg = *cfg;
    else hwt_surr_config_init(&s->cfg);
    const hwt_surr_config &c = s->cfg;
    if (c.members < 1 || c.members > HWT_SURR_MAX_MEMBERS || c.batch < GEMM_ROWS || c.batch % GEMM_ROWS != 0 ||
        c.batch > (1 << 20) || !(c.lr_max > 0.0f) || !(c.beta1 >= 0.0f && c.beta1 < 1.0f) ||
        !(c.beta2 >= 0.0f && c.beta2 < 1.0f) || !(c.eps > 0.0f) || !(c.spread_threshold >= 0.0f)) {
        delete s;
        return fail(HWT_EINVAL, "hwt_surr_create: bad configuration");
    }
    s->members = c.members;
    s->B = c.batch;
    s->train_cap = train_capacity;
    s->infer_cap = infer_capacity;
    s->ycap = yield_capacity;
    s->rng.seed(c.seed * 2654435761ull + 12345ull);
    auto bail = [&](int code, const std::string &msg) {
        hwt_surr_destroy(s);
        return fail(code, msg);
    };
    if (hipGetDevice(&s->device) != hipSuccess) return bail(HWT_EHIP, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(HWT_EHIP, "hipStreamCreate failed");
    for (hipEvent_t *e : {&s->e0, &s->e1, &s->e2})
        if (hipEventCreate(e) != hipSuccess) return bail(HWT_EHIP, "hipEventCreate failed");
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) == hipSuccess && khz > 0)
            s->wall_khz = khz;
        (void)hipGetLastError();
    }
    const size_t M = s->members, B = s->B, T = std::max(train_capacity, 1), I = std::max(infer_capacity, 1);
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t oP = carve(M * PSTRIDE * 4), oWb = carve(M * 3 * LAYER_W * 2), oWbT = carve(M * 3 * LAYER_W * 2);
    const size_t om1 = carve(M * PSTRIDE * 4), om2 = carve(M * PSTRIDE * 4);
    const size_t oxb = carve(M * B * NIN * 4), otb = carve(M * B * 4), oz = carve(M * 4 * B * HID * 4);
    const size_t oa = carve(M * 3 * B * HID * 2), oaT = carve(M * 3 * B * HID * 2), oa4 = carve(M * B * HID * 4);
    const size_t odz = carve(M * 3 * B * HID * 2), odzT = carve(M * 3 * B * HID * 2), odz1 = carve(M * HID * B * 4);
    const size_t oy = carve(M * B * 4), ogw = carve(M * 3 * DW_SPLIT * LAYER_W * 4), ogs = carve(M * PSTRIDE * 4);
    const size_t ohp = carve(M * (B / HEAD_ROWS) * (HID + 1) * 4), oloss = carve(M * 8), og = carve(M * PSTRIDE * 4);
    const size_t oX = carve(T * NIN * 4), otg = carve(T * 4), operm = carve(M * T * 4), orows = carve(B * 4);
    const size_t obad = carve(4), ofbi = carve(I * 4), ofbc = carve(16 * 4), ofbt = carve(I * 8);
    const size_t oinr = carve(I * 4);
    const size_t ocs = carve(I * 8), ocm = carve(I * 8), ocl = carve(I * 8), oca = carve(I * 8), ocg = carve(I * 8);
    const size_t ocu = carve(I * 2), ocst = carve(I * 2), octy = carve(I);
    s->bytes = off;
    if (hipMalloc(&s->mem, s->bytes) != hipSuccess) return bail(HWT_ENOMEM, "hipMalloc surrogate workspace failed");
    char *b = (char *)s->mem;
    s->net.P = (float *)(b + oP);
    s->net.Wb = (uint16_t *)(b + oWb);
    s->net.WbT = (uint16_t *)(b + oWbT);
    s->m1 = (float *)(b + om1);
    s->m2 = (float *)(b + om2);
    Acts &a = s->acts;
    a.B = (int)B;
    a.xb = (float *)(b + oxb);
    a.tb = (float *)(b + otb);
    a.z = (float *)(b + oz);
    a.a = (uint16_t *)(b + oa);
    a.aT = (uint16_t *)(b + oaT);
    a.a4 = (float *)(b + oa4);
    a.dz = (uint16_t *)(b + odz);
    a.dzT = (uint16_t *)(b + odzT);
    a.dz1T = (float *)(b + odz1);
    a.y = (float *)(b + oy);
    a.gw = (float *)(b + ogw);
    a.gs = (float *)(b + ogs);
    a.hp = (float *)(b + ohp);
    a.loss = (double *)(b + oloss);
    a.g = (float *)(b + og);
    s->X = (float *)(b + oX);
    s->target = (float *)(b + otg);
    s->perm = (int32_t *)(b + operm);
    s->rows = (int32_t *)(b + orows);
    s->bad = (int32_t *)(b + obad);
    s->fb_idx = (int32_t *)(b + ofbi);
    s->inf_rows = (int32_t *)(b + oinr);
    s->fb_count = (int32_t *)(b + ofbc);
    s->fb_tmp = (double *)(b + ofbt);
    s->comp = Compact{(double *)(b + ocs), (double *)(b + ocm), (double *)(b + ocl), (double *)(b + oca),
                      (double *)(b + ocg), (uint16_t *)(b + ocu), (uint16_t *)(b + ocst), (int8_t *)(b + octy)};
    if (hipHostMalloc((void **)&s->h_count, 16 * sizeof(int32_t), 0) != hipSuccess ||
        hipHostMalloc((void **)&s->h_loss, M * sizeof(double), 0) != hipSuccess)
        return bail(HWT_ENOMEM, "hipHostMalloc failed");
    // zero everything once (gradient partials of unused rows stay defined)
    if (hipMemsetAsync(s->mem, 0, s->bytes, s->stream) != hipSuccess) return bail(HWT_EHIP, "hipMemsetAsync failed");
    std::vector<float> P;
    init_params(s, P);
    if (hipMemcpyAsync(s->net.P, P.data(), P.size() * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return bail(HWT_EHIP, "hipMemcpy parameters failed");
    k_refresh_bf16<<<dim3((3 * LAYER_W + 255) / 256, s->members), 256, 0, s->stream>>>(s->net, s->members);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(HWT_EHIP, "initialisation failed");
    *out = s;
    return HWT_OK;
}

int hwt_surr_set_training_data(hwt_surrogate *s, const hwt_options_f64 *o, const hwt_yield_f64 *y, const double *prices)
{
    if (!s || !valid_opts(o) || !valid_yield(y) || (o->n > 0 && !prices))
        return fail(HWT_EINVAL, "hwt_surr_set_training_data: null or negative-size argument");
    if (o->n > s->train_cap) return fail(HWT_ENOMEM, "hwt_surr_set_training_data: more samples than the capacity");
    if (o->n < s->B) return fail(HWT_EINVAL, "hwt_surr_set_training_data: fewer samples than one batch");
    DeviceGuardS dg(s->device);
    k_zero_i32<<<1, 32, 0, s->stream>>>(s->bad, 1);
    k_train_inputs<<<(o->n + 255) / 256, 256, 0, s->stream>>>(sopts(*o), SYield{y->n, y->rate, y->day}, prices, s->X,
                                                              s->target, s->bad);
    SURR_HIP(hipMemcpyAsync(s->h_count, s->bad, 4, hipMemcpyDeviceToHost, s->stream));
    SURR_HIP(hipStreamSynchronize(s->stream));
    if (s->h_count[0] > 0)
        return fail(HWT_EINVAL, std::to_string(s->h_count[0]) + " training sample(s) outside the training box");
    s->ntrain = o->n;
    return HWT_OK;
}

int hwt_surr_train(hwt_surrogate *s, int32_t epochs, int32_t total_epochs, hwt_surr_stats *stats)
{
    if (!s || epochs < 0 || total_epochs < s->epochs_done + epochs)
        return fail(HWT_EINVAL, "hwt_surr_train: bad arguments");
    if (s->ntrain < s->B) return fail(HWT_EINVAL, "hwt_surr_train: no training data");
    DeviceGuardS dg(s->device);
    const int spe = s->ntrain / s->B;  // steps per epoch (the remainder rotates through the shuffles)
    const int64_t total = (int64_t)total_epochs * spe;
    std::vector<int32_t> perm((size_t)s->members * s->ntrain);
    SURR_HIP(hipEventRecord(s->e0, s->stream));
    for (int ep = 0; ep < epochs; ++ep) {
        for (int m = 0; m < s->members; ++m) {
            int32_t *p = perm.data() + (size_t)m * s->ntrain;
            for (int i = 0; i < s->ntrain; ++i) p[i] = i;
            for (int i = s->ntrain - 1; i > 0; --i) {
                const int j = (int)(s->rng() % (uint64_t)(i + 1));
                std::swap(p[i], p[j]);
            }
        }
        SURR_HIP(hipStreamSynchronize(s->stream));  // the previous epoch's reads of perm are done
        SURR_HIP(hipMemcpyAsync(s->perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, s->stream));
        k_zero_f64<<<1, 64, 0, s->stream>>>(s->acts.loss, s->members);
        for (int it = 0; it < spe; ++it) {
            k_gather<<<dim3(s->B / 256 + (s->B % 256 != 0), s->members), 256, 0, s->stream>>>(
                s->X, s->target, s->perm, s->ntrain, (int64_t)it * s->B, s->B, s->B, s->acts.xb, s->acts.tb, 0);
            const int rc = forward_backward(s, 0, s->members, s->B);
            if (rc) return rc;
            const int arc = adam_launch(s, lr_at(s, s->step, total));
            if (arc) return arc;
        }
        ++s->epochs_done;
    }
    SURR_HIP(hipEventRecord(s->e1, s->stream));
    SURR_HIP(hipMemcpyAsync(s->h_loss, s->acts.loss, s->members * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SURR_HIP(hipStreamSynchronize(s->stream));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->e0, s->e1) == hipSuccess) stats->ms = ms;
        double l = 0.0;
        for (int m = 0; m < s->members; ++m) l += s->h_loss[m];
        stats->loss = epochs > 0 ? l / ((double)s->members * spe * s->B) : 0.0;
        stats->steps = s->step;
    }
    return HWT_OK;
}

int hwt_surr_eval_rows(hwt_surrogate *s, int32_t member, const int32_t *host_rows, int32_t n, float *host_out,
                       float *host_loss, float *host_grad)
{
    if (!s || member < 0 || member >= s->members || !host_rows || n < GEMM_ROWS || n % GEMM_ROWS != 0 || n > s->B)
        return fail(HWT_EINVAL, "hwt_surr_eval_rows: bad arguments");
    for (int i = 0; i < n; ++i)
        if (host_rows[i] < 0 || host_rows[i] >= s->ntrain) return fail(HWT_EINVAL, "hwt_surr_eval_rows: row out of range");
    DeviceGuardS dg(s->device);
    // Rows go into the member's slot of the batch buffers; the kernels use the
    // batch pitch s->B and the actual row count n.
    SURR_HIP(hipMemcpyAsync(s->rows, host_rows, (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
    k_gather<<<dim3((n + 255) / 256, 1), 256, 0, s->stream>>>(s->X, s->target, s->rows, 0, 0, n, s->B, s->acts.xb,
                                                              s->acts.tb, member);
    k_zero_f64<<<1, 64, 0, s->stream>>>(s->acts.loss, s->members);
    const int rc = forward_backward(s, member, 1, n);
    if (rc) return rc;
    SURR_HIP(hipMemcpyAsync(s->h_loss, s->acts.loss, s->members * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    if (host_out)
        SURR_HIP(hipMemcpyAsync(host_out, s->acts.y + (size_t)member * s->B, (size_t)n * 4, hipMemcpyDeviceToHost,
                                s->stream));
    if (host_grad)
        SURR_HIP(hipMemcpyAsync(host_grad, s->acts.g + (size_t)member * PSTRIDE, (size_t)NPAR * 4,
                                hipMemcpyDeviceToHost, s->stream));
    SURR_HIP(hipStreamSynchronize(s->stream));
    if (host_loss) *host_loss = (float)(s->h_loss[member] / n);
    return HWT_OK;
}

int hwt_surr_get_inputs(hwt_surrogate *s, const int32_t *host_rows, int32_t n, float *host_x, float *host_target)
{
    if (!s || !host_rows || n < 0) return fail(HWT_EINVAL, "hwt_surr_get_inputs: bad arguments");
    DeviceGuardS dg(s->device);
    std::vector<float> X((size_t)s->ntrain * NIN), t(s->ntrain);
    SURR_HIP(hipMemcpy(X.data(), s->X, X.size() * 4, hipMemcpyDeviceToHost));
    SURR_HIP(hipMemcpy(t.data(), s->target, t.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        const int r = host_rows[i];
        if (r < 0 || r >= s->ntrain) return fail(HWT_EINVAL, "hwt_surr_get_inputs: row out of range");
        if (host_x) std::memcpy(host_x + (size_t)i * NIN, X.data() + (size_t)r * NIN, NIN * 4);
        if (host_target) host_target[i] = t[r];
    }
    return HWT_OK;
}

int hwt_surr_adam_step(hwt_surrogate *s, const float *host_grad, float lr)
{
    if (!s || !host_grad || !(lr >= 0.0f)) return fail(HWT_EINVAL, "hwt_surr_adam_step: bad arguments");
    DeviceGuardS dg(s->device);
    for (int m = 0; m < s->members; ++m)
        SURR_HIP(hipMemcpyAsync(s->acts.g + (size_t)m * PSTRIDE, host_grad + (size_t)m * NPAR, (size_t)NPAR * 4,
                                hipMemcpyHostToDevice, s->stream));
    const int rc = adam_launch(s, lr);
    if (rc) return rc;
    SURR_HIP(hipStreamSynchronize(s->stream));
    return HWT_OK;
}

int hwt_surr_get_params(hwt_surrogate *s, float *host)
{
    if (!s || !host) return fail(HWT_EINVAL, "hwt_surr_get_params: bad arguments");
    DeviceGuardS dg(s->device);
    SURR_HIP(hipStreamSynchronize(s->stream));
    for (int m = 0; m < s->members; ++m)
        SURR_HIP(hipMemcpy(host + (size_t)m * NPAR, s->net.P + (size_t)m * PSTRIDE, (size_t)NPAR * 4,
                           hipMemcpyDeviceToHost));
    return HWT_OK;
}

int hwt_surr_set_params(hwt_surrogate *s, const float *host)
{
    if (!s || !host) return fail(HWT_EINVAL, "hwt_surr_set_params: bad arguments");
    DeviceGuardS dg(s->device);
    SURR_HIP(hipStreamSynchronize(s->stream));
    for (int m = 0; m < s->members; ++m)
        SURR_HIP(hipMemcpy(s->net.P + (size_t)m * PSTRIDE, host + (size_t)m * NPAR, (size_t)NPAR * 4,
                           hipMemcpyHostToDevice));
    k_refresh_bf16<<<dim3((3 * LAYER_W + 255) / 256, s->members), 256, 0, s->stream>>>(s->net, s->members);
    SURR_HIP(hipGetLastError());
    SURR_HIP(hipStreamSynchronize(s->stream));
    return HWT_OK;
}

int hwt_surr_predict(hwt_surrogate *s, const hwt_options_f64 *o, const hwt_yield_f64 *y, double *mean, double *spread,
                     int8_t *in_box)
{
    if (!s || !valid_opts(o) || !valid_yield(y) || (o->n > 0 && (!mean || !spread || !in_box)))
        return fail(HWT_EINVAL, "hwt_surr_predict: null or negative-size argument");
    if (o->n == 0) return HWT_OK;
    DeviceGuardS dg(s->device);
    InferArgs g{};
    g.o = sopts(*o);
    g.y = SYield{y->n, y->rate, y->day};
    g.net = s->net;
    g.members = s->members;
    g.threshold = s->cfg.spread_threshold;
    g.mean = mean;
    g.spread = spread;
    g.in_box = in_box;
    k_infer<<<(o->n + INF_ROWS - 1) / INF_ROWS, 256, 0, s->stream>>>(g);
    SURR_HIP(hipGetLastError());
    SURR_HIP(hipStreamSynchronize(s->stream));
    return HWT_OK;
}

// One pricing pass, enqueued on st (for capture): the box split, the ensemble
// on the in-box rows, then the whole fallback list on the FP64 tree in one
// pass and its prices back to input order.  e1 marks the end of inference.
static int enqueue_pass(hwt_surrogate *s, hwt_plan *tree, const hwt_options_f64 *o, const hwt_yield_f64 *y,
                        double *out, hipStream_t st)
{
    uint64_t *stamps = reinterpret_cast<uint64_t *>(s->fb_count + 8);
    k_stamp<<<1, 1, 0, st>>>(stamps, 0);
    k_zero_i32<<<1, 32, 0, st>>>(s->fb_count, 8);
    if (o->n == 0) {
        k_stamp<<<1, 1, 0, st>>>(stamps, 1);
        k_stamp<<<1, 1, 0, st>>>(stamps, 2);
        SURR_HIP(hipGetLastError());
        return HWT_OK;
    }
    const int blocks = (o->n + 255) / 256;
    // split: out-of-box options to the fallback list, in-box ones to the
    // inference rows (running the tree on the out-of-box part beside the
    // inference was measured slower: both compete for the same CUs)
    k_box<<<(o->n + BOX_CHUNK - 1) / BOX_CHUNK, 256, 0, st>>>(sopts(*o), s->fb_idx, s->fb_count, s->inf_rows,
                                                            s->fb_count + 4);
    // the ensemble on the in-box rows; options it is unsure of join the list
    InferArgs g{};
    g.o = sopts(*o);
    g.y = SYield{y->n, y->rate, y->day};
    g.net = s->net;
    g.members = s->members;
    g.threshold = s->cfg.spread_threshold;
    g.out = out;
    g.fb_idx = s->fb_idx;
    g.fb_count = s->fb_count;
    g.rows = s->inf_rows;
    g.row_count = s->fb_count + 4;
    k_infer<<<(o->n + INF_ROWS - 1) / INF_ROWS, 256, 0, st>>>(g);
    k_stamp<<<1, 1, 0, st>>>(stamps, 1);
    SURR_HIP(hipGetLastError());
    // the whole list on the FP64 tree in one pass
    k_fb_gather<<<blocks, 256, 0, st>>>(sopts(*o), s->fb_idx, s->fb_count, s->comp);
    const hwt_options_f64 co{o->n,          s->comp.strike, s->comp.maturity, s->comp.length, s->comp.tu,
                             s->comp.steps, s->comp.a,      s->comp.sigma,    s->comp.type};
    const int rc = hwt_internal_plan_enqueue_counted(tree, &co, s->fb_count, y, s->fb_tmp, st);
    if (rc) return rc;
    k_fb_scatter<<<blocks, 256, 0, st>>>(s->fb_idx, s->fb_count, s->fb_tmp, out);
    k_stamp<<<1, 1, 0, st>>>(stamps, 2);
    SURR_HIP(hipGetLastError());
    return HWT_OK;
}

int hwt_surr_price(hwt_surrogate *s, hwt_plan *tree, const hwt_options_f64 *o, const hwt_yield_f64 *y, double *out,
                   hwt_surr_stats *stats)
{
    if (!s || !tree || !valid_opts(o) || !valid_yield(y) || (o->n > 0 && !out))
        return fail(HWT_EINVAL, "hwt_surr_price: null or negative-size argument");
    if (o->n > s->infer_cap) return fail(HWT_ENOMEM, "hwt_surr_price: more options than the capacity");
    if (y->n > s->ycap) return fail(HWT_ENOMEM, "hwt_surr_price: yield curve longer than the capacity");
    // one member has no spread (always 0): a finite threshold would accept
    // every in-box price with no error estimate at all
    if (s->members < 2 && std::isfinite(s->cfg.spread_threshold))
        return fail(HWT_EINVAL, "hwt_surr_price: a one-member ensemble has no spread to gate on; "
                                "set spread_threshold to +inf to take every in-box price from the network");
    DeviceGuardS dg(s->device);
    hipStream_t st = s->stream;
    // The whole pass -- box split, ensemble inference and gate, and the FP64
    // tree on the fallback list -- is one graph, captured on the first call
    // and again when its arguments change, then replayed.
    hwt_surrogate::PassKey key;
    std::memset(&key, 0, sizeof(key));  // padding too: the key is compared bytewise
    key.o = *o;
    key.y = *y;
    key.out = out;
    key.tree = tree;
    key.tree_serial = hwt_internal_plan_serial(tree);
    key.threshold = s->cfg.spread_threshold;
    key.members = s->members;
    if (!s->exec || std::memcmp(&key, &s->key, sizeof(key)) != 0) {
        if (s->exec) (void)hipGraphExecDestroy(s->exec);
        if (s->graph) (void)hipGraphDestroy(s->graph);
        s->exec = nullptr;
        s->graph = nullptr;
        (void)hipGetLastError();
        SURR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_pass(s, tree, o, y, out, st);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(HWT_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        s->graph = g;
        SURR_HIP(hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0));
        s->key = key;
    }
    SURR_HIP(hipEventRecord(s->e0, st));
    SURR_HIP(hipGraphLaunch(s->exec, st));
    if (o->n > 0) hwt_internal_plan_replayed_on(tree, st);
    SURR_HIP(hipEventRecord(s->e2, st));
    SURR_HIP(hipMemcpyAsync(s->h_count, s->fb_count, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SURR_HIP(hipStreamSynchronize(st));
    const int trc = o->n > 0 ? hwt_plan_sync(tree, nullptr) : HWT_OK;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->e0, s->e2) == hipSuccess) stats->ms = ms;
        uint64_t t[3];
        std::memcpy(t, s->h_count + 8, sizeof(t));
        stats->surrogate_ms = (double)(t[1] - t[0]) / s->wall_khz;
        stats->tree_ms = (double)(t[2] - t[1]) / s->wall_khz;
        stats->fallback = s->h_count[0];
        stats->out_of_box = s->h_count[1];
        stats->steps = s->step;
    }
    return trc;
}

int hwt_surr_save(hwt_surrogate *s, const char *path)
{
    if (!s || !path) return fail(HWT_EINVAL, "hwt_surr_save: bad arguments");
    DeviceGuardS dg(s->device);
    const size_t nm = (size_t)s->members * PSTRIDE;
    std::vector<float> P(nm), m1(nm), m2(nm);
    SURR_HIP(hipStreamSynchronize(s->stream));
    SURR_HIP(hipMemcpy(P.data(), s->net.P, nm * 4, hipMemcpyDeviceToHost));
    SURR_HIP(hipMemcpy(m1.data(), s->m1, nm * 4, hipMemcpyDeviceToHost));
    SURR_HIP(hipMemcpy(m2.data(), s->m2, nm * 4, hipMemcpyDeviceToHost));
    std::ostringstream rs;
    rs << s->rng;
    const std::string rng = rs.str();
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(HWT_EINVAL, std::string("hwt_surr_save: cannot open ") + path);
    const char magic[8] = {'H', 'W', 'T', 'S', 'U', 'R', 'R', '1'};
    const int32_t hdr[4] = {s->members, NPAR, PSTRIDE, s->epochs_done};
    const int64_t step = s->step, rl = (int64_t)rng.size();
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 && std::fwrite(&step, 8, 1, f) == 1 &&
              std::fwrite(&rl, 8, 1, f) == 1 && std::fwrite(rng.data(), 1, rng.size(), f) == rng.size() &&
              std::fwrite(P.data(), 4, nm, f) == nm && std::fwrite(m1.data(), 4, nm, f) == nm &&
              std::fwrite(m2.data(), 4, nm, f) == nm;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? HWT_OK : fail(HWT_EINVAL, std::string("hwt_surr_save: write failed: ") + path);
}

int hwt_surr_load(hwt_surrogate *s, const char *path)
{
    if (!s || !path) return fail(HWT_EINVAL, "hwt_surr_load: bad arguments");
    DeviceGuardS dg(s->device);
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(HWT_EINVAL, std::string("hwt_surr_load: cannot open ") + path);
    char magic[8];
    int32_t hdr[4];
    int64_t step = 0, rl = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "HWTSURR1", 8) == 0 &&
              std::fread(hdr, 4, 4, f) == 4 && hdr[0] == s->members && hdr[1] == NPAR && hdr[2] == PSTRIDE &&
              std::fread(&step, 8, 1, f) == 1 && std::fread(&rl, 8, 1, f) == 1 && rl > 0 && rl < (1 << 20);
    std::string rng;
    const size_t nm = (size_t)s->members * PSTRIDE;
    std::vector<float> P(nm), m1(nm), m2(nm);
    if (ok) {
        rng.resize((size_t)rl);
        ok = std::fread(&rng[0], 1, rng.size(), f) == rng.size() && std::fread(P.data(), 4, nm, f) == nm &&
             std::fread(m1.data(), 4, nm, f) == nm && std::fread(m2.data(), 4, nm, f) == nm;
    }
    std::fclose(f);
    if (!ok) return fail(HWT_EINVAL, std::string("hwt_surr_load: not a compatible checkpoint: ") + path);
    std::istringstream is(rng);
    is >> s->rng;
    s->step = step;
    s->epochs_done = hdr[3];
    SURR_HIP(hipStreamSynchronize(s->stream));
    SURR_HIP(hipMemcpy(s->net.P, P.data(), nm * 4, hipMemcpyHostToDevice));
    SURR_HIP(hipMemcpy(s->m1, m1.data(), nm * 4, hipMemcpyHostToDevice));
    SURR_HIP(hipMemcpy(s->m2, m2.data(), nm * 4, hipMemcpyHostToDevice));
    k_refresh_bf16<<<dim3((3 * LAYER_W + 255) / 256, s->members), 256, 0, s->stream>>>(s->net, s->members);
    SURR_HIP(hipGetLastError());
    SURR_HIP(hipStreamSynchronize(s->stream));
    return HWT_OK;
}

} // extern "C"
