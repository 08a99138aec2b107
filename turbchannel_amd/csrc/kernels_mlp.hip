// Learned subgrid closure on gfx950: forward, backward and Adam for the
// 10 -> 64 -> 64 -> 64 -> 6 GELU network with v_mfma_f32_16x16x32_bf16.
//
// One wave owns 64 points. Layers run as D[f_out][p] = sum_k W[f_out][k] act[p][k]:
// the A operand is 8 consecutive bfloat16 weights of one row (16 B from global,
// L1/L2-resident), the B operand 8 consecutive activations of one point (16 B
// from LDS). An accumulator lane holds 4 consecutive output features of one
// point, which is exactly an 8 B run of the next layer's [point][feature] LDS
// row, so layers chain through LDS without shuffles. Weight gradients sum over
// the 64 points (K = points), gathering their operands down LDS columns.
//
// Rounding model (the CPU restatement follows it): inputs, hidden activations,
// GELU derivatives and back-propagated gradients are bfloat16 (nearest even);
// weights are bfloat16 copies of FP32 master weights; products accumulate in
// FP32; biases, outputs and the loss are FP32.
#include <algorithm>
#include <cstring>

#include "kernels.h"
#include "mlp.h"

namespace tc {

namespace {

typedef unsigned short u16;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kLda = 72;          // LDS row stride of a [64 points][<=64] bfloat16 tile (16 B pad)
constexpr int kTile = 64 * kLda;  // elements per tile

// FP32 -> bfloat16, nearest even: gfx950's v_cvt_pk_bf16_f32 (two values per
// instruction when converted in pairs).
__device__ __forceinline__ u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }
__device__ __forceinline__ unsigned f2bf2(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

constexpr float kGc = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGa = 0.044715f;
constexpr float kLog2e = 1.4426950408889634f;
// GELU, tanh form: 0.5 z (1 + tanh u) = z s with s = 1 / (1 + exp(-2u)),
// u = sqrt(2/pi) (z + 0.044715 z^3), evaluated with the hardware exp2 and
// reciprocal (no libm call) as s = 1 / (1 + 2^(z (A + B z^2))), A = -2 log2(e)
// sqrt(2/pi), B = 0.044715 A; derivative s + s (1 - s) z (2 sqrt(2/pi) +
// 6 sqrt(2/pi) 0.044715 z^2).
constexpr float kGeA = -2.f * kLog2e * kGc, kGeB = kGeA * kGa;
constexpr float kGeD1 = 2.f * kGc, kGeD3 = 6.f * kGc * kGa;
__device__ __forceinline__ float gelu_s(float z, float z2) {
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z * __builtin_fmaf(kGeB, z2, kGeA)));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z, z * z); }
__device__ __forceinline__ float gelu(float z, float& der) {
    const float z2 = z * z, s = gelu_s(z, z2);
    der = __builtin_fmaf(s * (1.f - s), z * __builtin_fmaf(kGeD3, z2, kGeD1), s);
    return z * s;
}

__device__ __forceinline__ int lane() { return threadIdx.x & 63; }

// acc[mt][nt] = W[16 mt + .][:] . act[16 nt + .][:], K = 32 KS.
// W: global [MT*16][32 KS] bfloat16; act: LDS tile of this wave.
template <int MT, int KS>
__device__ __forceinline__ void mfma_layer(const u16* __restrict__ W, const u16* act, f32x4 (&acc)[MT][4]) {
    const int r = lane() & 15, g = lane() >> 4;
    u16x8 b[4][KS];
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) b[nt][ks] = *(const u16x8*)(act + (16 * nt + r) * kLda + 32 * ks + 8 * g);
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        u16x8 a[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) a[ks] = *(const u16x8*)(W + (16 * mt + r) * (32 * KS) + 32 * ks + 8 * g);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ks++) c = mfma(a[ks], b[nt][ks], c);
            acc[mt][nt] = c;
        }
    }
}

__device__ __forceinline__ void store4(u16* p, const u16 (&v)[4]) {
    uint2 q;
    q.x = (unsigned)v[0] | ((unsigned)v[1] << 16);
    q.y = (unsigned)v[2] | ((unsigned)v[3] << 16);
    *(uint2*)p = q;
}
__device__ __forceinline__ void load4(const u16* p, u16 (&v)[4]) {
    const uint2 q = *(const uint2*)p;
    v[0] = (u16)(q.x & 0xFFFF);
    v[1] = (u16)(q.x >> 16);
    v[2] = (u16)(q.y & 0xFFFF);
    v[3] = (u16)(q.y >> 16);
}

// Hidden-layer epilogue: act = bf16(gelu(acc + b)), der = bf16(gelu'(acc + b));
// bias == nullptr: the accumulators already hold the bias.
template <bool DER>
__device__ __forceinline__ void hidden_out(const f32x4 (&acc)[4][4], const float* __restrict__ bias, u16* act,
                                           u16* der) {
    const int r = lane() & 15, g = lane() >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const int f0 = 16 * mt + 4 * g;
        float bb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) bb[i] = bias ? bias[f0 + i] : 0.f;  // folded away without bias
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            float a[4], d[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = bias ? acc[mt][nt][i] + bb[i] : acc[mt][nt][i];
                if (DER)
                    a[i] = gelu(z, d[i]);
                else
                    a[i] = gelu(z);
            }
            *(uint2*)(act + (16 * nt + r) * kLda + f0) = make_uint2(f2bf2(a[0], a[1]), f2bf2(a[2], a[3]));
            if (DER) *(uint2*)(der + (16 * nt + r) * kLda + f0) = make_uint2(f2bf2(d[0], d[1]), f2bf2(d[2], d[3]));
        }
    }
}

// Inputs of 64 points into a tile: lane = point, columns 0..9 then zeros to 31.
__device__ __forceinline__ void stage_inputs(u16* xs, const float (&x)[kMlpIn]) {
    u16 v[32];
#pragma unroll
    for (int k = 0; k < 32; k++) v[k] = k < kMlpIn ? f2bf(x[k]) : (u16)0;
    u16* row = xs + lane() * kLda;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        u16x8 w;
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = v[8 * q + k];
        *(u16x8*)(row + 8 * q) = w;
    }
}

// Inference keeps the forward weights (A operands, 88 VGPRs) in registers for
// the life of a persistent wave, which walks over 64-point chunks: per chunk only
// the inputs and activations move.
struct FwdRegs {
    u16x8 a1[4][1], a2[4][2], a3[4][2], a4[1][2];
};

__device__ __forceinline__ void load_fwd_regs(const MlpDev& w, FwdRegs& R) {
    const int r = lane() & 15, g = lane() >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        R.a1[mt][0] = *(const u16x8*)(w.w + kW1p + (16 * mt + r) * 32 + 8 * g);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            R.a2[mt][ks] = *(const u16x8*)(w.w + kW2p + (16 * mt + r) * 64 + 32 * ks + 8 * g);
            R.a3[mt][ks] = *(const u16x8*)(w.w + kW3p + (16 * mt + r) * 64 + 32 * ks + 8 * g);
        }
    }
#pragma unroll
    for (int ks = 0; ks < 2; ks++) R.a4[0][ks] = *(const u16x8*)(w.w + kW4p + r * 64 + 32 * ks + 8 * g);
}

// The accumulators start from the layer's bias (rows 16 mt + 4 g + i).
template <int MT, int KS>
__device__ __forceinline__ void mfma_layer_r(const u16x8 (&a)[MT][KS], const u16* act, const float* __restrict__ bias,
                                             f32x4 (&acc)[MT][4]) {
    const int r = lane() & 15, g = lane() >> 4;
    u16x8 b[4][KS];
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) b[nt][ks] = *(const u16x8*)(act + (16 * nt + r) * kLda + 32 * ks + 8 * g);
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const float* bm = bias + 16 * mt + 4 * g;
            f32x4 c = {bm[0], bm[1], bm[2], bm[3]};
#pragma unroll
            for (int ks = 0; ks < KS; ks++) c = mfma(a[mt][ks], b[nt][ks], c);
            acc[mt][nt] = c;
        }
}

// Forward pass of one wave (a block of one wave) on tile `buf`; see forward_wave.
__device__ __forceinline__ void forward_wave_r(const FwdRegs& R, const float* __restrict__ bias, u16* buf,
                                               f32x4 (&y)[4]) {
    f32x4 acc[4][4];
    mfma_layer_r<4, 1>(R.a1, buf, bias, acc);
    __syncthreads();
    hidden_out<false>(acc, nullptr, buf, nullptr);
    __syncthreads();
    mfma_layer_r<4, 2>(R.a2, buf, bias + 64, acc);
    __syncthreads();
    hidden_out<false>(acc, nullptr, buf, nullptr);
    __syncthreads();
    mfma_layer_r<4, 2>(R.a3, buf, bias + 128, acc);
    __syncthreads();
    hidden_out<false>(acc, nullptr, buf, nullptr);
    __syncthreads();
    f32x4 o[1][4];
    mfma_layer_r<1, 2>(R.a4, buf, bias + 192, o);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) y[nt] = o[0][nt];
}

__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev w, const float* __restrict__ x, long long n,
                                                    float* __restrict__ y) {
    __shared__ __align__(16) u16 buf[kTile];
    FwdRegs R;
    load_fwd_regs(w, R);
    const long long nch = (n + 63) / 64;
    const int r = lane() & 15, g = lane() >> 4;
    for (long long c = blockIdx.x; c < nch; c += gridDim.x) {
        const long long p0 = c * 64, pi = p0 + lane();
        float xv[kMlpIn];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) xv[k] = pi < n ? x[pi * kMlpIn + k] : 0.f;
        stage_inputs(buf, xv);
        __syncthreads();
        f32x4 o[4];
        forward_wave_r(R, w.b, buf, o);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const long long p = p0 + 16 * nt + r;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int f = 4 * g + i;
                if (f < kMlpOut && p < n) y[p * kMlpOut + f] = o[nt][i];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Training: forward + backward of 64-point chunks, one wave per block; each
// block accumulates its chunks' gradient into its own partial record (fixed
// order, so the result does not depend on scheduling).
// ---------------------------------------------------------------------------
__device__ __forceinline__ u16x8 gather_col(const u16* t, int col, int ks, int g) {
    u16x8 v;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = t[(32 * ks + 8 * g + i) * kLda + col];
    return v;
}

// acc[mt][nt][i] = sum_p G[p][16 mt + 4 g + i] * A[p][16 nt + r]
template <int MT, int NT>
__device__ __forceinline__ void mfma_wgrad(const u16* G, const u16* A, f32x4 (&acc)[MT][NT]) {
    const int r = lane() & 15, g = lane() >> 4;
    u16x8 b[NT][2];
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int ks = 0; ks < 2; ks++) b[nt][ks] = gather_col(A, 16 * nt + r, ks, g);
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        u16x8 a[2];
#pragma unroll
        for (int ks = 0; ks < 2; ks++) a[ks] = gather_col(G, 16 * mt + r, ks, g);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            c = mfma(a[0], b[nt][0], c);
            c = mfma(a[1], b[nt][1], c);
            acc[mt][nt] = c;
        }
    }
}

__device__ __forceinline__ void rec_add(float* rec, int j, float v, bool first) { rec[j] = first ? v : rec[j] + v; }

// dW[out][in] (+)= sum over the chunk, for out < nout, in < nin; db[out] likewise.
template <int MT, int NT>
__device__ __forceinline__ void weight_grad(const u16* G, const u16* A, int nout, int nin, int offW, int offB,
                                            float* rec, bool first) {
    const int r = lane() & 15, g = lane() >> 4;
    f32x4 acc[MT][NT];
    mfma_wgrad<MT, NT>(G, A, acc);
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int o = 16 * mt + 4 * g + i, in = 16 * nt + r;
                if (o < nout && in < nin) rec_add(rec, offW + o * nin + in, acc[mt][nt][i], first);
            }
    if (lane() < nout) {
        float s = 0.f;
        for (int p = 0; p < 64; p++) s += bf2f(G[p * kLda + lane()]);
        rec_add(rec, offB + lane(), s, first);
    }
}

// G[p][in] = bf16((sum_out Wt[in][out] G[p][out]) * der[p][in]), in place.
template <int KS>
__device__ __forceinline__ void back_data(const u16* __restrict__ Wt, u16* G, const u16* der) {
    const int r = lane() & 15, g = lane() >> 4;
    f32x4 acc[4][4];
    mfma_layer<4, KS>(Wt, G, acc);
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const int off = (16 * nt + r) * kLda + 16 * mt + 4 * g;
            u16 d4[4], g4[4];
            load4(der + off, d4);
#pragma unroll
            for (int i = 0; i < 4; i++) g4[i] = f2bf(acc[mt][nt][i] * bf2f(d4[i]));
            store4(G + off, g4);
        }
    __syncthreads();
}

__global__ __launch_bounds__(64) void k_mlp_train(MlpDev w, const float* __restrict__ x, const float* __restrict__ t,
                                                  long long n, float gscale, float* __restrict__ partial) {
    // tiles: inputs, a1, a2, a3, d1, d2, d3, gradient
    __shared__ __align__(16) u16 sm[8 * kTile];
    u16 *xs = sm, *a1 = sm + kTile, *a2 = sm + 2 * kTile, *a3 = sm + 3 * kTile;
    u16 *d1 = sm + 4 * kTile, *d2 = sm + 5 * kTile, *d3 = sm + 6 * kTile, *G = sm + 7 * kTile;
    float* rec = partial + (size_t)blockIdx.x * kMlpRec;
    const int r = lane() & 15, g = lane() >> 4;
    const long long nch = (n + 63) / 64;
    float lsum = 0.f;
    bool first = true;
    for (long long c = blockIdx.x; c < nch; c += gridDim.x) {
        const long long p0 = c * 64;
        {
            const long long pi = p0 + lane();
            float xv[kMlpIn];
#pragma unroll
            for (int k = 0; k < kMlpIn; k++) xv[k] = pi < n ? x[pi * kMlpIn + k] : 0.f;
            stage_inputs(xs, xv);
            u16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
            *(u16x8*)(G + lane() * kLda + 16) = z;  // output-gradient columns 16..31 stay zero
            *(u16x8*)(G + lane() * kLda + 24) = z;
        }
        __syncthreads();
        f32x4 acc[4][4];
        mfma_layer<4, 1>(w.w + kW1p, xs, acc);
        hidden_out<true>(acc, w.b, a1, d1);
        __syncthreads();
        mfma_layer<4, 2>(w.w + kW2p, a1, acc);
        hidden_out<true>(acc, w.b + 64, a2, d2);
        __syncthreads();
        mfma_layer<4, 2>(w.w + kW3p, a2, acc);
        hidden_out<true>(acc, w.b + 128, a3, d3);
        __syncthreads();
        {
            f32x4 o[1][4];
            mfma_layer<1, 2>(w.w + kW4p, a3, o);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const long long p = p0 + 16 * nt + r;
                u16 g4[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int f = 4 * g + i;
                    float gy = 0.f;
                    if (f < kMlpOut && p < n) {
                        const float res = (o[0][nt][i] + w.b[192 + f]) - t[p * kMlpOut + f];
                        lsum += res * res;
                        gy = res * gscale;
                    }
                    g4[i] = f2bf(gy);
                }
                store4(G + (16 * nt + r) * kLda + 4 * g, g4);
            }
        }
        __syncthreads();
        weight_grad<1, 4>(G, a3, kMlpOut, kMlpHid, kOffW4, kOffB4, rec, first);
        __syncthreads();
        back_data<1>(w.w + kT4, G, d3);
        weight_grad<4, 4>(G, a2, kMlpHid, kMlpHid, kOffW3, kOffB3, rec, first);
        __syncthreads();
        back_data<2>(w.w + kT3, G, d2);
        weight_grad<4, 4>(G, a1, kMlpHid, kMlpHid, kOffW2, kOffB2, rec, first);
        __syncthreads();
        back_data<2>(w.w + kT2, G, d1);
        weight_grad<4, 1>(G, xs, kMlpHid, kMlpIn, kOffW1, kOffB1, rec, first);
        __syncthreads();
        first = false;
    }
    // loss of this block's chunks (sum of squared residuals)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane() == 0) rec[kMlpParams] = first ? 0.f : lsum;
    if (first) {  // no chunk: an empty record
        for (int j = lane(); j < kMlpParams; j += 64) rec[j] = 0.f;
    }
}

// Stage 1: tmp[s][j] = sum of records s, s + S, s + 2S, ... of column j (FP64).
__global__ void k_mlp_reduce_split(const float* __restrict__ partial, int nrec, double* __restrict__ tmp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, sp = blockIdx.y;
    if (j > kMlpParams) return;
    double s = 0.0;
    for (int k = sp; k < nrec; k += kMlpReduceSplit) s += partial[(size_t)k * kMlpRec + j];
    tmp[(size_t)sp * kMlpRec + j] = s;
}

// Where parameter j lives in the bfloat16 blocks (forward copy, transposed copy) or the bias block.
__device__ void param_slots(int j, int& wf, int& wt, int& bi) {
    wf = wt = bi = -1;
    if (j < kOffB1) {
        const int o = j / kMlpIn, i = j - o * kMlpIn;
        wf = kW1p + o * 32 + i;
    } else if (j < kOffW2) {
        bi = j - kOffB1;
    } else if (j < kOffB2) {
        const int q = j - kOffW2, o = q / 64, i = q - o * 64;
        wf = kW2p + o * 64 + i;
        wt = kT2 + i * 64 + o;
    } else if (j < kOffW3) {
        bi = 64 + j - kOffB2;
    } else if (j < kOffB3) {
        const int q = j - kOffW3, o = q / 64, i = q - o * 64;
        wf = kW3p + o * 64 + i;
        wt = kT3 + i * 64 + o;
    } else if (j < kOffW4) {
        bi = 128 + j - kOffB3;
    } else if (j < kOffB4) {
        const int q = j - kOffW4, o = q / 64, i = q - o * 64;
        wf = kW4p + o * 64 + i;
        wt = kT4 + i * 32 + o;
    } else {
        bi = 192 + j - kOffB4;
    }
}

// Stage 2: gradient, loss and (optionally) Adam with the device copies refreshed.
__global__ void k_mlp_reduce_adam(const double* __restrict__ tmp, float* __restrict__ grad, double* __restrict__ loss,
                                  AdamArgs a, float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                                  u16* __restrict__ wdev, float* __restrict__ bdev) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > kMlpParams) return;
    double s = 0.0;
    for (int k = 0; k < kMlpReduceSplit; k++) s += tmp[(size_t)k * kMlpRec + j];
    if (j == kMlpParams) {
        *loss = s;
        return;
    }
    const float gj = (float)s;
    grad[j] = gj;
    if (!a.apply) return;
    const float mj = a.b1 * m[j] + (1.f - a.b1) * gj;
    const float vj = a.b2 * v[j] + (1.f - a.b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float pj = params[j] - a.lr * (mj / a.bc1) / (sqrtf(vj / a.bc2) + a.eps);
    params[j] = pj;
    int wf, wt, bi;
    param_slots(j, wf, wt, bi);
    if (bi >= 0) bdev[bi] = pj;
    if (wf >= 0) wdev[wf] = f2bf(pj);
    if (wt >= 0) wdev[wt] = f2bf(pj);
}

// ---------------------------------------------------------------------------
// Closure stress at the nodes of a collocated field (the learned LES model)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// POST (kSgsPost*): what is kept of the network's stress at a node.
// kSgsPostClip: backscatter clipping, tau = 0 at a node where the stress would
// move energy from the sub-filter scales to the resolved ones (tau_ij S_ij > 0,
// S the resolved strain of the node's own gradient inputs).
// kSgsPostProject: the dissipative part along the deviatoric strain D, an eddy
// viscosity read off the network: tau' = (tau_kl D_kl / D_kl D_kl) D where
// tau_kl D_kl < 0, else 0 (tau' = -2 nu_t D with nu_t >= 0).
// Both need a node's six outputs, spread over the MFMA accumulator lanes, in the
// lane of the node's inputs: they meet through LDS.
template <typename T, int POST>
__global__ __launch_bounds__(64) void k_sgs_tau(Mesh<T> m, MlpDev w, const T* __restrict__ vel, T* __restrict__ tau,
                                                T u_tau2, T gscale, T yscale) {
    __shared__ __align__(16) u16 buf[kTile];
    __shared__ float tq[POST != kSgsPostNone ? 64 * 7 : 1];
    FwdRegs R;
    load_fwd_regs(w, R);
    const int nxy = m.nxy;
    const long long nch = ((long long)m.n3d + 63) / 64;
    const int r = lane() & 15, g = lane() >> 4;
    const T hx2 = T(0.5) * m.inv_hx, hy2 = T(0.5) * m.inv_hy;
    for (long long c = blockIdx.x; c < nch; c += gridDim.x) {
        const long long p0 = c * 64;
        const long long node = p0 + lane();
        const int iz = (int)(node / nxy);
        const bool interior = node < m.n3d && iz > 0 && iz < m.nz - 1;
        float xv[kMlpIn];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) xv[k] = 0.f;
        if (interior) {
            const int rem = (int)(node - (long long)iz * nxy), iy = rem / m.nx, ix = rem - iy * m.nx;
            const size_t b = (size_t)iz * nxy;
            const size_t xr = b + iy * m.nx + wrapi(ix + 1, m.nx), xl = b + iy * m.nx + wrapi(ix - 1, m.nx);
            const size_t yr = b + wrapi(iy + 1, m.ny) * m.nx + ix, yl = b + wrapi(iy - 1, m.ny) * m.nx + ix;
            const size_t c0 = (size_t)node, zr = c0 + nxy, zl = c0 - nxy;
            const T hm = m.DZ[iz - 1], hp = m.DZ[iz];
            const T den = hp * hm * (hp + hm);
            const T cp = hm * hm / den, cm = hp * hp / den;
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const T u0 = vel[3 * c0 + cc];
                const T gx = hx2 * (vel[3 * xr + cc] - vel[3 * xl + cc]);
                const T gy = hy2 * (vel[3 * yr + cc] - vel[3 * yl + cc]);
                const T gz = cp * (vel[3 * zr + cc] - u0) + cm * (u0 - vel[3 * zl + cc]);
                xv[3 * cc + 0] = (float)(gx * gscale);
                xv[3 * cc + 1] = (float)(gy * gscale);
                xv[3 * cc + 2] = (float)(gz * gscale);
            }
            const T z = m.ZZ[iz];
            const T d = (z - m.zlo) < (m.zhi - z) ? (z - m.zlo) : (m.zhi - z);
            xv[9] = (float)(d * yscale);
        }
        stage_inputs(buf, xv);
        __syncthreads();
        f32x4 o[4];
        forward_wave_r(R, w.b, buf, o);
        if constexpr (POST != kSgsPostNone) {
            // outputs of point 16 nt + r, feature 4 g + i -> tq[point][feature]
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (4 * g + i < kMlpOut) tq[(16 * nt + r) * 7 + 4 * g + i] = o[nt][i];
            __syncthreads();
            float t6[kMlpOut];
#pragma unroll
            for (int f = 0; f < kMlpOut; f++) t6[f] = tq[lane() * 7 + f];
            if constexpr (POST == kSgsPostClip) {
                // tau_ij S_ij with S_ij = (g_ij + g_ji) / 2, g_ij = xv[3 i + j] (positive scale)
                const float ts = t6[0] * xv[0] + t6[1] * xv[4] + t6[2] * xv[8] + t6[3] * (xv[1] + xv[3]) +
                                 t6[4] * (xv[2] + xv[6]) + t6[5] * (xv[5] + xv[7]);
                const bool keep = interior && !(ts > 0.f);
                if (node < m.n3d)
#pragma unroll
                    for (int f = 0; f < kMlpOut; f++) tau[node * kMlpOut + f] = keep ? T(t6[f]) * u_tau2 : T(0);
            } else {
                // D = S - tr(S) / 3 I in the inputs' (positive) gradient scale, which
                // cancels in (tau:D / D:D) D
                const float tr = (xv[0] + xv[4] + xv[8]) * (1.f / 3.f);
                float d[kMlpOut];
                d[0] = xv[0] - tr;
                d[1] = xv[4] - tr;
                d[2] = xv[8] - tr;
                d[3] = 0.5f * (xv[1] + xv[3]);
                d[4] = 0.5f * (xv[2] + xv[6]);
                d[5] = 0.5f * (xv[5] + xv[7]);
                const float td = t6[0] * d[0] + t6[1] * d[1] + t6[2] * d[2] +
                                 2.f * (t6[3] * d[3] + t6[4] * d[4] + t6[5] * d[5]);
                const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] +
                                 2.f * (d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
                const float c = (interior && td < 0.f && dd > 0.f) ? td / dd : 0.f;
                if (node < m.n3d)
#pragma unroll
                    for (int f = 0; f < kMlpOut; f++) tau[node * kMlpOut + f] = T(c * d[f]) * u_tau2;
            }
        } else {
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const long long p = p0 + 16 * nt + r;
                if (p >= m.n3d) continue;
                const int pz = (int)(p / nxy);
                const bool in = pz > 0 && pz < m.nz - 1;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int f = 4 * g + i;
                    if (f < kMlpOut) tau[p * kMlpOut + f] = in ? T(o[nt][i]) * u_tau2 : T(0);
                }
            }
        }
        __syncthreads();
    }
}

// Training samples at the interior nodes of the coarse grid from the 9 filtered
// fields q[f][iz][J][I] (f: u, v, w, uu, vv, ww, uv, uw, vw): inputs as the
// closure's, targets the deviatoric filt(u_i u_j) - filt(u_i) filt(u_j) over u_tau^2.
template <typename T>
__global__ void k_sgs_samples(const T* __restrict__ q, int nxc, int nyc, int nz, const T* __restrict__ ZZ,
                              const T* __restrict__ DZ, T hx2, T hy2, T zlo, T zhi, T gscale, T yscale, T inv_ut2,
                              float* __restrict__ X, float* __restrict__ Y) {
    const long long sidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long)nxc * nyc;
    if (sidx >= plane * (nz - 2)) return;
    const int I = (int)(sidx % nxc), J = (int)((sidx / nxc) % nyc), iz = (int)(sidx / plane) + 1;
    const size_t fs = (size_t)plane * nz;
    const size_t c0 = (size_t)iz * plane + (size_t)J * nxc + I;
    const size_t xr = (size_t)iz * plane + (size_t)J * nxc + wrapi(I + 1, nxc);
    const size_t xl = (size_t)iz * plane + (size_t)J * nxc + wrapi(I - 1, nxc);
    const size_t yr = (size_t)iz * plane + (size_t)wrapi(J + 1, nyc) * nxc + I;
    const size_t yl = (size_t)iz * plane + (size_t)wrapi(J - 1, nyc) * nxc + I;
    const size_t zr = c0 + plane, zl = c0 - plane;
    const T hm = DZ[iz - 1], hp = DZ[iz];
    const T den = hp * hm * (hp + hm);
    const T cp = hm * hm / den, cm = hp * hp / den;
    float* x = X + sidx * kMlpIn;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const T* f = q + c * fs;
        const T u0 = f[c0];
        x[3 * c + 0] = (float)(hx2 * (f[xr] - f[xl]) * gscale);
        x[3 * c + 1] = (float)(hy2 * (f[yr] - f[yl]) * gscale);
        x[3 * c + 2] = (float)((cp * (f[zr] - u0) + cm * (u0 - f[zl])) * gscale);
    }
    const T z = ZZ[iz];
    x[9] = (float)(((z - zlo) < (zhi - z) ? (z - zlo) : (zhi - z)) * yscale);
    constexpr int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    T tau[6];
#pragma unroll
    for (int k = 0; k < 6; k++) tau[k] = q[(3 + k) * fs + c0] - q[pa[k] * fs + c0] * q[pb[k] * fs + c0];
    const T tr = (tau[0] + tau[1] + tau[2]) / T(3);
    tau[0] -= tr;
    tau[1] -= tr;
    tau[2] -= tr;
    float* y = Y + sidx * kMlpOut;
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] = (float)(tau[k] * inv_ut2);
}

}  // namespace

void mlp_pack_device(const float* P, std::vector<unsigned short>& w, std::vector<float>& b) {
    auto bf = [](float f) {
        unsigned u;
        memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    };
    w.assign(kWBlock, 0);
    b.assign(kBBlock, 0.f);
    for (int o = 0; o < 64; o++)
        for (int i = 0; i < kMlpIn; i++) w[kW1p + o * 32 + i] = bf(P[kOffW1 + o * kMlpIn + i]);
    for (int o = 0; o < 64; o++)
        for (int i = 0; i < 64; i++) {
            w[kW2p + o * 64 + i] = w[kT2 + i * 64 + o] = bf(P[kOffW2 + o * 64 + i]);
            w[kW3p + o * 64 + i] = w[kT3 + i * 64 + o] = bf(P[kOffW3 + o * 64 + i]);
        }
    for (int o = 0; o < kMlpOut; o++)
        for (int i = 0; i < 64; i++) w[kW4p + o * 64 + i] = w[kT4 + i * 32 + o] = bf(P[kOffW4 + o * 64 + i]);
    for (int o = 0; o < 64; o++) {
        b[o] = P[kOffB1 + o];
        b[64 + o] = P[kOffB2 + o];
        b[128 + o] = P[kOffB3 + o];
    }
    for (int o = 0; o < kMlpOut; o++) b[192 + o] = P[kOffB4 + o];
}

namespace {
// Persistent grid: as many one-wave blocks as are resident on the device at once.
template <typename K>
long long resident_blocks(K kernel) {
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
        ncu = 256;
    return (long long)per_cu * ncu;
}
}  // namespace

hipError_t launch_mlp_forward(hipStream_t s, MlpDev w, const float* x, long long n, float* y) {
    if (n <= 0) return hipSuccess;
    static const long long resident = resident_blocks(k_mlp_forward);
    const long long nb = std::min((n + 63) / 64, resident);
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nb), dim3(64), 0, s, w, x, n, y);
    return hipGetLastError();
}

int mlp_train_blocks(long long n) {
    const long long nch = (n + 63) / 64;
    return (int)(nch < kMlpTrainMaxBlocks ? (nch < 1 ? 1 : nch) : kMlpTrainMaxBlocks);
}

hipError_t launch_mlp_partials(hipStream_t s, MlpDev w, const float* x, const float* t, long long n, float* partial) {
    const float gscale = (float)(2.0 / ((double)n * kMlpOut));
    hipLaunchKernelGGL(k_mlp_train, dim3(mlp_train_blocks(n)), dim3(64), 0, s, w, x, t, n, gscale, partial);
    return hipGetLastError();
}

hipError_t launch_mlp_reduce(hipStream_t s, const float* partial, int nrec, double* tmp, float* grad, double* loss,
                             const AdamArgs& a, float* params, float* m, float* v, unsigned short* wdev, float* bdev) {
    const int nb = (kMlpParams + 1 + 255) / 256;
    hipLaunchKernelGGL(k_mlp_reduce_split, dim3(nb, kMlpReduceSplit), dim3(256), 0, s, partial, nrec, tmp);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(nb), dim3(256), 0, s, tmp, grad, loss, a, params, m, v, wdev, bdev);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sgs_tau(hipStream_t s, const Mesh<T>& m, MlpDev w, const T* vel, T* tau, int post) {
    const double u_tau = (double)m.visc * (double)m.inv_len;
    const T ut2 = T(u_tau * u_tau), gs = T(1.0 / (u_tau * (double)m.inv_len)), ys = T((double)m.inv_len * kYPlusScale);
    const long long nch = ((long long)m.n3d + 63) / 64;
#define LSGS(P)                                                                                                   \
    {                                                                                                             \
        static const long long resident = resident_blocks(k_sgs_tau<T, P>);                                       \
        hipLaunchKernelGGL((k_sgs_tau<T, P>), dim3((unsigned)std::min(nch, resident)), dim3(64), 0, s, m, w, vel, \
                           tau, ut2, gs, ys);                                                                     \
    }
    if (post == kSgsPostClip) LSGS(kSgsPostClip)
    else if (post == kSgsPostProject) LSGS(kSgsPostProject)
    else LSGS(kSgsPostNone)
#undef LSGS
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sgs_samples(hipStream_t s, const Mesh<T>& m, const T* q, int nxc, int nyc, T hxc, T hyc, float* x,
                              float* y) {
    const double u_tau = (double)m.visc * (double)m.inv_len;
    const long long n = (long long)nxc * nyc * (m.nz - 2);
    hipLaunchKernelGGL((k_sgs_samples<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q, nxc, nyc, m.nz, m.ZZ,
                       m.DZ, T(0.5) / hxc, T(0.5) / hyc, m.zlo, m.zhi, T(1.0 / (u_tau * (double)m.inv_len)),
                       T((double)m.inv_len * kYPlusScale), T(1.0 / (u_tau * u_tau)), x, y);
    return hipGetLastError();
}

template hipError_t launch_sgs_samples<double>(hipStream_t, const Mesh<double>&, const double*, int, int, double, double,
                                               float*, float*);
template hipError_t launch_sgs_samples<float>(hipStream_t, const Mesh<float>&, const float*, int, int, float, float,
                                              float*, float*);
template hipError_t launch_sgs_tau<double>(hipStream_t, const Mesh<double>&, MlpDev, const double*, double*, int);
template hipError_t launch_sgs_tau<float>(hipStream_t, const Mesh<float>&, MlpDev, const float*, float*, int);

}  // namespace tc
