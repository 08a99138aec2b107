// Per-cell terms of the fused step diagnostics, shared by the diagnostics
// kernel and by the projection kernel, which can form them while it writes the
// new velocity: the per-plane sums of the profile statistics
// (staggered/func_global.h:712-808; collocated_AB3/func_global.h:207-289), the
// sums of w and w^2 at w's own points (flow rate, kinetic energy), and the
// maxima of the convective, viscous and total CFL numbers over tau
// (staggered/func_global.h:692-710; collocated_AB3/func_global.h:183-205).
// A partial record is NS sums followed by 3 maxima (diag_record(scheme) values).
#pragma once

#include "tc_internal.h"

namespace tc {

constexpr int kDiagThreads = 256;

template <int SCHEME>
struct DiagSums {
    static constexpr int NV = SCHEME == TC_SCHEME_AB3 ? 10 : 12;  // u[3], uu[6], nu_t (+ p, pp)
    static constexpr int NS = NV + 2, NX = NS + 3;
    double s[NS], mx[3];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int k = 0; k < NS; k++) s[k] = 0.0;
        mx[0] = mx[1] = mx[2] = 0.0;
    }
    // u, v: the cell's velocity; w: w at the cell centre (staggered) or the node;
    // wr: w at its own point; nut: the cell's nu_t (summed when has_nut); p: the
    // pressure (staggered); vc: the viscosity of the viscous CFL; ihx, ihy,
    // inv_hz and kv: the plane's mesh factors.
    __device__ __forceinline__ void add(double u, double v, double w, double wr, bool has_nut, double nut, double p,
                                        double vc, double ihx, double ihy, double inv_hz, double kv) {
        s[0] += u;
        s[1] += v;
        s[2] += w;
        s[3] += u * u;
        s[4] += v * v;
        s[5] += w * w;
        s[6] += u * v;
        s[7] += u * w;
        s[8] += v * w;
        if (has_nut) s[9] += nut;
        if (SCHEME != TC_SCHEME_AB3) {
            s[10] += p;
            s[11] += p * p;
        }
        s[NV] += wr;
        s[NV + 1] += wr * wr;
        const double cc = fabs(u) * ihx + fabs(v) * ihy + fabs(wr) * inv_hz;
        const double cv = kv * vc;
        mx[0] = fmax(mx[0], cc);
        mx[1] = fmax(mx[1], cv);
        mx[2] = fmax(mx[2], cc + cv);
    }
};

// Sum (or max) of each of NV values over a block of kDiagThreads threads; the
// result is in thread 0. Fixed order, so results do not depend on scheduling.
template <int NV>
__device__ __forceinline__ void block_reduce_diag(double (&v)[NV], double (*sh)[kDiagThreads / 64], bool is_max) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_down(x, o, 64);
            x = is_max ? fmax(x, y) : x + y;
        }
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double r = sh[k][0];
            for (int w = 1; w < kDiagThreads / 64; w++) r = is_max ? fmax(r, sh[k][w]) : r + sh[k][w];
            v[k] = r;
        }
}

// Reduces P values (P a power of two, 2 <= P <= 64) over a wavefront: value k
// is summed when bit k of summask is set, else maximised. Each of the first
// log2(P) steps gives away half of a lane's remaining values to the lane 32,
// 16, ... apart and combines the other half with what that lane gave, so that
// after P - 1 exchanges lane L holds one value, index L >> (6 - log2 P), over
// the lanes that share L's high bits; log2(64 / P) more exchanges finish it.
// The result: lane L with its low 6 - log2(P) bits clear holds value
// L >> (6 - log2 P) over the wave (P - 1 + log2(64 / P) exchanges instead of
// 6 per value). Fixed order, so results do not depend on scheduling.
// One halving step of wave_reduce_transpose: 2H values per lane -> H, with the
// lane OFF = 64 H / P apart; every index a compile-time constant (a dynamically
// indexed array would live in scratch memory).
template <int P, int H>
__device__ __forceinline__ void transpose_halve(double (&v)[P], int& base, unsigned long long summask, int lane) {
    if constexpr (H >= 1) {
        constexpr int OFF = 64 * H / P;
        const bool hi = (lane & OFF) != 0;
#pragma unroll
        for (int j = 0; j < H; j++) {
            const double keep = hi ? v[H + j] : v[j];
            const double give = hi ? v[j] : v[H + j];
            const double got = __shfl_xor(give, OFF, 64);
            const int idx = base + (hi ? H : 0) + j;
            v[j] = ((summask >> idx) & 1) ? keep + got : fmax(keep, got);
        }
        if (hi) base += H;
        transpose_halve<P, H / 2>(v, base, summask, lane);
    }
}

template <int P>
__device__ __forceinline__ double wave_reduce_transpose(double (&v)[P], unsigned long long summask) {
    static_assert(P >= 2 && P <= 64 && (P & (P - 1)) == 0, "P must be a power of two in [2, 64]");
    const int lane = threadIdx.x & 63;
    int base = 0;  // index of v[0] among the values
    transpose_halve<P, P / 2>(v, base, summask, lane);
    double x = v[0];
    const bool sum = (summask >> base) & 1;
#pragma unroll
    for (int off = 32 / P; off >= 1; off >>= 1) {
        const double y = __shfl_xor(x, off, 64);
        x = sum ? x + y : fmax(x, y);
    }
    return x;
}

// visc_fac of the viscous CFL number (CalcCFL_loc: 4 (nu + nu_t) per unit
// squared mesh step, doubled for the Galerkin / NAIVE viscous terms).
inline double diag_visc_fac(int scheme, int visc_type) {
    return scheme == TC_SCHEME_AB3 ? 4.0 * (visc_type + 1) : 4.0 * ((visc_type != 0) + 1);
}

}  // namespace tc
