// Stencil formulas of the explicit term, written once against a field accessor
// acc(c, dx, dy, dz) -> component c at offset (dx, dy, dz) from the cell, so the
// same arithmetic serves the LDS-tiled kernels and plain global-memory access.
// nu(dx, dy, dz) gives the turbulent viscosity at an offset.
#pragma once

#include "tc_internal.h"

namespace tc {

// Field accessors for real3 AoS arrays.
#define F3(a, c, dx, dy, dz) acc(c, dx, dy, dz)
#define UU(dx, dy, dz) acc(0, dx, dy, dz)
#define VV(dx, dy, dz) acc(1, dx, dy, dz)
#define WW(dx, dy, dz) acc(2, dx, dy, dz)

// The near-wall rules below read every operand unconditionally and then select:
// the plane index is uniform across a block, so a conditional read would become
// a branch with its own wait for the LDS or global load, once per rule and plane.
// Offsets past a wall read a stale or clamped value that the select discards;
// mesh array indices are clamped into the array.
__device__ __forceinline__ int zcl(int i, int nz) { return min(max(i, 0), nz - 1); }

// ---------------------------------------------------------------------------
// Convection, Harlow-Welch (staggered/func_global.h:345-370)
// ---------------------------------------------------------------------------
template <bool INT, typename T, class A>
__device__ __forceinline__ void conv_hw(const Mesh<T>& m, const A& acc, int iz, T f[3]) {
    // wall tests: on the interior planes (INT) every one is false
    const int zt = INT ? 3 : iz, NZt = INT ? 7 : m.nz;
    const int NZ = m.nz;
    const bool bot = zt == 0, top = zt == NZt - 1;
    const T dz = m.DZ[iz], dzm = bot ? T(0) : m.DZ[zcl(iz - 1, NZ)];
    const T inv_hz = m.IDZ[iz];        // 1 / DZ[iz]
    const T inv_hz_alt = m.IHDN[iz];   // 2 / (DZ[iz] + DZ[iz-1]), DZ[-1] = 0
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    const T up = UU(0, 0, 1), um = UU(0, 0, -1), vp = VV(0, 0, 1), vm = VV(0, 0, -1);
    const T wp = WW(0, 0, 1), wm = WW(0, 0, -1), wxp = WW(-1, 0, 1), wyp = WW(0, -1, 1);
    T f0 = 0, f1 = 0, f2 = 0;
    f0 -= ((u0 + UU(1, 0, 0)) * (u0 + UU(1, 0, 0)) - (u0 + UU(-1, 0, 0)) * (u0 + UU(-1, 0, 0))) * m.inv_hx;
    f0 -= ((u0 + UU(0, 1, 0)) * (VV(0, 1, 0) + VV(-1, 1, 0)) - (u0 + UU(0, -1, 0)) * (v0 + VV(-1, 0, 0))) * m.inv_hy;
    const T f0t = f0 - (u0 + up) * (wp + wxp) * inv_hz;
    f0 = top ? f0 : f0t;
    const T f0b = f0 + (u0 + um) * (w0 + WW(-1, 0, 0)) * inv_hz;
    f0 = bot ? f0 : f0b;

    f1 -= ((v0 + VV(1, 0, 0)) * (UU(1, 0, 0) + UU(1, -1, 0)) - (v0 + VV(-1, 0, 0)) * (u0 + UU(0, -1, 0))) * m.inv_hx;
    f1 -= ((v0 + VV(0, 1, 0)) * (v0 + VV(0, 1, 0)) - (v0 + VV(0, -1, 0)) * (v0 + VV(0, -1, 0))) * m.inv_hy;
    const T f1t = f1 - (v0 + vp) * (wp + wyp) * inv_hz;
    f1 = top ? f1 : f1t;
    const T f1b = f1 + (v0 + vm) * (w0 + WW(0, -1, 0)) * inv_hz;
    f1 = bot ? f1 : f1b;

    if (!bot) {
        f2 -= ((w0 + WW(1, 0, 0)) * (UU(1, 0, 0) * dz + UU(1, 0, -1) * dzm) - (w0 + WW(-1, 0, 0)) * (u0 * dz + um * dzm)) *
              m.inv_hx * inv_hz_alt;
        f2 -= ((w0 + WW(0, 1, 0)) * (VV(0, 1, 0) * dz + VV(0, 1, -1) * dzm) - (w0 + WW(0, -1, 0)) * (v0 * dz + vm * dzm)) *
              m.inv_hy * inv_hz_alt;
        const T wr = top ? T(0) : wp;
        f2 -= (w0 + wr) * (w0 + wr) * inv_hz_alt;
        f2 += (w0 + wm) * (w0 + wm) * inv_hz_alt;
    }
    f[0] = f0 * T(0.25);
    f[1] = f1 * T(0.25);
    f[2] = f2 * T(0.25);
}

// ---------------------------------------------------------------------------
// Convection, Verstappen-Veldman fourth order (staggered/func_global.h:374-529)
// Mass fluxes through faces: "bar" (9/-1 interpolation, weight DZ) and
// "barbar" (3h, weight DZW); near-wall mirror rules at iz in {0,1,2,NZ-3..NZ-1}.
// ---------------------------------------------------------------------------
// 4th-order interpolation of a transverse component along x at (dy, dz)
#define XBAR(A, dy, dz) (T(9) * (A(0, dy, dz) + A(-1, dy, dz)) - (A(-2, dy, dz) + A(1, dy, dz)))
#define YBAR(A, dx, dz) (T(9) * (A(dx, 0, dz) + A(dx, -1, dz)) - (A(dx, -2, dz) + A(dx, 1, dz)))

template <bool INT, typename T, class AC>
__device__ __forceinline__ T zbar_vv(const Mesh<T>& m, const AC& acc, int c, int dx, int dy, int iz,
                                     const T* __restrict__ A) {
    // wall tests: on the interior planes (INT) every one is false
    const int zt = INT ? 3 : iz, NZt = INT ? 7 : m.nz;
    // 9*(F(0)*A[iz] + F(-1)*A[iz-1]) - F(-2)*A[iz-2] - F(+1)*A[iz+1] with the wall rules
    const int NZ = m.nz;
    const T f0 = F3(vel, c, dx, dy, 0), fm = F3(vel, c, dx, dy, -1);
    const T fmm = F3(vel, c, dx, dy, -2), fp = F3(vel, c, dx, dy, 1);
    const T amm = A[zcl(iz - 2, NZ)], ap = A[zcl(iz + 1, NZ)], a_first = A[0], a_last = A[NZ - 1];
    T r = T(9) * (f0 * A[iz] + fm * A[iz - 1]);
    const T lo = fmm * amm, lo_wall = -fm * a_first;
    r -= (zt == 1) ? lo_wall : lo;
    const T hi = fp * ap, hi_wall = -f0 * a_last;
    r -= (zt == NZt - 1) ? hi_wall : hi;
    return r;
}

// only: -1 all components, else the one component (wave-uniform) the caller keeps.
// Each direction's term is a difference of face fluxes (interpolated mass flux
// times the sum of the two velocities), the regular faces weighted by DZ (or hy,
// hx hy), the 3h-faces by DZW / 27; those weights are per-plane constants and are
// applied once per direction, after the differences.
template <bool INT, typename T, class A>
__device__ __forceinline__ void conv_vv(const Mesh<T>& m, const A& acc, int iz, T f[3], int only = -1) {
    // wall tests: on the interior planes (INT) every one is false
    const int zt = INT ? 3 : iz, NZt = INT ? 7 : m.nz;
    const T C27 = T(1.0 / 27.0);
    const T hx = m.hx, hy = m.hy, hxy = hx * hy, hxy27 = C27 * hxy;
    const T dz = m.DZ[iz], dzw = m.DZW[iz];
    const T dzhy = dz * hy, dzhx = dz * hx, dzwhy27 = C27 * (dzw * hy), dzwhx27 = C27 * (dzw * hx);
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    T f0 = 0, f1 = 0, f2 = 0;
    // ---- u-equation
    if (only < 0 || only == 0) {
    {
        const T sp = u0 + UU(1, 0, 0), sm = UU(-1, 0, 0) + u0;
        const T s3p = u0 + UU(3, 0, 0), s3m = UU(-3, 0, 0) + u0;
        const T er = sp * (T(9) * sp - (UU(-1, 0, 0) + UU(2, 0, 0))) - sm * (T(9) * sm - (UU(-2, 0, 0) + UU(1, 0, 0)));
        const T ew = s3p * (T(9) * (UU(1, 0, 0) + UU(2, 0, 0)) - s3p) - s3m * (T(9) * (UU(-2, 0, 0) + UU(-1, 0, 0)) - s3m);
        f0 += dzhy * er - dzwhy27 * ew;
    }
    {
        const T er = XBAR(VV, 1, 0) * (u0 + UU(0, 1, 0)) - XBAR(VV, 0, 0) * (UU(0, -1, 0) + u0);
        const T ew = XBAR(VV, 2, 0) * (u0 + UU(0, 3, 0)) - XBAR(VV, -1, 0) * (UU(0, -3, 0) + u0);
        f0 += dzhx * er - dzwhx27 * ew;
    }
    {
        const T x0 = XBAR(WW, 0, 0), x1 = XBAR(WW, 0, 1);
        const T xm1 = XBAR(WW, 0, -1), x2 = XBAR(WW, 0, 2);
        const T p1 = UU(0, 0, 1), p2 = UU(0, 0, 2), p3 = UU(0, 0, 3);
        const T m1 = UU(0, 0, -1), m2 = UU(0, 0, -2), m3 = UU(0, 0, -3);
        const T bl = zt == 0 ? T(0) : x0;
        const T br = zt == NZt - 1 ? T(0) : x1;
        const T bbl = zt == 1 ? T(0) : zt == 0 ? br : xm1;
        const T bbr = zt == NZt - 2 ? T(0) : zt == NZt - 1 ? bl : x2;
        const T q_r = zt == NZt - 1 ? T(0) : u0 + p1;
        const T q_rr = u0 + (zt == NZt - 3 ? -p2 : zt == NZt - 2 ? T(0) : zt == NZt - 1 ? m2 : p3);
        const T q_l = zt == 0 ? T(0) : m1 + u0;
        const T q_ll = u0 + (zt == 2 ? -m2 : zt == 1 ? T(0) : zt == 0 ? p2 : m3);
        f0 += hxy * (br * q_r - bl * q_l) - hxy27 * (bbr * q_rr - bbl * q_ll);
    }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the equations' live ranges apart
    // ---- v-equation
    if (only < 0 || only == 1) {
    {
        const T er = YBAR(UU, 1, 0) * (v0 + VV(1, 0, 0)) - YBAR(UU, 0, 0) * (VV(-1, 0, 0) + v0);
        const T ew = YBAR(UU, 2, 0) * (v0 + VV(3, 0, 0)) - YBAR(UU, -1, 0) * (VV(-3, 0, 0) + v0);
        f1 += dzhy * er - dzwhy27 * ew;
    }
    {
        const T sp = v0 + VV(0, 1, 0), sm = VV(0, -1, 0) + v0;
        const T s3p = v0 + VV(0, 3, 0), s3m = VV(0, -3, 0) + v0;
        const T er = sp * (T(9) * sp - (VV(0, -1, 0) + VV(0, 2, 0))) - sm * (T(9) * sm - (VV(0, -2, 0) + VV(0, 1, 0)));
        const T ew = s3p * (T(9) * (VV(0, 1, 0) + VV(0, 2, 0)) - s3p) - s3m * (T(9) * (VV(0, -2, 0) + VV(0, -1, 0)) - s3m);
        f1 += dzhx * er - dzwhx27 * ew;
    }
    {
        const T y0 = YBAR(WW, 0, 0), y1 = YBAR(WW, 0, 1);
        const T ym1 = YBAR(WW, 0, -1), y2 = YBAR(WW, 0, 2);
        const T p1 = VV(0, 0, 1), p2 = VV(0, 0, 2), p3 = VV(0, 0, 3);
        const T m1 = VV(0, 0, -1), m2 = VV(0, 0, -2), m3 = VV(0, 0, -3);
        const T bl = zt == 0 ? T(0) : y0;
        const T br = zt == NZt - 1 ? T(0) : y1;
        const T bbl = zt == 1 ? T(0) : zt == 0 ? br : ym1;
        const T bbr = zt == NZt - 2 ? T(0) : zt == NZt - 1 ? bl : y2;
        const T q_r = zt == NZt - 1 ? T(0) : v0 + p1;
        const T q_rr = v0 + (zt == NZt - 3 ? -p2 : zt == NZt - 2 ? T(0) : zt == NZt - 1 ? m2 : p3);
        const T q_l = zt == 0 ? T(0) : m1 + v0;
        const T q_ll = v0 + (zt == 2 ? -m2 : zt == 1 ? T(0) : zt == 0 ? p2 : m3);
        f1 += hxy * (br * q_r - bl * q_l) - hxy27 * (bbr * q_rr - bbl * q_ll);
    }
    }
    __builtin_amdgcn_sched_barrier(0);
    // ---- w-equation (no unknown at iz = 0)
    if (zt != 0 && (only < 0 || only == 2)) {
        {
            const T br = zbar_vv<INT>(m, acc, 0, 1, 0, iz, m.DZ);
            const T bl = zbar_vv<INT>(m, acc, 0, 0, 0, iz, m.DZ);
            const T bbr = zbar_vv<INT>(m, acc, 0, 2, 0, iz, m.DZW);
            const T bbl = zbar_vv<INT>(m, acc, 0, -1, 0, iz, m.DZW);
            f2 += hy * ((br * (w0 + WW(1, 0, 0)) - bl * (WW(-1, 0, 0) + w0)) -
                        C27 * (bbr * (w0 + WW(3, 0, 0)) - bbl * (WW(-3, 0, 0) + w0)));
        }
        {
            // The reference scales this y-face flux by hy (staggered/func_global.h:475-487); kept.
            const T br = zbar_vv<INT>(m, acc, 1, 0, 1, iz, m.DZ);
            const T bl = zbar_vv<INT>(m, acc, 1, 0, 0, iz, m.DZ);
            const T bbr = zbar_vv<INT>(m, acc, 1, 0, 2, iz, m.DZW);
            const T bbl = zbar_vv<INT>(m, acc, 1, 0, -1, iz, m.DZW);
            f2 += hy * ((br * (w0 + WW(0, 1, 0)) - bl * (WW(0, -1, 0) + w0)) -
                        C27 * (bbr * (w0 + WW(0, 3, 0)) - bbl * (WW(0, -3, 0) + w0)));
        }
        {
            const T wzl = WW(0, 0, -1);
            const T p1 = WW(0, 0, 1), p2 = WW(0, 0, 2), p3 = WW(0, 0, 3);
            const T m2 = WW(0, 0, -2), m3 = WW(0, 0, -3);
            const T wzr = zt == NZt - 1 ? T(0) : p1;
            const T wzll = zt == 1 ? w0 : m2;
            const T wzrr = zt == NZt - 2 ? T(0) : zt == NZt - 1 ? w0 : p2;
            const T wzlll = zt == 1 ? p1 : zt == 2 ? wzl : m3;
            const T wzrrr = zt == NZt - 3 ? T(0) : zt == NZt - 2 ? p1 : zt == NZt - 1 ? wzl : p3;
            const T sl = wzl + w0, sr = w0 + wzr;
            const T bl = T(9) * sl - (wzll + wzr);
            const T br = T(9) * sr - (wzl + wzrr);
            const T bbl = T(9) * (wzll + wzl) - (wzlll + w0);
            const T bbr = T(9) * (wzr + wzrr) - (w0 + wzrrr);
            const T qr = zt < NZt - 2 ? (w0 + wzrrr) : zt == NZt - 2 ? (-p1 + w0) : (wzl + w0);
            const T ql = zt > 2 ? (wzlll + w0) : zt == 2 ? (-wzl + w0) : (p1 + w0);
            f2 += hxy * (br * sr - bl * sl) - hxy27 * (bbr * qr - bbl * ql);
        }
    }
    const T minus_inv_vol = m.IVOL[iz];  // -1 / (32 (dz - dzw/9) hx hy)
    f[0] = f0 * minus_inv_vol;
    f[1] = f1 * minus_inv_vol;
    f[2] = f2 * (-T(1. / 32.) * m.inv_hx * m.inv_hy * m.GCZ[iz]);
}

// ---------------------------------------------------------------------------
// Viscous term, constant viscosity (staggered/func_global.h:533-562)
// ---------------------------------------------------------------------------
template <bool INT, typename T, class A>
__device__ __forceinline__ void visc_cross(const Mesh<T>& m, const A& acc, int iz, T f[3], int only = -1) {
    // wall tests: on the interior planes (INT) every one is false
    const int zt = INT ? 3 : iz, NZt = INT ? 7 : m.nz;
    const bool bot = zt == 0, top = zt == NZt - 1;
    const T inv_hz = m.IDZ[iz];        // 1 / DZ[iz]
    const T inv_hz_alt = m.IHDN[iz];   // 2 / (DZ[iz] + DZ[iz-1]), DZ[-1] = 0
    const T ihx2 = m.inv_hx * m.inv_hx, ihy2 = m.inv_hy * m.inv_hy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (only >= 0 && c != only) {
            f[c] = T(0);
            continue;
        }
        const T a0 = F3(vel, c, 0, 0, 0);
        f[c] = ihx2 * (F3(vel, c, 1, 0, 0) + F3(vel, c, -1, 0, 0) - T(2) * a0);
        f[c] += ihy2 * (F3(vel, c, 0, 1, 0) + F3(vel, c, 0, -1, 0) - T(2) * a0);
    }
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    const T up = UU(0, 0, 1), vp = VV(0, 0, 1), wp = WW(0, 0, 1);
    const T um = UU(0, 0, -1), vm = VV(0, 0, -1), wm = WW(0, 0, -1);
    {  // upper neighbour; past the top wall the no-slip ghost (-u) enters as 0 - u
        const T inv_hz_up = m.IHUP[iz];  // 2 / (DZ[iz] + DZ[iz+1]), DZ[nz] = 0
        f[0] += ((top ? T(0) : up) - u0) * inv_hz_up * inv_hz;
        f[1] += ((top ? T(0) : vp) - v0) * inv_hz_up * inv_hz;
        f[2] += ((top ? T(0) : wp) - w0) * inv_hz * inv_hz_alt;
    }
    {  // lower neighbour; at the bottom wall there is no w-equation
        f[0] += ((bot ? T(0) : um) - u0) * inv_hz_alt * inv_hz;
        f[1] += ((bot ? T(0) : vm) - v0) * inv_hz_alt * inv_hz;
        const T f2b = f[2] + (wm - w0) * m.IDZM[iz] * inv_hz_alt;  // 1 / DZ[iz-1]
        f[2] = bot ? f[2] : f2b;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] *= m.visc;
}

// ---------------------------------------------------------------------------
// Viscous term, div(2 (nu + nu_t) S) (staggered/func_global.h:565-625),
// including the edge-average of sxy_190 at :598 as the reference has it.
// ---------------------------------------------------------------------------
template <bool INT, typename T, bool TURB, class A, class NU>
__device__ __forceinline__ void visc_naive(const Mesh<T>& m, const A& acc, const NU& nu, int iz, T f[3], int only = -1) {
    // wall tests: on the interior planes (INT) every one is false
    const int zt = INT ? 3 : iz, NZt = INT ? 7 : m.nz;
    const T visc = m.visc;
    auto mu = [&](int dx, int dy, int dz) -> T { return TURB ? visc + nu(dx, dy, dz) : visc; };
    const bool bot = zt == 0, top = zt == NZt - 1;
    const T dz = m.DZ[iz];
    const T inv_hz = m.IDZ[iz];         // 1 / DZ[iz]
    const T inv_hz_up = m.IHUP[iz];     // 2 / (DZ[iz] + DZ[iz+1]), DZ[nz] = 0
    const T inv_hz_down = m.IHDN[iz];   // 2 / (DZ[iz] + DZ[iz-1]), DZ[-1] = 0
    const T ihx = m.inv_hx, ihy = m.inv_hy;
    const T q = T(0.25);
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    const T mu0 = mu(0, 0, 0);
    // stresses shared by two equations (recomputed where a thread keeps one component)
    auto sxy_990 = [&]() {
        return ((u0 - UU(0, -1, 0)) * ihy + (v0 - VV(-1, 0, 0)) * ihx) * q * (mu(0, -1, 0) + mu(-1, -1, 0) + mu0 + mu(-1, 0, 0));
    };
    auto sxz_909 = [&]() {
        const T wall = T(2) * u0 * inv_hz * visc;
        const T in = ((u0 - UU(0, 0, -1)) * inv_hz_down + (w0 - WW(-1, 0, 0)) * ihx) * q *
                     (mu(0, 0, -1) + mu(-1, 0, -1) + mu0 + mu(-1, 0, 0));
        return bot ? wall : in;
    };
    auto syz_099 = [&]() {
        const T wall = T(2) * v0 * inv_hz * visc;
        const T in = ((v0 - VV(0, 0, -1)) * inv_hz_down + (w0 - WW(0, -1, 0)) * ihy) * q *
                     (mu(0, 0, -1) + mu(0, -1, -1) + mu0 + mu(0, -1, 0));
        return bot ? wall : in;
    };
    T f0 = 0, f1 = 0, f2 = 0;
    if (only < 0 || only == 0) {
        const T sxx_000 = T(2) * (UU(1, 0, 0) - u0) * ihx * mu0;
        const T sxx_800 = T(2) * (u0 - UU(-1, 0, 0)) * ihx * mu(-1, 0, 0);
        f0 += (sxx_000 - sxx_800) * ihx;
        const T sxy_910 = ((UU(0, 1, 0) - u0) * ihy + (VV(0, 1, 0) - VV(-1, 1, 0)) * ihx) * q * (mu0 + mu(-1, 0, 0) + mu(0, 1, 0) + mu(-1, 1, 0));
        f0 += (sxy_910 - sxy_990()) * ihy;
        const T wall = T(-2) * u0 * inv_hz * visc;
        const T in = ((UU(0, 0, 1) - u0) * inv_hz_up + (WW(0, 0, 1) - WW(-1, 0, 1)) * ihx) * q *
                     (mu0 + mu(-1, 0, 0) + mu(0, 0, 1) + mu(-1, 0, 1));
        const T sxz_901 = top ? wall : in;
        f0 += (sxz_901 - sxz_909()) * inv_hz;
    }
    if (only < 0 || only == 1) {
        const T sxy_190 = ((UU(1, 0, 0) - UU(1, -1, 0)) * ihy + (VV(1, 0, 0) - v0) * ihx) * q * (mu0 + mu(1, 0, 0) + mu(0, -1, 0) + mu(1, 1, 0));
        f1 += (sxy_190 - sxy_990()) * ihx;
        const T syy_000 = T(2) * (VV(0, 1, 0) - v0) * ihy * mu0;
        const T syy_080 = T(2) * (v0 - VV(0, -1, 0)) * ihy * mu(0, -1, 0);
        f1 += (syy_000 - syy_080) * ihy;
        const T wall = T(-2) * v0 * inv_hz * visc;
        const T in = ((VV(0, 0, 1) - v0) * inv_hz_up + (WW(0, 0, 1) - WW(0, -1, 1)) * ihy) * q *
                     (mu0 + mu(0, -1, 0) + mu(0, 0, 1) + mu(0, -1, 1));
        const T syz_091 = top ? wall : in;
        f1 += (syz_091 - syz_099()) * inv_hz;
    }
    if (zt != 0 && (only < 0 || only == 2)) {
        const T sxz_109 = ((UU(1, 0, 0) - UU(1, 0, -1)) * inv_hz_down + (WW(1, 0, 0) - w0) * ihx) * q *
                          (mu(0, 0, -1) + mu(1, 0, -1) + mu0 + mu(1, 0, 0));
        f2 += (sxz_109 - sxz_909()) * ihx;
        const T syz_019 = ((VV(0, 1, 0) - VV(0, 1, -1)) * inv_hz_down + (WW(0, 1, 0) - w0) * ihy) * q *
                          (mu(0, 0, -1) + mu(0, 1, -1) + mu0 + mu(0, 1, 0));
        f2 += (syz_019 - syz_099()) * ihy;
        const T wp = WW(0, 0, 1);
        const T szz_000 = T(2) * ((top ? T(0) : wp) - w0) * inv_hz * mu0;
        const T szz_008 = T(2) * (w0 - WW(0, 0, -1)) * mu(0, 0, -1) / m.DZ[iz - 1];
        f2 += (szz_000 - szz_008) * inv_hz_down;
    }
    f[0] = f0;
    f[1] = f1;
    f[2] = f2;
}

// ---------------------------------------------------------------------------
// Collocated nodes: skew-symmetric central convection + viscous term
// (collocated_AB3/func_fluxes.h:21-120). Interior nodes only.
// ---------------------------------------------------------------------------
template <typename T, class A>
__device__ __forceinline__ void conv_ab3(const Mesh<T>& m, const A& acc, int iz, T inv_hz, T f[3]) {
    const T ih[3] = {m.inv_hx, m.inv_hy, inv_hz};
    const T u0[3] = {UU(0, 0, 0), VV(0, 0, 0), WW(0, 0, 0)};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int ox = d == 0, oy = d == 1, oz = d == 2;
        const T uj = acc(d, -ox, -oy, -oz), uk = acc(d, ox, oy, oz);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T ck = acc(c, ox, oy, oz), cj = acc(c, -ox, -oy, -oz);
            const T conv_div = T(0.5) * (ck * uk - cj * uj);
            const T conv_adv = u0[d] * T(0.5) * (ck - cj);
            f[c] -= T(0.5) * (conv_div + conv_adv) * ih[d];
        }
    }
}

// Q1-Galerkin viscous term with element nu_t (collocated_AB3/func_fluxes.h:21-85).
template <typename T, bool TURB, class A, class NU>
__device__ __forceinline__ void visc_galerkin(const Mesh<T>& m, const A& acc, const NU& nu, int iz, T g_out[3]) {
    const T C9 = T(1.0 / 9.0), C18 = T(1.0 / 18.0), C36 = T(1.0 / 36.0), C3 = T(1.0 / 3.0), C6 = T(1.0 / 6.0);
    const T hx = m.hx, hy = m.hy;
    T sum[3] = {0, 0, 0};
#pragma unroll
    for (int oz = 0; oz <= 1; oz++) {
        const int sz = oz ? 1 : -1;
        const T hz = oz == 0 ? m.DZ[iz - 1] : m.DZ[iz];
        const T inv_hz = T(1) / hz;
#pragma unroll
        for (int oy = 0; oy <= 1; oy++) {
            const int sy = oy ? 1 : -1;
#pragma unroll
            for (int ox = 0; ox <= 1; ox++) {
                const int sx = ox ? 1 : -1;
                const T mx = m.inv_hx * hy * hz;
                const T my = m.inv_hy * hx * hz;
                const T mz = inv_hz * hx * hy;
                T U[8][3];  // corners: bit0 x, bit1 y, bit2 z
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int dx = (k & 1) ? sx : 0, dy = (k & 2) ? sy : 0, dz = (k & 4) ? sz : 0;
                    U[k][0] = acc(0, dx, dy, dz);
                    U[k][1] = acc(1, dx, dy, dz);
                    U[k][2] = acc(2, dx, dy, dz);
                }
                // names as in the reference: in, in_x, in_y, in_xy, in_z, in_xz, in_yz, in_xyz
#define Q(name, c) U[name][c]
                const int I0 = 0, IX_ = 1, IY_ = 2, IXY = 3, IZ_ = 4, IXZ = 5, IYZ = 6, IXYZ = 7;
                T g[3];
#pragma unroll
                for (int c = 0; c < 3; c++)
                    g[c] = (C9 * (Q(IX_, c) - Q(I0, c)) + C18 * (Q(IXY, c) - Q(IY_, c)) + C18 * (Q(IXZ, c) - Q(IZ_, c)) + C36 * (Q(IXYZ, c) - Q(IYZ, c))) * mx +
                           (C9 * (Q(IY_, c) - Q(I0, c)) + C18 * (Q(IXY, c) - Q(IX_, c)) + C18 * (Q(IYZ, c) - Q(IZ_, c)) + C36 * (Q(IXYZ, c) - Q(IXZ, c))) * my +
                           (C9 * (Q(IZ_, c) - Q(I0, c)) + C18 * (Q(IXZ, c) - Q(IX_, c)) + C18 * (Q(IYZ, c) - Q(IY_, c)) + C36 * (Q(IXYZ, c) - Q(IXY, c))) * mz;
                T gxx = C9 * (Q(IX_, 0) - Q(I0, 0)) + C18 * (Q(IXY, 0) - Q(IY_, 0)) + C18 * (Q(IXZ, 0) - Q(IZ_, 0)) + C36 * (Q(IXYZ, 0) - Q(IYZ, 0));
                T gxy = T(0.25) * (C3 * (Q(IX_, 1) - Q(I0, 1) + Q(IXY, 1) - Q(IY_, 1)) + C6 * (Q(IXZ, 1) - Q(IZ_, 1) + Q(IXYZ, 1) - Q(IYZ, 1)));
                T gxz = T(0.25) * (C3 * (Q(IX_, 2) - Q(I0, 2) + Q(IXZ, 2) - Q(IZ_, 2)) + C6 * (Q(IXY, 2) - Q(IY_, 2) + Q(IXYZ, 2) - Q(IYZ, 2)));
                T gyy = C9 * (Q(IY_, 1) - Q(I0, 1)) + C18 * (Q(IXY, 1) - Q(IX_, 1)) + C18 * (Q(IYZ, 1) - Q(IZ_, 1)) + C36 * (Q(IXYZ, 1) - Q(IXZ, 1));
                T gyx = T(0.25) * (C3 * (Q(IY_, 0) - Q(I0, 0) + Q(IXY, 0) - Q(IX_, 0)) + C6 * (Q(IYZ, 0) - Q(IZ_, 0) + Q(IXYZ, 0) - Q(IXZ, 0)));
                T gyz = T(0.25) * (C3 * (Q(IY_, 2) - Q(I0, 2) + Q(IYZ, 2) - Q(IZ_, 2)) + C6 * (Q(IXY, 2) - Q(IX_, 2) + Q(IXYZ, 2) - Q(IXZ, 2)));
                T gzz = C9 * (Q(IZ_, 2) - Q(I0, 2)) + C18 * (Q(IYZ, 2) - Q(IY_, 2)) + C18 * (Q(IXZ, 2) - Q(IX_, 2)) + C36 * (Q(IXYZ, 2) - Q(IXY, 2));
                T gzx = T(0.25) * (C3 * (Q(IZ_, 0) - Q(I0, 0) + Q(IXZ, 0) - Q(IX_, 0)) + C6 * (Q(IYZ, 0) - Q(IY_, 0) + Q(IXYZ, 0) - Q(IXY, 0)));
                T gzy = T(0.25) * (C3 * (Q(IZ_, 1) - Q(I0, 1) + Q(IYZ, 1) - Q(IY_, 1)) + C6 * (Q(IXZ, 1) - Q(IX_, 1) + Q(IXYZ, 1) - Q(IXY, 1)));
#undef Q
                gxx *= mx;
                gxy *= hz * T(sx * sy);
                gxz *= hy * T(sx * sz);
                gyy *= my;
                gyx *= hz * T(sx * sy);
                gyz *= hx * T(sy * sz);
                gzz *= mz;
                gzx *= hy * T(sx * sz);
                gzy *= hx * T(sy * sz);
                g[0] += gxx + gxy + gxz;
                g[1] += gyy + gyx + gyz;
                g[2] += gzz + gzx + gzy;
                T vc = m.visc;
                if (TURB) vc += nu(ox ? 0 : -1, oy ? 0 : -1, oz ? 0 : -1);
#pragma unroll
                for (int c = 0; c < 3; c++) sum[c] += g[c] * vc;
            }
        }
    }
    g_out[0] = sum[0];
    g_out[1] = sum[1];
    g_out[2] = sum[2];
}

#undef F3
#undef UU
#undef VV
#undef WW
#undef XBAR
#undef YBAR

// Periodic-neighbour accessor over a real3 AoS array in global memory for
// offsets -R..R in x and y (any offset in z).
template <typename T, int R>
struct NbVec {
    const T* __restrict__ v;
    int x[2 * R + 1], y[2 * R + 1], iz, nz, nxy;
    // z neighbours clamped to the grid: every load is valid, so the wall rules
    // select among values loaded unconditionally (no branch around a load)
    __device__ __forceinline__ T operator()(int c, int dx, int dy, int dz) const {
        const int zz = min(max(iz + dz, 0), nz - 1);
        return v[3 * (zz * nxy + y[dy + R] + x[dx + R]) + c];
    }
};
template <typename T, int R>
__device__ __forceinline__ NbVec<T, R> make_nbvec(const T* v, int ix, int iy, int iz, int nx, int ny, int nz) {
    NbVec<T, R> a;
    a.v = v;
#pragma unroll
    for (int d = -R; d <= R; d++) {
        int xx = ix + d, yy = iy + d;
        xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
        yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
        a.x[d + R] = xx;
        a.y[d + R] = yy * nx;
    }
    a.nxy = nx * ny;
    a.iz = iz;
    a.nz = nz;
    return a;
}

// The values the divergence at a cell reads: one set of 12 offsets covers the
// three schemes. A kernel loads them all first (load_taps) and evaluates
// div_cell on the loaded copy, so every load of a batch of cells is in flight
// before any is used. Offsets are literal at every use, so the lookup folds to
// a register.
__host__ __device__ constexpr int tap_key(int c, int dx, int dy, int dz) {
    return ((c * 8 + dx + 3) * 8 + dy + 3) * 8 + dz + 3;
}
template <typename T>
struct TapVals {
    T v[12];
    __device__ __forceinline__ T operator()(int c, int dx, int dy, int dz) const {
        switch (tap_key(c, dx, dy, dz)) {
            case tap_key(0, 1, 0, 0): return v[0];
            case tap_key(0, 0, 0, 0): return v[1];
            case tap_key(0, 2, 0, 0): return v[2];
            case tap_key(0, -1, 0, 0): return v[3];
            case tap_key(1, 0, 1, 0): return v[4];
            case tap_key(1, 0, 0, 0): return v[5];
            case tap_key(1, 0, 2, 0): return v[6];
            case tap_key(1, 0, -1, 0): return v[7];
            case tap_key(2, 0, 0, 0): return v[8];
            case tap_key(2, 0, 0, 1): return v[9];
            case tap_key(2, 0, 0, 2): return v[10];
            case tap_key(2, 0, 0, -1): return v[11];
            default: return T(0);
        }
    }
};
template <typename T, int SCHEME, class A>
__device__ __forceinline__ void load_taps(const A& acc, TapVals<T>& r) {
    r.v[0] = acc(0, 1, 0, 0);
    r.v[4] = acc(1, 0, 1, 0);
    r.v[8] = acc(2, 0, 0, 0);
    r.v[9] = acc(2, 0, 0, 1);
    if (SCHEME != TC_SCHEME_AB3) {
        r.v[1] = acc(0, 0, 0, 0);
        r.v[5] = acc(1, 0, 0, 0);
    }
    if (SCHEME != TC_SCHEME_HW) {
        r.v[3] = acc(0, -1, 0, 0);
        r.v[7] = acc(1, 0, -1, 0);
        r.v[11] = acc(2, 0, 0, -1);
    }
    if (SCHEME == TC_SCHEME_VV) {
        r.v[2] = acc(0, 2, 0, 0);
        r.v[6] = acc(1, 0, 2, 0);
        r.v[10] = acc(2, 0, 0, 2);
    }
}

// Volume-weighted discrete divergence at one cell (staggered/func_global.h:306-312
// HW, :325-333 VV; collocated_AB3/func_global.h:310-325). dz = DZ[iz], dzw = DZW[iz]
// (staggered), ch = the collocated node's cell height.
template <typename T, int SCHEME, class A>
__device__ __forceinline__ T div_cell(const Mesh<T>& m, const A& acc, int iz, T dz, T dzw, T ch) {
    const int NZ = m.nz;
    if (SCHEME == TC_SCHEME_HW) {
        const T hz = dz;
        const T w0 = acc(2, 0, 0, 0), wp = acc(2, 0, 0, 1);
        return (acc(0, 1, 0, 0) - acc(0, 0, 0, 0)) * m.inv_hx * hz + (acc(1, 0, 1, 0) - acc(1, 0, 0, 0)) * m.inv_hy * hz +
               ((iz < NZ - 1) ? (wp - w0) : (-w0));
    } else if (SCHEME == TC_SCHEME_VV) {
        const T c27 = T(1.0 / 27.0);
        const T dudx = m.inv_hx * (dz * (acc(0, 1, 0, 0) - acc(0, 0, 0, 0)) - c27 * dzw * (acc(0, 2, 0, 0) - acc(0, -1, 0, 0)));
        const T dvdy = m.inv_hy * (dz * (acc(1, 0, 1, 0) - acc(1, 0, 0, 0)) - c27 * dzw * (acc(1, 0, 2, 0) - acc(1, 0, -1, 0)));
        const T w0 = acc(2, 0, 0, 0), wp1 = acc(2, 0, 0, 1), wp2 = acc(2, 0, 0, 2), wm1 = acc(2, 0, 0, -1);
        const T wr = (iz == NZ - 1) ? T(0) : wp1;
        const T wrr = (iz == NZ - 1) ? w0 : (iz == NZ - 2) ? T(0) : wp2;
        const T wll = (iz == 0) ? wp1 : (iz == 1) ? T(0) : wm1;
        return dudx + dvdy + (wr - w0) - c27 * (wrr - wll);
    } else {
        const T hz = ch;
        T d = T(0.5) * (acc(0, 1, 0, 0) - acc(0, -1, 0, 0)) * m.hy * hz;
        d += T(0.5) * (acc(1, 0, 1, 0) - acc(1, 0, -1, 0)) * m.hx * hz;
        const T w0 = acc(2, 0, 0, 0), wp = acc(2, 0, 0, 1), wm = acc(2, 0, 0, -1);
        const T dw = iz == 0 ? (wp - w0) : iz == NZ - 1 ? (w0 - wm) : (wp - wm);
        d += T(0.5) * dw * m.hx * m.hy;
        return d;
    }
}

}  // namespace tc
