// Stencil kernels: explicit term + BDF3/EX3 right-hand side, turbulent
// viscosity, divergence, gradient and the velocity correction.
//
// Launch shape: one thread per cell, blockIdx.y = wall-normal plane, so every
// branch on iz (the near-wall closures of the fourth-order scheme) is uniform
// across a wavefront and every mesh coefficient is a scalar load.
#include "kernels.h"

namespace tc {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Periodic neighbours of one cell: linear x index and y row offset for -3..3.
struct Nb {
    int x[7], y[7], base, nxy;
    __device__ __forceinline__ int at(int dx, int dy, int dz) const { return base + dz * nxy + y[dy + 3] + x[dx + 3]; }
};

__device__ __forceinline__ bool cell_of_thread(int nx, int nxy, int& ix, int& iy) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nxy) return false;
    iy = t / nx;
    ix = t - iy * nx;
    return true;
}

template <int R>
__device__ __forceinline__ void make_nb(Nb& nb, int ix, int iy, int iz, int nx, int ny) {
#pragma unroll
    for (int d = -R; d <= R; d++) {
        nb.x[d + 3] = wrap(ix + d, nx);
        nb.y[d + 3] = wrap(iy + d, ny) * nx;
    }
    nb.nxy = nx * ny;
    nb.base = iz * nb.nxy;
}

// Field accessors for real3 AoS arrays.
#define F3(a, c, dx, dy, dz) a[3 * nb.at(dx, dy, dz) + (c)]
#define UU(dx, dy, dz) F3(vel, 0, dx, dy, dz)
#define VV(dx, dy, dz) F3(vel, 1, dx, dy, dz)
#define WW(dx, dy, dz) F3(vel, 2, dx, dy, dz)

// ---------------------------------------------------------------------------
// Convection, Harlow-Welch (staggered/func_global.h:345-370)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void conv_hw(const Mesh<T>& m, const T* __restrict__ vel, const Nb& nb, int iz, T f[3]) {
    const int NZ = m.nz;
    const T dz = m.DZ[iz], dzm = iz == 0 ? T(0) : m.DZ[iz - 1];
    const T inv_hz = T(1) / dz;
    const T inv_hz_alt = T(2) / (dz + dzm);
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    T f0 = 0, f1 = 0, f2 = 0;
    f0 -= ((u0 + UU(1, 0, 0)) * (u0 + UU(1, 0, 0)) - (u0 + UU(-1, 0, 0)) * (u0 + UU(-1, 0, 0))) * m.inv_hx;
    f0 -= ((u0 + UU(0, 1, 0)) * (VV(0, 1, 0) + VV(-1, 1, 0)) - (u0 + UU(0, -1, 0)) * (v0 + VV(-1, 0, 0))) * m.inv_hy;
    if (iz != NZ - 1) f0 -= (u0 + UU(0, 0, 1)) * (WW(0, 0, 1) + WW(-1, 0, 1)) * inv_hz;
    if (iz != 0) f0 += (u0 + UU(0, 0, -1)) * (w0 + WW(-1, 0, 0)) * inv_hz;

    f1 -= ((v0 + VV(1, 0, 0)) * (UU(1, 0, 0) + UU(1, -1, 0)) - (v0 + VV(-1, 0, 0)) * (u0 + UU(0, -1, 0))) * m.inv_hx;
    f1 -= ((v0 + VV(0, 1, 0)) * (v0 + VV(0, 1, 0)) - (v0 + VV(0, -1, 0)) * (v0 + VV(0, -1, 0))) * m.inv_hy;
    if (iz != NZ - 1) f1 -= (v0 + VV(0, 0, 1)) * (WW(0, 0, 1) + WW(0, -1, 1)) * inv_hz;
    if (iz != 0) f1 += (v0 + VV(0, 0, -1)) * (w0 + WW(0, -1, 0)) * inv_hz;

    if (iz != 0) {
        f2 -= ((w0 + WW(1, 0, 0)) * (UU(1, 0, 0) * dz + UU(1, 0, -1) * dzm) - (w0 + WW(-1, 0, 0)) * (u0 * dz + UU(0, 0, -1) * dzm)) *
              m.inv_hx * inv_hz_alt;
        f2 -= ((w0 + WW(0, 1, 0)) * (VV(0, 1, 0) * dz + VV(0, 1, -1) * dzm) - (w0 + WW(0, -1, 0)) * (v0 * dz + VV(0, 0, -1) * dzm)) *
              m.inv_hy * inv_hz_alt;
        const T wr = iz == NZ - 1 ? T(0) : WW(0, 0, 1);
        f2 -= (w0 + wr) * (w0 + wr) * inv_hz_alt;
        f2 += (w0 + WW(0, 0, -1)) * (w0 + WW(0, 0, -1)) * inv_hz_alt;
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

template <typename T>
__device__ __forceinline__ T zbar_vv(const Mesh<T>& m, const T* __restrict__ vel, const Nb& nb, int c, int dx, int dy,
                                     int iz, const T* __restrict__ A) {
    // 9*(F(0)*A[iz] + F(-1)*A[iz-1]) - F(-2)*A[iz-2] - F(+1)*A[iz+1] with the wall rules
    const int NZ = m.nz;
    const T f0 = F3(vel, c, dx, dy, 0), fm = F3(vel, c, dx, dy, -1);
    T r = T(9) * (f0 * A[iz] + fm * A[iz - 1]);
    r -= (iz == 1) ? -fm * A[0] : F3(vel, c, dx, dy, -2) * A[iz - 2];
    r -= (iz == NZ - 1) ? -f0 * A[NZ - 1] : F3(vel, c, dx, dy, 1) * A[iz + 1];
    return r;
}

template <typename T>
__device__ __forceinline__ void conv_vv(const Mesh<T>& m, const T* __restrict__ vel, const Nb& nb, int iz, T f[3]) {
    const int NZ = m.nz;
    const T C27 = T(1.0 / 27.0);
    const T hx = m.hx, hy = m.hy, hxy = hx * hy;
    const T dz = m.DZ[iz], dzw = m.DZW[iz];
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    T f0 = 0, f1 = 0, f2 = 0;
    // ---- u-equation
    {
        const T br = dz * hy * (T(9) * (u0 + UU(1, 0, 0)) - (UU(-1, 0, 0) + UU(2, 0, 0)));
        const T bl = dz * hy * (T(9) * (UU(-1, 0, 0) + u0) - (UU(-2, 0, 0) + UU(1, 0, 0)));
        const T bbr = dzw * hy * (T(9) * (UU(1, 0, 0) + UU(2, 0, 0)) - (u0 + UU(3, 0, 0)));
        const T bbl = dzw * hy * (T(9) * (UU(-2, 0, 0) + UU(-1, 0, 0)) - (UU(-3, 0, 0) + u0));
        f0 += br * (u0 + UU(1, 0, 0)) - C27 * bbr * (u0 + UU(3, 0, 0)) - bl * (UU(-1, 0, 0) + u0) + C27 * bbl * (UU(-3, 0, 0) + u0);
    }
    {
        const T br = dz * hx * XBAR(VV, 1, 0);
        const T bl = dz * hx * XBAR(VV, 0, 0);
        const T bbr = dzw * hx * XBAR(VV, 2, 0);
        const T bbl = dzw * hx * XBAR(VV, -1, 0);
        f0 += br * (u0 + UU(0, 1, 0)) - C27 * bbr * (u0 + UU(0, 3, 0)) - bl * (UU(0, -1, 0) + u0) + C27 * bbl * (UU(0, -3, 0) + u0);
    }
    {
        const T bl = iz == 0 ? T(0) : hxy * XBAR(WW, 0, 0);
        const T br = iz == NZ - 1 ? T(0) : hxy * XBAR(WW, 0, 1);
        const T bbl = iz == 1 ? T(0) : iz == 0 ? br : hxy * XBAR(WW, 0, -1);
        const T bbr = iz == NZ - 2 ? T(0) : iz == NZ - 1 ? bl : hxy * XBAR(WW, 0, 2);
        const T q_r = iz == NZ - 1 ? T(0) : u0 + UU(0, 0, 1);
        const T q_rr = u0 + (iz == NZ - 3 ? -UU(0, 0, 2) : iz == NZ - 2 ? T(0) : iz == NZ - 1 ? UU(0, 0, -2) : UU(0, 0, 3));
        const T q_l = iz == 0 ? T(0) : UU(0, 0, -1) + u0;
        const T q_ll = u0 + (iz == 2 ? -UU(0, 0, -2) : iz == 1 ? T(0) : iz == 0 ? UU(0, 0, 2) : UU(0, 0, -3));
        f0 += br * q_r - C27 * bbr * q_rr - bl * q_l + C27 * bbl * q_ll;
    }
    // ---- v-equation
    {
        const T br = dz * hy * YBAR(UU, 1, 0);
        const T bl = dz * hy * YBAR(UU, 0, 0);
        const T bbr = dzw * hy * YBAR(UU, 2, 0);
        const T bbl = dzw * hy * YBAR(UU, -1, 0);
        f1 += br * (v0 + VV(1, 0, 0)) - C27 * bbr * (v0 + VV(3, 0, 0)) - bl * (VV(-1, 0, 0) + v0) + C27 * bbl * (VV(-3, 0, 0) + v0);
    }
    {
        const T br = dz * hx * (T(9) * (v0 + VV(0, 1, 0)) - (VV(0, -1, 0) + VV(0, 2, 0)));
        const T bl = dz * hx * (T(9) * (VV(0, -1, 0) + v0) - (VV(0, -2, 0) + VV(0, 1, 0)));
        const T bbr = dzw * hx * (T(9) * (VV(0, 1, 0) + VV(0, 2, 0)) - (v0 + VV(0, 3, 0)));
        const T bbl = dzw * hx * (T(9) * (VV(0, -2, 0) + VV(0, -1, 0)) - (VV(0, -3, 0) + v0));
        f1 += br * (v0 + VV(0, 1, 0)) - C27 * bbr * (v0 + VV(0, 3, 0)) - bl * (VV(0, -1, 0) + v0) + C27 * bbl * (VV(0, -3, 0) + v0);
    }
    {
        const T bl = iz == 0 ? T(0) : hxy * YBAR(WW, 0, 0);
        const T br = iz == NZ - 1 ? T(0) : hxy * YBAR(WW, 0, 1);
        const T bbl = iz == 1 ? T(0) : iz == 0 ? br : hxy * YBAR(WW, 0, -1);
        const T bbr = iz == NZ - 2 ? T(0) : iz == NZ - 1 ? bl : hxy * YBAR(WW, 0, 2);
        const T q_r = iz == NZ - 1 ? T(0) : v0 + VV(0, 0, 1);
        const T q_rr = v0 + (iz == NZ - 3 ? -VV(0, 0, 2) : iz == NZ - 2 ? T(0) : iz == NZ - 1 ? VV(0, 0, -2) : VV(0, 0, 3));
        const T q_l = iz == 0 ? T(0) : VV(0, 0, -1) + v0;
        const T q_ll = v0 + (iz == 2 ? -VV(0, 0, -2) : iz == 1 ? T(0) : iz == 0 ? VV(0, 0, 2) : VV(0, 0, -3));
        f1 += br * q_r - C27 * bbr * q_rr - bl * q_l + C27 * bbl * q_ll;
    }
    // ---- w-equation (no unknown at iz = 0)
    if (iz != 0) {
        {
            const T br = hy * zbar_vv(m, vel, nb, 0, 1, 0, iz, m.DZ);
            const T bl = hy * zbar_vv(m, vel, nb, 0, 0, 0, iz, m.DZ);
            const T bbr = hy * zbar_vv(m, vel, nb, 0, 2, 0, iz, m.DZW);
            const T bbl = hy * zbar_vv(m, vel, nb, 0, -1, 0, iz, m.DZW);
            f2 += br * (w0 + WW(1, 0, 0)) - C27 * bbr * (w0 + WW(3, 0, 0)) - bl * (WW(-1, 0, 0) + w0) + C27 * bbl * (WW(-3, 0, 0) + w0);
        }
        {
            // The reference scales this y-face flux by hy (staggered/func_global.h:475-487); kept.
            const T br = hy * zbar_vv(m, vel, nb, 1, 0, 1, iz, m.DZ);
            const T bl = hy * zbar_vv(m, vel, nb, 1, 0, 0, iz, m.DZ);
            const T bbr = hy * zbar_vv(m, vel, nb, 1, 0, 2, iz, m.DZW);
            const T bbl = hy * zbar_vv(m, vel, nb, 1, 0, -1, iz, m.DZW);
            f2 += br * (w0 + WW(0, 1, 0)) - C27 * bbr * (w0 + WW(0, 3, 0)) - bl * (WW(0, -1, 0) + w0) + C27 * bbl * (WW(0, -3, 0) + w0);
        }
        {
            const T wzl = WW(0, 0, -1);
            const T wzr = iz == NZ - 1 ? T(0) : WW(0, 0, 1);
            const T wzll = iz == 1 ? w0 : WW(0, 0, -2);
            const T wzrr = iz == NZ - 2 ? T(0) : iz == NZ - 1 ? w0 : WW(0, 0, 2);
            const T wzlll = iz == 1 ? WW(0, 0, 1) : iz == 2 ? WW(0, 0, -1) : WW(0, 0, -3);
            const T wzrrr = iz == NZ - 3 ? T(0) : iz == NZ - 2 ? WW(0, 0, 1) : iz == NZ - 1 ? WW(0, 0, -1) : WW(0, 0, 3);
            const T bl = hxy * (T(9) * (wzl + w0) - (wzll + wzr));
            const T br = hxy * (T(9) * (w0 + wzr) - (wzl + wzrr));
            const T bbl = hxy * (T(9) * (wzll + wzl) - (wzlll + w0));
            const T bbr = hxy * (T(9) * (wzr + wzrr) - (w0 + wzrrr));
            f2 += br * (w0 + wzr) - bl * (wzl + w0);
            if (iz < NZ - 2) f2 -= C27 * bbr * (w0 + wzrrr);
            else if (iz == NZ - 2) f2 -= C27 * bbr * (-WW(0, 0, 1) + w0);
            else f2 -= C27 * bbr * (WW(0, 0, -1) + w0);
            if (iz > 2) f2 += C27 * bbl * (wzlll + w0);
            else if (iz == 2) f2 += C27 * bbl * (-WW(0, 0, -1) + w0);
            else f2 += C27 * bbl * (WW(0, 0, 1) + w0);
        }
    }
    const T minus_inv_vol = T(-1) / (T(32) * (dz - T(1.0 / 9.0) * dzw) * hx * hy);
    f[0] = f0 * minus_inv_vol;
    f[1] = f1 * minus_inv_vol;
    f[2] = f2 * (-T(1. / 32.) * m.inv_hx * m.inv_hy * m.GCZ[iz]);
}

// ---------------------------------------------------------------------------
// Viscous term, constant viscosity (staggered/func_global.h:533-562)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void visc_cross(const Mesh<T>& m, const T* __restrict__ vel, const Nb& nb, int iz, T f[3]) {
    const int NZ = m.nz;
    const T dz = m.DZ[iz];
    const T inv_hz = T(1) / dz;
    const T inv_hz_alt = T(2) / (dz + (iz == 0 ? T(0) : m.DZ[iz - 1]));
    const T ihx2 = m.inv_hx * m.inv_hx, ihy2 = m.inv_hy * m.inv_hy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const T a0 = F3(vel, c, 0, 0, 0);
        f[c] = ihx2 * (F3(vel, c, 1, 0, 0) + F3(vel, c, -1, 0, 0) - T(2) * a0);
        f[c] += ihy2 * (F3(vel, c, 0, 1, 0) + F3(vel, c, 0, -1, 0) - T(2) * a0);
    }
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    if (iz != NZ - 1) {
        const T inv_hz_up = T(2) / (dz + m.DZ[iz + 1]);
        f[0] += (UU(0, 0, 1) - u0) * inv_hz_up * inv_hz;
        f[1] += (VV(0, 0, 1) - v0) * inv_hz_up * inv_hz;
        f[2] += (WW(0, 0, 1) - w0) * inv_hz * inv_hz_alt;
    } else {
        const T inv_hz_up = T(2) / dz;
        f[0] += (-u0) * inv_hz_up * inv_hz;
        f[1] += (-v0) * inv_hz_up * inv_hz;
        f[2] += (-w0) * inv_hz * inv_hz_alt;
    }
    if (iz != 0) {
        f[0] += (UU(0, 0, -1) - u0) * inv_hz_alt * inv_hz;
        f[1] += (VV(0, 0, -1) - v0) * inv_hz_alt * inv_hz;
        f[2] += (WW(0, 0, -1) - w0) * (T(1) / m.DZ[iz - 1]) * inv_hz_alt;
    } else {
        f[0] += (-u0) * inv_hz_alt * inv_hz;
        f[1] += (-v0) * inv_hz_alt * inv_hz;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] *= m.visc;
}

// ---------------------------------------------------------------------------
// Viscous term, div(2 (nu + nu_t) S) (staggered/func_global.h:565-625),
// including the edge-average of sxy_190 at :598 as the reference has it.
// ---------------------------------------------------------------------------
template <typename T, bool TURB>
__device__ __forceinline__ void visc_naive(const Mesh<T>& m, const T* __restrict__ vel, const T* __restrict__ nut,
                                           const Nb& nb, int iz, T f[3]) {
    const int NZ = m.nz;
    const T visc = m.visc;
    auto mu = [&](int dx, int dy, int dz) -> T { return TURB ? visc + nut[nb.at(dx, dy, dz)] : visc; };
    const T dz = m.DZ[iz];
    const T inv_hz = T(1) / dz;
    const T inv_hz_up = T(2) / (dz + (iz == NZ - 1 ? T(0) : m.DZ[iz + 1]));
    const T inv_hz_down = T(2) / (dz + (iz == 0 ? T(0) : m.DZ[iz - 1]));
    const T ihx = m.inv_hx, ihy = m.inv_hy;
    const T q = T(0.25);
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    const T mu0 = mu(0, 0, 0);
    T f0 = 0, f1 = 0, f2 = 0;
    const T sxx_000 = T(2) * (UU(1, 0, 0) - u0) * ihx * mu0;
    const T sxx_800 = T(2) * (u0 - UU(-1, 0, 0)) * ihx * mu(-1, 0, 0);
    f0 += (sxx_000 - sxx_800) * ihx;
    const T sxy_910 = ((UU(0, 1, 0) - u0) * ihy + (VV(0, 1, 0) - VV(-1, 1, 0)) * ihx) * q * (mu0 + mu(-1, 0, 0) + mu(0, 1, 0) + mu(-1, 1, 0));
    const T sxy_990 = ((u0 - UU(0, -1, 0)) * ihy + (v0 - VV(-1, 0, 0)) * ihx) * q * (mu(0, -1, 0) + mu(-1, -1, 0) + mu0 + mu(-1, 0, 0));
    f0 += (sxy_910 - sxy_990) * ihy;
    const T sxz_901 = (iz == NZ - 1) ? (T(-2) * u0 * inv_hz * visc)
                                     : ((UU(0, 0, 1) - u0) * inv_hz_up + (WW(0, 0, 1) - WW(-1, 0, 1)) * ihx) * q *
                                           (mu0 + mu(-1, 0, 0) + mu(0, 0, 1) + mu(-1, 0, 1));
    const T sxz_909 = (iz == 0) ? (T(2) * u0 * inv_hz * visc)
                                : ((u0 - UU(0, 0, -1)) * inv_hz_down + (w0 - WW(-1, 0, 0)) * ihx) * q *
                                      (mu(0, 0, -1) + mu(-1, 0, -1) + mu0 + mu(-1, 0, 0));
    f0 += (sxz_901 - sxz_909) * inv_hz;

    const T sxy_190 = ((UU(1, 0, 0) - UU(1, -1, 0)) * ihy + (VV(1, 0, 0) - v0) * ihx) * q * (mu0 + mu(1, 0, 0) + mu(0, -1, 0) + mu(1, 1, 0));
    f1 += (sxy_190 - sxy_990) * ihx;
    const T syy_000 = T(2) * (VV(0, 1, 0) - v0) * ihy * mu0;
    const T syy_080 = T(2) * (v0 - VV(0, -1, 0)) * ihy * mu(0, -1, 0);
    f1 += (syy_000 - syy_080) * ihy;
    const T syz_091 = (iz == NZ - 1) ? (T(-2) * v0 * inv_hz * visc)
                                     : ((VV(0, 0, 1) - v0) * inv_hz_up + (WW(0, 0, 1) - WW(0, -1, 1)) * ihy) * q *
                                           (mu0 + mu(0, -1, 0) + mu(0, 0, 1) + mu(0, -1, 1));
    const T syz_099 = (iz == 0) ? (T(2) * v0 * inv_hz * visc)
                                : ((v0 - VV(0, 0, -1)) * inv_hz_down + (w0 - WW(0, -1, 0)) * ihy) * q *
                                      (mu(0, 0, -1) + mu(0, -1, -1) + mu0 + mu(0, -1, 0));
    f1 += (syz_091 - syz_099) * inv_hz;

    if (iz != 0) {
        const T sxz_109 = ((UU(1, 0, 0) - UU(1, 0, -1)) * inv_hz_down + (WW(1, 0, 0) - w0) * ihx) * q *
                          (mu(0, 0, -1) + mu(1, 0, -1) + mu0 + mu(1, 0, 0));
        f2 += (sxz_109 - sxz_909) * ihx;
        const T syz_019 = ((VV(0, 1, 0) - VV(0, 1, -1)) * inv_hz_down + (WW(0, 1, 0) - w0) * ihy) * q *
                          (mu(0, 0, -1) + mu(0, 1, -1) + mu0 + mu(0, 1, 0));
        f2 += (syz_019 - syz_099) * ihy;
        const T szz_000 = T(2) * ((iz == NZ - 1 ? T(0) : WW(0, 0, 1)) - w0) * inv_hz * mu0;
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
template <typename T>
__device__ __forceinline__ void conv_ab3(const Mesh<T>& m, const T* __restrict__ vel, const Nb& nb, int iz, T inv_hz, T f[3]) {
    const T ih[3] = {m.inv_hx, m.inv_hy, inv_hz};
    const T u0[3] = {UU(0, 0, 0), VV(0, 0, 0), WW(0, 0, 0)};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int jn = nb.at(d == 0 ? -1 : 0, d == 1 ? -1 : 0, d == 2 ? -1 : 0);
        const int kn = nb.at(d == 0 ? 1 : 0, d == 1 ? 1 : 0, d == 2 ? 1 : 0);
        const T uj = vel[3 * jn + d], uk = vel[3 * kn + d];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T conv_div = T(0.5) * (vel[3 * kn + c] * uk - vel[3 * jn + c] * uj);
            const T conv_adv = u0[d] * T(0.5) * (vel[3 * kn + c] - vel[3 * jn + c]);
            f[c] -= T(0.5) * (conv_div + conv_adv) * ih[d];
        }
    }
}

// Q1-Galerkin viscous term with element nu_t (collocated_AB3/func_fluxes.h:21-85).
template <typename T, bool TURB>
__device__ __forceinline__ void visc_galerkin(const Mesh<T>& m, const T* __restrict__ vel, const T* __restrict__ nut,
                                              const Nb& nb, int iz, T g_out[3]) {
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
                    const int id = nb.at((k & 1) ? sx : 0, (k & 2) ? sy : 0, (k & 4) ? sz : 0);
                    U[k][0] = vel[3 * id];
                    U[k][1] = vel[3 * id + 1];
                    U[k][2] = vel[3 * id + 2];
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
                if (TURB) vc += nut[nb.at(ox ? 0 : -1, oy ? 0 : -1, oz ? 0 : -1)];
#pragma unroll
                for (int c = 0; c < 3; c++) sum[c] += g[c] * vc;
            }
        }
    }
    g_out[0] = sum[0];
    g_out[1] = sum[1];
    g_out[2] = sum[2];
}

// Gradient at one cell (staggered/func_global.h:297-303, :316-322; collocated_AB3/func_global.h:298-303).
template <typename T, int SCHEME>
__device__ __forceinline__ void grad_at(const Mesh<T>& m, const T* __restrict__ a, const Nb& nb, int iz, T g[3]) {
    const int NZ = m.nz;
    if (SCHEME == TC_SCHEME_HW) {
        const T a0 = a[nb.at(0, 0, 0)];
        g[0] = m.inv_hx * (a0 - a[nb.at(-1, 0, 0)]);
        g[1] = m.inv_hy * (a0 - a[nb.at(0, -1, 0)]);
        g[2] = iz == 0 ? T(0) : (a0 - a[nb.at(0, 0, -1)]) * (T(2) / (m.DZ[iz] + m.DZ[iz - 1]));
    } else if (SCHEME == TC_SCHEME_VV) {
        const T a0 = a[nb.at(0, 0, 0)];
        const T g1 = m.GC1[iz], g2 = m.GC2[iz];
        g[0] = m.inv_hx * (g1 * (a0 - a[nb.at(-1, 0, 0)]) + g2 * (a[nb.at(1, 0, 0)] - a[nb.at(-2, 0, 0)]));
        g[1] = m.inv_hy * (g1 * (a0 - a[nb.at(0, -1, 0)]) + g2 * (a[nb.at(0, 1, 0)] - a[nb.at(0, -2, 0)]));
        if (iz == 0) {
            g[2] = T(0);
        } else {
            const T am = a[nb.at(0, 0, -1)];
            const T ar = iz == NZ - 1 ? -a0 : a[nb.at(0, 0, 1)];
            const T all = iz == 1 ? -am : a[nb.at(0, 0, -2)];
            g[2] = ((a0 - am) - T(1.0 / 27.0) * (ar - all)) * m.GCZ[iz];
        }
    } else {
        if (iz == 0 || iz == NZ - 1) {
            g[0] = g[1] = g[2] = T(0);
            return;
        }
        g[0] = T(0.5) * m.inv_hx * (a[nb.at(1, 0, 0)] - a[nb.at(-1, 0, 0)]);
        g[1] = T(0.5) * m.inv_hy * (a[nb.at(0, 1, 0)] - a[nb.at(0, -1, 0)]);
        g[2] = T(0.5) * (a[nb.at(0, 0, 1)] - a[nb.at(0, 0, -1)]) / m.cellh[iz];
    }
}

// Volume of the streamwise-velocity unknown of one cell (flow-rate sums).
template <typename T, int SCHEME>
__device__ __forceinline__ double u_volume(const Mesh<T>& m, int iz) {
    if (SCHEME == TC_SCHEME_AB3) return (iz == 0 || iz == m.nz - 1) ? 0.0 : (double)m.cellh[iz] * (double)m.hx * (double)m.hy;
    return (double)m.DZ[iz] * (double)m.hx * (double)m.hy;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
    return r;
}

// ---------------------------------------------------------------------------
// K1: explicit term (+ BDF3/EX3 right-hand side)
// ---------------------------------------------------------------------------
template <typename T, int SCHEME, int VISC, bool TURB>
__global__ __launch_bounds__(kBlock) void k_flux(Mesh<T> m, FluxArgs<T> a) {
    __shared__ double sh[kBlock / 64];
    const int iz = blockIdx.y;
    int ix, iy;
    const bool active = cell_of_thread(m.nx, m.nxy, ix, iy);
    double cfr = 0.0;
    if (active) {
        Nb nb;
        make_nb<3>(nb, ix, iy, iz, m.nx, m.ny);
        const T* __restrict__ vel = a.vel;
        const int in = nb.at(0, 0, 0);
        T k[3] = {0, 0, 0};
        if (SCHEME == TC_SCHEME_AB3) {
            if (iz != 0 && iz != m.nz - 1) {
                const T inv_hz = T(2) / (m.DZ[iz - 1] + m.DZ[iz]);
                conv_ab3(m, vel, nb, iz, inv_hz, k);
                if (VISC == 1) {
                    T g[3];
                    visc_galerkin<T, TURB>(m, vel, a.nut, nb, iz, g);
                    const T s = m.inv_hx * m.inv_hy * inv_hz;
#pragma unroll
                    for (int c = 0; c < 3; c++) k[c] += g[c] * s;
                } else {
                    const T ihx2 = m.visc * m.inv_hx * m.inv_hx, ihy2 = m.visc * m.inv_hy * m.inv_hy;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const T a0 = F3(vel, c, 0, 0, 0);
                        k[c] += ihx2 * (F3(vel, c, 1, 0, 0) + F3(vel, c, -1, 0, 0) - T(2) * a0);
                        k[c] += ihy2 * (F3(vel, c, 0, 1, 0) + F3(vel, c, 0, -1, 0) - T(2) * a0);
                        k[c] += m.visc * ((F3(vel, c, 0, 0, 1) - a0) / m.DZ[iz] - (a0 - F3(vel, c, 0, 0, -1)) / m.DZ[iz - 1]) * inv_hz;
                    }
                }
                k[0] += m.sour;
            }
        } else {
            T c[3], v[3];
            if (SCHEME == TC_SCHEME_HW) conv_hw(m, vel, nb, iz, c);
            else conv_vv(m, vel, nb, iz, c);
            if (VISC == 1) visc_naive<T, TURB>(m, vel, a.nut, nb, iz, v);
            else visc_cross(m, vel, nb, iz, v);
            k[0] = (m.sour + c[0]) + v[0];
            k[1] = (T(0) + c[1]) + v[1];
            k[2] = iz == 0 ? T(0) : (T(0) + c[2]) + v[2];
        }
        T* fc = a.fcur + 3 * in;
        fc[0] = k[0];
        fc[1] = k[1];
        fc[2] = k[2];
        if (a.order > 0) {
            T r[3];
            const T it = a.inv_tau;
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const int j = 3 * in + cc;
                if (a.order == 1) r[cc] = k[cc] + it * vel[j];
                else if (a.order == 2) r[cc] = T(2) * k[cc] - a.fprev[j] + it * (T(2) * vel[j] - T(0.5) * a.vprev[j]);
                else
                    r[cc] = T(3) * k[cc] - T(3) * a.fprev[j] + a.fprev2[j] +
                            it * (T(3) * vel[j] - T(1.5) * a.vprev[j] + T(1.0 / 3.0) * a.vprev2[j]);
            }
            if (SCHEME == TC_SCHEME_AB3 && a.stab) {
                T g[3];
                grad_at<T, TC_SCHEME_AB3>(m, a.pcur, nb, iz, g);
                r[0] -= g[0];
                r[1] -= g[1];
                r[2] -= g[2];
            }
            T* fn = a.fnext + 3 * in;
            fn[0] = r[0];
            fn[1] = r[1];
            fn[2] = r[2];
            if (a.cfr_partial) cfr = (double)r[0] * u_volume<T, SCHEME>(m, iz);
        }
    }
    if (a.cfr_partial) {
        const double s = block_sum(cfr, sh);
        if (threadIdx.x == 0) a.cfr_partial[blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Turbulent viscosity
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_turb_staggered(Mesh<T> m, const T* __restrict__ vel, T* __restrict__ nut) {
    // CalcAbsS + RecalcTurbVisc (staggered/func_global.h:823-890)
    const int iz = blockIdx.y, NZ = m.nz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    Nb nb;
    make_nb<1>(nb, ix, iy, iz, m.nx, m.ny);
    const T dz = m.DZ[iz];
    const T inv_hz = T(1) / dz;
    const T inv_hz_up = T(2) / (dz + (iz == NZ - 1 ? T(0) : m.DZ[iz + 1]));
    const T inv_hz_down = T(2) / (dz + (iz == 0 ? T(0) : m.DZ[iz - 1]));
    const T ihx = m.inv_hx, ihy = m.inv_hy;
    const T u0 = UU(0, 0, 0), v0 = VV(0, 0, 0), w0 = WW(0, 0, 0);
    T S;
    {
        const T sxx = (UU(1, 0, 0) - u0) * ihx;
        const T syy = (VV(0, 1, 0) - v0) * ihy;
        const T szz = ((iz == NZ - 1 ? T(0) : WW(0, 0, 1)) - w0) * inv_hz;
        S = sxx * sxx + syy * syy + szz * szz;
    }
    {
        const T a1 = (UU(0, 1, 0) - u0) * ihy + (VV(0, 1, 0) - VV(-1, 1, 0)) * ihx;
        const T a2 = (u0 - UU(0, -1, 0)) * ihy + (v0 - VV(-1, 0, 0)) * ihx;
        const T a3 = (UU(1, 1, 0) - UU(1, 0, 0)) * ihy + (VV(1, 1, 0) - VV(0, 1, 0)) * ihx;
        const T a4 = (UU(1, 0, 0) - UU(1, -1, 0)) * ihy + (VV(1, 0, 0) - v0) * ihx;
        S += T(0.125) * (a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4);
    }
    {
        const T a1 = (iz == NZ - 1) ? (T(-2) * u0 * inv_hz) : (UU(0, 0, 1) - u0) * inv_hz_up + (WW(0, 0, 1) - WW(-1, 0, 1)) * ihx;
        const T a2 = (iz == 0) ? (T(2) * u0 * inv_hz) : (u0 - UU(0, 0, -1)) * inv_hz_down + (w0 - WW(-1, 0, 0)) * ihx;
        const T a3 = (iz == NZ - 1) ? (T(-2) * UU(1, 0, 0) * inv_hz) : (UU(1, 0, 1) - UU(1, 0, 0)) * inv_hz_up + (WW(1, 0, 1) - WW(0, 0, 1)) * ihx;
        const T a4 = (iz == 0) ? (T(2) * UU(1, 0, 0) * inv_hz) : (UU(1, 0, 0) - UU(1, 0, -1)) * inv_hz_down + (WW(1, 0, 0) - w0) * ihx;
        S += T(0.125) * (a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4);
    }
    {
        const T a1 = (iz == NZ - 1) ? (T(-2) * v0 * inv_hz) : (VV(0, 0, 1) - v0) * inv_hz_up + (WW(0, 0, 1) - WW(0, -1, 1)) * ihy;
        const T a2 = (iz == 0) ? (T(2) * v0 * inv_hz) : (v0 - VV(0, 0, -1)) * inv_hz_down + (w0 - WW(0, -1, 0)) * ihy;
        const T a3 = (iz == NZ - 1) ? (T(-2) * VV(0, 1, 0) * inv_hz) : (VV(0, 1, 1) - VV(0, 1, 0)) * inv_hz_up + (WW(0, 1, 1) - WW(0, 0, 1)) * ihy;
        const T a4 = (iz == 0) ? (T(2) * VV(0, 1, 0) * inv_hz) : (VV(0, 1, 0) - VV(0, 1, -1)) * inv_hz_down + (WW(0, 1, 0) - w0) * ihy;
        S += T(0.125) * (a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4);
    }
    const T absS = sqrt(T(2) * S);
    const T H = T(0.5) * (m.zhi - m.zlo);
    const T Z = T(0.5) * (m.ZZ[iz] + m.ZZ[iz + 1]) - m.zlo;
    const T dist = (Z < H) ? Z : T(2) * H - Z;
    const T yp = dist * m.inv_len;
    const T fvd = T(1) - exp(-T(0.04 * 0.04 * 0.04) * yp * yp * yp);
    T delta = T(0.1) * pow(m.hx * m.hy * dz, T(1.0 / 3.0));
    if (delta > T(0.41) * dist) delta = T(0.41) * dist;
    nut[nb.at(0, 0, 0)] = fvd * delta * delta * absS;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_turb_ab3(Mesh<T> m, const T* __restrict__ vel, T* __restrict__ nut) {
    // element strain rate (collocated_AB3/func_global.h:342-384); plane nz-1 has no element
    const int iz = blockIdx.y, NZ = m.nz;
    int ix, iy;
    if (iz >= NZ - 1 || !cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    Nb nb;
    make_nb<1>(nb, ix, iy, iz, m.nx, m.ny);
    const T hz = m.DZ[iz];
    T gu[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const T p000 = F3(vel, c, 0, 0, 0), p100 = F3(vel, c, 1, 0, 0), p010 = F3(vel, c, 0, 1, 0), p110 = F3(vel, c, 1, 1, 0);
        const T p001 = F3(vel, c, 0, 0, 1), p101 = F3(vel, c, 1, 0, 1), p011 = F3(vel, c, 0, 1, 1), p111 = F3(vel, c, 1, 1, 1);
        gu[0][c] = ((p100 + p110 + p101 + p111) - (p000 + p010 + p001 + p011)) * (T(0.25) * m.inv_hx);
        gu[1][c] = ((p010 + p110 + p011 + p111) - (p000 + p100 + p001 + p101)) * (T(0.25) * m.inv_hy);
        gu[2][c] = ((p001 + p101 + p011 + p111) - (p000 + p100 + p010 + p110)) / (T(4) * hz);
    }
    gu[0][1] = T(0.5) * (gu[0][1] + gu[1][0]);
    gu[0][2] = T(0.5) * (gu[0][2] + gu[2][0]);
    gu[1][2] = T(0.5) * (gu[1][2] + gu[2][1]);
    T S = gu[0][0] * gu[0][0] + gu[1][1] * gu[1][1] + gu[2][2] * gu[2][2] +
          T(2) * (gu[0][1] * gu[0][1] + gu[0][2] * gu[0][2] + gu[1][2] * gu[1][2]);
    S = sqrt(T(2) * S);
    const T H = T(0.5) * (m.zhi - m.zlo);
    const T Z = T(0.5) * (m.ZZ[iz] + m.ZZ[iz + 1]);
    const T dist = (Z < H) ? Z : T(2) * H - Z;
    const T yp = dist * m.inv_len;
    const T fvd = T(1) - exp(-T(0.04 * 0.04 * 0.04) * yp * yp * yp);
    T delta = T(0.1) * pow(m.hx * m.hy * hz, T(1.0 / 3.0));
    if (delta > T(0.41) * dist) delta = T(0.41) * dist;
    nut[nb.at(0, 0, 0)] = fvd * delta * delta * S;
}

// ---------------------------------------------------------------------------
// Divergence (staggered/func_global.h:306-312, :325-333; collocated_AB3/func_global.h:310-325)
// ---------------------------------------------------------------------------
template <typename T, int SCHEME>
__global__ __launch_bounds__(kBlock) void k_div(Mesh<T> m, const T* __restrict__ v, T* __restrict__ dv) {
    const int iz = blockIdx.y, NZ = m.nz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    Nb nb;
    make_nb<2>(nb, ix, iy, iz, m.nx, m.ny);
    const int in = nb.at(0, 0, 0);
    T d;
    if (SCHEME == TC_SCHEME_HW) {
        const T hz = m.DZ[iz];
        d = (v[3 * nb.at(1, 0, 0)] - v[3 * in]) * m.inv_hx * hz + (v[3 * nb.at(0, 1, 0) + 1] - v[3 * in + 1]) * m.inv_hy * hz +
            ((iz < NZ - 1) ? (v[3 * nb.at(0, 0, 1) + 2] - v[3 * in + 2]) : (-v[3 * in + 2]));
    } else if (SCHEME == TC_SCHEME_VV) {
        const T dz = m.DZ[iz], dzw = m.DZW[iz], c27 = T(1.0 / 27.0);
        const T dudx = m.inv_hx * (dz * (v[3 * nb.at(1, 0, 0)] - v[3 * in]) - c27 * dzw * (v[3 * nb.at(2, 0, 0)] - v[3 * nb.at(-1, 0, 0)]));
        const T dvdy = m.inv_hy * (dz * (v[3 * nb.at(0, 1, 0) + 1] - v[3 * in + 1]) -
                                   c27 * dzw * (v[3 * nb.at(0, 2, 0) + 1] - v[3 * nb.at(0, -1, 0) + 1]));
        const T w0 = v[3 * in + 2];
        const T wr = (iz == NZ - 1) ? T(0) : v[3 * nb.at(0, 0, 1) + 2];
        const T wrr = (iz == NZ - 1) ? w0 : (iz == NZ - 2) ? T(0) : v[3 * nb.at(0, 0, 2) + 2];
        const T wll = (iz == 0) ? v[3 * nb.at(0, 0, 1) + 2] : (iz == 1) ? T(0) : v[3 * nb.at(0, 0, -1) + 2];
        d = dudx + dvdy + (wr - w0) - c27 * (wrr - wll);
    } else {
        const T hz = m.cellh[iz];
        d = T(0.5) * (v[3 * nb.at(1, 0, 0)] - v[3 * nb.at(-1, 0, 0)]) * m.hy * hz;
        d += T(0.5) * (v[3 * nb.at(0, 1, 0) + 1] - v[3 * nb.at(0, -1, 0) + 1]) * m.hx * hz;
        if (iz == 0) d += T(0.5) * (v[3 * nb.at(0, 0, 1) + 2] - v[3 * in + 2]) * m.hx * m.hy;
        else if (iz == NZ - 1) d += T(0.5) * (v[3 * in + 2] - v[3 * nb.at(0, 0, -1) + 2]) * m.hx * m.hy;
        else d += T(0.5) * (v[3 * nb.at(0, 0, 1) + 2] - v[3 * nb.at(0, 0, -1) + 2]) * m.hx * m.hy;
    }
    dv[in] = d;
}

template <typename T, int SCHEME>
__global__ __launch_bounds__(kBlock) void k_grad(Mesh<T> m, const T* __restrict__ p, T* __restrict__ g) {
    const int iz = blockIdx.y;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    Nb nb;
    make_nb<2>(nb, ix, iy, iz, m.nx, m.ny);
    T r[3];
    grad_at<T, SCHEME>(m, p, nb, iz, r);
    T* o = g + 3 * nb.at(0, 0, 0);
    o[0] = r[0];
    o[1] = r[1];
    o[2] = r[2];
}

// K7: velocity correction (staggered/func_global.h:937-941; collocated_AB3/func_global.h:428-449)
template <typename T, int SCHEME, bool STAB>
__global__ __launch_bounds__(kBlock) void k_project(Mesh<T> m, ProjectArgs<T> a) {
    const int iz = blockIdx.y;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    Nb nb;
    make_nb<2>(nb, ix, iy, iz, m.nx, m.ny);
    const int in = nb.at(0, 0, 0);
    T g[3];
    grad_at<T, SCHEME>(m, a.p, nb, iz, g);
    const T* r = a.r + 3 * in;
    T* o = a.vnext + 3 * in;
    T u = (r[0] - g[0]) * a.tau_over_bd0;
    if (a.shift && !(SCHEME == TC_SCHEME_AB3 && (iz == 0 || iz == m.nz - 1))) u += T(*a.shift);
    o[0] = u;
    o[1] = (r[1] - g[1]) * a.tau_over_bd0;
    o[2] = (r[2] - g[2]) * a.tau_over_bd0;
    if (STAB) a.pnext[in] = a.p[in] + a.pcur[in];
}

template <typename T, int SCHEME>
__global__ __launch_bounds__(kBlock) void k_add_u(Mesh<T> m, T* __restrict__ v, const double* __restrict__ shift) {
    const int iz = blockIdx.y;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    if (SCHEME == TC_SCHEME_AB3 && (iz == 0 || iz == m.nz - 1)) return;
    v[3 * (iz * m.nxy + iy * m.nx + ix)] += T(*shift);
}

// f[in] -= (sum * coef) * w[iz] (w == nullptr: w = 1); used by NormalizePressure.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_shift_plane(Mesh<T> m, T* __restrict__ f, const double* __restrict__ sum,
                                                         double coef, const T* __restrict__ w) {
    const int iz = blockIdx.y;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy)) return;
    const int in = iz * m.nxy + iy * m.nx + ix;
    const T s = T(sum[0] * coef);
    f[in] -= w ? s * w[iz] : s;
}

template <typename T>
dim3 cell_grid(const Mesh<T>& m) {
    return dim3((m.nxy + kBlock - 1) / kBlock, m.nz);
}

}  // namespace

int flux_grid_blocks(int nxy, int nz) { return ((nxy + kBlock - 1) / kBlock) * nz; }

template <typename T>
hipError_t launch_flux(hipStream_t s, const Mesh<T>& m, const FluxArgs<T>& a) {
    const dim3 g = cell_grid(m);
    if (a.scheme == TC_SCHEME_HW) {
        if (a.visc_type == 1) {
            if (a.turb) hipLaunchKernelGGL((k_flux<T, TC_SCHEME_HW, 1, true>), g, dim3(kBlock), 0, s, m, a);
            else hipLaunchKernelGGL((k_flux<T, TC_SCHEME_HW, 1, false>), g, dim3(kBlock), 0, s, m, a);
        } else {
            hipLaunchKernelGGL((k_flux<T, TC_SCHEME_HW, 0, false>), g, dim3(kBlock), 0, s, m, a);
        }
    } else if (a.scheme == TC_SCHEME_VV) {
        if (a.visc_type == 1) {
            if (a.turb) hipLaunchKernelGGL((k_flux<T, TC_SCHEME_VV, 1, true>), g, dim3(kBlock), 0, s, m, a);
            else hipLaunchKernelGGL((k_flux<T, TC_SCHEME_VV, 1, false>), g, dim3(kBlock), 0, s, m, a);
        } else {
            hipLaunchKernelGGL((k_flux<T, TC_SCHEME_VV, 0, false>), g, dim3(kBlock), 0, s, m, a);
        }
    } else {
        if (a.visc_type == 1) {
            if (a.turb) hipLaunchKernelGGL((k_flux<T, TC_SCHEME_AB3, 1, true>), g, dim3(kBlock), 0, s, m, a);
            else hipLaunchKernelGGL((k_flux<T, TC_SCHEME_AB3, 1, false>), g, dim3(kBlock), 0, s, m, a);
        } else {
            hipLaunchKernelGGL((k_flux<T, TC_SCHEME_AB3, 0, false>), g, dim3(kBlock), 0, s, m, a);
        }
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_turb_visc(hipStream_t s, const Mesh<T>& m, int scheme, const T* vel, T* nut) {
    const dim3 g = cell_grid(m);
    if (scheme == TC_SCHEME_AB3) hipLaunchKernelGGL((k_turb_ab3<T>), g, dim3(kBlock), 0, s, m, vel, nut);
    else hipLaunchKernelGGL((k_turb_staggered<T>), g, dim3(kBlock), 0, s, m, vel, nut);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_divergence(hipStream_t s, const Mesh<T>& m, int scheme, const T* v, T* dv) {
    const dim3 g = cell_grid(m);
    if (scheme == TC_SCHEME_HW) hipLaunchKernelGGL((k_div<T, TC_SCHEME_HW>), g, dim3(kBlock), 0, s, m, v, dv);
    else if (scheme == TC_SCHEME_VV) hipLaunchKernelGGL((k_div<T, TC_SCHEME_VV>), g, dim3(kBlock), 0, s, m, v, dv);
    else hipLaunchKernelGGL((k_div<T, TC_SCHEME_AB3>), g, dim3(kBlock), 0, s, m, v, dv);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_gradient(hipStream_t s, const Mesh<T>& m, int scheme, const T* p, T* gr) {
    const dim3 g = cell_grid(m);
    if (scheme == TC_SCHEME_HW) hipLaunchKernelGGL((k_grad<T, TC_SCHEME_HW>), g, dim3(kBlock), 0, s, m, p, gr);
    else if (scheme == TC_SCHEME_VV) hipLaunchKernelGGL((k_grad<T, TC_SCHEME_VV>), g, dim3(kBlock), 0, s, m, p, gr);
    else hipLaunchKernelGGL((k_grad<T, TC_SCHEME_AB3>), g, dim3(kBlock), 0, s, m, p, gr);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_project(hipStream_t s, const Mesh<T>& m, const ProjectArgs<T>& a) {
    const dim3 g = cell_grid(m);
    if (a.scheme == TC_SCHEME_HW) hipLaunchKernelGGL((k_project<T, TC_SCHEME_HW, false>), g, dim3(kBlock), 0, s, m, a);
    else if (a.scheme == TC_SCHEME_VV) hipLaunchKernelGGL((k_project<T, TC_SCHEME_VV, false>), g, dim3(kBlock), 0, s, m, a);
    else if (a.pnext) hipLaunchKernelGGL((k_project<T, TC_SCHEME_AB3, true>), g, dim3(kBlock), 0, s, m, a);
    else hipLaunchKernelGGL((k_project<T, TC_SCHEME_AB3, false>), g, dim3(kBlock), 0, s, m, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add_u(hipStream_t s, const Mesh<T>& m, int scheme, T* v, const double* shift) {
    const dim3 g = cell_grid(m);
    if (scheme == TC_SCHEME_AB3) hipLaunchKernelGGL((k_add_u<T, TC_SCHEME_AB3>), g, dim3(kBlock), 0, s, m, v, shift);
    else hipLaunchKernelGGL((k_add_u<T, TC_SCHEME_HW>), g, dim3(kBlock), 0, s, m, v, shift);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add_scaled_plane(hipStream_t s, const Mesh<T>& m, T* f, const double* sum, double coef, const T* plane_w) {
    hipLaunchKernelGGL((k_shift_plane<T>), cell_grid(m), dim3(kBlock), 0, s, m, f, sum, coef, plane_w);
    return hipGetLastError();
}

#define INST(T)                                                                                             \
    template hipError_t launch_flux<T>(hipStream_t, const Mesh<T>&, const FluxArgs<T>&);                    \
    template hipError_t launch_turb_visc<T>(hipStream_t, const Mesh<T>&, int, const T*, T*);                \
    template hipError_t launch_divergence<T>(hipStream_t, const Mesh<T>&, int, const T*, T*);               \
    template hipError_t launch_gradient<T>(hipStream_t, const Mesh<T>&, int, const T*, T*);                 \
    template hipError_t launch_project<T>(hipStream_t, const Mesh<T>&, const ProjectArgs<T>&);              \
    template hipError_t launch_add_u<T>(hipStream_t, const Mesh<T>&, int, T*, const double*);               \
    template hipError_t launch_add_scaled_plane<T>(hipStream_t, const Mesh<T>&, T*, const double*, double, const T*);
INST(double)
INST(float)
#undef INST

}  // namespace tc
