// Stencil kernels: explicit term + BDF3/EX3 right-hand side, turbulent
// viscosity, divergence, gradient and the velocity correction.
//
// Every block works on one wall-normal plane at a time, so every branch on iz
// (the near-wall closures of the fourth-order scheme) is uniform across a
// wavefront and every mesh coefficient is a scalar load.
#include <algorithm>
#include <type_traits>

#include <type_traits>
#include "diag.h"
#include "kernels.h"
#include "stencil.h"

namespace tc {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// A value no other thread of the kernel reads (BDF history, the projection's
// right-hand side). FP32: a streaming load, so it does not displace the rows that
// neighbouring blocks read as halos or taps (VV 192x160x128 step 0.233 -> 0.224
// ms); in FP64 the same hint made the flux and projection kernels slower.
template <typename T>
__device__ __forceinline__ T load_once(const T* p) {
    if constexpr (sizeof(T) == 4) return __builtin_nontemporal_load(p);
    else return *p;
}

// Periodic neighbours of one cell: linear x index and y row offset for -3..3.
struct Nb {
    int x[7], y[7], base, nxy;
    __device__ __forceinline__ int at(int dx, int dy, int dz) const { return base + dz * nxy + y[dy + 3] + x[dx + 3]; }
};

// Kernels over nodes launch a (plane chunk, plane) grid; with swz the logical
// block is XCD-swizzled so that each XCD walks a contiguous range of planes.
struct PlaneBlock {
    int bx, iz;
};
__device__ __forceinline__ PlaneBlock plane_block(int swz) {
    if (!swz) return {(int)blockIdx.x, (int)blockIdx.y};
    const int lb = xcd_swizzle(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    return {lb % (int)gridDim.x, lb / (int)gridDim.x};
}

__device__ __forceinline__ bool cell_of_thread(int nx, int nxy, int& ix, int& iy, int bx) {
    const int t = bx * kBlock + threadIdx.x;
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

// z offset dz from plane ib, clamped into the grid: every load is valid, and the
// wall rules select away what a clamped load reads (no branch around a load).
__device__ __forceinline__ int zclamp(int ib, int dz, int nz) { return min(max(dz, -ib), nz - 1 - ib); }

// z component of the staggered gradient at the cell dzo planes above nb's
// cell; iz is that cell's plane (staggered/func_global.h:302, :319-321). Past
// the top plane (iz == nz) the value is meaningless; the caller discards it.
template <typename T, int SCHEME>
__device__ __forceinline__ T grad_z_stag(const Mesh<T>& m, const T* __restrict__ a, const Nb& nb, int iz, int dzo) {
    const int NZ = m.nz, ib = iz - dzo, it = min(iz, NZ - 1);
    const T a0 = a[nb.at(0, 0, zclamp(ib, dzo, NZ))];
    const T am = a[nb.at(0, 0, zclamp(ib, dzo - 1, NZ))];
    T g;
    if (SCHEME == TC_SCHEME_HW) {
        g = (a0 - am) * m.IHDN[it];  // 2 / (DZ[iz] + DZ[iz-1])
    } else {
        const T arl = a[nb.at(0, 0, zclamp(ib, dzo + 1, NZ))];
        const T alll = a[nb.at(0, 0, zclamp(ib, dzo - 2, NZ))];
        const T ar = iz == NZ - 1 ? -a0 : arl;
        const T all = iz == 1 ? -am : alll;
        g = ((a0 - am) - T(1.0 / 27.0) * (ar - all)) * m.GCZ[it];
    }
    return iz == 0 ? T(0) : g;
}

// Gradient at one cell (staggered/func_global.h:297-303, :316-322; collocated_AB3/func_global.h:298-303).
template <typename T, int SCHEME>
__device__ __forceinline__ void grad_at(const Mesh<T>& m, const T* __restrict__ a, const Nb& nb, int iz, T g[3]) {
    const int NZ = m.nz;
    if (SCHEME == TC_SCHEME_HW) {
        const T a0 = a[nb.at(0, 0, 0)];
        g[0] = m.inv_hx * (a0 - a[nb.at(-1, 0, 0)]);
        g[1] = m.inv_hy * (a0 - a[nb.at(0, -1, 0)]);
        g[2] = grad_z_stag<T, SCHEME>(m, a, nb, iz, 0);
    } else if (SCHEME == TC_SCHEME_VV) {
        const T a0 = a[nb.at(0, 0, 0)];
        const T g1 = m.GC1[iz], g2 = m.GC2[iz];
        g[0] = m.inv_hx * (g1 * (a0 - a[nb.at(-1, 0, 0)]) + g2 * (a[nb.at(1, 0, 0)] - a[nb.at(-2, 0, 0)]));
        g[1] = m.inv_hy * (g1 * (a0 - a[nb.at(0, -1, 0)]) + g2 * (a[nb.at(0, 1, 0)] - a[nb.at(0, -2, 0)]));
        g[2] = grad_z_stag<T, SCHEME>(m, a, nb, iz, 0);
    } else {
        const bool wall = iz == 0 || iz == NZ - 1;
        const T gx = T(0.5) * m.inv_hx * (a[nb.at(1, 0, 0)] - a[nb.at(-1, 0, 0)]);
        const T gy = T(0.5) * m.inv_hy * (a[nb.at(0, 1, 0)] - a[nb.at(0, -1, 0)]);
        const T gz = T(0.5) * (a[nb.at(0, 0, zclamp(iz, 1, NZ))] - a[nb.at(0, 0, zclamp(iz, -1, NZ))]) / m.cellh[iz];
        g[0] = wall ? T(0) : gx;
        g[1] = wall ? T(0) : gy;
        g[2] = wall ? T(0) : gz;
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
// K1: explicit term (+ BDF3/EX3 right-hand side), 2.5D-blocked.
// A block owns a kTX x TY tile of (x, y) columns and marches along z over a
// chunk of planes. The velocity planes iz-RZ..iz+RZ of the tile plus a halo of
// RXY cells sit in an LDS ring (real3 AoS); RZ covers the stencils' reach in z
// (3 for the fourth-order scheme, whose z+-3 terms read the cell's own column
// only). Blocks are ordered so that each XCD works on neighbouring tiles and
// shares their halos in its L2.
// ---------------------------------------------------------------------------
constexpr int kTX = 32;
constexpr int kFluxMeshArrays = 10;  // DZ, DZW, GCZ, cellh and the six per-plane quotients
// Tile height (one thread per column of the tile).
template <typename T, int SCHEME, int VISC>
constexpr int flux_ty() {
    return 8;
}
// Threads per cell: in FP64 the fourth-order scheme gives each momentum
// component its own wave (the component is uniform across a wave), so that a
// thread holds one equation's registers and three waves instead of one share a
// SIMD to hide LDS latency. In FP32 the whole cell fits twice per SIMD already,
// and the split only adds waiting at the per-plane barrier for the costlier
// w-equation.
template <typename T, int SCHEME>
constexpr int flux_nc() {
    return SCHEME == TC_SCHEME_VV && sizeof(T) == 8 ? 3 : 1;
}
template <typename T, int SCHEME, int VISC>
constexpr int flux_threads() {
    return kTX * flux_ty<T, SCHEME, VISC>() * flux_nc<T, SCHEME>();
}
// Planes of the LDS ring either side of the computed one. The fourth-order
// scheme reaches z+-3 on the cell's own column only; a ring of +-3 planes holds
// those too, so that no global load sits in the stencil.
template <typename T, int SCHEME>
constexpr int flux_rz() {
    return SCHEME == TC_SCHEME_VV ? 3 : 1;
}
// Waves per SIMD the register allocation must allow.
template <typename T, int SCHEME, int VISC>
constexpr int flux_min_waves() {
    return 1;
}


// nu_t at an offset; an offset past a wall reads the wall plane (the stencils
// read such values unconditionally and discard them).
template <typename T>
struct GlobalNu {
    const T* nut;
    const Nb* nb;
    int iz, nz;
    __device__ __forceinline__ T operator()(int dx, int dy, int dz) const {
        return nut[nb->at(dx, dy, min(max(iz + dz, 0), nz - 1) - iz)];
    }
};

// The unit of the LDS ring fill: one real (4- or 8-byte loads and stores).
template <typename T>
using RingWord = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;

// Ring accessor. With a ring narrower than the fourth-order reach (RZ < 3)
// the own column's values at dz = +-(RZ+1) come from registers (cp, cm).
template <typename T, int RXY, int RZ, int TY>
struct RingAcc {
    static constexpr int W = kTX + 2 * RXY, H = TY + 2 * RXY, PS = W * H;
    // LDS planes iz-RZ..iz+RZ of the padded tile (real3 AoS), each pointing at the
    // corner (dx, dy) = (-RXY, -RXY) of the cell's window, so that every offset
    // is non-negative and becomes an immediate of the LDS read
    const T* pl[2 * RZ + 1];
    T cp[3], cm[3];  // own column at dz = +(RZ+1) and -(RZ+1)
    __device__ __forceinline__ T operator()(int c, int dx, int dy, int dz) const {
        if (dz == RZ + 1) return cp[c];
        if (dz == -RZ - 1) return cm[c];
        return pl[dz + RZ][3 * ((dy + RXY) * W + dx + RXY) + c];
    }
};

// Learned closure (new numerics, no reference counterpart): k -= d tau_ij / d x_j at
// an interior collocated node, central differences in x, y and the 3-point
// formula on the nonuniform z nodes; tau is [node][XX YY ZZ XY XZ YZ].
// The 21 stress values it needs are loaded ahead (sgs_load) into q[c][7] =
// (x+, x-, y+, y-, z+, own, z-) of component c's x-, y- and z-stresses.
template <typename T>
__device__ __forceinline__ void sgs_load(const T* __restrict__ t, const Nb& nb, int iz, int nz, T q[3][7]) {
    constexpr int ax[3] = {0, 3, 4}, ay[3] = {3, 1, 5}, az[3] = {4, 5, 2};
    const int dzp = iz < nz - 1 ? 1 : 0, dzm = iz > 0 ? -1 : 0;  // the wall planes never use these
    const int c0 = 6 * nb.at(0, 0, 0), xr = 6 * nb.at(1, 0, 0), xl = 6 * nb.at(-1, 0, 0);
    const int yr = 6 * nb.at(0, 1, 0), yl = 6 * nb.at(0, -1, 0), zr = 6 * nb.at(0, 0, dzp), zl = 6 * nb.at(0, 0, dzm);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        q[c][0] = t[xr + ax[c]];
        q[c][1] = t[xl + ax[c]];
        q[c][2] = t[yr + ay[c]];
        q[c][3] = t[yl + ay[c]];
        q[c][4] = t[zr + az[c]];
        q[c][5] = t[c0 + az[c]];
        q[c][6] = t[zl + az[c]];
    }
}
template <typename T>
__device__ __forceinline__ void sgs_divergence(const Mesh<T>& m, const T q[3][7], int iz, T k[3]) {
    const T hm = m.DZ[iz - 1], hp = m.DZ[iz];
    const T den = hp * hm * (hp + hm);
    const T cp = hm * hm / den, cm = hp * hp / den;
    const T hx2 = T(0.5) * m.inv_hx, hy2 = T(0.5) * m.inv_hy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        T d = hx2 * (q[c][0] - q[c][1]);
        d += hy2 * (q[c][2] - q[c][3]);
        const T t0 = q[c][5];
        d += cp * (q[c][4] - t0) + cm * (t0 - q[c][6]);
        k[c] -= d;
    }
}

// The plane loop of K1 for one block; COMP is the momentum component of the
// calling wave when a cell's three equations go to three waves (else -1), a
// compile-time value so that each wave's registers hold only its own equation.
// ALLINT: every plane of the block's chunk lies in 3..nz-4, so the staggered
// stencils run with their wall tests folded away.
template <typename T, int SCHEME, int VISC, bool TURB, int COMP, bool ALLINT = false>
__device__ __forceinline__ void flux_planes(const Mesh<T>& m, const FluxArgs<T>& a, int ntx, int ntiles, int zchunk,
                                            T* ring, double* sh) {
    constexpr int TY = flux_ty<T, SCHEME, VISC>();
    constexpr int RXY = SCHEME == TC_SCHEME_VV ? 3 : 1;
    constexpr int RZ = flux_rz<T, SCHEME>();
    constexpr bool COLUMN = SCHEME == TC_SCHEME_VV && RZ < 3;  // z+-3 of the own column from global memory
    constexpr int NS = 2 * RZ + 2;  // planes iz-RZ..iz+RZ plus the one in flight
    using Acc = RingAcc<T, RXY, RZ, TY>;
    constexpr int W = Acc::W, PS = Acc::PS;
    constexpr int NC = flux_nc<T, SCHEME>();
    constexpr int NT = kTX * TY * NC;
    using Word = RingWord<T>;                   // one real: a fill moves whole reals
    constexpr int DPC = 3;                      // words per real3 cell
    constexpr int DPS = PS * DPC;               // words per plane of the padded tile
    constexpr int ND = (DPS + NT - 1) / NT;
    // slot: the plane's words, then 64 spare words that the lanes past the plane
    // write (one per lane position, so no two lanes of a wave collide); every
    // lane writes ND words per fill, the same instruction stream for every plane
    constexpr int SLOT = (DPS + 64 + 1) / 2 * 2;  // reals per ring slot (16-byte multiple)
    const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tile = lb % ntiles, chunk = lb / ntiles;
    const int x0 = (tile % ntx) * kTX, y0 = (tile / ntx) * TY;
    const int z0 = chunk * zchunk, z1 = min(m.nz, z0 + zchunk);
    const int tcell = threadIdx.x % (kTX * TY);
    const int comp = COMP;
    const int tx = tcell % kTX, ty = tcell / kTX;
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix < m.nx && iy < m.ny;
    Nb nb;  // global neighbours for nu_t and the collocated pressure gradient
    make_nb<1>(nb, active ? ix : 0, active ? iy : 0, 0, m.nx, m.ny);

    // The padded tile of a plane is copied into LDS word by word (a word is one
    // real): word d of the slot (cell d / DPC, component d % DPC of the real3 AoS
    // cell) comes from the cell's periodic image; each lane holds the global word
    // offsets of its ND words within a plane, the same for every plane.
    int gdw[ND];
#pragma unroll
    for (int k = 0; k < ND; k++) {
        const int d = min(threadIdx.x + k * NT, DPS - 1);  // surplus lanes re-read a valid word
        const int e = d / DPC, wd = d - e * DPC;
        const int lx = e % W, ly = e / W;
        int gx = (x0 + lx - RXY) % m.nx;
        if (gx < 0) gx += m.nx;
        int gy = (y0 + ly - RXY) % m.ny;
        if (gy < 0) gy += m.ny;
        gdw[k] = (gy * m.nx + gx) * DPC + wd;
    }
    // A plane fill is ND word loads per lane into registers, written to the
    // ring slot after the plane in between has been computed: the loads in
    // flight target registers, so no LDS read of the computation waits for them.
    // Lanes past the plane re-read a valid word into the slot's spare words.
    Word pf[ND];
    auto fetch = [&](int p) {  // plane min(p, nz-1) -> registers
        const Word* __restrict__ src =
            reinterpret_cast<const Word*>(a.vel) + (size_t)min(p, m.nz - 1) * m.nxy * DPC;
#pragma unroll
        for (int k = 0; k < ND; k++) pf[k] = src[gdw[k]];
    };
    auto commit = [&](int p) {  // registers -> the ring slot of plane p
        Word* dst = reinterpret_cast<Word*>(ring + (p % NS) * SLOT);
#pragma unroll
        for (int k = 0; k < ND; k++) {
            const int d = threadIdx.x + k * NT;
            dst[(k + 1) * NT <= DPS || d < DPS ? d : DPS + (threadIdx.x & 63)] = pf[k];
        }
    };
    for (int p = max(0, z0 - RZ); p <= min(m.nz - 1, z0 + RZ); p++) {
        fetch(p);
        commit(p);
    }
    __syncthreads();

    // Plane loop. Per plane: (1) store the previous plane's results, (2) issue
    // this plane's own global loads, (3) issue the next plane's LDS fill, (4)
    // compute from LDS and registers only. The vector-memory counter is in
    // order, so waiting for (2) leaves (3) in flight; no global load sits inside
    // the computation, and every thread computes (outside cells on clamped
    // indices, results not stored) so that the mesh arrays are read uniformly.
    double cfr = 0.0;
    const int cx = active ? ix : 0, cy = active ? iy : 0;
    // NC == 3: a thread keeps component `comp` only (kc, rc; history vp[0] etc.)
    T kq[3] = {T(0), T(0), T(0)}, rq[3] = {T(0), T(0), T(0)};
    int in_q = 0;
    auto store_results = [&]() {
        if (NC == 3) {
            a.fcur[3 * in_q + comp] = kq[0];
            if (a.order > 0) a.fnext[3 * in_q + comp] = rq[0];
        } else {
            T* fc = a.fcur + 3 * in_q;
            fc[0] = kq[0];
            fc[1] = kq[1];
            fc[2] = kq[2];
            if (a.order > 0) {
                T* fn = a.fnext + 3 * in_q;
                fn[0] = rq[0];
                fn[1] = rq[1];
                fn[2] = rq[2];
            }
        }
    };
    constexpr int NH = NC == 3 ? 1 : 3;  // history components a thread loads
    for (int iz = z0; iz < z1; iz++) {
        if (iz > z0 && active) store_results();
        const int in = iz * m.nxy + cy * m.nx + cx;
        const int c0 = NC == 3 ? comp : 0;
        T vp[3] = {T(0), T(0), T(0)}, vp2[3] = {T(0), T(0), T(0)}, fp[3] = {T(0), T(0), T(0)}, fp2[3] = {T(0), T(0), T(0)};
        if (a.order >= 2) {
#pragma unroll
            for (int c = 0; c < NH; c++) {
                vp[c] = load_once(a.vprev + 3 * in + c0 + c);
                fp[c] = load_once(a.fprev + 3 * in + c0 + c);
            }
        }
        if (a.order == 3) {
#pragma unroll
            for (int c = 0; c < NH; c++) {
                vp2[c] = load_once(a.vprev2 + 3 * in + c0 + c);
                fp2[c] = load_once(a.fprev2 + 3 * in + c0 + c);
            }
        }
        Acc acc;
        if (COLUMN) {
            const int zp = min(iz + RZ + 1, m.nz - 1), zm = max(iz - RZ - 1, 0);
            const T* cpp = a.vel + 3 * ((size_t)zp * m.nxy + cy * m.nx + cx);
            const T* cmp = a.vel + 3 * ((size_t)zm * m.nxy + cy * m.nx + cx);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                acc.cp[c] = cpp[c];
                acc.cm[c] = cmp[c];
            }
        }
        nb.base = iz * m.nxy;
        T pg[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};  // collocated StabType 1: p at x+-, y+-, z+-
        const bool stab = SCHEME == TC_SCHEME_AB3 && a.order > 0 && a.stab && iz != 0 && iz != m.nz - 1;
        if (stab) {
            pg[0] = a.pcur[nb.at(1, 0, 0)];
            pg[1] = a.pcur[nb.at(-1, 0, 0)];
            pg[2] = a.pcur[nb.at(0, 1, 0)];
            pg[3] = a.pcur[nb.at(0, -1, 0)];
            pg[4] = a.pcur[nb.at(0, 0, 1)];
            pg[5] = a.pcur[nb.at(0, 0, -1)];
        }
        T tq[3][7];
        const bool sgs = SCHEME == TC_SCHEME_AB3 && a.tau != nullptr && iz != 0 && iz != m.nz - 1;
        if (sgs) sgs_load(a.tau, nb, iz, m.nz, tq);
        __builtin_amdgcn_sched_barrier(0);
        // planes iz-RZ..iz+RZ are in LDS; plane iz+RZ+1 (next iteration) goes into
        // the slot of plane iz-RZ-1, which no thread reads after the last barrier
        fetch(iz + RZ + 1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int d = -RZ; d <= RZ; d++) acc.pl[d + RZ] = ring + (((iz + d) % NS + NS) % NS) * SLOT + 3 * (ty * W + tx);
        GlobalNu<T> nu{a.nut, &nb, iz, m.nz};
        T k[3] = {0, 0, 0};
        if (SCHEME == TC_SCHEME_AB3) {
            if (iz != 0 && iz != m.nz - 1) {
                const T inv_hz = m.IHC[iz];  // 2 / (DZ[iz-1] + DZ[iz])
                conv_ab3(m, acc, iz, inv_hz, k);
                if (VISC == 1) {
                    T g[3];
                    visc_galerkin<T, TURB>(m, acc, nu, iz, g);
                    const T s = m.inv_hx * m.inv_hy * inv_hz;
#pragma unroll
                    for (int c = 0; c < 3; c++) k[c] += g[c] * s;
                } else {  // naive Laplacian (collocated_AB3/func_fluxes.h:110-113)
                    const T ihx2 = m.visc * m.inv_hx * m.inv_hx, ihy2 = m.visc * m.inv_hy * m.inv_hy;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const T a0 = acc(c, 0, 0, 0);
                        k[c] += ihx2 * (acc(c, 1, 0, 0) + acc(c, -1, 0, 0) - T(2) * a0);
                        k[c] += ihy2 * (acc(c, 0, 1, 0) + acc(c, 0, -1, 0) - T(2) * a0);
                        k[c] += m.visc * ((acc(c, 0, 0, 1) - a0) * m.IDZ[iz] - (a0 - acc(c, 0, 0, -1)) * m.IDZM[iz]) * inv_hz;
                    }
                }
                if (sgs) sgs_divergence(m, tq, iz, k);
                k[0] += m.sour;
            }
        } else {  // ExplicitTerm (staggered/func_global.h:642-649)
            auto explicit_term = [&](auto interior) {
                constexpr bool INT = decltype(interior)::value;
                T c[3], v[3];
                if (SCHEME == TC_SCHEME_HW) conv_hw<INT>(m, acc, iz, c);
                else conv_vv<INT>(m, acc, iz, c, comp);
                if (VISC == 1) visc_naive<INT, T, TURB>(m, acc, nu, iz, v, comp);
                else visc_cross<INT>(m, acc, iz, v, comp);
                k[0] = (m.sour + c[0]) + v[0];
                k[1] = (T(0) + c[1]) + v[1];
                k[2] = !INT && iz == 0 ? T(0) : (T(0) + c[2]) + v[2];
            };
            // FP64 constant viscosity: planes 3..nz-4 take a copy of the stencils
            // with every wall test folded away (a uniform branch)
            if constexpr (ALLINT) {
                explicit_term(std::true_type{});
            } else if constexpr (sizeof(T) == 8 && VISC == 0 && !TURB) {
                if (iz >= 3 && iz < m.nz - 3) explicit_term(std::true_type{});
                else explicit_term(std::false_type{});
            } else {
                T c[3], v[3];
                if (SCHEME == TC_SCHEME_HW) conv_hw<false>(m, acc, iz, c);
                else conv_vv<false>(m, acc, iz, c, comp);
                if (VISC == 1) visc_naive<false, T, TURB>(m, acc, nu, iz, v, comp);
                else visc_cross<false>(m, acc, iz, v, comp);
                k[0] = (m.sour + c[0]) + v[0];
                k[1] = (T(0) + c[1]) + v[1];
                k[2] = iz == 0 ? T(0) : (T(0) + c[2]) + v[2];
            }
        }
        if (NC == 3) {  // the kept component moves to slot 0
            k[0] = comp == 0 ? k[0] : comp == 1 ? k[1] : k[2];
        }
        if (a.order > 0) {  // BDF/EX combine (staggered/func_global.h:909-932)
            const T it = a.inv_tau;
#pragma unroll
            for (int cc = 0; cc < NH; cc++) {
                const T u0 = acc(c0 + cc, 0, 0, 0);
                if (a.order == 1) rq[cc] = k[cc] + it * u0;
                else if (a.order == 2) rq[cc] = T(2) * k[cc] - fp[cc] + it * (T(2) * u0 - T(0.5) * vp[cc]);
                else
                    rq[cc] = T(3) * k[cc] - T(3) * fp[cc] + fp2[cc] +
                             it * (T(3) * u0 - T(1.5) * vp[cc] + T(1.0 / 3.0) * vp2[cc]);
            }
            if (stab) {  // r -= G p_n (collocated_AB3/func_global.h:435-437; ApplyGradient :298-303)
                rq[0] -= T(0.5) * m.inv_hx * (pg[0] - pg[1]);
                rq[1] -= T(0.5) * m.inv_hy * (pg[2] - pg[3]);
                rq[2] -= T(0.5) * (pg[4] - pg[5]) / m.cellh[iz];
            }
            if (a.cfr_partial && active && c0 == 0) cfr += (double)rq[0] * u_volume<T, SCHEME>(m, iz);
        }
        kq[0] = k[0];
        kq[1] = k[1];
        kq[2] = k[2];
        in_q = in;
        commit(iz + RZ + 1);  // the slot of plane iz-RZ-1, which no thread reads after the last barrier
        __syncthreads();
    }
    if (z1 > z0 && active) store_results();
    if (a.cfr_partial) {
        const double s = block_sum(cfr, sh);
        if (threadIdx.x == 0) a.cfr_partial[blockIdx.x] = s;
    }
}

template <typename T, int SCHEME, int VISC, bool TURB>
__global__ __launch_bounds__((flux_threads<T, SCHEME, VISC>()), (flux_min_waves<T, SCHEME, VISC>())) void k_flux(
    Mesh<T> m, FluxArgs<T> a, int ntx, int ntiles, int zchunk) {
    constexpr int TY = flux_ty<T, SCHEME, VISC>();
    constexpr int RXY = SCHEME == TC_SCHEME_VV ? 3 : 1;
    constexpr int RZ = flux_rz<T, SCHEME>();
    constexpr int NS = 2 * RZ + 2;  // planes iz-RZ..iz+RZ plus the one in flight
    using Acc = RingAcc<T, RXY, RZ, TY>;
    constexpr int W = Acc::W, PS = Acc::PS;
    constexpr int NC = flux_nc<T, SCHEME>();
    constexpr int NT = kTX * TY * NC;
    using Word = RingWord<T>;                   // one real: a fill moves whole reals
    constexpr int DPC = 3;                      // words per real3 cell
    constexpr int DPS = PS * DPC;               // words per plane of the padded tile
    constexpr int ND = (DPS + NT - 1) / NT;
    // slot: the plane's words, then 64 spare words that the lanes past the plane
    // write (one per lane position, so no two lanes of a wave collide); every
    // lane writes ND words per fill, the same instruction stream for every plane
    constexpr int SLOT = (DPS + 64 + 1) / 2 * 2;  // reals per ring slot (16-byte multiple)
    __shared__ __align__(16) T ring[NS * SLOT];
    __shared__ double sh[NT / 64];
    // The mesh arrays the stencils read, mirrored in LDS (dynamic, kFluxMeshArrays
    // nz reals): LDS reads wait on their own counter, never on the global loads
    // in flight.
    extern __shared__ __align__(16) unsigned char mesh_raw[];
    {
        T* ml = reinterpret_cast<T*>(mesh_raw);
        const T* src[kFluxMeshArrays] = {m.DZ, m.DZW, m.GCZ, m.cellh, m.IDZ, m.IHDN, m.IHUP, m.IDZM, m.IVOL, m.IHC};
        for (int i = threadIdx.x; i < m.nz; i += NT)
#pragma unroll
            for (int k = 0; k < kFluxMeshArrays; k++) ml[k * m.nz + i] = src[k][i];
        m.DZ = ml;
        m.DZW = ml + m.nz;
        m.GCZ = ml + 2 * m.nz;
        m.cellh = ml + 3 * m.nz;
        m.IDZ = ml + 4 * m.nz;
        m.IHDN = ml + 5 * m.nz;
        m.IHUP = ml + 6 * m.nz;
        m.IDZM = ml + 7 * m.nz;
        m.IVOL = ml + 8 * m.nz;
        m.IHC = ml + 9 * m.nz;
    }
    if (NC == 3) {
        const int comp = threadIdx.x / (kTX * TY);  // wave-uniform
        if (comp == 0) flux_planes<T, SCHEME, VISC, TURB, 0>(m, a, ntx, ntiles, zchunk, ring, sh);
        else if (comp == 1) flux_planes<T, SCHEME, VISC, TURB, 1>(m, a, ntx, ntiles, zchunk, ring, sh);
        else flux_planes<T, SCHEME, VISC, TURB, 2>(m, a, ntx, ntiles, zchunk, ring, sh);
    } else if constexpr (sizeof(T) == 4 && SCHEME != TC_SCHEME_AB3 && VISC == 0 && !TURB) {
        // FP32 constant viscosity: a block whose chunk lies in planes 3..nz-4
        // runs the copy of the plane loop without wall tests; the choice is per
        // block, so each copy is scheduled as a loop of its own (a per-plane
        // choice made the FP32 loop wait for its history loads)
        const int z0 = xcd_swizzle(blockIdx.x, gridDim.x) / ntiles * zchunk, z1 = min(m.nz, z0 + zchunk);
        if (z0 >= 3 && z1 <= m.nz - 3) flux_planes<T, SCHEME, VISC, TURB, -1, true>(m, a, ntx, ntiles, zchunk, ring, sh);
        else flux_planes<T, SCHEME, VISC, TURB, -1>(m, a, ntx, ntiles, zchunk, ring, sh);
    } else {
        flux_planes<T, SCHEME, VISC, TURB, -1>(m, a, ntx, ntiles, zchunk, ring, sh);
    }
}

struct FluxGrid {
    int ntx, ntiles, zchunk, nblocks;
};
// Blocks resident at once on the device for one flux kernel instantiation.
template <typename T, int SCHEME, int VISC, bool TURB>
int flux_slots(size_t dyn_lds) {
    static int slots = 0;
    static size_t slots_lds = 0;
    if (slots == 0 || slots_lds != dyn_lds) {
        int per_cu = 0, dev = 0, ncu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_flux<T, SCHEME, VISC, TURB>,
                                                         flux_threads<T, SCHEME, VISC>(), dyn_lds) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            ncu < 1)
            ncu = 256;
        slots = per_cu * ncu;
        slots_lds = dyn_lds;
    }
    return slots;
}
// z-chunks: the count whose grid, in whole rounds of resident blocks, takes the
// fewest plane-steps, each block paying its chunk plus the ring's start-up
// planes (rounds x (zchunk + halo)); ties go to fewer chunks.
inline FluxGrid flux_grid(int nx, int ny, int nz, int ty, int slots, int halo) {
    FluxGrid g;
    g.ntx = (nx + kTX - 1) / kTX;
    g.ntiles = g.ntx * ((ny + ty - 1) / ty);
    const int max_nch = std::max(1, nz / 4);
    long best = -1;
    int best_nch = 1;
    for (int nch = 1; nch <= max_nch; nch++) {
        const int zc = (nz + nch - 1) / nch;
        const long blocks = (long)g.ntiles * ((nz + zc - 1) / zc);
        const long rounds = (blocks + slots - 1) / slots;
        const long cost = rounds * (zc + halo);
        if (best < 0 || cost < best) {
            best = cost;
            best_nch = nch;
        }
    }
    g.zchunk = (nz + best_nch - 1) / best_nch;
    g.nblocks = g.ntiles * ((nz + g.zchunk - 1) / g.zchunk);
    return g;
}

// local accessors for the remaining direct-access kernels
#define F3(a, c, dx, dy, dz) a[3 * nb.at(dx, dy, dz) + (c)]
#define UU(dx, dy, dz) F3(vel, 0, dx, dy, dz)
#define VV(dx, dy, dz) F3(vel, 1, dx, dy, dz)
#define WW(dx, dy, dz) F3(vel, 2, dx, dy, dz)

// ---------------------------------------------------------------------------
// Turbulent viscosity
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_turb_staggered(Mesh<T> m, const T* __restrict__ vel, T* __restrict__ nut) {
    // CalcAbsS + RecalcTurbVisc (staggered/func_global.h:823-890)
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz, NZ = m.nz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
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
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz, NZ = m.nz;
    int ix, iy;
    if (iz >= NZ - 1 || !cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
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

template <typename T, int SCHEME>
__global__ __launch_bounds__(kBlock) void k_grad(Mesh<T> m, const T* __restrict__ p, T* __restrict__ g) {
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
    Nb nb;
    make_nb<2>(nb, ix, iy, iz, m.nx, m.ny);
    T r[3];
    grad_at<T, SCHEME>(m, p, nb, iz, r);
    T* o = g + 3 * nb.at(0, 0, 0);
    o[0] = r[0];
    o[1] = r[1];
    o[2] = r[2];
}

// Per-plane factors of the diagnostics' CFL terms.
struct DiagPlane {
    double ihx, ihy, inv_hz, kv;
};

// K7 at one cell: velocity correction (staggered/func_global.h:937-941;
// collocated_AB3/func_global.h:428-449), in two parts: project_eval does every
// load and the arithmetic, project_store the stores (and the diagnostics
// terms), so that a thread can issue the loads of several cells before its
// first store. DIAG: also the values of the cell's diagnostics terms (the
// terms k_diag adds per cell, diag.h); the staggered cell-centre w needs the
// new w one plane up, formed here as the thread that stores it forms it (r
// must not alias the output then, as in a step).
template <typename T>
struct CellProj {
    int in;
    T u, v, w, pn, wup;
    double nt, vc, ps;
};

template <typename T, int SCHEME, bool STAB, bool DIAG>
__device__ __forceinline__ void project_eval(const Mesh<T>& m, const ProjectArgs<T>& a, int ix, int iy, int iz,
                                             CellProj<T>& c) {
    Nb nb;
    make_nb<2>(nb, ix, iy, iz, m.nx, m.ny);
    const int in = nb.at(0, 0, 0);
    c.in = in;
    const T* r = a.r + 3 * in;
    const T r0 = load_once(r), r1 = load_once(r + 1), r2 = load_once(r + 2);  // p: other cells' taps too
    const T sh = a.shift ? T(*a.shift) : T(0);
    c.pn = T(0);
    if (STAB) c.pn = a.p[in] + a.pcur[in];
    T g[3];
    grad_at<T, SCHEME>(m, a.p, nb, iz, g);
    c.wup = T(0);  // staggered: the new w one plane up (0 past the top wall)
    if (DIAG && SCHEME != TC_SCHEME_AB3) {
        const T rup = a.r[3 * nb.at(0, 0, zclamp(iz, 1, m.nz)) + 2];
        const T wu = (rup - grad_z_stag<T, SCHEME>(m, a.p, nb, iz + 1, 1)) * a.tau_over_bd0;
        c.wup = iz < m.nz - 1 ? wu : T(0);
    }
    c.nt = 0.0;
    c.vc = m.visc;
    c.ps = 0.0;
    if (DIAG) {
        if (a.nut) c.nt = a.nut[in];
        if (SCHEME == TC_SCHEME_AB3) {
            if (a.nut) {  // GetNuTAtNode: mean of the 8 elements around the node (interior nodes)
                const T* n = a.nut;
                const int c0 = in, cx = nb.at(-1, 0, 0), cy = nb.at(0, -1, 0), cxy = nb.at(-1, -1, 0);
                const int dz = zclamp(iz, -1, m.nz) * m.nxy;
                const double s8 = (double)n[c0] + n[cx] + n[cy] + n[cxy] + n[c0 + dz] + n[cx + dz] + n[cy + dz] +
                                  n[cxy + dz];
                if (iz != 0 && iz != m.nz - 1) c.vc += 0.125 * s8;
            }
        } else {
            if (a.nut) c.vc = (double)m.visc + c.nt;
            c.ps = a.p[in];
        }
    }
    T u = (r0 - g[0]) * a.tau_over_bd0;
    if (a.shift && !(SCHEME == TC_SCHEME_AB3 && (iz == 0 || iz == m.nz - 1))) u += sh;
    c.u = u;
    c.v = (r1 - g[1]) * a.tau_over_bd0;
    c.w = (r2 - g[2]) * a.tau_over_bd0;
}

template <typename T, int SCHEME, bool STAB, bool DIAG>
__device__ __forceinline__ void project_store(const ProjectArgs<T>& a, const CellProj<T>& c, DiagSums<SCHEME>* acc,
                                              const DiagPlane& dp) {
    // streaming stores: the new fields are not read again here, so they should not
    // displace the p and r other blocks still read as stencil taps
    T* o = a.vnext + 3 * c.in;
    __builtin_nontemporal_store(c.u, o);
    __builtin_nontemporal_store(c.v, o + 1);
    __builtin_nontemporal_store(c.w, o + 2);
    if (STAB) __builtin_nontemporal_store(c.pn, a.pnext + c.in);
    if (DIAG) {
        const double wc = SCHEME != TC_SCHEME_AB3 ? 0.5 * ((double)c.w + (double)c.wup) : (double)c.w;
        acc->add(c.u, c.v, wc, c.w, a.nut != nullptr, c.nt, c.ps, c.vc, dp.ihx, dp.ihy, dp.inv_hz, dp.kv);
    }
}

// Cells per thread of the projection that also forms the diagnostics: the
// loads of all of them go out before the first store, and one reduction of the
// record serves kProjDiagCells * kBlock cells of a plane.
constexpr int kProjDiagCells = 4;

// K7 over a (plane chunk, plane) grid, one cell per thread; DIAG: kProjDiagCells
// chunks of kBlock cells of a plane per block and one partial diagnostics record
// per block (each wave's sums by transposing exchanges, then the block's four
// waves through LDS).
template <typename T, int SCHEME, bool STAB, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_project(Mesh<T> m, ProjectArgs<T> a) {
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz;
    DiagPlane dp;
    if (!DIAG) {
        int ix, iy;
        if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
        CellProj<T> c;
        project_eval<T, SCHEME, STAB, false>(m, a, ix, iy, iz, c);
        project_store<T, SCHEME, STAB, false>(a, c, nullptr, dp);
        return;
    }
    using D = DiagSums<SCHEME>;
    constexpr int P = D::NX <= 16 ? 16 : 32;
    __shared__ double shr[kBlock / 64][P];
    dp.ihx = m.inv_hx;
    dp.ihy = m.inv_hy;
    dp.inv_hz = SCHEME == TC_SCHEME_AB3 ? 1.0 / (double)m.cellh[iz] : 1.0 / (double)m.DZ[iz];
    dp.kv = a.visc_fac * (dp.ihx * dp.ihx + dp.ihy * dp.ihy + dp.inv_hz * dp.inv_hz);
    CellProj<T> c[kProjDiagCells];
    bool act[kProjDiagCells];
#pragma unroll
    for (int k = 0; k < kProjDiagCells; k++) {  // a cell past the plane computes on the plane's last cell
        const int t = (pb.bx * kProjDiagCells + k) * kBlock + threadIdx.x;
        act[k] = t < m.nxy;
        const int tt = act[k] ? t : m.nxy - 1;
        const int iy = tt / m.nx, ix = tt - iy * m.nx;
        project_eval<T, SCHEME, STAB, true>(m, a, ix, iy, iz, c[k]);
    }
    D acc;
    acc.init();
#pragma unroll
    for (int k = 0; k < kProjDiagCells; k++)
        if (act[k]) project_store<T, SCHEME, STAB, true>(a, c[k], &acc, dp);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (P > 16 && D::NX == 17 && !a.nut) {
        // Without nu_t every cell's viscous CFL term is kv * visc, so that maximum
        // needs no exchange, and the other 14 sums and 2 maxima fit 16 slots: half
        // the exchanges of the 32-slot reduction. Each block has an active cell.
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = k < D::NS ? acc.s[k] : k == D::NS ? acc.mx[0] : acc.mx[2];
        const double r = wave_reduce_transpose<16>(v, (1ull << D::NS) - 1);
        if (lane % 4 == 0) shr[wid][lane / 4 + (lane / 4 > D::NS)] = r;  // slot 15 -> record value NS + 2
        if (lane == 0) shr[wid][D::NS + 1] = dp.kv * (double)m.visc;
    } else {
        double v[P];
#pragma unroll
        for (int k = 0; k < P; k++) v[k] = k < D::NS ? acc.s[k] : k < D::NX ? acc.mx[k - D::NS] : 0.0;
        const unsigned long long summask = (1ull << D::NS) - 1;  // then 3 maxima (and padding)
        const double r = wave_reduce_transpose<P>(v, summask);
        constexpr int LS = 64 / P;  // lanes per value
        if (lane % LS == 0) shr[wid][lane / LS] = r;
    }
    __syncthreads();
    if (threadIdx.x < D::NX) {
        const int k = threadIdx.x;
        double x = shr[0][k];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++) x = k < D::NS ? x + shr[w][k] : fmax(x, shr[w][k]);
        a.diag[((size_t)iz * gridDim.x + pb.bx) * D::NX + k] = x;
    }
}

template <typename T, int SCHEME>
__global__ __launch_bounds__(kBlock) void k_add_u(Mesh<T> m, T* __restrict__ v, const double* __restrict__ shift) {
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
    if (SCHEME == TC_SCHEME_AB3 && (iz == 0 || iz == m.nz - 1)) return;
    v[3 * (iz * m.nxy + iy * m.nx + ix)] += T(*shift);
}

// f[in] -= (sum * coef) * w[iz] (w == nullptr: w = 1); used by NormalizePressure.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_shift_plane(Mesh<T> m, T* __restrict__ f, const double* __restrict__ sum,
                                                         double coef, const T* __restrict__ w) {
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
    const int in = iz * m.nxy + iy * m.nx + ix;
    const T s = T(sum[0] * coef);
    f[in] -= w ? s * w[iz] : s;
}

template <typename T>
dim3 cell_grid(const Mesh<T>& m) {
    return dim3((m.nxy + kBlock - 1) / kBlock, m.nz);
}

// ---------------------------------------------------------------------------
// SRK stage updates (collocated_SRK/func_global.h:317-408). Elementwise over
// nodes except for the collocated gradient of one scalar; every read of a
// node comes before its writes, since outputs alias inputs at the same node.
// ---------------------------------------------------------------------------
// Collocated gradient (the AB3 branch of grad_at) of s = S + S2 (S2 may be nullptr).
template <typename T>
__device__ __forceinline__ void grad_sum_ab3(const Mesh<T>& m, const T* __restrict__ S, const T* __restrict__ S2,
                                             const Nb& nb, int iz, T g[3]) {
    if (iz == 0 || iz == m.nz - 1) {
        g[0] = g[1] = g[2] = T(0);
        return;
    }
    auto at = [&](int dx, int dy, int dz) {
        const int i = nb.at(dx, dy, dz);
        return S2 ? S[i] + S2[i] : S[i];
    };
    g[0] = T(0.5) * m.inv_hx * (at(1, 0, 0) - at(-1, 0, 0));
    g[1] = T(0.5) * m.inv_hy * (at(0, 1, 0) - at(0, -1, 0));
    g[2] = T(0.5) * (at(0, 0, 1) - at(0, 0, -1)) / m.cellh[iz];
}

template <typename T, bool GRAD>
__global__ __launch_bounds__(kBlock) void k_srk_vec(Mesh<T> m, SrkVec<T> a) {
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
    Nb nb;
    make_nb<1>(nb, ix, iy, iz, m.nx, m.ny);
    const int in = nb.at(0, 0, 0);
    T kc[3] = {T(0), T(0), T(0)};
    if (a.KC) {
        const T* k = a.KC + 3 * in;
        kc[0] = load_once(k);
        kc[1] = load_once(k + 1);
        kc[2] = load_once(k + 2);
    }
    if (GRAD) {
        T g[3];
        grad_sum_ab3<T>(m, a.S, a.S2, nb, iz, g);
        kc[0] -= g[0];
        kc[1] -= g[1];
        kc[2] -= g[2];
    }
    T pj = T(0), qn = T(0);
    if (a.pj_out || a.qj_out) {
        pj = a.S2 ? a.S[in] + a.S2[in] : a.S[in];
        if (a.qj_out) qn = a.inv_tau_stage * (pj - load_once(a.pjhat + in));
    }
    T o[3] = {T(0), T(0), T(0)};
    if (a.out) {
        if (a.X) {
            const T* x = a.X + 3 * in;
            o[0] = load_once(x);
            o[1] = load_once(x + 1);
            o[2] = load_once(x + 2);
        }
#pragma unroll
        for (int j = 0; j < kMaxImexStages; j++)
            if (j < a.nk) {
                const T* k = a.K[j] + 3 * in;
                const T cj = a.c[j];
                o[0] += cj * load_once(k);
                o[1] += cj * load_once(k + 1);
                o[2] += cj * load_once(k + 2);
            }
        if (a.KC) {
            o[0] += a.ckc * kc[0];
            o[1] += a.ckc * kc[1];
            o[2] += a.ckc * kc[2];
        }
        if (a.A) {
            const T* x = a.A + 3 * in;
            o[0] += a.sa * load_once(x);
            o[1] += a.sa * load_once(x + 1);
            o[2] += a.sa * load_once(x + 2);
        }
        if (a.B) {
            const T* x = a.B + 3 * in;
            o[0] += a.sb * load_once(x);
            o[1] += a.sb * load_once(x + 1);
            o[2] += a.sb * load_once(x + 2);
        }
    }
    // streaming stores: nothing here reads them again (the gradient taps of S stay in L2)
    if (a.KCout) {
        T* k = a.KCout + 3 * in;
#pragma unroll
        for (int c = 0; c < 3; c++) __builtin_nontemporal_store(kc[c], k + c);
    }
    if (a.out) {
        T* x = a.out + 3 * in;
#pragma unroll
        for (int c = 0; c < 3; c++) __builtin_nontemporal_store(o[c], x + c);
    }
    if (a.pj_out) __builtin_nontemporal_store(pj, a.pj_out + in);
    if (a.qj_out) __builtin_nontemporal_store(qn, a.qj_out + in);
}

template <typename T, bool STAB, bool CK>
__global__ __launch_bounds__(kBlock) void k_srk_pre(Mesh<T> m, SrkPre<T> a) {
    const PlaneBlock pb = plane_block(m.xcd_swz);
    const int iz = pb.iz;
    int ix, iy;
    if (!cell_of_thread(m.nx, m.nxy, ix, iy, pb.bx)) return;
    const int in = iz * m.nxy + iy * m.nx + ix;
    const T p = a.p[in];
    if (!STAB) {
        const T fix = a.cfix * p;
        T muj = p;
        if (CK) muj *= a.m0;
#pragma unroll
        for (int j = 0; j < kMaxImexStages; j++)
            if (j < a.nj) muj += a.cm[j] * (a.mujt[j][in] - fix);
        a.bs[in] = muj - a.clast * p;
    } else {
        const T q = a.q[in];
        const T fix = a.cfix * (a.tau_stage * q);
        T muj = a.tau_stage * q;
        if (CK) muj *= a.m0;
        T ph = p;
        if (CK) ph += a.ph0 * q;
#pragma unroll
        for (int j = 0; j < kMaxImexStages; j++)
            if (j < a.nj) {
                muj += a.cm[j] * (a.mujt[j][in] - fix);
                ph += a.cq[j] * a.qj[j][in];
            }
        a.pjhat[in] = ph;
        a.bs[in] = muj + ph - a.clast * q;
    }
}

}  // namespace

// Upper bound of the per-block partial sums written by launch_flux (at least one z-plane per block).
int flux_grid_blocks(int nx, int ny, int nz) { return ((nx + kTX - 1) / kTX) * ((ny + 7) / 8) * nz; }

template <typename T>
hipError_t launch_flux(hipStream_t s, const Mesh<T>& m, const FluxArgs<T>& a, int* nblocks) {
#define L(SCH, VIS, TRB)                                                                                       \
    {                                                                                                          \
        constexpr int ty = flux_ty<T, SCH, VIS>();                                                             \
        const size_t lds = kFluxMeshArrays * (size_t)m.nz * sizeof(T);                                       \
        const FluxGrid fg = flux_grid(m.nx, m.ny, m.nz, ty, flux_slots<T, SCH, VIS, TRB>(lds), 2 * flux_rz<T, SCH>() + 1);               \
        if (nblocks) *nblocks = fg.nblocks;                                                                    \
        hipLaunchKernelGGL((k_flux<T, SCH, VIS, TRB>), dim3(fg.nblocks), dim3(flux_threads<T, SCH, VIS>()), lds, s, m, a, fg.ntx, \
                           fg.ntiles, fg.zchunk);                                                              \
    }
    if (a.scheme == TC_SCHEME_HW) {
        if (a.visc_type == 1) {
            if (a.turb) L(TC_SCHEME_HW, 1, true) else L(TC_SCHEME_HW, 1, false)
        } else {
            L(TC_SCHEME_HW, 0, false)
        }
    } else if (a.scheme == TC_SCHEME_VV) {
        if (a.visc_type == 1) {
            if (a.turb) L(TC_SCHEME_VV, 1, true) else L(TC_SCHEME_VV, 1, false)
        } else {
            L(TC_SCHEME_VV, 0, false)
        }
    } else {
        if (a.visc_type == 1) {
            if (a.turb) L(TC_SCHEME_AB3, 1, true) else L(TC_SCHEME_AB3, 1, false)
        } else {
            L(TC_SCHEME_AB3, 0, false)
        }
    }
#undef L
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
hipError_t launch_gradient(hipStream_t s, const Mesh<T>& m, int scheme, const T* p, T* gr) {
    const dim3 g = cell_grid(m);
    if (scheme == TC_SCHEME_HW) hipLaunchKernelGGL((k_grad<T, TC_SCHEME_HW>), g, dim3(kBlock), 0, s, m, p, gr);
    else if (scheme == TC_SCHEME_VV) hipLaunchKernelGGL((k_grad<T, TC_SCHEME_VV>), g, dim3(kBlock), 0, s, m, p, gr);
    else hipLaunchKernelGGL((k_grad<T, TC_SCHEME_AB3>), g, dim3(kBlock), 0, s, m, p, gr);
    return hipGetLastError();
}

int project_diag_chunks(int nxy) { return (nxy + kBlock * kProjDiagCells - 1) / (kBlock * kProjDiagCells); }

template <typename T>
hipError_t launch_project(hipStream_t s, const Mesh<T>& mesh, const ProjectArgs<T>& a) {
    dim3 g = cell_grid(mesh);
    // XCD-contiguous block order at every plane size: the y taps of p sit in the
    // rows of the neighbouring chunks of the plane, which then share an L2. The
    // diagnostics variant (4 cells of a plane per thread) keeps the plain order
    // on large planes: there the contiguous order fetched more (VV 384 x 384 x 192
    // FP32 581 -> 1181 MB) and ran slower.
    Mesh<T> m = mesh;
    if (!a.diag) m.xcd_swz = 1;
#define LP(SCH, STAB, DIAG) hipLaunchKernelGGL((k_project<T, SCH, STAB, DIAG>), g, dim3(kBlock), 0, s, m, a)
    if (a.diag) {
        g.x = project_diag_chunks(m.nxy);
        if (a.scheme == TC_SCHEME_HW) LP(TC_SCHEME_HW, false, true);
        else if (a.scheme == TC_SCHEME_VV) LP(TC_SCHEME_VV, false, true);
        else if (a.pnext) LP(TC_SCHEME_AB3, true, true);
        else LP(TC_SCHEME_AB3, false, true);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return launch_diag_final_wide(s, a.scheme, a.diag, (int)g.x, m.nz, a.diag_out);
    }
    if (a.scheme == TC_SCHEME_HW) LP(TC_SCHEME_HW, false, false);
    else if (a.scheme == TC_SCHEME_VV) LP(TC_SCHEME_VV, false, false);
    else if (a.pnext) LP(TC_SCHEME_AB3, true, false);
    else LP(TC_SCHEME_AB3, false, false);
#undef LP
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

template <typename T>
hipError_t launch_srk_vec(hipStream_t s, const Mesh<T>& m, const SrkVec<T>& a) {
    if (a.S && !a.KC) return hipErrorInvalidValue;
    if (a.S) hipLaunchKernelGGL((k_srk_vec<T, true>), cell_grid(m), dim3(kBlock), 0, s, m, a);
    else hipLaunchKernelGGL((k_srk_vec<T, false>), cell_grid(m), dim3(kBlock), 0, s, m, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_srk_pre(hipStream_t s, const Mesh<T>& m, const SrkPre<T>& a) {
    const dim3 g = cell_grid(m);
    if (a.stab) {
        if (a.ck) hipLaunchKernelGGL((k_srk_pre<T, true, true>), g, dim3(kBlock), 0, s, m, a);
        else hipLaunchKernelGGL((k_srk_pre<T, true, false>), g, dim3(kBlock), 0, s, m, a);
    } else {
        if (a.ck) hipLaunchKernelGGL((k_srk_pre<T, false, true>), g, dim3(kBlock), 0, s, m, a);
        else hipLaunchKernelGGL((k_srk_pre<T, false, false>), g, dim3(kBlock), 0, s, m, a);
    }
    return hipGetLastError();
}

#define INST(T)                                                                                             \
    template hipError_t launch_srk_vec<T>(hipStream_t, const Mesh<T>&, const SrkVec<T>&);                  \
    template hipError_t launch_srk_pre<T>(hipStream_t, const Mesh<T>&, const SrkPre<T>&);                  \
    template hipError_t launch_flux<T>(hipStream_t, const Mesh<T>&, const FluxArgs<T>&, int*);              \
    template hipError_t launch_turb_visc<T>(hipStream_t, const Mesh<T>&, int, const T*, T*);                \
    template hipError_t launch_gradient<T>(hipStream_t, const Mesh<T>&, int, const T*, T*);                 \
    template hipError_t launch_project<T>(hipStream_t, const Mesh<T>&, const ProjectArgs<T>&);              \
    template hipError_t launch_add_u<T>(hipStream_t, const Mesh<T>&, int, T*, const double*);               \
    template hipError_t launch_add_scaled_plane<T>(hipStream_t, const Mesh<T>&, T*, const double*, double, const T*);
INST(double)
INST(float)
#undef INST

}  // namespace tc
