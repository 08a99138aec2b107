// Reductions with FP64 accumulation: volume integrals, kinetic energy, the
// pressure-normalisation sum, CFL and per-plane statistics
// (staggered/func_global.h:657-808; collocated_AB3/func_global.h:162-289).
// Two stages with a fixed grid, so results do not depend on scheduling.
#include "diag.h"
#include "kernels.h"

namespace tc {

namespace {

constexpr int kThreads = 256;

template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double (*sh)[kThreads / 64], bool is_max) {
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
            for (int w = 1; w < kThreads / 64; w++) r = is_max ? fmax(r, sh[k][w]) : r + sh[k][w];
            v[k] = r;
        }
}

template <typename T>
__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <typename T, int OP, int SCHEME>
__global__ __launch_bounds__(kThreads) void k_reduce(Mesh<T> m, RedArgs<T> a, double* partial) {
    __shared__ double sh[3][kThreads / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    const double hxy = (double)m.hx * (double)m.hy;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < m.n3d; i += gridDim.x * kThreads) {
        const int iz = i / m.nxy;
        if (OP == RED_INTEGRAL || OP == RED_KE) {
            const double u = a.v[3 * i], v = a.v[3 * i + 1], w = a.v[3 * i + 2];
            if (SCHEME == TC_SCHEME_AB3) {
                const double vol = (double)m.cellh[iz] * hxy;
                if (OP == RED_INTEGRAL) {
                    acc[0] += u * vol;
                    acc[1] += v * vol;
                    acc[2] += w * vol;
                } else {
                    acc[0] += (u * u + v * v + w * w) * vol;
                }
            } else {
                const double dz = m.DZ[iz];
                const double dzw = iz == 0 ? 0.0 : 0.5 * ((double)m.DZ[iz] + (double)m.DZ[iz - 1]);
                if (OP == RED_INTEGRAL) {
                    acc[0] += u * dz * hxy;
                    acc[1] += v * dz * hxy;
                    acc[2] += w * dzw * hxy;
                } else {
                    acc[0] += ((u * u + v * v) * dz + w * w * dzw) * hxy;
                }
            }
        } else if (OP == RED_PSUM) {
            acc[0] += (double)a.p[i] * a.wplane[iz];
        } else {  // RED_CFL (staggered/func_global.h:692-704; collocated_AB3/func_global.h:188-199)
            double inv_hz, vc;
            if (SCHEME == TC_SCHEME_AB3) {
                inv_hz = 1.0 / (double)m.cellh[iz];
                vc = m.visc;
                if (a.nut && iz != 0 && iz != m.nz - 1) {  // GetNuTAtNode
                    const int ix = i % m.nx, iy = (i / m.nx) % m.ny;
                    const int xl = wrapi<T>(ix - 1, m.nx), yl = wrapi<T>(iy - 1, m.ny);
                    const int p0 = iz * m.nxy, p1 = (iz - 1) * m.nxy;
                    const T* n = a.nut;
                    vc += 0.125 * ((double)n[p0 + iy * m.nx + ix] + n[p0 + iy * m.nx + xl] + n[p0 + yl * m.nx + ix] +
                                   n[p0 + yl * m.nx + xl] + n[p1 + iy * m.nx + ix] + n[p1 + iy * m.nx + xl] +
                                   n[p1 + yl * m.nx + ix] + n[p1 + yl * m.nx + xl]);
                }
            } else {
                inv_hz = 1.0 / (double)m.DZ[iz];
                vc = a.nut ? (double)m.visc + (double)a.nut[i] : (double)m.visc;
            }
            double c = 0;
            const double ihx = m.inv_hx, ihy = m.inv_hy;
            if (a.mode & 1)
                c += a.tau * (fabs((double)a.v[3 * i]) * ihx + fabs((double)a.v[3 * i + 1]) * ihy + fabs((double)a.v[3 * i + 2]) * inv_hz);
            if (a.mode & 2) c += a.visc_fac * a.tau * vc * (ihx * ihx + ihy * ihy + inv_hz * inv_hz);
            acc[0] = fmax(acc[0], c);
        }
    }
    block_reduce<3>(acc, sh, OP == RED_CFL);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) partial[blockIdx.x * 3 + k] = acc[k];
}

__global__ __launch_bounds__(kThreads) void k_reduce_final(const double* partial, int nb, int is_max, double* out) {
    __shared__ double sh[3][kThreads / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += kThreads)
        for (int k = 0; k < 3; k++) acc[k] = is_max ? fmax(acc[k], partial[b * 3 + k]) : acc[k] + partial[b * 3 + k];
    block_reduce<3>(acc, sh, is_max);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[k] = acc[k];
}

// Per-plane statistics (CalcProfiles / CalcVelocityProfile).
template <typename T, int SCHEME>
__global__ __launch_bounds__(kThreads) void k_profiles(Mesh<T> m, const T* __restrict__ vel, const T* __restrict__ p,
                                                       const T* __restrict__ nut, double* partial) {
    constexpr int NV = SCHEME == TC_SCHEME_AB3 ? 10 : 12;
    __shared__ double sh[NV][kThreads / 64];
    const int iz = blockIdx.y, nch = gridDim.x;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] = 0.0;
    const int base = iz * m.nxy;
    for (int j = blockIdx.x * kThreads + threadIdx.x; j < m.nxy; j += nch * kThreads) {
        const int in = base + j;
        const double u = vel[3 * in], v = vel[3 * in + 1];
        double w = vel[3 * in + 2];
        if (SCHEME != TC_SCHEME_AB3) w = 0.5 * (w + (iz == m.nz - 1 ? 0.0 : (double)vel[3 * (in + m.nxy) + 2]));
        s[0] += u;
        s[1] += v;
        s[2] += w;
        s[3] += u * u;
        s[4] += v * v;
        s[5] += w * w;
        s[6] += u * v;
        s[7] += u * w;
        s[8] += v * w;
        if (nut) s[9] += nut[in];
        if (SCHEME != TC_SCHEME_AB3) {
            const double pp = p[in];
            s[10] += pp;
            s[11] += pp * pp;
        }
    }
    block_reduce<NV>(s, sh, false);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) partial[((size_t)iz * nch + blockIdx.x) * NV + k] = s[k];
}

__global__ void k_profiles_final(const double* partial, int nch, int nv, double inv_nxy, double* out) {
    const int iz = blockIdx.x, k = threadIdx.x;
    if (k >= nv) return;
    double r = 0.0;
    for (int c = 0; c < nch; c++) r += partial[((size_t)iz * nch + c) * nv + k];
    out[iz * nv + k] = r * inv_nxy;
}

// Fused per-step diagnostics: one pass over a velocity field (and p, nu_t) per
// plane chunk accumulates the plane statistics of k_profiles, the raw sums of w
// and w^2 the volume integrals need, and the plane maxima of the CFL numbers
// divided by tau (convective, viscous, both). Column layout of a plane record:
// [0, NV) as k_profiles, NV: sum w, NV+1: sum w^2, NV+2..NV+4: maxima.
template <typename T, int SCHEME>
__global__ __launch_bounds__(kThreads) void k_diag(Mesh<T> m, const T* __restrict__ vel, const T* __restrict__ p,
                                                   const T* __restrict__ nut, double visc_fac, double* partial) {
    using D = DiagSums<SCHEME>;
    __shared__ double sh[D::NS][kThreads / 64];
    __shared__ double shm[3][kThreads / 64];
    const int iz = blockIdx.y, nch = gridDim.x;
    D acc;
    acc.init();
    const int base = iz * m.nxy;
    const double ihx = m.inv_hx, ihy = m.inv_hy;
    const double inv_hz = SCHEME == TC_SCHEME_AB3 ? 1.0 / (double)m.cellh[iz] : 1.0 / (double)m.DZ[iz];
    const double kv = visc_fac * (ihx * ihx + ihy * ihy + inv_hz * inv_hz);
    const bool top = iz == m.nz - 1;
    const bool nut_node = SCHEME == TC_SCHEME_AB3 && nut && iz != 0 && !top;
    // U cells per thread per pass: every load of a pass is issued before the
    // first sum (the cells of a thread are visited in the same order as a plain
    // strided loop, so the sums are bitwise those of one; a cell past the plane
    // reads the pass's first cell and adds nothing)
    constexpr int U = 4;
    const int stride = nch * kThreads;
    for (int j0 = blockIdx.x * kThreads + threadIdx.x; j0 < m.nxy; j0 += U * stride) {
        double u_[U], v_[U], wr_[U], wn_[U], p_[U], nt_[U], vc_[U];
        bool ok[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int j = j0 + k * stride;
            ok[k] = j < m.nxy;
            const int jj = ok[k] ? j : j0;
            const int in = base + jj;
            u_[k] = vel[3 * in];
            v_[k] = vel[3 * in + 1];
            wr_[k] = vel[3 * in + 2];
            wn_[k] = (SCHEME != TC_SCHEME_AB3 && !top) ? (double)vel[3 * (in + m.nxy) + 2] : 0.0;
            p_[k] = SCHEME != TC_SCHEME_AB3 ? (double)p[in] : 0.0;
            nt_[k] = nut ? (double)nut[in] : 0.0;
            vc_[k] = m.visc;
            if (SCHEME == TC_SCHEME_AB3) {
                if (nut_node) {  // GetNuTAtNode: mean of the 8 elements around the node
                    const int iy = jj / m.nx, ix = jj - iy * m.nx;
                    const int xl = wrapi<T>(ix - 1, m.nx), yl = wrapi<T>(iy - 1, m.ny);
                    const int p0 = iz * m.nxy, p1 = (iz - 1) * m.nxy;
                    const T* n = nut;
                    vc_[k] += 0.125 * ((double)n[p0 + iy * m.nx + ix] + n[p0 + iy * m.nx + xl] + n[p0 + yl * m.nx + ix] +
                                       n[p0 + yl * m.nx + xl] + n[p1 + iy * m.nx + ix] + n[p1 + iy * m.nx + xl] +
                                       n[p1 + yl * m.nx + ix] + n[p1 + yl * m.nx + xl]);
                }
            } else if (nut) {
                vc_[k] = (double)m.visc + nt_[k];
            }
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            if (!ok[k]) continue;
            const double w = SCHEME != TC_SCHEME_AB3 ? 0.5 * (wr_[k] + wn_[k]) : wr_[k];
            acc.add(u_[k], v_[k], w, wr_[k], nut != nullptr, nt_[k], p_[k], vc_[k], ihx, ihy, inv_hz, kv);
        }
    }
    block_reduce_diag<D::NS>(acc.s, sh, false);
    block_reduce_diag<3>(acc.mx, shm, true);
    if (threadIdx.x == 0) {
        double* o = partial + ((size_t)iz * nch + blockIdx.x) * D::NX;
#pragma unroll
        for (int k = 0; k < D::NS; k++) o[k] = acc.s[k];
        for (int k = 0; k < 3; k++) o[D::NS + k] = acc.mx[k];
    }
}

// Chunks combined in order, 8 loads issued before they are added (a chunk
// past nch reads chunk 0 and is skipped).
__global__ void k_diag_final(const double* partial, int nch, int nx_rec, int nsum, double* out) {
    const int iz = blockIdx.x, k = threadIdx.x;
    if (k >= nx_rec) return;
    constexpr int U = 8;
    const double* src = partial + (size_t)iz * nch * nx_rec + k;
    double r = 0.0;
    for (int c0 = 0; c0 < nch; c0 += U) {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = src[(size_t)(c0 + u < nch ? c0 + u : 0) * nx_rec];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (c0 + u < nch) r = k < nsum ? r + x[u] : fmax(r, x[u]);
    }
    out[iz * nx_rec + k] = r;
}

// The same reduction for many chunks per plane (the projection's records: one
// per block of a plane): a block per plane, chunks strided over its threads.
template <int SCHEME>
__global__ __launch_bounds__(kThreads) void k_diag_final_wide(const double* partial, int nch, double* out) {
    using D = DiagSums<SCHEME>;
    __shared__ double sh[D::NS][kThreads / 64];
    __shared__ double shm[3][kThreads / 64];
    const int iz = blockIdx.x;
    D acc;
    acc.init();
    for (int c = threadIdx.x; c < nch; c += kThreads) {
        const double* r = partial + ((size_t)iz * nch + c) * D::NX;
#pragma unroll
        for (int k = 0; k < D::NS; k++) acc.s[k] += r[k];
#pragma unroll
        for (int k = 0; k < 3; k++) acc.mx[k] = fmax(acc.mx[k], r[D::NS + k]);
    }
    block_reduce_diag<D::NS>(acc.s, sh, false);
    block_reduce_diag<3>(acc.mx, shm, true);
    if (threadIdx.x == 0) {
        double* o = out + (size_t)iz * D::NX;
#pragma unroll
        for (int k = 0; k < D::NS; k++) o[k] = acc.s[k];
        for (int k = 0; k < 3; k++) o[D::NS + k] = acc.mx[k];
    }
}

}  // namespace

int diag_record(int scheme) { return (scheme == TC_SCHEME_AB3 ? 10 : 12) + 5; }

hipError_t launch_diag_final_wide(hipStream_t s, int scheme, const double* partial, int nch, int nz, double* out) {
    if (scheme == TC_SCHEME_AB3)
        hipLaunchKernelGGL((k_diag_final_wide<TC_SCHEME_AB3>), dim3(nz), dim3(kThreads), 0, s, partial, nch, out);
    else
        hipLaunchKernelGGL((k_diag_final_wide<TC_SCHEME_HW>), dim3(nz), dim3(kThreads), 0, s, partial, nch, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_diag(hipStream_t s, const Mesh<T>& m, int scheme, const T* v, const T* p, const T* nut,
                       double visc_fac, double* partial, double* out) {
    const int nch = profiles_chunks(m.nxy);
    const dim3 g(nch, m.nz);
    const int nx_rec = diag_record(scheme);
    if (scheme == TC_SCHEME_AB3)
        hipLaunchKernelGGL((k_diag<T, TC_SCHEME_AB3>), g, dim3(kThreads), 0, s, m, v, p, nut, visc_fac, partial);
    else
        hipLaunchKernelGGL((k_diag<T, TC_SCHEME_HW>), g, dim3(kThreads), 0, s, m, v, p, nut, visc_fac, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(m.nz), dim3(64), 0, s, partial, nch, nx_rec, nx_rec - 3, out);
    return hipGetLastError();
}

int profiles_chunks(int nxy) {
    int c = (nxy + 4 * kThreads - 1) / (4 * kThreads);
    return c < 1 ? 1 : (c > 64 ? 64 : c);
}

template <typename T>
hipError_t launch_reduce(hipStream_t s, const Mesh<T>& m, const RedArgs<T>& a, double* partial, double* out) {
    const dim3 g(kRedBlocks), b(kThreads);
    const bool ab3 = a.scheme == TC_SCHEME_AB3;
#define LAUNCH(OP)                                                                                    \
    if (ab3) hipLaunchKernelGGL((k_reduce<T, OP, TC_SCHEME_AB3>), g, b, 0, s, m, a, partial);       \
    else hipLaunchKernelGGL((k_reduce<T, OP, TC_SCHEME_HW>), g, b, 0, s, m, a, partial);
    switch (a.op) {
        case RED_INTEGRAL: LAUNCH(RED_INTEGRAL); break;
        case RED_KE: LAUNCH(RED_KE); break;
        case RED_PSUM: LAUNCH(RED_PSUM); break;
        default: LAUNCH(RED_CFL); break;
    }
#undef LAUNCH
    hipLaunchKernelGGL(k_reduce_final, dim3(1), b, 0, s, partial, kRedBlocks, a.op == RED_CFL ? 1 : 0, out);
    return hipGetLastError();
}


template <typename T>
hipError_t launch_profiles(hipStream_t s, const Mesh<T>& m, int scheme, const T* v, const T* p, const T* nut,
                           double* partial, double* out) {
    const int nch = profiles_chunks(m.nxy);
    const dim3 g(nch, m.nz);
    int nv;
    if (scheme == TC_SCHEME_AB3) {
        hipLaunchKernelGGL((k_profiles<T, TC_SCHEME_AB3>), g, dim3(kThreads), 0, s, m, v, p, nut, partial);
        nv = 10;
    } else {
        hipLaunchKernelGGL((k_profiles<T, TC_SCHEME_HW>), g, dim3(kThreads), 0, s, m, v, p, nut, partial);
        nv = 12;
    }
    hipLaunchKernelGGL(k_profiles_final, dim3(m.nz), dim3(64), 0, s, partial, nch, nv, 1.0 / (double)m.nxy, out);
    return hipGetLastError();
}

#define INST(T)                                                                                                 \
    template hipError_t launch_diag<T>(hipStream_t, const Mesh<T>&, int, const T*, const T*, const T*, double,   \
                                       double*, double*);                                                         \
    template hipError_t launch_reduce<T>(hipStream_t, const Mesh<T>&, const RedArgs<T>&, double*, double*);     \
    template hipError_t launch_profiles<T>(hipStream_t, const Mesh<T>&, int, const T*, const T*, const T*, double*, \
                                           double*);
INST(double)
INST(float)
#undef INST

}  // namespace tc
