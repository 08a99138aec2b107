// Spectral pressure solve: batched mixed-radix FFTs held in LDS and the
// wall-normal banded solves per horizontal mode.
//
// Spectrum layout (complex, half spectrum in x, as the reference's cuFFT R2C
// output, staggered/func_global.h:197-206): spec[(iz*NY + ky)*NXH + kx],
// NXH = NX/2 + 1. The x transforms pack two real rows into one complex
// transform; the y transforms run on tiles of TK adjacent kx columns so that
// every global access is a contiguous run of TK complex values.
#include "kernels.h"
#include "stencil.h"

namespace tc {

namespace {

template <typename T>
__device__ __forceinline__ Cplx<T> cmul(Cplx<T> a, Cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ Cplx<T> cadd(Cplx<T> a, Cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cplx<T> csub(Cplx<T> a, Cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by s*i (s = +-1)
template <typename T>
__device__ __forceinline__ Cplx<T> cmul_si(Cplx<T> a, T s) { return {-s * a.y, s * a.x}; }

// In-register DFT of length R (2, 3, 4, 5); s = -1 forward, +1 inverse.
template <typename T, int R>
__device__ __forceinline__ void bfly(Cplx<T>* v, T s) {
    if (R == 2) {
        const Cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        const Cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const Cplx<T> t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), s);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if (R == 3) {
        const T c = T(-0.5), sn = T(0.86602540378443864676);  // cos, sin of 2 pi / 3
        const Cplx<T> b = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const Cplx<T> a = v[0];
        const Cplx<T> t = {a.x + c * b.x, a.y + c * b.y};
        const Cplx<T> r = cmul_si(Cplx<T>{sn * d.x, sn * d.y}, s);
        v[0] = cadd(a, b);
        v[1] = cadd(t, r);
        v[2] = csub(t, r);
    } else {  // R == 5
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);  // cos 2pi/5, cos 4pi/5
        const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);   // sin 2pi/5, sin 4pi/5
        const Cplx<T> a = v[0];
        const Cplx<T> b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        const Cplx<T> b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        const Cplx<T> t1 = {a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y};
        const Cplx<T> t2 = {a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y};
        const Cplx<T> r1 = cmul_si(Cplx<T>{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y}, s);
        const Cplx<T> r2 = cmul_si(Cplx<T>{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y}, s);
        v[0] = cadd(a, cadd(b1, b2));
        v[1] = cadd(t1, r1);
        v[4] = csub(t1, r1);
        v[2] = cadd(t2, r2);
        v[3] = csub(t2, r2);
    }
}

// One Stockham pass of radix R over nlines lines of length N (Ns = product of
// the radices already applied). Element e of line l lives at
// buf[l*lstride + e*estride]; tw[k] = exp(-2 pi i k/N). Butterflies are
// numbered so that consecutive threads touch consecutive LDS words:
// line-fastest when the lines are interleaved (estride > 1).
template <typename T, int R>
__device__ __forceinline__ void stockham_pass(const Cplx<T>* __restrict__ src, Cplx<T>* __restrict__ dst, int N, int Ns,
                                              int nlines, int lstride, int estride, const Cplx<T>* __restrict__ tw, T s) {
    const int NR = N / R;
    const int total = NR * nlines;
    const int tstep = N / (Ns * R);
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int l, j;
        if (estride == 1) {
            l = t / NR;
            j = t - l * NR;
        } else {
            j = t / nlines;
            l = t - j * nlines;
        }
        Cplx<T> v[R];
        const Cplx<T>* sp = src + l * lstride;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = sp[(j + r * NR) * estride];
        const int jm = j % Ns;
        if (jm != 0) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                Cplx<T> w = tw[r * jm * tstep];
                w.y *= -s;  // forward table; conjugate for the inverse
                v[r] = cmul(v[r], w);
            }
        }
        bfly<T, R>(v, s);
        const int o = (j - jm) * R + jm;
        Cplx<T>* dp = dst + l * lstride;
#pragma unroll
        for (int r = 0; r < R; r++) dp[(o + r * Ns) * estride] = v[r];
    }
}

// Stockham autosort transform held in LDS; returns the buffer with the result.
template <typename T>
__device__ Cplx<T>* lds_fft(Cplx<T>* src, Cplx<T>* dst, const FFTPlan& pl, int nlines, int lstride, int estride,
                            const Cplx<T>* __restrict__ tw, T s) {
    const int N = pl.n;
    int Ns = 1;
    for (int st = 0; st < pl.nst; st++) {
        const int R = pl.radix[st];
        switch (R) {
            case 2: stockham_pass<T, 2>(src, dst, N, Ns, nlines, lstride, estride, tw, s); break;
            case 3: stockham_pass<T, 3>(src, dst, N, Ns, nlines, lstride, estride, tw, s); break;
            case 4: stockham_pass<T, 4>(src, dst, N, Ns, nlines, lstride, estride, tw, s); break;
            default: stockham_pass<T, 5>(src, dst, N, Ns, nlines, lstride, estride, tw, s); break;
        }
        __syncthreads();
        Cplx<T>* tmp = src;
        src = dst;
        dst = tmp;
        Ns *= R;
    }
    return src;
}

constexpr int kFftThreads = 256;

// Forward x transform of the divergence of v: rows 2q, 2q+1 (row = iy + NY*iz)
// form the real and imaginary parts of line q; the divergence is evaluated
// while the lines are staged into LDS.
template <typename T, int SCHEME>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_fwd_div(FFTPlan pl, const Cplx<T>* __restrict__ tw, Mesh<T> m,
                                                                const T* __restrict__ v, Cplx<T>* __restrict__ spec,
                                                                int lpb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n, nxh = N / 2 + 1, nrows = m.ny * m.nz;
    Cplx<T>* b1 = b0 + lpb * N;
    const int q0 = blockIdx.x * lpb;
    constexpr int R = SCHEME == TC_SCHEME_VV ? 2 : 1;
    for (int t = threadIdx.x; t < lpb * N; t += blockDim.x) {
        const int l = t / N, n = t - l * N;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        Cplx<T> z = {T(0), T(0)};
        if (ra < nrows) {
            const int iz = ra / m.ny, iy = ra - iz * m.ny;
            z.x = div_cell<T, SCHEME>(m, make_nbvec<T, R>(v, n, iy, iz, m.nx, m.ny), iz);
        }
        if (rb < nrows) {
            const int iz = rb / m.ny, iy = rb - iz * m.ny;
            z.y = div_cell<T, SCHEME>(m, make_nbvec<T, R>(v, n, iy, iz, m.nx, m.ny), iz);
        }
        b0[t] = z;
    }
    __syncthreads();
    const Cplx<T>* res = lds_fft(b0, b1, pl, lpb, N, 1, tw, T(-1));
    for (int t = threadIdx.x; t < lpb * nxh; t += blockDim.x) {
        const int l = t / nxh, k = t - l * nxh;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        if (ra >= nrows) continue;
        const Cplx<T> z = res[l * N + k];
        const Cplx<T> zc = res[l * N + (k == 0 ? 0 : N - k)];
        // A = (Z + conj Zc)/2, B = -i (Z - conj Zc)/2
        spec[(size_t)ra * nxh + k] = {T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y)};
        if (rb < nrows) spec[(size_t)rb * nxh + k] = {T(0.5) * (z.y + zc.y), T(-0.5) * (z.x - zc.x)};
    }
}

// Inverse x transform: half spectra of rows 2q, 2q+1 -> real rows (times scale).
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_inv(FFTPlan pl, const Cplx<T>* __restrict__ tw,
                                                            const Cplx<T>* __restrict__ spec, T* __restrict__ out,
                                                            int nrows, int lpb, T scale) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n, nxh = N / 2 + 1, half = N / 2;
    Cplx<T>* b1 = b0 + lpb * N;
    const int q0 = blockIdx.x * lpb;
    for (int t = threadIdx.x; t < lpb * N; t += blockDim.x) {
        const int l = t / N, k = t - l * N;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        const int kk = k <= half ? k : N - k;
        Cplx<T> A = {T(0), T(0)}, B = {T(0), T(0)};
        if (ra < nrows) A = spec[(size_t)ra * nxh + kk];
        if (rb < nrows) B = spec[(size_t)rb * nxh + kk];
        if (kk == 0 || kk == half) {  // real-to-complex symmetry: these bins are real
            A.y = T(0);
            B.y = T(0);
        }
        if (k > half) {
            A.y = -A.y;
            B.y = -B.y;
        }
        b0[t] = {A.x - B.y, A.y + B.x};  // A + iB
    }
    __syncthreads();
    const Cplx<T>* res = lds_fft(b0, b1, pl, lpb, N, 1, tw, T(1));
    for (int t = threadIdx.x; t < lpb * N; t += blockDim.x) {
        const int l = t / N, n = t - l * N;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        if (ra < nrows) out[(size_t)ra * N + n] = res[t].x * scale;
        if (rb < nrows) out[(size_t)rb * N + n] = res[t].y * scale;
    }
}

// y transform of TK adjacent kx columns of one plane, in place.
// With psum != nullptr (inverse transform of the pressure solve) the (0,0) mode
// of each plane first gets the normalisation X00[iz] -= psum * C[iz] (FP64).
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_y(FFTPlan pl, const Cplx<T>* __restrict__ tw, Cplx<T>* spec,
                                                        int nxh, int tk, T s, const double* __restrict__ psum,
                                                        const double* __restrict__ C) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n;
    Cplx<T>* b1 = b0 + tk * N;
    const int kx0 = blockIdx.x * tk, iz = blockIdx.y;
    const int ncol = min(tk, nxh - kx0);
    Cplx<T>* base = spec + (size_t)iz * N * nxh + kx0;
    for (int t = threadIdx.x; t < tk * N; t += blockDim.x) {
        const int n = t / tk, k = t - n * tk;
        Cplx<T> v = k < ncol ? base[(size_t)n * nxh + k] : Cplx<T>{T(0), T(0)};
        if (psum && t == 0 && kx0 == 0) v.x = T((double)v.x - psum[0] * C[iz]);
        b0[t] = v;
    }
    __syncthreads();
    const Cplx<T>* res = lds_fft(b0, b1, pl, tk, 1, tk, tw, s);
    for (int t = threadIdx.x; t < tk * N; t += blockDim.x) {
        const int n = t / tk, k = t - n * tk;
        if (k < ncol) base[(size_t)n * nxh + k] = res[t];
    }
}

// ---------------------------------------------------------------------------
// Wall-normal banded solves (ThomasForPressure: staggered/func_global.h:71-189,
// collocated_AB3/func_global.h:8-61) with factors precomputed on the host.
// One thread per (mode, real/imaginary part): the matrix is real.
// ---------------------------------------------------------------------------
template <typename T, int W>
__device__ void zsolve_low_lane(Cplx<T>* spec, const ZFacDev<double>& f, int li, int part, int nm, int mode,
                                double* ylow, int nlow, const ZNorm& norm);

template <typename T, int W>
__global__ __launch_bounds__(256) void k_zsolve(Cplx<T>* spec, ZFacDev<T> f, ZFacDev<double> flow, const int* lowidx,
                                                 int nlow, double* ylow, ZNorm norm) {
    constexpr int B = 8;  // rows whose loads are issued together, ahead of the recurrence
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nm = f.ncols, nz = f.nz;
    const int mode = t >> 1, part = t & 1;
    if (mode >= nm) return;
    if (lowidx) {  // FP32-bulk: the lowest wavenumbers take the FP64 path
        const int li = lowidx[mode];
        if (li >= 0) {
            zsolve_low_lane<T, W>(spec, flow, li, part, nm, mode, ylow, nlow, norm);
            return;
        }
    }
    T* __restrict__ x = reinterpret_cast<T*>(spec) + 2 * mode + part;
    const size_t xs = 2 * (size_t)nm;  // stride between planes, in reals
    const T* __restrict__ M1 = f.M1 + mode;
    const T* __restrict__ M2 = f.M2 + mode;
    const T* __restrict__ M3 = f.M3 + mode;
    const T* __restrict__ D = f.D + mode;
    const T* __restrict__ U1 = f.U1 + mode;
    const T* __restrict__ U2 = f.U2 + mode;
    const T* __restrict__ U3 = f.U3 + mode;
    T y1 = 0, y2 = 0, y3 = 0;
    // forward substitution
    for (int i0 = 0; i0 < nz; i0 += B) {
        T b[B], m1[B], m2[B], m3[B], d[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 + k;
            if (i < nz) {
                const size_t fi = (size_t)i * nm;
                b[k] = x[i * xs];
                m1[k] = M1[fi];
                d[k] = D[fi];
                if (W == 3) {
                    m2[k] = M2[fi];
                    m3[k] = M3[fi];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 + k;
            if (i < nz) {
                T v = b[k];
                if (W == 3) v = v + m3[k] * y3;
                if (W == 3) v = v + m2[k] * y2;
                v = v + m1[k] * y1;
                v = v * d[k];
                x[i * xs] = v;
                y3 = y2;
                y2 = y1;
                y1 = v;
            }
        }
    }
    // backward substitution
    const bool n00 = (mode == 0 && part == 0 && norm.s00 != nullptr);
    double psum = 0.0;
    T z1 = 0, z2 = 0, z3 = 0;
    for (int i0 = nz - 1; i0 >= 0; i0 -= B) {
        T b[B], u1[B], u2[B], u3[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 - k;
            if (i >= 0) {
                const size_t fi = (size_t)i * nm;
                b[k] = x[i * xs];
                u1[k] = U1[fi];
                if (W == 3) {
                    u2[k] = U2[fi];
                    u3[k] = U3[fi];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 - k;
            if (i >= 0) {
                T v = b[k];
                v = v - u1[k] * z1;
                if (W == 3) v = v - u2[k] * z2;
                if (W == 3) v = v - u3[k] * z3;
                x[i * xs] = v;
                if (n00) psum += norm.Wt[i] * (double)v;
                z3 = z2;
                z2 = z1;
                z1 = v;
            }
        }
    }
    if (n00) norm.s00[0] = psum;  // applied by the inverse y transform (NormalizePressure)
}

// FP64 solve of one listed mode (the FP32-bulk engine's low wavenumbers). The
// intermediate of the forward pass stays in double (ylow: nz * nlow * 2).
template <typename T, int W>
__device__ void zsolve_low_lane(Cplx<T>* spec, const ZFacDev<double>& f, int li, int part, int nm, int mode,
                                double* ylow, int nlow, const ZNorm& norm) {
    constexpr int B = 8;
    const int nz = f.nz;
    T* __restrict__ x = reinterpret_cast<T*>(spec) + 2 * mode + part;
    const size_t xs = 2 * (size_t)nm;
    double* __restrict__ y = ylow + 2 * li + part;
    const size_t ys = 2 * (size_t)nlow;
    double y1 = 0, y2 = 0, y3 = 0;
    for (int i0 = 0; i0 < nz; i0 += B) {
        double b[B], m1[B], m2[B], m3[B], d[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 + k;
            if (i < nz) {
                const size_t fi = (size_t)i * nlow + li;
                b[k] = (double)x[i * xs];
                m1[k] = f.M1[fi];
                d[k] = f.D[fi];
                if (W == 3) {
                    m2[k] = f.M2[fi];
                    m3[k] = f.M3[fi];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 + k;
            if (i < nz) {
                double v = b[k];
                if (W == 3) v = v + m3[k] * y3;
                if (W == 3) v = v + m2[k] * y2;
                v = v + m1[k] * y1;
                v = v * d[k];
                y[i * ys] = v;
                y3 = y2;
                y2 = y1;
                y1 = v;
            }
        }
    }
    const bool n00 = (mode == 0 && part == 0);
    double psum = 0.0, z1 = 0, z2 = 0, z3 = 0;
    for (int i0 = nz - 1; i0 >= 0; i0 -= B) {
        double b[B], u1[B], u2[B], u3[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 - k;
            if (i >= 0) {
                const size_t fi = (size_t)i * nlow + li;
                b[k] = y[i * ys];
                u1[k] = f.U1[fi];
                if (W == 3) {
                    u2[k] = f.U2[fi];
                    u3[k] = f.U3[fi];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int i = i0 - k;
            if (i >= 0) {
                double v = b[k];
                v = v - u1[k] * z1;
                if (W == 3) v = v - u2[k] * z2;
                if (W == 3) v = v - u3[k] * z3;
                x[i * xs] = T(v);
                if (n00) psum += norm.Wt[i] * v;
                z3 = z2;
                z2 = z1;
                z1 = v;
            }
        }
    }
    if (n00) norm.s00[0] = psum;  // applied by the inverse y transform (NormalizePressure)
}

// Forward x transform of one component (cb < 0) or of the product of two
// components of a real3 field (training samples of the learned closure).
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_fwd_q(FFTPlan pl, const Cplx<T>* __restrict__ tw,
                                                              const T* __restrict__ v, int ca, int cb, int nrows,
                                                              Cplx<T>* __restrict__ spec, int lpb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n, nxh = N / 2 + 1;
    Cplx<T>* b1 = b0 + lpb * N;
    const int q0 = blockIdx.x * lpb;
    auto q = [&](int row, int n) {
        const T* c = v + 3 * ((size_t)row * N + n);
        return cb < 0 ? c[ca] : c[ca] * c[cb];
    };
    for (int t = threadIdx.x; t < lpb * N; t += blockDim.x) {
        const int l = t / N, n = t - l * N;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        Cplx<T> z = {T(0), T(0)};
        if (ra < nrows) z.x = q(ra, n);
        if (rb < nrows) z.y = q(rb, n);
        b0[t] = z;
    }
    __syncthreads();
    const Cplx<T>* res = lds_fft(b0, b1, pl, lpb, N, 1, tw, T(-1));
    for (int t = threadIdx.x; t < lpb * nxh; t += blockDim.x) {
        const int l = t / nxh, k = t - l * nxh;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        if (ra >= nrows) continue;
        const Cplx<T> z = res[l * N + k];
        const Cplx<T> zc = res[l * N + (k == 0 ? 0 : N - k)];
        spec[(size_t)ra * nxh + k] = {T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y)};
        if (rb < nrows) spec[(size_t)rb * nxh + k] = {T(0.5) * (z.y + zc.y), T(-0.5) * (z.x - zc.x)};
    }
}

// Half spectrum of the coarse grid holding the modes a sharp cut-off keeps:
// |kx| < nxc/2 and |ky| < nyc/2 (the coarse Nyquist modes are zero).
template <typename T>
__global__ void k_spec_crop(const Cplx<T>* __restrict__ fine, int nxh, int ny, Cplx<T>* __restrict__ coarse, int nxhc,
                            int nyc, int nz) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nxhc * nyc * nz) return;
    const int kx = (int)(t % nxhc);
    const long long rest = t / nxhc;
    const int kyc = (int)(rest % nyc), iz = (int)(rest / nyc);
    const int nxc = 2 * (nxhc - 1), hy = nyc / 2;
    Cplx<T> v = {T(0), T(0)};
    if (kx < nxc / 2) {
        const int ky = kyc < hy ? kyc : (kyc > hy ? ny - (nyc - kyc) : -1);
        if (ky >= 0) v = fine[((size_t)iz * ny + ky) * nxh + kx];
    }
    coarse[t] = v;
}

template <typename T>
int lines_per_block(int n) {
    // two ping-pong buffers per line; keep a block near 24 KiB of LDS
    const int per_line = 2 * n * (int)sizeof(Cplx<T>);
    int l = (24 * 1024) / per_line;
    return l < 1 ? 1 : (l > 16 ? 16 : l);
}

}  // namespace

template <typename T>
hipError_t launch_fft_x_fwd_div(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const Mesh<T>& m, int scheme,
                                const T* v, Cplx<T>* spec) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (m.ny * m.nz + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = 2 * (size_t)lpb * pl.n * sizeof(Cplx<T>);
    if (scheme == TC_SCHEME_HW)
        hipLaunchKernelGGL((k_fft_x_fwd_div<T, TC_SCHEME_HW>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, m, v, spec, lpb);
    else if (scheme == TC_SCHEME_VV)
        hipLaunchKernelGGL((k_fft_x_fwd_div<T, TC_SCHEME_VV>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, m, v, spec, lpb);
    else
        hipLaunchKernelGGL((k_fft_x_fwd_div<T, TC_SCHEME_AB3>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, m, v, spec, lpb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_x_inv(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const Cplx<T>* spec, T* out, int nrows,
                            T scale) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (nrows + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = 2 * (size_t)lpb * pl.n * sizeof(Cplx<T>);
    hipLaunchKernelGGL((k_fft_x_inv<T>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, spec, out, nrows, lpb, scale);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_y(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, Cplx<T>* spec, int nxh, int nz, int inverse,
                        const double* psum, const double* C) {
    // 128-byte runs (8 complex doubles, 16 complex floats) while the block stays within 48 KiB of LDS
    int tk = 128 / (int)sizeof(Cplx<T>);
    while (tk > 1 && 2 * (size_t)tk * pl.n * sizeof(Cplx<T>) > 48 * 1024) tk >>= 1;
    const dim3 grid((nxh + tk - 1) / tk, nz);
    const size_t sm = 2 * (size_t)tk * pl.n * sizeof(Cplx<T>);
    hipLaunchKernelGGL((k_fft_y<T>), grid, dim3(kFftThreads), sm, s, pl, tw, spec, nxh, tk, T(inverse ? 1 : -1), psum, C);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zsolve(hipStream_t s, Cplx<T>* spec, const ZFacDev<T>& f, const ZFacDev<double>& flow,
                         const int* lowidx, int nlow, double* ylow, ZNorm norm) {
    const int nthreads = 2 * f.ncols;
    const int grid = (nthreads + 255) / 256;
    if (f.W == 3) hipLaunchKernelGGL((k_zsolve<T, 3>), dim3(grid), dim3(256), 0, s, spec, f, flow, lowidx, nlow, ylow, norm);
    else hipLaunchKernelGGL((k_zsolve<T, 1>), dim3(grid), dim3(256), 0, s, spec, f, flow, lowidx, nlow, ylow, norm);
    return hipGetLastError();
}


template <typename T>
hipError_t launch_fft_x_fwd_q(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const T* v, int ca, int cb,
                              int nrows, Cplx<T>* spec) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (nrows + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = 2 * (size_t)lpb * pl.n * sizeof(Cplx<T>);
    hipLaunchKernelGGL((k_fft_x_fwd_q<T>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, v, ca, cb, nrows, spec, lpb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_spec_crop(hipStream_t s, const Cplx<T>* fine, int nxh, int ny, Cplx<T>* coarse, int nxhc, int nyc,
                            int nz) {
    const long long n = (long long)nxhc * nyc * nz;
    hipLaunchKernelGGL((k_spec_crop<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fine, nxh, ny, coarse, nxhc,
                       nyc, nz);
    return hipGetLastError();
}

#define INST(T)                                                                                                      \
    template hipError_t launch_fft_x_fwd_div<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const Mesh<T>&, int,    \
                                                const T*, Cplx<T>*);                                                 \
    template hipError_t launch_fft_x_inv<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const Cplx<T>*, T*, int, T); \
    template hipError_t launch_fft_y<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, Cplx<T>*, int, int, int,       \
                                        const double*, const double*);                                              \
    template hipError_t launch_zsolve<T>(hipStream_t, Cplx<T>*, const ZFacDev<T>&, const ZFacDev<double>&,         \
                                         const int*, int, double*, ZNorm);                                   \
    template hipError_t launch_fft_x_fwd_q<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const T*, int, int, int,  \
                                              Cplx<T>*);                                                             \
    template hipError_t launch_spec_crop<T>(hipStream_t, const Cplx<T>*, int, int, Cplx<T>*, int, int, int);
INST(double)
INST(float)
#undef INST

}  // namespace tc
