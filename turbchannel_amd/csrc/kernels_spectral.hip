// Spectral pressure solve: batched mixed-radix FFTs held in LDS and the
// wall-normal banded solves per horizontal mode.
//
// Spectrum layout (complex, half spectrum in x, as the reference's cuFFT R2C
// output, staggered/func_global.h:197-206): spec[(iz*NY + ky)*NXH + kx],
// NXH = NX/2 + 1. The x transforms pack two real rows into one complex
// transform; the y transforms run on tiles of TK adjacent kx columns so that
// every global access is a contiguous run of TK complex values.
#include "kernels.h"

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
template <typename T>
__device__ __forceinline__ void bfly(Cplx<T>* v, int R, T s) {
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

// Stockham autosort transform of nlines lines of length N held in LDS.
// Element e of line l lives at buf[l*lstride + e*estride]. tw[k] = exp(-2 pi i k/N).
// Butterflies are numbered so that consecutive threads touch consecutive LDS
// words: line-fastest when the lines are interleaved (estride > 1).
template <typename T>
__device__ Cplx<T>* lds_fft(Cplx<T>* src, Cplx<T>* dst, const FFTPlan& pl, int nlines, int lstride, int estride,
                            const Cplx<T>* __restrict__ tw, T s) {
    const int N = pl.n;
    int Ns = 1;
    for (int st = 0; st < pl.nst; st++) {
        const int R = pl.radix[st];
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
            Cplx<T> v[5];
            const Cplx<T>* sp = src + l * lstride;
            for (int r = 0; r < R; r++) v[r] = sp[(j + r * NR) * estride];
            const int jm = j % Ns;
            if (jm != 0) {
                for (int r = 1; r < R; r++) {
                    Cplx<T> w = tw[r * jm * tstep];
                    w.y *= -s;  // forward table; conjugate for the inverse
                    v[r] = cmul(v[r], w);
                }
            }
            bfly(v, R, s);
            const int o = (j - jm) * R + jm;
            Cplx<T>* dp = dst + l * lstride;
            for (int r = 0; r < R; r++) dp[(o + r * Ns) * estride] = v[r];
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

// Forward x transform: rows 2q, 2q+1 of the real input form line q.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_fwd(FFTPlan pl, const Cplx<T>* __restrict__ tw,
                                                            const T* __restrict__ in, Cplx<T>* __restrict__ spec,
                                                            int nrows, int lpb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n, nxh = N / 2 + 1;
    Cplx<T>* b1 = b0 + lpb * N;
    const int q0 = blockIdx.x * lpb;
    for (int t = threadIdx.x; t < lpb * N; t += blockDim.x) {
        const int l = t / N, n = t - l * N;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        Cplx<T> z = {T(0), T(0)};
        if (ra < nrows) z.x = in[(size_t)ra * N + n];
        if (rb < nrows) z.y = in[(size_t)rb * N + n];
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
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_y(FFTPlan pl, const Cplx<T>* __restrict__ tw, Cplx<T>* spec,
                                                        int nxh, int tk, T s) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n;
    Cplx<T>* b1 = b0 + tk * N;
    const int kx0 = blockIdx.x * tk, iz = blockIdx.y;
    const int ncol = min(tk, nxh - kx0);
    Cplx<T>* base = spec + (size_t)iz * N * nxh + kx0;
    for (int t = threadIdx.x; t < tk * N; t += blockDim.x) {
        const int n = t / tk, k = t - n * tk;
        b0[t] = k < ncol ? base[(size_t)n * nxh + k] : Cplx<T>{T(0), T(0)};
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
__device__ __forceinline__ bool is_low_mode(int kx, int ky, int ny, int klow) {
    return kx <= klow && (ky <= klow || ky >= ny - klow);
}

template <typename T, int W>
__global__ __launch_bounds__(256) void k_zsolve(Cplx<T>* spec, ZFacDev<T> f, int nxh, int ny, int skip_low, int klow,
                                                 ZNorm norm) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nm = f.ncols, nz = f.nz;
    const int mode = t >> 1, part = t & 1;
    if (mode >= nm) return;
    const int kx = mode % nxh, ky = mode / nxh;
    if (skip_low && is_low_mode(kx, ky, ny, klow)) return;
    T* x = reinterpret_cast<T*>(spec) + 2 * mode + part;
    const size_t xs = 2 * (size_t)nm;  // stride between planes, in reals
    T y1 = 0, y2 = 0, y3 = 0;
    // forward substitution
    for (int i = 0; i < nz; i++) {
        const size_t fi = (size_t)i * nm + mode;
        T v = x[i * xs];
        if (W == 3) v = v + f.M3[fi] * y3;
        if (W == 3) v = v + f.M2[fi] * y2;
        v = v + f.M1[fi] * y1;
        v = v * f.D[fi];
        x[i * xs] = v;
        y3 = y2;
        y2 = y1;
        y1 = v;
    }
    // backward substitution
    const bool n00 = (mode == 0 && part == 0 && norm.s00 != nullptr);
    double psum = 0.0;
    T z1 = 0, z2 = 0, z3 = 0;
    for (int i = nz - 1; i >= 0; i--) {
        const size_t fi = (size_t)i * nm + mode;
        T v = x[i * xs];
        v = v - f.U1[fi] * z1;
        if (W == 3) v = v - f.U2[fi] * z2;
        if (W == 3) v = v - f.U3[fi] * z3;
        x[i * xs] = v;
        if (n00) {
            psum += norm.Wt[i] * (double)v;
            norm.s00[i] = (double)v;
        }
        z3 = z2;
        z2 = z1;
        z1 = v;
    }
    if (n00)  // NormalizePressure of the (0,0) mode, in spectral space
        for (int i = 0; i < nz; i++) x[i * xs] = T(norm.s00[i] - psum * norm.C[i]);
}

// FP64 solves of a short list of modes (the FP32-bulk engine's low wavenumbers).
template <typename T, int W>
__global__ __launch_bounds__(64) void k_zsolve_low(Cplx<T>* spec, ZFacDev<double> f, const int* __restrict__ modes,
                                                   int nlow, int nm, double* ylow, ZNorm norm) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int li = t >> 1, part = t & 1;
    if (li >= nlow) return;
    const int mode = modes[li];
    const int nz = f.nz;
    T* x = reinterpret_cast<T*>(spec) + 2 * mode + part;
    const size_t xs = 2 * (size_t)nm;
    double* y = ylow + t;
    const size_t ys = 2 * (size_t)nlow;
    double y1 = 0, y2 = 0, y3 = 0;
    for (int i = 0; i < nz; i++) {
        const size_t fi = (size_t)i * nlow + li;
        double v = (double)x[i * xs];
        if (W == 3) v = v + f.M3[fi] * y3;
        if (W == 3) v = v + f.M2[fi] * y2;
        v = v + f.M1[fi] * y1;
        v = v * f.D[fi];
        y[i * ys] = v;
        y3 = y2;
        y2 = y1;
        y1 = v;
    }
    const bool n00 = (mode == 0 && part == 0);
    double psum = 0.0, z1 = 0, z2 = 0, z3 = 0;
    for (int i = nz - 1; i >= 0; i--) {
        const size_t fi = (size_t)i * nlow + li;
        double v = y[i * ys];
        v = v - f.U1[fi] * z1;
        if (W == 3) v = v - f.U2[fi] * z2;
        if (W == 3) v = v - f.U3[fi] * z3;
        y[i * ys] = v;
        if (n00) psum += norm.Wt[i] * v;
        z3 = z2;
        z2 = z1;
        z1 = v;
    }
    for (int i = 0; i < nz; i++) {
        double v = y[i * ys];
        if (n00) v -= psum * norm.C[i];
        x[i * xs] = T(v);
    }
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
hipError_t launch_fft_x_fwd(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const T* in, Cplx<T>* spec, int nrows) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (nrows + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = 2 * (size_t)lpb * pl.n * sizeof(Cplx<T>);
    hipLaunchKernelGGL((k_fft_x_fwd<T>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, in, spec, nrows, lpb);
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
hipError_t launch_fft_y(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, Cplx<T>* spec, int nxh, int nz, int inverse) {
    // 128-byte runs (8 complex doubles, 16 complex floats) while the block stays within 48 KiB of LDS
    int tk = 128 / (int)sizeof(Cplx<T>);
    while (tk > 1 && 2 * (size_t)tk * pl.n * sizeof(Cplx<T>) > 48 * 1024) tk >>= 1;
    const dim3 grid((nxh + tk - 1) / tk, nz);
    const size_t sm = 2 * (size_t)tk * pl.n * sizeof(Cplx<T>);
    hipLaunchKernelGGL((k_fft_y<T>), grid, dim3(kFftThreads), sm, s, pl, tw, spec, nxh, tk, T(inverse ? 1 : -1));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zsolve(hipStream_t s, Cplx<T>* spec, const ZFacDev<T>& f, int nxh, int ny, int skip_low, int klow,
                         ZNorm norm) {
    const int nthreads = 2 * f.ncols;
    const int grid = (nthreads + 255) / 256;
    if (f.W == 3) hipLaunchKernelGGL((k_zsolve<T, 3>), dim3(grid), dim3(256), 0, s, spec, f, nxh, ny, skip_low, klow, norm);
    else hipLaunchKernelGGL((k_zsolve<T, 1>), dim3(grid), dim3(256), 0, s, spec, f, nxh, ny, skip_low, klow, norm);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zsolve_low(hipStream_t s, Cplx<T>* spec, const ZFacDev<double>& f, const int* modes, int nlow, int nm,
                             double* ylow, ZNorm norm) {
    const int grid = (2 * nlow + 63) / 64;
    if (f.W == 3) hipLaunchKernelGGL((k_zsolve_low<T, 3>), dim3(grid), dim3(64), 0, s, spec, f, modes, nlow, nm, ylow, norm);
    else hipLaunchKernelGGL((k_zsolve_low<T, 1>), dim3(grid), dim3(64), 0, s, spec, f, modes, nlow, nm, ylow, norm);
    return hipGetLastError();
}

#define INST(T)                                                                                                      \
    template hipError_t launch_fft_x_fwd<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const T*, Cplx<T>*, int);   \
    template hipError_t launch_fft_x_inv<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const Cplx<T>*, T*, int, T); \
    template hipError_t launch_fft_y<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, Cplx<T>*, int, int, int);       \
    template hipError_t launch_zsolve<T>(hipStream_t, Cplx<T>*, const ZFacDev<T>&, int, int, int, int, ZNorm);      \
    template hipError_t launch_zsolve_low<T>(hipStream_t, Cplx<T>*, const ZFacDev<double>&, const int*, int, int,   \
                                             double*, ZNorm);
INST(double)
INST(float)
#undef INST

}  // namespace tc
