// Spectral pressure solve: batched mixed-radix FFTs held in LDS and the
// wall-normal banded solves per horizontal mode.
//
// Spectrum layout (complex, half spectrum in x as the reference's cuFFT R2C
// output, staggered/func_global.h:197-206, stored ky-fastest):
// spec[(iz*NXH + kx)*NY + ky], NXH = NX/2 + 1, mode = kx*NY + ky. The x
// transforms pack two real rows into one complex transform and exchange the
// spectra with runs of consecutive ky; the y transforms then run on contiguous
// lines like the x transforms.
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

// In-register DFT of length R (2, 3, 4, 5, 6, 8); s = -1 forward, +1 inverse.
template <typename T, int R>
__device__ __forceinline__ void bfly(Cplx<T>* v, T s);

// Length 2P from two length-P DFTs of the even and odd inputs (decimation in
// time): X[k] = A[k] + w^k B[k], X[k+P] = A[k] - w^k B[k], w = exp(s 2 pi i / 2P).
template <typename T, int P>
__device__ __forceinline__ void bfly_split(Cplx<T>* v, T s, const Cplx<T>* w) {
    Cplx<T> a[P], b[P];
#pragma unroll
    for (int k = 0; k < P; k++) {
        a[k] = v[2 * k];
        b[k] = v[2 * k + 1];
    }
    bfly<T, P>(a, s);
    bfly<T, P>(b, s);
#pragma unroll
    for (int k = 0; k < P; k++) {
        const Cplx<T> t = k == 0 ? b[0] : cmul(b[k], w[k]);
        v[k] = cadd(a[k], t);
        v[k + P] = csub(a[k], t);
    }
}

template <typename T, int R>
__device__ __forceinline__ void bfly(Cplx<T>* v, T s) {
    if (R == 8) {
        const T c = T(0.70710678118654752440);  // cos(pi/4)
        const Cplx<T> w[4] = {{T(1), T(0)}, {c, s * c}, {T(0), s}, {-c, s * c}};
        bfly_split<T, 4>(v, s, w);
    } else if (R == 6) {
        const T sn = T(0.86602540378443864676);  // sin(pi/3)
        const Cplx<T> w[3] = {{T(1), T(0)}, {T(0.5), s * sn}, {T(-0.5), s * sn}};
        bfly_split<T, 3>(v, s, w);
    } else if (R == 2) {
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

// LDS element index with one pad element every 4: the passes read and write
// with strides of the sub-transform spans (48, 6 and 1 elements at 384 =
// 8 * 8 * 6), which would otherwise fall on few banks. One pad every 16 left
// 2-4 bank-conflict cycles per LDS instruction (SQ counters); every 4 takes the
// 384 x 384 x 192 FP32 transforms 141 -> 135 us (forward x with the divergence,
// collocated), 168 -> 158 us (fourth order), 40.5 -> 38.5 us (each y), 43.5 ->
// 41.4 us (inverse x); every 8 and every 32 gained less.
__device__ __forceinline__ int ldsi(int i) { return i + (i >> 2); }
__host__ __device__ constexpr size_t lds_elems(size_t n) { return n + (n >> 2) + 1; }

// One in-place decimation-in-frequency pass of radix R over nlines lines of
// length N. Element e of line l lives at buf[ldsi(l*lstride + e*estride)];
// tw[ldsi(k)] = exp(-2 pi i k/N). A butterfly reads the R elements j + r*S of
// its sub-transform (length L = R*S), applies the length-R DFT, multiplies
// output r by exp(s 2 pi i j r / L) and writes back to the same places, so
// one buffer serves the whole transform. Butterflies are numbered so that
// consecutive threads touch consecutive LDS words: line-fastest when the lines
// are interleaved (estride > 1).
template <typename T, int R>
__device__ __forceinline__ void dif_pass(Cplx<T>* buf, const FFTPlan& pl, int st, int nlines, FastDiv dl, int lstride,
                                         int estride, const Cplx<T>* __restrict__ tw, T s) {
    const int NR = (int)pl.dnr[st].d, S = (int)pl.dspan[st].d, L = S * R, tws = pl.twstride[st];
    const int total = NR * nlines;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int l, b;
        if (estride == 1) {
            l = pl.dnr[st].div(t);
            b = t - l * NR;
        } else {
            b = dl.div(t);
            l = t - b * nlines;
        }
        const int blk = pl.dspan[st].div(b), j = b - blk * S;
        const int e0 = l * lstride + (blk * L + j) * estride, es = S * estride;
        Cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = buf[ldsi(e0 + r * es)];
        bfly<T, R>(v, s);
        if (j != 0) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                Cplx<T> w = tw[ldsi(r * j * tws)];
                w.y *= -s;  // forward table; conjugate for the inverse
                v[r] = cmul(v[r], w);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) buf[ldsi(e0 + r * es)] = v[r];
    }
}

// Transform held in one LDS buffer, in place; the result of frequency k sits
// at position dif_pos(pl, k) (mixed-radix digit reversal). Call after the
// barrier that ends the staging; ends with a barrier.
template <typename T>
__device__ void lds_fft(Cplx<T>* buf, const FFTPlan& pl, int nlines, FastDiv dl, int lstride, int estride,
                        const Cplx<T>* __restrict__ tw, T s) {
    for (int st = 0; st < pl.nst; st++) {
        switch (pl.radix[st]) {
            case 8: dif_pass<T, 8>(buf, pl, st, nlines, dl, lstride, estride, tw, s); break;
            case 6: dif_pass<T, 6>(buf, pl, st, nlines, dl, lstride, estride, tw, s); break;
            case 4: dif_pass<T, 4>(buf, pl, st, nlines, dl, lstride, estride, tw, s); break;
            case 2: dif_pass<T, 2>(buf, pl, st, nlines, dl, lstride, estride, tw, s); break;
            case 3: dif_pass<T, 3>(buf, pl, st, nlines, dl, lstride, estride, tw, s); break;
            default: dif_pass<T, 5>(buf, pl, st, nlines, dl, lstride, estride, tw, s); break;
        }
        __syncthreads();
    }
}

// Position of frequency k after lds_fft: its digits in the radices of the
// stages, least significant first, weight each by that stage's span.
__device__ __forceinline__ int dif_pos(const FFTPlan& pl, int k) {
    int p = 0;
    for (int st = 0; st < pl.nst; st++) {
        const int q = pl.drad[st].div(k);
        p += (k - q * pl.radix[st]) * (int)pl.dspan[st].d;
        k = q;
    }
    return p;
}

constexpr int kFftThreadsC = 256;  // threads of the transform blocks
// ---------------------------------------------------------------------------
// The same transform with the length known at compile time (the lengths of the
// benchmark grids): the radix sequence make_plan chooses, as constants, so
// that the butterfly numbering, the LDS addresses, the twiddle indices and the
// digit reversal of the output are constant divisions and offsets instead of
// run-time FastDiv chains. The arithmetic of every butterfly is that of
// dif_pass.
// ---------------------------------------------------------------------------
struct CPlan {
    int nst = 0;
    int radix[kMaxStages] = {}, span[kMaxStages] = {}, tws[kMaxStages] = {};
};
constexpr CPlan cplan(int n) {
    CPlan p;
    int r = n;
    while (r % 8 == 0) {
        p.radix[p.nst++] = 8;
        r /= 8;
    }
    if (r % 4 == 0) {
        p.radix[p.nst++] = 4;
        r /= 4;
    }
    if (r % 6 == 0) {
        p.radix[p.nst++] = 6;
        r /= 6;
    }
    for (int f = 2; f <= 5; f++) {
        if (f == 4) continue;
        while (r % f == 0) {
            p.radix[p.nst++] = f;
            r /= f;
        }
    }
    int L = n;
    for (int st = 0; st < p.nst; st++) {
        p.tws[st] = n / L;
        p.span[st] = L / p.radix[st];
        L /= p.radix[st];
    }
    return p;
}

template <typename T, int N, int ST>
__device__ __forceinline__ void dif_pass_c(Cplx<T>* buf, int nlines, const Cplx<T>* __restrict__ tw, T s) {
    constexpr CPlan P = cplan(N);
    constexpr int R = P.radix[ST], S = P.span[ST], L = S * R, TWS = P.tws[ST], NR = N / R;
    const int total = NR * nlines;
    for (int t = threadIdx.x; t < total; t += kFftThreadsC) {
        const int l = t / NR, b = t - l * NR;
        const int blk = b / S, j = b - blk * S;
        const int e0 = l * N + blk * L + j;
        Cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = buf[ldsi(e0 + r * S)];
        bfly<T, R>(v, s);
        if (S > 1 && j != 0) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                Cplx<T> w = tw[ldsi(r * j * TWS)];
                w.y *= -s;  // forward table; conjugate for the inverse
                v[r] = cmul(v[r], w);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) buf[ldsi(e0 + r * S)] = v[r];
    }
}

template <typename T, int N, int ST = 0>
__device__ __forceinline__ void lds_fft_c(Cplx<T>* buf, int nlines, const Cplx<T>* __restrict__ tw, T s) {
    constexpr CPlan P = cplan(N);
    if constexpr (ST < P.nst) {
        dif_pass_c<T, N, ST>(buf, nlines, tw, s);
        __syncthreads();
        lds_fft_c<T, N, ST + 1>(buf, nlines, tw, s);
    }
}

template <int N>
__device__ __forceinline__ int dif_pos_c(int k) {
    constexpr CPlan P = cplan(N);
    int p = 0;
#pragma unroll
    for (int st = 0; st < P.nst; st++) {
        const int q = k / P.radix[st];
        p += (k - q * P.radix[st]) * P.span[st];
        k = q;
    }
    return p;
}

// NC > 0: the compile-time length NC; NC = 0: the run-time plan.
template <typename T, int NC>
__device__ __forceinline__ void fft_lines(Cplx<T>* buf, const FFTPlan& pl, int nlines, const Cplx<T>* __restrict__ tw,
                                          T s) {
    if constexpr (NC > 0)
        lds_fft_c<T, NC>(buf, nlines, tw, s);
    else
        lds_fft(buf, pl, nlines, FastDiv(), pl.n, 1, tw, s);
}
template <int NC>
__device__ __forceinline__ int fft_pos(const FFTPlan& pl, int k) {
    if constexpr (NC > 0)
        return dif_pos_c<NC>(k);
    else
        return dif_pos(pl, k);
}
template <int NC>
__device__ __forceinline__ int div_n(const FFTPlan& pl, int t) {
    if constexpr (NC > 0)
        return t / NC;
    else
        return pl.dn.div(t);
}

// Line lengths with a compile-time transform (the benchmark grids' NX and NY).
#define TC_FFT_LENGTHS(X)          \
    case 64: X(64); break;         \
    case 96: X(96); break;         \
    case 128: X(128); break;       \
    case 160: X(160); break;       \
    case 192: X(192); break;       \
    case 384: X(384); break;       \
    case 512: X(512); break;

// Lines per transform block: 12 KiB of LDS for the lines (fwd_div adds as much
// for its u rows); 12 and 48 KiB measured slower at 192x160 and 384x384.
template <typename T>
constexpr int lines_per_block(int n) {
    const int per_line = 2 * n * (int)sizeof(Cplx<T>);
    const int l = (24 * 1024) / per_line;
    return l < 1 ? 1 : (l > 16 ? 16 : l);
}

constexpr int kFftThreads = kFftThreadsC;
constexpr int kTwRegs = 4;  // twiddle loads per thread issued with the data loads (N <= 1024 in one round)

// Staging into LDS: every global load of a thread's share is issued before any
// of the values is used (unconditional, clamped addresses; a conditional load
// per element would cost a full memory round trip per element), and the
// twiddle table goes to LDS with them, so the Stockham passes read no global
// memory.
template <typename T>
struct TwStage {
    Cplx<T> r[kTwRegs];
    __device__ __forceinline__ void load(const Cplx<T>* __restrict__ tw, int N) {
#pragma unroll
        for (int u = 0; u < kTwRegs; u++) r[u] = tw[min((int)threadIdx.x + u * kFftThreads, N - 1)];
    }
    __device__ __forceinline__ void store(Cplx<T>* twl, const Cplx<T>* __restrict__ tw, int N) const {
        // unconditional (surplus lanes store the same last entry), so that the
        // loads are not sunk into a branch
#pragma unroll
        for (int u = 0; u < kTwRegs; u++) twl[ldsi(min((int)threadIdx.x + u * kFftThreads, N - 1))] = r[u];
        for (int i = threadIdx.x + kTwRegs * kFftThreads; i < N; i += kFftThreads) twl[ldsi(i)] = tw[i];
    }
};

// The half spectra of the block's 2*lpb rows (row = iy + NY*iz; line l holds
// rows 2l and 2l+1 as real and imaginary parts) to spec[(iz*NXH + kx)*NY + iy]:
// consecutive threads take consecutive rows of one kx, so each store
// instruction writes runs of 2*lpb consecutive values.
template <typename T, int NC>
__device__ __forceinline__ void store_half_spectra(const Cplx<T>* res, const FFTPlan& pl, int ny, FastDiv dny,
                                                   FastDiv d2l, int q0, int nrows, Cplx<T>* __restrict__ spec) {
    const int N = NC > 0 ? NC : pl.n, nl = NC > 0 ? lines_per_block<T>(NC) : (int)d2l.d / 2, nxh = N / 2 + 1;
    // one thread per (kx, line): the line's two rows are neighbours in y (NY is
    // even), so their two spectral values are one 2-complex store
    for (int t = threadIdx.x; t < nxh * nl; t += kFftThreads) {
        const int k = NC > 0 ? t / nl : d2l.div(2 * t), l = t - k * nl;
        const int row = 2 * (q0 + l);
        if (row >= nrows) continue;
        const Cplx<T> z = res[ldsi(l * N + fft_pos<NC>(pl, k))];
        const Cplx<T> zc = res[ldsi(l * N + fft_pos<NC>(pl, k == 0 ? 0 : N - k))];
        // row 2l: A = (Z + conj Zc)/2; row 2l+1: B = -i (Z - conj Zc)/2
        const Cplx<T> a = {T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y)};
        const Cplx<T> b = {T(0.5) * (z.y + zc.y), T(-0.5) * (z.x - zc.x)};
        const int iz = dny.div(row), iy = row - iz * ny;
        // streaming stores: the spectrum is not read again by this kernel, so it
        // should not displace the rows whose y and z taps other blocks still read
        T* sp = reinterpret_cast<T*>(spec + ((size_t)iz * nxh + k) * ny + iy);
        if (row + 1 < nrows) {
            using V4 = __attribute__((ext_vector_type(4))) T;
            __builtin_nontemporal_store(V4{a.x, a.y, b.x, b.y}, reinterpret_cast<V4*>(sp));
        } else {
            __builtin_nontemporal_store(a.x, sp);
            __builtin_nontemporal_store(a.y, sp + 1);
        }
    }
}

// Forward x transform of the divergence of v: rows 2q, 2q+1 (row = iy + NY*iz)
// form the real and imaginary parts of line q; the divergence is evaluated
// while the lines are staged into LDS.
// One real3 cell (12 or 24 bytes, aligned to one real): a single wider load.
template <typename T>
struct Cell3 {
    T x, y, z;
};

// Block 0 of the divergence transform: the flow-rate shift (CfrShiftArgs). Thread
// t adds partials t, t + 256, ... in that order, their loads issued in batches of
// 32 before any is added (a loop that waits for each load in turn is a chain of
// memory round trips); then the waves' sums by shuffles, and those in wave order.
__device__ __forceinline__ void cfr_shift_block(const CfrShiftArgs& c) {
    __shared__ double sh[kFftThreads / 64];
    constexpr int B = 32;
    double acc = 0.0;
    for (int b0 = 0; b0 < c.nb; b0 += B * kFftThreads) {
        double q[B];
#pragma unroll
        for (int i = 0; i < B; i++) q[i] = c.partial[min(b0 + (int)threadIdx.x + i * kFftThreads, c.nb - 1)];
#pragma unroll
        for (int i = 0; i < B; i++)
            if (b0 + (int)threadIdx.x + i * kFftThreads < c.nb) acc += q[i];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < kFftThreads / 64; w++) r += sh[w];
        c.shift[0] = (c.ub_vol - r * c.scale) / c.vol_u;
    }
}

// Elements per batch of the divergence's tap loads. Four waves per SIMD (the
// block's LDS) leave 128 VGPRs: the second-order collocated taps (6 loads per
// element) keep 6 elements' loads in flight there; the staggered ones 2
// (measured: 6 at fourth order is slower).
constexpr int div_batch_width(int scheme) { return scheme == TC_SCHEME_AB3 ? 6 : 2; }

template <typename T, int SCHEME, int IT, int NC>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_fwd_div(FFTPlan pl, const Cplx<T>* __restrict__ tw, Mesh<T> m,
                                                                const T* __restrict__ v, Cplx<T>* __restrict__ spec,
                                                                int lpb_, FastDiv dny, FastDiv d2l, CfrShiftArgs cfr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if (cfr.partial != nullptr && blockIdx.x == 0) cfr_shift_block(cfr);  // block-uniform
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = NC > 0 ? NC : pl.n, nrows = m.ny * m.nz;
    const int lpb = NC > 0 ? lines_per_block<T>(NC) : lpb_;
    Cplx<T>* b1 = b0 + lds_elems(lpb * N);
    Cplx<T>* twl = b1 + lds_elems(lpb * N + 2 * lpb);
    // XCD-swizzled: the y and z taps of the divergence read rows of neighbouring
    // blocks, which then run on the same XCD
    const int q0 = (m.xcd_swz ? xcd_swizzle(blockIdx.x, gridDim.x) : (int)blockIdx.x) * lpb;
    TwStage<T> tws;
    tws.load(tw, N);
    const int tot = lpb * N;  // elements (line, x); at most IT * U * kFftThreads
    // Rows 2l, 2l+1 of line l are neighbours in y within one plane (NY is even).
    // (1) each element's two own cells (u, v, w of rows 2l and 2l+1 at x) in one
    // load per cell: u goes to LDS (ub[row * N + x]) for the x taps of the
    // divergence, v and w stay in registers; (2) the v of the rows y-1, y+2,
    // y+3 around the pair (the pair shares its own v as each other's y taps) and
    // the w of the z taps, all loaded before any is used.
    T* ub = reinterpret_cast<T*>(b1);
    constexpr int U = div_batch_width(SCHEME);  // elements per batch of the divergence
    const int ny = m.ny, nz = m.nz;
    T va[IT][U], wa[IT][U], vb[IT][U], wb[IT][U];
    {
        Cell3<T> ca[IT][U], cb[IT][U];
#pragma unroll
        for (int it = 0; it < IT; it++)
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = min((int)threadIdx.x + (it * U + u) * kFftThreads, tot - 1);
                const int l = div_n<NC>(pl, t), n = t - l * N;
                const int ra = min(2 * (q0 + l), nrows - 2);  // rows ra, ra + 1 of one plane
                ca[it][u] = *reinterpret_cast<const Cell3<T>*>(v + 3 * ((size_t)ra * N + n));
                cb[it][u] = *reinterpret_cast<const Cell3<T>*>(v + 3 * ((size_t)(ra + 1) * N + n));
            }
#pragma unroll
        for (int it = 0; it < IT; it++)
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = (int)threadIdx.x + (it * U + u) * kFftThreads;
                if (t < tot) {
                    const int l = div_n<NC>(pl, t), n = t - l * N;
                    ub[2 * l * N + n] = ca[it][u].x;
                    ub[(2 * l + 1) * N + n] = cb[it][u].x;
                }
                va[it][u] = ca[it][u].y;
                wa[it][u] = ca[it][u].z;
                vb[it][u] = cb[it][u].y;
                wb[it][u] = cb[it][u].z;
            }
    }
    tws.store(twl, tw, N);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t0 = (int)threadIdx.x + it * U * kFftThreads;
        if (t0 >= tot) break;
        T vm1[U] = {}, vp2[U], vp3[U] = {};
        T wa1[U], wb1[U], wa2[U] = {}, wb2[U] = {}, wam[U] = {}, wbm[U] = {};
        T dzv[U] = {}, dzwv[U] = {}, chv[U] = {};
        int iz[U], l_[U], n_[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = min(t0 + u * kFftThreads, tot - 1);
            const int l = div_n<NC>(pl, t), n = t - l * N;
            const int ra = min(2 * (q0 + l), nrows - 2);  // rows ra, ra + 1 of one plane
            const int z = dny.div(ra), iy = ra - z * ny;
            iz[u] = z;
            l_[u] = l;
            n_[u] = n;
            auto at = [&](int zz, int yy) { return 3 * (((size_t)zz * ny + yy) * N + n); };
            auto wrapy = [&](int yy) { return yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy); };
            vp2[u] = v[at(z, wrapy(iy + 2)) + 1];
            if (SCHEME != TC_SCHEME_HW) vm1[u] = v[at(z, wrapy(iy - 1)) + 1];
            if (SCHEME == TC_SCHEME_VV) vp3[u] = v[at(z, wrapy(iy + 3)) + 1];
            const int zp = min(z + 1, nz - 1), zm = max(z - 1, 0), zpp = min(z + 2, nz - 1);
            wa1[u] = v[at(zp, iy) + 2];
            wb1[u] = v[at(zp, iy + 1) + 2];
            if (SCHEME != TC_SCHEME_HW) {
                wam[u] = v[at(zm, iy) + 2];
                wbm[u] = v[at(zm, iy + 1) + 2];
            }
            if (SCHEME == TC_SCHEME_VV) {
                wa2[u] = v[at(zpp, iy) + 2];
                wb2[u] = v[at(zpp, iy + 1) + 2];
            }
            if (SCHEME == TC_SCHEME_AB3) {
                chv[u] = m.cellh[z];
            } else {
                dzv[u] = m.DZ[z];
                if (SCHEME == TC_SCHEME_VV) dzwv[u] = m.DZW[z];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int l = l_[u], n = n_[u];
            const int xm = n == 0 ? N - 1 : n - 1, xp = n + 1 >= N ? n + 1 - N : n + 1, xpp = n + 2 >= N ? n + 2 - N : n + 2;
            const T* ua = ub + 2 * l * N;
            const T* ubb = ua + N;
            TapVals<T> qa, qb;
#pragma unroll
            for (int k = 0; k < 12; k++) qa.v[k] = qb.v[k] = T(0);
            qa.v[0] = ua[xp];
            qb.v[0] = ubb[xp];
            qa.v[1] = ua[n];
            qb.v[1] = ubb[n];
            qa.v[3] = ua[xm];
            qb.v[3] = ubb[xm];
            if (SCHEME == TC_SCHEME_VV) {
                qa.v[2] = ua[xpp];
                qb.v[2] = ubb[xpp];
            }
            // v taps y+1, y, y+2, y-1
            qa.v[4] = vb[it][u];
            qa.v[5] = va[it][u];
            qa.v[6] = vp2[u];
            qa.v[7] = vm1[u];
            qb.v[4] = vp2[u];
            qb.v[5] = vb[it][u];
            qb.v[6] = vp3[u];
            qb.v[7] = va[it][u];
            // w taps z, z+1, z+2, z-1
            qa.v[8] = wa[it][u];
            qa.v[9] = wa1[u];
            qa.v[10] = wa2[u];
            qa.v[11] = wam[u];
            qb.v[8] = wb[it][u];
            qb.v[9] = wb1[u];
            qb.v[10] = wb2[u];
            qb.v[11] = wbm[u];
            const T zx = div_cell<T, SCHEME>(m, qa, iz[u], dzv[u], dzwv[u], chv[u]);
            const T zy = div_cell<T, SCHEME>(m, qb, iz[u], dzv[u], dzwv[u], chv[u]);
            const int t = t0 + u * kFftThreads;
            if (t < tot) {
                const int ra = 2 * (q0 + l);
                b0[ldsi(t)] = {ra < nrows ? zx : T(0), ra + 1 < nrows ? zy : T(0)};
            }
        }
    }
    __syncthreads();
    fft_lines<T, NC>(b0, pl, lpb, twl, T(-1));
    store_half_spectra<T, NC>(b0, pl, m.ny, dny, d2l, q0, nrows, spec);
}

// Inverse x transform: half spectra of rows 2q, 2q+1 -> real rows (times scale).
// A thread reads the two rows' values of one kx together (consecutive ky), so
// each load instruction covers runs of lpb value pairs, and writes the
// complex line A + iB at k and, by the conjugate symmetry, at N - k.
template <typename T, int NC>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_inv(FFTPlan pl, const Cplx<T>* __restrict__ tw,
                                                            const Cplx<T>* __restrict__ spec, T* __restrict__ out,
                                                            int nrows, int ny, FastDiv dny, int lpb_, FastDiv dlpb,
                                                            T scale) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = NC > 0 ? NC : pl.n, nxh = N / 2 + 1, half = N / 2;
    const int lpb = NC > 0 ? lines_per_block<T>(NC) : lpb_;
    Cplx<T>* twl = b0 + lds_elems(lpb * N);
    // XCD-contiguous block order: a block reads 2*lpb values of each kx, part of a
    // cache line whose rest the next block reads, which then runs on the same XCD
    const int q0 = xcd_swizzle(blockIdx.x, gridDim.x) * lpb;
    struct alignas(2 * sizeof(Cplx<T>)) Pair { Cplx<T> a, b; };
    constexpr int U = 4;
    TwStage<T> tws;
    tws.load(tw, N);
    const int tst = nxh * lpb;  // items (kk, l), l fastest
    for (int t0 = threadIdx.x; t0 < tst; t0 += U * kFftThreads) {
        Pair q[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = min(t0 + u * kFftThreads, tst - 1);
            const int kk = (NC > 0 ? t / lpb : dlpb.div(t)), l = t - kk * lpb;
            const int ra = min(2 * (q0 + l), nrows - 2);  // rows ra, ra + 1: one plane, consecutive ky
            const int iz = dny.div(ra), iy = ra - iz * ny;
            q[u] = *reinterpret_cast<const Pair*>(spec + ((size_t)iz * nxh + kk) * ny + iy);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * kFftThreads;
            if (t >= tst) continue;
            const int kk = (NC > 0 ? t / lpb : dlpb.div(t)), l = t - kk * lpb;
            const int ra = 2 * (q0 + l);
            Cplx<T> a = ra < nrows ? q[u].a : Cplx<T>{T(0), T(0)};
            Cplx<T> b = ra + 1 < nrows ? q[u].b : Cplx<T>{T(0), T(0)};
            if (kk == 0 || kk == half) {  // real-to-complex symmetry: these bins are real
                a.y = T(0);
                b.y = T(0);
            }
            b0[ldsi(l * N + kk)] = {a.x - b.y, a.y + b.x};  // A + iB
            if (kk != 0 && kk != half) b0[ldsi(l * N + N - kk)] = {a.x + b.y, b.x - a.y};  // conj A + i conj B
        }
    }
    tws.store(twl, tw, N);
    __syncthreads();
    fft_lines<T, NC>(b0, pl, lpb, twl, T(1));
    for (int t = threadIdx.x; t < lpb * N; t += kFftThreads) {
        const int l = div_n<NC>(pl, t), n = t - l * N;
        const int ra = 2 * (q0 + l), rb = ra + 1;
        const Cplx<T> z = b0[ldsi(l * N + fft_pos<NC>(pl, n))];
        if (ra < nrows) out[(size_t)ra * N + n] = z.x * scale;
        if (rb < nrows) out[(size_t)rb * N + n] = z.y * scale;
    }
}

// y transform of lpb contiguous lines (line L = iz*NXH + kx, NY values), in place.
// With psum != nullptr (inverse transform of the pressure solve) the (0,0) mode
// of each plane first gets the normalisation X00[iz] -= psum * C[iz] (FP64).
template <typename T, int NC>
__global__ __launch_bounds__(kFftThreads) void k_fft_y(FFTPlan pl, const Cplx<T>* __restrict__ tw, Cplx<T>* spec,
                                                        int nlines, FastDiv dnxh, int lpb_, T s,
                                                        const double* __restrict__ psum, const double* __restrict__ C) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = NC > 0 ? NC : pl.n;
    const int lpb = NC > 0 ? lines_per_block<T>(NC) : lpb_;
    Cplx<T>* twl = b0 + lds_elems(lpb * N);
    const int L0 = blockIdx.x * lpb;
    const int nl = min(lpb, nlines - L0);
    Cplx<T>* base = spec + (size_t)L0 * N;
    constexpr int U = 6;  // all of a 4 x 384 block's loads in one batch (2 and 4 measured slower)
    TwStage<T> tws;
    tws.load(tw, N);
    const int tot = lpb * N, tin = nl * N;
    for (int t0 = threadIdx.x; t0 < tot; t0 += U * kFftThreads) {
        Cplx<T> v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = base[min(t0 + u * kFftThreads, tin - 1)];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * kFftThreads;
            if (t >= tot) continue;
            Cplx<T> w = t < tin ? v[u] : Cplx<T>{T(0), T(0)};
            if (psum) {
                const int l = div_n<NC>(pl, t), n = t - l * N, L = L0 + l;
                const int iz = dnxh.div(L);
                if (n == 0 && L == iz * (int)dnxh.d && l < nl) w.x = T((double)w.x - psum[0] * C[iz]);
            }
            b0[ldsi(t)] = w;
        }
    }
    tws.store(twl, tw, N);
    __syncthreads();
    fft_lines<T, NC>(b0, pl, lpb, twl, s);
    for (int t = threadIdx.x; t < tin; t += kFftThreads) {
        const int l = div_n<NC>(pl, t), n = t - l * N;
        base[t] = b0[ldsi(l * N + fft_pos<NC>(pl, n))];
    }
}

// ---------------------------------------------------------------------------
// Wall-normal banded solves (ThomasForPressure: staggered/func_global.h:71-189,
// collocated_AB3/func_global.h:8-61) with factors precomputed on the host.
// One thread per (mode, real/imaginary part): the matrix is real.
// ---------------------------------------------------------------------------
// One row of packed factors (see ZFacDev).
template <typename CT, int W>
__device__ __forceinline__ void load_fwd_row(const CT* __restrict__ F, CT& m1, CT& m2, CT& m3, CT& d) {
    if (W == 3) {
        struct alignas(4 * sizeof(CT)) V4 { CT a, b, c, d; };
        const V4 q = *reinterpret_cast<const V4*>(F);
        m1 = q.a;
        m2 = q.b;
        m3 = q.c;
        d = q.d;
    } else {
        struct alignas(2 * sizeof(CT)) V2 { CT a, b; };
        const V2 q = *reinterpret_cast<const V2*>(F);
        m1 = q.a;
        d = q.b;
    }
}
template <typename CT, int W>
__device__ __forceinline__ void load_bwd_row(const CT* __restrict__ U, CT& u1, CT& u2, CT& u3) {
    if (W == 3) {
        struct V3 { CT a, b, c; };
        const V3 q = *reinterpret_cast<const V3*>(U);
        u1 = q.a;
        u2 = q.b;
        u3 = q.c;
    } else {
        u1 = *U;
    }
}

// The forward elimination of one mode formed in-lane (see ZFactors::R): the
// row's constants are mode-independent (the same address in every lane, so
// scalar loads), the mode enters through its wavenumber terms, and the lane
// carries the normalised upper band of its last W rows. Same steps as the
// host's elimination (build_zfactors); pivots inverted by v_rcp (FP32 lanes).
template <typename CT, int W>
struct ZInl {
    CT k0 = 0, k1 = 0, k2 = 0, k3 = 0;  // W = 1: MM (k0); W = 3: sx, s3x, sy, s3y
    CT b1[W] = {}, b2[W] = {}, b3[W] = {};  // upper band of rows i-1 (W = 1: U), i-2, i-3
    __device__ __forceinline__ void init(const CT* __restrict__ KX, const CT* __restrict__ KY, int kx, int aky) {
        if constexpr (W == 3) {
            k0 = KX[2 * kx];
            k1 = KX[2 * kx + 1];
            k2 = KY[2 * aky];
            k3 = KY[2 * aky + 1];
        } else {
            k0 = KX[kx] + KY[aky];
        }
    }
    // factors of the row whose constants are at R: multipliers and inverse pivot
    __device__ __forceinline__ void row(const CT* __restrict__ R, CT& m1, CT& m2, CT& m3, CT& d) {
        static_assert(sizeof(CT) == 4, "in-lane factors for FP32 lanes");
        if constexpr (W == 3) {
            CT l2 = R[1], l3 = R[2];
            const CT cx = R[7] * k0 - R[8] * k1, cy = R[7] * k2 - R[8] * k3;
            CT dg = R[3] - R[9] * cx * cx;
            dg = dg - R[10] * cy * cy;
            CT u1 = R[4], u2 = R[5];
            const CT q3 = -R[0];
            l2 = l2 + q3 * b3[0];
            l3 = l3 + q3 * b3[1];
            dg = dg + q3 * b3[2];
            const CT q2 = -l2;
            l3 = l3 + q2 * b2[0];
            dg = dg + q2 * b2[1];
            u1 = u1 + q2 * b2[2];
            const CT q1 = -l3;
            dg = dg + q1 * b1[0];
            u1 = u1 + q1 * b1[1];
            u2 = u2 + q1 * b1[2];
            d = __builtin_amdgcn_rcpf(dg);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                b3[k] = b2[k];
                b2[k] = b1[k];
            }
            b1[0] = u1 * d;
            b1[1] = u2 * d;
            b1[2] = R[6] * d;
            m1 = q1;
            m2 = q2;
            m3 = q3;
        } else {
            d = __builtin_amdgcn_rcpf((R[2] + R[3] * k0) - R[0] * b1[0]);
            m1 = -R[0];
            m2 = m3 = 0;
            b1[0] = R[1] * d;
        }
    }
};

// One lane's solve: forward sweep x -> y, backward sweep y -> x, in CT
// arithmetic. Rows move in full batches of B with no per-row conditions (a
// condition per row splits the code into blocks at whose joins the compiler
// waits for every load in flight, one memory round trip per row). The loads of
// a batch are issued NB-1 batches ahead of the recurrence and before the stores
// of the batch being finished, since the vector-memory counter counts loads and
// stores together. The nz % B remaining rows follow one at a time.
// INL (W = 1): the forward factors are formed in-lane by zi from the row
// constants R (F is not read).
template <typename T, typename CT, int W, int B, int NB, bool INL = false>
__device__ __forceinline__ void lane_solve(T* __restrict__ x, size_t xs, const CT* __restrict__ F, size_t fs,
                                           const CT* __restrict__ U, size_t us, CT* __restrict__ y, size_t ys, int nz,
                                           bool n00, const ZNorm& norm, const CT* __restrict__ R = nullptr,
                                           ZInl<CT, W>* zi = nullptr) {
    const int nfull = nz / B;
    CT y1 = 0, y2 = 0, y3 = 0;
    auto fwd_row = [&](CT b, CT m1, CT m2, CT m3, CT d, int i) {
        CT v = b;
        if (W == 3) v = v + m3 * y3;
        if (W == 3) v = v + m2 * y2;
        v = v + m1 * y1;
        v = v * d;
        y[i * ys] = v;
        y3 = y2;
        y2 = y1;
        y1 = v;
    };
    // forward substitution: y[i] = (((x[i] + M3 y[i-3]) + M2 y[i-2]) + M1 y[i-1]) * D
    if (nfull > 0) {
        CT b[NB][B], m1[NB][B], m2[NB][B], m3[NB][B], d[NB][B];
        auto load = [&](int kb, CT(&bb)[B], CT(&q1)[B], CT(&q2)[B], CT(&q3)[B], CT(&qd)[B]) {
            const int i0 = (kb < nfull ? kb : nfull - 1) * B;
#pragma unroll
            for (int k = 0; k < B; k++) {
                bb[k] = (CT)x[(i0 + k) * xs];
                if constexpr (!INL) load_fwd_row<CT, W>(F + (i0 + k) * fs, q1[k], q2[k], q3[k], qd[k]);
            }
        };
#pragma unroll
        for (int sb = 0; sb < NB - 1; sb++) load(sb, b[sb], m1[sb], m2[sb], m3[sb], d[sb]);
        for (int kb0 = 0; kb0 < nfull; kb0 += NB) {
#pragma unroll
            for (int sb = 0; sb < NB; sb++) {
                const int kb = kb0 + sb;
                if (kb >= nfull) break;
                const int nx = (sb + NB - 1) % NB;
                load(kb + NB - 1, b[nx], m1[nx], m2[nx], m3[nx], d[nx]);
#pragma unroll
                for (int k = 0; k < B; k++) {
                    if constexpr (INL) {
                        CT q1, q2, q3, qd;
                        zi->row(R + (size_t)(kb * B + k) * zrow_w(W), q1, q2, q3, qd);
                        fwd_row(b[sb][k], q1, q2, q3, qd, kb * B + k);
                    } else {
                        fwd_row(b[sb][k], m1[sb][k], m2[sb][k], m3[sb][k], d[sb][k], kb * B + k);
                    }
                }
            }
        }
    }
    for (int i = nfull * B; i < nz; i++) {
        CT m1, m2 = 0, m3 = 0, d;
        if constexpr (INL) zi->row(R + (size_t)i * zrow_w(W), m1, m2, m3, d);
        else load_fwd_row<CT, W>(F + i * fs, m1, m2, m3, d);
        fwd_row((CT)x[i * xs], m1, m2, m3, d, i);
    }
    // backward substitution: x[i] = y[i] - U1 x[i+1] - U2 x[i+2] - U3 x[i+3]
    double psum = 0.0;
    CT z1 = 0, z2 = 0, z3 = 0;
    auto bwd_row = [&](CT b, CT u1, CT u2, CT u3, double wt, int i) {
        CT v = b;
        v = v - u1 * z1;
        if (W == 3) v = v - u2 * z2;
        if (W == 3) v = v - u3 * z3;
        x[i * xs] = T(v);
        psum += wt * (double)v;
        z3 = z2;
        z2 = z1;
        z1 = v;
    };
    if (nfull > 0) {
        CT b[NB][B], u1[NB][B], u2[NB][B], u3[NB][B];
        double wt[NB][B];  // normalisation weights (used by the (0,0) lane; a broadcast load)
        auto load = [&](int kb, CT(&bb)[B], CT(&q1)[B], CT(&q2)[B], CT(&q3)[B], double(&qw)[B]) {
            const int i0 = nz - 1 - (kb < nfull ? kb : nfull - 1) * B;
#pragma unroll
            for (int k = 0; k < B; k++) {
                bb[k] = y[(i0 - k) * ys];
                load_bwd_row<CT, W>(U + (i0 - k) * us, q1[k], q2[k], q3[k]);
                qw[k] = norm.Wt[i0 - k];
            }
        };
#pragma unroll
        for (int sb = 0; sb < NB - 1; sb++) load(sb, b[sb], u1[sb], u2[sb], u3[sb], wt[sb]);
        for (int kb0 = 0; kb0 < nfull; kb0 += NB) {
#pragma unroll
            for (int sb = 0; sb < NB; sb++) {
                const int kb = kb0 + sb;
                if (kb >= nfull) break;
                const int nx = (sb + NB - 1) % NB;
                load(kb + NB - 1, b[nx], u1[nx], u2[nx], u3[nx], wt[nx]);
#pragma unroll
                for (int k = 0; k < B; k++)
                    bwd_row(b[sb][k], u1[sb][k], u2[sb][k], u3[sb][k], wt[sb][k], nz - 1 - (kb * B + k));
            }
        }
    }
    for (int i = nz - 1 - nfull * B; i >= 0; i--) {
        CT u1, u2 = 0, u3 = 0;
        load_bwd_row<CT, W>(U + i * us, u1, u2, u3);
        bwd_row(y[i * ys], u1, u2, u3, norm.Wt[i], i);
    }
    if (n00) norm.s00[0] = psum;  // applied by the inverse y transform (NormalizePressure)
}

// lane_solve with nz = NZ known at compile time and the forward intermediate
// kept on chip: the row loops are unrolled, so each intermediate row has a
// place fixed at compile time, and the backward sweep reads it there instead of
// a second pass over the spectrum. The last NR rows (the first the backward
// sweep needs) are registers, the NL rows before them LDS (yl: lane-fastest,
// 64 reals a row), the first NZ - NR - NL rows go over the spectrum. Same
// arithmetic as lane_solve.
template <typename T, int W, int B, int NB, int NZ, bool INL = false, int NR = NZ, int NL = 0>
__device__ __forceinline__ void lane_solve_reg(T* __restrict__ x, size_t xs, const T* __restrict__ F, size_t fs,
                                               const T* __restrict__ U, size_t us, bool n00, const ZNorm& norm,
                                               const T* __restrict__ R = nullptr, ZInl<T, W>* zi = nullptr,
                                               T* __restrict__ yl = nullptr) {
    static_assert(NZ % B == 0 && NR % B == 0 && NL % B == 0 && NR + NL <= NZ, "rows in whole batches");
    constexpr int NH = NZ - NR - NL;  // rows over the spectrum
    constexpr int NBAT = NZ / B;
    T yr[NR];
    T y1 = 0, y2 = 0, y3 = 0;
    {
        T b[NB][B], m1[NB][B], m2[NB][B], m3[NB][B], d[NB][B];
        auto load = [&](int kb, T(&bb)[B], T(&q1)[B], T(&q2)[B], T(&q3)[B], T(&qd)[B]) {
            const int i0 = (kb < NBAT ? kb : NBAT - 1) * B;
#pragma unroll
            for (int k = 0; k < B; k++) {
                bb[k] = x[(i0 + k) * xs];
                if constexpr (!INL) load_fwd_row<T, W>(F + (i0 + k) * fs, q1[k], q2[k], q3[k], qd[k]);
            }
        };
#pragma unroll
        for (int sb = 0; sb < NB - 1; sb++) load(sb, b[sb], m1[sb], m2[sb], m3[sb], d[sb]);
#pragma unroll
        for (int kb = 0; kb < NBAT; kb++) {
            const int sb = kb % NB, nx = (kb + NB - 1) % NB;
            load(kb + NB - 1, b[nx], m1[nx], m2[nx], m3[nx], d[nx]);
#pragma unroll
            for (int k = 0; k < B; k++) {
                T f1, f2 = 0, f3 = 0, fd;
                if constexpr (INL) {
                    zi->row(R + (size_t)(kb * B + k) * zrow_w(W), f1, f2, f3, fd);
                } else {
                    f1 = m1[sb][k];
                    if (W == 3) f2 = m2[sb][k];
                    if (W == 3) f3 = m3[sb][k];
                    fd = d[sb][k];
                }
                T v = b[sb][k];
                if (W == 3) v = v + f3 * y3;
                if (W == 3) v = v + f2 * y2;
                v = v + f1 * y1;
                v = v * fd;
                const int i = kb * B + k;
                if (i < NH) x[i * xs] = v;
                else if (i < NH + NL) yl[(i - NH) * 64] = v;
                else yr[i - NH - NL >= 0 ? i - NH - NL : 0] = v;
                y3 = y2;
                y2 = y1;
                y1 = v;
            }
            __builtin_amdgcn_sched_barrier(0);  // keep each batch's loads in its place
        }
    }
    double psum = 0.0;
    T z1 = 0, z2 = 0, z3 = 0;
    // the stores' addresses from a pointer the compiler cannot equate with x, so
    // that it recomputes them instead of keeping the forward sweep's 2 NZ
    // address registers alive across the sweeps
    T* xw = x;
    asm volatile("" : "+v"(xw));
    // the intermediate rows over the spectrum come back through another such
    // pointer, after every store of the forward sweep (memory clobber)
    const T* xr = x;
    asm volatile("" : "+v"(xr)::"memory");
    {
        T u1[NB][B], u2[NB][B], u3[NB][B], yy[NB][B];
        auto load = [&](int kb, T(&q1)[B], T(&q2)[B], T(&q3)[B], T(&qy)[B]) {
            const int i0 = NZ - 1 - (kb < NBAT ? kb : NBAT - 1) * B;
#pragma unroll
            for (int k = 0; k < B; k++) {
                load_bwd_row<T, W>(U + (i0 - k) * us, q1[k], q2[k], q3[k]);
                if (i0 - k < NH) qy[k] = xr[(i0 - k) * xs];
                else if (i0 - k < NH + NL) qy[k] = yl[(i0 - k - NH) * 64];
            }
        };
#pragma unroll
        for (int sb = 0; sb < NB - 1; sb++) load(sb, u1[sb], u2[sb], u3[sb], yy[sb]);
#pragma unroll
        for (int kb = 0; kb < NBAT; kb++) {
            const int sb = kb % NB, nx = (kb + NB - 1) % NB;
            load(kb + NB - 1, u1[nx], u2[nx], u3[nx], yy[nx]);
#pragma unroll
            for (int k = 0; k < B; k++) {
                const int i = NZ - 1 - (kb * B + k);
                T v;
                if (i < NH + NL) v = yy[sb][k];
                else v = yr[i - NH - NL >= 0 ? i - NH - NL : 0];
                v = v - u1[sb][k] * z1;
                if (W == 3) v = v - u2[sb][k] * z2;
                if (W == 3) v = v - u3[sb][k] * z3;
                xw[i * xs] = v;
                if (n00) psum += norm.Wt[i] * (double)v;
                z3 = z2;
                z2 = z1;
                z1 = v;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (n00) norm.s00[0] = psum;  // applied by the inverse y transform (NormalizePressure)
}

// Rows of the nz = 192 z-solve's forward intermediate kept in registers and in
// LDS (the rest go over the spectrum). With the factors formed in-lane: 128 in
// registers (165-166 VGPRs, three waves per SIMD, so all 2 317 waves of a
// 384 x 384 x 192 solve are resident at once), 48 in LDS (12 KiB per one-wave
// block, 144 KiB per CU), 16 over the spectrum (AB3 384 x 384 x 192 80 -> 69 us;
// seven-band 105 -> 90 us).
template <int W>
constexpr int zreg_rows(int nz) { return nz == 192 ? 128 : nz; }
template <int W>
constexpr int zlds_rows(int nz) { return nz == 192 ? 48 : 0; }

// Few independent systems (2 per mode) for the latency of every row: one wave
// per block so the waves spread over all CUs, and pipelined row batches. The
// FP32-bulk engine's FP64 low modes run in blocks of their own, after the bulk
// blocks, so that no FP32 wave carries a divergent FP64 lane; their forward
// intermediate stays in double (ylow: nz * nlow * 2). NBL = 0: the bulk
// forward intermediate is written over the spectrum; NBL > 0: it stays in LDS
// (lane-fastest, nz x 64 reals), with NBL batches of rows in the pipeline;
// NZR > 0 (nz == NZR): it stays in registers (lane_solve_reg).
template <typename T, int W, int NBL, int NZR = 0, bool INL = false>
__global__ __launch_bounds__(64) void k_zsolve(Cplx<T>* spec, ZFacDev<T> f, ZFacDev<double> flow, const int* lowidx,
                                                const int* lowmode, int nlow, double* ylow, ZNorm norm, int nbulk,
                                                const T* __restrict__ zrow) {
    constexpr int FW = zfac_fw(W), UW = zfac_uw(W);
    const int nm = f.nmodes, nz = f.nz;
    const size_t xs = 2 * (size_t)nm;  // stride between planes, in reals
    if ((int)blockIdx.x >= nbulk) {    // FP64 low-mode blocks
        const int t = (blockIdx.x - nbulk) * blockDim.x + threadIdx.x;
        const int li = t >> 1, part = t & 1;
        if (li >= nlow) return;
        const int mode = lowmode[li];
        lane_solve<T, double, W, 8, 2>(reinterpret_cast<T*>(spec) + 2 * mode + part, xs, flow.F + (size_t)li * FW,
                                        (size_t)nlow * FW, flow.U + (size_t)li * UW, (size_t)nlow * UW,
                                        ylow + 2 * li + part, 2 * (size_t)nlow, nz, mode == 0 && part == 0, norm);
        return;
    }
    // XCD-contiguous block order: the blocks of one kx line run on one XCD, so
    // modes ky and ny - ky (one factor column) read the factors through one L2
    const int t = xcd_swizzle(blockIdx.x, nbulk) * blockDim.x + threadIdx.x;
    const int mode = t >> 1, part = t & 1;
    if (mode >= nm) return;
    if (lowidx && lowidx[mode] >= 0) return;  // solved in FP64 by the low-mode blocks
    T* x = reinterpret_cast<T*>(spec) + 2 * mode + part;
    const int col = zfac_col(mode, f.ny);  // the factors of ky and ny - ky are one column
    const size_t nc = f.ncols;
    ZInl<T, W> zi;
    if constexpr (INL) {
        const int kx = mode / f.ny, ky = mode - kx * f.ny;
        zi.init(f.KX, f.KY, kx, min(ky, f.ny - ky));
    }
    if constexpr (NZR > 0) {  // the intermediate on chip (nz == NZR)
        // one batch of rows ahead: 8 rows (W = 1), 4 rows (W = 3: 5 reals per row)
        extern __shared__ __align__(16) unsigned char smem_raw[];
        T* yl = reinterpret_cast<T*>(smem_raw) + threadIdx.x;
        lane_solve_reg<T, W, W == 3 ? 4 : 8, 2, NZR, INL, zreg_rows<W>(NZR), zlds_rows<W>(NZR)>(
            x, xs, f.F + (size_t)col * FW, nc * FW, f.U + (size_t)col * UW, nc * UW,
            mode == 0 && part == 0 && norm.s00 != nullptr, norm, zrow, &zi, yl);
    } else if constexpr (NBL == 0) {
        // batches of 8 rows in flight: NB - 1 (FP64 lanes hold 5 reals per row of W = 3)
        constexpr int NB = 3;
        lane_solve<T, T, W, 8, NB, INL>(x, xs, f.F + (size_t)col * FW, nc * FW, f.U + (size_t)col * UW, nc * UW, x,
                                        xs, nz, mode == 0 && part == 0 && norm.s00 != nullptr, norm, zrow, &zi);
    } else {
        extern __shared__ __align__(16) unsigned char smem_raw[];
        T* yl = reinterpret_cast<T*>(smem_raw) + threadIdx.x;
        lane_solve<T, T, W, 8, NBL, INL>(x, xs, f.F + (size_t)col * FW, nc * FW, f.U + (size_t)col * UW, nc * UW, yl,
                                         64, nz, mode == 0 && part == 0 && norm.s00 != nullptr, norm, zrow, &zi);
    }
}

// Forward x transform of one component (cb < 0) or of the product of two
// components of a real3 field (training samples of the learned closure).
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_fwd_q(FFTPlan pl, const Cplx<T>* __restrict__ tw,
                                                              const T* __restrict__ v, int ca, int cb, int nrows,
                                                              Cplx<T>* __restrict__ spec, int lpb, int ny, FastDiv dny,
                                                              FastDiv d2l) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cplx<T>* b0 = reinterpret_cast<Cplx<T>*>(smem_raw);
    const int N = pl.n;
    Cplx<T>* twl = b0 + lds_elems(lpb * N);
    const int q0 = blockIdx.x * lpb;
    auto q = [&](int row, int n) {
        const T* c = v + 3 * ((size_t)row * N + n);
        const T x = c[ca], y = c[cb < 0 ? ca : cb];
        return cb < 0 ? x : x * y;
    };
    constexpr int U = 4;
    TwStage<T> tws;
    tws.load(tw, N);
    const int tot = lpb * N;
    for (int t0 = threadIdx.x; t0 < tot; t0 += U * kFftThreads) {
        T zx[U], zy[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = min(t0 + u * kFftThreads, tot - 1);
            const int l = pl.dn.div(t), n = t - l * N;
            zx[u] = q(min(2 * (q0 + l), nrows - 1), n);
            zy[u] = q(min(2 * (q0 + l) + 1, nrows - 1), n);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * kFftThreads;
            if (t < tot) {
                const int ra = 2 * (q0 + pl.dn.div(t));
                b0[ldsi(t)] = {ra < nrows ? zx[u] : T(0), ra + 1 < nrows ? zy[u] : T(0)};
            }
        }
    }
    tws.store(twl, tw, N);
    __syncthreads();
    lds_fft(b0, pl, lpb, FastDiv(), N, 1, twl, T(-1));
    store_half_spectra<T, 0>(b0, pl, ny, dny, d2l, q0, nrows, spec);
}

// Half spectrum of the coarse grid holding the modes a sharp cut-off keeps:
// |kx| < nxc/2 and |ky| < nyc/2 (the coarse Nyquist modes are zero).
template <typename T>
__global__ void k_spec_crop(const Cplx<T>* __restrict__ fine, int nxh, int ny, Cplx<T>* __restrict__ coarse, int nxhc,
                            int nyc, int nz) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nxhc * nyc * nz) return;
    const int kyc = (int)(t % nyc);
    const long long rest = t / nyc;
    const int kx = (int)(rest % nxhc), iz = (int)(rest / nxhc);
    const int nxc = 2 * (nxhc - 1), hy = nyc / 2;
    Cplx<T> v = {T(0), T(0)};
    if (kx < nxc / 2) {
        const int ky = kyc < hy ? kyc : (kyc > hy ? ny - (nyc - kyc) : -1);
        if (ky >= 0) v = fine[((size_t)iz * nxh + kx) * ny + ky];
    }
    coarse[t] = v;
}


}  // namespace

// Batches of 2 elements per thread in the divergence staging: up to 4 (lines of
// up to 2048 points).
constexpr int fwd_div_batches(int lpb, int n, int u) { return (lpb * n + u * kFftThreads - 1) / (u * kFftThreads); }

template <typename T, int SCHEME>
hipError_t launch_fft_x_fwd_div_s(hipStream_t s, int it, dim3 grid, size_t sm, const FFTPlan& pl, const Cplx<T>* tw,
                                  const Mesh<T>& m, const T* v, Cplx<T>* spec, int lpb, FastDiv dny, FastDiv d2l,
                                  const CfrShiftArgs& cfr) {
#define LFD(IT, NC_) hipLaunchKernelGGL((k_fft_x_fwd_div<T, SCHEME, IT, NC_>), grid, dim3(kFftThreads), sm, s, pl, tw, m, \
                                        v, spec, lpb, dny, d2l, cfr)
#define LFDC(NC_) LFD(fwd_div_batches(lines_per_block<T>(NC_), NC_, div_batch_width(SCHEME)), NC_)
    switch (pl.n) {
        TC_FFT_LENGTHS(LFDC)
        default:
            switch (it) {
                case 1: LFD(1, 0); break;
                case 2: LFD(2, 0); break;
                case 3: LFD(3, 0); break;
                case 4: LFD(4, 0); break;
                default: return hipErrorInvalidValue;
            }
    }
#undef LFDC
#undef LFD
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_x_fwd_div(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const Mesh<T>& m, int scheme,
                                const T* v, Cplx<T>* spec, const CfrShiftArgs& cfr) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (m.ny * m.nz + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = (lds_elems((size_t)lpb * pl.n) + lds_elems((size_t)lpb * pl.n + 2 * lpb) + lds_elems(pl.n)) *
                      sizeof(Cplx<T>);
    const FastDiv dny((unsigned)m.ny), d2l((unsigned)(2 * lpb));
    const int it = fwd_div_batches(lpb, pl.n, div_batch_width(scheme));
    if (scheme == TC_SCHEME_HW)
        return launch_fft_x_fwd_div_s<T, TC_SCHEME_HW>(s, it, dim3(grid), sm, pl, tw, m, v, spec, lpb, dny, d2l, cfr);
    if (scheme == TC_SCHEME_VV)
        return launch_fft_x_fwd_div_s<T, TC_SCHEME_VV>(s, it, dim3(grid), sm, pl, tw, m, v, spec, lpb, dny, d2l, cfr);
    return launch_fft_x_fwd_div_s<T, TC_SCHEME_AB3>(s, it, dim3(grid), sm, pl, tw, m, v, spec, lpb, dny, d2l, cfr);
}

template <typename T>
hipError_t launch_fft_x_inv(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const Cplx<T>* spec, T* out, int ny,
                            int nrows, T scale) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (nrows + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = (lds_elems((size_t)lpb * pl.n) + lds_elems(pl.n)) * sizeof(Cplx<T>);
#define LFI(NC_) hipLaunchKernelGGL((k_fft_x_inv<T, NC_>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, spec, out, \
                                   nrows, ny, FastDiv((unsigned)ny), lpb, FastDiv((unsigned)lpb), scale)
    switch (pl.n) {
        TC_FFT_LENGTHS(LFI)
        default: LFI(0);
    }
#undef LFI
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_y(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, Cplx<T>* spec, int nxh, int nz, int inverse,
                        const double* psum, const double* C) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = nxh * nz;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = (lds_elems((size_t)lpb * pl.n) + lds_elems(pl.n)) * sizeof(Cplx<T>);
#define LFY(NC_) hipLaunchKernelGGL((k_fft_y<T, NC_>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, spec, nlines, \
                                   FastDiv((unsigned)nxh), lpb, T(inverse ? 1 : -1), psum, C)
    switch (pl.n) {
        TC_FFT_LENGTHS(LFY)
        default: LFY(0);
    }
#undef LFY
    return hipGetLastError();
}

// Blocks of the LDS-intermediate z-solve resident at once on the device.
template <typename T, int W>
int zsolve_slots(size_t dyn_lds) {
    static int slots = 0;
    static size_t slots_lds = 0;
    if (slots == 0 || slots_lds != dyn_lds) {
        int per_cu = 0, dev = 0, ncu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_zsolve<T, W, 3>, 64, dyn_lds) != hipSuccess ||
            per_cu < 1)
            per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            ncu < 1)
            ncu = 0;
        slots = per_cu * ncu;
        slots_lds = dyn_lds;
    }
    return slots;
}

template <typename T>
hipError_t launch_zsolve(hipStream_t s, Cplx<T>* spec, const ZFacDev<T>& f, const ZFacDev<double>& flow,
                         const int* lowidx, const int* lowmode, int nlow, double* ylow, ZNorm norm) {
    const int nbulk = (2 * f.nmodes + 63) / 64;
    const int nlowb = lowidx ? (2 * nlow + 63) / 64 : 0;
    const dim3 grid(nbulk + nlowb);
    // The forward intermediate goes to LDS (64 lanes x nz reals per block)
    // when every block of the launch is resident at once; otherwise the LDS
    // would cut the resident waves below what the row latency needs, and the
    // intermediate goes over the spectrum instead.
    const size_t sm = 64 * (size_t)f.nz * sizeof(T);
    const bool lds = sm <= 64 * 1024 && (int)grid.x <= (f.W == 3 ? zsolve_slots<T, 3>(sm) : zsolve_slots<T, 1>(sm));
    // FP32-bulk engines form the bulk modes' forward factors in-lane (the (0,0)
    // mode, which the host pins, is one of the FP64 low modes): the
    // tridiagonal systems on every path (AB3 384 x 384 x 192: 90 -> 78 us), the
    // seven-band ones with the on-chip intermediate at nz = 192 (three waves
    // per SIMD: 105 -> 90 us; with two waves per SIMD the elimination's chain
    // per row made the solve slower, 105 -> 115-137 us).
    const bool inl = sizeof(T) == 4 && f.R != nullptr && lowidx != nullptr && (f.W == 1 || f.nz == 192);
#define LZS(W_, N_)                                                                                                     \
    do {                                                                                                                \
        if constexpr (sizeof(T) == 4 && W_ == 1) {                                                                      \
            if (inl) {                                                                                                  \
                hipLaunchKernelGGL((k_zsolve<T, W_, N_, 0, true>), grid, dim3(64), N_ ? sm : 0, s, spec, f, flow,      \
                                   lowidx, lowmode, nlow, ylow, norm, nbulk, f.R);                                      \
                break;                                                                                                  \
            }                                                                                                           \
        }                                                                                                               \
        hipLaunchKernelGGL((k_zsolve<T, W_, N_>), grid, dim3(64), N_ ? sm : 0, s, spec, f, flow, lowidx, lowmode,      \
                           nlow, ylow, norm, nbulk, (const T*)nullptr);                                                 \
    } while (0)
    // FP32 with nz = 192 (the 384 x 384 x 192 grids) when not every block is
    // resident: the factors in-lane and the intermediate on chip (zreg_rows,
    // zlds_rows), which removes the spectrum's second pass (AB3 118 -> 88 us,
    // fourth order 140 -> 105 us at 384 x 384 x 192 with all rows in
    // registers and streamed factors)
    const bool reg = sizeof(T) == 4 && inl && f.nz == 192 && !lds;
#define LZR(W_)                                                                                                         \
    do {                                                                                                                \
        if constexpr (sizeof(T) == 4)                                                                                   \
            hipLaunchKernelGGL((k_zsolve<T, W_, 0, 192, true>), grid, dim3(64), zlds_rows<W_>(192) * 64 * sizeof(T), s, \
                               spec, f, flow, lowidx, lowmode, nlow, ylow, norm, nbulk, f.R);                           \
        else                                                                                                            \
            LZS(W_, 0);                                                                                                 \
    } while (0)
    if (f.W == 3) {
        if (lds) LZS(3, 3);
        else if (reg) LZR(3);
        else LZS(3, 0);
    } else {
        if (lds) LZS(1, 3);
        else if (reg) LZR(1);
        else LZS(1, 0);
    }
#undef LZR
#undef LZS
    return hipGetLastError();
}


template <typename T>
hipError_t launch_fft_x_fwd_q(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const T* v, int ca, int cb, int ny,
                              int nrows, Cplx<T>* spec) {
    const int lpb = lines_per_block<T>(pl.n);
    const int nlines = (nrows + 1) / 2;
    const int grid = (nlines + lpb - 1) / lpb;
    const size_t sm = (lds_elems((size_t)lpb * pl.n) + lds_elems(pl.n)) * sizeof(Cplx<T>);
    hipLaunchKernelGGL((k_fft_x_fwd_q<T>), dim3(grid), dim3(kFftThreads), sm, s, pl, tw, v, ca, cb, nrows, spec, lpb,
                       ny, FastDiv((unsigned)ny), FastDiv((unsigned)(2 * lpb)));
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
                                                const T*, Cplx<T>*, const CfrShiftArgs&);                            \
    template hipError_t launch_fft_x_inv<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const Cplx<T>*, T*, int, int, T); \
    template hipError_t launch_fft_y<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, Cplx<T>*, int, int, int,       \
                                        const double*, const double*);                                              \
    template hipError_t launch_zsolve<T>(hipStream_t, Cplx<T>*, const ZFacDev<T>&, const ZFacDev<double>&,         \
                                         const int*, const int*, int, double*, ZNorm);                                   \
    template hipError_t launch_fft_x_fwd_q<T>(hipStream_t, const FFTPlan&, const Cplx<T>*, const T*, int, int, int, int, \
                                              Cplx<T>*);                                                             \
    template hipError_t launch_spec_crop<T>(hipStream_t, const Cplx<T>*, int, int, Cplx<T>*, int, int, int);
INST(double)
INST(float)
#undef INST

}  // namespace tc
