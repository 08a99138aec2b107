/*
 * tc_oracle.c -- CPU restatement of bahvalo/TurbChannel's time stepper (FP64, OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP engine. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function cites the reference file:line it restates. The reference cannot
 * be compiled in this image (its CPU path includes cufft.h and fftw3.h, and its
 * driver calls the CUDA runtime), so this file is the oracle; it is pinned by the
 * reference's own self-tests (staggered/selftest.cpp) restated in tests/, by the
 * DFT (numpy), by dense solves on tiny grids and by closed forms.
 *
 * Deliberate generalisations, none of which changes results at the reference's
 * sizes: runtime NX/NY/NZ (periodic wrap by comparison instead of bitmask; the
 * two agree at powers of two, see or_index_selfcheck); the FFT is a plain
 * mixed-radix DFT in place of FFTW (same unnormalised C2C transforms);
 * GCZ[0], which the reference computes from DZ[-1] and DZ[-2] and never uses
 * (staggered/main.cpp:127), is set to 0; nu_t starts zeroed.
 */
#include "tc_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define C1_36 (1.0 / 36.0)
#define C1_27 (1.0 / 27.0)
#define C1_18 (1.0 / 18.0)
#define C1_9 (1.0 / 9.0)
#define C1_6 (1.0 / 6.0)
#define C1_3 (1.0 / 3.0)
#define PI_D 3.14159265358979323846

static char g_err[256] = "";
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char* or_last_error(void) { return g_err; }

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* FFT: recursive mixed-radix decimation in time, table-driven twiddles.      */
/* Replaces FFTW's unnormalised C2C (staggered/func_global.h:250-275).        */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n;
    cplx* tw[2]; /* tw[0]: exp(-2 pi i k/n) (forward), tw[1]: exp(+2 pi i k/n) */
} fft_plan;

static void plan_init(fft_plan* p, int n) {
    p->n = n;
    for (int d = 0; d < 2; d++) {
        p->tw[d] = (cplx*)malloc(sizeof(cplx) * (size_t)n);
        double sgn = d == 0 ? -1.0 : 1.0;
        for (int k = 0; k < n; k++) {
            double a = 2.0 * PI_D * (double)k / (double)n;
            p->tw[d][k] = cos(a) + I * (sgn * sin(a));
        }
    }
}
static void plan_free(fft_plan* p) {
    free(p->tw[0]);
    free(p->tw[1]);
}

static int smallest_factor(int n) {
    for (int f = 2; f * f <= n; f++)
        if (n % f == 0) return f;
    return n;
}

/* out[k] = sum_j in[j*stride] * W_n^{jk}, W_n = tw[N/n] ; N = table length */
static void dft_rec(const cplx* in, int stride, cplx* out, int n, const cplx* tw, int N) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor(n), m = n / p, step = N / n;
    for (int r = 0; r < p; r++) dft_rec(in + (size_t)r * stride, stride * p, out + (size_t)r * m, m, tw, N);
    cplx t[64];
    cplx* tt = p <= 64 ? t : (cplx*)malloc(sizeof(cplx) * (size_t)p);
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) tt[r] = out[(size_t)r * m + k] * tw[(size_t)((long)r * k % n) * step];
        for (int q = 0; q < p; q++) {
            cplx acc = 0;
            for (int r = 0; r < p; r++) acc += tt[r] * tw[(size_t)((long)r * q % p) * m * step];
            out[(size_t)k + (size_t)m * q] = acc;
        }
    }
    if (tt != t) free(tt);
}

static void fft_line(const fft_plan* p, int dir, const cplx* in, int istride, cplx* out, int ostride, cplx* work) {
    /* work: 2n */
    cplx* a = work;
    cplx* b = work + p->n;
    for (int i = 0; i < p->n; i++) a[i] = in[(size_t)i * istride];
    dft_rec(a, 1, b, p->n, p->tw[dir], p->n);
    for (int i = 0; i < p->n; i++) out[(size_t)i * ostride] = b[i];
}

void or_fft(int n, const double* in_re_im, double* out_re_im, int sign) {
    fft_plan p;
    plan_init(&p, n);
    cplx* w = (cplx*)malloc(sizeof(cplx) * 2 * (size_t)n);
    fft_line(&p, sign < 0 ? 0 : 1, (const cplx*)in_re_im, 1, (cplx*)out_re_im, 1, w);
    free(w);
    plan_free(&p);
}

/* ------------------------------------------------------------------------- */
/* Solver state (staggered/data_global_inc.h:102-167, config.h:47-73;         */
/* collocated_AB3/config.h:39-56, data_global.h:125-226)                      */
/* ------------------------------------------------------------------------- */
struct or_solver {
    or_params P;
    int NX, NY, NZ, NXY, N3D;
    double inv_hx, inv_hy, area;
    double *ZZ, *DZ, *DZW, *GC1, *GC2, *GCZ, *MAT_A, MAT_D[6], *KERG, KERG_norm2;
    double *vel[4], *flux[4], *pres[4], *divv, *nu_t, *tau;
    cplx* spec;
    double* cosphi[2];
    fft_plan px, py;
    double tmr[OR_NTIMERS]; /* seconds per channel (or_timers) */
};

/* Generalised periodic neighbours (staggered/data_global_inc.h:4-9). */
static inline int XR(const or_solver* s, int i) { return (i % s->NX == s->NX - 1) ? i - (s->NX - 1) : i + 1; }
static inline int XL(const or_solver* s, int i) { return (i % s->NX == 0) ? i + (s->NX - 1) : i - 1; }
static inline int YR(const or_solver* s, int i) {
    return ((i / s->NX) % s->NY == s->NY - 1) ? i - s->NX * (s->NY - 1) : i + s->NX;
}
static inline int YL(const or_solver* s, int i) {
    return ((i / s->NX) % s->NY == 0) ? i + s->NX * (s->NY - 1) : i - s->NX;
}
static inline int ZR(const or_solver* s, int i) { return i + s->NXY; }
static inline int ZL(const or_solver* s, int i) { return i - s->NXY; }

/* Power-of-2 bitmask forms, restated from staggered/data_global_inc.h:4-7. */
static inline int XRb(int NX, int i) { int j = i + 1; if ((j & (NX - 1)) == 0) j -= NX; return j; }
static inline int XLb(int NX, int i) { int j = i; if ((j & (NX - 1)) == 0) j += NX; return j - 1; }
static inline int YRb(int NX, int NY, int i) { int j = i + NX; if ((j & (NX * (NY - 1))) == 0) j -= NX * NY; return j; }
static inline int YLb(int NX, int NY, int i) { int j = i; if ((j & (NX * (NY - 1))) == 0) j += NX * NY; return j - NX; }

int or_index_selfcheck(int nx, int ny, int nz) {
    or_solver s;
    memset(&s, 0, sizeof s);
    s.NX = nx; s.NY = ny; s.NZ = nz; s.NXY = nx * ny; s.N3D = nx * ny * nz;
    int bad = 0;
    for (int i = 0; i < s.N3D; i++) {
        bad += XR(&s, i) != XRb(nx, i);
        bad += XL(&s, i) != XLb(nx, i);
        bad += YR(&s, i) != YRb(nx, ny, i);
        bad += YL(&s, i) != YLb(nx, ny, i);
    }
    return bad;
}

#define SQR(x) ((x) * (x))
/* collocated node layout: the AB3 and SRK solvers */
#define IS_COLL(sc) ((sc) == OR_SCHEME_AB3 || (sc) == OR_SCHEME_SRK)

/* ------------------------------------------------------------------------- */
/* Mesh and matrix set-up                                                     */
/* ------------------------------------------------------------------------- */

/* solve_banded: Gauss elimination without pivoting (staggered/main.cpp:73-109). */
int or_solve_banded(int N, int NB, double* M, double* x, double tiny) {
    if (!(NB & 1)) return 1;
    const int w = NB / 2;
    for (int i = 0; i < N; i++) {
        for (int j = -w; j < 0; j++) {
            if (i + j < 0) continue;
            double mult = -M[i * NB + w + j];
            for (int k = 0; k <= w; k++) {
                if (i + j + k >= N) continue;
                M[i * NB + (w + j) + k] += mult * M[(i + j) * NB + w + k];
            }
            x[i] += mult * x[i + j];
        }
        if (fabs(M[i * NB + w]) < tiny) return 2;
        double mult = 1.0 / M[i * NB + w];
        for (int k = 0; k <= w; k++) {
            if (i + k >= N) continue;
            M[i * NB + w + k] *= mult;
        }
        x[i] *= mult;
    }
    for (int i = N - 1; i >= 0; i--)
        for (int j = 1; j <= w; j++) {
            if (i + j >= N) continue;
            x[i] -= M[i * NB + w + j] * x[i + j];
        }
    return 0;
}

/* FillMatrix: 7-band D*G of the VV scheme for the (0,0) mode (staggered/main.cpp:12-48). */
static int fill_matrix(const or_solver* s, double* M) {
    const double* inv_hz = s->GCZ;
    const int w = 3, NB = 2 * w + 1, NZ = s->NZ;
    const double grad_coef[4] = {1. / 27., -1., 1., -1. / 27.};
    for (int j = 0; j < NZ; j++) {
        for (int pos = 0; pos < NB; pos++) M[j * NB + pos] = 0.;
        for (int ll = -1; ll <= 2; ll++) {
            double coeff1 = grad_coef[ll + 1];
            int l = j + ll;
            if (l == 0 || l == NZ) continue;
            if (l < 0) l = -l;
            if (l > NZ) l = 2 * NZ - l;
            for (int mm = -2; mm <= 1; mm++) {
                double coeff2 = grad_coef[mm + 2];
                int m = l + mm;
                if (m < 0) { m = -m - 1; coeff2 *= -1.; }
                if (m >= NZ) { m = 2 * NZ - m - 1; coeff2 *= -1.; }
                int pos = m - j + w;
                if (pos < 0 || pos >= NB || l <= 0 || l >= NZ) return 1;
                M[j * NB + pos] += coeff1 * coeff2 * inv_hz[l];
            }
        }
    }
    return 0;
}

void or_fill_matrix(const or_solver* s, double* M) { fill_matrix(s, M); }

/* InitDerivativeMeshData (staggered/main.cpp:112-156). */
static int init_staggered_mesh(or_solver* s) {
    const int NZ = s->NZ;
    for (int i = 0; i < NZ; i++) s->DZ[i] = s->ZZ[i + 1] - s->ZZ[i];
    if (s->P.scheme == OR_SCHEME_HW) {
        s->MAT_A[0] = 0.;
        for (int j = 1; j <= NZ - 1; j++) s->MAT_A[j] = 2. / (s->DZ[j - 1] + s->DZ[j]);
        return 0;
    }
    double* DZ = s->DZ;
    for (int iz = 0; iz < NZ; iz++) {
        s->DZW[iz] = (DZ[iz] + DZ[iz == 0 ? iz : iz - 1] + DZ[iz == NZ - 1 ? iz : iz + 1]) / 3.;
        s->GC1[iz] = DZ[iz] / (DZ[iz] - s->DZW[iz] / 9.);
        s->GC2[iz] = -(1. / 27.) * s->DZW[iz] / (DZ[iz] - s->DZW[iz] / 9.);
        if (iz == 0) {
            s->GCZ[iz] = 0.; /* reference reads DZ[-1], DZ[-2]; value never used */
        } else {
            s->GCZ[iz] = 1 / (0.5 * (DZ[iz] + DZ[iz - 1]) -
                              (1. / 27.) * (DZ[iz] + DZ[iz - 1] +
                                            0.5 * (DZ[iz == NZ - 1 ? iz : iz + 1] + DZ[iz == 1 ? 0 : iz - 2])));
        }
    }
    double* MAT = (double*)calloc((size_t)7 * NZ, sizeof(double));
    if (fill_matrix(s, MAT)) { free(MAT); set_err("Internal error"); return 1; }
    for (int iz = 0; iz < NZ; iz++)
        for (int i = 0; i < 3; i++) s->MAT_A[iz * 3 + i] = (iz + i - 3 >= 0) ? MAT[iz * 7 + i] : 0.;
    s->MAT_D[0] = MAT[3];
    s->MAT_D[1] = MAT[7 + 3];
    s->MAT_D[2] = MAT[7 * 2 + 3];
    s->MAT_D[3] = MAT[(NZ - 3) * 7 + 3];
    s->MAT_D[4] = MAT[(NZ - 2) * 7 + 3];
    s->MAT_D[5] = MAT[(NZ - 1) * 7 + 3];

    /* kernel of the gradient operator */
    const double eps = 1. / 27.;
    MAT[1] = -1;
    MAT[2] = eps;
    for (int iz = 1; iz < NZ - 1; iz++) {
        MAT[iz * 3 + 0] = MAT[iz * 3 + 2] = eps;
        MAT[iz * 3 + 1] = -1 + eps;
    }
    MAT[(NZ - 1) * 3] = eps;
    MAT[(NZ - 1) * 3 + 1] = -1;
    for (int iz = 0; iz < NZ; iz++) s->KERG[iz] = 1;
    int errcode = or_solve_banded(NZ, 3, MAT, s->KERG, 1e-10);
    free(MAT);
    if (errcode) { set_err("Internal error in solving system for K"); return 1; }
    s->KERG_norm2 = 0;
    for (int iz = 0; iz < NZ; iz++) s->KERG_norm2 += s->KERG[iz] * s->KERG[iz] * DZ[iz];
    return 0;
}

/* collocated mesh data (collocated_AB3/main.cpp:238-241, 295-300) */
static void init_ab3_mesh(or_solver* s) {
    const int NZ = s->NZ;
    for (int i = 0; i < NZ - 1; i++) s->DZ[i] = s->ZZ[i + 1] - s->ZZ[i];
    double* A = s->MAT_A;
    A[0] = 0.;
    for (int j = 1; j < NZ; j++) A[j] = s->P.hx * s->P.hy / s->DZ[j - 1];
    if (s->P.fd_mode) { A[1] *= 0.5; A[NZ - 1] *= 0.5; }
}

or_solver* or_create(const or_params* p, const double* zz) {
    if (!p || !zz || p->nx < 4 || p->ny < 4 || p->nz < 6 || p->scheme < 0 || p->scheme > 3 ||
        (p->sgs_model && !IS_COLL(p->scheme)) ||
        (p->scheme == OR_SCHEME_SRK && (p->visc_type || p->turb_visc || p->sgs_model || p->cfr_ubulk > 0))) {
        set_err("bad parameters");
        return NULL;
    }
    or_solver* s = (or_solver*)calloc(1, sizeof(or_solver));
    s->P = *p;
    s->NX = p->nx; s->NY = p->ny; s->NZ = p->nz;
    s->NXY = s->NX * s->NY;
    s->N3D = s->NXY * s->NZ;
    s->inv_hx = 1. / p->hx;
    s->inv_hy = 1. / p->hy;
    int NZ = s->NZ, nzz = IS_COLL(p->scheme) ? NZ : NZ + 1;
    s->ZZ = (double*)calloc((size_t)nzz, sizeof(double));
    memcpy(s->ZZ, zz, sizeof(double) * (size_t)nzz);
    s->DZ = (double*)calloc((size_t)NZ, sizeof(double));
    s->DZW = (double*)calloc((size_t)NZ, sizeof(double));
    s->GC1 = (double*)calloc((size_t)NZ, sizeof(double));
    s->GC2 = (double*)calloc((size_t)NZ, sizeof(double));
    s->GCZ = (double*)calloc((size_t)NZ, sizeof(double));
    s->MAT_A = (double*)calloc((size_t)3 * NZ, sizeof(double));
    s->KERG = (double*)calloc((size_t)NZ, sizeof(double));
    if (IS_COLL(p->scheme)) {
        init_ab3_mesh(s);
        s->area = p->hx * s->NX * p->hy * s->NY * (s->ZZ[NZ - 1] - s->ZZ[0]);
    } else {
        if (init_staggered_mesh(s)) { or_destroy(s); return NULL; }
        s->area = p->hx * s->NX * p->hy * s->NY * (s->ZZ[NZ] - s->ZZ[0]);
    }
    size_t n3 = (size_t)s->N3D;
    for (int i = 0; i < 4; i++) {
        s->vel[i] = (double*)calloc(3 * n3, sizeof(double));
        s->flux[i] = (double*)calloc(3 * n3, sizeof(double));
        s->pres[i] = (double*)calloc(n3, sizeof(double));
    }
    s->divv = (double*)calloc(n3, sizeof(double));
    s->nu_t = p->turb_visc ? (double*)calloc(n3, sizeof(double)) : NULL;
    s->tau = p->sgs_model ? (double*)calloc(6 * n3, sizeof(double)) : NULL;
    s->spec = (cplx*)calloc(n3, sizeof(cplx));
    s->cosphi[0] = (double*)calloc((size_t)s->NX, sizeof(double));
    s->cosphi[1] = (double*)calloc((size_t)s->NY, sizeof(double));
    /* InitFourier, CPU branch (staggered/func_global.h:250-258) */
    for (int i = 0; i < s->NX; i++) s->cosphi[0][i] = cos(2. * PI_D * i / s->NX);
    for (int i = 0; i < s->NY; i++) s->cosphi[1][i] = cos(2. * PI_D * i / s->NY);
    plan_init(&s->px, s->NX);
    plan_init(&s->py, s->NY);
    return s;
}

void or_destroy(or_solver* s) {
    if (!s) return;
    free(s->ZZ); free(s->DZ); free(s->DZW); free(s->GC1); free(s->GC2); free(s->GCZ);
    free(s->MAT_A); free(s->KERG);
    for (int i = 0; i < 4; i++) { free(s->vel[i]); free(s->flux[i]); free(s->pres[i]); }
    free(s->divv); free(s->nu_t); free(s->tau); free(s->spec); free(s->cosphi[0]); free(s->cosphi[1]);
    plan_free(&s->px);
    plan_free(&s->py);
    free(s);
}

int or_mesh_array(const or_solver* s, const char* name, double* out, int cap) {
    const double* src = NULL;
    int n = 0, NZ = s->NZ, ab3 = IS_COLL(s->P.scheme);
    if (!strcmp(name, "ZZ")) { src = s->ZZ; n = ab3 ? NZ : NZ + 1; }
    else if (!strcmp(name, "DZ")) { src = s->DZ; n = ab3 ? NZ - 1 : NZ; }
    else if (!strcmp(name, "MAT_A")) { src = s->MAT_A; n = s->P.scheme == OR_SCHEME_VV ? 3 * NZ : NZ; }
    else if (s->P.scheme == OR_SCHEME_VV) {
        if (!strcmp(name, "DZW")) { src = s->DZW; n = NZ; }
        else if (!strcmp(name, "GC1")) { src = s->GC1; n = NZ; }
        else if (!strcmp(name, "GC2")) { src = s->GC2; n = NZ; }
        else if (!strcmp(name, "GCZ")) { src = s->GCZ; n = NZ; }
        else if (!strcmp(name, "MAT_D")) { src = s->MAT_D; n = 6; }
        else if (!strcmp(name, "KERG")) { src = s->KERG; n = NZ; }
    }
    if (!src) return -1;
    if (out) memcpy(out, src, sizeof(double) * (size_t)(n < cap ? n : cap));
    return n;
}

double or_kerg_norm2(const or_solver* s) { return s->KERG_norm2; }

/* ------------------------------------------------------------------------- */
/* Staggered operators                                                        */
/* ------------------------------------------------------------------------- */
#define u_(X) vel[3 * (X) + 0]
#define v_(X) vel[3 * (X) + 1]
#define w_(X) vel[3 * (X) + 2]
#define XR_(i) XR(s, i)
#define XL_(i) XL(s, i)
#define YR_(i) YR(s, i)
#define YL_(i) YL(s, i)
#define ZR_(i) ZR(s, i)
#define ZL_(i) ZL(s, i)
#define XRR_(i) XR_(XR_(i))
#define XLL_(i) XL_(XL_(i))
#define YRR_(i) YR_(YR_(i))
#define YLL_(i) YL_(YL_(i))
#define ZRR_(i) ZR_(ZR_(i))
#define ZLL_(i) ZL_(ZL_(i))
#define XRRR_(i) XR_(XR_(XR_(i)))
#define XLLL_(i) XL_(XL_(XL_(i)))
#define YRRR_(i) YR_(YR_(YR_(i)))
#define YLLL_(i) YL_(YL_(YL_(i)))
#define ZRRR_(i) ZR_(ZR_(ZR_(i)))
#define ZLLL_(i) ZL_(ZL_(ZL_(i)))

/* ConvFlux, Harlow-Welch (staggered/func_global.h:345-370). */
static void conv_flux_hw(const or_solver* s, int in, const double* vel, double* out) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double* DZ = s->DZ;
    const double inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    double f0 = 0, f1 = 0, f2 = 0;
    double inv_hz = 1 / DZ[iz];
    double inv_hz_alt = 2.0 / (DZ[iz] + (iz == 0 ? 0.0 : DZ[iz - 1]));

    f0 -= (SQR(u_(in) + u_(XR_(in))) - SQR(u_(in) + u_(XL_(in)))) * inv_hx;
    f0 -= ((u_(in) + u_(YR_(in))) * (v_(YR_(in)) + v_(YR_(XL_(in)))) - (u_(in) + u_(YL_(in))) * (v_(in) + v_(XL_(in)))) * inv_hy;
    if (iz != NZ - 1) f0 -= (u_(in) + u_(ZR_(in))) * (w_(ZR_(in)) + w_(ZR_(XL_(in)))) * inv_hz;
    if (iz != 0) f0 += (u_(in) + u_(ZL_(in))) * (w_(in) + w_(XL_(in))) * inv_hz;

    f1 -= ((v_(in) + v_(XR_(in))) * (u_(XR_(in)) + u_(XR_(YL_(in)))) - (v_(in) + v_(XL_(in))) * (u_(in) + u_(YL_(in)))) * inv_hx;
    f1 -= (SQR(v_(in) + v_(YR_(in))) - SQR(v_(in) + v_(YL_(in)))) * inv_hy;
    if (iz != NZ - 1) f1 -= (v_(in) + v_(ZR_(in))) * (w_(ZR_(in)) + w_(ZR_(YL_(in)))) * inv_hz;
    if (iz != 0) f1 += (v_(in) + v_(ZL_(in))) * (w_(in) + w_(YL_(in))) * inv_hz;

    if (iz != 0) {
        f2 -= ((w_(in) + w_(XR_(in))) * (u_(XR_(in)) * DZ[iz] + u_(XR_(ZL_(in))) * DZ[iz - 1]) -
               (w_(in) + w_(XL_(in))) * (u_(in) * DZ[iz] + u_(ZL_(in)) * DZ[iz - 1])) * inv_hx * inv_hz_alt;
        f2 -= ((w_(in) + w_(YR_(in))) * (v_(YR_(in)) * DZ[iz] + v_(YR_(ZL_(in))) * DZ[iz - 1]) -
               (w_(in) + w_(YL_(in))) * (v_(in) * DZ[iz] + v_(ZL_(in)) * DZ[iz - 1])) * inv_hy * inv_hz_alt;
        f2 -= SQR(w_(in) + (iz == NZ - 1 ? 0.0 : w_(ZR_(in)))) * (inv_hz_alt);
        f2 += SQR(w_(in) + w_(ZL_(in))) * (inv_hz_alt);
    }
    out[0] = f0 * 0.25;
    out[1] = f1 * 0.25;
    out[2] = f2 * 0.25;
}

/* ConvFlux, Verstappen-Veldman (staggered/func_global.h:374-529). */
static void conv_flux_vv(const or_solver* s, int in, const double* vel, double* out) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double *DZ = s->DZ, *DZW = s->DZW, *GCZ = s->GCZ;
    const double hx = s->P.hx, hy = s->P.hy, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    double f0 = 0, f1 = 0, f2 = 0;
    {
        double ubar_xr = DZ[iz] * hy * (9 * (u_(in) + u_(XR_(in))) - (u_(XL_(in)) + u_(XRR_(in))));
        double ubar_xl = DZ[iz] * hy * (9 * (u_(XL_(in)) + u_(in)) - (u_(XLL_(in)) + u_(XR_(in))));
        double ubarbar_xr = DZW[iz] * hy * (9 * (u_(XR_(in)) + u_(XRR_(in))) - (u_(in) + u_(XRRR_(in))));
        double ubarbar_xl = DZW[iz] * hy * (9 * (u_(XLL_(in)) + u_(XL_(in))) - (u_(XLLL_(in)) + u_(in)));
        f0 += ubar_xr * (u_(in) + u_(XR_(in))) - C1_27 * ubarbar_xr * (u_(in) + u_(XRRR_(in))) -
              ubar_xl * (u_(XL_(in)) + u_(in)) + C1_27 * ubarbar_xl * (u_(XLLL_(in)) + u_(in));
    }
    {
        double vbar_yr = DZ[iz] * hx * (9 * (v_(YR_(in)) + v_(XL_(YR_(in)))) - (v_(XLL_(YR_(in))) + v_(XR_(YR_(in)))));
        double vbar_yl = DZ[iz] * hx * (9 * (v_(in) + v_(XL_(in))) - (v_(XLL_(in)) + v_(XR_(in))));
        double vbarbar_yr = DZW[iz] * hx * (9 * (v_(YRR_(in)) + v_(XL_(YRR_(in)))) - (v_(XLL_(YRR_(in))) + v_(XR_(YRR_(in)))));
        double vbarbar_yl = DZW[iz] * hx * (9 * (v_(YL_(in)) + v_(XL_(YL_(in)))) - (v_(XLL_(YL_(in))) + v_(XR_(YL_(in)))));
        f0 += vbar_yr * (u_(in) + u_(YR_(in))) - C1_27 * vbarbar_yr * (u_(in) + u_(YRRR_(in))) -
              vbar_yl * (u_(YL_(in)) + u_(in)) + C1_27 * vbarbar_yl * (u_(YLLL_(in)) + u_(in));
    }
    {
        double wbar_zr, wbar_zl, wbarbar_zr, wbarbar_zl;
        if (iz == 0) wbar_zl = 0;
        else wbar_zl = hx * hy * (9 * (w_(in) + w_(XL_(in))) - (w_(XLL_(in)) + w_(XR_(in))));
        if (iz == NZ - 1) wbar_zr = 0;
        else wbar_zr = hx * hy * (9 * (w_(ZR_(in)) + w_(XL_(ZR_(in)))) - (w_(XLL_(ZR_(in))) + w_(XR_(ZR_(in)))));
        if (iz == 1) wbarbar_zl = 0;
        else if (iz == 0) wbarbar_zl = wbar_zr;
        else wbarbar_zl = hx * hy * (9 * (w_(ZL_(in)) + w_(XL_(ZL_(in)))) - (w_(XLL_(ZL_(in))) + w_(XR_(ZL_(in)))));
        if (iz == NZ - 2) wbarbar_zr = 0;
        else if (iz == NZ - 1) wbarbar_zr = wbar_zl;
        else wbarbar_zr = hx * hy * (9 * (w_(ZRR_(in)) + w_(XL_(ZRR_(in)))) - (w_(XLL_(ZRR_(in))) + w_(XR_(ZRR_(in)))));
        double u_zr = iz == NZ - 1 ? 0.0 : u_(in) + u_(ZR_(in));
        double u_zrr = u_(in) + (iz == NZ - 3 ? -u_(ZRR_(in)) : iz == NZ - 2 ? 0.0 : iz == NZ - 1 ? u_(ZLL_(in)) : u_(ZRRR_(in)));
        double u_zl = iz == 0 ? 0.0 : u_(ZL_(in)) + u_(in);
        double u_zll = u_(in) + (iz == 2 ? -u_(ZLL_(in)) : iz == 1 ? 0.0 : iz == 0 ? u_(ZRR_(in)) : u_(ZLLL_(in)));
        f0 += wbar_zr * u_zr - C1_27 * wbarbar_zr * u_zrr - wbar_zl * u_zl + C1_27 * wbarbar_zl * u_zll;
    }
    {
        double ubar_xr = DZ[iz] * hy * (9 * (u_(XR_(in)) + u_(YL_(XR_(in)))) - (u_(YLL_(XR_(in))) + u_(YR_(XR_(in)))));
        double ubar_xl = DZ[iz] * hy * (9 * (u_(in) + u_(YL_(in))) - (u_(YLL_(in)) + u_(YR_(in))));
        double ubarbar_xr = DZW[iz] * hy * (9 * (u_(XRR_(in)) + u_(YL_(XRR_(in)))) - (u_(YLL_(XRR_(in))) + u_(YR_(XRR_(in)))));
        double ubarbar_xl = DZW[iz] * hy * (9 * (u_(XL_(in)) + u_(YL_(XL_(in)))) - (u_(YLL_(XL_(in))) + u_(YR_(XL_(in)))));
        f1 += ubar_xr * (v_(in) + v_(XR_(in))) - C1_27 * ubarbar_xr * (v_(in) + v_(XRRR_(in))) -
              ubar_xl * (v_(XL_(in)) + v_(in)) + C1_27 * ubarbar_xl * (v_(XLLL_(in)) + v_(in));
    }
    {
        double vbar_yr = DZ[iz] * hx * (9 * (v_(in) + v_(YR_(in))) - (v_(YL_(in)) + v_(YRR_(in))));
        double vbar_yl = DZ[iz] * hx * (9 * (v_(YL_(in)) + v_(in)) - (v_(YLL_(in)) + v_(YR_(in))));
        double vbarbar_yr = DZW[iz] * hx * (9 * (v_(YR_(in)) + v_(YRR_(in))) - (v_(in) + v_(YRRR_(in))));
        double vbarbar_yl = DZW[iz] * hx * (9 * (v_(YLL_(in)) + v_(YL_(in))) - (v_(YLLL_(in)) + v_(in)));
        f1 += vbar_yr * (v_(in) + v_(YR_(in))) - C1_27 * vbarbar_yr * (v_(in) + v_(YRRR_(in))) -
              vbar_yl * (v_(YL_(in)) + v_(in)) + C1_27 * vbarbar_yl * (v_(YLLL_(in)) + v_(in));
    }
    {
        double wbar_zr, wbar_zl, wbarbar_zr, wbarbar_zl;
        if (iz == 0) wbar_zl = 0;
        else wbar_zl = hx * hy * (9 * (w_(in) + w_(YL_(in))) - (w_(YLL_(in)) + w_(YR_(in))));
        if (iz == NZ - 1) wbar_zr = 0;
        else wbar_zr = hx * hy * (9 * (w_(ZR_(in)) + w_(YL_(ZR_(in)))) - (w_(YLL_(ZR_(in))) + w_(YR_(ZR_(in)))));
        if (iz == 1) wbarbar_zl = 0;
        else if (iz == 0) wbarbar_zl = wbar_zr;
        else wbarbar_zl = hx * hy * (9 * (w_(ZL_(in)) + w_(YL_(ZL_(in)))) - (w_(YLL_(ZL_(in))) + w_(YR_(ZL_(in)))));
        if (iz == NZ - 2) wbarbar_zr = 0;
        else if (iz == NZ - 1) wbarbar_zr = wbar_zl;
        else wbarbar_zr = hx * hy * (9 * (w_(ZRR_(in)) + w_(YL_(ZRR_(in)))) - (w_(YLL_(ZRR_(in))) + w_(YR_(ZRR_(in)))));
        double v_zr = iz == NZ - 1 ? 0.0 : v_(in) + v_(ZR_(in));
        double v_zrr = v_(in) + (iz == NZ - 3 ? -v_(ZRR_(in)) : iz == NZ - 2 ? 0.0 : iz == NZ - 1 ? v_(ZLL_(in)) : v_(ZRRR_(in)));
        double v_zl = iz == 0 ? 0.0 : v_(ZL_(in)) + v_(in);
        double v_zll = v_(in) + (iz == 2 ? -v_(ZLL_(in)) : iz == 1 ? 0.0 : iz == 0 ? v_(ZRR_(in)) : v_(ZLLL_(in)));
        f1 += wbar_zr * v_zr - C1_27 * wbarbar_zr * v_zrr - wbar_zl * v_zl + C1_27 * wbarbar_zl * v_zll;
    }
    if (iz != 0) {
        {
            double ubar_xr = 9 * (u_(XR_(in)) * DZ[iz] + u_(ZL_(XR_(in))) * DZ[iz - 1]);
            ubar_xr -= (iz == 1) ? -u_(ZL_(XR_(in))) * DZ[0] : u_(ZLL_(XR_(in))) * DZ[iz - 2];
            ubar_xr -= (iz == NZ - 1) ? -u_(XR_(in)) * DZ[NZ - 1] : u_(ZR_(XR_(in))) * DZ[iz + 1];
            ubar_xr *= hy;
            double ubar_xl = 9 * (u_(in) * DZ[iz] + u_(ZL_(in)) * DZ[iz - 1]);
            ubar_xl -= (iz == 1) ? -u_(ZL_(in)) * DZ[0] : u_(ZLL_(in)) * DZ[iz - 2];
            ubar_xl -= (iz == NZ - 1) ? -u_(in) * DZ[NZ - 1] : u_(ZR_(in)) * DZ[iz + 1];
            ubar_xl *= hy;
            double ubarbar_xr = 9 * (u_(XRR_(in)) * DZW[iz] + u_(ZL_(XRR_(in))) * DZW[iz - 1]);
            ubarbar_xr -= (iz == 1) ? -u_(ZL_(XRR_(in))) * DZW[0] : u_(ZLL_(XRR_(in))) * DZW[iz - 2];
            ubarbar_xr -= (iz == NZ - 1) ? -u_(XRR_(in)) * DZW[NZ - 1] : u_(ZR_(XRR_(in))) * DZW[iz + 1];
            ubarbar_xr *= hy;
            double ubarbar_xl = 9 * (u_(XL_(in)) * DZW[iz] + u_(ZL_(XL_(in))) * DZW[iz - 1]);
            ubarbar_xl -= (iz == 1) ? -u_(ZL_(XL_(in))) * DZW[0] : u_(ZLL_(XL_(in))) * DZW[iz - 2];
            ubarbar_xl -= (iz == NZ - 1) ? -u_(XL_(in)) * DZW[NZ - 1] : u_(ZR_(XL_(in))) * DZW[iz + 1];
            ubarbar_xl *= hy;
            f2 += ubar_xr * (w_(in) + w_(XR_(in))) - C1_27 * ubarbar_xr * (w_(in) + w_(XRRR_(in))) -
                  ubar_xl * (w_(XL_(in)) + w_(in)) + C1_27 * ubarbar_xl * (w_(XLLL_(in)) + w_(in));
        }
        {
            double vbar_yr = 9 * (v_(YR_(in)) * DZ[iz] + v_(ZL_(YR_(in))) * DZ[iz - 1]);
            vbar_yr -= (iz == 1) ? -v_(ZL_(YR_(in))) * DZ[0] : v_(ZLL_(YR_(in))) * DZ[iz - 2];
            vbar_yr -= (iz == NZ - 1) ? -v_(YR_(in)) * DZ[NZ - 1] : v_(ZR_(YR_(in))) * DZ[iz + 1];
            vbar_yr *= hy;
            double vbar_yl = 9 * (v_(in) * DZ[iz] + v_(ZL_(in)) * DZ[iz - 1]);
            vbar_yl -= (iz == 1) ? -v_(ZL_(in)) * DZ[0] : v_(ZLL_(in)) * DZ[iz - 2];
            vbar_yl -= (iz == NZ - 1) ? -v_(in) * DZ[NZ - 1] : v_(ZR_(in)) * DZ[iz + 1];
            vbar_yl *= hy;
            double vbarbar_yr = 9 * (v_(YRR_(in)) * DZW[iz] + v_(ZL_(YRR_(in))) * DZW[iz - 1]);
            vbarbar_yr -= (iz == 1) ? -v_(ZL_(YRR_(in))) * DZW[0] : v_(ZLL_(YRR_(in))) * DZW[iz - 2];
            vbarbar_yr -= (iz == NZ - 1) ? -v_(YRR_(in)) * DZW[NZ - 1] : v_(ZR_(YRR_(in))) * DZW[iz + 1];
            vbarbar_yr *= hy;
            double vbarbar_yl = 9 * (v_(YL_(in)) * DZW[iz] + v_(ZL_(YL_(in))) * DZW[iz - 1]);
            vbarbar_yl -= (iz == 1) ? -v_(ZL_(YL_(in))) * DZW[0] : v_(ZLL_(YL_(in))) * DZW[iz - 2];
            vbarbar_yl -= (iz == NZ - 1) ? -v_(YL_(in)) * DZW[NZ - 1] : v_(ZR_(YL_(in))) * DZW[iz + 1];
            vbarbar_yl *= hy;
            f2 += vbar_yr * (w_(in) + w_(YR_(in))) - C1_27 * vbarbar_yr * (w_(in) + w_(YRRR_(in))) -
                  vbar_yl * (w_(YL_(in)) + w_(in)) + C1_27 * vbarbar_yl * (w_(YLLL_(in)) + w_(in));
        }
        {
            double wzl = w_(ZL_(in));
            double wzr = iz == NZ - 1 ? 0.0 : w_(ZR_(in));
            double wzll = w_(iz == 1 ? in : ZLL_(in));
            double wzrr = iz == NZ - 2 ? 0.0 : iz == NZ - 1 ? w_(in) : w_(ZRR_(in));
            double wzlll = iz == 1 ? w_(ZR_(in)) : iz == 2 ? w_(ZL_(in)) : w_(ZLLL_(in));
            double wzrrr = iz == NZ - 3 ? 0.0 : iz == NZ - 2 ? w_(ZR_(in)) : iz == NZ - 1 ? w_(ZL_(in)) : w_(ZRRR_(in));
            double wbar_zl = hx * hy * (9 * (wzl + w_(in)) - (wzll + wzr));
            double wbar_zr = hx * hy * (9 * (w_(in) + wzr) - (wzl + wzrr));
            double wbarbar_zl = hx * hy * (9 * (wzll + wzl) - (wzlll + w_(in)));
            double wbarbar_zr = hx * hy * (9 * (wzr + wzrr) - (w_(in) + wzrrr));
            f2 += wbar_zr * (w_(in) + wzr) - wbar_zl * (wzl + w_(in));
            if (iz < NZ - 2) f2 -= C1_27 * wbarbar_zr * (w_(in) + wzrrr);
            else if (iz == NZ - 2) f2 -= C1_27 * wbarbar_zr * (-w_(ZR_(in)) + w_(in));
            else f2 -= C1_27 * wbarbar_zr * (w_(ZL_(in)) + w_(in));
            if (iz > 2) f2 += C1_27 * wbarbar_zl * (wzlll + w_(in));
            else if (iz == 2) f2 += C1_27 * wbarbar_zl * (-w_(ZL_(in)) + w_(in));
            else f2 += C1_27 * wbarbar_zl * (w_(ZR_(in)) + w_(in));
        }
    }
    double minus_inv_vol = -1.0 / (32 * (DZ[iz] - C1_9 * DZW[iz]) * hx * hy);
    out[0] = f0 * minus_inv_vol;
    out[1] = f1 * minus_inv_vol;
    out[2] = f2 * (-(1. / 32.) * inv_hx * inv_hy * GCZ[iz]);
}

/* ViscFlux<CROSS> (staggered/func_global.h:533-562). */
static void visc_flux_cross(const or_solver* s, int in, const double* vel, double* out) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double* DZ = s->DZ;
    const double inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    double inv_hz = 1 / DZ[iz];
    double inv_hz_alt = 2.0 / (DZ[iz] + (iz == 0 ? 0.0 : DZ[iz - 1]));
    double f[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) f[c] += inv_hx * inv_hx * (vel[3 * XR_(in) + c] + vel[3 * XL_(in) + c] - 2. * vel[3 * in + c]);
    for (int c = 0; c < 3; c++) f[c] += inv_hy * inv_hy * (vel[3 * YR_(in) + c] + vel[3 * YL_(in) + c] - 2. * vel[3 * in + c]);
    if (iz != NZ - 1) {
        double inv_hz_up = 2.0 / (DZ[iz] + DZ[iz + 1]);
        f[0] += (u_(ZR_(in)) - u_(in)) * inv_hz_up * inv_hz;
        f[1] += (v_(ZR_(in)) - v_(in)) * inv_hz_up * inv_hz;
        f[2] += (w_(ZR_(in)) - w_(in)) * inv_hz * inv_hz_alt;
    } else {
        double inv_hz_up = 2.0 / (DZ[iz]);
        f[0] += (-u_(in)) * inv_hz_up * inv_hz;
        f[1] += (-v_(in)) * inv_hz_up * inv_hz;
        f[2] += (-w_(in)) * inv_hz * inv_hz_alt;
    }
    if (iz != 0) {
        f[0] += (u_(ZL_(in)) - u_(in)) * inv_hz_alt * inv_hz;
        f[1] += (v_(ZL_(in)) - v_(in)) * inv_hz_alt * inv_hz;
        f[2] += (w_(ZL_(in)) - w_(in)) * (1 / DZ[iz - 1]) * inv_hz_alt;
    } else {
        f[0] += (-u_(in)) * inv_hz_alt * inv_hz;
        f[1] += (-v_(in)) * inv_hz_alt * inv_hz;
    }
    const double visc = s->P.visc;
    out[0] = visc * f[0];
    out[1] = visc * f[1];
    out[2] = visc * f[2];
}

/* ViscFlux<NAIVE> with nu_t (staggered/func_global.h:565-625), incl. the
 * mu(XR(YR(in))) term of sxy_190 at :598, kept as in the reference. */
static void visc_flux_naive(const or_solver* s, int in, const double* vel, double* out) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double* DZ = s->DZ;
    const double inv_hx = s->inv_hx, inv_hy = s->inv_hy, visc = s->P.visc;
    const double* nu_t = s->nu_t;
#define mu(X) (nu_t ? visc + nu_t[X] : visc)
    double inv_hz = 1 / DZ[iz];
    double inv_hz_up = 2 / (DZ[iz] + (iz == NZ - 1 ? 0.0 : DZ[iz + 1]));
    double inv_hz_down = 2 / (DZ[iz] + (iz == 0 ? 0.0 : DZ[iz - 1]));
    double f0 = 0, f1 = 0, f2 = 0;
    double sxx_000 = 2 * (u_(XR_(in)) - u_(in)) * inv_hx * mu(in);
    double sxx_800 = 2 * (u_(in) - u_(XL_(in))) * inv_hx * mu(XL_(in));
    f0 += (sxx_000 - sxx_800) * inv_hx;
    double sxy_910 = ((u_(YR_(in)) - u_(in)) * inv_hy + (v_(YR_(in)) - v_(XL_(YR_(in)))) * inv_hx) *
                     0.25 * (mu(in) + mu(XL_(in)) + mu(YR_(in)) + mu(XL_(YR_(in))));
    double sxy_990 = ((u_(in) - u_(YL_(in))) * inv_hy + (v_(in) - v_(XL_(in))) * inv_hx) *
                     0.25 * (mu(YL_(in)) + mu(YL_(XL_(in))) + mu(in) + mu(XL_(in)));
    f0 += (sxy_910 - sxy_990) * inv_hy;
    double sxz_901 = (iz == NZ - 1) ? (-2 * u_(in) * inv_hz * visc)
                                    : ((u_(ZR_(in)) - u_(in)) * inv_hz_up + (w_(ZR_(in)) - w_(XL_(ZR_(in)))) * inv_hx) *
                                          0.25 * (mu(in) + mu(XL_(in)) + mu(ZR_(in)) + mu(XL_(ZR_(in))));
    double sxz_909 = (iz == 0) ? (2 * u_(in) * inv_hz * visc)
                               : ((u_(in) - u_(ZL_(in))) * inv_hz_down + (w_(in) - w_(XL_(in))) * inv_hx) *
                                     0.25 * (mu(ZL_(in)) + mu(ZL_(XL_(in))) + mu(in) + mu(XL_(in)));
    f0 += (sxz_901 - sxz_909) * inv_hz;

    double sxy_190 = ((u_(XR_(in)) - u_(XR_(YL_(in)))) * inv_hy + (v_(XR_(in)) - v_(in)) * inv_hx) *
                     0.25 * (mu(in) + mu(XR_(in)) + mu(YL_(in)) + mu(XR_(YR_(in))));
    f1 += (sxy_190 - sxy_990) * inv_hx;
    double syy_000 = 2 * (v_(YR_(in)) - v_(in)) * inv_hy * mu(in);
    double syy_080 = 2 * (v_(in) - v_(YL_(in))) * inv_hy * mu(YL_(in));
    f1 += (syy_000 - syy_080) * inv_hy;
    double syz_091 = (iz == NZ - 1) ? (-2 * v_(in) * inv_hz * visc)
                                    : ((v_(ZR_(in)) - v_(in)) * inv_hz_up + (w_(ZR_(in)) - w_(YL_(ZR_(in)))) * inv_hy) *
                                          0.25 * (mu(in) + mu(YL_(in)) + mu(ZR_(in)) + mu(YL_(ZR_(in))));
    double syz_099 = (iz == 0) ? (2 * v_(in) * inv_hz * visc)
                               : ((v_(in) - v_(ZL_(in))) * inv_hz_down + (w_(in) - w_(YL_(in))) * inv_hy) *
                                     0.25 * (mu(ZL_(in)) + mu(ZL_(YL_(in))) + mu(in) + mu(YL_(in)));
    f1 += (syz_091 - syz_099) * inv_hz;

    if (iz != 0) {
        double sxz_109 = ((u_(XR_(in)) - u_(XR_(ZL_(in)))) * inv_hz_down + (w_(XR_(in)) - w_(in)) * inv_hx) *
                         0.25 * (mu(ZL_(in)) + mu(ZL_(XR_(in))) + mu(in) + mu(XR_(in)));
        f2 += (sxz_109 - sxz_909) * inv_hx;
        double syz_019 = ((v_(YR_(in)) - v_(YR_(ZL_(in)))) * inv_hz_down + (w_(YR_(in)) - w_(in)) * inv_hy) *
                         0.25 * (mu(ZL_(in)) + mu(ZL_(YR_(in))) + mu(in) + mu(YR_(in)));
        f2 += (syz_019 - syz_099) * inv_hy;
        double szz_000 = 2 * ((iz == NZ - 1 ? 0.0 : w_(ZR_(in))) - w_(in)) * inv_hz * mu(in);
        double szz_008 = 2 * (w_(in) - (iz == 0 ? 0.0 : w_(ZL_(in)))) * mu(ZL_(in)) / DZ[iz - 1];
        f2 += (szz_000 - szz_008) * inv_hz_down;
    }
#undef mu
    out[0] = f0;
    out[1] = f1;
    out[2] = f2;
}

/* ApplyGradient (staggered/func_global.h:297-303 HW, :316-322 VV). */
static void grad_staggered(const or_solver* s, int in, const double* a, double* ga) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double *DZ = s->DZ, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    if (s->P.scheme == OR_SCHEME_HW) {
        ga[3 * in + 0] = inv_hx * (a[in] - a[XL_(in)]);
        ga[3 * in + 1] = inv_hy * (a[in] - a[YL_(in)]);
        if (iz == 0) ga[3 * in + 2] = 0;
        else ga[3 * in + 2] = (a[in] - a[ZL_(in)]) * (2 / (DZ[iz] + DZ[iz - 1]));
    } else {
        const double *GC1 = s->GC1, *GC2 = s->GC2, *GCZ = s->GCZ;
        ga[3 * in + 0] = inv_hx * (GC1[iz] * (a[in] - a[XL_(in)]) + GC2[iz] * (a[XR_(in)] - a[XL_(XL_(in))]));
        ga[3 * in + 1] = inv_hy * (GC1[iz] * (a[in] - a[YL_(in)]) + GC2[iz] * (a[YR_(in)] - a[YL_(YL_(in))]));
        if (iz == 0) ga[3 * in + 2] = 0;
        else
            ga[3 * in + 2] = ((a[in] - a[ZL_(in)]) - C1_27 * ((iz == NZ - 1 ? -a[in] : a[ZR_(in)]) -
                                                              (iz == 1 ? -a[ZL_(in)] : a[ZL_(ZL_(in))]))) * GCZ[iz];
    }
}

/* ApplyDivergence (staggered/func_global.h:306-312 HW, :325-333 VV). */
static void div_staggered(const or_solver* s, int in, const double* v, double* dv) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double *DZ = s->DZ, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    if (s->P.scheme == OR_SCHEME_HW) {
        double hz = DZ[iz];
        dv[in] = (v[3 * XR_(in) + 0] - v[3 * in + 0]) * inv_hx * hz + (v[3 * YR_(in) + 1] - v[3 * in + 1]) * inv_hy * hz +
                 ((in < (NZ - 1) * s->NXY) ? (v[3 * ZR_(in) + 2] - v[3 * in + 2]) : (-v[3 * in + 2]));
    } else {
        const double* DZW = s->DZW;
        double dudx = inv_hx * (DZ[iz] * (v[3 * XR_(in)] - v[3 * in]) - C1_27 * DZW[iz] * (v[3 * XR_(XR_(in))] - v[3 * XL_(in)]));
        double dvdy = inv_hy * (DZ[iz] * (v[3 * YR_(in) + 1] - v[3 * in + 1]) -
                                C1_27 * DZW[iz] * (v[3 * YR_(YR_(in)) + 1] - v[3 * YL_(in) + 1]));
        double wr = (iz == NZ - 1) ? 0.0 : v[3 * ZR_(in) + 2];
        double wrr = (iz == NZ - 1) ? v[3 * in + 2] : (iz == NZ - 2) ? 0.0 : v[3 * ZR_(ZR_(in)) + 2];
        double wll = (iz == 0) ? v[3 * ZR_(in) + 2] : (iz == 1) ? 0.0 : v[3 * ZL_(in) + 2];
        dv[in] = dudx + dvdy + (wr - v[3 * in + 2]) - C1_27 * (wrr - wll);
    }
}

/* CalcAbsS (staggered/func_global.h:823-870). */
static double calc_abs_s(const or_solver* s, int in, const double* vel) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double *DZ = s->DZ, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    double inv_hz = 1 / DZ[iz];
    double inv_hz_up = 2 / (DZ[iz] + (iz == NZ - 1 ? 0.0 : DZ[iz + 1]));
    double inv_hz_down = 2 / (DZ[iz] + (iz == 0 ? 0.0 : DZ[iz - 1]));
    double AbsS = 0;
    {
        double sxx = (u_(XR_(in)) - u_(in)) * inv_hx;
        double syy = (v_(YR_(in)) - v_(in)) * inv_hy;
        double szz = ((iz == NZ - 1 ? 0.0 : w_(ZR_(in))) - w_(in)) * inv_hz;
        AbsS = sxx * sxx + syy * syy + szz * szz;
    }
    {
        double sxy_910 = (u_(YR_(in)) - u_(in)) * inv_hy + (v_(YR_(in)) - v_(XL_(YR_(in)))) * inv_hx;
        double sxy_990 = (u_(in) - u_(YL_(in))) * inv_hy + (v_(in) - v_(XL_(in))) * inv_hx;
        double sxy_110 = (u_(XR_(YR_(in))) - u_(XR_(in))) * inv_hy + (v_(XR_(YR_(in))) - v_(YR_(in))) * inv_hx;
        double sxy_190 = (u_(XR_(in)) - u_(XR_(YL_(in)))) * inv_hy + (v_(XR_(in)) - v_(in)) * inv_hx;
        AbsS += 0.125 * (sxy_910 * sxy_910 + sxy_990 * sxy_990 + sxy_110 * sxy_110 + sxy_190 * sxy_190);
    }
    {
        double sxz_901 = (iz == NZ - 1) ? (-2 * u_(in) * inv_hz) : (u_(ZR_(in)) - u_(in)) * inv_hz_up + (w_(ZR_(in)) - w_(XL_(ZR_(in)))) * inv_hx;
        double sxz_909 = (iz == 0) ? (2 * u_(in) * inv_hz) : (u_(in) - u_(ZL_(in))) * inv_hz_down + (w_(in) - w_(XL_(in))) * inv_hx;
        double sxz_101 = (iz == NZ - 1) ? (-2 * u_(XR_(in)) * inv_hz)
                                        : (u_(XR_(ZR_(in))) - u_(XR_(in))) * inv_hz_up + (w_(XR_(ZR_(in))) - w_(ZR_(in))) * inv_hx;
        double sxz_109 = (iz == 0) ? (2 * u_(XR_(in)) * inv_hz)
                                   : (u_(XR_(in)) - u_(ZL_(XR_(in)))) * inv_hz_down + (w_(XR_(in)) - w_(in)) * inv_hx;
        AbsS += 0.125 * (sxz_901 * sxz_901 + sxz_909 * sxz_909 + sxz_101 * sxz_101 + sxz_109 * sxz_109);
    }
    {
        double syz_091 = (iz == NZ - 1) ? (-2 * v_(in) * inv_hz) : (v_(ZR_(in)) - v_(in)) * inv_hz_up + (w_(ZR_(in)) - w_(YL_(ZR_(in)))) * inv_hy;
        double syz_099 = (iz == 0) ? (2 * v_(in) * inv_hz) : (v_(in) - v_(ZL_(in))) * inv_hz_down + (w_(in) - w_(YL_(in))) * inv_hy;
        double syz_011 = (iz == NZ - 1) ? (-2 * v_(YR_(in)) * inv_hz)
                                        : (v_(YR_(ZR_(in))) - v_(YR_(in))) * inv_hz_up + (w_(YR_(ZR_(in))) - w_(ZR_(in))) * inv_hy;
        double syz_019 = (iz == 0) ? (2 * v_(YR_(in)) * inv_hz)
                                   : (v_(YR_(in)) - v_(ZL_(YR_(in)))) * inv_hz_down + (w_(YR_(in)) - w_(in)) * inv_hy;
        AbsS += 0.125 * (syz_091 * syz_091 + syz_099 * syz_099 + syz_011 * syz_011 + syz_019 * syz_019);
    }
    return sqrt(2 * AbsS);
}

/* RecalcTurbVisc (staggered/func_global.h:875-890). */
static void turb_visc_staggered(or_solver* s, int in, const double* vel) {
    const int NZ = s->NZ, iz = in / s->NXY;
    const double* ZZ = s->ZZ;
    const double H = 0.5 * (ZZ[NZ] - ZZ[0]);
    const double Z = 0.5 * (ZZ[iz] + ZZ[iz + 1]) - ZZ[0];
    const double dist_to_wall = (Z < H) ? Z : 2 * H - Z;
    const double y_plus = dist_to_wall * s->P.inv_length_scale;
    const double C_SMAG = 0.1;
    const double C_exp = 0.04 * 0.04 * 0.04;
    const double AbsS = calc_abs_s(s, in, vel);
    const double f_VD = 1 - exp(-C_exp * y_plus * y_plus * y_plus);
    const double V = s->P.hx * s->P.hy * s->DZ[iz];
    double Delta_LES = C_SMAG * pow(V, C1_3);
    if (Delta_LES > 0.41 * dist_to_wall) Delta_LES = 0.41 * dist_to_wall;
    s->nu_t[in] = f_VD * Delta_LES * Delta_LES * AbsS;
}

/* ------------------------------------------------------------------------- */
/* Collocated-AB3 operators                                                   */
/* ------------------------------------------------------------------------- */
static inline int is_wall(const or_solver* s, int i) { return i < s->NXY || i >= s->N3D - s->NXY; }

/* GetCellHeight (collocated_AB3/func_global.h:150-155). */
static inline double cell_height(const or_solver* s, int in) {
    int iz = in / s->NXY;
    if (iz == 0) return 0.5 * s->DZ[0];
    if (iz == s->NZ - 1) return 0.5 * s->DZ[s->NZ - 2];
    return 0.5 * (s->DZ[iz - 1] + s->DZ[iz]);
}
static inline double cell_volume(const or_solver* s, int in) { return cell_height(s, in) * s->P.hx * s->P.hy; }

/* GetElementIndex (collocated_AB3/func_fluxes.h:4-9). */
static inline int element_index(const or_solver* s, int in, int ox, int oy, int oz) {
    if (!ox) in = XL_(in);
    if (!oy) in = YL_(in);
    if (!oz) in = ZL_(in);
    return in;
}

/* CalcViscTermGalerkin (collocated_AB3/func_fluxes.h:21-85). */
static void visc_galerkin(const or_solver* s, int in, const double* u, double* out) {
    const int NZ = s->NZ, NXY = s->NXY;
    const double hx = s->P.hx, hy = s->P.hy, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    double sum[3] = {0, 0, 0};
    for (int oz = 0; oz <= 1; oz++) {
        if (s->P.fd_mode) {
            if (is_wall(s, in)) continue;
        } else {
            if ((oz == 0 && in < NXY) || (oz == 1 && in >= (NZ - 1) * NXY)) continue;
        }
        int in_z = oz ? ZR_(in) : ZL_(in);
        int iz = in / NXY;
        double hz = oz == 0 ? s->DZ[iz == 0 ? 0 : iz - 1] : s->DZ[iz];
        double inv_hz = 1. / hz;
        for (int oy = 0; oy <= 1; oy++) {
            int in_y = oy ? YR_(in) : YL_(in);
            int in_yz = oy ? YR_(in_z) : YL_(in_z);
            for (int ox = 0; ox <= 1; ox++) {
                int in_x = ox ? XR_(in) : XL_(in);
                int in_xy = ox ? XR_(in_y) : XL_(in_y);
                int in_xz = ox ? XR_(in_z) : XL_(in_z);
                int in_xyz = ox ? XR_(in_yz) : XL_(in_yz);
                double mx = inv_hx * hy * hz;
                double my = inv_hy * hx * hz;
                double mz = inv_hz * hx * hy;
#define U(n, c) u[3 * (n) + (c)]
                double g[3];
                for (int c = 0; c < 3; c++)
                    g[c] = (C1_9 * (U(in_x, c) - U(in, c)) + C1_18 * (U(in_xy, c) - U(in_y, c)) + C1_18 * (U(in_xz, c) - U(in_z, c)) +
                            C1_36 * (U(in_xyz, c) - U(in_yz, c))) * mx +
                           (C1_9 * (U(in_y, c) - U(in, c)) + C1_18 * (U(in_xy, c) - U(in_x, c)) + C1_18 * (U(in_yz, c) - U(in_z, c)) +
                            C1_36 * (U(in_xyz, c) - U(in_xz, c))) * my +
                           (C1_9 * (U(in_z, c) - U(in, c)) + C1_18 * (U(in_xz, c) - U(in_x, c)) + C1_18 * (U(in_yz, c) - U(in_y, c)) +
                            C1_36 * (U(in_xyz, c) - U(in_xy, c))) * mz;
                double gxx = C1_9 * (U(in_x, 0) - U(in, 0)) + C1_18 * (U(in_xy, 0) - U(in_y, 0)) + C1_18 * (U(in_xz, 0) - U(in_z, 0)) + C1_36 * (U(in_xyz, 0) - U(in_yz, 0));
                double gxy = 0.25 * (C1_3 * (U(in_x, 1) - U(in, 1) + U(in_xy, 1) - U(in_y, 1)) + C1_6 * (U(in_xz, 1) - U(in_z, 1) + U(in_xyz, 1) - U(in_yz, 1)));
                double gxz = 0.25 * (C1_3 * (U(in_x, 2) - U(in, 2) + U(in_xz, 2) - U(in_z, 2)) + C1_6 * (U(in_xy, 2) - U(in_y, 2) + U(in_xyz, 2) - U(in_yz, 2)));
                double gyy = C1_9 * (U(in_y, 1) - U(in, 1)) + C1_18 * (U(in_xy, 1) - U(in_x, 1)) + C1_18 * (U(in_yz, 1) - U(in_z, 1)) + C1_36 * (U(in_xyz, 1) - U(in_xz, 1));
                double gyx = 0.25 * (C1_3 * (U(in_y, 0) - U(in, 0) + U(in_xy, 0) - U(in_x, 0)) + C1_6 * (U(in_yz, 0) - U(in_z, 0) + U(in_xyz, 0) - U(in_xz, 0)));
                double gyz = 0.25 * (C1_3 * (U(in_y, 2) - U(in, 2) + U(in_yz, 2) - U(in_z, 2)) + C1_6 * (U(in_xy, 2) - U(in_x, 2) + U(in_xyz, 2) - U(in_xz, 2)));
                double gzz = C1_9 * (U(in_z, 2) - U(in, 2)) + C1_18 * (U(in_yz, 2) - U(in_y, 2)) + C1_18 * (U(in_xz, 2) - U(in_x, 2)) + C1_36 * (U(in_xyz, 2) - U(in_xy, 2));
                double gzx = 0.25 * (C1_3 * (U(in_z, 0) - U(in, 0) + U(in_xz, 0) - U(in_x, 0)) + C1_6 * (U(in_yz, 0) - U(in_y, 0) + U(in_xyz, 0) - U(in_xy, 0)));
                double gzy = 0.25 * (C1_3 * (U(in_z, 1) - U(in, 1) + U(in_yz, 1) - U(in_y, 1)) + C1_6 * (U(in_xz, 1) - U(in_x, 1) + U(in_xyz, 1) - U(in_xy, 1)));
#undef U
                gxx *= mx;
                gxy *= hz * (ox ^ oy ? -1 : 1);
                gxz *= hy * (ox ^ oz ? -1 : 1);
                gyy *= my;
                gyx *= hz * (ox ^ oy ? -1 : 1);
                gyz *= hx * (oy ^ oz ? -1 : 1);
                gzz *= mz;
                gzx *= hy * (ox ^ oz ? -1 : 1);
                gzy *= hx * (oy ^ oz ? -1 : 1);
                g[0] += gxx + gxy + gxz;
                g[1] += gyy + gyx + gyz;
                g[2] += gzz + gzx + gzy;
                double visc_coeff = s->P.visc;
                if (s->nu_t) visc_coeff += s->nu_t[element_index(s, in, ox, oy, oz)];
                for (int c = 0; c < 3; c++) sum[c] += g[c] * visc_coeff;
            }
        }
    }
    out[0] = sum[0]; out[1] = sum[1]; out[2] = sum[2];
}

/* CalcFluxTerm_FD at one node (collocated_AB3/func_fluxes.h:87-120). */
static void flux_ab3(const or_solver* s, int in, const double* u, int do_conv, int do_visc, double* kuhat) {
    double sum[3] = {0, 0, 0};
    if (!is_wall(s, in)) {
        int iz = in / s->NXY;
        double inv_hz = 2 / (s->DZ[iz - 1] + s->DZ[iz]);
        if (do_conv) {
            const double inv_h[3] = {s->inv_hx, s->inv_hy, inv_hz};
            for (int idir = 0; idir < 3; idir++) {
                const double invh = inv_h[idir];
                int jn = idir == 0 ? XL_(in) : idir == 1 ? YL_(in) : ZL_(in);
                int kn = idir == 0 ? XR_(in) : idir == 1 ? YR_(in) : ZR_(in);
                for (int c = 0; c < 3; c++) {
                    double conv_div = 0.5 * (u[3 * kn + c] * u[3 * kn + idir] - u[3 * jn + c] * u[3 * jn + idir]);
                    double conv_adv = u[3 * in + idir] * 0.5 * (u[3 * kn + c] - u[3 * jn + c]);
                    sum[c] -= 0.5 * (conv_div + conv_adv) * invh;
                }
            }
        }
        if (do_visc) {
            if (s->P.visc_type == 1) {
                double g[3];
                visc_galerkin(s, in, u, g);
                for (int c = 0; c < 3; c++) sum[c] += g[c] * s->inv_hx * s->inv_hy * inv_hz;
            } else {
                const double visc = s->P.visc, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
                for (int c = 0; c < 3; c++) sum[c] += visc * inv_hx * inv_hx * (u[3 * XR_(in) + c] + u[3 * XL_(in) + c] - 2. * u[3 * in + c]);
                for (int c = 0; c < 3; c++) sum[c] += visc * inv_hy * inv_hy * (u[3 * YR_(in) + c] + u[3 * YL_(in) + c] - 2. * u[3 * in + c]);
                for (int c = 0; c < 3; c++)
                    sum[c] += visc * ((u[3 * ZR_(in) + c] - u[3 * in + c]) / s->DZ[iz] - (u[3 * in + c] - u[3 * ZL_(in) + c]) / s->DZ[iz - 1]) * inv_hz;
            }
        }
        if (do_visc && s->tau) {
            /* learned closure (new numerics): -d tau_ij / d x_j, central differences in x, y
             * and the 3-point nonuniform formula in z; tau is zero on the wall planes */
            static const int TC[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
            const double* t = s->tau;
            const double hm = s->DZ[iz - 1], hp = s->DZ[iz];
            const double cp = hm * hm / (hp * hm * (hp + hm)), cm = hp * hp / (hp * hm * (hp + hm));
            for (int c = 0; c < 3; c++) {
                const int ax = TC[c][0], ay = TC[c][1], az = TC[c][2];
                double d = 0.5 * s->inv_hx * (t[6 * XR_(in) + ax] - t[6 * XL_(in) + ax]);
                d += 0.5 * s->inv_hy * (t[6 * YR_(in) + ay] - t[6 * YL_(in) + ay]);
                d += cp * (t[6 * ZR_(in) + az] - t[6 * in + az]) + cm * (t[6 * in + az] - t[6 * ZL_(in) + az]);
                sum[c] -= d;
            }
        }
        sum[0] += s->P.sour_dpdx;
    }
    kuhat[3 * in + 0] = sum[0];
    kuhat[3 * in + 1] = sum[1];
    kuhat[3 * in + 2] = sum[2];
}

/* ApplyGradient (collocated_AB3/func_global.h:298-303). */
static void grad_ab3(const or_solver* s, int in, const double* a, double* R) {
    if (is_wall(s, in)) { R[3 * in] = R[3 * in + 1] = R[3 * in + 2] = 0; return; }
    R[3 * in + 0] = 0.5 * s->inv_hx * (a[XR_(in)] - a[XL_(in)]);
    R[3 * in + 1] = 0.5 * s->inv_hy * (a[YR_(in)] - a[YL_(in)]);
    R[3 * in + 2] = 0.5 * (a[ZR_(in)] - a[ZL_(in)]) / cell_height(s, in);
}

/* ApplyDivergence (collocated_AB3/func_global.h:310-325). */
static void div_ab3(const or_solver* s, int in, const double* v, double* dv) {
    const double hx = s->P.hx, hy = s->P.hy;
    double d = 0, hz = cell_height(s, in);
    d += 0.5 * (v[3 * XR_(in)] - v[3 * XL_(in)]) * hy * hz;
    d += 0.5 * (v[3 * YR_(in) + 1] - v[3 * YL_(in) + 1]) * hx * hz;
    if (in < s->NXY) d += 0.5 * (v[3 * ZR_(in) + 2] - v[3 * in + 2]) * hx * hy;
    else if (in >= (s->NZ - 1) * s->NXY) d += 0.5 * (v[3 * in + 2] - v[3 * ZL_(in) + 2]) * hx * hy;
    else d += 0.5 * (v[3 * ZR_(in) + 2] - v[3 * ZL_(in) + 2]) * hx * hy;
    dv[in] = d;
}

/* RecalcTurbVisc on elements (collocated_AB3/func_global.h:342-384). */
static void turb_visc_ab3(or_solver* s, int ie, const double* u) {
    const int NZ = s->NZ, iz = ie / s->NXY;
    if (iz >= NZ - 1) return;
    const double hz = s->DZ[iz], inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    double AbsS;
    {
        int in = ie, in_x = XR_(in), in_y = YR_(in), in_xy = YR_(in_x), in_z = ZR_(in), in_xz = ZR_(in_x), in_yz = ZR_(in_y), in_xyz = ZR_(in_xy);
        double gu[3][3];
#define U(n, c) u[3 * (n) + (c)]
        for (int c = 0; c < 3; c++) {
            gu[0][c] = ((U(in_x, c) + U(in_xy, c) + U(in_xz, c) + U(in_xyz, c)) - (U(in, c) + U(in_y, c) + U(in_z, c) + U(in_yz, c))) * (0.25 * inv_hx);
            gu[1][c] = ((U(in_y, c) + U(in_xy, c) + U(in_yz, c) + U(in_xyz, c)) - (U(in, c) + U(in_x, c) + U(in_z, c) + U(in_xz, c))) * (0.25 * inv_hy);
            gu[2][c] = ((U(in_z, c) + U(in_xz, c) + U(in_yz, c) + U(in_xyz, c)) - (U(in, c) + U(in_x, c) + U(in_y, c) + U(in_xy, c))) / (4 * hz);
        }
#undef U
        gu[0][1] = 0.5 * (gu[0][1] + gu[1][0]);
        gu[0][2] = 0.5 * (gu[0][2] + gu[2][0]);
        gu[1][2] = 0.5 * (gu[1][2] + gu[2][1]);
        AbsS = SQR(gu[0][0]) + SQR(gu[1][1]) + SQR(gu[2][2]) + 2 * (SQR(gu[0][1]) + SQR(gu[0][2]) + SQR(gu[1][2]));
        AbsS = sqrt(2 * AbsS);
    }
    const double* ZZ = s->ZZ;
    const double H = 0.5 * (ZZ[NZ - 1] - ZZ[0]);
    const double Z = 0.5 * (ZZ[iz] + ZZ[iz + 1]);
    const double dist_to_wall = (Z < H) ? Z : 2 * H - Z;
    const double y_plus = dist_to_wall * s->P.inv_length_scale;
    const double C_SMAG = 0.1, C_exp = 0.04 * 0.04 * 0.04;
    const double f_VD = 1 - exp(-C_exp * y_plus * y_plus * y_plus);
    const double V = s->P.hx * s->P.hy * hz;
    double Delta_LES = C_SMAG * pow(V, C1_3);
    if (Delta_LES > 0.41 * dist_to_wall) Delta_LES = 0.41 * dist_to_wall;
    s->nu_t[ie] = f_VD * Delta_LES * Delta_LES * AbsS;
}

/* GetNuTAtNode (collocated_AB3/func_global.h:183-186). */
static double nut_at_node(const or_solver* s, int i) {
    if (is_wall(s, i) || !s->nu_t) return 0.0;
    const double* n = s->nu_t;
    return 0.125 * (n[i] + n[XL_(i)] + n[YL_(i)] + n[XL_(YL_(i))] + n[ZL_(i)] + n[XL_(ZL_(i))] + n[YL_(ZL_(i))] + n[XL_(YL_(ZL_(i)))]);
}

/* ------------------------------------------------------------------------- */
/* Pressure solver                                                            */
/* ------------------------------------------------------------------------- */

/* ThomasForPressure, HW (staggered/func_global.h:71-106), CPU stride NX. */
static void thomas_hw(or_solver* s, int ixy, double* p) {
    const int NX = s->NX, NY = s->NY, NZ = s->NZ, stride = NX * NY;
    const double *DZ = s->DZ, *MAT_A = s->MAT_A, inv_hx = s->inv_hx, inv_hy = s->inv_hy;
    int ix = ixy % NX, iy = ixy / NX;
    double MM = 2 * ((1 - s->cosphi[0][ix]) * inv_hx * inv_hx + (1 - s->cosphi[1][iy]) * inv_hy * inv_hy);
    cplx* x = s->spec + ixy;
    double b0 = -DZ[0] * MM - MAT_A[1];
    if (ixy == 0) b0 *= 2;
    double inv_b0 = 1 / b0;
    x[0] *= inv_b0;
    p[0] = MAT_A[1] * inv_b0;
    for (int i = 1; i <= NZ - 1; i++) {
        double delta = -MAT_A[i] - DZ[i] * MM;
        if (i != NZ - 1) delta -= MAT_A[i + 1];
        delta -= MAT_A[i] * p[i - 1];
        delta = 1 / delta;
        if (i != NZ - 1) p[i] = MAT_A[i + 1] * delta;
        x[(size_t)i * stride] = (x[(size_t)i * stride] - MAT_A[i] * x[(size_t)(i - 1) * stride]) * delta;
    }
    for (int i = NZ - 2; i >= 0; i--) x[(size_t)i * stride] -= x[(size_t)(i + 1) * stride] * p[i];
}

/* ThomasForPressure, VV 7-band (staggered/func_global.h:110-189), CPU stride NX. */
static void thomas_vv(or_solver* s, int ixy, double* buf) {
    const int NX = s->NX, NY = s->NY, N = s->NZ, w = 3, stride = NX * NY;
    const double *DZ = s->DZ, *DZW = s->DZW, *MAT_A = s->MAT_A, *MAT_D = s->MAT_D;
    const double inv_hx = s->inv_hx, inv_hy = s->inv_hy, tiny = 1e-30;
    cplx* x = s->spec + ixy;
    for (int i = 0; i < N; i++) {
        double cur_line[7] = {0, 0, 0, 0, 0, 0, 0};
        double sum = 0;
        for (int k = 0; k < w; k++) {
            if (i + k - w < 0 || i + k - w >= N) continue;
            cur_line[k] = MAT_A[i * w + k];
            sum += MAT_A[i * w + k];
        }
        for (int k = 1; k <= w; k++) {
            if (i + k < 0 || i + k >= N) continue;
            cur_line[w + k] = MAT_A[(i + k) * w + w - k];
            sum += MAT_A[(i + k) * w + w - k];
        }
        if (i <= 2) cur_line[w] = MAT_D[i];
        else if (i >= N - 3) cur_line[w] = MAT_D[i - (N - 3) + 3];
        else cur_line[w] = -sum;
        if (ixy == 0 && i == 0) cur_line[w] *= 2;
        {
            int ix = ixy % NX, iy = ixy / NX;
            double sin_phi_2 = sin(PI_D * ix / NX);
            double sin_3phi_2 = 3 * sin_phi_2 - 4 * sin_phi_2 * sin_phi_2 * sin_phi_2;
            double ci = DZ[i] * sin_phi_2 - C1_27 * DZW[i] * sin_3phi_2;
            cur_line[w] -= 4 * inv_hx * inv_hx * ci * ci / (DZ[i] - C1_9 * DZW[i]);
            sin_phi_2 = sin(PI_D * iy / NY);
            sin_3phi_2 = 3 * sin_phi_2 - 4 * sin_phi_2 * sin_phi_2 * sin_phi_2;
            ci = DZ[i] * sin_phi_2 - C1_27 * DZW[i] * sin_3phi_2;
            cur_line[w] -= 4 * inv_hy * inv_hy * ci * ci / (DZ[i] - C1_9 * DZW[i]);
        }
        for (int j = -w; j < 0; j++) {
            if (i + j < 0) continue;
            double mult = -cur_line[w + j];
            for (int k = 1; k <= w; k++) {
                if (i + j + k >= N) continue;
                cur_line[w + j + k] += mult * buf[(i + j) * w + k - 1];
            }
            x[(size_t)i * stride] += mult * x[(size_t)(i + j) * stride];
        }
        if (fabs(cur_line[w]) < tiny) { x[0] = -1e50; return; }
        double mult = 1.0 / cur_line[w];
        for (int k = 1; k <= w; k++) {
            if (i + k >= N) continue;
            buf[i * w + k - 1] = cur_line[w + k] * mult;
        }
        x[(size_t)i * stride] *= mult;
    }
    for (int i = N - 1; i >= 0; i--)
        for (int j = 1; j <= w; j++) {
            if (i + j >= N) continue;
            x[(size_t)i * stride] -= buf[i * w + j - 1] * x[(size_t)(i + j) * stride];
        }
}

/* ThomasForPressure, collocated (collocated_AB3/func_global.h:8-61), CPU stride NX. */
static void thomas_ab3(or_solver* s, int ixy, double* p) {
    const int NX = s->NX, NY = s->NY, NZ = s->NZ, stride = NX * NY, FDmode = s->P.fd_mode;
    const double *DZ = s->DZ, *MAT_A = s->MAT_A, hx = s->P.hx, hy = s->P.hy;
    int ix = ixy % NX, iy = ixy / NX;
    double MM = (1 - s->cosphi[0][ix]) * hy / hx + (1 - s->cosphi[1][iy]) * hx / hy;
    cplx* x = s->spec + ixy;
    double b0 = -MAT_A[1];
    if (!FDmode) b0 += DZ[0] * MM;
    if (ixy == 0) b0 *= 2;
    double inv_b0 = 1 / b0;
    x[0] *= inv_b0;
    p[0] = MAT_A[1] * inv_b0;
    for (int i = 1; i <= NZ - 1; i++) {
        const double _MM = FDmode && (i == 0 || i == NZ - 1) ? 0.0 : MM;
        double delta = 0;
        if (i != 0) delta -= MAT_A[i] + DZ[i - 1] * _MM;
        if (i != NZ - 1) delta -= MAT_A[i + 1] + DZ[i] * _MM;
        delta -= MAT_A[i] * p[i - 1];
        delta = 1 / delta;
        if (i != NZ - 1) p[i] = MAT_A[i + 1] * delta;
        x[(size_t)i * stride] = (x[(size_t)i * stride] - MAT_A[i] * x[(size_t)(i - 1) * stride]) * delta;
    }
    for (int i = NZ - 2; i >= 0; i--) x[(size_t)i * stride] -= x[(size_t)(i + 1) * stride] * p[i];
}

static void fft2d_planes(or_solver* s, int dir) {
    const int NX = s->NX, NY = s->NY, NZ = s->NZ;
    int nmax = NX > NY ? NX : NY;
#pragma omp parallel
    {
        cplx* work = (cplx*)malloc(sizeof(cplx) * 2 * (size_t)nmax);
        cplx* line = (cplx*)malloc(sizeof(cplx) * (size_t)nmax);
#pragma omp for
        for (int iz = 0; iz < NZ; iz++) {
            cplx* pl = s->spec + (size_t)iz * NX * NY;
            for (int iy = 0; iy < NY; iy++) {
                fft_line(&s->px, dir, pl + (size_t)iy * NX, 1, line, 1, work);
                memcpy(pl + (size_t)iy * NX, line, sizeof(cplx) * (size_t)NX);
            }
            for (int ix = 0; ix < NX; ix++) fft_line(&s->py, dir, pl + ix, NX, pl + ix, NX, work);
        }
        free(work);
        free(line);
    }
}

/* NormalizePressure (staggered/func_global.h:657-675; collocated_AB3/func_global.h:163-170,
 * including the collocated hx*hy factor of :167). */
void or_normalize_pressure(or_solver* s, double* f) {
    const int N3D = s->N3D, NXY = s->NXY;
    double psum = 0;
    if (s->P.scheme == OR_SCHEME_HW) {
#pragma omp parallel for reduction(+ : psum)
        for (int i = 0; i < N3D; i++) psum += f[i] * s->DZ[i / NXY];
        double p_shift = psum / ((s->ZZ[s->NZ] - s->ZZ[0]) * NXY);
#pragma omp parallel for
        for (int i = 0; i < N3D; i++) f[i] -= p_shift;
    } else if (s->P.scheme == OR_SCHEME_VV) {
#pragma omp parallel for reduction(+ : psum)
        for (int i = 0; i < N3D; i++) psum += f[i] * s->KERG[i / NXY] * s->DZ[i / NXY];
        double p_shift = psum / (s->KERG_norm2 * NXY);
#pragma omp parallel for
        for (int i = 0; i < N3D; i++) f[i] -= p_shift * s->KERG[i / NXY];
    } else {
#pragma omp parallel for reduction(+ : psum)
        for (int i = 0; i < N3D; i++) psum += f[i] * cell_volume(s, i);
        /* SRK: psum / area (collocated_SRK/func_global.h:130-137) */
        double p_shift = s->P.scheme == OR_SCHEME_SRK ? psum / s->area : psum * s->P.hx * s->P.hy / s->area;
#pragma omp parallel for
        for (int i = 0; i < N3D; i++) f[i] -= p_shift;
    }
}

/* LinearSystemSolveFFT, CPU branch (staggered/func_global.h:260-280;
 * collocated_AB3/func_global.h:127-147). */
void or_solve_pressure(or_solver* s, double* XX, const double* f) {
    const int N3D = s->N3D, NXY = s->NXY, NZ = s->NZ;
    double t0 = omp_get_wtime();
    for (int in = 0; in < N3D; in++) s->spec[in] = f[in];
    fft2d_planes(s, 0);
    double t1 = omp_get_wtime();
    s->tmr[OR_TMR_FFT] += t1 - t0;
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 3 * (size_t)NZ);
#pragma omp for
        for (int ixy = 0; ixy < NXY; ixy++) {
            if (s->P.scheme == OR_SCHEME_HW) thomas_hw(s, ixy, buf);
            else if (s->P.scheme == OR_SCHEME_VV) thomas_vv(s, ixy, buf);
            else thomas_ab3(s, ixy, buf);
        }
        free(buf);
    }
    t0 = omp_get_wtime();
    s->tmr[OR_TMR_THOMAS] += t0 - t1;
    fft2d_planes(s, 1);
    double inv_NN = 1. / NXY;
    for (int in = 0; in < N3D; in++) XX[in] = creal(s->spec[in]) * inv_NN;
    t1 = omp_get_wtime();
    s->tmr[OR_TMR_FFT] += t1 - t0;
    or_normalize_pressure(s, XX);
    s->tmr[OR_TMR_NORMALIZE] += omp_get_wtime() - t1;
}

void or_timers(or_solver* s, double* out, int reset) {
    for (int k = 0; k < OR_NTIMERS; k++) {
        if (out) out[k] = s->tmr[k];
        if (reset) s->tmr[k] = 0.0;
    }
}

/* ------------------------------------------------------------------------- */
/* Public field operators                                                     */
/* ------------------------------------------------------------------------- */
void or_flux_term(or_solver* s, const double* u, int do_conv, int do_visc, double* kuhat) {
    const int N3D = s->N3D, NXY = s->NXY;
    if (IS_COLL(s->P.scheme)) {
#pragma omp parallel for
        for (int in = 0; in < N3D; in++) flux_ab3(s, in, u, do_conv, do_visc, kuhat);
        return;
    }
    /* CalcFluxTerm_FD (staggered/func_global.h:632-640) */
#pragma omp parallel for
    for (int in = 0; in < N3D; in++) {
        double c[3] = {0, 0, 0}, v[3] = {0, 0, 0}, r[3] = {0, 0, 0};
        if (do_conv) {
            if (s->P.scheme == OR_SCHEME_HW) conv_flux_hw(s, in, u, c);
            else conv_flux_vv(s, in, u, c);
            for (int k = 0; k < 3; k++) r[k] += c[k];
        }
        if (do_visc) {
            if (s->P.visc_type == 1) visc_flux_naive(s, in, u, v);
            else visc_flux_cross(s, in, u, v);
            for (int k = 0; k < 3; k++) r[k] += v[k];
        }
        if (in < NXY) r[2] = 0;
        for (int k = 0; k < 3; k++) kuhat[3 * in + k] = r[k];
    }
}

void or_explicit_term(or_solver* s, const double* u, double* kuhat) {
    const int N3D = s->N3D, NXY = s->NXY;
    if (IS_COLL(s->P.scheme)) {
        or_flux_term(s, u, 1, 1, kuhat);
        return;
    }
    /* ExplicitTerm (staggered/func_global.h:642-649) */
#pragma omp parallel for
    for (int in = 0; in < N3D; in++) {
        double c[3], v[3];
        if (s->P.scheme == OR_SCHEME_HW) conv_flux_hw(s, in, u, c);
        else conv_flux_vv(s, in, u, c);
        if (s->P.visc_type == 1) visc_flux_naive(s, in, u, v);
        else visc_flux_cross(s, in, u, v);
        kuhat[3 * in + 0] = (s->P.sour_dpdx + c[0]) + v[0];
        kuhat[3 * in + 1] = (0.0 + c[1]) + v[1];
        kuhat[3 * in + 2] = (0.0 + c[2]) + v[2];
        if (in < NXY) kuhat[3 * in + 2] = 0;
    }
}

void or_visc_flux_naive(or_solver* s, const double* u, double* f) {
#pragma omp parallel for
    for (int in = 0; in < s->N3D; in++) visc_flux_naive(s, in, u, f + 3 * in);
}

void or_apply_gradient(or_solver* s, const double* p, double* g) {
#pragma omp parallel for
    for (int in = 0; in < s->N3D; in++) {
        if (IS_COLL(s->P.scheme)) grad_ab3(s, in, p, g);
        else grad_staggered(s, in, p, g);
    }
}

void or_apply_divergence(or_solver* s, const double* v, double* dv) {
#pragma omp parallel for
    for (int in = 0; in < s->N3D; in++) {
        if (IS_COLL(s->P.scheme)) div_ab3(s, in, v, dv);
        else div_staggered(s, in, v, dv);
    }
}

/* UnapplyGradient (staggered/func_global.h:290-293). */
void or_unapply_gradient(or_solver* s, const double* v, double* p) {
    double t0 = omp_get_wtime();
    or_apply_divergence(s, v, s->divv);
    s->tmr[OR_TMR_DIVERGENCE] += omp_get_wtime() - t0;
    or_solve_pressure(s, p, s->divv);
}

void or_recalc_turb_visc(or_solver* s, const double* u) {
    if (!s->nu_t) return;
#pragma omp parallel for
    for (int in = 0; in < s->N3D; in++) {
        if (IS_COLL(s->P.scheme)) turb_visc_ab3(s, in, u);
        else turb_visc_staggered(s, in, u);
    }
}

void or_calc_abs_s(or_solver* s, const double* u, double* absS) {
#pragma omp parallel for
    for (int in = 0; in < s->N3D; in++) absS[in] = calc_abs_s(s, in, u);
}

/* CalcIntegral (staggered/func_global.h:685-690; collocated_AB3/func_global.h:178-181). */
void or_integral(or_solver* s, const double* u, double out[3]) {
    double a = 0, b = 0, c = 0;
    const int NXY = s->NXY;
    const double hx = s->P.hx, hy = s->P.hy;
#pragma omp parallel for reduction(+ : a, b, c)
    for (int i = 0; i < s->N3D; i++) {
        if (IS_COLL(s->P.scheme)) {
            double vol = cell_volume(s, i);
            a += u[3 * i] * vol; b += u[3 * i + 1] * vol; c += u[3 * i + 2] * vol;
        } else {
            int iz = i / NXY;
            a += u[3 * i] * s->DZ[iz] * hx * hy;
            b += u[3 * i + 1] * s->DZ[iz] * hx * hy;
            c += (i < NXY ? 0.0 : u[3 * i + 2] * 0.5 * (s->DZ[iz] + s->DZ[iz - 1])) * hx * hy;
        }
    }
    out[0] = a; out[1] = b; out[2] = c;
}

/* CalcKineticEnergy (staggered/func_global.h:678-683; collocated_AB3/func_global.h:173-176). */
double or_kinetic_energy(or_solver* s, const double* u) {
    double e = 0;
    const int NXY = s->NXY;
#pragma omp parallel for reduction(+ : e)
    for (int i = 0; i < s->N3D; i++) {
        if (IS_COLL(s->P.scheme)) {
            e += (u[3 * i] * u[3 * i] + u[3 * i + 1] * u[3 * i + 1] + u[3 * i + 2] * u[3 * i + 2]) * cell_volume(s, i);
        } else {
            int iz = i / NXY;
            e += ((u[3 * i] * u[3 * i] + u[3 * i + 1] * u[3 * i + 1]) * s->DZ[iz] +
                  u[3 * i + 2] * u[3 * i + 2] * (i < NXY ? 0.0 : 0.5 * (s->DZ[iz] + s->DZ[iz - 1]))) * s->P.hx * s->P.hy;
        }
    }
    return 0.5 * e;
}

/* CalcProfiles, CPU branch (staggered/func_global.h:784-807; collocated_AB3/func_global.h:272-288). */
int or_profiles(or_solver* s, const double* vel, const double* p, double* av) {
    const int NZ = s->NZ, NXY = s->NXY, ab3 = IS_COLL(s->P.scheme);
    const int W = ab3 ? 10 : 12;
#pragma omp parallel for
    for (int iz = 0; iz < NZ; iz++) {
        double sm[12] = {0};
        for (int in = iz * NXY; in < (iz + 1) * NXY; in++) {
            double u = vel[3 * in], v = vel[3 * in + 1];
            double w = ab3 ? vel[3 * in + 2] : 0.5 * (vel[3 * in + 2] + (iz == NZ - 1 ? 0.0 : vel[3 * (in + NXY) + 2]));
            sm[0] += u; sm[1] += v; sm[2] += w;
            sm[3] += u * u; sm[4] += v * v; sm[5] += w * w;
            sm[6] += u * v; sm[7] += u * w; sm[8] += v * w;
            if (s->nu_t) sm[9] += s->nu_t[in];
            if (!ab3) { sm[10] += p[in]; sm[11] += p[in] * p[in]; }
        }
        for (int k = 0; k < W; k++) av[iz * W + k] = sm[k] * (1.0 / NXY);
    }
    return W;
}

/* CalcCFL (staggered/func_global.h:692-710; collocated_AB3/func_global.h:188-205). */
double or_cfl(or_solver* s, const double* u, double tau, int mode) {
    double cfl_max = 0;
    const int NXY = s->NXY;
    const double inv_hx = s->inv_hx, inv_hy = s->inv_hy;
#pragma omp parallel for reduction(max : cfl_max)
    for (int i = 0; i < s->N3D; i++) {
        double inv_hz, visc_coeff, fac;
        if (IS_COLL(s->P.scheme)) {
            inv_hz = 1 / cell_height(s, i);
            visc_coeff = s->P.visc + nut_at_node(s, i);
            fac = 4 * (s->P.visc_type + 1);
        } else {
            inv_hz = 1 / s->DZ[i / NXY];
            visc_coeff = s->nu_t ? s->P.visc + s->nu_t[i] : s->P.visc;
            fac = 4 * ((s->P.visc_type != 0) + 1);
        }
        double CFL = 0;
        if (mode & 1) CFL += tau * (fabs(u[3 * i]) * inv_hx + fabs(u[3 * i + 1]) * inv_hy + fabs(u[3 * i + 2]) * inv_hz);
        if (mode & 2) CFL += fac * tau * visc_coeff * (inv_hx * inv_hx + inv_hy * inv_hy + inv_hz * inv_hz);
        if (CFL > cfl_max) cfl_max = CFL;
    }
    return cfl_max;
}

/* ------------------------------------------------------------------------- */
/* Time step                                                                  */
/* ------------------------------------------------------------------------- */
int or_set_field(or_solver* s, int kind, int slot, const double* src) {
    size_t n3 = (size_t)s->N3D;
    if (slot < 0 || slot > 3) return 1;
    if (kind == OR_KIND_VEL) memcpy(s->vel[slot], src, 3 * n3 * sizeof(double));
    else if (kind == OR_KIND_FLUX) memcpy(s->flux[slot], src, 3 * n3 * sizeof(double));
    else if (kind == OR_KIND_PRES) memcpy(s->pres[slot], src, n3 * sizeof(double));
    else if (kind == OR_KIND_NUT && s->nu_t) memcpy(s->nu_t, src, n3 * sizeof(double));
    else if (kind == OR_KIND_TAU && s->tau) memcpy(s->tau, src, 6 * n3 * sizeof(double));
    else return 1;
    return 0;
}

int or_get_field(or_solver* s, int kind, int slot, double* dst) {
    size_t n3 = (size_t)s->N3D;
    if (slot < 0 || slot > 3) return 1;
    if (kind == OR_KIND_VEL) memcpy(dst, s->vel[slot], 3 * n3 * sizeof(double));
    else if (kind == OR_KIND_FLUX) memcpy(dst, s->flux[slot], 3 * n3 * sizeof(double));
    else if (kind == OR_KIND_PRES) memcpy(dst, s->pres[slot], n3 * sizeof(double));
    else if (kind == OR_KIND_NUT && s->nu_t) memcpy(dst, s->nu_t, n3 * sizeof(double));
    else if (kind == OR_KIND_TAU && s->tau) memcpy(dst, s->tau, 6 * n3 * sizeof(double));
    else return 1;
    return 0;
}

/* Constant-flow-rate correction (new numerics, no reference counterpart):
 * shift the streamwise velocity of every u degree of freedom by one constant so
 * that the bulk velocity equals cfr_ubulk. A uniform shift has zero discrete
 * divergence in the periodic direction, so the projection is not disturbed. */
double or_cfr_correct(or_solver* s, double* u) {
    double q[3];
    or_integral(s, u, q);
    double vol_u = 0;
    if (IS_COLL(s->P.scheme)) {
        for (int iz = 1; iz < s->NZ - 1; iz++) vol_u += 0.5 * (s->DZ[iz - 1] + s->DZ[iz]);
        vol_u *= s->P.hx * s->P.hy * s->NXY;
    } else {
        vol_u = s->area;
    }
    double shift = (s->P.cfr_ubulk * s->area - q[0]) / vol_u;
#pragma omp parallel for
    for (int i = 0; i < s->N3D; i++) {
        if (IS_COLL(s->P.scheme) && is_wall(s, i)) continue;
        u[3 * i] += shift;
    }
    return shift;
}

/* Step (staggered/func_global.h:897-942; collocated_AB3/func_global.h:387-450). */
int or_step(or_solver* s, double t_start, double tau, int cur, int bdf_order) {
    (void)t_start;
    if (bdf_order < 1 || bdf_order > 3) { set_err("Wrong bdf_order"); return 1; }
    if (s->P.scheme == OR_SCHEME_SRK) { set_err("the SRK solver steps with or_srk_step"); return 1; }
    const int N3D = s->N3D;
    int next = (cur + 1) % 4, prev = (cur + 3) % 4, prev2 = (cur + 2) % 4;
    const double inv_tau = 1. / tau;
    const double bd0 = (bdf_order == 3) ? (1.5 + C1_3) : (bdf_order == 2) ? 1.5 : 1.0;
    const double tau_over_bd0 = tau / bd0;
    /* timer channels: the reference's explicit / grad / pressure F, T, N */
    double t0 = omp_get_wtime(), t1;
#define OR_LAP(ch) (t1 = omp_get_wtime(), s->tmr[ch] += t1 - t0, t0 = t1)
    if (s->nu_t) or_recalc_turb_visc(s, s->vel[cur]);
    or_explicit_term(s, s->vel[cur], s->flux[cur]);
    OR_LAP(OR_TMR_EXPLICIT);
    double *Fn = s->flux[next], *Fc = s->flux[cur], *Fp = s->flux[prev], *Fp2 = s->flux[prev2];
    double *Vc = s->vel[cur], *Vp = s->vel[prev], *Vp2 = s->vel[prev2], *Vn = s->vel[next];
#pragma omp parallel for
    for (int k = 0; k < 3 * N3D; k++) {
        if (bdf_order == 1) Fn[k] = Fc[k] + inv_tau * Vc[k];
        else if (bdf_order == 2) Fn[k] = 2. * Fc[k] - Fp[k] + inv_tau * (2 * Vc[k] - 0.5 * Vp[k]);
        else Fn[k] = 3 * Fc[k] - 3 * Fp[k] + Fp2[k] + inv_tau * (3 * Vc[k] - 1.5 * Vp[k] + C1_3 * Vp2[k]);
    }
    OR_LAP(OR_TMR_OTHER);
    if (IS_COLL(s->P.scheme) && s->P.stab_type != 0) {
        or_apply_gradient(s, s->pres[cur], Vn);
#pragma omp parallel for
        for (int k = 0; k < 3 * N3D; k++) Fn[k] -= Vn[k];
        OR_LAP(OR_TMR_GRADIENT);
        or_unapply_gradient(s, Fn, s->pres[next]); /* divergence and pressure channels inside */
        t0 = omp_get_wtime();
        or_apply_gradient(s, s->pres[next], Vn);
#pragma omp parallel for
        for (int i = 0; i < N3D; i++) s->pres[next][i] += s->pres[cur][i];
        OR_LAP(OR_TMR_GRADIENT);
    } else {
        or_unapply_gradient(s, Fn, s->pres[next]);
        t0 = omp_get_wtime();
        or_apply_gradient(s, s->pres[next], Vn);
        OR_LAP(OR_TMR_GRADIENT);
    }
#pragma omp parallel for
    for (int k = 0; k < 3 * N3D; k++) Vn[k] = (Fn[k] - Vn[k]) * tau_over_bd0;
    if (s->P.cfr_ubulk > 0) or_cfr_correct(s, Vn);
    OR_LAP(OR_TMR_OTHER);
#undef OR_LAP
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Collocated SRK: IMEX Runge-Kutta stepper (collocated_SRK/)                 */
/* ------------------------------------------------------------------------- */
#define OR_MAX_STAGES 6

/* Butcher tables, restated from collocated_SRK/imex.cpp:48-239 (published
 * coefficients: Ascher-Ruuth-Spiteri 1997, Kennedy-Carpenter 2003,
 * Boscarino-Russo 2007, Boscarino 2009). Row-major, (s+1) rows of s entries;
 * the last row holds the weights. Returns s, 0 for an unknown method. */
static int imex_rows(int method, int* type, double* BE, double* BI) {
    const double gL = 0.435866521508458999416019451193556842529; /* middle root of 6x^3-18x^2+9x-1 */
    memset(BE, 0, sizeof(double) * 42);
    memset(BI, 0, sizeof(double) * 42);
#define ROW(T, r, ...) do { const double v_[] = {__VA_ARGS__}; memcpy(T + (r) * S, v_, sizeof v_); } while (0)
    int S = 0;
    switch (method) {
    case OR_IMEX_ARS_121: /* :48-62 */
        S = 2; *type = OR_IMEX_TYPE_ARS;
        ROW(BE, 1, 1., 0.); ROW(BE, 2, 0., 1.);
        ROW(BI, 1, 0., 1.); ROW(BI, 2, 0., 1.);
        break;
    case OR_IMEX_ARS_232: { /* :65-83 */
        S = 3; *type = OR_IMEX_TYPE_ARS;
        const double g = 1. - 0.5 * sqrt(2.), dl = -2. * sqrt(2.) / 3.;
        ROW(BE, 1, g, 0., 0.); ROW(BE, 2, dl, 1. - dl, 0.); ROW(BE, 3, 0., 1. - g, g);
        ROW(BI, 1, 0., g, 0.); ROW(BI, 2, 0., 1. - g, g); ROW(BI, 3, 0., 1. - g, g);
        break;
    }
    case OR_IMEX_ARS_343: { /* :86-111 */
        S = 4; *type = OR_IMEX_TYPE_ARS;
        const double g = gL;
        const double b1 = -1.5 * g * g + 4 * g - 0.25, b2 = 1.5 * g * g - 5. * g + 1.25;
        const double a42 = 0.5529291479, a43 = 0.5529291479;
        const double c1 = 1. - 4.5 * g + 1.5 * g * g, c2 = 2.75 - 10.5 * g + 3.75 * g * g;
        ROW(BI, 1, 0., g, 0., 0.); ROW(BI, 2, 0., (1. - g) / 2., g, 0.);
        ROW(BI, 3, 0., b1, b2, g); ROW(BI, 4, 0., b1, b2, g);
        ROW(BE, 1, g, 0., 0., 0.);
        ROW(BE, 2, c1 * a42 + c2 * a43 - 3.5 + 13. * g - 4.5 * g * g,
            -c1 * a42 - c2 * a43 + 4. - 12.5 * g + 4.5 * g * g, 0., 0.);
        ROW(BE, 3, 1. - a42 - a43, a42, a43, 0.);
        ROW(BE, 4, 0., b1, b2, g);
        break;
    }
    case OR_IMEX_MARS_343: { /* :115-135 */
        S = 4; *type = OR_IMEX_TYPE_ARS;
        const double g = gL, b2 = 1.20849664917601;
        ROW(BI, 1, 0., g, 0., 0.); ROW(BI, 2, 0., 0.28206673924577, g, 0.);
        ROW(BI, 3, 0., b2, 1. - g - b2, g); ROW(BI, 4, 0., b2, 1. - g - b2, g);
        ROW(BE, 1, g, 0., 0., 0.); ROW(BE, 2, 0.535396540307354, 0.182536720446875, 0., 0.);
        ROW(BE, 3, 0.63041255815287, -0.83193390106308, 1.20152134291021, 0.);
        ROW(BE, 4, 0., b2, 1. - g - b2, g);
        break;
    }
    case OR_IMEX_ARK3: { /* :138-160 */
        S = 4; *type = OR_IMEX_TYPE_CK;
        const double g = gL;
        const double b1 = 1471266399579. / 7840856788654., b2 = -4482444167858. / 7529755066697.;
        const double b3 = 1. - g - b1 - b2;
        ROW(BI, 1, g, g, 0., 0.);
        ROW(BI, 2, 2746238789719. / 10658868560708., -640167445237. / 6845629431997., g, 0.);
        ROW(BI, 3, b1, b2, b3, g); ROW(BI, 4, b1, b2, b3, g);
        ROW(BE, 1, 2 * g, 0., 0., 0.);
        ROW(BE, 2, 5535828885825. / 10492691773637., 788022342437. / 10882634858940., 0., 0.);
        ROW(BE, 3, 6485989280629. / 16251701735622., -4246266847089. / 9704473918619.,
            10755448449292. / 10357097424841., 0.);
        ROW(BE, 4, b1, b2, b3, g);
        break;
    }
    case OR_IMEX_ARK4: { /* :163-186 */
        S = 6; *type = OR_IMEX_TYPE_CK;
        const double b0 = 82889. / 524892., b2 = 15625. / 83664., b3 = 69875. / 102672., b4 = -2260. / 8211., b5 = 0.25;
        ROW(BI, 1, 0.25, 0.25, 0., 0., 0., 0.);
        ROW(BI, 2, 8611. / 62500., -1743. / 31250., 0.25, 0., 0., 0.);
        ROW(BI, 3, 5012029. / 34652500., -654441. / 2922500., 174375. / 388108., 0.25, 0., 0.);
        ROW(BI, 4, 15267082809. / 155376265600., -71443401. / 120774400., 730878875. / 902184768.,
            2285395. / 8070912., 0.25, 0.);
        ROW(BI, 5, b0, 0., b2, b3, b4, b5);
        ROW(BI, 6, b0, 0., b2, b3, b4, b5);
        ROW(BE, 1, 0.5, 0., 0., 0., 0., 0.);
        ROW(BE, 2, 13861. / 62500., 6889. / 62500., 0., 0., 0., 0.);
        ROW(BE, 3, -116923316275. / 2393684061468., -2731218467317. / 15368042101831.,
            9408046702089. / 11113171139209., 0., 0., 0.);
        ROW(BE, 4, -451086348788. / 2902428689909., -2682348792572. / 7519795681897.,
            12662868775082. / 11960479115383., 3355817975965. / 11060851509271., 0., 0.);
        ROW(BE, 5, 647845179188. / 3216320057751., 73281519250. / 8382639484533., 552539513391. / 3454668386233.,
            3354512671639. / 8306763924573., 4040. / 17871., 0.);
        ROW(BE, 6, b0, 0., b2, b3, b4, b5);
        break;
    }
    case OR_IMEX_MARK3: { /* :189-209 */
        S = 4; *type = OR_IMEX_TYPE_CK;
        const double g = gL, b1 = 0.60424832458800;
        ROW(BI, 1, g, g, 0., 0.); ROW(BI, 2, -4.30002662176923, 2.26541338346372, g, 0.);
        ROW(BI, 3, b1, 0., 1. - g - b1, g); ROW(BI, 4, b1, 0., 1. - g - b1, g);
        ROW(BE, 1, 2 * g, 0., 0., 0.); ROW(BE, 2, -3.06478674186224, 1.46604002506519, 0., 0.);
        ROW(BE, 3, 0.21444560762133, 0.71075364965269, 0.07480074272597, 0.);
        ROW(BE, 4, b1, 0., 1. - g - b1, g);
        break;
    }
    case OR_IMEX_BHR_553: { /* :213-239 */
        S = 5; *type = OR_IMEX_TYPE_CK;
        const double g = 424782. / 974569.;
        const double i41 = 487698502336740678603511. / 1181159636928185920260208.;
        const double i43 = 302987763081184622639300143137943089. / 1535359944203293318639180129368156500.;
        const double i44 = -105235928335100616072938218863. / 2282554452064661756575727198000.;
        ROW(BI, 1, g, g, 0., 0., 0.);
        ROW(BI, 2, g, -31733082319927313. / 455705377221960889379854647102., g, 0., 0.);
        ROW(BI, 3, -3012378541084922027361996761794919360516301377809610. /
                      45123394056585269977907753045030512597955897345819349.,
            -62865589297807153294268. / 102559673441610672305587327019095047.,
            418769796920855299603146267001414900945214277000. / 212454360385257708555954598099874818603217167139.,
            g, 0.);
        ROW(BI, 4, i41, 0., i43, i44, g);
        ROW(BI, 5, i41, 0., i43, i44, g);
        ROW(BE, 1, 2. * g, 0., 0., 0., 0.);
        ROW(BE, 2, g, g, 0., 0., 0.);
        ROW(BE, 3, -475883375220285986033264. / 594112726933437845704163., 0.,
            1866233449822026827708736. / 594112726933437845704163., 0., 0.);
        ROW(BE, 4, 62828845818073169585635881686091391737610308247. / 176112910684412105319781630311686343715753056000.,
            -i43, 262315887293043739337088563996093207. / 297427554730376353252081786906492000.,
            -987618231894176581438124717087. / 23877337660202969319526901856000., 0.);
        ROW(BE, 5, i41, 0., i43, i44, g);
        break;
    }
    default:
        return 0;
    }
#undef ROW
    return S;
}

int or_imex_table(int method, int* type, double* be, double* bi, double* d, double* alpha_tau_max) {
    double BE[42], BI[42];
    int ty = 0;
    const int S = imex_rows(method, &ty, BE, BI);
    if (!S) return 0;
    if (type) *type = ty;
    if (be) memcpy(be, BE, sizeof(double) * (size_t)(S * (S + 1)));
    if (bi) memcpy(bi, BI, sizeof(double) * (size_t)(S * (S + 1)));
    if (d) { /* CalcD (collocated_SRK/imex.cpp:34-44) */
        const double ass = BI[S * S - 1];
        d[0] = 1.;
        for (int j = 1; j < S; j++) {
            d[j] = 0.;
            for (int k = 0; k < j; k++) d[j] -= BI[j * S + k] * d[k] / ass;
        }
    }
    if (alpha_tau_max) alpha_tau_max[0] = alpha_tau_max[1] = 2.;
    return S;
}

/* data::Step (collocated_SRK/func_global.h:317-408) with ImplicitTerm = 0 and
 * ImplicitStage = copy (:310-316). Ku is kept, as the reference keeps it. */
int or_srk_step(or_solver* s, int method, int stab_type, double alpha_tau, double t_start, double tau,
                double* velocity, double* pressure, double* dpdt) {
    (void)t_start;
    if (s->P.scheme != OR_SCHEME_SRK) { set_err("or_srk_step needs scheme OR_SCHEME_SRK"); return 1; }
    double BE[42], BI[42], d[OR_MAX_STAGES];
    int type = 0;
    const int S = or_imex_table(method, &type, BE, BI, d, NULL);
    if (!S) { set_err("unknown IMEX method"); return 1; }
    if (stab_type && !dpdt) { set_err("stab_type 1 needs dpdt"); return 1; }
    const int N = s->N3D, N3 = 3 * N;
    const int ck = type == OR_IMEX_TYPE_CK;
    const double ass = BI[S * S - 1], inv_ass = 1. / ass;
    const double tau_stage = tau * ass, inv_tau_stage = 1. / tau_stage;
    const double alpha = alpha_tau / tau;
    double *Ku[OR_MAX_STAGES], *Kuhat[OR_MAX_STAGES], *mujt[OR_MAX_STAGES], *qj[OR_MAX_STAGES];
    for (int i = 0; i < S; i++) {
        Ku[i] = (double*)calloc((size_t)N3, sizeof(double));
        Kuhat[i] = (double*)calloc((size_t)N3, sizeof(double));
        mujt[i] = (double*)calloc((size_t)N, sizeof(double));
        qj[i] = (double*)calloc((size_t)N, sizeof(double));
    }
    double* nu1 = (double*)calloc((size_t)N3, sizeof(double));
    double* bv = (double*)calloc((size_t)N3, sizeof(double));
    double* gp = (double*)calloc((size_t)N3, sizeof(double));
    double* u = (double*)calloc((size_t)N3, sizeof(double));
    double* bs = (double*)calloc((size_t)N, sizeof(double));
    double* pj = (double*)calloc((size_t)N, sizeof(double));
    double* muj = (double*)calloc((size_t)N, sizeof(double));
    double* pjhat = (double*)calloc((size_t)N, sizeof(double));

    if (type == OR_IMEX_TYPE_CK || type == OR_IMEX_TYPE_ARS) {
        or_explicit_term(s, velocity, Kuhat[0]);
        or_apply_gradient(s, pressure, gp);
        for (int k = 0; k < N3; k++) { Ku[0][k] = 0.; Kuhat[0][k] -= gp[k]; }
    }
    if (ck)
        for (int k = 0; k < N3; k++) nu1[k] = Ku[0][k] + Kuhat[0][k] + alpha * velocity[k];
    const int j0 = type == OR_IMEX_TYPE_A ? 0 : 1;
    for (int i = j0; i < S; i++) {
        for (int k = 0; k < N3; k++) {
            double b = velocity[k];
            for (int j = 0; j < i; j++) b += tau * BI[i * S + j] * Ku[j][k] + tau * BE[i * S + j] * Kuhat[j][k];
            bv[k] = b;
        }
        memcpy(u, bv, sizeof(double) * (size_t)N3);                             /* ImplicitStage */
        for (int k = 0; k < N3; k++) Ku[i][k] = (u[k] - bv[k]) * inv_tau_stage;
        or_explicit_term(s, u, Kuhat[i]);
        for (int n = 0; n < N; n++) {
            if (stab_type == 0) {
                pjhat[n] = 0.;
                muj[n] = pressure[n];
                if (ck) muj[n] *= 1. - alpha_tau * BI[i * S];
                for (int j = j0; j < i; j++) muj[n] += inv_ass * BI[i * S + j] * mujt[j][n];
                bs[n] = muj[n] - alpha * tau_stage * pressure[n];
            } else {
                muj[n] = tau_stage * dpdt[n];
                if (ck) muj[n] *= 1. - alpha_tau * BI[i * S];
                pjhat[n] = pressure[n];
                if (ck) pjhat[n] += tau * BI[i * S] * dpdt[n];
                for (int j = j0; j < i; j++) {
                    muj[n] += inv_ass * BI[i * S + j] * mujt[j][n];
                    pjhat[n] += tau * BI[i * S + j] * qj[j][n];
                }
                bs[n] = muj[n] + pjhat[n] - alpha * tau_stage * tau_stage * dpdt[n];
            }
        }
        or_apply_gradient(s, bs, bv);
        for (int k = 0; k < N3; k++) {
            bv[k] = Ku[i][k] + Kuhat[i][k] - bv[k] + alpha * velocity[k];
            if (ck) bv[k] -= d[i] * nu1[k];
        }
        or_unapply_gradient(s, bv, mujt[i]);
        for (int n = 0; n < N; n++) {
            pj[n] = mujt[i][n] + bs[n];
            mujt[i][n] -= alpha * tau_stage * (stab_type ? tau_stage * dpdt[n] : pressure[n]);
            if (stab_type) qj[i][n] = inv_tau_stage * (pj[n] - pjhat[n]);
        }
        or_apply_gradient(s, pj, bv);
        for (int k = 0; k < N3; k++) Kuhat[i][k] -= bv[k];
    }
    for (int k = 0; k < N3; k++)
        for (int j = 0; j < S; j++) velocity[k] += tau * BI[S * S + j] * Ku[j][k] + tau * BE[S * S + j] * Kuhat[j][k];
    for (int n = 0; n < N; n++) {
        pressure[n] = pj[n];
        if (stab_type) dpdt[n] = qj[S - 1][n];
    }
    for (int i = 0; i < S; i++) { free(Ku[i]); free(Kuhat[i]); free(mujt[i]); free(qj[i]); }
    free(nu1); free(bv); free(gp); free(u); free(bs); free(pj); free(muj); free(pjhat);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* New numerics: tanh mesh and the seeded initial field                       */
/* ------------------------------------------------------------------------- */

/* Faces z_k = 1 + tanh(gamma (2k/n - 1)) / tanh(gamma), k = 0..n, on [0, 2]. */
void or_tanh_mesh(int n, double gamma, double* zz) {
    for (int k = 0; k <= n; k++) zz[k] = 1.0 + tanh(gamma * (2.0 * k / n - 1.0)) / tanh(gamma);
    zz[0] = 0.0;
    zz[n] = 2.0;
}

static uint64_t splitmix64(uint64_t* st) {
    uint64_t z = (*st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double gauss(uint64_t* st) {
    double u1 = 1.0 - (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
    double u2 = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * PI_D * u2);
}

/* Seeded initial field: laminar parabola with bulk velocity ubulk plus the
 * discrete curl of a smoothed Gaussian vector potential, scaled to noise_rms
 * (root mean square over the three components). Not projected. */
void or_initial_field(or_solver* s, unsigned long long seed, double ubulk, double noise_rms, double* u) {
    const int NX = s->NX, NY = s->NY, NZ = s->NZ, NXY = s->NXY, ab3 = IS_COLL(s->P.scheme);
    const int NP = ab3 ? NZ : NZ + 1; /* potential planes: faces (staggered) or nodes (AB3) */
    const double hx = s->P.hx, hy = s->P.hy;
    const double z0 = s->ZZ[0], z1 = ab3 ? s->ZZ[NZ - 1] : s->ZZ[NZ];
    const double zc = 0.5 * (z0 + z1), H = 0.5 * (z1 - z0);
    size_t np = (size_t)NXY * NP;
    double* psi = (double*)malloc(sizeof(double) * 3 * np);
    double* tmp = (double*)malloc(sizeof(double) * np);
    uint64_t st = (uint64_t)seed;
    for (int c = 0; c < 3; c++)
        for (size_t i = 0; i < np; i++) psi[c * np + i] = gauss(&st);
    for (int c = 0; c < 3; c++) {
        double* a = psi + c * np;
        for (int pass = 0; pass < 2; pass++) {
            for (size_t i = 0; i < np; i++) {
                int ix = (int)(i % NX);
                size_t row = i - ix;
                tmp[i] = 0.25 * a[row + (ix + NX - 1) % NX] + 0.5 * a[i] + 0.25 * a[row + (ix + 1) % NX];
            }
            for (size_t i = 0; i < np; i++) {
                int iy = (int)((i / NX) % NY);
                size_t base = i - (size_t)iy * NX;
                a[i] = 0.25 * tmp[base + (size_t)((iy + NY - 1) % NY) * NX] + 0.5 * tmp[i] + 0.25 * tmp[base + (size_t)((iy + 1) % NY) * NX];
            }
        }
        for (int k = 0; k < NP; k++) {
            double z;
            if (ab3 || c < 2) z = s->ZZ[k];
            else z = k < NZ ? 0.5 * (s->ZZ[k] + s->ZZ[k + 1]) : z1;
            double zeta = (z - zc) / H;
            double env = (1 - zeta * zeta) * (1 - zeta * zeta);
            if (!ab3 && c == 2 && k == NZ) env = 0;
            for (int j = 0; j < NXY; j++) a[(size_t)k * NXY + j] *= env;
        }
    }
    double *px = psi, *py = psi + np, *pz = psi + 2 * np;
#define P(a, i, j, k) a[(size_t)(k) * NXY + (size_t)(((j) + NY) % NY) * NX + (size_t)(((i) + NX) % NX)]
    double ss = 0;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                size_t in = (size_t)k * NXY + (size_t)j * NX + i;
                double a, b, c;
                if (!ab3) {
                    double dz = s->DZ[k];
                    a = (P(pz, i, j + 1, k) - P(pz, i, j, k)) / hy - (P(py, i, j, k + 1) - P(py, i, j, k)) / dz;
                    b = (P(px, i, j, k + 1) - P(px, i, j, k)) / dz - (P(pz, i + 1, j, k) - P(pz, i, j, k)) / hx;
                    c = (P(py, i + 1, j, k) - P(py, i, j, k)) / hx - (P(px, i, j + 1, k) - P(px, i, j, k)) / hy;
                    if (k == 0) c = 0;
                } else if (k == 0 || k == NZ - 1) {
                    a = b = c = 0;
                } else {
                    double dz2 = s->ZZ[k + 1] - s->ZZ[k - 1];
                    a = (P(pz, i, j + 1, k) - P(pz, i, j - 1, k)) / (2 * hy) - (P(py, i, j, k + 1) - P(py, i, j, k - 1)) / dz2;
                    b = (P(px, i, j, k + 1) - P(px, i, j, k - 1)) / dz2 - (P(pz, i + 1, j, k) - P(pz, i - 1, j, k)) / (2 * hx);
                    c = (P(py, i + 1, j, k) - P(py, i - 1, j, k)) / (2 * hx) - (P(px, i, j + 1, k) - P(px, i, j - 1, k)) / (2 * hy);
                }
                u[3 * in] = a; u[3 * in + 1] = b; u[3 * in + 2] = c;
                ss += a * a + b * b + c * c;
            }
#undef P
    double scale = ss > 0 ? noise_rms / sqrt(ss / (3.0 * s->N3D)) : 0.0;
    for (int k = 0; k < NZ; k++) {
        double z = ab3 ? s->ZZ[k] : 0.5 * (s->ZZ[k] + s->ZZ[k + 1]);
        double zeta = (z - zc) / H;
        double ulam = 1.5 * ubulk * (1 - zeta * zeta);
        if (ab3 && (k == 0 || k == NZ - 1)) ulam = 0;
        for (int j = 0; j < NXY; j++) {
            size_t in = (size_t)k * NXY + j;
            u[3 * in] = ulam + scale * u[3 * in];
            u[3 * in + 1] *= scale;
            u[3 * in + 2] *= scale;
        }
    }
    free(psi);
    free(tmp);
}
