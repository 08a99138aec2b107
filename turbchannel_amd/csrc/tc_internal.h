// Internal declarations shared by the host set-up, the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/turbchannel.h"

namespace tc {

// ---------------------------------------------------------------------------
// Host-side mesh data (FP64), derived from tc_params.
// ---------------------------------------------------------------------------
struct HostMesh {
    int scheme = 0, nx = 0, ny = 0, nz = 0;
    std::vector<double> ZZ, DZ, DZW, GC1, GC2, GCZ, MAT_A, KERG;
    double MAT_D[6] = {0, 0, 0, 0, 0, 0};
    double KERG_norm2 = 0;
    double area = 0;   // box volume (the reference's `area`)
    double vol_u = 0;  // volume carried by streamwise-velocity unknowns (flow-rate correction)
    std::vector<double> cellh;  // collocated cell heights per node plane
};

// Returns "" on success, an error text otherwise.
std::string build_mesh(const tc_params& p, HostMesh& m);
int solve_banded(int N, int NB, double* M, double* x, double tiny);

// Precomputed elimination of the wall-normal pressure systems, one column per
// (kx, |ky|): kx in [0, nx/2], ky in [0, ny/2] (the system of ky and ny - ky is
// the same: it depends on ky only through cos(2 pi ky / ny), or sin(pi ky / ny)
// for the fourth-order scheme); zfac_col maps a mode to its column.
// W = bandwidth (1 or 3).
// Forward:  x[i] = (((x[i] + M3 x[i-3]) + M2 x[i-2]) + M1 x[i-1]) * D[i]
// Backward: x[i] -= U1 x[i+1];  x[i] -= U2 x[i+2];  x[i] -= U3 x[i+3]
// Arrays are [coef][iz][mode]. Mode-independent M1 of the tridiagonal solvers is
// stored per mode too, to keep one kernel.
struct ZFactors {
    int W = 1, nz = 0, ncols = 0;
    std::vector<double> M[3], D, U[3];
    // The forward elimination formed in-lane instead (FP32-bulk engines): the
    // mode-independent per-row constants R[iz][zrow_w(W)] and the wavenumber
    // terms KX[kx][zk_w(W)], KY[|ky|][zk_w(W)] (layout below). Not used for the
    // pinned (0,0) mode.
    std::vector<double> R, KX, KY;
    // normalisation of the (0,0) mode in spectral space: X00[iz] -= psum * C[iz],
    // psum = sum_iz Wt[iz] * X00[iz]
    std::vector<double> Wt, C;
};
// In-lane forward elimination, per row iz of R (zrow_w(W) reals):
//   W = 1: AM (entry at iz-1, 0 on row 0), AP (entry at iz+1, 0 on the last
//          row), C0, CQ; KX[kx] + KY[|ky|] = the mode's MM; with U[-1] = 0:
//          D = 1 / ((C0 + CQ MM) - AM U[iz-1]), M1 = -AM, U = AP D.
//   W = 3: LO1, LO2, LO3 (the row's entries at iz-3, iz-2, iz-1), DG (diagonal
//          without the wavenumber term), UP1, UP2, UP3 (entries at iz+1..iz+3),
//          DZ, DZW / 27, EX = 4 / (hx^2 (DZ - DZW / 9)), EY (with hy), 0;
//          KX[kx] = (sin(pi kx / nx), sin(3 pi kx / nx)), KY likewise: the
//          diagonal is DG - EX cx^2 - EY cy^2 with c = DZ s - DZW / 27 s3, and
//          the elimination with the normalised upper band of the last 3 rows
//          follows build_zfactors.
constexpr int zrow_w(int W) { return W == 3 ? 12 : 4; }
constexpr int zk_w(int W) { return W == 3 ? 2 : 1; }
std::string build_zfactors(const tc_params& p, const HostMesh& m, ZFactors& f);

void make_initial_field(const tc_params& p, const HostMesh& m, unsigned long long seed, double ubulk,
                        double noise_rms, double* out);

// IMEX Runge-Kutta method of the SRK stepper: E and I are (s+1) x s row-major
// (row s = weights), d the coefficients of CalcD (collocated_SRK/imex.cpp:34-44).
constexpr int kMaxImexStages = 6;
struct ImexTable {
    int type = 0, s = 0;
    std::vector<double> E, I, d;
    double alpha_tau_max[2] = {2., 2.};
};
bool imex_table(int method, ImexTable& t);

// The collocated SRK solver shares the AB3 mesh, operators and pressure solver;
// only its pressure normalisation differs (zero mean instead of the AB3
// hx*hy-scaled shift). Mesh and factor builders see it as AB3.
inline tc_params as_kernel_params(const tc_params& p) {
    tc_params q = p;
    if (q.scheme == TC_SCHEME_SRK) q.scheme = TC_SCHEME_AB3;
    return q;
}

// ---------------------------------------------------------------------------
// Device-side types
// ---------------------------------------------------------------------------
template <typename T>
struct Cplx {
    T x, y;
};

// Division by a constant d >= 1 for 0 <= n < 2^31: q = (mulhi(n, m) + n) >> s
// with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1 (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994).
// The kernels' index arithmetic uses it in place of integer division.
struct FastDiv {
    unsigned d = 1, m = 1;
    int s = 0;
    FastDiv() = default;
    explicit FastDiv(unsigned dd) : d(dd) {
        while ((1ull << s) < dd) s++;
        m = (unsigned)(((1ull << 32) * ((1ull << s) - dd)) / dd + 1);
    }
    __host__ __device__ __forceinline__ int div(int n) const {
        const unsigned hi = (unsigned)(((unsigned long long)(unsigned)n * m) >> 32);
        return (int)((hi + (unsigned)n) >> s);
    }
};

// Device mesh arrays per engine: ZZ, DZ, DZW, GC1, GC2, GCZ, cellh and the six
// per-plane quotients of Mesh.
constexpr int kMeshArrays = 13;

constexpr int kMaxStages = 12;
// Mixed-radix in-place decimation-in-frequency plan. Stage st has radix
// radix[st] and works on sub-transforms of length L = n / (product of the
// radices before it) with butterfly span S = L / radix[st]; twstride[st] =
// n / L. Divisors: by n / radix[st] (butterflies per line), by S, by the radix;
// dn divides by n.
struct FFTPlan {
    int n = 0, nst = 0;
    int radix[kMaxStages] = {0}, twstride[kMaxStages] = {0};
    FastDiv dnr[kMaxStages], dspan[kMaxStages], drad[kMaxStages], dn;
};
bool make_plan(int n, FFTPlan& pl);

// Constant data handed to kernels by value.
template <typename T>
struct Mesh {
    int nx, ny, nz, nxy, n3d, nxh;
    T hx, hy, inv_hx, inv_hy, visc, sour, inv_len;
    T zlo, zhi;  // wall coordinates
    const T *ZZ, *DZ, *DZW, *GC1, *GC2, *GCZ, *cellh;
    // Per-plane quotients the stencils use, formed on the host in T (the same
    // correctly rounded values the device division gives):
    // IDZ = 1/DZ[iz]; IHDN = 2/(DZ[iz] + DZ[iz-1]) (DZ[-1] = 0);
    // IHUP = 2/(DZ[iz] + DZ[iz+1]) (DZ[nz] = 0); IDZM = 1/DZ[max(iz-1, 0)];
    // IVOL = -1/(32 (DZ[iz] - DZW[iz]/9) hx hy) (fourth-order convection);
    // IHC = 2/(DZ[iz-1] + DZ[iz]) at the interior collocated nodes.
    // A zero denominator gives 0 (those entries are never used).
    const T *IDZ, *IHDN, *IHUP, *IDZM, *IVOL, *IHC;
    // 1: kernels that read neighbouring rows and planes renumber their blocks
    // so that each XCD works on a contiguous range of planes (set when four
    // planes of a real3 field fit one XCD's 4 MiB L2, so the neighbours hit
    // there); 0: plain order, every XCD streaming the same region
    int xcd_swz = 0;
};

}  // namespace tc
