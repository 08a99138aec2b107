// Launchers of the engine's HIP kernels (one translation unit per kernel family).
#pragma once

#include "tc_internal.h"

namespace tc {

// Block b of a launch runs on XCD b % 8. Renumbering b -> xcd_swizzle(b, nb)
// gives each XCD one contiguous range of logical blocks, so blocks that read
// each other's neighbourhoods (adjacent rows and planes) share an L2.
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
    const int q = nb / 8, r = nb % 8, x = b % 8, i = b / 8;
    return x * q + min(x, r) + i;
}

// Explicit term (convection + viscosity + source) fused with the BDF3/EX3
// right-hand side. order == 0: write only kuhat = F(vel) to fcur.
template <typename T>
struct FluxArgs {
    int scheme = 0, visc_type = 0, turb = 0, order = 0, stab = 0;
    const T* vel = nullptr;   // u_n
    const T* nut = nullptr;   // nu_t (turb)
    T* fcur = nullptr;        // F_n (written)
    const T *vprev = nullptr, *vprev2 = nullptr, *fprev = nullptr, *fprev2 = nullptr;
    T* fnext = nullptr;       // right-hand side r (written when order > 0)
    const T* pcur = nullptr;  // collocated StabType 1: r -= G p_n
    T inv_tau = 0;
    double* cfr_partial = nullptr;  // per-block sums of r_u * volume (nullptr: off)
    const T* tau = nullptr;         // collocated learned closure: subtract div tau (nullptr: off)
};
// nblocks (may be nullptr) receives the number of per-block partial sums written to cfr_partial.
template <typename T>
hipError_t launch_flux(hipStream_t s, const Mesh<T>& m, const FluxArgs<T>& a, int* nblocks);
int flux_grid_blocks(int nx, int ny, int nz);  // upper bound of that number (buffer sizing)

template <typename T>
hipError_t launch_turb_visc(hipStream_t s, const Mesh<T>& m, int scheme, const T* vel, T* nut);
template <typename T>
hipError_t launch_gradient(hipStream_t s, const Mesh<T>& m, int scheme, const T* p, T* g);

// u_next = (r - G p) * tau_over_bd0 (+ flow-rate shift on u); collocated StabType 1
// also forms p_next = dp + p_cur.
template <typename T>
struct ProjectArgs {
    int scheme = 0;
    const T* p = nullptr;       // pressure (or pressure increment)
    const T* r = nullptr;       // right-hand side
    T* vnext = nullptr;
    T tau_over_bd0 = 1;
    const double* shift = nullptr;  // flow-rate shift on u (nullptr: none)
    const T* pcur = nullptr;        // StabType 1
    T* pnext = nullptr;             // StabType 1
    // The step's diagnostics (nullptr: off): the projection also writes one
    // partial diagnostics record per block for the new velocity (and pressure,
    // nu_t), which diag_out then receives per plane, as launch_diag would.
    double* diag = nullptr;
    double* diag_out = nullptr;
    double visc_fac = 0;
    const T* nut = nullptr;
};
template <typename T>
hipError_t launch_project(hipStream_t s, const Mesh<T>& m, const ProjectArgs<T>& a);

// Adds a device-held shift to the streamwise velocity (collocated walls excluded).
template <typename T>
hipError_t launch_add_u(hipStream_t s, const Mesh<T>& m, int scheme, T* v, const double* shift);
// f[cell] -= (sum[0] * coef) * plane_w[iz]  (plane_w == nullptr: 1); NormalizePressure's shift.
template <typename T>
hipError_t launch_add_scaled_plane(hipStream_t s, const Mesh<T>& m, T* f, const double* sum, double coef, const T* plane_w);

// ---------------------------------------------------------------------------
// Stage updates of the collocated IMEX Runge-Kutta (SRK) step, one thread per node
// ---------------------------------------------------------------------------
// kc = KC - G s with s = S (+ S2 when set; S == nullptr: no gradient), stored to
// KCout when that is set; out = X + sum_{j<nk} c[j] K[j] + ckc kc + sa A + sb B,
// absent terms skipped. With pj_out: pj_out = S + S2 at the node, and with
// qj_out: qj_out = inv_tau_stage (S + S2 - pjhat) (the stage pressure and its
// rate, collocated_SRK/func_global.h:389-393). out may alias X, KCout KC.
template <typename T>
struct SrkVec {
    const T* KC = nullptr;
    const T* S = nullptr;
    const T* S2 = nullptr;
    T* KCout = nullptr;
    T* out = nullptr;
    const T* X = nullptr;
    int nk = 0;
    const T* K[kMaxImexStages] = {};
    T c[kMaxImexStages] = {};
    T ckc = 0;
    const T* A = nullptr;
    T sa = 0;
    const T* B = nullptr;
    T sb = 0;
    T* pj_out = nullptr;
    T* qj_out = nullptr;
    const T* pjhat = nullptr;
    T inv_tau_stage = 0;
};
template <typename T>
hipError_t launch_srk_vec(hipStream_t s, const Mesh<T>& m, const SrkVec<T>& a);

// Right-hand side scalar of the stage pressure system (collocated_SRK/func_global.h:357-384),
// with the stored mujtilde[j] taken as mujt[j] - cfix X, X = (stab ? tau_stage q : p)
// (the update of :391, applied here instead of in place):
// stab 0: bs = (m0 p + sum_j cm[j] mujt~[j]) - clast p          (m0 only for ck)
// stab 1: muj = m0 (tau_stage q) + sum_j cm[j] mujt~[j]; pjhat = p + ph0 q + sum_j cq[j] qj[j];
//         bs = muj + pjhat - clast q                             (m0, ph0 only for ck)
template <typename T>
struct SrkPre {
    int stab = 0, ck = 0;
    const T* p = nullptr;
    const T* q = nullptr;
    T* bs = nullptr;
    T* pjhat = nullptr;
    T m0 = 1, ph0 = 0, tau_stage = 0, clast = 0, cfix = 0;
    int nj = 0;
    const T* mujt[kMaxImexStages] = {};
    T cm[kMaxImexStages] = {};
    const T* qj[kMaxImexStages] = {};
    T cq[kMaxImexStages] = {};
};
template <typename T>
hipError_t launch_srk_pre(hipStream_t s, const Mesh<T>& m, const SrkPre<T>& a);

// ---------------------------------------------------------------------------
// Spectral pressure solve
// ---------------------------------------------------------------------------
// The flow-rate correction's shift = (ub V - q scale) / V_u, q the sum of the
// explicit-term kernel's per-block partial sums. The divergence transform that
// follows that kernel forms it in its block 0 (partial != nullptr), so it needs
// no launch of its own.
struct CfrShiftArgs {
    const double* partial = nullptr;
    int nb = 0;
    double scale = 0.0, ub_vol = 0.0, vol_u = 1.0;
    double* shift = nullptr;
};
// Forward x transform of the discrete divergence of v (computed on the fly; ApplyDivergence).
template <typename T>
hipError_t launch_fft_x_fwd_div(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const Mesh<T>& m, int scheme,
                                const T* v, Cplx<T>* spec, const CfrShiftArgs& cfr = CfrShiftArgs());
// psum/C (inverse only, may be nullptr): X00[iz] -= psum[0] * C[iz] before the transform.
template <typename T>
hipError_t launch_fft_y(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, Cplx<T>* spec, int nxh, int nz,
                        int inverse, const double* psum, const double* C);
template <typename T>
hipError_t launch_fft_x_inv(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const Cplx<T>* spec, T* out,
                            int ny, int nrows, T scale);

// Training samples of the learned closure: forward x transform of component ca
// (cb < 0) or of the product of components ca, cb of a real3 field, and the
// coarse half spectrum of the modes a sharp cut-off keeps.
template <typename T>
hipError_t launch_fft_x_fwd_q(hipStream_t s, const FFTPlan& pl, const Cplx<T>* tw, const T* v, int ca, int cb,
                              int ny, int nrows, Cplx<T>* spec);
template <typename T>
hipError_t launch_spec_crop(hipStream_t s, const Cplx<T>* fine, int nxh, int ny, Cplx<T>* coarse, int nxhc, int nyc,
                            int nz);

// Factor arrays on the device, packed per row so that a row costs one or two
// load instructions: F[iz][col][FW] = (M1, M2, M3, D) for W = 3, (M1, D) for
// W = 1; U[iz][col][UW] = (U1, U2, U3) or (U1). col = mode (or index in the
// low-mode list).
// Column of the z-solve factors of spectrum mode kx * ny + ky (ny even).
__host__ __device__ inline int zfac_col(int mode, int ny) {
    const int kx = mode / ny, ky = mode - kx * ny;
    return kx * (ny / 2 + 1) + (ky <= ny - ky ? ky : ny - ky);
}
template <typename CT>
struct ZFacDev {
    int W = 1, nz = 0, ncols = 0;  // factor columns
    int nmodes = 0, ny = 0;        // spectrum modes, and modes per kx (columns by zfac_col)
    const CT* F = nullptr;
    const CT* U = nullptr;
    // In-lane forward elimination (R != nullptr): per-row constants
    // R[iz][zrow_w(W)] and wavenumber terms KX, KY (see ZFactors); the forward
    // sweep then reads no factors, the backward sweep still reads U.
    const CT* R = nullptr;
    const CT* KX = nullptr;
    const CT* KY = nullptr;
};
constexpr int zfac_fw(int W) { return W == 3 ? 4 : 2; }
constexpr int zfac_uw(int W) { return W == 3 ? 3 : 1; }
struct ZNorm {
    const double *Wt = nullptr, *C = nullptr;
    double* s00 = nullptr;  // receives the (0,0)-mode normalisation sum (one double)
};
// Solve every mode. FP32-bulk engines pass lowidx (mode -> index in the FP64
// factor list, -1 for the FP32 modes), lowmode (index -> mode), the FP64 factors
// of those modes and ylow (nz * nlow * 2 doubles of scratch); FP64 engines pass
// lowidx = lowmode = nullptr. The lane of the (0,0) mode writes its
// normalisation sum to norm.s00[0].
template <typename T>
hipError_t launch_zsolve(hipStream_t s, Cplx<T>* spec, const ZFacDev<T>& f, const ZFacDev<double>& flow,
                         const int* lowidx, const int* lowmode, int nlow, double* ylow, ZNorm norm);

// ---------------------------------------------------------------------------
// Reductions (FP64 accumulation, deterministic two-stage)
// ---------------------------------------------------------------------------
enum RedOp { RED_INTEGRAL = 0, RED_KE = 1, RED_PSUM = 2, RED_CFL = 3 };
template <typename T>
struct RedArgs {
    int op = 0, scheme = 0;
    const T* v = nullptr;    // vector field
    const T* p = nullptr;    // scalar field
    const T* nut = nullptr;
    const double* wplane = nullptr;  // per-plane weights (RED_PSUM)
    double tau = 0, visc_fac = 0;
    int mode = 3;
};
constexpr int kRedBlocks = 1024;
// out: 3 doubles (integral), 1 (KE, PSUM, CFL). partial: kRedBlocks*3 doubles.
template <typename T>
hipError_t launch_reduce(hipStream_t s, const Mesh<T>& m, const RedArgs<T>& a, double* partial, double* out);
// Plane statistics: out[iz*NV + k], NV = 12 (staggered) or 10 (collocated).
template <typename T>
hipError_t launch_profiles(hipStream_t s, const Mesh<T>& m, int scheme, const T* v, const T* p, const T* nut,
                           double* partial, double* out);
int profiles_chunks(int nxy);
// Fused diagnostics: per plane a record of diag_record(scheme) doubles: the plane
// sums of the profile statistics (NV values as launch_profiles, not divided by
// the plane size), sum w, sum w^2 (w at its own points), and the maxima of the
// convective, viscous and total CFL numbers over tau. partial: profiles_chunks *
// nz * diag_record doubles.
int diag_record(int scheme);
template <typename T>
hipError_t launch_diag(hipStream_t s, const Mesh<T>& m, int scheme, const T* v, const T* p, const T* nut,
                       double visc_fac, double* partial, double* out);
// Per-plane records from nch partial records per plane (any nch).
hipError_t launch_diag_final_wide(hipStream_t s, int scheme, const double* partial, int nch, int nz, double* out);
// Partial records the projection writes when it forms the diagnostics (see ProjectArgs::diag).
int project_diag_chunks(int nxy);

}  // namespace tc
