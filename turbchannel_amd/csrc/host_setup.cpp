// Host-side set-up of the engine: wall-normal mesh data, the precomputed
// elimination of the pressure systems, the seeded initial field and FFT plans.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "tc_internal.h"

namespace tc {

static const double kPi = 3.14159265358979323846;

// Gauss elimination without pivoting on a banded matrix stored row-wise with NB
// entries per row (staggered/main.cpp:73-109). Used for the gradient kernel.
int solve_banded(int N, int NB, double* M, double* x, double tiny) {
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
        if (std::fabs(M[i * NB + w]) < tiny) return 2;
        double mult = 1.0 / M[i * NB + w];
        for (int k = 0; k <= w; k++)
            if (i + k < N) M[i * NB + w + k] *= mult;
        x[i] *= mult;
    }
    for (int i = N - 1; i >= 0; i--)
        for (int j = 1; j <= w; j++)
            if (i + j < N) x[i] -= M[i * NB + w + j] * x[i + j];
    return 0;
}

// 7-band matrix of D*G of the fourth-order scheme for the mode constant in x and y,
// with the mirror rules at the walls (staggered/main.cpp:12-48).
static bool vv_matrix(const HostMesh& m, std::vector<double>& M) {
    const int nz = m.nz, w = 3, nb = 7;
    const double gc[4] = {1. / 27., -1., 1., -1. / 27.};
    M.assign((size_t)nb * nz, 0.0);
    for (int j = 0; j < nz; j++) {
        for (int ll = -1; ll <= 2; ll++) {
            int l = j + ll;  // face of the gradient
            if (l == 0 || l == nz) continue;
            if (l < 0) l = -l;
            if (l > nz) l = 2 * nz - l;
            for (int mm = -2; mm <= 1; mm++) {
                double c2 = gc[mm + 2];
                int c = l + mm;  // cell of the pressure
                if (c < 0) { c = -c - 1; c2 = -c2; }
                if (c >= nz) { c = 2 * nz - c - 1; c2 = -c2; }
                int pos = c - j + w;
                if (pos < 0 || pos >= nb || l <= 0 || l >= nz) return false;
                M[(size_t)j * nb + pos] += gc[ll + 1] * c2 * m.GCZ[l];
            }
        }
    }
    return true;
}

std::string build_mesh(const tc_params& p, HostMesh& m) {
    if (p.scheme < 0 || p.scheme > 2) return "unknown scheme";
    if (p.nx < 4 || p.ny < 4 || (p.nx & 1) || (p.ny & 1)) return "nx, ny must be even and >= 4";
    if (p.nz < 6) return "nz must be >= 6";
    if (!p.zz) return "zz is NULL";
    if (!(p.hx > 0) || !(p.hy > 0)) return "hx, hy must be positive";
    const int nz = p.nz;
    m.scheme = p.scheme;
    m.nx = p.nx;
    m.ny = p.ny;
    m.nz = nz;
    const bool ab3 = p.scheme == TC_SCHEME_AB3;
    const int nzz = ab3 ? nz : nz + 1;
    m.ZZ.assign(p.zz, p.zz + nzz);
    for (int i = 1; i < nzz; i++)
        if (!(m.ZZ[i] > m.ZZ[i - 1])) return "zz must be strictly increasing";
    m.DZ.assign(nz, 0.0);
    m.DZW.assign(nz, 0.0);
    m.GC1.assign(nz, 0.0);
    m.GC2.assign(nz, 0.0);
    m.GCZ.assign(nz, 0.0);
    m.KERG.assign(nz, 0.0);
    m.cellh.assign(nz, 0.0);
    const double Lxy = p.hx * p.nx * p.hy * p.ny;
    if (ab3) {
        // collocated nodes (collocated_AB3/main.cpp:238-241, 295-300)
        for (int i = 0; i < nz - 1; i++) m.DZ[i] = m.ZZ[i + 1] - m.ZZ[i];
        m.MAT_A.assign(nz, 0.0);
        for (int j = 1; j < nz; j++) m.MAT_A[j] = p.hx * p.hy / m.DZ[j - 1];
        if (p.fd_mode) {
            m.MAT_A[1] *= 0.5;
            m.MAT_A[nz - 1] *= 0.5;
        }
        for (int iz = 0; iz < nz; iz++)  // GetCellHeight (collocated_AB3/func_global.h:150-155)
            m.cellh[iz] = iz == 0 ? 0.5 * m.DZ[0] : iz == nz - 1 ? 0.5 * m.DZ[nz - 2] : 0.5 * (m.DZ[iz - 1] + m.DZ[iz]);
        m.area = Lxy * (m.ZZ[nz - 1] - m.ZZ[0]);
        double v = 0;
        for (int iz = 1; iz < nz - 1; iz++) v += m.cellh[iz];
        m.vol_u = v * Lxy;
        return "";
    }
    for (int i = 0; i < nz; i++) m.DZ[i] = m.ZZ[i + 1] - m.ZZ[i];
    m.area = Lxy * (m.ZZ[nz] - m.ZZ[0]);
    m.vol_u = m.area;
    if (p.scheme == TC_SCHEME_HW) {
        // tridiagonal coefficients (staggered/main.cpp:117-120)
        m.MAT_A.assign(nz, 0.0);
        for (int j = 1; j <= nz - 1; j++) m.MAT_A[j] = 2. / (m.DZ[j - 1] + m.DZ[j]);
        return "";
    }
    // fourth-order scheme (staggered/main.cpp:122-155)
    const std::vector<double>& DZ = m.DZ;
    for (int iz = 0; iz < nz; iz++) {
        m.DZW[iz] = (DZ[iz] + DZ[iz == 0 ? iz : iz - 1] + DZ[iz == nz - 1 ? iz : iz + 1]) / 3.;
        m.GC1[iz] = DZ[iz] / (DZ[iz] - m.DZW[iz] / 9.);
        m.GC2[iz] = -(1. / 27.) * m.DZW[iz] / (DZ[iz] - m.DZW[iz] / 9.);
        // GCZ[0] is never used (the gradient has no z-component at iz = 0)
        m.GCZ[iz] = iz == 0 ? 0.0
                            : 1 / (0.5 * (DZ[iz] + DZ[iz - 1]) -
                                   (1. / 27.) * (DZ[iz] + DZ[iz - 1] +
                                                 0.5 * (DZ[iz == nz - 1 ? iz : iz + 1] + DZ[iz == 1 ? 0 : iz - 2])));
    }
    std::vector<double> M;
    if (!vv_matrix(m, M)) return "Internal error";
    m.MAT_A.assign((size_t)3 * nz, 0.0);
    for (int iz = 0; iz < nz; iz++)
        for (int i = 0; i < 3; i++) m.MAT_A[iz * 3 + i] = (iz + i - 3 >= 0) ? M[(size_t)iz * 7 + i] : 0.0;
    m.MAT_D[0] = M[3];
    m.MAT_D[1] = M[7 + 3];
    m.MAT_D[2] = M[14 + 3];
    m.MAT_D[3] = M[(size_t)(nz - 3) * 7 + 3];
    m.MAT_D[4] = M[(size_t)(nz - 2) * 7 + 3];
    m.MAT_D[5] = M[(size_t)(nz - 1) * 7 + 3];
    // kernel of the gradient: tridiagonal system with eps = 1/27 (staggered/main.cpp:140-154)
    std::vector<double> T3((size_t)3 * nz, 0.0);
    const double eps = 1. / 27.;
    T3[1] = -1;
    T3[2] = eps;
    for (int iz = 1; iz < nz - 1; iz++) {
        T3[iz * 3 + 0] = T3[iz * 3 + 2] = eps;
        T3[iz * 3 + 1] = -1 + eps;
    }
    T3[(size_t)(nz - 1) * 3] = eps;
    T3[(size_t)(nz - 1) * 3 + 1] = -1;
    for (int iz = 0; iz < nz; iz++) m.KERG[iz] = 1;
    if (solve_banded(nz, 3, T3.data(), m.KERG.data(), 1e-10)) return "Internal error in solving system for K";
    m.KERG_norm2 = 0;
    for (int iz = 0; iz < nz; iz++) m.KERG_norm2 += m.KERG[iz] * m.KERG[iz] * DZ[iz];
    return "";
}

// ---------------------------------------------------------------------------
// Elimination factors of the wall-normal systems, per horizontal mode.
// The systems are those of ThomasForPressure: staggered/func_global.h:71-106
// (HW), :110-189 (VV), collocated_AB3/func_global.h:8-61. The arithmetic of each
// factor is the reference's, so substitution with them reproduces its solves.
// ---------------------------------------------------------------------------
std::string build_zfactors(const tc_params& p, const HostMesh& m, ZFactors& f) {
    const int nx = p.nx, ny = p.ny, nz = p.nz, nxh = nx / 2 + 1, nyh = ny / 2 + 1;
    const int nm = nxh * nyh;  // columns (kx, |ky|)
    const int W = p.scheme == TC_SCHEME_VV ? 3 : 1;
    f.W = W;
    f.nz = nz;
    f.ncols = nm;
    for (int k = 0; k < 3; k++) {
        f.M[k].assign(k < W ? (size_t)nz * nm : 0, 0.0);
        f.U[k].assign(k < W ? (size_t)nz * nm : 0, 0.0);
    }
    f.D.assign((size_t)nz * nm, 0.0);
    const double inv_hx = 1.0 / p.hx, inv_hy = 1.0 / p.hy;
    const double* DZ = m.DZ.data();
    const double* A = m.MAT_A.data();
    std::string err;
    for (int mode = 0; mode < nm; mode++) {
        const int ix = mode / nyh, iy = mode % nyh;  // column (kx, |ky|), ky <= ny / 2
        auto at = [&](int iz) { return (size_t)iz * nm + mode; };
        if (p.scheme == TC_SCHEME_HW || p.scheme == TC_SCHEME_AB3) {
            const double cphi = std::cos(2. * kPi * ix / nx), cpsi = std::cos(2. * kPi * iy / ny);
            double pv_prev;
            if (p.scheme == TC_SCHEME_HW) {
                const double MM = 2 * ((1 - cphi) * inv_hx * inv_hx + (1 - cpsi) * inv_hy * inv_hy);
                double b0 = -DZ[0] * MM - A[1];
                if (mode == 0) b0 *= 2;  // pin
                double inv_b0 = 1 / b0;
                f.D[at(0)] = inv_b0;
                pv_prev = A[1] * inv_b0;
                f.U[0][at(0)] = pv_prev;
                for (int i = 1; i <= nz - 1; i++) {
                    double delta = -A[i] - DZ[i] * MM;
                    if (i != nz - 1) delta -= A[i + 1];
                    delta -= A[i] * pv_prev;
                    delta = 1 / delta;
                    pv_prev = (i != nz - 1) ? A[i + 1] * delta : 0.0;
                    f.U[0][at(i)] = pv_prev;
                    f.M[0][at(i)] = -A[i];
                    f.D[at(i)] = delta;
                }
            } else {
                const double MM = (1 - cphi) * p.hy / p.hx + (1 - cpsi) * p.hx / p.hy;
                const int FD = p.fd_mode;
                double b0 = -A[1];
                if (!FD) b0 += DZ[0] * MM;
                if (mode == 0) b0 *= 2;  // pin
                double inv_b0 = 1 / b0;
                f.D[at(0)] = inv_b0;
                pv_prev = A[1] * inv_b0;
                f.U[0][at(0)] = pv_prev;
                for (int i = 1; i <= nz - 1; i++) {
                    const double mm = FD && (i == 0 || i == nz - 1) ? 0.0 : MM;
                    double delta = 0;
                    if (i != 0) delta -= A[i] + DZ[i - 1] * mm;
                    if (i != nz - 1) delta -= A[i + 1] + DZ[i] * mm;
                    delta -= A[i] * pv_prev;
                    delta = 1 / delta;
                    pv_prev = (i != nz - 1) ? A[i + 1] * delta : 0.0;
                    f.U[0][at(i)] = pv_prev;
                    f.M[0][at(i)] = -A[i];
                    f.D[at(i)] = delta;
                }
            }
            continue;
        }
        // seven-band system of the fourth-order scheme
        const int w = 3;
        std::vector<double> buf((size_t)w * nz, 0.0);
        const double sx = std::sin(kPi * ix / nx), sy = std::sin(kPi * iy / ny);
        const double s3x = 3 * sx - 4 * sx * sx * sx, s3y = 3 * sy - 4 * sy * sy * sy;
        bool singular = false;
        for (int i = 0; i < nz && !singular; i++) {
            double line[7] = {0, 0, 0, 0, 0, 0, 0};
            double sum = 0;
            for (int k = 0; k < w; k++) {
                if (i + k - w < 0 || i + k - w >= nz) continue;
                line[k] = A[i * w + k];
                sum += A[i * w + k];
            }
            for (int k = 1; k <= w; k++) {
                if (i + k >= nz) continue;
                line[w + k] = A[(i + k) * w + w - k];
                sum += A[(i + k) * w + w - k];
            }
            if (i <= 2) line[w] = m.MAT_D[i];
            else if (i >= nz - 3) line[w] = m.MAT_D[i - (nz - 3) + 3];
            else line[w] = -sum;
            if (mode == 0 && i == 0) line[w] *= 2;  // pin
            double ci = DZ[i] * sx - (1. / 27.) * m.DZW[i] * s3x;
            line[w] -= 4 * inv_hx * inv_hx * ci * ci / (DZ[i] - (1. / 9.) * m.DZW[i]);
            ci = DZ[i] * sy - (1. / 27.) * m.DZW[i] * s3y;
            line[w] -= 4 * inv_hy * inv_hy * ci * ci / (DZ[i] - (1. / 9.) * m.DZW[i]);
            for (int j = -w; j < 0; j++) {
                if (i + j < 0) continue;
                double mult = -line[w + j];
                for (int k = 1; k <= w; k++)
                    if (i + j + k < nz) line[w + j + k] += mult * buf[(size_t)(i + j) * w + k - 1];
                f.M[-j - 1][at(i)] = mult;  // M1 <- j=-1, M2 <- j=-2, M3 <- j=-3
            }
            if (std::fabs(line[w]) < 1e-30) { singular = true; break; }
            double d = 1.0 / line[w];
            for (int k = 1; k <= w; k++)
                if (i + k < nz) {
                    buf[(size_t)i * w + k - 1] = line[w + k] * d;
                    f.U[k - 1][at(i)] = line[w + k] * d;
                }
            f.D[at(i)] = d;
        }
        if (singular) {
            err = "zero pivot in the pressure system";
        }
    }
    if (!err.empty()) return err;
    // The constants of the in-lane forward elimination (the loops above, split
    // into what depends on the row and what on the wavenumbers)
    if (W == 3) {
        const int RW = zrow_w(3);
        f.R.assign((size_t)nz * RW, 0.0);
        f.KX.assign((size_t)nxh * 2, 0.0);
        f.KY.assign((size_t)nyh * 2, 0.0);
        for (int i = 0; i < nz; i++) {
            double* r = &f.R[(size_t)i * RW];
            double sum = 0;
            for (int k = 0; k < 3; k++)
                if (i + k - 3 >= 0) {
                    r[k] = A[i * 3 + k];
                    sum += A[i * 3 + k];
                }
            for (int k = 1; k <= 3; k++)
                if (i + k < nz) {
                    r[3 + k] = A[(i + k) * 3 + 3 - k];
                    sum += A[(i + k) * 3 + 3 - k];
                }
            r[3] = i <= 2 ? m.MAT_D[i] : (i >= nz - 3 ? m.MAT_D[i - (nz - 3) + 3] : -sum);
            const double den = DZ[i] - (1. / 9.) * m.DZW[i];
            r[7] = DZ[i];
            r[8] = (1. / 27.) * m.DZW[i];
            r[9] = 4 * inv_hx * inv_hx / den;
            r[10] = 4 * inv_hy * inv_hy / den;
        }
        for (int ix = 0; ix < nxh; ix++) {
            const double sx = std::sin(kPi * ix / nx);
            f.KX[(size_t)ix * 2] = sx;
            f.KX[(size_t)ix * 2 + 1] = 3 * sx - 4 * sx * sx * sx;
        }
        for (int iy = 0; iy < nyh; iy++) {
            const double sy = std::sin(kPi * iy / ny);
            f.KY[(size_t)iy * 2] = sy;
            f.KY[(size_t)iy * 2 + 1] = 3 * sy - 4 * sy * sy * sy;
        }
    }
    if (W == 1) {
        const bool hw = p.scheme == TC_SCHEME_HW, fd = p.fd_mode != 0;
        f.R.assign((size_t)nz * zrow_w(1), 0.0);
        f.KX.assign(nxh, 0.0);
        f.KY.assign(nyh, 0.0);
        for (int i = 0; i < nz; i++) {
            double* r = &f.R[(size_t)i * zrow_w(1)];
            r[0] = i > 0 ? A[i] : 0.0;
            r[1] = i < nz - 1 ? A[i + 1] : 0.0;
            if (i == 0) {
                r[2] = -A[1];
                r[3] = hw ? -DZ[0] : (fd ? 0.0 : DZ[0]);
            } else {
                r[2] = -A[i] - (i < nz - 1 ? A[i + 1] : 0.0);
                if (hw) r[3] = -DZ[i];
                else r[3] = -((fd && i == nz - 1) ? 0.0 : DZ[i - 1]) - (i < nz - 1 ? DZ[i] : 0.0);
            }
        }
        for (int ix = 0; ix < nxh; ix++) {
            const double c = 1 - std::cos(2. * kPi * ix / nx);
            f.KX[ix] = hw ? 2 * c * inv_hx * inv_hx : c * p.hy / p.hx;
        }
        for (int iy = 0; iy < nyh; iy++) {
            const double c = 1 - std::cos(2. * kPi * iy / ny);
            f.KY[iy] = hw ? 2 * c * inv_hy * inv_hy : c * p.hx / p.hy;
        }
    }
    // (0,0)-mode normalisation in spectral space (NormalizePressure,
    // staggered/func_global.h:657-675; collocated_AB3/func_global.h:163-170).
    f.Wt.assign(nz, 0.0);
    f.C.assign(nz, 0.0);
    const double nxy = (double)nx * ny;
    for (int iz = 0; iz < nz; iz++) {
        if (p.scheme == TC_SCHEME_HW) {
            f.Wt[iz] = DZ[iz];
            f.C[iz] = 1.0 / (m.ZZ[nz] - m.ZZ[0]);
        } else if (p.scheme == TC_SCHEME_VV) {
            f.Wt[iz] = m.KERG[iz] * DZ[iz];
            f.C[iz] = m.KERG[iz] / m.KERG_norm2;
        } else {
            f.Wt[iz] = m.cellh[iz] * p.hx * p.hy;
            f.C[iz] = p.hx * p.hy / m.area * nxy;
        }
    }
    return "";
}

// ---------------------------------------------------------------------------
// Seeded initial field: parabola + curl of a smoothed Gaussian vector potential.
// ---------------------------------------------------------------------------
namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double gauss() {  // Box-Muller, one value per pair
        double u1 = 1.0 - uniform();
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};
}  // namespace

void make_initial_field(const tc_params& p, const HostMesh& m, unsigned long long seed, double ubulk,
                        double noise_rms, double* u) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const size_t nxy = (size_t)nx * ny;
    const bool ab3 = p.scheme == TC_SCHEME_AB3;
    const int np = ab3 ? nz : nz + 1;  // potential planes: faces (staggered) or nodes (collocated)
    const double hx = p.hx, hy = p.hy;
    const double z0 = m.ZZ[0], z1 = ab3 ? m.ZZ[nz - 1] : m.ZZ[nz];
    const double zc = 0.5 * (z0 + z1), H = 0.5 * (z1 - z0);
    const size_t npl = nxy * np;
    std::vector<double> psi(3 * npl), tmp(npl);
    SplitMix64 rng{(uint64_t)seed};
    for (size_t i = 0; i < 3 * npl; i++) psi[i] = rng.gauss();
    for (int c = 0; c < 3; c++) {
        double* a = psi.data() + c * npl;
        for (int pass = 0; pass < 2; pass++) {  // [1 2 1]/4 in x, then in y, twice
            for (size_t i = 0; i < npl; i++) {
                const int ix = (int)(i % nx);
                const size_t row = i - ix;
                tmp[i] = 0.25 * a[row + (ix + nx - 1) % nx] + 0.5 * a[i] + 0.25 * a[row + (ix + 1) % nx];
            }
            for (size_t i = 0; i < npl; i++) {
                const int iy = (int)((i / nx) % ny);
                const size_t base = i - (size_t)iy * nx;
                a[i] = 0.25 * tmp[base + (size_t)((iy + ny - 1) % ny) * nx] + 0.5 * tmp[i] +
                       0.25 * tmp[base + (size_t)((iy + 1) % ny) * nx];
            }
        }
        for (int k = 0; k < np; k++) {  // wall envelope (1 - zeta^2)^2
            double z = (ab3 || c < 2) ? m.ZZ[k] : (k < nz ? 0.5 * (m.ZZ[k] + m.ZZ[k + 1]) : z1);
            double zeta = (z - zc) / H;
            double env = (1 - zeta * zeta) * (1 - zeta * zeta);
            if (!ab3 && c == 2 && k == nz) env = 0;
            for (size_t j = 0; j < nxy; j++) a[(size_t)k * nxy + j] *= env;
        }
    }
    const double *px = psi.data(), *py = psi.data() + npl, *pz = psi.data() + 2 * npl;
    auto P = [&](const double* a, int i, int j, int k) {
        return a[(size_t)k * nxy + (size_t)((j + ny) % ny) * nx + (size_t)((i + nx) % nx)];
    };
    double ss = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t in = (size_t)k * nxy + (size_t)j * nx + i;
                double a, b, c;
                if (!ab3) {  // staggered curl: potential on cell edges, exactly solenoidal for the MAC divergence
                    const double dz = m.DZ[k];
                    a = (P(pz, i, j + 1, k) - P(pz, i, j, k)) / hy - (P(py, i, j, k + 1) - P(py, i, j, k)) / dz;
                    b = (P(px, i, j, k + 1) - P(px, i, j, k)) / dz - (P(pz, i + 1, j, k) - P(pz, i, j, k)) / hx;
                    c = (P(py, i + 1, j, k) - P(py, i, j, k)) / hx - (P(px, i, j + 1, k) - P(px, i, j, k)) / hy;
                    if (k == 0) c = 0;
                } else if (k == 0 || k == nz - 1) {
                    a = b = c = 0;
                } else {  // nodal central differences
                    const double dz2 = m.ZZ[k + 1] - m.ZZ[k - 1];
                    a = (P(pz, i, j + 1, k) - P(pz, i, j - 1, k)) / (2 * hy) - (P(py, i, j, k + 1) - P(py, i, j, k - 1)) / dz2;
                    b = (P(px, i, j, k + 1) - P(px, i, j, k - 1)) / dz2 - (P(pz, i + 1, j, k) - P(pz, i - 1, j, k)) / (2 * hx);
                    c = (P(py, i + 1, j, k) - P(py, i - 1, j, k)) / (2 * hx) - (P(px, i, j + 1, k) - P(px, i, j - 1, k)) / (2 * hy);
                }
                u[3 * in] = a;
                u[3 * in + 1] = b;
                u[3 * in + 2] = c;
                ss += a * a + b * b + c * c;
            }
    const double n3d = (double)nxy * nz;
    const double scale = ss > 0 ? noise_rms / std::sqrt(ss / (3.0 * n3d)) : 0.0;
    for (int k = 0; k < nz; k++) {
        const double z = ab3 ? m.ZZ[k] : 0.5 * (m.ZZ[k] + m.ZZ[k + 1]);
        const double zeta = (z - zc) / H;
        double ulam = 1.5 * ubulk * (1 - zeta * zeta);
        if (ab3 && (k == 0 || k == nz - 1)) ulam = 0;
        for (size_t j = 0; j < nxy; j++) {
            const size_t in = (size_t)k * nxy + j;
            u[3 * in] = ulam + scale * u[3 * in];
            u[3 * in + 1] *= scale;
            u[3 * in + 2] *= scale;
        }
    }
}

// Radix sequence of a mixed-radix transform, few stages: 8s first, then one 4,
// a 6 where a 2 and a 3 remain, then 2, 3, 5 (384 = 8 8 6, 160 = 8 4 5).
bool make_plan(int n, FFTPlan& pl) {
    pl = FFTPlan();
    pl.n = n;
    int r = n;
    auto push = [&](int f) {
        if (pl.nst >= kMaxStages) return false;
        pl.radix[pl.nst++] = f;
        r /= f;
        return true;
    };
    while (r % 8 == 0)
        if (!push(8)) return false;
    if (r % 4 == 0 && !push(4)) return false;
    if (r % 6 == 0 && !push(6)) return false;
    for (int f : {2, 3, 5})
        while (r % f == 0)
            if (!push(f)) return false;
    if (r != 1 || n < 1) return false;
    int L = n;
    for (int st = 0; st < pl.nst; st++) {
        const int R = pl.radix[st];
        pl.twstride[st] = n / L;
        pl.dnr[st] = FastDiv((unsigned)(n / R));
        pl.dspan[st] = FastDiv((unsigned)(L / R));
        pl.drad[st] = FastDiv((unsigned)R);
        L /= R;
    }
    pl.dn = FastDiv((unsigned)n);
    return true;
}

}  // namespace tc
