// C ABI of the MI355X time stepper (include/turbchannel.h).
//
// The engine owns all device buffers and one HIP stream (default priority).
// FP64 engines run everything in double. FP32-bulk engines keep fields,
// stencils and FFTs in float and use double for the lowest-wavenumber
// wall-normal solves, the (0,0)-mode normalisation, the flow-rate correction
// and all reductions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "diag.h"
#include "kernels.h"
#include "mlp.h"

using namespace tc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

void tc_set_error(const char* msg) { g_err = msg; }

namespace {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fail(TC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kLowK = 2;  // FP32-bulk: modes with kx <= 2 and |ky| <= 2 are solved in FP64

struct Timing {
    bool on = false;
    struct Rec {
        int k;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double total_ms[TC_K_COUNT] = {0};
    long long count[TC_K_COUNT] = {0};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};

}  // namespace

struct tc_engine {
    tc_params P{};
    std::vector<double> zz;
    HostMesh hm;
    ZFactors zf;
    int dev = 0;
    hipStream_t s = nullptr;
    bool fp32 = false;
    int nx = 0, ny = 0, nz = 0, nxy = 0, nxh = 0, nm = 0;
    size_t n3d = 0, rsz = 8;
    void* vel[4] = {};
    void* flux[4] = {};
    void* pres[4] = {};
    void* nut = nullptr;
    void* dp = nullptr;   // pressure increment (collocated StabType 1) / scratch pressure
    void* spec = nullptr;
    void* dmesh = nullptr;  // mesh arrays in T
    size_t mstride = 0;     // elements per mesh array
    double* d_wplane = nullptr;  // normalisation weights per plane
    void* d_kerg = nullptr;      // KERG in T (plane shift weight, VV)
    void* zfac = nullptr;        // factors in T
    void* zinl = nullptr;        // in-lane forward constants in T (FP32-bulk): R, KX, KY
    double* zfac_low = nullptr;  // factors of the low modes in double
    int* d_lowidx = nullptr;  // mode -> index in the FP64 low-mode list, or -1
    int* d_lowmode = nullptr; // index in the low-mode list -> mode
    int nlow = 0;
    double* ylow = nullptr;
    double *d_Wt = nullptr, *d_C = nullptr, *d_s00 = nullptr;
    FFTPlan plx, ply;
    void *twx = nullptr, *twy = nullptr;
    double* d_partial = nullptr;
    size_t npartial = 0;
    double* d_out = nullptr;    // small results (16 doubles)
    double* d_shift = nullptr;  // flow-rate shift
    double* d_prof = nullptr;   // profiles result (nz * 12)
    // diagnostics, nz records each, in pinned host memory that the final
    // reductions write through its device address (no copy to the host):
    // buffers 0 and 1 hold the records the last two steps formed (tc_step
    // alternates), buffer 2 the record of a separate pass
    double* h_diag[3] = {};
    double* hd_diag[3] = {};
    hipEvent_t diag_ev[2] = {};  // recorded after the step that wrote buffer 0 / 1
    // tc_step also forms the diagnostics of the new fields in its projection
    // (tc_step_diagnostics); buffer b then holds those of velocity (and
    // pressure) slot diag_rec[b] until a call changes a field, or -1
    bool step_diag = false;
    int diag_rec[2] = {-1, -1};
    int diag_cur = 0;  // the buffer the next step writes
    void drop_records() { diag_rec[0] = diag_rec[1] = -1; }
    // learned closure (collocated, sgs_model 1, 2 or 3)
    void* tau = nullptr;                // [node][6] closure stress
    unsigned short* sgs_w = nullptr;    // bfloat16 weight block
    float* sgs_b = nullptr;             // FP32 biases
    bool sgs_ready = false;
    // training-sample filter (allocated on first use)
    int f_factor = 0;
    FFTPlan fplx, fply;
    void *ftwx = nullptr, *ftwy = nullptr, *fspec = nullptr, *fq = nullptr;
    Timing tim;
    std::vector<void*> allocs;
    // collocated SRK stepper (scheme TC_SCHEME_SRK; P.scheme holds TC_SCHEME_AB3
    // for the shared kernels)
    bool srk = false;
    int srk_stab = 0;
    double alpha_tau = 0;
    ImexTable imex;
    void* sK[kMaxImexStages] = {};     // Kuhat per stage (real3)
    void* smujt[kMaxImexStages] = {};  // mujtilde per stage
    void* sqj[kMaxImexStages] = {};    // qj per stage (StabType 1)
    void* snu1 = nullptr;              // nu1tilde (methods of type CK)
    void* swork = nullptr;             // stage velocity u, then buf_vector (real3)
    void *sbs = nullptr, *spjhat = nullptr;
    // captured steps, replayed when tc_step is called again with the same
    // (current slot, order, tau): one graph launch instead of ~10 kernel launches
    bool use_graphs = false;
    std::map<std::tuple<int, int, unsigned long long>, hipGraphExec_t> graphs;
    size_t graph_cap = 48;  // captured steps kept; TC_GRAPH_CACHE_MAX (>= 1) overrides

    template <typename T>
    Mesh<T> mesh() const {
        Mesh<T> m;
        m.nx = nx;
        m.ny = ny;
        m.nz = nz;
        m.nxy = nxy;
        m.n3d = (int)n3d;
        m.nxh = nxh;
        m.hx = (T)P.hx;
        m.hy = (T)P.hy;
        m.inv_hx = (T)(1.0 / P.hx);
        m.inv_hy = (T)(1.0 / P.hy);
        m.visc = (T)P.visc;
        m.sour = (T)P.sour_dpdx;
        m.inv_len = (T)P.inv_length_scale;
        m.zlo = (T)hm.ZZ.front();
        m.zhi = (T)hm.ZZ.back();
        const T* b = (const T*)dmesh;
        m.ZZ = b;
        m.DZ = b + mstride;
        m.DZW = b + 2 * mstride;
        m.GC1 = b + 3 * mstride;
        m.GC2 = b + 4 * mstride;
        m.GCZ = b + 5 * mstride;
        m.cellh = b + 6 * mstride;
        m.IDZ = b + 7 * mstride;
        m.IHDN = b + 8 * mstride;
        m.IHUP = b + 9 * mstride;
        m.IDZM = b + 10 * mstride;
        m.IVOL = b + 11 * mstride;
        m.IHC = b + 12 * mstride;
        m.xcd_swz = (size_t)nxy * 3 * sizeof(T) * 4 <= ((size_t)4 << 20) ? 1 : 0;
        return m;
    }

    void tbeg(int k, hipEvent_t& a) {
        a = nullptr;
        if (!tim.on) return;
        a = tim.get();
        if (a) hipEventRecord(a, s);
    }
    void tend(int k, hipEvent_t a) {
        if (!tim.on || !a) return;
        hipEvent_t b = tim.get();
        if (!b) return;
        hipEventRecord(b, s);
        tim.recs.push_back({k, a, b});
    }
};

namespace {

template <typename T>
int dalloc(tc_engine* e, T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&q, n * sizeof(T)));
    e->allocs.push_back(q);  // freed with the engine, also if the clear below fails
    *p = (T*)q;
    HIPCHK(hipMemset(q, 0, n * sizeof(T)));
    return TC_OK;
}

int dalloc_bytes(tc_engine* e, void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    HIPCHK(hipMalloc(p, bytes));
    e->allocs.push_back(*p);
    HIPCHK(hipMemset(*p, 0, bytes));
    return TC_OK;
}

template <typename T>
int upload_as(tc_engine* e, void* dst, const double* src, size_t n) {
    std::vector<T> tmp(n);
    for (size_t i = 0; i < n; i++) tmp[i] = (T)src[i];
    HIPCHK(hipMemcpy(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
    return TC_OK;
}

void* field_ptr(tc_engine* e, int kind, int slot) {
    if (slot < 0 || slot > 3) return nullptr;
    switch (kind) {
        case TC_VEL: return e->vel[slot];
        case TC_FLUX: return e->flux[slot];
        case TC_PRES: return e->pres[slot];
        case TC_NUT: return slot == 0 ? e->nut : nullptr;
        case TC_TAU: return slot == 0 ? e->tau : nullptr;
        default: return nullptr;
    }
}
bool is_vec(int kind) { return kind == TC_VEL || kind == TC_FLUX; }
bool is_scal(int kind) { return kind == TC_PRES || kind == TC_NUT; }
size_t ncomp(int kind) { return kind == TC_TAU ? 6 : is_vec(kind) ? 3 : 1; }

// FFT twiddles exp(-2 pi i k / n) in T.
template <typename T>
int upload_twiddles(tc_engine* e, int n, void** dst) {
    std::vector<Cplx<T>> t(n);
    for (int k = 0; k < n; k++) {
        const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)n;
        t[k] = {(T)std::cos(a), (T)(-std::sin(a))};
    }
    if (int r = dalloc_bytes(e, dst, n * sizeof(Cplx<T>))) return r;
    HIPCHK(hipMemcpy(*dst, t.data(), n * sizeof(Cplx<T>), hipMemcpyHostToDevice));
    return TC_OK;
}

template <typename T>
int setup_device(tc_engine* e) {
    const size_t n3 = e->n3d;
    for (int i = 0; i < 4; i++) {
        T* p;
        if (int r = dalloc(e, &p, 3 * n3)) return r;
        e->vel[i] = p;
        if (int r = dalloc(e, &p, 3 * n3)) return r;
        e->flux[i] = p;
        if (int r = dalloc(e, &p, n3)) return r;
        e->pres[i] = p;
    }
    T* p;
    if (e->P.turb_visc) {
        if (int r = dalloc(e, &p, n3)) return r;
        e->nut = p;
    }
    if (int r = dalloc(e, &p, n3)) return r;
    e->dp = p;
    if (e->P.sgs_model) {
        if (int r = dalloc(e, &p, 6 * n3)) return r;
        e->tau = p;
        if (int r = dalloc(e, &e->sgs_w, (size_t)kWBlock)) return r;
        if (int r = dalloc(e, &e->sgs_b, (size_t)kBBlock)) return r;
    }
    if (e->srk) {
        const int S = e->imex.s;
        for (int i = 0; i < S; i++) {
            if (int r = dalloc(e, &p, 3 * n3)) return r;
            e->sK[i] = p;
            if (int r = dalloc(e, &p, n3)) return r;
            e->smujt[i] = p;
            if (e->srk_stab) {
                if (int r = dalloc(e, &p, n3)) return r;
                e->sqj[i] = p;
            }
        }
        if (e->imex.type == TC_IMEX_TYPE_CK) {
            if (int r = dalloc(e, &p, 3 * n3)) return r;
            e->snu1 = p;
        }
        if (int r = dalloc(e, &p, 3 * n3)) return r;
        e->swork = p;
        if (int r = dalloc(e, &p, n3)) return r;
        e->sbs = p;
        if (int r = dalloc(e, &p, n3)) return r;
        e->spjhat = p;
    }
    if (int r = dalloc_bytes(e, &e->spec, (size_t)e->nm * e->nz * sizeof(Cplx<T>))) return r;
    // mesh arrays: ZZ, DZ, DZW, GC1, GC2, GCZ, cellh, then the per-plane
    // quotients IDZ, IHDN, IHUP, IDZM, IVOL, IHC, each mstride long
    const HostMesh& hm = e->hm;
    e->mstride = (size_t)e->nz + 8;
    std::vector<double> mh(kMeshArrays * e->mstride, 0.0);
    auto put = [&](int k, const std::vector<double>& v) {
        for (size_t i = 0; i < v.size() && i < e->mstride; i++) mh[k * e->mstride + i] = v[i];
    };
    put(0, hm.ZZ);
    put(1, hm.DZ);
    put(2, hm.DZW);
    put(3, hm.GC1);
    put(4, hm.GC2);
    put(5, hm.GCZ);
    put(6, hm.cellh);
    {
        // quotients in T, as the device would form them (stored exactly in double)
        const int nz = e->nz;
        const bool coll = e->P.scheme == TC_SCHEME_AB3;
        auto dzt = [&](int i) -> T { return (i < 0 || i >= (int)hm.DZ.size()) ? T(0) : (T)hm.DZ[i]; };
        auto q = [](T num, T den) -> T { return den == T(0) ? T(0) : num / den; };
        const T hx = (T)e->P.hx, hy = (T)e->P.hy;
        for (int iz = 0; iz < nz; iz++) {
            const T dz = dzt(iz), dzm = dzt(iz - 1), dzp = iz + 1 < nz ? dzt(iz + 1) : T(0);
            const T dzw = iz < (int)hm.DZW.size() ? (T)hm.DZW[iz] : T(0);
            mh[7 * e->mstride + iz] = (double)q(T(1), dz);
            mh[8 * e->mstride + iz] = (double)q(T(2), dz + dzm);
            mh[9 * e->mstride + iz] = (double)q(T(2), dz + dzp);
            mh[10 * e->mstride + iz] = (double)q(T(1), dzt(std::max(iz - 1, 0)));
            mh[11 * e->mstride + iz] = (double)q(T(-1), T(32) * (dz - T(1.0 / 9.0) * dzw) * hx * hy);
            mh[12 * e->mstride + iz] = (coll && iz > 0 && iz < nz - 1) ? (double)q(T(2), dzm + dz) : 0.0;
        }
    }
    if (int r = dalloc_bytes(e, &e->dmesh, mh.size() * sizeof(T))) return r;
    if (int r = upload_as<T>(e, e->dmesh, mh.data(), mh.size())) return r;
    // normalisation weights (NormalizePressure)
    std::vector<double> wpl(e->nz);
    for (int iz = 0; iz < e->nz; iz++) {
        if (e->P.scheme == TC_SCHEME_HW) wpl[iz] = hm.DZ[iz];
        else if (e->P.scheme == TC_SCHEME_VV) wpl[iz] = hm.KERG[iz] * hm.DZ[iz];
        else wpl[iz] = hm.cellh[iz] * e->P.hx * e->P.hy;
    }
    if (int r = dalloc(e, &e->d_wplane, e->nz)) return r;
    HIPCHK(hipMemcpy(e->d_wplane, wpl.data(), e->nz * sizeof(double), hipMemcpyHostToDevice));
    if (int r = dalloc_bytes(e, &e->d_kerg, e->nz * sizeof(T))) return r;
    if (int r = upload_as<T>(e, e->d_kerg, hm.KERG.data(), e->nz)) return r;
    // wall-normal factors
    const ZFactors& zf = e->zf;
    const size_t fz = (size_t)e->nz * zf.ncols;
    const int FW = zfac_fw(zf.W), UW = zfac_uw(zf.W);
    // packed per row: F[i][FW] = (M1..M_W, D), U[i][UW] = (U1..U_W), i = iz * ncols + col
    auto pack = [&](auto* F, auto* U, size_t n, auto idx) {
        for (size_t i = 0; i < n; i++) {
            const size_t j = idx(i);
            for (int c = 0; c < zf.W; c++) F[i * FW + c] = zf.M[c][j];
            F[i * FW + FW - 1] = zf.D[j];
            for (int c = 0; c < zf.W; c++) U[i * UW + c] = zf.U[c][j];
        }
    };
    if (int r = dalloc_bytes(e, &e->zfac, (FW + UW) * fz * sizeof(T))) return r;
    {
        std::vector<T> tmp((FW + UW) * fz);
        pack(tmp.data(), tmp.data() + FW * fz, fz, [](size_t i) { return i; });
        HIPCHK(hipMemcpy(e->zfac, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    if (int r = dalloc(e, &e->d_Wt, e->nz)) return r;
    if (int r = dalloc(e, &e->d_C, e->nz)) return r;
    if (int r = dalloc(e, &e->d_s00, e->nz)) return r;
    HIPCHK(hipMemcpy(e->d_Wt, zf.Wt.data(), e->nz * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_C, zf.C.data(), e->nz * sizeof(double), hipMemcpyHostToDevice));
    if (e->fp32 && !zf.R.empty()) {  // the bulk modes form their forward elimination in-lane
        std::vector<double> all(zf.R);
        all.insert(all.end(), zf.KX.begin(), zf.KX.end());
        all.insert(all.end(), zf.KY.begin(), zf.KY.end());
        if (int r = dalloc_bytes(e, &e->zinl, all.size() * sizeof(T))) return r;
        if (int r = upload_as<T>(e, (T*)e->zinl, all.data(), all.size())) return r;
    }
    if (e->fp32) {  // low-mode list and its double factors
        std::vector<int> low;
        for (int kx = 0; kx <= kLowK && kx < e->nxh; kx++)
            for (int ky = 0; ky < e->ny; ky++)
                if (ky <= kLowK || ky >= e->ny - kLowK) low.push_back(kx * e->ny + ky);
        e->nlow = (int)low.size();
        std::vector<int> lowidx(e->nm, -1);
        for (int li = 0; li < e->nlow; li++) lowidx[low[li]] = li;
        if (int r = dalloc(e, &e->d_lowidx, lowidx.size())) return r;
        HIPCHK(hipMemcpy(e->d_lowidx, lowidx.data(), lowidx.size() * sizeof(int), hipMemcpyHostToDevice));
        if (int r = dalloc(e, &e->d_lowmode, low.size())) return r;
        HIPCHK(hipMemcpy(e->d_lowmode, low.data(), low.size() * sizeof(int), hipMemcpyHostToDevice));
        const size_t lz = (size_t)e->nz * e->nlow;
        std::vector<double> lf((FW + UW) * lz);
        const int nlow = e->nlow, ncols = zf.ncols, ny = e->ny;
        pack(lf.data(), lf.data() + FW * lz, lz,
             [&](size_t i) { return (i / nlow) * ncols + zfac_col(low[i % nlow], ny); });
        if (int r = dalloc(e, &e->zfac_low, lf.size())) return r;
        HIPCHK(hipMemcpy(e->zfac_low, lf.data(), lf.size() * sizeof(double), hipMemcpyHostToDevice));
        if (int r = dalloc(e, &e->ylow, 2 * lz)) return r;
    }
    if (int r = upload_twiddles<T>(e, e->nx, &e->twx)) return r;
    if (int r = upload_twiddles<T>(e, e->ny, &e->twy)) return r;
    e->npartial = std::max<size_t>((size_t)kRedBlocks * 3, (size_t)flux_grid_blocks(e->nx, e->ny, e->nz));
    e->npartial = std::max<size_t>(e->npartial, (size_t)profiles_chunks(e->nxy) * e->nz * diag_record(e->P.scheme));
    e->npartial = std::max<size_t>(e->npartial, (size_t)project_diag_chunks(e->nxy) * e->nz * diag_record(e->P.scheme));
    if (int r = dalloc(e, &e->d_partial, e->npartial)) return r;
    if (int r = dalloc(e, &e->d_out, 16)) return r;
    if (int r = dalloc(e, &e->d_shift, 4)) return r;
    if (int r = dalloc(e, &e->d_prof, (size_t)e->nz * 12)) return r;
    for (int k = 0; k < 3; k++) {
        HIPCHK(hipHostMalloc((void**)&e->h_diag[k], (size_t)e->nz * diag_record(e->P.scheme) * sizeof(double),
                             hipHostMallocMapped));
        HIPCHK(hipHostGetDevicePointer((void**)&e->hd_diag[k], e->h_diag[k], 0));
    }
    for (int k = 0; k < 2; k++) HIPCHK(hipEventCreateWithFlags(&e->diag_ev[k], hipEventDisableTiming));
    return TC_OK;
}

template <typename T>
ZFacDev<T> zfac_dev(tc_engine* e) {
    ZFacDev<T> f;
    f.W = e->zf.W;
    f.nz = e->nz;
    f.ncols = e->zf.ncols;
    f.nmodes = e->nm;
    f.ny = e->ny;
    const size_t fz = (size_t)e->nz * f.ncols;
    f.F = (const T*)e->zfac;
    f.U = f.F + zfac_fw(f.W) * fz;
    if (e->zinl) {
        f.R = (const T*)e->zinl;
        f.KX = f.R + e->zf.R.size();
        f.KY = f.KX + e->zf.KX.size();
    }
    return f;
}

ZFacDev<double> zfac_low_dev(tc_engine* e) {
    ZFacDev<double> f;
    f.W = e->zf.W;
    f.nz = e->nz;
    f.ncols = e->nlow;
    const size_t lz = (size_t)e->nz * e->nlow;
    f.F = e->zfac_low;
    f.U = f.F + zfac_fw(f.W) * lz;
    return f;
}

// UnapplyGradient: p = L^-1 D v (ApplyDivergence + LinearSystemSolveFFT incl.
// NormalizePressure, staggered/func_global.h:214-248, :290-293).
template <typename T>
int unapply(tc_engine* e, const T* v, T* x, const CfrShiftArgs& cfr = CfrShiftArgs()) {
    hipEvent_t ev;
    Cplx<T>* spec = (Cplx<T>*)e->spec;
    const int nrows = e->ny * e->nz;
    e->tbeg(TC_K_FFT_X, ev);
    HIPCHK(launch_fft_x_fwd_div<T>(e->s, e->plx, (const Cplx<T>*)e->twx, e->mesh<T>(), e->P.scheme, v, spec, cfr));
    e->tend(TC_K_FFT_X, ev);
    e->tbeg(TC_K_FFT_Y, ev);
    HIPCHK(launch_fft_y<T>(e->s, e->ply, (const Cplx<T>*)e->twy, spec, e->nxh, e->nz, 0, nullptr, nullptr));
    e->tend(TC_K_FFT_Y, ev);
    e->tbeg(TC_K_ZSOLVE, ev);
    ZNorm norm;
    norm.Wt = e->d_Wt;
    norm.C = e->d_C;
    norm.s00 = e->d_s00;
    if (e->fp32) {
        HIPCHK(launch_zsolve<T>(e->s, spec, zfac_dev<T>(e), zfac_low_dev(e), e->d_lowidx, e->d_lowmode, e->nlow, e->ylow,
                                norm));
    } else {
        HIPCHK(launch_zsolve<T>(e->s, spec, zfac_dev<T>(e), ZFacDev<double>{}, nullptr, nullptr, 0, nullptr, norm));
    }
    e->tend(TC_K_ZSOLVE, ev);
    e->tbeg(TC_K_IFFT_Y, ev);
    HIPCHK(launch_fft_y<T>(e->s, e->ply, (const Cplx<T>*)e->twy, spec, e->nxh, e->nz, 1, e->d_s00, e->d_C));
    e->tend(TC_K_IFFT_Y, ev);
    e->tbeg(TC_K_IFFT_X, ev);
    HIPCHK(launch_fft_x_inv<T>(e->s, e->plx, (const Cplx<T>*)e->twx, spec, x, e->ny, nrows,
                               (T)(1.0 / ((double)e->nx * e->ny))));
    e->tend(TC_K_IFFT_X, ev);
    return TC_OK;
}

MlpDev sgs_view(const tc_engine* e) {
    MlpDev d;
    d.w = e->sgs_w;
    d.b = e->sgs_b;
    return d;
}

template <typename T>
int step(tc_engine* e, double tau, int cur, int order) {
    const int next = (cur + 1) % 4, prev = (cur + 3) % 4, prev2 = (cur + 2) % 4;
    const double bd0 = (order == 3) ? (1.5 + 1.0 / 3.0) : (order == 2) ? 1.5 : 1.0;
    const Mesh<T> m = e->mesh<T>();
    const bool stab = e->P.scheme == TC_SCHEME_AB3 && e->P.stab_type != 0;
    const bool cfr = e->P.cfr_ubulk > 0;
    hipEvent_t ev;
    if (e->P.turb_visc) {
        e->tbeg(TC_K_TURB_VISC, ev);
        HIPCHK(launch_turb_visc<T>(e->s, m, e->P.scheme, (const T*)e->vel[cur], (T*)e->nut));
        e->tend(TC_K_TURB_VISC, ev);
    }
    if (e->P.sgs_model) {
        if (!e->sgs_ready) return fail(TC_ERR_STATE, "learned closure weights are not set (tc_set_sgs_params)");
        e->tbeg(TC_K_TURB_VISC, ev);
        HIPCHK(launch_sgs_tau<T>(e->s, m, sgs_view(e), (const T*)e->vel[cur], (T*)e->tau, e->P.sgs_model - 1));
        e->tend(TC_K_TURB_VISC, ev);
    }
    FluxArgs<T> a;
    a.scheme = e->P.scheme;
    a.visc_type = e->P.visc_type;
    a.turb = e->P.turb_visc;
    a.order = order;
    a.stab = stab;
    a.vel = (const T*)e->vel[cur];
    a.nut = (const T*)e->nut;
    a.fcur = (T*)e->flux[cur];
    a.vprev = (const T*)e->vel[prev];
    a.vprev2 = (const T*)e->vel[prev2];
    a.fprev = (const T*)e->flux[prev];
    a.fprev2 = (const T*)e->flux[prev2];
    a.fnext = (T*)e->flux[next];
    a.pcur = (const T*)e->pres[cur];
    a.inv_tau = (T)(1.0 / tau);
    a.cfr_partial = cfr ? e->d_partial : nullptr;
    a.tau = (const T*)e->tau;
    int nblk = 0;
    e->tbeg(TC_K_FLUX, ev);
    HIPCHK(launch_flux<T>(e->s, m, a, &nblk));
    e->tend(TC_K_FLUX, ev);
    T* ptarget = stab ? (T*)e->dp : (T*)e->pres[next];
    CfrShiftArgs cs;  // the flow-rate shift, formed by the divergence transform's block 0
    if (cfr) {
        cs.partial = e->d_partial;
        cs.nb = nblk;
        cs.scale = tau / bd0;
        cs.ub_vol = e->P.cfr_ubulk * e->hm.area;
        cs.vol_u = e->hm.vol_u;
        cs.shift = e->d_shift;
    }
    if (int r = unapply<T>(e, (const T*)e->flux[next], ptarget, cs)) return r;
    ProjectArgs<T> pa;
    pa.scheme = e->P.scheme;
    pa.p = ptarget;
    pa.r = (const T*)e->flux[next];
    pa.vnext = (T*)e->vel[next];
    pa.tau_over_bd0 = (T)(tau / bd0);
    pa.shift = cfr ? e->d_shift : nullptr;
    if (stab) {
        pa.pcur = (const T*)e->pres[cur];
        pa.pnext = (T*)e->pres[next];
    }
    if (e->step_diag) {
        pa.diag = e->d_partial;
        pa.diag_out = e->hd_diag[e->diag_cur];
        pa.visc_fac = diag_visc_fac(e->P.scheme, e->P.visc_type);
        pa.nut = (const T*)e->nut;
    }
    e->tbeg(TC_K_PROJECT, ev);
    HIPCHK(launch_project<T>(e->s, m, pa));
    e->tend(TC_K_PROJECT, ev);
    return TC_OK;
}

template <typename T>
int normalize(tc_engine* e, T* p) {
    const Mesh<T> m = e->mesh<T>();
    RedArgs<T> r;
    r.op = RED_PSUM;
    r.scheme = e->P.scheme;
    r.p = p;
    r.wplane = e->d_wplane;
    HIPCHK(launch_reduce<T>(e->s, m, r, e->d_partial, e->d_out));
    double coef;
    const T* w = nullptr;
    const double nxy = (double)e->nxy;
    if (e->P.scheme == TC_SCHEME_HW) coef = 1.0 / ((e->hm.ZZ[e->nz] - e->hm.ZZ[0]) * nxy);
    else if (e->P.scheme == TC_SCHEME_VV) {
        coef = 1.0 / (e->hm.KERG_norm2 * nxy);
        w = (const T*)e->d_kerg;
    } else if (e->srk) coef = 1.0 / e->hm.area;  // zero mean (collocated_SRK/func_global.h:130-137)
    else coef = e->P.hx * e->P.hy / e->hm.area;
    HIPCHK(launch_add_scaled_plane<T>(e->s, m, p, e->d_out, coef, w));
    return TC_OK;
}

// One step of the collocated SRK solver: data::Step (collocated_SRK/func_global.h:317-408)
// with ImplicitTerm = 0 and ImplicitStage = copy (:310-316), so Ku is zero and the
// stage velocity is buf_vector itself. Launches per stage: the stage velocity
// (which also finishes the previous stage: Kuhat -= G p_j, q_j, from p_j =
// mujtilde + buf_scalar formed at the nodes it reads), the explicit term, the
// pressure right-hand side (which applies the previous stages' mujtilde update
// of :391 as it reads them), buf_vector, and the five kernels of the pressure
// solve. The final update finishes the last stage the same way and writes p and
// dp/dt into the caller's slots. Terms with a zero Butcher coefficient are not read.
template <typename T>
int srk_step(tc_engine* e, double tau, int vs, int ps, int qs) {
    const ImexTable& tb = e->imex;
    const int S = tb.s;
    const double ass = tb.I[S * S - 1], inv_ass = 1.0 / ass;
    const double tau_stage = tau * ass, inv_tau_stage = 1.0 / tau_stage;
    const double alpha_tau = e->alpha_tau, alpha = alpha_tau / tau;
    const bool ck = tb.type == TC_IMEX_TYPE_CK, stab = e->srk_stab != 0;
    const int j0 = tb.type == TC_IMEX_TYPE_A ? 0 : 1;
    const Mesh<T> m = e->mesh<T>();
    T* vel = (T*)e->vel[vs];
    T* p = (T*)e->pres[ps];
    T* q = stab ? (T*)e->pres[qs] : nullptr;
    T* K[kMaxImexStages];
    for (int i = 0; i < S; i++) K[i] = (T*)e->sK[i];
    T* work = (T*)e->swork;
    T* nu1 = (T*)e->snu1;
    T* bs = (T*)e->sbs;
    T* pjhat = (T*)e->spjhat;
    hipEvent_t ev;
    auto flux = [&](const T* u, T* out) -> int {
        FluxArgs<T> a;
        a.scheme = TC_SCHEME_AB3;
        a.visc_type = 0;
        a.order = 0;
        a.vel = u;
        a.fcur = out;
        e->tbeg(TC_K_FLUX, ev);
        HIPCHK(launch_flux<T>(e->s, m, a, nullptr));
        e->tend(TC_K_FLUX, ev);
        return TC_OK;
    };
    auto vec = [&](const SrkVec<T>& a) -> int {
        e->tbeg(TC_K_PROJECT, ev);
        HIPCHK(launch_srk_vec<T>(e->s, m, a));
        e->tend(TC_K_PROJECT, ev);
        return TC_OK;
    };
    // sum_{j<n} c_row[j] Kuhat[j] with the zero coefficients left out
    auto combo = [&](SrkVec<T>& a, const double* row, int n) {
        a.nk = 0;
        for (int j = 0; j < n; j++)
            if (row[j] != 0.0) {
                a.K[a.nk] = K[j];
                a.c[a.nk++] = (T)(tau * row[j]);
            }
    };
    // finish stage i: Kuhat[i] -= G p_i with p_i = mujtilde[i] + buf_scalar; q_i or dp/dt
    auto finish = [&](SrkVec<T>& a, int i, T* pj_out, T* qj_out) {
        a.KC = K[i];
        a.S = (const T*)e->smujt[i];
        a.S2 = bs;
        a.pj_out = pj_out;
        if (stab) {
            a.qj_out = qj_out;
            a.pjhat = pjhat;
            a.inv_tau_stage = (T)inv_tau_stage;
        }
    };
    if (j0 == 1) {  // Kuhat[0] = F(u_n) - G p_n; CK: nu1tilde = Kuhat[0] + alpha u_n
        if (int r = flux(vel, K[0])) return r;
        SrkVec<T> a;
        a.KC = K[0];
        a.S = p;
        a.KCout = K[0];
        if (ck) {
            a.out = nu1;
            a.ckc = T(1);
            a.A = vel;
            a.sa = (T)alpha;
        }
        if (int r = vec(a)) return r;
    }
    for (int i = j0; i < S; i++) {
        {  // u = u_n + sum_j tau E_ij Kuhat[j], finishing stage i - 1 on the way
            SrkVec<T> a;
            a.out = work;
            a.X = vel;
            const bool fuse = i > j0;
            combo(a, &tb.E[i * S], fuse ? i - 1 : i);
            if (fuse) {
                finish(a, i - 1, nullptr, (T*)e->sqj[i - 1]);
                a.KCout = K[i - 1];
                a.ckc = (T)(tau * tb.E[i * S + i - 1]);
            }
            if (int r = vec(a)) return r;
        }
        if (int r = flux(work, K[i])) return r;
        {
            SrkPre<T> a;
            a.stab = stab;
            a.ck = ck;
            a.p = p;
            a.q = q;
            a.bs = bs;
            a.pjhat = pjhat;
            a.m0 = (T)(1. - alpha_tau * tb.I[i * S]);
            a.ph0 = (T)(tau * tb.I[i * S]);
            a.tau_stage = (T)tau_stage;
            a.clast = stab ? (T)(alpha * tau_stage * tau_stage) : (T)(alpha * tau_stage);
            a.cfix = (T)(alpha * tau_stage);
            a.nj = i - j0;
            for (int j = j0; j < i; j++) {
                a.mujt[j - j0] = (const T*)e->smujt[j];
                a.cm[j - j0] = (T)(inv_ass * tb.I[i * S + j]);
                a.qj[j - j0] = (const T*)e->sqj[j];
                a.cq[j - j0] = (T)(tau * tb.I[i * S + j]);
            }
            e->tbeg(TC_K_OTHER, ev);
            HIPCHK(launch_srk_pre<T>(e->s, m, a));
            e->tend(TC_K_OTHER, ev);
        }
        {  // buf_vector = Kuhat[i] - G bs + alpha u_n (- d_i nu1tilde)
            SrkVec<T> a;
            a.out = work;
            a.KC = K[i];
            a.S = bs;
            a.ckc = T(1);
            a.A = vel;
            a.sa = (T)alpha;
            if (ck) {
                a.B = nu1;
                a.sb = (T)(-tb.d[i]);
            }
            if (int r = vec(a)) return r;
        }
        if (int r = unapply<T>(e, work, (T*)e->smujt[i])) return r;
    }
    // u_{n+1} = u_n + sum_j tau b_j Kuhat[j], finishing the last stage: p = p_s, dp/dt = q_s
    SrkVec<T> a;
    a.out = vel;
    a.X = vel;
    combo(a, &tb.E[S * S], S - 1);
    finish(a, S - 1, p, q);
    a.ckc = (T)(tau * tb.E[S * S + S - 1]);
    return vec(a);
}

template <typename T>
int reduce_to_host(tc_engine* e, RedArgs<T> r, double* out, int n) {
    HIPCHK(launch_reduce<T>(e->s, e->mesh<T>(), r, e->d_partial, e->d_out));
    HIPCHK(hipMemcpyAsync(out, e->d_out, n * sizeof(double), hipMemcpyDeviceToHost, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
    return TC_OK;
}

template <typename T>
int project(tc_engine* e, T* v) {
    if (int r = unapply<T>(e, v, (T*)e->dp)) return r;
    ProjectArgs<T> pa;
    pa.scheme = e->P.scheme;
    pa.p = (const T*)e->dp;
    pa.r = v;
    pa.vnext = v;
    pa.tau_over_bd0 = T(1);
    HIPCHK(launch_project<T>(e->s, e->mesh<T>(), pa));
    return TC_OK;
}

// Training samples of the learned closure from one collocated velocity field:
// sharp spectral cut-off by `factor` in x and y, sampled on the coarse grid.
template <typename T>
int sgs_samples(tc_engine* e, const T* v, int factor, float* x, float* y) {
    const int nxc = e->nx / factor, nyc = e->ny / factor, nz = e->nz;
    const int nxhc = nxc / 2 + 1;
    if (e->f_factor != factor) {
        if (e->f_factor != 0) return fail(TC_ERR_STATE, "the sample filter factor cannot change on one engine");
        if (!make_plan(nxc, e->fplx) || !make_plan(nyc, e->fply))
            return fail(TC_ERR_ARG, "coarse sizes must have only the prime factors 2, 3 and 5");
        if (int r = upload_twiddles<T>(e, nxc, &e->ftwx)) return r;
        if (int r = upload_twiddles<T>(e, nyc, &e->ftwy)) return r;
        if (int r = dalloc_bytes(e, &e->fspec, (size_t)nxhc * nyc * nz * sizeof(Cplx<T>))) return r;
        if (int r = dalloc_bytes(e, &e->fq, 9 * (size_t)nxc * nyc * nz * sizeof(T))) return r;
        e->f_factor = factor;
    }
    const Mesh<T> m = e->mesh<T>();
    Cplx<T>* spec = (Cplx<T>*)e->spec;
    Cplx<T>* cs = (Cplx<T>*)e->fspec;
    T* q = (T*)e->fq;
    const size_t fs = (size_t)nxc * nyc * nz;
    static const int ca[9] = {0, 1, 2, 0, 1, 2, 0, 0, 1}, cb[9] = {-1, -1, -1, 0, 1, 2, 1, 2, 2};
    for (int f = 0; f < 9; f++) {
        HIPCHK(launch_fft_x_fwd_q<T>(e->s, e->plx, (const Cplx<T>*)e->twx, v, ca[f], cb[f], e->ny, e->ny * nz, spec));
        HIPCHK(launch_fft_y<T>(e->s, e->ply, (const Cplx<T>*)e->twy, spec, e->nxh, nz, 0, nullptr, nullptr));
        HIPCHK(launch_spec_crop<T>(e->s, spec, e->nxh, e->ny, cs, nxhc, nyc, nz));
        HIPCHK(launch_fft_y<T>(e->s, e->fply, (const Cplx<T>*)e->ftwy, cs, nxhc, nz, 1, nullptr, nullptr));
        HIPCHK(launch_fft_x_inv<T>(e->s, e->fplx, (const Cplx<T>*)e->ftwx, cs, q + f * fs, nyc, nyc * nz,
                                   (T)(1.0 / ((double)e->nx * e->ny))));
    }
    HIPCHK(launch_sgs_samples<T>(e->s, m, q, nxc, nyc, (T)(e->P.hx * factor), (T)(e->P.hy * factor), x, y));
    HIPCHK(hipStreamSynchronize(e->s));
    return TC_OK;
}

#define DISPATCH(e, call) ((e)->fp32 ? call<float> : call<double>)

int check(tc_engine* e) {
    if (!e) return fail(TC_ERR_ARG, "engine is NULL");
    if (hipSetDevice(e->dev) != hipSuccess) return fail(TC_ERR_HIP, "hipSetDevice failed");
    return TC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* tc_last_error(void) { return g_err.c_str(); }

int tc_mesh_array(const tc_params* pin, const char* name, double* out, int cap) {
    if (!pin || !name) return -1;
    const tc_params kp = as_kernel_params(*pin);
    const tc_params* p = &kp;
    HostMesh m;
    if (!build_mesh(*p, m).empty()) return -1;
    std::vector<double> v;
    if (!strcmp(name, "ZZ")) v = m.ZZ;
    else if (!strcmp(name, "DZ")) v.assign(m.DZ.begin(), m.DZ.begin() + (p->scheme == TC_SCHEME_AB3 ? p->nz - 1 : p->nz));
    else if (!strcmp(name, "MAT_A")) v = m.MAT_A;
    else if (!strcmp(name, "KERG_norm2") && p->scheme == TC_SCHEME_VV) v = {m.KERG_norm2};
    else if (!strncmp(name, "ZF_", 3)) {  // wall-normal elimination factors
        ZFactors f;
        if (!build_zfactors(*p, m, f).empty()) return -1;
        const char* q = name + 3;
        if (!strcmp(q, "M") || !strcmp(q, "U"))
            for (int c = 0; c < f.W; c++) v.insert(v.end(), (*q == 'M' ? f.M : f.U)[c].begin(), (*q == 'M' ? f.M : f.U)[c].end());
        else if (!strcmp(q, "D")) v = f.D;
        else if (!strcmp(q, "R")) v = f.R;
        else if (!strcmp(q, "KX")) v = f.KX;
        else if (!strcmp(q, "KY")) v = f.KY;
        else return -1;
    } else if (p->scheme == TC_SCHEME_VV) {
        if (!strcmp(name, "DZW")) v = m.DZW;
        else if (!strcmp(name, "GC1")) v = m.GC1;
        else if (!strcmp(name, "GC2")) v = m.GC2;
        else if (!strcmp(name, "GCZ")) v = m.GCZ;
        else if (!strcmp(name, "MAT_D")) v.assign(m.MAT_D, m.MAT_D + 6);
        else if (!strcmp(name, "KERG")) v = m.KERG;
        else return -1;
    } else return -1;
    if (out)
        for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
    return (int)v.size();
}

int tc_make_initial_field(const tc_params* pin, unsigned long long seed, double ubulk, double noise_rms, double* out) {
    if (!pin || !out) return fail(TC_ERR_ARG, "NULL argument");
    const tc_params p = as_kernel_params(*pin);
    HostMesh m;
    std::string err = build_mesh(p, m);
    if (!err.empty()) return fail(TC_ERR_ARG, err);
    make_initial_field(p, m, seed, ubulk, noise_rms, out);
    return TC_OK;
}

int tc_create(const tc_params* pin, int device, tc_engine** out) {
    if (!pin || !out) return fail(TC_ERR_ARG, "NULL argument");
    *out = nullptr;
    ImexTable imex;
    if (pin->scheme == TC_SCHEME_SRK) {
        if (pin->visc_type != 0 || pin->turb_visc || pin->sgs_model || pin->cfr_ubulk > 0)
            return fail(TC_ERR_ARG, "the SRK solver has the FD viscous term only, no turbulence model and no flow-rate control");
        if (pin->stab_type != 0 && pin->stab_type != 1) return fail(TC_ERR_ARG, "StabType must be 0 or 1");
        if (!imex_table(pin->imex, imex)) return fail(TC_ERR_ARG, "unknown IMEX method");
    }
    const tc_params kp = as_kernel_params(*pin);
    const tc_params* p = &kp;
    if (p->precision != TC_FP64 && p->precision != TC_FP32_BULK) return fail(TC_ERR_ARG, "unknown precision");
    if (p->visc_type < 0 || p->visc_type > 1) return fail(TC_ERR_ARG, "visc_type must be 0 or 1");
    if (p->scheme != TC_SCHEME_AB3 && p->turb_visc && p->visc_type == 0)
        return fail(TC_ERR_ARG, "tViscApproximation::CROSS does not work for a turbulence model");
    if (p->sgs_model < 0 || p->sgs_model > 3) return fail(TC_ERR_ARG, "sgs_model must be 0, 1, 2 or 3");
    if (p->sgs_model && (p->scheme != TC_SCHEME_AB3 || p->turb_visc))
        return fail(TC_ERR_ARG, "the learned closure runs on the collocated scheme, without nu_t");
    const double n3 = (double)p->nx * p->ny * p->nz;
    if (n3 * 3 > 2.0e9) return fail(TC_ERR_ARG, "grid too large for 32-bit cell indexing");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(TC_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(TC_ERR_ARG, "DEVICE ID OUT OF RANGE");
    std::unique_ptr<tc_engine> e(new tc_engine());
    e->P = *p;
    e->dev = device;
    std::string err = build_mesh(*p, e->hm);
    if (!err.empty()) return fail(TC_ERR_ARG, err);
    e->zz = e->hm.ZZ;
    e->P.zz = e->zz.data();
    err = build_zfactors(e->P, e->hm, e->zf);
    if (!err.empty()) return fail(TC_ERR_NUMERIC, err);
    if (pin->scheme == TC_SCHEME_SRK) {
        e->srk = true;
        e->srk_stab = pin->stab_type;
        e->imex = imex;
        e->alpha_tau = pin->alpha_tau > 0 ? pin->alpha_tau : 0.5 * imex.alpha_tau_max[pin->stab_type];
        // zero-mean normalisation: the AB3 shift without its hx*hy factor
        for (double& c : e->zf.C) c /= p->hx * p->hy;
    }
    if (!make_plan(p->nx, e->plx) || !make_plan(p->ny, e->ply)) return fail(TC_ERR_ARG, "nx and ny must have only the prime factors 2, 3 and 5");
    if (p->nx > 2048) return fail(TC_ERR_ARG, "nx must be at most 2048 (lines of the x transforms)");
    e->fp32 = p->precision == TC_FP32_BULK;
    e->rsz = e->fp32 ? 4 : 8;
    e->nx = p->nx;
    e->ny = p->ny;
    e->nz = p->nz;
    e->nxy = p->nx * p->ny;
    e->nxh = p->nx / 2 + 1;
    e->nm = e->nxh * p->ny;
    e->n3d = (size_t)e->nxy * p->nz;
    if (const char* cap = std::getenv("TC_GRAPH_CACHE_MAX")) {
        const long c = std::strtol(cap, nullptr, 10);
        if (c >= 1) e->graph_cap = (size_t)c;
    }
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
    int r = e->fp32 ? setup_device<float>(e.get()) : setup_device<double>(e.get());
    if (r) {
        tc_destroy(e.release());
        return r;
    }
    if (const hipError_t se = hipDeviceSynchronize(); se != hipSuccess) {
        tc_destroy(e.release());  // frees what setup_device allocated
        return fail(TC_ERR_HIP, std::string("hipDeviceSynchronize: ") + hipGetErrorString(se));
    }
    *out = e.release();
    return TC_OK;
}

int tc_destroy(tc_engine* e) {
    if (!e) return TC_OK;
    hipSetDevice(e->dev);
    if (e->s) hipStreamSynchronize(e->s);
    for (auto& rc : e->tim.recs) {
        hipEventDestroy(rc.a);
        hipEventDestroy(rc.b);
    }
    for (auto ev : e->tim.pool) hipEventDestroy(ev);
    for (auto& g : e->graphs) hipGraphExecDestroy(g.second);
    for (void* p : e->allocs) hipFree(p);
    for (int k = 0; k < 3; k++)
        if (e->h_diag[k]) hipHostFree(e->h_diag[k]);
    for (int k = 0; k < 2; k++)
        if (e->diag_ev[k]) hipEventDestroy(e->diag_ev[k]);
    if (e->s) hipStreamDestroy(e->s);
    delete e;
    return TC_OK;
}

int tc_real_size(const tc_engine* e) { return e ? (int)e->rsz : 0; }

int tc_sync(tc_engine* e) {
    if (int r = check(e)) return r;
    HIPCHK(hipStreamSynchronize(e->s));
    return TC_OK;
}

int tc_set_field(tc_engine* e, int kind, int slot, const void* host, size_t bytes) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* d = field_ptr(e, kind, slot);
    if (!d || !host) return fail(TC_ERR_ARG, "bad field");
    const size_t need = ncomp(kind) * e->n3d * e->rsz;
    if (bytes != need) return fail(TC_ERR_ARG, "size mismatch: expected " + std::to_string(need) + " bytes");
    HIPCHK(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
    return TC_OK;
}

int tc_get_field(tc_engine* e, int kind, int slot, void* host, size_t bytes) {
    if (int r = check(e)) return r;
    void* d = field_ptr(e, kind, slot);
    if (!d || !host) return fail(TC_ERR_ARG, "bad field");
    const size_t need = ncomp(kind) * e->n3d * e->rsz;
    if (bytes != need) return fail(TC_ERR_ARG, "size mismatch: expected " + std::to_string(need) + " bytes");
    HIPCHK(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
    return TC_OK;
}

int tc_field_device_ptr(tc_engine* e, int kind, int slot, void** dev_ptr) {
    if (!e || !dev_ptr) return fail(TC_ERR_ARG, "NULL argument");
    void* d = field_ptr(e, kind, slot);
    if (!d) return fail(TC_ERR_ARG, "bad field");
    e->drop_records();  // the caller may write through it
    *dev_ptr = d;
    return TC_OK;
}

}  // extern "C"

namespace {

// Runs one step through the engine's graph cache: captured on first use of
// (a, b, tau), replayed afterwards.
template <class F>
int run_graph(tc_engine* e, int a, int b, double tau, F&& run) {
    unsigned long long tb;
    std::memcpy(&tb, &tau, sizeof tb);
    const auto key = std::make_tuple(a, b, tb);
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
        if (e->graphs.size() >= e->graph_cap) {  // step sizes changed many times: start over
            // a step launched from one of these graphs may still run (the driver
            // launches the next step before it reads the last one's record)
            HIPCHK(hipStreamSynchronize(e->s));
            for (auto& g : e->graphs) hipGraphExecDestroy(g.second);
            e->graphs.clear();
        }
        HIPCHK(hipStreamBeginCapture(e->s, hipStreamCaptureModeRelaxed));
        const int r = run();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(e->s, &g);
        if (r) {
            if (g) hipGraphDestroy(g);
            return r;
        }
        HIPCHK(ec);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HIPCHK(ei);
        it = e->graphs.emplace(key, x).first;
    }
    HIPCHK(hipGraphLaunch(it->second, e->s));
    return TC_OK;
}

}  // namespace

extern "C" {

int tc_step(tc_engine* e, double t, double tau, int cur, int bdf_order) {
    (void)t;
    if (int r = check(e)) return r;
    if (e->srk) return fail(TC_ERR_STATE, "the SRK solver steps with tc_srk_step");
    if (bdf_order < 1 || bdf_order > 3) return fail(TC_ERR_ARG, "Wrong bdf_order");
    if (cur < 0 || cur > 3) return fail(TC_ERR_ARG, "bad current_field");
    if (!(tau > 0)) return fail(TC_ERR_ARG, "tau must be positive");
    const int b = e->diag_cur;  // the record buffer this step writes; the other keeps the last step's
    e->diag_rec[b] = -1;
    auto run = [&]() { return e->fp32 ? step<float>(e, tau, cur, bdf_order) : step<double>(e, tau, cur, bdf_order); };
    int r;
    if (!e->use_graphs || e->tim.on) {
        r = run();
    } else {
        if (e->P.sgs_model && !e->sgs_ready)
            return fail(TC_ERR_STATE, "learned closure weights are not set (tc_set_sgs_params)");
        r = run_graph(e, cur, bdf_order + 8 * e->step_diag + 16 * b, tau, run);
    }
    if (r != TC_OK) {
        e->drop_records();
        return r;
    }
    if (e->step_diag) {
        HIPCHK(hipEventRecord(e->diag_ev[b], e->s));
        const int next = (cur + 1) % 4;
        if (e->diag_rec[1 - b] == next) e->diag_rec[1 - b] = -1;  // that slot is rewritten
        e->diag_rec[b] = next;
        e->diag_cur = 1 - b;
    }
    return r;
}

int tc_step_diagnostics(tc_engine* e, int on) {
    if (int r = check(e)) return r;
    if (e->srk) return fail(TC_ERR_STATE, "the SRK solver has no fused step diagnostics");
    e->step_diag = on != 0;
    e->drop_records();
    return TC_OK;
}

int tc_srk_step(tc_engine* e, double t, double tau, int vel_slot, int p_slot, int dpdt_slot) {
    (void)t;
    if (int r = check(e)) return r;
    if (!e->srk) return fail(TC_ERR_STATE, "tc_srk_step needs an engine of scheme TC_SCHEME_SRK");
    if (vel_slot < 0 || vel_slot > 3 || p_slot < 0 || p_slot > 3) return fail(TC_ERR_ARG, "bad slot");
    if (e->srk_stab && (dpdt_slot < 0 || dpdt_slot > 3 || dpdt_slot == p_slot)) return fail(TC_ERR_ARG, "bad dpdt slot");
    if (!(tau > 0)) return fail(TC_ERR_ARG, "tau must be positive");
    e->drop_records();
    auto run = [&]() {
        return e->fp32 ? srk_step<float>(e, tau, vel_slot, p_slot, dpdt_slot)
                       : srk_step<double>(e, tau, vel_slot, p_slot, dpdt_slot);
    };
    if (!e->use_graphs || e->tim.on) return run();
    const int slots = vel_slot + 4 * p_slot + 16 * (e->srk_stab ? dpdt_slot : 0);
    return run_graph(e, slots, -1, tau, run);
}

int tc_set_graphs(tc_engine* e, int on) {
    if (int r = check(e)) return r;
    e->use_graphs = on != 0;
    return TC_OK;
}

int tc_explicit_term(tc_engine* e, int in_kind, int in_slot, int out_kind, int out_slot) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* in = field_ptr(e, in_kind, in_slot);
    void* out = field_ptr(e, out_kind, out_slot);
    if (!in || !out || !is_vec(in_kind) || !is_vec(out_kind) || in == out) return fail(TC_ERR_ARG, "bad field");
    auto run = [&](auto zero) -> int {
        using T = decltype(zero);
        FluxArgs<T> a;
        a.scheme = e->P.scheme;
        a.visc_type = e->P.visc_type;
        a.turb = e->P.turb_visc;
        a.order = 0;
        a.vel = (const T*)in;
        a.nut = (const T*)e->nut;
        a.tau = (const T*)e->tau;
        a.fcur = (T*)out;
        hipEvent_t ev;
        e->tbeg(TC_K_FLUX, ev);
        HIPCHK(launch_flux<T>(e->s, e->mesh<T>(), a, nullptr));
        e->tend(TC_K_FLUX, ev);
        return TC_OK;
    };
    return e->fp32 ? run(0.0f) : run(0.0);
}

int tc_apply_gradient(tc_engine* e, int p_kind, int p_slot, int out_kind, int out_slot) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* p = field_ptr(e, p_kind, p_slot);
    void* out = field_ptr(e, out_kind, out_slot);
    if (!p || !out || !is_scal(p_kind) || !is_vec(out_kind)) return fail(TC_ERR_ARG, "bad field");
    if (e->fp32) HIPCHK(launch_gradient<float>(e->s, e->mesh<float>(), e->P.scheme, (const float*)p, (float*)out));
    else HIPCHK(launch_gradient<double>(e->s, e->mesh<double>(), e->P.scheme, (const double*)p, (double*)out));
    return TC_OK;
}

int tc_unapply_gradient(tc_engine* e, int v_kind, int v_slot, int p_kind, int p_slot) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* v = field_ptr(e, v_kind, v_slot);
    void* p = field_ptr(e, p_kind, p_slot);
    if (!v || !p || !is_vec(v_kind) || !is_scal(p_kind)) return fail(TC_ERR_ARG, "bad field");
    return e->fp32 ? unapply<float>(e, (const float*)v, (float*)p) : unapply<double>(e, (const double*)v, (double*)p);
}

int tc_normalize_pressure(tc_engine* e, int kind, int slot) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* p = field_ptr(e, kind, slot);
    if (!p || !is_scal(kind)) return fail(TC_ERR_ARG, "bad field");
    return e->fp32 ? normalize<float>(e, (float*)p) : normalize<double>(e, (double*)p);
}

int tc_recalc_turb_visc(tc_engine* e, int kind, int slot) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind)) return fail(TC_ERR_ARG, "bad field");
    if (e->P.sgs_model) {  // the learned closure's counterpart: tau from this field
        if (!e->sgs_ready) return fail(TC_ERR_STATE, "learned closure weights are not set (tc_set_sgs_params)");
        const int post = e->P.sgs_model - 1;
        if (e->fp32)
            HIPCHK(launch_sgs_tau<float>(e->s, e->mesh<float>(), sgs_view(e), (const float*)v, (float*)e->tau, post));
        else
            HIPCHK(launch_sgs_tau<double>(e->s, e->mesh<double>(), sgs_view(e), (const double*)v, (double*)e->tau, post));
        return TC_OK;
    }
    if (!e->nut) return TC_OK;  // RecalcTurbVisc returns when nu_t is not allocated
    if (e->fp32) HIPCHK(launch_turb_visc<float>(e->s, e->mesh<float>(), e->P.scheme, (const float*)v, (float*)e->nut));
    else HIPCHK(launch_turb_visc<double>(e->s, e->mesh<double>(), e->P.scheme, (const double*)v, (double*)e->nut));
    return TC_OK;
}

int tc_integral(tc_engine* e, int kind, int slot, double out[3]) {
    if (int r = check(e)) return r;
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind) || !out) return fail(TC_ERR_ARG, "bad field");
    auto run = [&](auto zero) -> int {
        using T = decltype(zero);
        RedArgs<T> a;
        a.op = RED_INTEGRAL;
        a.scheme = e->P.scheme;
        a.v = (const T*)v;
        return reduce_to_host<T>(e, a, out, 3);
    };
    return e->fp32 ? run(0.0f) : run(0.0);
}

int tc_kinetic_energy(tc_engine* e, int kind, int slot, double* out) {
    if (int r = check(e)) return r;
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind) || !out) return fail(TC_ERR_ARG, "bad field");
    auto run = [&](auto zero) -> int {
        using T = decltype(zero);
        RedArgs<T> a;
        a.op = RED_KE;
        a.scheme = e->P.scheme;
        a.v = (const T*)v;
        double o;
        if (int r = reduce_to_host<T>(e, a, &o, 1)) return r;
        *out = 0.5 * o;
        return TC_OK;
    };
    return e->fp32 ? run(0.0f) : run(0.0);
}

int tc_cfl(tc_engine* e, int kind, int slot, double tau, int mode, double* out) {
    if (int r = check(e)) return r;
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind) || !out) return fail(TC_ERR_ARG, "bad field");
    auto run = [&](auto zero) -> int {
        using T = decltype(zero);
        RedArgs<T> a;
        a.op = RED_CFL;
        a.scheme = e->P.scheme;
        a.v = (const T*)v;
        a.nut = (const T*)e->nut;
        a.tau = tau;
        a.mode = mode;
        a.visc_fac = diag_visc_fac(e->P.scheme, e->P.visc_type);
        return reduce_to_host<T>(e, a, out, 1);
    };
    return e->fp32 ? run(0.0f) : run(0.0);
}

int tc_profiles(tc_engine* e, int vel_kind, int vel_slot, int p_slot, double* host_av, size_t count) {
    if (int r = check(e)) return r;
    void* v = field_ptr(e, vel_kind, vel_slot);
    const int nv = e->P.scheme == TC_SCHEME_AB3 ? 10 : 12;
    void* p = e->P.scheme == TC_SCHEME_AB3 ? nullptr : field_ptr(e, TC_PRES, p_slot);
    if (!v || !is_vec(vel_kind) || !host_av) return fail(TC_ERR_ARG, "bad field");
    if (e->P.scheme != TC_SCHEME_AB3 && !p) return fail(TC_ERR_ARG, "bad pressure slot");
    if (count != (size_t)nv * e->nz) return fail(TC_ERR_ARG, "count must be nz*" + std::to_string(nv));
    if (e->fp32)
        HIPCHK(launch_profiles<float>(e->s, e->mesh<float>(), e->P.scheme, (const float*)v, (const float*)p, (const float*)e->nut,
                                      e->d_partial, e->d_prof));
    else
        HIPCHK(launch_profiles<double>(e->s, e->mesh<double>(), e->P.scheme, (const double*)v, (const double*)p,
                                       (const double*)e->nut, e->d_partial, e->d_prof));
    HIPCHK(hipMemcpyAsync(host_av, e->d_prof, count * sizeof(double), hipMemcpyDeviceToHost, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
    return TC_OK;
}

int tc_diagnostics(tc_engine* e, int vel_slot, int p_slot, double tau, int cfl_mode, double integral[3],
                   double* kinetic_energy, double* cfl, double* host_av, size_t count) {
    if (int r = check(e)) return r;
    const bool ab3 = e->P.scheme == TC_SCHEME_AB3;
    const int nv = ab3 ? 10 : 12, nrec = diag_record(e->P.scheme);
    void* v = field_ptr(e, TC_VEL, vel_slot);
    void* p = ab3 ? nullptr : field_ptr(e, TC_PRES, p_slot);
    if (!v) return fail(TC_ERR_ARG, "bad velocity slot");
    if (!ab3 && !p) return fail(TC_ERR_ARG, "bad pressure slot");
    if (host_av && count != (size_t)nv * e->nz) return fail(TC_ERR_ARG, "count must be nz*" + std::to_string(nv));
    if (cfl && (cfl_mode < 1 || cfl_mode > 3)) return fail(TC_ERR_ARG, "cfl_mode must be 1, 2 or 3");
    // the record a step formed of these fields (waiting for that step only, so a
    // later step may already run), else one pass now into buffer 2; either final
    // reduction writes the host buffer through its device mapping (no copy)
    int k = -1;
    for (int j = 0; j < 2; j++)
        if (e->diag_rec[j] == vel_slot && (ab3 || p_slot == vel_slot)) k = j;
    if (k >= 0) {
        HIPCHK(hipEventSynchronize(e->diag_ev[k]));
    } else {
        k = 2;
        const double visc_fac = diag_visc_fac(e->P.scheme, e->P.visc_type);
        if (e->fp32)
            HIPCHK(launch_diag<float>(e->s, e->mesh<float>(), e->P.scheme, (const float*)v, (const float*)p,
                                      (const float*)e->nut, visc_fac, e->d_partial, e->hd_diag[2]));
        else
            HIPCHK(launch_diag<double>(e->s, e->mesh<double>(), e->P.scheme, (const double*)v, (const double*)p,
                                       (const double*)e->nut, visc_fac, e->d_partial, e->hd_diag[2]));
        HIPCHK(hipStreamSynchronize(e->s));
    }
    // volumes as the device kernels see them: the mesh in the engine's precision
    auto r = [&](double x) { return e->fp32 ? (double)(float)x : x; };
    const double hxy = r(e->P.hx) * r(e->P.hy), inv_nxy = 1.0 / (double)e->nxy;
    double q[3] = {0, 0, 0}, ke = 0, cmax = 0;
    for (int iz = 0; iz < e->nz; iz++) {
        const double* d = e->h_diag[k] + (size_t)iz * nrec;
        const double sw = d[nv], sww = d[nv + 1];
        if (ab3) {
            const double vol = r(e->hm.cellh[iz]) * hxy;
            q[0] += d[0] * vol;
            q[1] += d[1] * vol;
            q[2] += sw * vol;
            ke += (d[3] + d[4] + sww) * vol;
        } else {
            const double dz = r(e->hm.DZ[iz]);
            const double dzw = iz == 0 ? 0.0 : 0.5 * (dz + r(e->hm.DZ[iz - 1]));
            q[0] += d[0] * dz * hxy;
            q[1] += d[1] * dz * hxy;
            q[2] += sw * dzw * hxy;
            ke += ((d[3] + d[4]) * dz + sww * dzw) * hxy;
        }
        if (cfl) cmax = std::max(cmax, d[nv + 1 + cfl_mode]);
        if (host_av)
            for (int k = 0; k < nv; k++) host_av[(size_t)iz * nv + k] = d[k] * inv_nxy;
    }
    if (integral)
        for (int k = 0; k < 3; k++) integral[k] = q[k];
    if (kinetic_energy) *kinetic_energy = 0.5 * ke;
    if (cfl) *cfl = tau * cmax;
    return TC_OK;
}

int tc_cfr_correct(tc_engine* e, int kind, int slot, double* shift) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind)) return fail(TC_ERR_ARG, "bad field");
    if (!(e->P.cfr_ubulk > 0)) return fail(TC_ERR_STATE, "cfr_ubulk is not set");
    double q[3];
    if (int r = tc_integral(e, kind, slot, q)) return r;
    const double sh = (e->P.cfr_ubulk * e->hm.area - q[0]) / e->hm.vol_u;
    HIPCHK(hipMemcpyAsync(e->d_shift, &sh, sizeof(double), hipMemcpyHostToDevice, e->s));
    if (e->fp32) HIPCHK(launch_add_u<float>(e->s, e->mesh<float>(), e->P.scheme, (float*)v, e->d_shift));
    else HIPCHK(launch_add_u<double>(e->s, e->mesh<double>(), e->P.scheme, (double*)v, e->d_shift));
    HIPCHK(hipStreamSynchronize(e->s));
    if (shift) *shift = sh;
    return TC_OK;
}

int tc_project(tc_engine* e, int kind, int slot) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind)) return fail(TC_ERR_ARG, "bad field");
    return e->fp32 ? project<float>(e, (float*)v) : project<double>(e, (double*)v);
}

int tc_initial_field(tc_engine* e, int slot, unsigned long long seed, double ubulk, double noise_rms) {
    if (int r = check(e)) return r;
    e->drop_records();
    void* v = field_ptr(e, TC_VEL, slot);
    if (!v) return fail(TC_ERR_ARG, "bad slot");
    std::vector<double> u(3 * e->n3d);
    make_initial_field(e->P, e->hm, seed, ubulk, noise_rms, u.data());
    if (e->fp32) {
        if (int r = upload_as<float>(e, v, u.data(), u.size())) return r;
    } else {
        HIPCHK(hipMemcpy(v, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return tc_project(e, TC_VEL, slot);
}

int tc_set_sgs_params(tc_engine* e, const float* params, size_t count) {
    if (int r = check(e)) return r;
    if (!e->P.sgs_model) return fail(TC_ERR_STATE, "engine created without the learned closure (sgs_model = 0)");
    if (!params || count != (size_t)kMlpParams)
        return fail(TC_ERR_ARG, "expected " + std::to_string(kMlpParams) + " parameters");
    std::vector<unsigned short> w;
    std::vector<float> b;
    mlp_pack_device(params, w, b);
    HIPCHK(hipMemcpyAsync(e->sgs_w, w.data(), w.size() * sizeof(unsigned short), hipMemcpyHostToDevice, e->s));
    HIPCHK(hipMemcpyAsync(e->sgs_b, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
    e->sgs_ready = true;
    return TC_OK;
}

long long tc_sgs_sample_count(const tc_engine* e, int factor) {
    if (!e || factor < 1 || e->nx % factor || e->ny % factor) return -1;
    return (long long)(e->nx / factor) * (e->ny / factor) * (e->nz - 2);
}

int tc_sgs_samples(tc_engine* e, int kind, int slot, int factor, float* dev_x, float* dev_t, long long cap,
                   long long* count) {
    if (int r = check(e)) return r;
    if (e->P.scheme != TC_SCHEME_AB3) return fail(TC_ERR_ARG, "training samples are taken from collocated fields");
    void* v = field_ptr(e, kind, slot);
    if (!v || !is_vec(kind) || !dev_x || !dev_t) return fail(TC_ERR_ARG, "bad field or output");
    if (factor < 1 || e->nx % (2 * factor) || e->ny % (2 * factor))
        return fail(TC_ERR_ARG, "factor must divide nx/2 and ny/2");
    const long long n = tc_sgs_sample_count(e, factor);
    if (cap < n) return fail(TC_ERR_ARG, "output capacity below " + std::to_string(n) + " samples");
    int r = e->fp32 ? sgs_samples<float>(e, (const float*)v, factor, dev_x, dev_t)
                    : sgs_samples<double>(e, (const double*)v, factor, dev_x, dev_t);
    if (r) return r;
    if (count) *count = n;
    return TC_OK;
}

int tc_timing_enable(tc_engine* e, int on) {
    if (!e) return fail(TC_ERR_ARG, "engine is NULL");
    e->tim.on = on != 0;
    return TC_OK;
}

int tc_timing_read(tc_engine* e, int kernel, double* total_ms, long long* launches) {
    if (int r = check(e)) return r;
    if (kernel < 0 || kernel >= TC_K_COUNT) return fail(TC_ERR_ARG, "bad kernel id");
    HIPCHK(hipStreamSynchronize(e->s));
    for (auto& rc : e->tim.recs) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, rc.a, rc.b));
        e->tim.total_ms[rc.k] += ms;
        e->tim.count[rc.k] += 1;
        e->tim.pool.push_back(rc.a);
        e->tim.pool.push_back(rc.b);
    }
    e->tim.recs.clear();
    if (total_ms) *total_ms = e->tim.total_ms[kernel];
    if (launches) *launches = e->tim.count[kernel];
    return TC_OK;
}

int tc_timing_reset(tc_engine* e) {
    if (int r = check(e)) return r;
    HIPCHK(hipStreamSynchronize(e->s));
    for (auto& rc : e->tim.recs) {
        e->tim.pool.push_back(rc.a);
        e->tim.pool.push_back(rc.b);
    }
    e->tim.recs.clear();
    for (int k = 0; k < TC_K_COUNT; k++) {
        e->tim.total_ms[k] = 0;
        e->tim.count[k] = 0;
    }
    return TC_OK;
}

}  // extern "C"
