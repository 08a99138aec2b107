/*
 * turbchannel.h -- C ABI of the MI355X time stepper for TurbChannel's
 * staggered (Harlow-Welch 2nd order, Verstappen-Veldman 4th order) and
 * collocated BDF3/extrapolation and IMEX Runge-Kutta (SRK) solvers.
 *
 * One opaque engine per device. It owns every device buffer, as the
 * reference's data_cuda::data does (staggered/func_global.h:944-974).
 * Host buffers use the reference's layout: real3 AoS, cell index
 * in = ix + NX*(iy + NY*iz), ix fastest (staggered/func_cuda.cpp:68), where
 * real is double for TC_FP64 and float for TC_FP32_BULK (tc_real_size()).
 *
 * Every call returns TC_OK or an error code; tc_last_error() gives the text.
 * The reference prints and calls exit(0) instead (staggered/func_cuda.cpp:23-35);
 * a driver maps a non-zero status to that. Work is queued on the engine's
 * stream; calls that return host data (tc_get_field, tc_integral, tc_profiles,
 * tc_cfl, ...) synchronise first, and tc_sync() waits for everything queued.
 *
 * A field is named by (kind, slot): the reference's vel[4], flux[4], pres[4]
 * ring buffers (staggered/data_global_inc.h:126-130) and nu_t (slot 0).
 */
#ifndef TURBCHANNEL_H
#define TURBCHANNEL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_engine tc_engine;
typedef struct tc_mlp tc_mlp;

enum tc_scheme { TC_SCHEME_HW = 0, TC_SCHEME_VV = 1, TC_SCHEME_AB3 = 2,
                 TC_SCHEME_SRK = 3 /* collocated IMEX Runge-Kutta (collocated_SRK/): AB3's mesh and
                                      FD operators, stepped by tc_srk_step */ };
/* IMEX Runge-Kutta methods of the SRK stepper (collocated_SRK/imex.h:26-42) */
enum tc_imex { TC_IMEX_ARS_121 = 0, TC_IMEX_ARS_232 = 1, TC_IMEX_ARS_343 = 2, TC_IMEX_MARS_343 = 3,
               TC_IMEX_ARK3 = 4, TC_IMEX_ARK4 = 5, TC_IMEX_MARK3 = 6, TC_IMEX_BHR_553 = 7 };
/* tIMEXtype (collocated_SRK/imex.h:9) */
enum tc_imex_type { TC_IMEX_TYPE_A = 0, TC_IMEX_TYPE_ARS = 1, TC_IMEX_TYPE_CK = 2 };
enum tc_precision { TC_FP64 = 0, TC_FP32_BULK = 1 };
enum tc_kind { TC_VEL = 0, TC_FLUX = 1, TC_PRES = 2, TC_NUT = 3,
               TC_TAU = 4 /* learned-closure stress, 6 reals per node: XX YY ZZ XY XZ YZ (slot 0) */ };
enum tc_status { TC_OK = 0, TC_ERR_ARG = 1, TC_ERR_HIP = 2, TC_ERR_STATE = 3, TC_ERR_NUMERIC = 4 };

/* Kernel classes reported by tc_timing_read(). */
enum tc_kernel {
    TC_K_TURB_VISC = 0, /* subgrid model: Smagorinsky-van Driest nu_t, or the learned closure's tau */
    TC_K_FLUX = 1,      /* explicit term + BDF3/EX3 right-hand side */
    TC_K_DIV = 2,       /* divergence of the right-hand side: fused into TC_K_FFT_X, no launch of its own */
    TC_K_FFT_X = 3,     /* divergence + forward real FFT along x */
    TC_K_FFT_Y = 4,     /* forward complex FFT along y */
    TC_K_ZSOLVE = 5,    /* banded wall-normal solve per (kx, ky) */
    TC_K_IFFT_Y = 6,    /* inverse complex FFT along y */
    TC_K_IFFT_X = 7,    /* inverse real FFT along x */
    TC_K_PROJECT = 8,   /* gradient + velocity correction (+ flow-rate shift) */
    TC_K_OTHER = 9,     /* reductions and small kernels */
    TC_K_COUNT = 10
};

/* Problem definition: the reference's main_params (staggered/config.h:47-73,
 * collocated_AB3/config.h:39-56) plus the wall-normal mesh and this engine's
 * scheme/precision switches, which the reference fixes at compile time
 * (CONV_APPR_TYPE, VISC_APPR_TYPE, ENABLE_TURB_VISC, USE_DOUBLE_PRECISION). */
typedef struct tc_params {
    int scheme;            /* tc_scheme */
    int precision;         /* tc_precision */
    int nx, ny, nz;        /* staggered: nz cells; collocated: nz nodes incl. walls. nx, ny even */
    int turb_visc;         /* allocate nu_t and run the Smagorinsky-van Driest model */
    int visc_type;         /* staggered: 0 CROSS, 1 NAIVE; collocated: 0 FD Laplacian, 1 Galerkin */
    int stab_type;         /* collocated: StabType 0 or 1 */
    int fd_mode;           /* collocated: FDmode */
    double visc, sour_dpdx, inv_length_scale;
    double hx, hy, xmin, ymin;
    double cfr_ubulk;      /* > 0: hold the bulk velocity at this value (constant flow rate) */
    const double* zz;      /* staggered: nz+1 face coordinates; collocated: nz node coordinates */
    int sgs_model;         /* collocated only, turb_visc = 0: 1 = learned closure (no reference
                              counterpart): each step evaluates tau from the velocity with the
                              network set by tc_set_sgs_params and subtracts its divergence;
                              2 = the same with backscatter clipping (tau = 0 at a node where
                              tau_ij S_ij > 0, S the resolved strain); 3 = the dissipative part
                              of the network's stress along the deviatoric strain D:
                              tau' = (tau_kl D_kl / D_kl D_kl) D where tau_kl D_kl < 0, else 0 */
    int imex;              /* TC_SCHEME_SRK: tc_imex method (main_params::IMEXtype, NStages,
                              ButcherE/I, d; collocated_SRK/main.cpp:342-350) */
    double alpha_tau;      /* TC_SCHEME_SRK: alpha = alpha_tau / tau; <= 0 selects the reference's
                              default 0.5 * alpha_tau_max[stab_type] (collocated_SRK/main.cpp:352) */
} tc_params;

/* Lifecycle. Replaces data::alloc_all + InitFourier + CopyDataToDevice
 * (staggered/func_global.h:944-963, :197-212; staggered/func_cuda.cpp:102-104)
 * and the mesh set-up InitDerivativeMeshData (staggered/main.cpp:112-156). */
int tc_create(const tc_params* p, int device, tc_engine** out);
int tc_destroy(tc_engine* e);
const char* tc_last_error(void);
int tc_real_size(const tc_engine* e);           /* sizeof(real) of host buffers: 8 or 4 */
int tc_sync(tc_engine* e);

/* Host <-> device copies of one field; replaces the cudaMemcpy calls on
 * vel[i], pres[i], nu_t in staggered/main.cpp:283-286, :348-349, :382-383. */
int tc_set_field(tc_engine* e, int kind, int slot, const void* host, size_t bytes);
int tc_get_field(tc_engine* e, int kind, int slot, void* host, size_t bytes);
int tc_field_device_ptr(tc_engine* e, int kind, int slot, void** dev_ptr);

/* One time step: data::Step(t, tau, current_field, bdf_order)
 * (staggered/func_global.h:897-942; collocated_AB3/func_global.h:387-450).
 * Reads slots cur, cur-1, cur-2 (mod 4) and writes slot (cur+1)%4. */
int tc_step(tc_engine* e, double t, double tau, int cur, int bdf_order);
/* One step of the collocated SRK solver: data::Step(t, tau, velocity, pressure, dpdt)
 * (collocated_SRK/func_global.h:317-408), in place on velocity slot vel_slot and
 * pressure slots p_slot and dpdt_slot (dpdt_slot is read and written for
 * stab_type 1 only and must differ from p_slot). Engines of TC_SCHEME_SRK only;
 * tc_step is refused on them. Graph capture as for tc_step, keyed by the slots and tau. */
int tc_srk_step(tc_engine* e, double t, double tau, int vel_slot, int p_slot, int dpdt_slot);
/* Butcher tables of one IMEX method: (s+1) x s row-major, row s = weights
 * (tIMEXMethod::ButcherE/ButcherI, d of CalcD and alpha_tau_max; collocated_SRK/imex.cpp).
 * Each output may be NULL; be and bi need s*(s+1) doubles, d s doubles. No device needed. */
int tc_imex_table(int method, int* nstages, int* type, double* be, double* bi, double* d, double* alpha_tau_max);

/* on != 0: tc_step captures each distinct (cur, bdf_order, tau) step into a HIP
 * graph on first use and replays it afterwards, so a step is one graph launch
 * instead of about ten kernel launches (the results are identical). Off by
 * default; ignored while kernel timing is on. No reference counterpart. */
int tc_set_graphs(tc_engine* e, int on);

/* Operators used by the reference's self-tests (staggered/selftest.cpp:360-431). */
int tc_explicit_term(tc_engine* e, int in_kind, int in_slot, int out_kind, int out_slot);     /* data::ExplicitTerm */
int tc_apply_gradient(tc_engine* e, int p_kind, int p_slot, int out_kind, int out_slot);      /* data::ApplyGradient */
int tc_unapply_gradient(tc_engine* e, int v_kind, int v_slot, int p_kind, int p_slot);        /* data::UnapplyGradient */
int tc_normalize_pressure(tc_engine* e, int kind, int slot);                                  /* data::NormalizePressure */
int tc_recalc_turb_visc(tc_engine* e, int kind, int slot);                                    /* data::RecalcTurbVisc */

/* Reductions, synchronous (staggered/func_global.h:677-808). */
int tc_integral(tc_engine* e, int kind, int slot, double out[3]);                            /* data::CalcIntegral */
int tc_kinetic_energy(tc_engine* e, int kind, int slot, double* out);                        /* data::CalcKineticEnergy */
int tc_cfl(tc_engine* e, int kind, int slot, double tau, int mode, double* out);             /* data::CalcCFL */
/* data::CalcProfiles / CalcVelocityProfile: per plane, FP64, staggered 12 values
 * (u[3], uu[6], nu_t, p, pp) and collocated 10 (u[3], uu[6], nu_t). */
int tc_profiles(tc_engine* e, int vel_kind, int vel_slot, int p_slot, double* host_av, size_t count);

/* The per-step diagnostics of the reference's time loop (staggered/main.cpp:312-338:
 * CalcIntegral, CalcProfiles, CalcCFL; collocated_AB3/main.cpp:414-441) from one
 * pass over velocity slot vel_slot (and pressure slot p_slot, nu_t): one kernel
 * and one copy to the host instead of three of each. Each output may be NULL.
 * integral: as tc_integral; kinetic_energy: as tc_kinetic_energy; cfl: as
 * tc_cfl(tau, cfl_mode); host_av/count: as tc_profiles. No reference counterpart
 * as a single call. */
int tc_diagnostics(tc_engine* e, int vel_slot, int p_slot, double tau, int cfl_mode, double integral[3],
                   double* kinetic_energy, double* cfl, double* host_av, size_t count);
/* on != 0: tc_step also forms these diagnostics of the fields it writes, in its
 * projection kernel (staggered/main.cpp:312-338 reads the new velocity three
 * more times; here the values are summed while they are stored), and a
 * following tc_diagnostics of the step's new velocity slot (and the same
 * pressure slot, staggered) copies them instead of reading the fields again.
 * The records of the last two steps are kept (each with the nu_t of its own
 * step), and tc_diagnostics waits only for the step whose record it reads, so a
 * caller may launch step n+1 before reading step n's diagnostics; a third step
 * replaces the older record. Any call that may change a field or nu_t
 * (tc_set_field, tc_field_device_ptr, tc_srk_step, ...) drops both records;
 * writes through a device pointer obtained earlier are not seen. Off by
 * default; refused on SRK engines. No reference counterpart. */
int tc_step_diagnostics(tc_engine* e, int on);

/* Constant-flow-rate correction of one velocity field (no reference counterpart;
 * tc_step applies it itself when cfr_ubulk > 0). shift may be NULL. */
int tc_cfr_correct(tc_engine* e, int kind, int slot, double* shift);
/* u <- u - G L^-1 D u: the projection of tc_step without the time-step terms. */
int tc_project(tc_engine* e, int kind, int slot);

/* Seeded initial field (no reference counterpart): laminar parabola with bulk
 * velocity ubulk plus the discrete curl of a smoothed Gaussian vector potential
 * scaled to noise_rms, on the host in FP64, not projected. out: 3*nx*ny*nz doubles.
 * tc_initial_field uploads it to a velocity slot and projects it. */
int tc_make_initial_field(const tc_params* p, unsigned long long seed, double ubulk, double noise_rms, double* out);
int tc_initial_field(tc_engine* e, int slot, unsigned long long seed, double ubulk, double noise_rms);

/* Mesh data derived on the host from tc_params (staggered/main.cpp:112-156,
 * collocated_AB3/main.cpp:295-300). name: "ZZ","DZ","DZW","GC1","GC2","GCZ",
 * "MAT_A","MAT_D","KERG","KERG_norm2"; and the wall-normal elimination of the
 * pressure systems (ThomasForPressure, staggered/func_global.h:71-189,
 * collocated_AB3/func_global.h:8-61) per (kx, |ky|) column: "ZF_M" (W x nz x
 * columns forward multipliers), "ZF_D" (inverse pivots), "ZF_U" (normalised upper
 * band), and the constants from which the FP32-bulk z-solve forms the forward
 * factors in-lane: "ZF_R" (nz x 12 for the fourth-order scheme, nz x 4
 * otherwise), "ZF_KX", "ZF_KY" (wavenumber terms per kx, |ky|). Returns the
 * length, or -1. No device needed. */
int tc_mesh_array(const tc_params* p, const char* name, double* out, int cap);

/* ---------------------------------------------------------------------------
 * Learned subgrid closure (no reference counterpart; the reference's closure is
 * RecalcTurbVisc, staggered/func_global.h:815-890, collocated_AB3/func_global.h:342-384).
 * Network 10 -> 64 -> 64 -> 64 -> 6 with GELU (tanh form), bfloat16 MFMA with FP32
 * accumulation. Inputs: the 9 velocity-gradient components d u_i / d x_j (i-major)
 * times nu / u_tau^2, and the wall distance times inv_length_scale / 100; outputs:
 * the deviatoric stress XX YY ZZ XY XZ YZ over u_tau^2, u_tau = visc * inv_length_scale.
 * Parameters: flat FP32 vector, per layer W (out x in, row-major) then b.
 * ------------------------------------------------------------------------- */
int tc_set_sgs_params(tc_engine* e, const float* params, size_t count);
/* Samples for training from one collocated velocity field: sharp spectral cut-off
 * by `factor` in x and y, at the coarse grid's interior nodes. dev_x: 10 floats and
 * dev_t: 6 floats per sample, device memory; tc_sgs_sample_count() samples. */
long long tc_sgs_sample_count(const tc_engine* e, int factor);
int tc_sgs_samples(tc_engine* e, int kind, int slot, int factor, float* dev_x, float* dev_t, long long cap,
                   long long* count);

/* Trainer: FP32 master weights, Adam (beta 0.9, 0.999, eps 1e-8) and bfloat16
 * device copies on one device. Loss: mean over samples and outputs of the squared
 * error. All calls return when the work is done. */
int tc_mlp_param_count(void);
int tc_mlp_create(int device, unsigned long long seed, tc_mlp** out);
int tc_mlp_destroy(tc_mlp* m);
int tc_mlp_set_params(tc_mlp* m, const float* host, size_t count);   /* also resets Adam */
int tc_mlp_get_params(tc_mlp* m, float* host, size_t count);
int tc_mlp_forward(tc_mlp* m, const float* dev_x, long long n, float* dev_y);
/* Loss and its gradient at the current weights (host_grad may be NULL). */
int tc_mlp_gradient(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float* host_grad, size_t count,
                    double* loss);
/* One Adam step on the batch; loss (may be NULL) is the batch loss before the step. */
int tc_mlp_train_step(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float lr, double* loss);
/* tc_mlp_train_step that allocates nothing: no counterpart in the reference. For
 * callers that cannot allocate device memory (training loops under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). The gradient kernel's
 * per-block partial records go to a static device array of this library (one per
 * GPU, 512 records = 19.3 MB, held until the process ends) instead of the
 * trainer's own buffer, which tc_mlp_train_step grows with hipMalloc when a batch
 * needs more records. Every batch size fits (the kernel writes at most 512
 * records; n above 2^31 - 1 is rejected). Results equal tc_mlp_train_step's bit
 * for bit. Calls that overlap on one GPU (any trainers) share the array and race. */
int tc_mlp_train_step_noalloc(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float lr,
                              double* loss);
int tc_mlp_steps(const tc_mlp* m, long long* steps);

/* Per-kernel timing with HIP events on the engine's stream. */
int tc_timing_enable(tc_engine* e, int on);
int tc_timing_read(tc_engine* e, int kernel, double* total_ms, long long* launches);
int tc_timing_reset(tc_engine* e);

#ifdef __cplusplus
}
#endif
#endif
