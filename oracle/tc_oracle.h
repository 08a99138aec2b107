/*
 * tc_oracle.h -- CPU restatement of bahvalo/TurbChannel's time stepper.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * engine in turbchannel_amd/. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Everything is FP64 and follows the reference's CPU (OpenMP) path:
 *   staggered/func_global.h, staggered/main.cpp:12-185,
 *   collocated_AB3/func_global.h, collocated_AB3/func_fluxes.h,
 *   collocated_AB3/main.cpp:197-301, collocated_SRK/func_global.h,
 *   collocated_SRK/imex.cpp.
 * Index helpers are generalised from the reference's power-of-2 bitmask form
 * (staggered/data_global_inc.h:4-9) to any NX, NY; or_index_selfcheck()
 * compares both forms at power-of-2 sizes.
 *
 * Field layout is the reference's: real3 AoS, cell index
 * in = ix + NX*(iy + NY*iz), ix fastest (staggered/func_cuda.cpp:68).
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_SCHEME_HW = 0, OR_SCHEME_VV = 1, OR_SCHEME_AB3 = 2,
       OR_SCHEME_SRK = 3 /* collocated_SRK: AB3's mesh and operators (FD only), zero-mean pressure */ };
/* IMEX Runge-Kutta methods of collocated_SRK/imex.h:26-42 and their types (imex.h:9) */
enum { OR_IMEX_ARS_121 = 0, OR_IMEX_ARS_232 = 1, OR_IMEX_ARS_343 = 2, OR_IMEX_MARS_343 = 3,
       OR_IMEX_ARK3 = 4, OR_IMEX_ARK4 = 5, OR_IMEX_MARK3 = 6, OR_IMEX_BHR_553 = 7 };
enum { OR_IMEX_TYPE_A = 0, OR_IMEX_TYPE_ARS = 1, OR_IMEX_TYPE_CK = 2 };
enum { OR_KIND_VEL = 0, OR_KIND_FLUX = 1, OR_KIND_PRES = 2, OR_KIND_NUT = 3, OR_KIND_TAU = 4 };

typedef struct or_params {
    int scheme;        /* OR_SCHEME_* */
    int nx, ny, nz;    /* staggered: nz cells; AB3: nz nodes (walls included) */
    int turb_visc;     /* 1: nu_t array + Smagorinsky-van Driest (ENABLE_TURB_VISC) */
    int visc_type;     /* staggered: 0 CROSS, 1 NAIVE; AB3: 0 FD Laplacian, 1 Galerkin */
    int stab_type;     /* AB3: 0 or 1 (collocated_AB3/func_global.h:424-449) */
    int fd_mode;       /* AB3: FDmode (collocated_AB3/config.h:55) */
    double visc, sour_dpdx, inv_length_scale;
    double hx, hy, xmin, ymin;
    double cfr_ubulk;  /* >0: hold the bulk velocity at this value (new numerics) */
    int sgs_model;     /* AB3 only. 1: learned closure; the explicit term subtracts
                          div tau of the tau field (OR_KIND_TAU, 6 per node, XX YY ZZ XY XZ YZ),
                          which the caller sets (new numerics) */
} or_params;

typedef struct or_solver or_solver;

/* zz: staggered nz+1 face coordinates, AB3 nz node coordinates. NULL on error. */
or_solver* or_create(const or_params* p, const double* zz);
void or_destroy(or_solver* s);
const char* or_last_error(void);

/* Mesh-derived arrays (staggered/main.cpp:112-156; collocated_AB3/main.cpp:295-300).
 * name: "ZZ","DZ","DZW","GC1","GC2","GCZ","MAT_A","MAT_D","KERG". Returns length, -1 if absent. */
int or_mesh_array(const or_solver* s, const char* name, double* out, int cap);
double or_kerg_norm2(const or_solver* s);
/* FillMatrix (staggered/main.cpp:12-48): NZ rows of 7 entries. */
void or_fill_matrix(const or_solver* s, double* M);
/* solve_banded (staggered/main.cpp:73-109). Returns its error code. */
int or_solve_banded(int N, int NB, double* M, double* x, double tiny);

/* Field operators on caller arrays (real3 AoS = 3*N3D doubles). */
void or_explicit_term(or_solver* s, const double* u, double* kuhat);
void or_flux_term(or_solver* s, const double* u, int do_conv, int do_visc, double* kuhat);
void or_visc_flux_naive(or_solver* s, const double* u, double* f);  /* staggered ViscFlux<NAIVE> */
void or_apply_gradient(or_solver* s, const double* p, double* g);
void or_apply_divergence(or_solver* s, const double* v, double* dv);
void or_solve_pressure(or_solver* s, double* x, const double* f);   /* LinearSystemSolveFFT */
void or_unapply_gradient(or_solver* s, const double* v, double* p);
void or_normalize_pressure(or_solver* s, double* p);
/* Seconds spent per channel since the last reset (or creation), mirroring the
 * reference's timers (staggered/main.cpp:376-378): explicit term (with nu_t),
 * gradient (the step's ApplyGradient calls), divergence, pressure FFT (with the
 * copies around it), pressure Thomas, pressure normalisation, and the rest of
 * or_step (BDF combination, projection, flow rate). out: OR_NTIMERS doubles. */
enum { OR_TMR_EXPLICIT, OR_TMR_GRADIENT, OR_TMR_DIVERGENCE, OR_TMR_FFT, OR_TMR_THOMAS, OR_TMR_NORMALIZE,
       OR_TMR_OTHER, OR_NTIMERS };
void or_timers(or_solver* s, double* out, int reset);
void or_recalc_turb_visc(or_solver* s, const double* u);            /* writes the internal nu_t */
void or_calc_abs_s(or_solver* s, const double* u, double* absS);    /* staggered CalcAbsS */
void or_integral(or_solver* s, const double* u, double out[3]);
double or_kinetic_energy(or_solver* s, const double* u);
/* Profiles: staggered 12 doubles/plane (u3, uu6, nu_t, p, pp); AB3 10 (u3, uu6, nu_t). */
int or_profiles(or_solver* s, const double* u, const double* p, double* av);
double or_cfl(or_solver* s, const double* u, double tau, int mode);

/* Ring-buffer state, as the reference's data::vel/flux/pres[4] and nu_t; plus tau. */
int or_set_field(or_solver* s, int kind, int slot, const double* src);
int or_get_field(or_solver* s, int kind, int slot, double* dst);
int or_step(or_solver* s, double t, double tau, int cur, int bdf_order);
/* SRK: Butcher tables ((s+1) x s, row s = weights), d of CalcD and alpha_tau_max
 * (collocated_SRK/imex.cpp). Arrays may be NULL; be/bi need 42 doubles, d 6.
 * Returns the number of stages, 0 for an unknown method. */
int or_imex_table(int method, int* type, double* be, double* bi, double* d, double* alpha_tau_max);
/* One step of data::Step (collocated_SRK/func_global.h:317-408) in place on
 * caller arrays: velocity (3*N3D), pressure and dpdt (N3D; dpdt is read and
 * written only for stab_type 1). The solver must have scheme OR_SCHEME_SRK. */
int or_srk_step(or_solver* s, int method, int stab_type, double alpha_tau, double t_start, double tau,
                double* velocity, double* pressure, double* dpdt);
/* Bulk-flow correction used by or_step when cfr_ubulk > 0 (returns the shift applied). */
double or_cfr_correct(or_solver* s, double* u);

/* New numerics: tanh wall-normal mesh and the seeded initial field (before projection). */
void or_tanh_mesh(int n_intervals, double gamma, double* zz);
void or_initial_field(or_solver* s, unsigned long long seed, double ubulk, double noise_rms, double* u);

/* Plain complex FFT (for pinning against numpy): n points, sign -1 forward / +1 backward. */
void or_fft(int n, const double* in_re_im, double* out_re_im, int sign);

/* Compares generalised and power-of-2 bitmask index helpers; returns #mismatches. */
int or_index_selfcheck(int nx, int ny, int nz);

void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
