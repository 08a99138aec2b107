// turbch -- command-line channel-flow solver on the MI355X engine.
//
// The host program of the reference, restated over the C ABI of
// include/turbchannel.h: the same options (staggered/io.cpp:107-126,
// collocated_AB3/io.cpp:104-123), the same time loop with the BDF order ramp and
// CFL-driven step control (staggered/main.cpp:288-371, collocated_AB3/main.cpp:392-476)
// or, for the IMEX Runge-Kutta solver, with a fixed step (collocated_SRK/main.cpp:472-550),
// the same screen and log.dat lines, the same files (qNNNNN.vtk, rNNNNN.dat,
// output.vtk, output.dat, res.dat; staggered/io.cpp:6-96, collocated_AB3/io.cpp:9-101)
// and the same start from input.vtk when it exists (staggered/io.cpp:31-69).
// Errors print a message and exit with status 0, as the reference does
// (staggered/func_cuda.cpp:23-35).
//
// The reference fixes the scheme, precision, grid and physics at compile time
// (staggered/config.h, collocated_AB3/config.h); here they are options with the
// reference's values as defaults. Options with no reference counterpart: bulk
// units and a constant flow rate, the box size, the tanh / uniform / geometric /
// file wall-normal meshes, a seeded divergence-free start, the learned subgrid
// closure, and an exact binary restart of the whole solver state.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "turbchannel.h"

namespace {

const double kPi = 3.14159265358979323846;

[[noreturn]] void die(const std::string& msg) {
    printf("%s\n", msg.c_str());
    fflush(stdout);
    exit(0);
}

void check(int status, const char* what) {
    if (status != TC_OK) die(std::string(what) + ": " + tc_last_error());
}

// ---------------------------------------------------------------------------
// Options

struct Options {
    // the reference's t_cmdline_params (staggered/config.h:75-87; collocated_AB3/io.h:10-21)
    int device = 0;
    int nz = -1;  // staggered: cells (default 32); collocated: nodes (default from the mesh)
    double time_max = -1;  // 100 (0.1 for srk)
    double time_start_averaging = 0.0;
    long long time_steps_max = 0x7FFFFFFF;
    double tau = -1;       // 1e-3 (0.8e-3 for srk)
    double cfl_min = 0.05, cfl_max = 0.5;
    int s_progress = -1, f_progress = -1, p_progress = -1;  // defaults depend on the solver
    // what the reference sets at compile time
    std::string scheme = "vv";  // hw | vv | ab3 | srk
    std::string precision = "fp64";
    int nx = -1, ny = -1;
    int les = -1;     // ENABLE_TURB_VISC: 1 (both reference configs) unless the learned closure is on
    int visc = -1;    // staggered 0 CROSS / 1 NAIVE, collocated 0 FD / 1 Galerkin; default follows les
    int stab = 1, fd_mode = 1;
    double re_tau = -1;  // 395 (180 for srk) unless given
    std::string imex = "ars343";  // srk: IMEX Runge-Kutta method (collocated_SRK/main.cpp:342)
    double alpha_tau = 0;         // srk: <= 0 selects 0.5 * alpha_tau_max[stab]
    // options with no reference counterpart
    double re_b = 0;     // > 0: bulk units, nu = 2 / Re_b
    double lx = -1, ly = -1;
    std::string mesh;    // sinh | tanh | uniform | geometric | file
    double stretch = 6.5, gamma = 1.8;
    double geo_y1 = -1, geo_ratio = -1, geo_ybulk = -1;  // wall units; 50, 1.05, 2.2 (srk: 1, 1.3, 10)
    std::string zfile;
    double cfr = 0;      // > 0: hold the bulk velocity at this value
    long long seed = -1; // >= 0: seeded divergence-free start instead of the analytic one
    double ubulk = 1.0, noise = 0.1;
    std::string sgs_file;
    int sgs_clip = 0;  // 1: backscatter clipping of the learned closure
    int sgs_proj = 0;  // 1: the learned closure's dissipative part along the strain rate
    long long checkpoint_every = 0;
    std::string checkpoint = "restart.bin";
    std::string restart;
    int timing = 0;
    int graphs = 1;
    int fused_diag = 1;
    int pipeline = 1;
    int print_mesh = 0;
};

struct OptSpec {
    const char* names;  // comma-separated aliases
    const char* help;
};

const OptSpec kOpts[] = {
    {"-d,--device", "device to use"},
    {"--nz", "number of nodes in the normal direction (staggered: cells)"},
    {"-t,--TimeMax", "maximal integration time"},
    {"-a,--TimeStartAveraging", "time where averaging should be started"},
    {"--TimeStepsMax", "maximal number of iterations"},
    {"--tau", "initial timestep size"},
    {"--CFLmin", "if CFL falls below this value, then the timestep will be increased"},
    {"--CFLmax", "if CFL lifts above this value, then the timestep will be decreased"},
    {"-s,--SProgress", "number of timesteps between writing to screen and log"},
    {"-f,--FProgress", "number of timesteps between writing a field"},
    {"-p,--PProgress", "number of timesteps between writing a profile"},
    {"--scheme", "hw | vv (staggered) | ab3 | srk (collocated BDF3 / IMEX Runge-Kutta); default vv"},
    {"--imex", "srk: ars121 | ars232 | ars343 | mars343 | ark3 | ark4 | mark3 | bhr553 (default ars343)"},
    {"--alpha_tau", "srk: pressure stabilisation alpha * tau (default 0.5 * the method's limit)"},
    {"--precision", "fp64 | fp32 (FP32 bulk, FP64 islands); default fp64"},
    {"--nx", "cells (nodes) in x; default 64 (srk 128)"},
    {"--ny", "cells (nodes) in y; default 32 staggered, 64 collocated"},
    {"--les", "1: Smagorinsky-van Driest eddy viscosity (default without --sgs), 0: none"},
    {"--visc", "viscous term: staggered 0 CROSS / 1 NAIVE, collocated 0 FD / 1 Galerkin"},
    {"--stab", "collocated StabType 0 | 1 (default 1)"},
    {"--Re_tau", "friction Reynolds number (default 395, srk 180)"},
    {"--Re_b", "bulk units: H = U_b = 1, nu = 2/Re_b"},
    {"--Lx", "box length in x over H (default 4 pi)"},
    {"--Ly", "box length in y over H (default 2 pi staggered, 4 pi/3 collocated)"},
    {"--mesh", "sinh | tanh | uniform | geometric | file"},
    {"--stretch", "sinh mesh stretching (default 6.5)"},
    {"--gamma", "tanh mesh parameter (default 1.8)"},
    {"--geo_y1", "geometric mesh: first step in wall units (default 50, srk 1)"},
    {"--geo_ratio", "geometric mesh: growth ratio (default 1.05, srk 1.3)"},
    {"--geo_ybulk", "geometric mesh: largest step in wall units (default 2.2, srk 10)"},
    {"--zfile", "wall-normal coordinates, one per line (faces staggered, nodes collocated)"},
    {"--cfr", "hold the bulk velocity at this value (constant flow rate)"},
    {"--seed", "seeded divergence-free start (laminar profile + random vortices)"},
    {"--ubulk", "bulk velocity of the seeded start (default 1)"},
    {"--noise", "rms of the seeded perturbation (default 0.1)"},
    {"--sgs", "learned closure: file of float32 network parameters (collocated, --les 0)"},
    {"--sgs_clip", "1: the learned closure's stress is zero where tau_ij S_ij > 0 (backscatter clipping)"},
    {"--sgs_proj", "1: only the learned closure's dissipative part along the deviatoric strain rate (eddy viscosity)"},
    {"--checkpoint-every", "write the restart file every this many steps (0: never)"},
    {"--checkpoint", "restart file to write (default restart.bin)"},
    {"--restart", "continue exactly from this restart file"},
    {"--timing", "1: report the time of each kernel class at the end"},
    {"--graphs", "1: replay repeated steps from captured HIP graphs (default), 0: launch each kernel"},
    {"--fused_diag", "1: the step's projection forms the per-step diagnostics (default), 0: a pass of its own"},
    {"--pipeline", "1: launch the next step before reading this step's diagnostics (default; needs --fused_diag 1)"},
    {"--print-mesh", "1: print the wall-normal coordinates and exit"},
};

void usage() {
    printf("Turbulent flow between parallel plates\nUsage: turbch [OPTIONS]\n\nOptions:\n");
    printf("  %-26s %s\n", "-h,--help", "Print this help message and exit");
    for (const OptSpec& o : kOpts) printf("  %-26s %s\n", o.names, o.help);
}

bool matches(const char* names, const std::string& key) {
    std::string s(names);
    size_t b = 0;
    while (b <= s.size()) {
        size_t e = s.find(',', b);
        if (e == std::string::npos) e = s.size();
        if (s.compare(b, e - b, key) == 0) return true;
        b = e + 1;
    }
    return false;
}

template <typename V>
void parse_num(const std::string& key, const std::string& val, V& out) {
    char* end = nullptr;
    errno = 0;
    if constexpr (std::is_integral<V>::value) {
        long long x = strtoll(val.c_str(), &end, 10);
        if (errno || end == val.c_str() || *end) die("Could not convert: " + key + " = " + val);
        out = (V)x;
    } else {
        double x = strtod(val.c_str(), &end);
        if (errno || end == val.c_str() || *end) die("Could not convert: " + key + " = " + val);
        out = x;
    }
}

// Returns false when the program should stop (help printed or an error).
bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; i++) {
        std::string key = argv[i], val;
        if (key == "-h" || key == "--help") {
            usage();
            return false;
        }
        const size_t eq = key.find('=');
        bool have_val = false;
        if (key.compare(0, 2, "--") == 0 && eq != std::string::npos) {
            val = key.substr(eq + 1);
            key = key.substr(0, eq);
            have_val = true;
        }
        const OptSpec* spec = nullptr;
        for (const OptSpec& s : kOpts)
            if (matches(s.names, key)) spec = &s;
        if (!spec) {
            printf("The following argument was not expected: %s\nRun with --help for more information.\n", argv[i]);
            return false;
        }
        if (!have_val) {
            if (i + 1 >= argc) {
                printf("%s: 1 required argument missing\nRun with --help for more information.\n", key.c_str());
                return false;
            }
            val = argv[++i];
        }
        auto is = [&](const char* name) { return matches(spec->names, name); };
        if (is("--device")) parse_num(key, val, o.device);
        else if (is("--nz")) parse_num(key, val, o.nz);
        else if (is("--TimeMax")) parse_num(key, val, o.time_max);
        else if (is("--TimeStartAveraging")) parse_num(key, val, o.time_start_averaging);
        else if (is("--TimeStepsMax")) parse_num(key, val, o.time_steps_max);
        else if (is("--tau")) parse_num(key, val, o.tau);
        else if (is("--CFLmin")) parse_num(key, val, o.cfl_min);
        else if (is("--CFLmax")) parse_num(key, val, o.cfl_max);
        else if (is("--SProgress")) parse_num(key, val, o.s_progress);
        else if (is("--FProgress")) parse_num(key, val, o.f_progress);
        else if (is("--PProgress")) parse_num(key, val, o.p_progress);
        else if (is("--scheme")) o.scheme = val;
        else if (is("--imex")) o.imex = val;
        else if (is("--alpha_tau")) parse_num(key, val, o.alpha_tau);
        else if (is("--precision")) o.precision = val;
        else if (is("--nx")) parse_num(key, val, o.nx);
        else if (is("--ny")) parse_num(key, val, o.ny);
        else if (is("--les")) parse_num(key, val, o.les);
        else if (is("--visc")) parse_num(key, val, o.visc);
        else if (is("--stab")) parse_num(key, val, o.stab);
        else if (is("--Re_tau")) parse_num(key, val, o.re_tau);
        else if (is("--Re_b")) parse_num(key, val, o.re_b);
        else if (is("--Lx")) parse_num(key, val, o.lx);
        else if (is("--Ly")) parse_num(key, val, o.ly);
        else if (is("--mesh")) o.mesh = val;
        else if (is("--stretch")) parse_num(key, val, o.stretch);
        else if (is("--gamma")) parse_num(key, val, o.gamma);
        else if (is("--geo_y1")) parse_num(key, val, o.geo_y1);
        else if (is("--geo_ratio")) parse_num(key, val, o.geo_ratio);
        else if (is("--geo_ybulk")) parse_num(key, val, o.geo_ybulk);
        else if (is("--zfile")) o.zfile = val;
        else if (is("--cfr")) parse_num(key, val, o.cfr);
        else if (is("--seed")) parse_num(key, val, o.seed);
        else if (is("--ubulk")) parse_num(key, val, o.ubulk);
        else if (is("--noise")) parse_num(key, val, o.noise);
        else if (is("--sgs")) o.sgs_file = val;
        else if (is("--sgs_clip")) parse_num(key, val, o.sgs_clip);
        else if (is("--sgs_proj")) parse_num(key, val, o.sgs_proj);
        else if (is("--checkpoint-every")) parse_num(key, val, o.checkpoint_every);
        else if (is("--checkpoint")) o.checkpoint = val;
        else if (is("--restart")) o.restart = val;
        else if (is("--timing")) parse_num(key, val, o.timing);
        else if (is("--graphs")) parse_num(key, val, o.graphs);
        else if (is("--fused_diag")) parse_num(key, val, o.fused_diag);
        else if (is("--pipeline")) parse_num(key, val, o.pipeline);
        else if (is("--print-mesh")) parse_num(key, val, o.print_mesh);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Problem set-up

struct Problem {
    int scheme = TC_SCHEME_VV, precision = TC_FP64;
    bool collocated = false;
    int nx = 0, ny = 0, nz = 0;
    double visc = 0, u_tau = 0, sour_dpdx = 0, re_tau = 0;
    double hx = 0, hy = 0, Lx = 0, Ly = 0, xmin = 0, ymin = 0, area = 0;
    std::vector<double> zz;  // staggered: nz+1 faces; collocated: nz nodes
    std::vector<double> dz;  // staggered: nz cell heights; collocated: nz-1 element heights
    tc_params params{};
};

std::vector<double> read_column(const std::string& fname) {
    std::ifstream f(fname);
    if (!f.is_open()) die("cannot open " + fname);
    std::vector<double> v;
    double x;
    while (f >> x) v.push_back(x);
    return v;
}

// Coordinates of the wall-normal mesh over [0, 2H], H = 1.
// sinh: staggered/main.cpp:219-223. Geometric progression from the wall, capped
// step, uniform core, mirrored: collocated_AB3/main.cpp:205-226 (in wall units).
// tanh: z_k = 1 + tanh(gamma (2k/n - 1)) / tanh(gamma), no reference counterpart.
// re = u_tau / nu: the half-height in wall units.
std::vector<double> make_mesh(const Options& o, bool collocated, int& nz, double re) {
    std::string kind = o.mesh.empty() ? (collocated ? "geometric" : "sinh") : o.mesh;
    std::vector<double> z;
    if (kind == "file") {
        if (o.zfile.empty()) die("--mesh file needs --zfile");
        z = read_column(o.zfile);
    } else if (kind == "geometric") {
        const double ref = re;
        std::vector<double> zp = {0.0, o.geo_y1};
        while (true) {
            const double h = (zp[zp.size() - 1] - zp[zp.size() - 2]) * o.geo_ratio;
            if (h >= o.geo_ybulk) break;
            zp.push_back(zp.back() + h);
        }
        const int n = int((ref - zp.back()) / o.geo_ybulk) + 1;
        const double hb = (ref - zp.back()) / n;
        for (int i = 0; i < n - 1; i++) zp.push_back(zp.back() + hb);
        zp.push_back(ref);
        for (int i = (int)zp.size() - 2; i >= 0; i--) zp.push_back(2.0 * ref - zp[i]);
        for (double& v : zp) v /= re;
        z = zp;
    } else {
        // intervals: staggered cells (NZ_MAX = 32 by default), collocated nodes - 1
        const int n = collocated ? (o.nz > 0 ? o.nz - 1 : 64) : (o.nz > 0 ? o.nz : 32);
        if (n < 2) die("--nz is too small");
        z.assign(n + 1, 0.0);
        for (int k = 0; k <= n; k++) {
            if (kind == "uniform") z[k] = 2.0 * k / n;
            else if (kind == "tanh") z[k] = 1.0 + std::tanh(o.gamma * (2.0 * k / n - 1.0)) / std::tanh(o.gamma);
            else if (kind != "sinh") die("unknown mesh " + kind);
        }
        if (kind == "sinh") {
            for (int i = 0; i <= n / 2; i++) {
                z[i] = std::sinh(o.stretch * i / n) / std::sinh(0.5 * o.stretch);
                z[n - i] = 2.0 - z[i];
            }
        }
        if (kind == "tanh") z[0] = 0.0, z[n] = 2.0;
    }
    const int npts = (int)z.size();
    nz = collocated ? npts : npts - 1;
    if (o.nz > 0 && o.nz != nz) die("--nz " + std::to_string(o.nz) + " does not match the mesh (" + std::to_string(nz) + ")");
    return z;
}

Problem setup(const Options& o) {
    Problem P;
    if (o.scheme == "hw") P.scheme = TC_SCHEME_HW;
    else if (o.scheme == "vv") P.scheme = TC_SCHEME_VV;
    else if (o.scheme == "ab3") P.scheme = TC_SCHEME_AB3;
    else if (o.scheme == "srk") P.scheme = TC_SCHEME_SRK;
    else die("unknown scheme " + o.scheme);
    const bool srk = P.scheme == TC_SCHEME_SRK;
    int imex = -1;
    {
        static const char* names[] = {"ars121", "ars232", "ars343", "mars343", "ark3", "ark4", "mark3", "bhr553"};
        std::string k;
        for (char c : o.imex)
            if (c != '_') k += (char)std::tolower((unsigned char)c);
        for (int i = 0; i < 8; i++)
            if (k == names[i]) imex = i;
        if (imex < 0) die("unknown IMEX method " + o.imex);
    }
    if (o.precision == "fp64") P.precision = TC_FP64;
    else if (o.precision == "fp32") P.precision = TC_FP32_BULK;
    else die("unknown precision " + o.precision);
    P.collocated = P.scheme == TC_SCHEME_AB3 || srk;
    P.nx = o.nx > 0 ? o.nx : (srk ? 128 : 64);
    P.ny = o.ny > 0 ? o.ny : (P.collocated ? 64 : 32);
    P.re_tau = o.re_tau > 0 ? o.re_tau : (srk ? 180.0 : 395.0);
    // reference units (staggered/main.cpp:200-205): u_tau = 1, nu = 1/Re_tau;
    // bulk units: U_b = 1, nu = 2/Re_b, u_tau = Re_tau nu
    P.visc = o.re_b > 0 ? 2.0 / o.re_b : 1.0 / P.re_tau;
    P.u_tau = P.re_tau * P.visc;
    P.sour_dpdx = P.u_tau * P.u_tau;
    P.Lx = o.lx > 0 ? o.lx : 4.0 * kPi;
    P.Ly = o.ly > 0 ? o.ly : (P.collocated ? 4.0 * kPi / 3.0 : 2.0 * kPi);
    P.hx = P.Lx / P.nx;
    P.hy = P.Ly / P.ny;
    P.xmin = 0.0;
    P.ymin = -0.5 * P.Ly;
    Options om = o;  // geometric mesh defaults: collocated_AB3/main.cpp:209-211, collocated_SRK/main.cpp:315-317
    if (om.geo_y1 <= 0) om.geo_y1 = srk ? 1.0 : 50.0;
    if (om.geo_ratio <= 0) om.geo_ratio = srk ? 1.3 : 1.05;
    if (om.geo_ybulk <= 0) om.geo_ybulk = srk ? 10.0 : 2.2;
    P.zz = make_mesh(om, P.collocated, P.nz, P.u_tau / P.visc);
    if (P.collocated) {
        for (int i = 0; i + 1 < P.nz; i++) P.dz.push_back(P.zz[i + 1] - P.zz[i]);
        P.area = P.Lx * P.Ly * (P.zz[P.nz - 1] - P.zz[0]);
    } else {
        for (int i = 0; i < P.nz; i++) P.dz.push_back(P.zz[i + 1] - P.zz[i]);
        P.area = P.Lx * P.Ly * (P.zz[P.nz] - P.zz[0]);
    }
    tc_params& p = P.params;
    p.scheme = P.scheme;
    p.precision = P.precision;
    p.nx = P.nx;
    p.ny = P.ny;
    p.nz = P.nz;
    const int les = o.les >= 0 ? o.les : (o.sgs_file.empty() && !srk ? 1 : 0);
    p.turb_visc = les;
    p.visc_type = o.visc >= 0 ? o.visc : (les ? 1 : 0);  // VISC_APPR_TYPE, VISC_TERM_GALERKIN
    p.stab_type = o.stab;
    p.fd_mode = o.fd_mode;
    p.visc = P.visc;
    p.sour_dpdx = P.sour_dpdx;
    p.inv_length_scale = P.u_tau / P.visc;
    p.hx = P.hx;
    p.hy = P.hy;
    p.xmin = P.xmin;
    p.ymin = P.ymin;
    p.cfr_ubulk = o.cfr;
    p.zz = P.zz.data();
    p.sgs_model = o.sgs_file.empty() ? 0 : o.sgs_proj ? 3 : o.sgs_clip ? 2 : 1;
    p.imex = imex;
    p.alpha_tau = o.alpha_tau;
    return P;
}

// ---------------------------------------------------------------------------
// The solver run, in the engine's real type

template <typename T>
struct Run {
    const Options& o;
    Problem& P;
    tc_engine* e = nullptr;
    size_t n3d;
    int nprof;                        // values per plane from tc_profiles
    std::vector<T> vel, pres, qres;   // host copies of one slot (real3 AoS, real); qres: srk dp/dt
    std::vector<T> av;                // last profiles, in real (the reference's av_u)
    std::vector<double> av_sum_d;     // time-integrated profiles, staggered (double)
    std::vector<T> av_sum_r;          // time-integrated profiles, collocated (real)
    double intergal_time = 0.0;
    double t = 0.0, tau;
    long long it = 0, it_start = 0;
    int cur = 0;
    double sec_step = 0, sec_red = 0, sec_out = 0;

    Run(const Options& o_, Problem& P_) : o(o_), P(P_), tau(o_.tau) {
        n3d = (size_t)P.nx * P.ny * P.nz;
        nprof = P.collocated ? 10 : 12;
        vel.assign(3 * n3d, T(0));
        pres.assign(n3d, T(0));
        qres.assign(n3d, T(0));
        av.assign((size_t)nprof * P.nz, T(0));
        av_sum_d.assign(av.size(), 0.0);
        av_sum_r.assign(av.size(), T(0));
    }

    // GetCoor (staggered/main.cpp:58-66): cell centre, with the variable's own
    // face coordinate in its direction.
    void coor(size_t in, int ivar, double c[3]) const {
        const int ix = in % P.nx, iy = (in / P.nx) % P.ny, iz = in / ((size_t)P.nx * P.ny);
        c[0] = (ix + 0.5) * P.hx;
        c[1] = (iy + 0.5) * P.hy;
        c[2] = 0.5 * (P.zz[iz] + P.zz[iz + 1]);
        if (ivar == 0) c[0] = ix * P.hx;
        if (ivar == 1) c[1] = iy * P.hy;
        if (ivar == 2) c[2] = P.zz[iz];
    }

    // The reference's analytic start: laminar parabola plus one pair of streamwise
    // vortex rolls (staggered/main.cpp:158-185, collocated_AB3/main.cpp:348-368).
    void analytic_start() {
        const double Hy = P.Ly;
        if (!P.collocated) {
            const double Hz = P.zz[P.nz] - P.zz[0], zmin = P.zz[0], H = 0.5 * Hz;
            for (size_t in = 0; in < n3d; in++) {
                double r[3];
                coor(in, 0, r);
                double zplus = std::min(r[2], Hz - r[2]) * P.u_tau / P.visc;
                // real = double expression, as the reference's compound assignments round
                T u0 = T(0.5 * (r[2] - zmin) * (zmin + Hz - r[2]) * P.sour_dpdx / P.visc);
                u0 = T(u0 - std::cos(r[0]) * std::sin(2.0 * kPi * r[1] / Hy) * zplus * std::exp(-0.01 * zplus * zplus));
                coor(in, 1, r);
                zplus = std::min(r[2], 2.0 * H - r[2]) * P.u_tau / P.visc;
                const T v0 = T(0.0 + std::sin(r[0]) * std::cos(2.0 * kPi * r[1] / Hy) * zplus *
                                         std::exp(-0.01 * zplus * zplus) / (2.0 * kPi / Hy));
                vel[3 * in] = u0, vel[3 * in + 1] = v0, vel[3 * in + 2] = T(0);
                pres[in] = T(0);
            }
        } else {
            const double zmin = 0.0, Hz = P.zz[P.nz - 1] - P.zz[0];
            const double re = P.u_tau / P.visc;
            for (size_t in = 0; in < n3d; in++) {
                const int ix = in % P.nx, iy = (in / P.nx) % P.ny, iz = in / ((size_t)P.nx * P.ny);
                const double x = ix * P.hx, y = iy * P.hy, z = P.zz[iz];
                const double zp = (P.zz[iz] - P.zz[0]) * re;
                const double zplus = std::min(zp, re * 2 - zp);
                T u0 = T(0.5 * (z - zmin) * (zmin + Hz - z) * P.sour_dpdx / P.visc);
                const T v0 = T(0.0 + std::sin(x) * std::cos(2.0 * kPi * y / Hy) * zplus *
                                         std::exp(-0.01 * zplus * zplus) / (2.0 * kPi / Hy));
                u0 = T(u0 - std::cos(x) * std::sin(2.0 * kPi * y / Hy) * zplus * std::exp(-0.01 * zplus * zplus));
                vel[3 * in] = u0, vel[3 * in + 1] = v0, vel[3 * in + 2] = T(0);
                pres[in] = T(0);
            }
        }
    }

    // The SRK start (collocated_SRK/main.cpp:441-451): laminar parabola, a
    // spanwise velocity 0.1 sin(x/2) z+ exp(-0.01 z+^2), and 0.1 rand()/RAND_MAX
    // added to every component away from the walls (the C library's rand(),
    // unseeded, as the reference).
    void srk_start() {
        const double zmin = 0.0, Hz = 2.0;
        const double re = P.u_tau / P.visc;
        for (size_t in = 0; in < n3d; in++) {
            const int ix = in % P.nx, iz = in / ((size_t)P.nx * P.ny);
            const double x = ix * P.hx, z = P.zz[iz];
            const double zp = (P.zz[iz] - P.zz[0]) * re;
            const double zplus = std::min(zp, re * 2 - zp);
            T u[3] = {T(0.5 * (z - zmin) * (zmin + Hz - z) * P.sour_dpdx / P.visc),
                      T(1e-1 * std::sin(0.5 * x) * zplus * std::exp(-0.01 * zplus * zplus)), T(0)};
            for (int d = 0; d < 3; d++)
                if (iz != 0 && iz != P.nz - 1) u[d] += T(1e-1 * double(rand()) / RAND_MAX);
            vel[3 * in] = u[0], vel[3 * in + 1] = u[1], vel[3 * in + 2] = u[2];
            pres[in] = T(0);
        }
    }

    // DumpData (staggered/io.cpp:6-29, collocated_AB3/io.cpp:9-32): ASCII VTK
    // rectilinear grid; z coordinates are the faces (staggered) or nodes. The SRK
    // solver's file has dp/dt as a fifth value (collocated_SRK/main.cpp:238-259).
    void dump_vtk(const char* fname) const {
        FILE* f = fopen(fname, "wt");
        if (!f) die(std::string("cannot write ") + fname);
        const int N[3] = {P.nx, P.ny, P.nz};
        fprintf(f, "# vtk DataFile Version 2.0\n");
        fprintf(f, "Volume example\n");
        fprintf(f, "ASCII\n");
        fprintf(f, "DATASET RECTILINEAR_GRID\n");
        fprintf(f, "DIMENSIONS %i %i %i\n", N[0], N[1], N[2]);
        for (int d = 0; d < 3; d++) {
            fprintf(f, "%c_COORDINATES %i float\n", 'X' + d, N[d]);
            for (int i = 0; i < N[d]; i++)
                fprintf(f, "%e ", d == 0 ? P.xmin + P.hx * i : d == 1 ? P.ymin + P.hy * i : P.zz[i]);
            fprintf(f, "\n");
        }
        const bool q = P.scheme == TC_SCHEME_SRK;
        fprintf(f, "POINT_DATA %i\n", (int)n3d);
        fprintf(f, "SCALARS volume_scalars float %i\n", q ? 5 : 4);
        fprintf(f, "LOOKUP_TABLE default\n");
        for (size_t in = 0; in < n3d; in++) {
            fprintf(f, "%e %e %e %e", (double)vel[3 * in], (double)vel[3 * in + 1], (double)vel[3 * in + 2],
                    (double)pres[in]);
            if (q) fprintf(f, " %e", (double)qres[in]);
            fprintf(f, "\n");
        }
        fclose(f);
    }

    // ReadData (staggered/io.cpp:31-69): a field stored on a grid coarser by
    // integer factors in x and y is repeated over the finer cells.
    bool read_vtk(const char* fname) {
        std::ifstream file(fname);
        if (!file.is_open()) return false;
        std::string line;
        int n[3] = {-1, -1, -1};
        bool data = false;
        while (std::getline(file, line)) {
            if (line.compare(0, 10, "DIMENSIONS") == 0) {
                if (sscanf(line.c_str() + 10, "%i%i%i", n, n + 1, n + 2) != 3) die("Unexpected EOF (no data read)");
                if (n[2] != P.nz) die("NZ mismatch");
                if (n[0] <= 0 || n[1] <= 0 || P.nx % n[0] || P.ny % n[1]) die("NX or NY is wrong");
            }
            if (line.compare(0, 10, "LOOKUP_TAB") == 0) {
                data = true;
                break;
            }
        }
        if (!data || n[0] < 0 || n[1] < 0 || n[2] < 0) die("Unexpected EOF (no data read)");
        printf("Stored data %ix%ix%i, reading to %ix%ix%i\n", n[0], n[1], n[2], P.nx, P.ny, P.nz);
        const int MX = P.nx / n[0], MY = P.ny / n[1];
        for (int iz = 0; iz < n[2]; iz++)
            for (int iy = 0; iy < n[1]; iy++)
                for (int ix = 0; ix < n[0]; ix++) {
                    if (!std::getline(file, line)) die("Unexpected EOF");
                    const size_t in = (size_t)iz * P.ny * P.nx + (size_t)iy * P.nx * MY + (size_t)ix * MX;
                    T q[4] = {T(0), T(0), T(0), T(0)};
                    const char* s = line.c_str();
                    for (int k = 0; k < 4; k++) {
                        char* end;
                        q[k] = sizeof(T) == 8 ? (T)strtod(s, &end) : (T)strtof(s, &end);
                        if (end == s) break;
                        s = end;
                    }
                    for (int jy = 0; jy < MY; jy++)
                        for (int jx = 0; jx < MX; jx++) {
                            const size_t jn = in + (size_t)jy * P.nx + jx;
                            vel[3 * jn] = q[0], vel[3 * jn + 1] = q[1], vel[3 * jn + 2] = q[2];
                            pres[jn] = q[3];
                        }
                }
        return true;
    }

    // DumpProfiles: staggered/io.cpp:71-96 (values at cell centres, u_fric = 1),
    // collocated_AB3/io.cpp:74-101 (values at nodes, scaled by u_tau; nu_t lives
    // on elements and is averaged to the nodes).
    template <typename A>
    void dump_profiles(const A* a, const char* fname) const {
        FILE* F = fopen(fname, "wt");
        if (!F) die(std::string("cannot write ") + fname);
        for (int iz = 0; iz < P.nz; iz++) {
            const A* r = a + (size_t)nprof * iz;
            double y_plus, ufr;
            if (!P.collocated) {
                ufr = 1.0;
                const double H = 0.5 * (P.zz[P.nz] - P.zz[0]);
                const double z = 0.5 * (P.zz[iz] + P.zz[iz + 1]) - P.zz[0];
                y_plus = (z < H ? z : 2.0 * H - z) * ufr / P.visc;
            } else {
                ufr = P.u_tau;
                const double H = 0.5 * (P.zz[P.nz - 1] - P.zz[0]);
                const double z = P.zz[iz];
                y_plus = (z < H ? z : 2.0 * H - z) * ufr / P.visc;
            }
            // staggered: the means in double; collocated: a real3 (real / real
            // and real products), as in the reference
            double u[3], uu[6], uu_sub[6];
            if (!P.collocated) {
                for (int i = 0; i < 3; i++) u[i] = r[i] / ufr;
                const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
                for (int i = 0; i < 6; i++) uu_sub[i] = u[pr[i][0]] * u[pr[i][1]];
            } else {
                A ua[3];
                for (int i = 0; i < 3; i++) ua[i] = r[i] * (A)(1.0 / (A)ufr);
                for (int i = 0; i < 3; i++) u[i] = ua[i];
                const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
                for (int i = 0; i < 6; i++) uu_sub[i] = (double)(A)(ua[pr[i][0]] * ua[pr[i][1]]);
            }
            for (int i = 0; i < 6; i++) uu[i] = r[3 + i] / (ufr * ufr);
            uu[0] = std::max(0., uu[0] - uu_sub[0]);
            uu[1] = std::max(0., uu[1] - uu_sub[1]);
            uu[2] = std::max(0., uu[2] - uu_sub[2]);
            uu[3] -= uu_sub[3];
            uu[4] -= uu_sub[4];
            uu[5] -= uu_sub[5];
            if (!P.collocated) {
                const double nu_t = r[9];
                const double pp = std::max(0., double(r[11]) - (double)(A)(r[10] * r[10]));
                fprintf(F, "%e %e %e %e %e %e %e %e\n", y_plus, u[0], std::sqrt(uu[0]), std::sqrt(uu[1]),
                        std::sqrt(uu[2]), uu[4], nu_t / P.visc, std::sqrt(pp));
            } else {
                double nu_t;
                if (iz == 0) nu_t = a[9];
                else if (iz == P.nz - 1) nu_t = a[(size_t)nprof * (iz - 1) + 9];
                else nu_t = 0.5 * (r[9] + a[(size_t)nprof * (iz - 1) + 9]);
                fprintf(F, "%e %e %e %e %e %e %e\n", y_plus, u[0], std::sqrt(uu[0]), std::sqrt(uu[1]),
                        std::sqrt(uu[2]), uu[4], nu_t / P.visc);
            }
        }
        fclose(F);
    }

    void download(int slot) {
        check(tc_get_field(e, TC_VEL, slot, vel.data(), vel.size() * sizeof(T)), "tc_get_field");
        check(tc_get_field(e, TC_PRES, slot, pres.data(), pres.size() * sizeof(T)), "tc_get_field");
        if (P.scheme == TC_SCHEME_SRK)  // dp/dt lives in pressure slot 1
            check(tc_get_field(e, TC_PRES, 1, qres.data(), qres.size() * sizeof(T)), "tc_get_field");
    }

    // ---- exact restart: every ring slot plus the loop state --------------
    struct Header {
        char magic[8];
        int32_t version, scheme, precision, real_size, nx, ny, nz, cur;
        int64_t it, it_start;
        double t, tau, intergal_time;
    };

    void write_restart() {
        const std::string tmp = o.checkpoint + ".tmp";
        FILE* f = fopen(tmp.c_str(), "wb");
        if (!f) die("cannot write " + tmp);
        Header h{};
        memcpy(h.magic, "TCRST01", 8);
        h.version = 1, h.scheme = P.scheme, h.precision = P.precision, h.real_size = (int)sizeof(T);
        h.nx = P.nx, h.ny = P.ny, h.nz = P.nz, h.cur = cur;
        h.it = it + 1, h.it_start = it_start;  // state before step it + 1
        h.t = t, h.tau = tau, h.intergal_time = intergal_time;
        bool ok = fwrite(&h, sizeof h, 1, f) == 1;
        const size_t nz1 = P.zz.size();
        ok = ok && fwrite(P.zz.data(), sizeof(double), nz1, f) == nz1;
        ok = ok && fwrite(av.data(), sizeof(T), av.size(), f) == av.size();
        ok = ok && fwrite(av_sum_d.data(), sizeof(double), av_sum_d.size(), f) == av_sum_d.size();
        ok = ok && fwrite(av_sum_r.data(), sizeof(T), av_sum_r.size(), f) == av_sum_r.size();
        std::vector<T> buf(3 * n3d);
        const int kinds[3] = {TC_VEL, TC_FLUX, TC_PRES};
        for (int k : kinds)
            for (int s = 0; s < 4; s++) {
                const size_t cnt = (k == TC_PRES ? 1 : 3) * n3d;
                check(tc_get_field(e, k, s, buf.data(), cnt * sizeof(T)), "tc_get_field");
                ok = ok && fwrite(buf.data(), sizeof(T), cnt, f) == cnt;
            }
        ok = (fclose(f) == 0) && ok;
        if (!ok) die("cannot write " + tmp);
        if (rename(tmp.c_str(), o.checkpoint.c_str()) != 0) die("cannot rename " + tmp);
    }

    void read_restart() {
        FILE* f = fopen(o.restart.c_str(), "rb");
        if (!f) die("cannot open " + o.restart);
        Header h{};
        bool ok = fread(&h, sizeof h, 1, f) == 1;
        if (!ok || memcmp(h.magic, "TCRST01", 8) != 0 || h.version != 1) die(o.restart + " is not a restart file");
        if (h.scheme != P.scheme || h.precision != P.precision || h.real_size != (int)sizeof(T) || h.nx != P.nx ||
            h.ny != P.ny || h.nz != P.nz)
            die(o.restart + " was written for another scheme, precision or grid");
        std::vector<double> zz(P.zz.size());
        ok = fread(zz.data(), sizeof(double), zz.size(), f) == zz.size();
        if (!ok || zz != P.zz) die(o.restart + " was written on another wall-normal mesh");
        ok = fread(av.data(), sizeof(T), av.size(), f) == av.size();
        ok = ok && fread(av_sum_d.data(), sizeof(double), av_sum_d.size(), f) == av_sum_d.size();
        ok = ok && fread(av_sum_r.data(), sizeof(T), av_sum_r.size(), f) == av_sum_r.size();
        std::vector<T> buf(3 * n3d);
        const int kinds[3] = {TC_VEL, TC_FLUX, TC_PRES};
        for (int k : kinds)
            for (int s = 0; s < 4 && ok; s++) {
                const size_t cnt = (k == TC_PRES ? 1 : 3) * n3d;
                ok = fread(buf.data(), sizeof(T), cnt, f) == cnt;
                if (ok) check(tc_set_field(e, k, s, buf.data(), cnt * sizeof(T)), "tc_set_field");
            }
        fclose(f);
        if (!ok) die(o.restart + " is truncated");
        cur = h.cur, it = h.it, it_start = h.it_start;
        t = h.t, tau = h.tau, intergal_time = h.intergal_time;
    }

    void load_sgs() {
        const int np = tc_mlp_param_count();
        std::vector<float> w(np);
        FILE* f = fopen(o.sgs_file.c_str(), "rb");
        if (!f) die("cannot open " + o.sgs_file);
        const size_t got = fread(w.data(), sizeof(float), np, f);
        const bool extra = fgetc(f) != EOF;
        fclose(f);
        if (got != (size_t)np || extra)
            die(o.sgs_file + ": expected " + std::to_string(np) + " float32 network parameters");
        check(tc_set_sgs_params(e, w.data(), w.size()), "tc_set_sgs_params");
    }

    void report_timing() {
        if (!o.timing) return;
        static const char* names[TC_K_COUNT] = {"turb_visc", "flux", "div", "fft_x", "fft_y",
                                                "zsolve", "ifft_y", "ifft_x", "project", "other"};
        printf("Kernels (ms):");
        for (int k = 0; k < TC_K_COUNT; k++) {
            double ms = 0;
            long long n = 0;
            check(tc_timing_read(e, k, &ms, &n), "tc_timing_read");
            if (n) printf(" %s %.3f", names[k], ms);
        }
        printf("\n");
    }

    // The SRK time loop (collocated_SRK/main.cpp:472-550): a fixed step, the bulk
    // velocity and the friction velocity from the first node plane, averaging
    // from t = 0 with t = (step + 1) tau, fields and profiles every FProgress
    // steps and at the final step. Velocity in slot 0, pressure in pressure slot
    // 0, dp/dt (StabType 1) in pressure slot 1.
    int srk_loop(FILE* log) {
        const int sprog = o.s_progress > 0 ? o.s_progress : 10;
        const int fprog = o.f_progress > 0 ? o.f_progress : 10000;
        const int pprog = o.p_progress > 0 ? o.p_progress : fprog;
        std::vector<double> prof((size_t)nprof * P.nz);
        printf("Start of time integration (N=%i,%i,%i)\n", P.nx, P.ny, P.nz);
        const double t_total0 = now();
        for (bool final_step = false; !final_step; it++) {
            const double t0 = now();
            check(tc_srk_step(e, it * tau, tau, 0, 0, 1), "tc_srk_step");
            check(tc_sync(e), "tc_sync");
            const double t1 = now();
            sec_step += t1 - t0;
            double q[3];
            check(tc_diagnostics(e, 0, 0, tau, 3, q, nullptr, nullptr, prof.data(), prof.size()), "tc_diagnostics");
            const double flow_rate = q[0] / P.area;
            for (size_t i = 0; i < prof.size(); i++) av[i] = T(prof[i]);
            for (size_t i = 0; i < av.size(); i++) av_sum_r[i] += T(tau) * av[i];
            const double t3 = now();
            sec_red += t3 - t1;

            t = (it + 1) * tau;
            intergal_time = t;
            // the reference's loop ends at TimeMax only (its step limit is fixed at 10^6);
            // here the step limit also makes the final step, as in the other loops
            if (t >= o.time_max || it + 1 == o.time_steps_max) final_step = true;
            const double dudy_cur = (av[nprof] - 0.) / P.dz[0];
            const double ufr_cur = dudy_cur < 0. ? 0. : std::sqrt(dudy_cur * P.visc);
            const double dudy_int = av_sum_r[nprof] / (P.dz[0] * t);
            const double ufr_int = dudy_int < 0. ? 0. : std::sqrt(dudy_int * P.visc);
            if (it % sprog == 0 || final_step) {
                printf("T=%f Ubulk=%.04f Ufric=%.04f UfricAv=%.04f\n", t, flow_rate, ufr_cur, ufr_int);
                fprintf(log, "%f %f %f %f\n", t, flow_rate, ufr_cur, ufr_int);
                fflush(stdout);
            }
            if ((it != 0 && it % fprog == 0) || final_step) {
                char fname[64];
                snprintf(fname, sizeof fname, "q%05lld.vtk", it);
                download(0);
                dump_vtk(fname);
            }
            if ((it != 0 && it % pprog == 0) || final_step) {
                char fname[64];
                snprintf(fname, sizeof fname, "r%05lld.dat", it);
                dump_profiles(av.data(), fname);
            }
            sec_out += now() - t3;
            if (o.checkpoint_every > 0 && ((it + 1) % o.checkpoint_every == 0 || final_step)) write_restart();
        }
        fclose(log);
        check(tc_sync(e), "tc_sync");
        printf("Time: total %.04f, step %.04f, diagnostics %.04f, output %.04f\n", now() - t_total0, sec_step, sec_red,
               sec_out);
        report_timing();
        download(0);
        dump_vtk("output.vtk");
        if (t > 0.) {
            for (T& v : av_sum_r) v *= T(1. / t);
            dump_profiles(av_sum_r.data(), "res.dat");
            const double a1 = av_sum_r[nprof], a2 = av_sum_r[2 * nprof];
            const double dudy = 1.5 * (a1 - 0.) / P.dz[0] - 0.5 * (a2 - a1) / P.dz[1];
            const double u_fric = dudy < 0. ? 0. : std::sqrt(P.visc * dudy);
            printf("u_fric: obtained=%.6f, expected=%.6f\n", u_fric, P.u_tau);
        }
        check(tc_destroy(e), "tc_destroy");
        return 0;
    }

    static double now() {
        using namespace std::chrono;
        return duration<double>(steady_clock::now().time_since_epoch()).count();
    }

    int main_loop() {
        check(tc_create(&P.params, o.device, &e), "tc_create");
        if ((size_t)tc_real_size(e) != sizeof(T)) die("engine real size mismatch");
        if (!o.sgs_file.empty()) load_sgs();
        if (o.timing) check(tc_timing_enable(e, 1), "tc_timing_enable");
        check(tc_set_graphs(e, o.graphs), "tc_set_graphs");

        FILE* log = nullptr;
        if (!o.restart.empty()) {
            read_restart();
            printf("Restarted from %s at step %lld, T=%f\n", o.restart.c_str(), it, t);
            log = fopen("log.dat", "at");
        } else {
            if (read_vtk("input.vtk")) {
                printf("Initial data read\n");
                for (int s = 0; s < 4; s++) {
                    std::vector<T> zv(3 * n3d, T(0)), zp(n3d, T(0));
                    check(tc_set_field(e, TC_VEL, s, s == 0 ? vel.data() : zv.data(), vel.size() * sizeof(T)),
                          "tc_set_field");
                    check(tc_set_field(e, TC_PRES, s, s == 0 ? pres.data() : zp.data(), pres.size() * sizeof(T)),
                          "tc_set_field");
                }
            } else {
                printf("input.vtk not found. Starting from the beginning\n");
                if (o.seed >= 0) {
                    std::vector<T> zv(3 * n3d, T(0)), zp(n3d, T(0));
                    for (int s = 0; s < 4; s++) {
                        check(tc_set_field(e, TC_VEL, s, zv.data(), zv.size() * sizeof(T)), "tc_set_field");
                        check(tc_set_field(e, TC_PRES, s, zp.data(), zp.size() * sizeof(T)), "tc_set_field");
                    }
                    check(tc_initial_field(e, 0, (unsigned long long)o.seed, o.ubulk, o.noise), "tc_initial_field");
                } else {
                    if (P.scheme == TC_SCHEME_SRK) srk_start();
                    else analytic_start();
                    std::vector<T> zv(3 * n3d, T(0)), zp(n3d, T(0));
                    for (int s = 0; s < 4; s++) {
                        check(tc_set_field(e, TC_VEL, s, s == 0 ? vel.data() : zv.data(), vel.size() * sizeof(T)),
                              "tc_set_field");
                        check(tc_set_field(e, TC_PRES, s, s == 0 ? pres.data() : zp.data(), pres.size() * sizeof(T)),
                              "tc_set_field");
                    }
                }
            }
            log = fopen("log.dat", "wt");
        }
        if (!log) die("cannot open log.dat");
        check(tc_sync(e), "tc_sync");
        if (P.scheme == TC_SCHEME_SRK) return srk_loop(log);
        check(tc_step_diagnostics(e, o.fused_diag), "tc_step_diagnostics");

        const int sprog = o.s_progress > 0 ? o.s_progress : (P.collocated ? 10 : 200);
        const int fprog = o.f_progress > 0 ? o.f_progress : (P.collocated ? 10000 : 1000000);
        const int pprog = o.p_progress > 0 ? o.p_progress : 100000;
        const int max_bdf_order = 3;
        std::vector<double> prof((size_t)nprof * P.nz);

        printf("Start of time integration (N=%i,%i,%i)\n", P.nx, P.ny, P.nz);
        const double t_total0 = now();
        // Pipelining: step it+1 is launched before step it's diagnostics are read
        // (they wait for step it only), unless step it ends the run or writes a
        // field or a checkpoint. It writes only the ring slot after it and what it
        // recomputes from its input, so if the CFL control then changes tau it is
        // launched again from the same slot with the new tau at order 1.
        const bool pipe = o.pipeline && o.fused_diag;
        bool launched = false;  // step it already launched
        for (bool final_step = false; !final_step; it++) {
            double t0 = now();
            const int order = (int)std::min<long long>(it - it_start + 1, max_bdf_order);
            if (!launched) check(tc_step(e, it * tau, tau, cur, order), "tc_step");
            launched = false;
            cur = (cur + 1) % 4;
            const bool ends = t + tau >= o.time_max || it + 1 == o.time_steps_max;
            const bool writes = (it != 0 && it % fprog == 0) ||
                                (o.checkpoint_every > 0 && (it + 1) % o.checkpoint_every == 0);
            if (pipe && !ends && !writes) {
                const int order1 = (int)std::min<long long>(it + 1 - it_start + 1, max_bdf_order);
                check(tc_step(e, (it + 1) * tau, tau, cur, order1), "tc_step");
                launched = true;
            } else {
                check(tc_sync(e), "tc_sync");
            }
            double t1 = now();
            sec_step += t1 - t0;

            // CalcIntegral, CalcProfiles and CalcCFL of the new field in one pass
            double q[3], cfl;
            check(tc_diagnostics(e, cur, cur, tau, 3, q, nullptr, &cfl, prof.data(), prof.size()), "tc_diagnostics");
            // CalcIntegral(...) / area: a real3 divided by a real
            const T flow_rate = T(T(q[0]) * T(1.0 / (double)T(P.area)));
            for (size_t i = 0; i < prof.size(); i++) av[i] = T(prof[i]);
            double t3 = now();
            sec_red += t3 - t1;

            if (t >= o.time_start_averaging) {
                for (size_t i = 0; i < av.size(); i++) {
                    if (P.collocated) av_sum_r[i] += T(tau) * av[i];
                    else av_sum_d[i] += (double)(T(tau) * av[i]);
                }
                intergal_time += tau;
            }

            t += tau;
            if (t >= o.time_max || it + 1 == o.time_steps_max) final_step = true;
            double dudy_cur, dudy_int = 0.0;
            if (!P.collocated) {
                dudy_cur = (av[0] - 0.) / (0.5 * P.dz[0]);
                if (intergal_time > 0.) dudy_int = av_sum_d[0] / (0.5 * P.dz[0] * intergal_time);
            } else {
                dudy_cur = (av[nprof] - 0.) / P.dz[0];
                if (intergal_time > 0) dudy_int = av_sum_r[nprof] / (P.dz[0] * intergal_time);
            }
            const double ufr_cur = dudy_cur < 0. ? 0. : std::sqrt(dudy_cur * P.visc);
            const double ufr_int = dudy_int < 0. ? 0. : std::sqrt(dudy_int * P.visc);
            if (it % sprog == 0 || final_step) {
                printf("T=%f Ubulk=%.04f tau=%.1e CFL=%.04f Ufric=%.04f UfricAv=%.04f\n", t, (double)flow_rate, tau,
                       cfl, ufr_cur, ufr_int);
                fprintf(log, "%f %f %e %f %f %f\n", t, (double)flow_rate, tau, cfl, ufr_cur, ufr_int);
                fflush(stdout);
            }
            if (it != 0 && it % fprog == 0) {
                char fname[64];
                snprintf(fname, sizeof fname, "q%05lld.vtk", it);
                download(cur);
                dump_vtk(fname);
            }
            if (it != 0 && it % pprog == 0) {
                char fname[64];
                snprintf(fname, sizeof fname, "r%05lld.dat", it);
                dump_profiles(av.data(), fname);
            }
            const double tau_used = tau;
            while (cfl > o.cfl_max) {
                cfl *= 0.5;
                tau *= 0.5;
                it_start = it + 1;
            }
            if (cfl < o.cfl_min) {
                cfl *= 1.5;
                tau *= 1.5;
                it_start = it + 1;
            }
            if (launched && tau != tau_used)  // step it+1 ran with the old tau: again, order 1
                check(tc_step(e, (it + 1) * tau, tau, cur, 1), "tc_step");
            sec_out += now() - t3;
            if (o.checkpoint_every > 0 && ((it + 1) % o.checkpoint_every == 0 || final_step)) write_restart();
        }
        fclose(log);
        check(tc_sync(e), "tc_sync");
        const double total = now() - t_total0;
        if (!P.collocated) {
            printf("Total time: %.04f\n", total);
            printf("Channels: step %.04f, diagnostics %.04f, output %.04f\n", sec_step, sec_red, sec_out);
        } else {
            printf("Time: total %.04f, step %.04f, diagnostics %.04f, output %.04f\n", total, sec_step, sec_red,
                   sec_out);
        }
        if (pipe)  // the host clock sees launches only; each wait lands in tc_diagnostics
            printf("(pipelined loop: the step channel times the launches, the diagnostics channel includes the "
                   "wait for each step's kernels; --pipeline 0 separates them)\n");
        report_timing();

        download(cur);
        dump_vtk("output.vtk");
        dump_profiles(av.data(), "output.dat");
        if (intergal_time > 0.) {
            double u_fric;
            if (!P.collocated) {
                for (double& v : av_sum_d) v *= 1. / intergal_time;
                dump_profiles(av_sum_d.data(), "res.dat");
                const double dudy = (av_sum_d[0] - 0.) / (0.5 * P.dz[0]);
                u_fric = dudy < 0. ? 0. : std::sqrt(P.visc * dudy);
            } else {
                for (T& v : av_sum_r) v *= T(1. / intergal_time);
                dump_profiles(av_sum_r.data(), "res.dat");
                const double a1 = av_sum_r[nprof], a2 = av_sum_r[2 * nprof];
                const double dudy = 1.5 * (a1 - 0.) / P.dz[0] - 0.5 * (a2 - a1) / P.dz[1];
                u_fric = dudy < 0. ? 0. : std::sqrt(P.visc * dudy);
            }
            printf("u_fric: obtained=%.6f, expected=%.6f\n", u_fric, P.u_tau);
        }
        check(tc_destroy(e), "tc_destroy");
        return 0;
    }
};

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) return 0;
    // collocated_SRK/main.cpp:437-438: TimeMax 0.1, tau 0.8e-3
    if (o.time_max < 0) o.time_max = o.scheme == "srk" ? 0.10000000001 : 100.0;
    if (o.tau < 0) o.tau = o.scheme == "srk" ? 0.8e-3 : 1e-3;
    if (o.tau <= 0 || o.cfl_min < 0 || o.cfl_max <= 0) die("tau and CFLmax must be positive");
    if (o.s_progress == 0 || o.f_progress == 0 || o.p_progress == 0) die("progress intervals must be positive");
    Problem P = setup(o);
    if (o.print_mesh) {
        printf("nz %d\n", P.nz);
        for (double z : P.zz) printf("%.17g\n", z);
        return 0;
    }
    if (P.precision == TC_FP64) return Run<double>(o, P).main_loop();
    return Run<float>(o, P).main_loop();
}
