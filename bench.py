"""Benchmark of the MI355X time stepper on one GPU.

Default workload: the collocated BDF3/EX3 ("AB3") DNS on a 384 x 384 x 192
grid (x, span, wall-normal nodes), tanh wall-normal stretching 2.0, box
2 pi x pi x 2, bulk Re 13750 held by the constant-flow-rate correction,
seeded divergence-free initial field, dt = 5e-4, FP32 bulk arithmetic (the
collocated solver's own `real`, collocated_AB3/config.h:4-10) with FP64 kept
for the low-mode z-solves, the normalisation, the flow rate and the
statistics. A step is one tc_step (explicit term + BDF3/EX3, pressure solve,
projection); the fields are resident in HBM for the whole timed region.

Prints one JSON line with the metric (grid-point steps per second), the
roofline of the dominant kernel (measured live with HIP events on the engine's
stream), the CPU baseline (the FP64 OpenMP oracle on the host cores, on a
bounded number of steps of the same workload) and, under "extra_workloads",
the step rates of the fourth-order staggered DNS at 384 x 384 x 192 (FP32
bulk) and 192 x 160 x 128 (FP64).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import turbchannel_amd as tca  # noqa: E402

PEAK_HBM_GBS = 8000.0
# Algorithmic reals moved per grid point by each kernel class (each array
# element counted once; halos and cache reuse not counted).
REALS_PER_POINT = {
    "flux": 21,      # reads u_n, u_n-1, u_n-2, F_n-1, F_n-2 (15); writes F_n, r (6)
    "fft_x": 4,      # divergence fused in: reads r (3); writes the half spectrum (~1)
    "fft_y": 2,
    "zsolve": 2,     # reads and writes the spectrum once each
    "ifft_y": 2,
    "ifft_x": 2,
    "project": 7,    # reads p (1), r (3); writes u_n+1 (3)
}
# Collocated BDF3/EX3 with StabType 1: the explicit term also reads p_n (its
# gradient enters r), the projection reads the increment, r and p_n and writes
# u_n+1 and p_n+1.
REALS_PER_POINT_AB3 = dict(REALS_PER_POINT, flux=22, project=9)

# The SRK stepper's explicit term is F(u) alone (reads u, writes F); its stage
# combinations (the "project" class) move a number of reals that depends on the
# stage and the method, so they get no figure.
REALS_PER_POINT_SRK = dict(REALS_PER_POINT, flux=6)
REALS_PER_POINT_SRK.pop("project")

PEAK_BF16_TFLOPS = 2500.0  # dense bfloat16 MFMA
# flops per sample of the closure network (10 -> 64 -> 64 -> 64 -> 6): forward, and
# forward + backward (weight gradients of every layer, data gradients of layers 2-4)
MLP_FWD_FLOP = 2 * (10 * 64 + 64 * 64 + 64 * 64 + 64 * 6)
MLP_TRAIN_FLOP = 2 * MLP_FWD_FLOP + 2 * (64 * 6 + 64 * 64 + 64 * 64)

WORKLOADS = {
    # name: scheme, nx, ny, nz, gamma, Re_b, Re_tau, dt, extra engine parameters
    "vv_192x160x128": (tca.SCHEME_VV, 192, 160, 128, 1.8, 5600.0, 180.0, 1e-3, {}),
    "hw_64x64x64": (tca.SCHEME_HW, 64, 64, 64, 1.8, 5600.0, 180.0, 2e-3, {}),
    "ab3_384x384x192": (tca.SCHEME_AB3, 384, 384, 192, 2.0, 13750.0, 395.0, 5e-4, {}),
    "vv_384x384x192": (tca.SCHEME_VV, 384, 384, 192, 2.0, 13750.0, 395.0, 5e-4, {}),
    # collocated LES at Re_tau 590, 64 x 64 x 96 nodes: learned closure, and the
    # reference's Smagorinsky-van Driest model on the same grid
    "les_learned_64x64x96": (tca.SCHEME_AB3, 64, 64, 96, 2.0, 21900.0, 590.0, 2e-3, dict(visc_type=1, sgs_model=1)),
    # the learned closure kept as its dissipative part along the strain rate (eddy viscosity)
    "les_learned_eddy_64x64x96": (tca.SCHEME_AB3, 64, 64, 96, 2.0, 21900.0, 590.0, 2e-3, dict(visc_type=1, sgs_model=3)),
    "les_smag_64x64x96": (tca.SCHEME_AB3, 64, 64, 96, 2.0, 21900.0, 590.0, 2e-3, dict(visc_type=1, turb_visc=1)),
    # collocated IMEX Runge-Kutta (SRK) stepper, ARS(3,4,3) with StabType 1 as in
    # collocated_SRK/main.cpp:342-352: the reference's grid size (128 x 64 x 49
    # nodes) and a large grid
    "srk_128x64x49": (tca.SCHEME_SRK, 128, 64, 49, 2.0, 5600.0, 180.0, 8e-4, dict(imex=2)),
    "srk_384x384x192": (tca.SCHEME_SRK, 384, 384, 192, 2.0, 13750.0, 395.0, 5e-4, dict(imex=2)),
}
TRAIN_WORKLOAD = "closure_train_b65536"


def problem(name, precision):
    scheme, nx, ny, nz, gamma, re_b, re_tau, dt, extra = WORKLOADS[name]
    collocated = scheme in (tca.SCHEME_AB3, tca.SCHEME_SRK)
    if collocated:
        k = np.arange(nz)
        zz = 1.0 + np.tanh(gamma * (2.0 * k / (nz - 1) - 1.0)) / math.tanh(gamma)
    else:
        k = np.arange(nz + 1)
        zz = 1.0 + np.tanh(gamma * (2.0 * k / nz - 1.0)) / math.tanh(gamma)
    zz[0], zz[-1] = 0.0, 2.0
    visc = 2.0 / re_b
    u_tau = re_tau * visc
    kw = dict(visc=visc, sour_dpdx=u_tau * u_tau, inv_length_scale=u_tau / visc, hx=2 * math.pi / nx,
              hy=math.pi / ny, cfr_ubulk=1.0, precision=precision)
    if collocated:
        kw.update(stab_type=1, fd_mode=1, visc_type=0)
    if scheme == tca.SCHEME_SRK:  # no flow-rate control in the SRK solver
        kw["cfr_ubulk"] = 0.0
    kw.update(extra)
    return scheme, nx, ny, nz, zz, kw, dt


def closure_weights(seed=7):
    """Closure parameters for timing runs: the trainer's seeded initial network
    with its output layer scaled by 1e-2 (same cost as a trained one)."""
    net = tca.Mlp(seed=seed)
    p = net.get_params()
    net.close()
    n_out = 6 * 64 + 6
    p[-n_out:] *= 1e-2
    return p


def bench_training(args):
    """Closure training throughput: Adam steps on batches of 65536 samples."""
    import torch  # device memory only

    batch, nbatch = 65536, 16
    rng = np.random.default_rng(0)
    X = rng.normal(size=(batch * nbatch, 10)).astype(np.float32)
    T = rng.normal(size=(batch * nbatch, 6)).astype(np.float32)
    x = torch.from_numpy(X).to("cuda")
    t = torch.from_numpy(T).to("cuda")
    torch.cuda.synchronize()
    net = tca.Mlp(seed=0)
    step = lambda k: net.train_step(x.data_ptr() + 40 * batch * (k % nbatch),  # noqa: E731
                                    t.data_ptr() + 24 * batch * (k % nbatch), batch, lr=1e-3)
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    sps = batch * args.steps / el
    tflops = sps * MLP_TRAIN_FLOP / 1e12
    rec = {
        "metric": "closure training samples/s", "value": sps, "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulation, fp32 Adam)", "data": "synthetic (normal inputs and targets)",
        "config": {"workload": TRAIN_WORKLOAD, "network": "10-64-64-64-6 GELU", "batch": batch,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "k_mlp_train + reduce/Adam (one step)", "achieved": round(tflops, 2),
                     "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / PEAK_BF16_TFLOPS, 5),
                     "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(rec))


def device_copy_rate(nbytes=256 << 20, reps=20):
    """Read + write rate of a plain device-to-device copy of nbytes (torch), the
    practical ceiling of a streaming kernel on this GPU; None without torch."""
    try:
        import torch

        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del a, b
        gbs = 2 * nbytes / (ms * 1e-3) / 1e9
        return {"bytes": nbytes, "gbs": round(gbs, 1), "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4)}
    except Exception as exc:  # no torch device: reported, not fatal
        return {"error": repr(exc)}


def host_cores():
    """(physical cores, logical CPUs) of the host from /proc/cpuinfo, and the CPUs
    this process may run on."""
    phys, logical = set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                if line.startswith("processor"):
                    logical += 1
                elif line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":")[1].strip()
                    phys.add((pid, cid))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"physical": len(phys) or None, "logical": logical or None, "allowed": allowed}


def baseline_threads():
    """Threads for the CPU baseline: every core this job may use. A GPU job on this
    host gets a share of 16 cores (OMP_NUM_THREADS there), whatever the machine has."""
    allowed = host_cores()["allowed"]
    env = os.environ.get("OMP_NUM_THREADS")
    share = int(env) if env and env.isdigit() and int(env) > 0 else 16
    return max(1, min(allowed, share))


def cpu_baseline(name, seconds_target=15.0, threads=None, shipped_threads=0, shipped_seconds=5.0):
    """FP64 OpenMP oracle on the host cores, on a bounded number of steps; with
    shipped_threads, also the same steps on the reference's own thread count."""
    import oracle

    scheme, nx, ny, nz, zz, kw, dt = problem(name, tca.FP64)
    kw.pop("precision")
    threads = threads or baseline_threads()
    oracle.set_threads(threads)
    srk = scheme == tca.SCHEME_SRK
    method = kw.pop("imex", 2)
    sol = oracle.Solver(scheme, nx, ny, nz, zz, **kw)
    u0 = sol.project(sol.initial_field(seed=42))
    if srk:  # SRK: the oracle's or_srk_step on the same state, timed the same way
        st = [u0, np.zeros(sol.n3d), np.zeros(sol.n3d)]
        t0 = time.perf_counter()
        n = 0
        while True:
            st = list(sol.srk_step(n * dt, dt, *st, method=method, stab_type=kw["stab_type"], alpha_tau=1.0))
            n += 1
            el = time.perf_counter() - t0
            if el > seconds_target:
                break
        return {"value": nx * ny * nz * n / el, "unit": "grid-pt steps/s", "cores": threads, "kind": "port",
                "sample": f"{n} SRK steps of {name} (FP64 oracle, OpenMP {threads} threads, {el:.1f} s)"}
    sol.set_field(oracle.KIND_VEL, 0, u0)
    closure = None
    if kw.get("sgs_model"):  # learned closure: the network forward (numpy) feeds each step's tau
        from oracle import mlp

        p = closure_weights()
        u_tau = kw["visc"] * kw["inv_length_scale"]
        closure = lambda u: mlp.node_tau(p, u, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau,  # noqa: E731
                                         kw["inv_length_scale"], emulate=False, clip=kw["sgs_model"] == 2,
                                         project=kw["sgs_model"] == 3)

    def step(it, cur, order):
        if closure is not None:
            sol.set_field(oracle.KIND_TAU, 0, closure(sol.get_field(oracle.KIND_VEL, cur)))
        sol.step(it * dt, dt, cur, order)

    cur = 0
    for it in range(3):  # order ramp, untimed
        step(it, cur, it + 1)
        cur = (cur + 1) % 4
    sol.timers(reset=True)
    t0 = time.perf_counter()
    n = 0
    while True:
        step(n, cur, 3)
        cur = (cur + 1) % 4
        n += 1
        el = time.perf_counter() - t0
        if el > seconds_target:
            break
    out = {"value": nx * ny * nz * n / el, "unit": "grid-pt steps/s", "cores": threads, "kind": "port",
           "sample": f"{n} BDF3 steps of {name} (FP64 oracle, OpenMP {threads} threads, {el:.1f} s)",
           # the reference's timer channels (explicit, grad, pressure F/T/N), seconds per step
           "channels_s_per_step": {k: round(v / n, 6) for k, v in sol.timers().items()}}
    if shipped_threads:
        # the reference's own OpenMP setting, timed on the same state for at least one step
        oracle.set_threads(shipped_threads)
        t0 = time.perf_counter()
        n1 = 0
        while True:
            step(n + n1, cur, 3)
            cur = (cur + 1) % 4
            n1 += 1
            el1 = time.perf_counter() - t0
            if el1 > shipped_seconds:
                break
        out["as_shipped"] = {"value": nx * ny * nz * n1 / el1, "cores": shipped_threads,
                             "sample": f"{n1} BDF3 steps, OpenMP {shipped_threads} threads, {el1:.1f} s"}
    return out


# Kernel-class name -> kernel symbol prefix in the rocprofv3 summaries.
PROFILE_SYMBOL = {"flux": "k_flux", "fft_x": "k_fft_x_fwd_div", "fft_y": "k_fft_y", "zsolve": "k_zsolve",
                  "ifft_x": "k_fft_x_inv", "project": "k_project"}


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/step_kernels_<workload>_<precision>.md: FETCH_SIZE and WRITE_SIZE,
    each scaled by its reading on a plain device copy), or None if absent."""
    path = os.path.join(ROOT, "profiles", f"step_kernels_{workload}_{precision}.md")
    sym = PROFILE_SYMBOL.get(kernel)
    if sym is None or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) >= 7 and cols[0].startswith(sym + "<"):
                try:
                    return {"bytes": (float(cols[5]) + float(cols[6])) * 1e6, "source": os.path.relpath(path, ROOT)}
                except ValueError:
                    return None
    return None


# Precision a workload runs in unless --precision says otherwise: the collocated
# solver's own real is float (collocated_AB3/config.h:4-10), the staggered one's
# double (staggered/config.h:5-10); the 384^2 staggered grid is the FP32-bulk case.
DEFAULT_PRECISION = {"ab3_384x384x192": "fp32", "vv_384x384x192": "fp32", "srk_384x384x192": "fp32",
                     "les_learned_64x64x96": "fp32", "les_learned_eddy_64x64x96": "fp32", "les_smag_64x64x96": "fp32"}
DEFAULT_WORKLOAD = "ab3_384x384x192"
# step rates reported beside the default line (no CPU baseline, no roofline of their own)
EXTRA_WORKLOADS = [("vv_384x384x192", "fp32"), ("vv_192x160x128", "fp64")]
# OpenMP threads of the reference's own programs (staggered/main.cpp:279,
# collocated_AB3/main.cpp:380)
SHIPPED_THREADS = {tca.SCHEME_HW: 1, tca.SCHEME_VV: 1, tca.SCHEME_AB3: 2, tca.SCHEME_SRK: 2}


def run_workload(name, precision_name, steps, warmup, graphs=True, diagnostics=True):
    """Times `steps` steps of one workload on the GPU. Returns the engine-side
    figures: steps/s, per-kernel durations from HIP events and the roofline of the
    dominant kernel."""
    precision = tca.FP64 if precision_name == "fp64" else tca.FP32_BULK
    scheme, nx, ny, nz, zz, kw, dt = problem(name, precision)
    n3d = nx * ny * nz
    eng = tca.Engine(scheme, nx, ny, nz, zz, device=0, **kw)
    if kw.get("sgs_model"):
        eng.set_sgs_params(closure_weights())
    eng.initial_field(0, seed=42, ubulk=1.0, noise_rms=0.1)
    eng.graphs(graphs)
    srk = scheme == tca.SCHEME_SRK
    reals = REALS_PER_POINT_SRK if srk else REALS_PER_POINT_AB3 if scheme == tca.SCHEME_AB3 else REALS_PER_POINT
    cur = 0

    def advance(it, order=3):
        """One step; returns the slot of the new velocity."""
        nonlocal cur
        if srk:  # in place: velocity slot 0, pressure slot 0, dp/dt in pressure slot 1
            eng.srk_step(it * dt, dt, 0, 0, 1)
            return 0
        eng.step(it * dt, dt, cur, order)
        cur = (cur + 1) % 4
        return cur

    # warm-up, long enough to capture the four order-3 steps (one per ring slot)
    for it in range(max(warmup, 7)):
        advance(it, min(it + 1, 3))
    eng.sync()

    try:
        import torch

        sync_torch = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync_torch = lambda: None  # noqa: E731

    # timed region: the steps only (replayed HIP graphs unless graphs=False)
    eng.sync()
    sync_torch()
    t0 = time.perf_counter()
    for it in range(steps):
        advance(it)
    eng.sync()
    sync_torch()
    el = time.perf_counter() - t0

    # per-kernel durations: the same steps launched kernel by kernel with HIP
    # events on the engine's stream (events bypass the graphs)
    eng.timing_reset()
    eng.timing(True)
    for it in range(min(steps, 50)):
        advance(it)
    eng.sync()
    eng.timing(False)
    kt = eng.timing_read()

    out = {"value": n3d * steps / el, "ms_per_step": 1e3 * el / steps, "steps": steps}
    # steps plus the reference loop's per-step diagnostics (flow rate, plane
    # statistics, CFL: staggered/main.cpp:312-338), formed by the step's
    # projection (the SRK stepper: one pass of their own)
    if diagnostics:
        nd = max(1, steps // 4)
        if not srk:
            eng.step_diagnostics(True)
        for it in range(4):  # captures of the steps that form them
            v = advance(it)
            eng.diagnostics(v, v, dt, 3)
        eng.sync()
        t1 = time.perf_counter()
        for it in range(nd):
            v = advance(it)
            eng.diagnostics(v, v, dt, 3)
        eng.sync()
        el_diag = time.perf_counter() - t1
        out["with_diagnostics"] = {"value": n3d * nd / el_diag, "ms_per_step": 1e3 * el_diag / nd, "steps": nd}

    rsz = 8 if precision == tca.FP64 else 4
    per_kernel = {}
    for k, (ms, n) in kt.items():
        if n == 0:
            continue
        avg = ms / n
        d = {"avg_us": round(1e3 * avg, 2), "launches": n}
        if k in reals:
            gbs = reals[k] * rsz * n3d / (avg * 1e-3) / 1e9
            d["gbs"] = round(gbs, 1)
            d["frac_hbm"] = round(gbs / PEAK_HBM_GBS, 4)
        per_kernel[k] = d
    dom = max((k for k in per_kernel if k in reals), key=lambda k: kt[k][0])
    dom_avg_s = kt[dom][0] / kt[dom][1] * 1e-3
    achieved = reals[dom] * rsz * n3d / dom_avg_s / 1e9
    traffic = measured_traffic(name, precision_name, dom)
    out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                       "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                       "algorithmic_bytes": reals[dom] * rsz * n3d,
                       "traffic": traffic["bytes"] if traffic else None,
                       "traffic_source": traffic["source"] if traffic else None}
    out["kernels"] = per_kernel
    if kw.get("sgs_model") and kt["turb_visc"][1]:
        avg_s = kt["turb_visc"][0] / kt["turb_visc"][1] * 1e-3
        tf = MLP_FWD_FLOP * nx * ny * (nz - 2) / avg_s / 1e12
        out["closure"] = {"kernel": "k_sgs_tau", "avg_us": round(avg_s * 1e6, 2), "bound": "mfma",
                          "achieved": round(tf, 2), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(tf / PEAK_BF16_TFLOPS, 5)}
    out["config"] = {"workload": name,
                     "scheme": "SRK " + tca.IMEX_METHODS[kw["imex"]] if srk else ["HW", "VV", "AB3"][scheme],
                     "grid": [nx, ny, nz], "points": n3d, "dt": dt, "precision": precision_name,
                     "parallelism": "single GPU", "graphs": graphs}
    out["dtype"] = "f64" if precision == tca.FP64 else \
        "f32 (f64 low-mode solves, normalisation, flow rate, statistics)"
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS) + [TRAIN_WORKLOAD])
    ap.add_argument("--precision", default=None, choices=["fp64", "fp32"],
                    help="default: fp32 for the 384^2 and LES workloads, fp64 otherwise")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the extra workloads' step rates")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel of a step (no HIP graph replay)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB device copy (torch) whose counter readings scale the traffic figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch

        a = torch.ones(256 * 2**20 // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    if args.workload == TRAIN_WORKLOAD:
        bench_training(args)
        return
    copy = device_copy_rate()  # before the engine: torch sees the GPU only if it starts HIP first
    precision_name = args.precision or DEFAULT_PRECISION.get(args.workload, "fp64")
    r = run_workload(args.workload, precision_name, args.steps, args.warmup, graphs=not args.no_graphs)

    extra = None
    if not args.no_extra and args.workload == DEFAULT_WORKLOAD and args.precision is None:
        extra = {}
        for name, prec in EXTRA_WORKLOADS:
            x = run_workload(name, prec, args.steps, args.warmup, graphs=not args.no_graphs, diagnostics=False)
            extra[f"{name}_{prec}"] = {"value": x["value"], "unit": "grid-pt steps/s",
                                       "ms_per_step": x["ms_per_step"], "dtype": x["dtype"],
                                       "config": x["config"], "roofline": x["roofline"], "kernels": x["kernels"]}

    cpu = None
    if not args.no_cpu_baseline:
        scheme = WORKLOADS[args.workload][0]
        cpu = cpu_baseline(args.workload, args.cpu_seconds, shipped_threads=SHIPPED_THREADS[scheme])
        cpu["host_cores"] = host_cores()
    rec = {
        "metric": "grid-pt steps/s",
        "value": r["value"],
        "unit": "grid-pt steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": r["dtype"],
        "data": "synthetic (seeded divergence-free initial field)",
        "config": r["config"],
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
        "device_copy": copy,
        "closure": r.get("closure"),
        "speedup_vs_cpu": (r["value"] / cpu["value"]) if cpu else None,
        "with_diagnostics": r.get("with_diagnostics"),
        "kernels": r["kernels"],
        "extra_workloads": extra,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
