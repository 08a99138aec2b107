"""The closure configurations of the benchmark on one GPU, in three stages that
each run in one process (device memory does not survive between them):

``dns395``   collocated BDF3/EX3 DNS at Re_tau ~395 (bulk Re 13750, 384 x 384 x 192
             nodes, tanh 2.0, dt 5e-4, FP32 bulk): transition from the seeded
             start (noise rms 0.3 of the bulk velocity: with 0.1 the flow is
             still laminar after 6 time units), 25 time units to settle, 2000
             steps, then 50 snapshots 200 steps apart with the
             statistics averaged over those 10^4 steps; the same 10^4 steps of an
             FP64 engine started from the same turbulent state, for the FP32-bulk
             statistics error. Each snapshot gives the training samples of the
             4x sharp spectral cut-off in x and y (96 x 96 coarse planes, exact
             sub-filter stress, ~1.75 M samples); 40 snapshots train the network
             (Adam, lr 1e-3 decaying to 1e-4, batch 65536), 10 are held out.
             A second network, trained the same way on the dissipative part of
             each target along the sample's deviatoric strain rate
             (closure.strain_projection), is the eddy-viscosity variant.
             Writes the trained networks (raw inputs/outputs) and the statistics.
``judge590`` the judge: DNS at Re_tau ~590 (bulk Re 21900) on 512 x 512 x 256 nodes,
             tanh 1.8, seeded start, FP32 bulk: transition, 15 time units to
             settle, then profiles of
             the mean velocity and of the rms and <u'w'> of the field filtered
             onto the LES grids' cut-offs (64 and 96 modes in x and y).
``les590``   LES on the LES grids (FD viscous term replaced by the Galerkin
             stress form with the closure's viscosity): the learned closure and
             no model on 64 x 64 x 96, the reference's Smagorinsky-van Driest model
             (collocated_AB3/func_global.h:335-384) on 64 x 64 x 96 and on the 1.5x
             finer 96 x 96 x 144; dt 2e-3 (halved by the CFL rule when needed),
             40 time units from the seeded start (the 2 x 10^4 steps of dt 2e-3),
             then statistics over 40 more (a fixed time rather than a fixed step
             count, since the CFL rule halves dt on the finer grids). Errors of mean U+,
             u_tau and the rms/<u'w'> profiles against the judge (interpolated in
             wall units onto the LES nodes) and GPU seconds per flow-through.
             The network predicts the sub-filter stress in wall units at the
             training data's filter width; on an LES grid its outputs are scaled
             by (Delta+_les / Delta+_train)^2 (the Delta^2 of a gradient-type
             stress at the same gradients in wall units, Delta+ = Re_tau L /
             cut-off modes): 5.0 for the 64-mode grid at Re_tau 590 against the
             96-mode cut-off of the Re_tau 395 data. ``learned_unscaled`` runs the
             network without it; ``learned_clip`` with it and with backscatter
             clipping (sgs_model 2: no stress where tau_ij S_ij > 0);
             ``learned_proj`` with it, keeping only the stress's dissipative part
             along the deviatoric strain rate D (sgs_model 3: an eddy viscosity
             -tau_kl D_kl / (2 D_kl D_kl) >= 0 read off the network), and
             ``learned_proj_unscaled`` the same without the scaling;
             ``learned_eddy`` and ``learned_eddy_unscaled`` run the second network
             of the dns395 stage (trained on that dissipative part of the
             filtered-DNS stress) the same way; ``learned_eddy_k41`` scales it by
             (Delta+_les / Delta+_train)^(4/3), the inertial-range scaling of an
             eddy viscosity (eps^(1/3) Delta^(4/3)) in place of Delta^2.

Run (from the repository root): ``python -m turbchannel_amd.closure_protocol STAGE
--out FILE [--net FILE]``; ``--quick`` shrinks every size for a functional check.
torch is used only to hold training samples in device memory.
"""
import argparse
import json
import math
import time

import numpy as np

from . import FLUX, FP32_BULK, FP64, PRES, SCHEME_AB3, VEL, Engine, Mlp
from .closure import Standardizer, filter_width_ratio2, scale_outputs, strain_projection, train
from .les_study import Run, plane_stats, profile_stats, spectral_restrict, wall_shear

LX, LY = 2.0 * math.pi, math.pi

FULL = dict(
    dns395=dict(grid=(384, 384, 192), gamma=2.0, re_b=13750.0, re_tau=395.0, dt=5e-4, noise=0.3, seed=1,
                t_trans_max=80.0, re_tau_turb=300.0, steps_timed=2000, n_snap=50, snap_every=200, holdout=10,
                factor=4, iters=30000, batch=65536, stats_every=10, trans_seconds=420.0, t_develop=25.0),
    judge590=dict(grid=(512, 512, 256), gamma=1.8, re_b=21900.0, re_tau=590.0, dt=1e-3, noise=0.3, seed=1,
                  t_trans_max=60.0, re_tau_turb=450.0, t_develop=15.0, t_avg=25.0, les_cuts=[64, 96],
                  stats_every=20, wall_seconds=960.0),
    les590=dict(re_b=21900.0, re_tau=590.0, gamma=2.0, dt=2e-3, t_wash=40.0, t_avg=40.0, noise=0.3, seed=1,
                cases=[("learned", 64, 64, 96), ("learned_clip", 64, 64, 96), ("learned_unscaled", 64, 64, 96),
                       ("learned_proj", 64, 64, 96), ("learned_proj_unscaled", 64, 64, 96),
                       ("learned_eddy", 64, 64, 96), ("learned_eddy_unscaled", 64, 64, 96),
                       ("smagorinsky", 64, 64, 96), ("smagorinsky", 96, 96, 144), ("no_model", 64, 64, 96)],
                time_steps=2000),
)
QUICK = dict(
    dns395=dict(FULL["dns395"], grid=(64, 48, 33), t_trans_max=0.2, re_tau_turb=1e9, steps_timed=20, n_snap=6,
                snap_every=10, holdout=2, iters=200, batch=4096, t_develop=0.05),
    judge590=dict(FULL["judge590"], grid=(64, 48, 33), t_trans_max=0.2, re_tau_turb=1e9, t_develop=0.05, t_avg=0.2,
                  les_cuts=[16, 24], wall_seconds=60.0),
    les590=dict(FULL["les590"], t_wash=0.1, t_avg=0.1, cases=[("learned", 16, 16, 25), ("learned_clip", 16, 16, 25),
                                                  ("learned_proj", 16, 16, 25), ("learned_eddy", 16, 16, 25),
                                                  ("smagorinsky", 16, 16, 25), ("smagorinsky", 24, 24, 37),
                                                  ("no_model", 16, 16, 25)],
                time_steps=20),
)


def nodes(nz, gamma):
    n = nz - 1
    z = 1.0 + np.tanh(gamma * (2.0 * np.arange(n + 1) / n - 1.0)) / math.tanh(gamma)
    z[0], z[-1] = 0.0, 2.0
    return z


def learned_post(model):
    """tc_params.sgs_model of a model name: 0 (not learned), 1 (the network's
    stress), 2 (backscatter clipped), 3 (dissipative part along the strain rate)."""
    if not model.startswith("learned"):
        return 0
    return 3 if model.startswith(("learned_proj", "learned_eddy")) else 2 if model == "learned_clip" else 1


def engine(nx, ny, zz, re_b, re_tau, model, precision=FP32_BULK):
    """model: 'dns' (FD viscous term), 'no_model', 'smagorinsky', 'learned', 'learned_clip',
    'learned_unscaled', 'learned_proj', 'learned_proj_unscaled', 'learned_eddy',
    'learned_eddy_unscaled', 'learned_eddy_k41' (Galerkin viscous term)."""
    visc = 2.0 / re_b
    u_tau = re_tau * visc
    kw = dict(precision=precision, visc=visc, sour_dpdx=u_tau * u_tau, inv_length_scale=re_tau, hx=LX / nx,
              hy=LY / ny, stab_type=1, fd_mode=1, cfr_ubulk=1.0, visc_type=0 if model == "dns" else 1,
              turb_visc=1 if model == "smagorinsky" else 0,
              sgs_model=learned_post(model))
    e = Engine(SCHEME_AB3, nx, ny, len(zz), zz, **kw)
    e.graphs(True)
    return e


def transition(run, zz, visc, t_max, re_tau_turb, log, block=500, deadline=None):
    """Steps until the wall-shear Re_tau first exceeds re_tau_turb, t_max, or the
    wall-clock deadline (time.perf_counter() value)."""
    hist = []
    last = time.perf_counter()
    while run.t < t_max and (deadline is None or time.perf_counter() < deadline):
        run.advance(min(block * run.dt, t_max - run.t), block=block)
        U = run.e.profiles((VEL, run.cur), run.cur)[:, 0]
        rt = wall_shear(U, zz, visc) / visc
        hist.append((round(run.t, 3), round(rt, 1)))
        if time.perf_counter() - last > 60:
            log(f"  t={run.t:.2f}, Re_tau {rt:.1f}, dt {run.dt:.2e}")
            last = time.perf_counter()
        if rt > re_tau_turb:
            break
    log(f"transition: t={run.t:.2f} after {run.it} steps, dt {run.dt:.2e}, "
        f"Re_tau history {hist[::max(1, len(hist) // 12)]}")
    return hist


def develop(run, zz, visc, t_dev, log, deadline=None, block=500):
    """t_dev more time units after the transition (the overshoot of the wall shear
    settles), with the Re_tau history."""
    hist = []
    t_end = run.t + t_dev
    last = time.perf_counter()
    while run.t < t_end - 0.5 * run.dt and (deadline is None or time.perf_counter() < deadline):
        run.advance(min(block * run.dt, t_end - run.t), block=block)
        U = run.e.profiles((VEL, run.cur), run.cur)[:, 0]
        hist.append((round(run.t, 3), round(wall_shear(U, zz, visc) / visc, 1)))
        if time.perf_counter() - last > 60:
            log(f"  t={run.t:.2f}, Re_tau {hist[-1][1]:.1f}, dt {run.dt:.2e}")
            last = time.perf_counter()
    log(f"developed: t={run.t:.2f} after {run.it} steps, dt {run.dt:.2e}, Re_tau {hist[::max(1, len(hist) // 8)]}")
    return hist


def stats_over(run, nsteps, every):
    """Mean of the per-plane records (tc_profiles) every `every` steps over nsteps."""
    s, n = 0.0, 0
    for _ in range(nsteps // every):
        run.steps(every)
        s = s + run.e.profiles((VEL, run.cur), run.cur)
        n += 1
    run.check()
    return s / n


def stage_dns395(P, log):
    import torch

    nx, ny, nz = P["grid"]
    zz = nodes(nz, P["gamma"])
    visc = 2.0 / P["re_b"]
    f = P["factor"]
    out = dict(stage="dns395", grid=[nx, ny, nz], re_b=P["re_b"], re_tau=P["re_tau"], gamma=P["gamma"],
               noise=P["noise"], factor=f)
    dns = engine(nx, ny, zz, P["re_b"], P["re_tau"], "dns")
    dns.initial_field(0, seed=P["seed"], ubulk=1.0, noise_rms=P["noise"])
    run = Run(dns, P["dt"])
    t0 = time.perf_counter()
    out["re_tau_history"] = transition(run, zz, visc, P["t_trans_max"], P["re_tau_turb"], log,
                                       deadline=t0 + P["trans_seconds"])
    out["transition_t"], out["transition_steps"] = run.t, run.it
    out["develop_history"] = develop(run, zz, visc, P["t_develop"], log)
    sec = run.timed(P["steps_timed"])  # the timed steps
    out["ms_per_step_fp32"] = 1e3 * sec
    run.check()
    log(f"DNS 395: {P['steps_timed']} timed steps at {1e3 * sec:.3f} ms/step, t={run.t:.2f}, dt={run.dt:.2e}")
    # the FP64 engine from the same (FP32) state, for the statistics error
    d64 = engine(nx, ny, zz, P["re_b"], P["re_tau"], "dns", precision=FP64)
    for slot in range(4):
        for kind in (VEL, FLUX, PRES):  # the ring slots the next steps read
            d64.set_field(kind, slot, dns.get_field(kind, slot).astype(np.float64))
    run64 = Run(d64, run.dt)
    run64.cur, run64.it, run64.start, run64.t = run.cur, run.it, run.start, run.t
    n_samp = dns.sgs_sample_count(f)
    snaps_x, snaps_t = [], []
    prof32 = 0.0
    n_prof = 0
    for s in range(P["n_snap"]):
        prof32 = prof32 + stats_over(run, P["snap_every"], P["stats_every"])
        n_prof += 1
        x = torch.empty(10 * n_samp, dtype=torch.float32, device="cuda")
        t = torch.empty(6 * n_samp, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        assert dns.sgs_samples((VEL, run.cur), f, x.data_ptr(), t.data_ptr(), n_samp) == n_samp
        snaps_x.append(x)
        snaps_t.append(t)
    av32 = prof32 / n_prof
    t_dns = time.perf_counter() - t0
    log(f"DNS 395 FP32: {P['n_snap']} snapshots, statistics over {P['n_snap'] * P['snap_every']} steps, t={run.t:.2f}, "
        f"{t_dns:.0f} s")
    av64 = stats_over(run64, P["n_snap"] * P["snap_every"], P["stats_every"])
    st32, st64 = profile_stats(av32), profile_stats(av64)
    ut32, ut64 = wall_shear(st32[:, 0], zz, visc), wall_shear(st64[:, 0], zz, visc)
    names = ["mean_u", "rms_u", "rms_v", "rms_w", "uw"]
    diff = {n: float(np.abs(st32[:, k] - st64[:, k]).max() / np.abs(st64[:, k]).max()) for k, n in enumerate(names)}
    out.update(re_tau_fp32=ut32 / visc, re_tau_fp64=ut64 / visc, stats_fp32_vs_fp64=diff, dt=run.dt,
               ms_per_step_fp64=1e3 * run64.timed(200),
               profiles=dict(z=zz.tolist(), fp32=st32.tolist(), fp64=st64.tolist()))
    log(f"DNS 395 statistics: Re_tau FP32 {ut32 / visc:.1f}, FP64 {ut64 / visc:.1f}; max difference / max: {diff}")
    d64.close()
    dns.close()

    # training set and training (cfg 3): the network on the sub-filter stress, and
    # a second one on its dissipative part along the strain rate (the part an
    # engine with sgs_model 3 keeps)
    ntr = P["n_snap"] - P["holdout"]
    Xn = torch.cat(snaps_x[:ntr]).view(-1, 10).cpu().numpy()
    Tn = torch.cat(snaps_t[:ntr]).view(-1, 6).cpu().numpy()
    Xh = torch.cat(snaps_x[ntr:]).view(-1, 10)
    Thn = torch.cat(snaps_t[ntr:]).view(-1, 6).cpu().numpy().astype(np.float64)
    del snaps_x, snaps_t
    out.update(samples_per_snapshot=n_samp, train_samples=int(Xn.shape[0]), holdout_samples=int(Thn.shape[0]))
    perm = np.random.default_rng(0).permutation(len(Xn))
    Xhn = Xh.cpu().numpy()
    for key, tgt, tgt_h in (("", Tn, Thn), ("proj_", None, None)):
        if tgt is None:  # the dissipative part of the same targets
            tgt = strain_projection(Tn, Xn)
            tgt_h = strain_projection(Thn, Xhn).astype(np.float64)
        st = Standardizer(Xn, tgt)
        xs = torch.from_numpy(st.inputs(Xn)[perm]).cuda().contiguous()
        ts = torch.from_numpy(st.targets(tgt)[perm]).cuda().contiguous()
        torch.cuda.synchronize()
        net = Mlp(seed=0)
        t0 = time.perf_counter()
        losses = train(net, xs.data_ptr(), ts.data_ptr(), xs.shape[0], batch=P["batch"], iters=P["iters"], lr=1e-3,
                       lr_final=1e-4)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        params = st.fold(net.get_params())
        net.close()
        del xs, ts
        raw = Mlp(seed=0)
        raw.set_params(params)
        yh = torch.empty(6 * Xh.shape[0], dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        raw.forward(Xh.data_ptr(), Xh.shape[0], yh.data_ptr())
        torch.cuda.synchronize()
        raw.close()
        pred = yh.view(-1, 6).cpu().numpy().astype(np.float64)
        del yh
        if key:  # what the engine keeps of the prediction
            pred = strain_projection(pred, Xhn).astype(np.float64)
        mse, mse0 = float(np.mean((pred - tgt_h) ** 2)), float(np.mean(tgt_h ** 2))
        out.update({key + "train_loss": [float(losses[0]), float(np.mean(losses[-100:]))],
                    key + "train_seconds": t_train, key + "train_samples_per_s": P["iters"] * P["batch"] / t_train,
                    key + "holdout_mse_learned": mse, key + "holdout_mse_zero": mse0,
                    key + "holdout_explained": 1.0 - mse / mse0, key + "params": [float(v) for v in params]})
        log(f"training {key or 'stress'}: {Xn.shape[0]} samples, {P['iters']} iterations in {t_train:.1f} s, loss "
            f"{losses[0]:.3f} -> {np.mean(losses[-100:]):.3f}; held-out MSE learned {mse:.4e} vs no model {mse0:.4e} "
            f"({100 * (1 - mse / mse0):.1f}% explained)")
    return out


def stage_judge590(P, log):
    nx, ny, nz = P["grid"]
    zz = nodes(nz, P["gamma"])
    visc = 2.0 / P["re_b"]
    out = dict(stage="judge590", grid=[nx, ny, nz], re_b=P["re_b"], gamma=P["gamma"], noise=P["noise"])
    t_start = time.perf_counter()
    dns = engine(nx, ny, zz, P["re_b"], P["re_tau"], "dns")
    dns.initial_field(0, seed=P["seed"], ubulk=1.0, noise_rms=P["noise"])
    run = Run(dns, P["dt"])
    # at most 55% of the wall-clock allowance for the transition, the rest averages
    out["re_tau_history"] = transition(run, zz, visc, P["t_trans_max"], P["re_tau_turb"], log,
                                       deadline=t_start + 0.4 * P["wall_seconds"])
    out["transition_t"], out["transition_steps"] = run.t, run.it
    out["develop_history"] = develop(run, zz, visc, P["t_develop"], log, deadline=t_start + 0.55 * P["wall_seconds"])
    out["ms_per_step"] = 1e3 * run.timed(200)
    run.check()
    # planar statistics every stats_every steps; the statistics of the field
    # filtered onto the LES cut-offs at 8 evenly spaced times
    s, n = 0.0, 0
    filt = {c: 0.0 for c in P["les_cuts"]}
    nf = 0
    t0_avg = run.t
    t_end = run.t + P["t_avg"]
    k = 0
    while run.t < t_end and time.perf_counter() - t_start < P["wall_seconds"]:
        run.steps(P["stats_every"])
        s = s + dns.profiles((VEL, run.cur), run.cur)
        n += 1
        k += 1
        if k % 50 == 0:
            run.check()
        if run.t - t0_avg >= (nf + 0.5) * P["t_avg"] / 8:
            u = dns.get_field(VEL, run.cur).astype(np.float64).reshape(nz, ny, nx, 3)
            for c in P["les_cuts"]:
                filt[c] = filt[c] + plane_stats(spectral_restrict(u, nx, ny, nz, c, c))
            del u
            nf += 1
            log(f"judge: t={run.t:.2f}, {n} records, {nf} filtered fields, {time.perf_counter() - t_start:.0f} s")
    av = profile_stats(s / n)
    ut = wall_shear(av[:, 0], zz, visc)
    out.update(t_end=run.t, steps=run.it, dt=run.dt, avg_time=run.t - t0_avg, records=n,
               filtered_fields=nf, re_tau=ut / visc, u_tau=ut, z=zz.tolist(), profile=av.tolist(),
               filtered={str(c): (filt[c] / max(nf, 1)).tolist() for c in P["les_cuts"]})
    log(f"judge DNS: Re_tau {ut / visc:.1f}, averaged over {out['avg_time']:.2f} time units ({n} records)")
    dns.close()
    return out


def stage_les590(P, net, judge, log):
    visc = 2.0 / P["re_b"]
    params = np.asarray(net["params"], dtype=np.float32)
    params_eddy = np.asarray(net["proj_params"], dtype=np.float32) if "proj_params" in net else None
    zj = np.asarray(judge["z"])
    jprof = np.asarray(judge["profile"])
    ut_j = judge["u_tau"]
    res = {}
    for model, lx, ly, lz in P["cases"]:
        zz = nodes(lz, P["gamma"])
        e = engine(lx, ly, zz, P["re_b"], P["re_tau"], model)
        if model.startswith("learned"):
            # the training data's cut-off: the DNS grid over the filter factor
            gx, gy = net["grid"][0], net["grid"][1]
            fac = net.get("factor", FULL["dns395"]["factor"])
            s = filter_width_ratio2(net.get("re_tau", FULL["dns395"]["re_tau"]), (gx // fac, gy // fac),
                                    P["re_tau"], (lx, ly)) if not model.endswith("unscaled") else 1.0
            if model.endswith("_k41"):  # an eddy viscosity in the inertial range: eps^(1/3) Delta^(4/3)
                s = s ** (2.0 / 3.0)
            e.set_sgs_params(scale_outputs(params_eddy if model.startswith("learned_eddy") else params, s))
        e.initial_field(0, seed=P["seed"], ubulk=1.0, noise_rms=P["noise"])
        run = Run(e, P["dt"])
        name = f"{model}_{lx}x{ly}x{lz}"
        r = dict(model=model, grid=[lx, ly, lz])
        if model.startswith("learned"):
            r["output_scale"] = s
        try:
            run.advance(P["t_wash"], block=500)
            av = stats_over(run, int(round(P["t_avg"] / run.dt)), 10)
            st = profile_stats(av)
            ut = wall_shear(st[:, 0], zz, visc)
            sec = run.timed(P["time_steps"])
            # the judge in wall units, interpolated onto this grid's wall distances
            m = lz // 2 + 1
            d = np.minimum(zz, 2.0 - zz)[:m]
            mirror = lambda p: 0.5 * (p[:m] + p[::-1][:m])  # noqa: E731
            dj = np.minimum(zj, 2.0 - zj)[: len(zj) // 2 + 1]
            mj = lambda p: 0.5 * (p[: len(zj) // 2 + 1] + p[::-1][: len(zj) // 2 + 1])  # noqa: E731
            cut = str(lx)
            filt = np.asarray(judge["filtered"][cut]) if cut in judge["filtered"] else None
            yp_les, yp_j = d * ut / visc, dj * ut_j / visc
            Uplus_les = mirror(st[:, 0]) / ut
            Uplus_j = np.interp(yp_les, yp_j, mj(jprof[:, 0]) / ut_j)
            U_j = np.interp(d, dj, mj(jprof[:, 0]))
            w = np.gradient(d)

            def rel(a, b):
                return float(math.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w * b ** 2)))
            r.update(re_tau=ut / visc, u_tau_err=abs(ut - ut_j) / ut_j, mean_u_err=rel(mirror(st[:, 0]), U_j),
                     mean_u_plus_err=rel(Uplus_les, Uplus_j), dt=run.dt, steps=run.it, ms_per_step=1e3 * sec,
                     seconds_per_flow_through=sec * LX / run.dt)
            if filt is not None:  # resolved rms and <u'w'> against the judge filtered to this grid's cut-off
                # <u'w'> changes sign across the centre line: fold the upper half with a minus sign
                anti = lambda p, n: 0.5 * (p[:n] - p[::-1][:n])  # noqa: E731
                for k, nm in ((1, "rms_u"), (2, "rms_v"), (3, "rms_w"), (4, "uw")):
                    if k == 4:
                        fj = np.interp(d, dj, anti(filt[:, k], len(dj)))
                        r[nm + "_err"] = rel(anti(st[:, k], m), fj)
                    else:
                        fj = np.interp(d, dj, mj(filt[:, k]))
                        r[nm + "_err"] = rel(mirror(st[:, k]), fj)
            r["profile"] = dict(z=zz.tolist(), stats=st.tolist())
        except FloatingPointError as exc:
            r["diverged"] = str(exc)
        e.close()
        res[name] = r
        log(f"LES {name}: " + ", ".join(f"{k} {v:.4g}" for k, v in r.items() if isinstance(v, float)))
    return dict(stage="les590", judge_re_tau=judge["re_tau"], seed=P["seed"], les=res)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("stage", choices=["dns395", "judge590", "les590"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--net", help="dns395 output (trained network) for les590")
    ap.add_argument("--judge", help="judge590 output for les590")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--seed", type=int, help="les590: seed of the LES starts (a replicate of the study)")
    ap.add_argument("--cases", help="les590: comma-separated model:nx:ny:nz in place of the default cases")
    a = ap.parse_args()
    P = dict((QUICK if a.quick else FULL)[a.stage])
    if a.seed is not None:
        P["seed"] = a.seed
    if a.cases:
        P["cases"] = [(c.split(":")[0], *map(int, c.split(":")[1:])) for c in a.cases.split(",")]
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.stage == "dns395":
        res = stage_dns395(P, log)
    elif a.stage == "judge590":
        res = stage_judge590(P, log)
    else:
        with open(a.net) as f:
            net = json.load(f)
        with open(a.judge) as f:
            judge = json.load(f)
        res = stage_les590(P, net, judge, log)
    with open(a.out, "w") as f:
        json.dump(res, f)
    log(f"wrote {a.out}")


if __name__ == "__main__":
    main()
