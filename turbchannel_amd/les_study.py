"""A-posteriori study of the learned subgrid closure on one GPU.

1. DNS of turbulent channel flow with the collocated BDF3/EX3 solver (FP32 bulk,
   constant flow rate, bulk units): seeded start, transition, then averaged
   profiles and training samples of the spectrally filtered field (sharp cut-off
   by `factor` in x and y, exact sub-filter stress) at evenly spaced times.
2. The closure network is trained on the GPU (Adam, standardised data) on the
   earlier snapshots; the later ones are held out (a-priori error).
3. LES on the filtered grid, started from the filtered final DNS field: the
   learned closure, the reference's Smagorinsky-van Driest model on the same grid
   and on a grid 1.5x finer in x and y, and no model. Each run washes out its
   start, then averages profiles.
4. Errors of the mean velocity and of the resolved rms and shear-stress profiles
   against the filtered DNS, and the GPU wall-clock per flow-through (Lx / U_b),
   from the time of a block of steps between two synchronisations.

Run: ``python -m turbchannel_amd.les_study --preset re180 --out study.json``.
torch is used only to hold the training samples in device memory.
"""
import argparse
import json
import math
import time

import numpy as np

from . import FP32_BULK, PRES, SCHEME_AB3, VEL, Engine, Mlp
from .closure import Standardizer, train

PRESETS = {
    # Re_tau 180 (bulk Re 5600), box 2 pi x pi x 2; DNS 256 x 192 x 129 nodes, LES on the 4x coarser
    # 64 x 48 grid in x and y (same wall-normal nodes), the finer plain LES at 96 x 72
    "re180": dict(dns=(256, 192), nz=129, gamma=1.8, factor=4, fine=(96, 72), dt=2e-3, t_trans=150.0,
                  t_dns_avg=100.0, n_snap=25, holdout=5, t_les_wash=40.0, t_les_avg=100.0, iters=12000,
                  batch=65536, time_steps=2000, seed=1, noise=0.3),
    # the same pipeline at toy sizes, for a functional check
    "quick": dict(dns=(64, 48), nz=33, gamma=1.8, factor=2, fine=(48, 36), dt=2e-3, t_trans=0.4,
                  t_dns_avg=0.4, n_snap=4, holdout=1, t_les_wash=0.2, t_les_avg=0.4, iters=200, batch=4096,
                  time_steps=50, seed=1, noise=0.3),
}
RE_B, RE_TAU = 5600.0, 180.0
LX, LY = 2.0 * math.pi, math.pi


def nodes(nz, gamma):
    """Collocated wall-normal nodes z_k = 1 + tanh(gamma (2k/n - 1)) / tanh(gamma), n = nz - 1."""
    n = nz - 1
    z = 1.0 + np.tanh(gamma * (2.0 * np.arange(n + 1) / n - 1.0)) / math.tanh(gamma)
    z[0], z[-1] = 0.0, 2.0
    return z


def make_engine(nx, ny, zz, model):
    """model: 'dns' (no model, FD viscous term), 'none', 'smag', 'learned' (Galerkin viscous term)."""
    visc = 2.0 / RE_B
    u_tau = RE_TAU * visc
    kw = dict(precision=FP32_BULK, visc=visc, sour_dpdx=u_tau * u_tau, inv_length_scale=RE_TAU, hx=LX / nx,
              hy=LY / ny, stab_type=1, fd_mode=1, cfr_ubulk=1.0, visc_type=0 if model == "dns" else 1,
              turb_visc=1 if model == "smag" else 0, sgs_model=1 if model == "learned" else 0)
    e = Engine(SCHEME_AB3, nx, ny, len(zz), zz, **kw)
    e.graphs(True)  # repeated steps replay a captured HIP graph
    return e


class Run:
    """The reference's stepping rules (BDF order ramp 1-2-3 after every change of
    the step, halving while CFL > CFLmax), with CFL checked between blocks of
    steps instead of after every step so that the GPU queue never drains."""

    def __init__(self, eng, dt, cfl_max=0.5):
        self.e, self.dt, self.cfl_max = eng, dt, cfl_max
        self.cur = self.it = self.start = 0
        self.t = 0.0

    def steps(self, n):
        for _ in range(n):
            self.e.step(self.t, self.dt, self.cur, min(self.it - self.start + 1, 3))
            self.cur = (self.cur + 1) % 4
            self.it += 1
            self.t += self.dt

    def check(self):
        c = self.e.cfl((VEL, self.cur), self.dt, 3)
        if not math.isfinite(c):
            raise FloatingPointError(f"non-finite CFL at t={self.t:.3f}")
        if c > self.cfl_max:
            while c > self.cfl_max:
                c *= 0.5
                self.dt *= 0.5
            self.start = self.it
        return c

    def advance(self, T, block=200, on_block=None):
        t_end = self.t + T
        while self.t < t_end - 0.5 * self.dt:
            self.steps(max(1, min(block, int(round((t_end - self.t) / self.dt)))))
            self.check()
            if on_block:
                on_block(self)

    def timed(self, n):
        """Seconds per step over n steps between two synchronisations."""
        self.e.sync()
        t0 = time.perf_counter()
        self.steps(n)
        self.e.sync()
        return (time.perf_counter() - t0) / n


def spectral_restrict(u, nx, ny, nz, cx, cy):
    """Sharp spectral cut-off of a (nz, ny, nx, ...) field onto the cx x cy grid
    (|kx| < cx/2, |ky| < cy/2 kept; the Nyquist modes dropped)."""
    U = np.fft.fft2(u, axes=(1, 2))
    kx, ky = np.fft.fftfreq(nx, 1.0 / nx), np.fft.fftfreq(ny, 1.0 / ny)
    keep_x = np.abs(kx) < cx / 2
    keep_y = np.abs(ky) < cy / 2
    C = np.zeros((nz, cy, cx) + u.shape[3:], dtype=complex)
    ix = np.where(keep_x)[0]
    iy = np.where(keep_y)[0]
    cxi = np.where(kx[ix] >= 0, kx[ix], kx[ix] + cx).astype(int)
    cyi = np.where(ky[iy] >= 0, ky[iy], ky[iy] + cy).astype(int)
    C[:, cyi[:, None], cxi[None, :]] = U[:, iy[:, None], ix[None, :]]
    return np.real(np.fft.ifft2(C, axes=(1, 2))) * (cx * cy) / (nx * ny)


def plane_stats(u):
    """Per-plane mean U and rms u', v', w' and <u'w'> of a (nz, ny, nx, 3) field."""
    m = u.mean(axis=(1, 2))
    f = u - m[:, None, None, :]
    rms = np.sqrt((f ** 2).mean(axis=(1, 2)))
    uw = (f[..., 0] * f[..., 2]).mean(axis=(1, 2))
    return np.column_stack([m[:, 0], rms, uw])


def profile_stats(av):
    """The same columns from tc_profiles rows (u[3], uu[6] XX YY ZZ XY XZ YZ, nu_t)."""
    u = av[:, :3]
    var = np.maximum(av[:, 3:6] - u ** 2, 0.0)
    uw = av[:, 7] - u[:, 0] * u[:, 2]
    return np.column_stack([u[:, 0], np.sqrt(var), uw])


def wall_shear(U, zz, visc):
    """u_tau from the mean velocity, second-order one-sided at both walls (averaged)."""
    def one(Ua, za):
        d0, d1 = za[1] - za[0], za[2] - za[1]
        dudz = 1.5 * (Ua[1] - Ua[0]) / d0 - 0.5 * (Ua[2] - Ua[1]) / d1
        return dudz
    g = 0.5 * (one(U, zz) + one(U[::-1], 2.0 - zz[::-1]))
    return math.sqrt(max(g, 0.0) * visc)


def errors(les, ref, zz, visc):
    """Relative L2 errors over the wall-normal nodes (trapezoidal weights)."""
    w = np.zeros_like(zz)
    w[1:] += 0.5 * np.diff(zz)
    w[:-1] += 0.5 * np.diff(zz)

    def rel(a, b):
        return float(math.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w * b ** 2)))
    ut_l, ut_r = wall_shear(les[:, 0], zz, visc), wall_shear(ref[:, 0], zz, visc)
    return dict(mean_u=rel(les[:, 0], ref[:, 0]),
                mean_u_plus=rel(les[:, 0] / ut_l, ref[:, 0] / ut_r) if ut_l > 0 else float("inf"),
                u_tau=abs(ut_l - ut_r) / ut_r, rms_u=rel(les[:, 1], ref[:, 1]), rms_v=rel(les[:, 2], ref[:, 2]),
                rms_w=rel(les[:, 3], ref[:, 3]), uw=rel(les[:, 4], ref[:, 4]))


def study(preset="re180", log=print):
    import torch  # device memory for the samples

    P = dict(PRESETS[preset])
    nx, ny = P["dns"]
    zz = nodes(P["nz"], P["gamma"])
    nz = len(zz)
    visc = 2.0 / RE_B
    f = P["factor"]
    cx, cy = nx // f, ny // f
    out = dict(preset=preset, re_b=RE_B, re_tau_nominal=RE_TAU, box=[LX, LY, 2.0], dns_grid=[nx, ny, nz],
               les_grid=[cx, cy, nz], fine_grid=[*P["fine"], nz], dt=P["dt"], factor=f)

    # ---- 1. DNS -------------------------------------------------------------------
    dns = make_engine(nx, ny, zz, "dns")
    dns.initial_field(0, seed=P["seed"], ubulk=1.0, noise_rms=P["noise"])
    run = Run(dns, P["dt"])
    hist = []

    def note(r):
        U = r.e.profiles((VEL, r.cur), r.cur)[:, 0]
        hist.append((round(r.t, 3), round(wall_shear(U, zz, visc) / visc, 2)))
    t0 = time.perf_counter()
    run.advance(P["t_trans"], block=500, on_block=note)
    log(f"DNS transition: t={run.t:.1f}, Re_tau history (every 20th block): {hist[::20]}")
    out["dns_re_tau_history"] = hist[::5]
    n_samp = dns.sgs_sample_count(f)
    snaps_x, snaps_t = [], []
    av_sum, av_n = np.zeros((nz, 10)), 0
    filt_sum = np.zeros((nz, 5))
    interval = P["t_dns_avg"] / P["n_snap"]
    for s in range(P["n_snap"]):
        def acc(r):
            nonlocal av_sum, av_n
            av_sum += r.e.profiles((VEL, r.cur), r.cur)
            av_n += 1
        run.advance(interval, block=50, on_block=acc)
        x = torch.empty(10 * n_samp, dtype=torch.float32, device="cuda")
        t = torch.empty(6 * n_samp, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        got = dns.sgs_samples((VEL, run.cur), f, x.data_ptr(), t.data_ptr(), n_samp)
        assert got == n_samp
        snaps_x.append(x)
        snaps_t.append(t)
        u = dns.get_field(VEL, run.cur).astype(np.float64).reshape(nz, ny, nx, 3)
        filt_sum += plane_stats(spectral_restrict(u, nx, ny, nz, cx, cy))
    dns_time = time.perf_counter() - t0
    ref_full = profile_stats(av_sum / av_n)
    ref_filt = filt_sum / P["n_snap"]
    ref_filt[:, 0] = ref_full[:, 0]  # the planar mean is unchanged by the filter; use the densely sampled one
    ut = wall_shear(ref_full[:, 0], zz, visc)
    log(f"DNS: Re_tau {ut / visc:.1f}, {run.it} steps, dt {run.dt:.2e}, {dns_time:.1f} s")
    out.update(dns_re_tau=ut / visc, dns_steps=run.it, dns_dt=run.dt, dns_ms_per_step=1e3 * run.timed(P["time_steps"]))
    out["dns_profile"] = dict(z=zz.tolist(), U=ref_full[:, 0].tolist(), rms_filtered=ref_filt[:, 1:4].tolist(),
                              uw_filtered=ref_filt[:, 4].tolist())
    u_dns = dns.get_field(VEL, run.cur).astype(np.float64).reshape(nz, ny, nx, 3)
    dns.close()

    # ---- 2. training -------------------------------------------------------------
    ntr = P["n_snap"] - P["holdout"]
    X = torch.cat(snaps_x[:ntr]).view(-1, 10)
    T = torch.cat(snaps_t[:ntr]).view(-1, 6)
    Xh = torch.cat(snaps_x[ntr:]).view(-1, 10)
    Th = torch.cat(snaps_t[ntr:]).view(-1, 6)
    Xn, Tn = X.cpu().numpy(), T.cpu().numpy()
    st = Standardizer(Xn, Tn)
    perm = torch.from_numpy(np.random.default_rng(0).permutation(len(Xn))).cuda()
    xs = torch.from_numpy(st.inputs(Xn)).cuda()[perm].contiguous()
    ts = torch.from_numpy(st.targets(Tn)).cuda()[perm].contiguous()
    torch.cuda.synchronize()
    net = Mlp(seed=0)
    t0 = time.perf_counter()
    losses = train(net, xs.data_ptr(), ts.data_ptr(), len(Xn), batch=P["batch"], iters=P["iters"], lr=1e-3,
                   lr_final=1e-4)
    train_time = time.perf_counter() - t0
    params = st.fold(net.get_params())
    raw = Mlp(seed=0)
    raw.set_params(params)
    yh = torch.empty(6 * len(Xh), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    raw.forward(Xh.data_ptr(), len(Xh), yh.data_ptr())
    torch.cuda.synchronize()
    pred, Thn = yh.view(-1, 6).cpu().numpy().astype(np.float64), Th.cpu().numpy().astype(np.float64)
    mse_net, mse_zero = float(np.mean((pred - Thn) ** 2)), float(np.mean(Thn ** 2))
    log(f"training: {len(Xn)} samples, loss {losses[0]:.3f} -> {losses[-1]:.3f}, {train_time:.1f} s; held-out "
        f"MSE learned {mse_net:.4e} vs no model {mse_zero:.4e}")
    out.update(train_samples=len(Xn), holdout_samples=len(Thn), train_loss=[float(losses[0]), float(losses[-1])],
               train_seconds=train_time, holdout_mse_learned=mse_net, holdout_mse_zero=mse_zero)
    del X, T, xs, ts, snaps_x, snaps_t

    # ---- 3. LES ------------------------------------------------------------------
    results = {}
    cases = [("learned", cx, cy), ("smagorinsky", cx, cy), ("smagorinsky_fine", *P["fine"]), ("no_model", cx, cy)]
    for name, lx, ly in cases:
        model = {"learned": "learned", "no_model": "none"}.get(name, "smag")
        e = make_engine(lx, ly, zz, model)
        if model == "learned":
            e.set_sgs_params(params)
        u0 = spectral_restrict(u_dns, nx, ny, nz, lx, ly).reshape(-1, 3)
        e.set_field(VEL, 0, u0)
        e.project((VEL, 0))
        r = Run(e, P["dt"])
        res = dict(grid=[lx, ly, nz])
        try:
            r.advance(P["t_les_wash"], block=500)
            s, n = np.zeros((nz, 10)), 0

            def acc(rr):
                nonlocal s, n
                s += rr.e.profiles((VEL, rr.cur), rr.cur)
                n += 1
            r.advance(P["t_les_avg"], block=50, on_block=acc)
            st_les = profile_stats(s / n)
            sec = r.timed(P["time_steps"])
            res.update(errors(st_les, ref_filt, zz, visc), re_tau=wall_shear(st_les[:, 0], zz, visc) / visc,
                       dt=r.dt, ms_per_step=1e3 * sec, seconds_per_flow_through=sec * (LX / 1.0) / r.dt,
                       profile=dict(U=st_les[:, 0].tolist(), rms=st_les[:, 1:4].tolist(), uw=st_les[:, 4].tolist()))
        except FloatingPointError as exc:
            res.update(diverged=str(exc))
        e.close()
        results[name] = res
        log(f"LES {name} {lx}x{ly}x{nz}: " + ", ".join(f"{k} {v:.4g}" for k, v in res.items()
                                                       if isinstance(v, float)))
    out["les"] = results
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--preset", default="re180", choices=sorted(PRESETS))
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    res = study(a.preset, log=lambda s: print(s, flush=True))
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    summary = {k: {kk: vv for kk, vv in v.items() if kk != "profile"} for k, v in res["les"].items()}
    print(json.dumps(dict(dns_re_tau=res["dns_re_tau"], les=summary)))


if __name__ == "__main__":
    main()
