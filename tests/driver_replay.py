"""Helpers for the command-line driver tests: run turbchannel_amd/turbch, read
the files it writes, and replay its time loop on the oracle.

The replay restates the reference's host loop (staggered/main.cpp:306-371,
collocated_AB3/main.cpp:392-476): BDF order ramp after every change of the step,
bulk velocity from CalcIntegral / area, profiles, CFL-driven halving and 1.5x
growth of the step, and the friction velocity from the first plane's mean."""
import os
import subprocess

import numpy as np

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "turbchannel_amd", "turbch")


def run_driver(args, cwd, timeout=120, env=None):
    """Runs the command-line solver; env adds variables to the inherited environment."""
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    full_env = dict(os.environ, **env) if env else None
    r = subprocess.run([DRIVER, *[str(a) for a in args]], cwd=cwd, capture_output=True, text=True, timeout=timeout,
                       env=full_env)
    return r


def write_vtk(path, nx, ny, nz, vel, pres, x=None, y=None, z=None):
    """ASCII VTK in the layout the driver reads (DumpData, staggered/io.cpp:6-29);
    returns the values as the reader parses them back."""
    x = np.zeros(nx) if x is None else x
    y = np.zeros(ny) if y is None else y
    z = np.zeros(nz) if z is None else z
    lines = ["# vtk DataFile Version 2.0", "Volume example", "ASCII", "DATASET RECTILINEAR_GRID",
             f"DIMENSIONS {nx} {ny} {nz}"]
    for c, arr in zip("XYZ", (x, y, z)):
        lines.append(f"{c}_COORDINATES {len(arr)} float")
        lines.append("".join("%e " % v for v in arr))
    n = nx * ny * nz
    lines += [f"POINT_DATA {n}", "SCALARS volume_scalars float 4", "LOOKUP_TABLE default"]
    vel = np.asarray(vel).reshape(n, 3)
    pres = np.asarray(pres).reshape(n)
    rows = ["%e %e %e %e" % (vel[i, 0], vel[i, 1], vel[i, 2], pres[i]) for i in range(n)]
    with open(path, "w") as f:
        f.write("\n".join(lines + rows) + "\n")
    back = np.array([[float(t) for t in r.split()] for r in rows])
    return back[:, :3].copy(), back[:, 3].copy()


def read_vtk(path):
    """(dims, coordinates, values[n, 4]) of a file written by the driver."""
    with open(path) as f:
        lines = f.read().splitlines()
    dims = tuple(int(v) for v in lines[4].split()[1:4])
    coords = [np.array([float(v) for v in lines[6 + 2 * d].split()]) for d in range(3)]
    start = lines.index("LOOKUP_TABLE default") + 1
    vals = np.array([[float(v) for v in r.split()] for r in lines[start:start + int(np.prod(dims))]])
    return dims, coords, vals


def read_table(path):
    return np.loadtxt(path, ndmin=2)


def replay(sol, u0, p0, *, tau, steps_max, area, dz, visc, collocated, time_max=100.0, t_avg=0.0,
           cfl_min=0.05, cfl_max=0.5, sprog=1):
    """The driver's loop on the oracle. Returns the log rows, the final velocity
    and pressure, the last profiles and the time-averaged profiles."""
    zero3 = np.zeros_like(u0)
    for s in range(4):
        sol.set_field(oracle.KIND_VEL, s, u0 if s == 0 else zero3)
        sol.set_field(oracle.KIND_PRES, s, p0 if s == 0 else np.zeros_like(p0))
    t, it, it_start, cur = 0.0, 0, 0, 0
    rows = []
    av_sum, t_int = None, 0.0
    while True:
        order = min(it - it_start + 1, 3)
        sol.step(it * tau, tau, cur, order)
        cur = (cur + 1) % 4
        u = sol.get_field(oracle.KIND_VEL, cur)
        p = sol.get_field(oracle.KIND_PRES, cur)
        flow_rate = sol.integral(u)[0] / area
        av = sol.profiles(u, p)
        if t >= t_avg:
            av_sum = tau * av if av_sum is None else av_sum + tau * av
            t_int += tau
        t += tau
        final = t >= time_max or it + 1 == steps_max
        if collocated:
            dudy = av[1, 0] / dz[0]
            dudy_int = av_sum[1, 0] / (dz[0] * t_int) if t_int > 0 else 0.0
        else:
            dudy = av[0, 0] / (0.5 * dz[0])
            dudy_int = av_sum[0, 0] / (0.5 * dz[0] * t_int) if t_int > 0 else 0.0
        ufr = 0.0 if dudy < 0 else np.sqrt(dudy * visc)
        ufr_int = 0.0 if dudy_int < 0 else np.sqrt(dudy_int * visc)
        cfl = sol.cfl(u, tau, 3)
        if it % sprog == 0 or final:
            rows.append((t, flow_rate, tau, cfl, ufr, ufr_int))
        while cfl > cfl_max:
            cfl *= 0.5
            tau *= 0.5
            it_start = it + 1
        if cfl < cfl_min:
            cfl *= 1.5
            tau *= 1.5
            it_start = it + 1
        it += 1
        if final:
            break
    return np.array(rows), u, p, av, (av_sum / t_int if t_int > 0 else None)


def analytic_start(nx, ny, nz, hx, hy, zz, visc, sour_dpdx, u_tau, collocated):
    """The reference's start without input.vtk: laminar parabola plus a pair of
    streamwise rolls (staggered/main.cpp:158-185 with GetCoor :58-66;
    collocated_AB3/main.cpp:348-368)."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    iz, iy, ix = iz.ravel(), iy.ravel(), ix.ravel()
    Hy = hy * ny
    u = np.zeros((nx * ny * nz, 3))

    def rolls(z, zplus, x, y):
        g = zplus * np.exp(-0.01 * zplus * zplus)
        return -np.cos(x) * np.sin(2 * np.pi * y / Hy) * g, np.sin(x) * np.cos(2 * np.pi * y / Hy) * g / (2 * np.pi / Hy)

    if collocated:
        z = zz[iz]
        Hz = zz[-1] - zz[0]
        re = u_tau / visc
        zp = (z - zz[0]) * re
        zplus = np.minimum(zp, 2 * re - zp)
        du, dv = rolls(z, zplus, ix * hx, iy * hy)
        u[:, 0] = 0.5 * z * (Hz - z) * sour_dpdx / visc + du
        u[:, 1] = dv
    else:
        Hz, zmin = zz[-1] - zz[0], zz[0]
        zc = 0.5 * (zz[iz] + zz[iz + 1])
        zplus = np.minimum(zc, Hz - zc) * u_tau / visc
        du, _ = rolls(zc, zplus, ix * hx, (iy + 0.5) * hy)
        u[:, 0] = 0.5 * (zc - zmin) * (zmin + Hz - zc) * sour_dpdx / visc + du
        _, dv = rolls(zc, zplus, (ix + 0.5) * hx, iy * hy)
        u[:, 1] = dv
    return u


def profile_columns(av, zz, visc, collocated, u_tau=1.0):
    """The columns of DumpProfiles (staggered/io.cpp:71-96; collocated_AB3/io.cpp:74-101)
    from per-plane means, in FP64."""
    nz = av.shape[0]
    out = []
    for iz in range(nz):
        if collocated:
            ufr = u_tau
            H = 0.5 * (zz[nz - 1] - zz[0])
            z = zz[iz]
        else:
            ufr = 1.0
            H = 0.5 * (zz[nz] - zz[0])
            z = 0.5 * (zz[iz] + zz[iz + 1]) - zz[0]
        yp = (z if z < H else 2 * H - z) * ufr / visc
        u = av[iz, :3] / ufr
        uu = av[iz, 3:9] / ufr ** 2
        var = [max(0.0, uu[k] - u[k] ** 2) for k in range(3)]
        uw = uu[4] - u[0] * u[2]
        if collocated:
            nut = av[0, 9] if iz == 0 else av[iz - 1, 9] if iz == nz - 1 else 0.5 * (av[iz, 9] + av[iz - 1, 9])
            out.append([yp, u[0], *np.sqrt(var), uw, nut / visc])
        else:
            pp = max(0.0, av[iz, 11] - av[iz, 10] ** 2)
            out.append([yp, u[0], *np.sqrt(var), uw, av[iz, 9] / visc, np.sqrt(pp)])
    return np.array(out)
