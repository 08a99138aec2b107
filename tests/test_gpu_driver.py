"""The command-line driver on the GPU: its screen/log lines, VTK fields and
profile files against the same loop replayed on the oracle, the start from a
coarser input.vtk, step-size control in both directions, the exact restart, and
the learned closure read from a file."""
import math
import re

import numpy as np
import pytest

import turbchannel_amd as tca
from tests import cases, parity
from tests.driver_replay import (analytic_start, profile_columns, read_table, read_vtk, replay, run_driver,
                                 write_vtk)

pytestmark = pytest.mark.gpu

STAG = dict(nx=16, ny=12, nz=16)  # the "vv_les" parity case: reference set-up, sinh mesh, LES (NAIVE)
AB3 = dict(nx=16, ny=12, nz=18)   # the "ab3_les" parity case: bulk units, tanh mesh, LES (Galerkin)
AB3_ARGS = ["--scheme", "ab3", "--Re_b", 5600, "--Re_tau", 180, "--Lx", repr(2 * math.pi), "--Ly", repr(math.pi),
            "--mesh", "tanh", "--gamma", 1.8, "--nx", 16, "--ny", 12, "--nz", 18, "--les", 1, "--visc", 1]
STAG_ARGS = ["--scheme", "vv", "--nx", 16, "--ny", 12, "--nz", 16]


def _log_rows(path):
    return read_table(path)


def _check_log(rows_drv, rows_ref):
    assert rows_drv.shape == rows_ref.shape, (rows_drv.shape, rows_ref.shape)
    # columns printed with %f: t, bulk velocity, CFL, u_fric, averaged u_fric
    for c in (0, 1, 3, 4, 5):
        np.testing.assert_allclose(rows_drv[:, c], rows_ref[:, c], rtol=0, atol=1.5e-6, err_msg=f"column {c}")
    np.testing.assert_allclose(rows_drv[:, 2], rows_ref[:, 2], rtol=1e-6)  # tau, %e


def _check_field(path, u, p):
    dims, _, vals = read_vtk(path)
    su, sp = np.abs(u).max(), max(np.abs(p).max(), 1e-30)
    np.testing.assert_allclose(vals[:, :3], u, rtol=1e-6, atol=1e-6 * su)
    np.testing.assert_allclose(vals[:, 3], p, rtol=1e-6, atol=1e-6 * sp)


def _check_profiles(path, cols):
    got = read_table(path)
    assert got.shape == cols.shape
    # %e keeps 7 significant digits; a variance near zero loses digits to the
    # cancellation uu - u^2, hence the small absolute term per column
    scale = np.maximum(np.abs(cols).max(axis=0), 1e-30)
    bad = np.abs(got - cols) > 2e-6 * np.abs(cols) + 1e-7 * scale
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], cols[bad][:5])


def test_driver_staggered_matches_oracle_replay(tmp_path):
    sol = parity.make_oracle("vv_les")
    nx, ny, nz = STAG["nx"], STAG["ny"], STAG["nz"]
    zz = cases.sinh_mesh(nz)
    kw = cases.ref_staggered_kwargs(nx, ny)
    u0 = sol.project(sol.initial_field(seed=11, ubulk=15.0, noise_rms=1.5))
    p0 = np.zeros(nx * ny * nz)
    u0, p0 = write_vtk(tmp_path / "input.vtk", nx, ny, nz, u0, p0)
    cfl_max, cfl_min = 0.3, 0.02
    # above the ceiling already without the eddy viscosity (which the first step adds)
    tau = 1.5 * cfl_max / sol.cfl(u0, 1.0, 3)
    steps = 12
    r = run_driver([*STAG_ARGS, "--tau", repr(tau), "--TimeStepsMax", steps, "-s", 1, "-f", 6, "-p", 6,
                    "--CFLmax", cfl_max, "--CFLmin", cfl_min], tmp_path)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "Initial data read" in r.stdout, r.stdout + r.stderr
    area = kw["hx"] * nx * kw["hy"] * ny * (zz[-1] - zz[0])
    rows, u, p, av, av_mean = replay(sol, u0, p0, tau=tau, steps_max=steps, area=area, dz=np.diff(zz),
                                     visc=kw["visc"], collocated=False, cfl_min=cfl_min, cfl_max=cfl_max)
    assert rows[0, 2] == tau and rows[-1, 2] <= tau / 2  # the control halved the step
    _check_log(_log_rows(tmp_path / "log.dat"), rows)
    assert len(re.findall(r"^T=", r.stdout, re.M)) == steps
    _check_field(tmp_path / "output.vtk", u, p)
    dims, coords, _ = read_vtk(tmp_path / "output.vtk")
    assert dims == (nx, ny, nz)
    np.testing.assert_allclose(coords[0], np.arange(nx) * kw["hx"], rtol=1e-6)
    np.testing.assert_allclose(coords[1], -math.pi + np.arange(ny) * kw["hy"], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(coords[2], zz[:nz], rtol=1e-6, atol=1e-7)
    assert (tmp_path / "q00006.vtk").exists() and (tmp_path / "r00006.dat").exists()
    _check_profiles(tmp_path / "output.dat", profile_columns(av, zz, kw["visc"], False))
    _check_profiles(tmp_path / "res.dat", profile_columns(av_mean, zz, kw["visc"], False))
    m = re.search(r"u_fric: obtained=([0-9.]+), expected=([0-9.]+)", r.stdout)
    dudy = av_mean[0, 0] / (0.5 * (zz[1] - zz[0]))
    assert float(m.group(1)) == pytest.approx(math.sqrt(kw["visc"] * dudy), abs=1.5e-6)
    assert float(m.group(2)) == pytest.approx(1.0)


def test_driver_collocated_from_coarse_input_matches_oracle_replay(tmp_path):
    sol = parity.make_oracle("ab3_les")
    nx, ny, nz = AB3["nx"], AB3["ny"], AB3["nz"]
    zz = cases.ab3_mesh_tanh(nz, 1.8)
    kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    # a field on a grid coarser by 2 in x and y: the reader repeats every value
    fine = sol.initial_field(seed=5, ubulk=1.0, noise_rms=0.1).reshape(nz, ny, nx, 3)
    coarse = fine[:, ::2, ::2, :]
    pc = np.zeros((nz, ny // 2, nx // 2))
    cu, cp = write_vtk(tmp_path / "input.vtk", nx // 2, ny // 2, nz, coarse, pc)
    u0 = np.repeat(np.repeat(cu.reshape(nz, ny // 2, nx // 2, 3), 2, axis=1), 2, axis=2).reshape(-1, 3)
    p0 = np.repeat(np.repeat(cp.reshape(nz, ny // 2, nx // 2), 2, axis=1), 2, axis=2).ravel()
    cfl_max, cfl_min = 0.5, 0.05
    tau = 0.3 * cfl_min / sol.cfl(u0, 1.0, 3)  # CFL below the floor: the step grows by 1.5 repeatedly
    steps = 10
    r = run_driver([*AB3_ARGS, "--tau", repr(tau), "--TimeStepsMax", steps, "-s", 2, "-f", 100, "-p", 100,
                    "--CFLmax", cfl_max, "--CFLmin", cfl_min, "-a", 2.5 * tau], tmp_path)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Stored data {nx // 2}x{ny // 2}x{nz}, reading to {nx}x{ny}x{nz}" in r.stdout
    area = kw["hx"] * nx * kw["hy"] * ny * (zz[-1] - zz[0])
    u_tau = 180 * kw["visc"]
    rows, u, p, av, av_mean = replay(sol, u0, p0, tau=tau, steps_max=steps, area=area, dz=np.diff(zz),
                                     visc=kw["visc"], collocated=True, cfl_min=cfl_min, cfl_max=cfl_max,
                                     t_avg=2.5 * tau, sprog=2)
    assert rows[-1, 2] > 2 * tau  # the control grew the step
    _check_log(_log_rows(tmp_path / "log.dat"), rows)
    _check_field(tmp_path / "output.vtk", u, p)
    _, coords, _ = read_vtk(tmp_path / "output.vtk")
    np.testing.assert_allclose(coords[2], zz, rtol=1e-6, atol=1e-7)
    _check_profiles(tmp_path / "output.dat", profile_columns(av, zz, kw["visc"], True, u_tau))
    _check_profiles(tmp_path / "res.dat", profile_columns(av_mean, zz, kw["visc"], True, u_tau))
    dz = np.diff(zz)
    dudy = 1.5 * av_mean[1, 0] / dz[0] - 0.5 * (av_mean[2, 0] - av_mean[1, 0]) / dz[1]
    m = re.search(r"u_fric: obtained=([0-9.]+), expected=([0-9.]+)", r.stdout)
    assert float(m.group(1)) == pytest.approx(math.sqrt(kw["visc"] * max(dudy, 0)), abs=1.5e-6)


@pytest.mark.parametrize("name", ["vv_les", "ab3_les"])
def test_driver_analytic_start(tmp_path, name):
    # no input.vtk: the reference's analytic start, then two steps
    collocated = name == "ab3_les"
    scheme, nx, ny, nz, zz, kw = parity.problem(name)
    sol = parity.make_oracle(name)
    u_tau = kw["inv_length_scale"] * kw["visc"]
    u0 = analytic_start(nx, ny, nz, kw["hx"], kw["hy"], zz, kw["visc"], kw["sour_dpdx"], u_tau, collocated)
    p0 = np.zeros(nx * ny * nz)
    tau = 0.2 / sol.cfl(u0, 1.0, 3)
    r = run_driver([*(AB3_ARGS if collocated else STAG_ARGS), "--tau", repr(tau), "--TimeStepsMax", 2, "-s", 1,
                    "--CFLmin", 0.01, "--CFLmax", 0.9], tmp_path)
    assert r.returncode == 0 and "input.vtk not found. Starting from the beginning" in r.stdout, r.stdout + r.stderr
    hz = zz[-1] - zz[0]
    rows, u, p, av, _ = replay(sol, u0, p0, tau=tau, steps_max=2, area=kw["hx"] * nx * kw["hy"] * ny * hz,
                               dz=np.diff(zz), visc=kw["visc"], collocated=collocated, cfl_min=0.01, cfl_max=0.9)
    _check_log(_log_rows(tmp_path / "log.dat"), rows)
    _check_field(tmp_path / "output.vtk", u, p)


SRK_ARGS = ["--scheme", "srk", "--Re_b", 5600, "--Re_tau", 180, "--Lx", repr(2 * math.pi), "--Ly", repr(math.pi),
            "--mesh", "tanh", "--gamma", 1.8, "--nx", 16, "--ny", 12, "--nz", 18, "--imex", "ark3", "--stab", 1]


@pytest.mark.parametrize("args", [STAG_ARGS + ["--precision", "fp64"], AB3_ARGS + ["--precision", "fp32"],
                                  SRK_ARGS + ["--precision", "fp64"]],
                         ids=["vv_fp64", "ab3_fp32", "srk_fp64"])
def test_driver_restart_is_exact(tmp_path, args):
    common = [*args, "--seed", 3, "--tau", "2e-3", "-s", 1, "-f", 1000, "-p", 1000]
    full, first = tmp_path / "full", tmp_path / "split"
    full.mkdir(), first.mkdir()
    r = run_driver([*common, "--TimeStepsMax", 10], full)
    assert r.returncode == 0 and ("Total time" in r.stdout or "Time: total" in r.stdout), r.stdout + r.stderr
    r = run_driver([*common, "--TimeStepsMax", 5, "--checkpoint-every", 5], first)
    assert r.returncode == 0 and (first / "restart.bin").exists(), r.stdout + r.stderr
    r = run_driver([*common, "--TimeStepsMax", 10, "--restart", "restart.bin"], first)
    assert r.returncode == 0 and "Restarted from restart.bin at step 5" in r.stdout, r.stdout + r.stderr
    names = ("output.vtk", "res.dat", "log.dat") if "srk" in args else ("output.vtk", "output.dat", "res.dat", "log.dat")
    for name in names:  # the SRK loop writes no output.dat (collocated_SRK/main.cpp:543-551)
        assert (full / name).read_bytes() == (first / name).read_bytes(), name


def test_driver_restart_rejects_other_grid(tmp_path):
    r = run_driver([*STAG_ARGS, "--seed", 1, "--TimeStepsMax", 1, "--checkpoint-every", 1], tmp_path)
    assert r.returncode == 0 and (tmp_path / "restart.bin").exists(), r.stdout
    r = run_driver(["--scheme", "vv", "--nx", 16, "--ny", 16, "--nz", 16, "--restart", "restart.bin"], tmp_path)
    assert "written for another scheme, precision or grid" in r.stdout


def test_driver_learned_closure_from_file(tmp_path):
    # zero weights: the closure's stress is zero, so the run equals the run
    # without a subgrid model; random weights: the run differs and stays finite
    npar = tca.mlp_param_count()
    base = ["--scheme", "ab3", "--Re_b", 5600, "--Re_tau", 180, "--Lx", repr(2 * math.pi), "--Ly", repr(math.pi),
            "--mesh", "tanh", "--nx", 32, "--ny", 16, "--nz", 18, "--seed", 2, "--tau", "2e-3", "--TimeStepsMax", 6,
            "-s", 1, "--precision", "fp64"]
    (tmp_path / "none").mkdir(), (tmp_path / "zero").mkdir(), (tmp_path / "rand").mkdir()
    np.zeros(npar, dtype="<f4").tofile(tmp_path / "zero.bin")
    rng = np.random.default_rng(0)
    w = (rng.standard_normal(npar) * 0.05).astype("<f4")
    w.tofile(tmp_path / "rand.bin")
    r0 = run_driver([*base, "--les", 0], tmp_path / "none")
    r1 = run_driver([*base, "--sgs", "../zero.bin"], tmp_path / "zero")
    r2 = run_driver([*base, "--sgs", "../rand.bin"], tmp_path / "rand")
    for r in (r0, r1, r2):
        assert r.returncode == 0 and "Time: total" in r.stdout, r.stdout + r.stderr
    a = read_vtk(tmp_path / "none" / "output.vtk")[2]
    b = read_vtk(tmp_path / "zero" / "output.vtk")[2]
    c = read_vtk(tmp_path / "rand" / "output.vtk")[2]
    np.testing.assert_allclose(b, a, rtol=1e-6, atol=1e-6 * np.abs(a).max())
    assert np.isfinite(c).all() and np.abs(c - a).max() > 1e-6 * np.abs(a).max()
    # backscatter clipping (sgs_model 2): another run again, finite
    (tmp_path / "clip").mkdir()
    r3 = run_driver([*base, "--sgs", "../rand.bin", "--sgs_clip", 1], tmp_path / "clip")
    assert r3.returncode == 0 and "Time: total" in r3.stdout, r3.stdout + r3.stderr
    d = read_vtk(tmp_path / "clip" / "output.vtk")[2]
    assert np.isfinite(d).all() and np.abs(d - c).max() > 1e-6 * np.abs(a).max()
    # dissipative part along the strain rate (sgs_model 3): finite, differs from both
    (tmp_path / "proj").mkdir()
    r4 = run_driver([*base, "--sgs", "../rand.bin", "--sgs_proj", 1], tmp_path / "proj")
    assert r4.returncode == 0 and "Time: total" in r4.stdout, r4.stdout + r4.stderr
    f = read_vtk(tmp_path / "proj" / "output.vtk")[2]
    assert np.isfinite(f).all() and min(np.abs(f - c).max(), np.abs(f - d).max()) > 1e-6 * np.abs(a).max()
    np.zeros(npar - 1, dtype="<f4").tofile(tmp_path / "short.bin")
    r = run_driver([*base, "--sgs", "short.bin"], tmp_path)
    assert f"expected {npar} float32 network parameters" in r.stdout


def test_driver_srk_matches_oracle_replay(tmp_path):
    """--scheme srk: the fixed-step loop of collocated_SRK/main.cpp:472-550 against
    or_srk_step replayed on the oracle (log lines, 5-value VTK with dp/dt,
    profiles, time-averaged profiles from t = 0 and the friction velocity)."""
    import oracle
    nx, ny, nz = 16, 12, 18
    zz = cases.ab3_mesh_tanh(nz, 1.8)
    kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    sol = oracle.Solver(oracle.SCHEME_SRK, nx, ny, nz, zz, stab_type=1, fd_mode=1, visc_type=0, **kw)
    u0 = sol.project(sol.initial_field(seed=9, ubulk=1.0, noise_rms=0.1))
    u0, p0 = write_vtk(tmp_path / "input.vtk", nx, ny, nz, u0, np.zeros(nx * ny * nz))
    tau, steps, alpha_tau = 2e-3, 9, 0.8
    r = run_driver(["--scheme", "srk", "--Re_b", 5600, "--Re_tau", 180, "--Lx", repr(2 * math.pi), "--Ly",
                    repr(math.pi), "--mesh", "tanh", "--gamma", 1.8, "--nx", nx, "--ny", ny, "--nz", nz, "--imex",
                    "ark3", "--stab", 1, "--alpha_tau", alpha_tau, "--tau", repr(tau), "--TimeStepsMax", steps,
                    "-s", 2, "-f", 4], tmp_path)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "Initial data read" in r.stdout, r.stdout + r.stderr
    area = kw["hx"] * nx * kw["hy"] * ny * (zz[-1] - zz[0])
    dz = np.diff(zz)
    u, p, q = u0, p0, np.zeros_like(p0)
    rows, av_sum = [], 0.0
    for it in range(steps):
        u, p, q = sol.srk_step(it * tau, tau, u, p, q, method=4, stab_type=1, alpha_tau=alpha_tau)
        av = sol.profiles(u, p)
        av_sum = av_sum + tau * av
        t = (it + 1) * tau
        dudy, dudy_int = av[1, 0] / dz[0], av_sum[1, 0] / (dz[0] * t)
        if it % 2 == 0 or it == steps - 1:
            rows.append((t, sol.integral(u)[0] / area, math.sqrt(max(dudy, 0) * kw["visc"]),
                         math.sqrt(max(dudy_int, 0) * kw["visc"])))
        if it == 4:
            u4, p4, q4 = u, p, q
    got = read_table(tmp_path / "log.dat")
    assert got.shape == (len(rows), 4)
    np.testing.assert_allclose(got, np.array(rows), rtol=0, atol=1.5e-6)
    for name, (uu, pp, qq) in (("q00004.vtk", (u4, p4, q4)), ("output.vtk", (u, p, q))):
        dims, _, vals = read_vtk(tmp_path / name)
        assert dims == (nx, ny, nz) and vals.shape[1] == 5
        np.testing.assert_allclose(vals[:, :3], uu, rtol=1e-6, atol=1e-6 * np.abs(uu).max())
        np.testing.assert_allclose(vals[:, 3], pp, rtol=1e-6, atol=1e-6 * np.abs(pp).max())
        np.testing.assert_allclose(vals[:, 4], qq, rtol=1e-6, atol=1e-6 * np.abs(qq).max())
    assert (tmp_path / "q00008.vtk").exists() and (tmp_path / "r00008.dat").exists()  # the final step
    u_tau = 180 * kw["visc"]
    _check_profiles(tmp_path / "r00008.dat", profile_columns(av, zz, kw["visc"], True, u_tau))
    _check_profiles(tmp_path / "res.dat", profile_columns(av_sum / (steps * tau), zz, kw["visc"], True, u_tau))
    m = re.search(r"u_fric: obtained=([0-9.]+), expected=([0-9.]+)", r.stdout)
    a = av_sum / (steps * tau)
    dudy = 1.5 * a[1, 0] / dz[0] - 0.5 * (a[2, 0] - a[1, 0]) / dz[1]
    assert float(m.group(1)) == pytest.approx(math.sqrt(kw["visc"] * max(dudy, 0)), abs=1.5e-6)
    assert float(m.group(2)) == pytest.approx(u_tau, abs=1e-6)


# a narrow CFL band: the control changes tau on most steps, in both directions
@pytest.mark.parametrize("args", [STAG_ARGS + ["--les", 1, "--CFLmax", 0.06, "--CFLmin", 0.05],
                                  AB3_ARGS + ["--CFLmax", 0.3, "--CFLmin", 0.25]], ids=["vv", "ab3"])
def test_driver_pipelined_loop_writes_the_same_files(tmp_path, args):
    """Launching step n+1 before reading step n's diagnostics (and launching it
    again when the step-size control changes tau) changes no output byte."""
    common = [*args, "--seed", 3, "--tau", "2e-3", "-s", 1, "-f", 4, "-p", 3, "--TimeStepsMax", 11,
              "--checkpoint-every", 5]
    dirs = {}
    for pipe in (0, 1):
        d = tmp_path / f"pipe{pipe}"
        d.mkdir()
        r = run_driver([*common, "--pipeline", pipe], d)
        assert r.returncode == 0, r.stdout + r.stderr
        dirs[pipe] = d
    taus = _log_rows(dirs[1] / "log.dat")[:, 2]
    assert len(set(taus.tolist())) > 1  # the control changed the step
    names = sorted(p.name for p in dirs[0].iterdir())
    assert names == sorted(p.name for p in dirs[1].iterdir()) and "q00004.vtk" in names
    for name in names:
        assert (dirs[0] / name).read_bytes() == (dirs[1] / name).read_bytes(), name


@pytest.mark.parametrize("args", [STAG_ARGS + ["--les", 1, "--CFLmax", 0.06, "--CFLmin", 0.05]], ids=["vv"])
def test_driver_graph_cache_clear_with_a_step_in_flight(tmp_path, args):
    """With the graph cache capped at 2 entries, the step-size control makes the
    engine drop its captured graphs on most steps while the pipelined loop has the
    next step in flight; the files are those of the unpipelined loop."""
    common = [*args, "--seed", 3, "--tau", "2e-3", "-s", 1, "-f", 4, "-p", 3, "--TimeStepsMax", 11]
    dirs = {}
    for pipe in (0, 1):
        d = tmp_path / f"pipe{pipe}"
        d.mkdir()
        r = run_driver([*common, "--pipeline", pipe], d, env={"TC_GRAPH_CACHE_MAX": "2"})
        assert r.returncode == 0, r.stdout + r.stderr
        dirs[pipe] = d
    assert len(set(_log_rows(dirs[1] / "log.dat")[:, 2].tolist())) > 2
    names = sorted(p.name for p in dirs[0].iterdir())
    assert names == sorted(p.name for p in dirs[1].iterdir())
    for name in names:
        assert (dirs[0] / name).read_bytes() == (dirs[1] / name).read_bytes(), name
