"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle on the
same seeded inputs. FP64 bar: max-norm relative error <= 1e-12 per operator and
per step (the FFTs and reduction orders differ, so results are not bitwise).
FP32-bulk bar: <= 1e-5 relative per step against the FP64 engine."""
import numpy as np
import pytest

import oracle
import turbchannel_amd as tca
from tests import cases, parity

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL_TRAJ = 1e-10
TOL32 = 1e-5
# closure cases: test_gpu_closure.py; the deep cases: FP32-bulk per-step error only
ALL = [c for c in parity.CASES if "sgs" not in c and "deep" not in c]


def fields(sol, seed=3):
    u = parity.seeded_velocity(sol, seed=seed)
    n = sol.n3d
    p = np.sin(np.arange(n, dtype=np.float64) ** 2 + 87.0)
    return u, p


@pytest.mark.parametrize("name", ALL)
def test_explicit_term(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    u, _ = fields(sol)
    if sol.params.turb_visc:  # non-trivial nu_t from the model itself
        sol.recalc_turb_visc(u)
        nut = sol.get_field(tca.NUT, 0)
        eng.set_field(tca.NUT, 0, nut)
    eng.set_field(tca.VEL, 0, u)
    eng.explicit_term((tca.VEL, 0), (tca.VEL, 1))
    got = eng.get_field(tca.VEL, 1)
    ref = sol.explicit_term(u)
    assert parity.rel(got, ref) < TOL64


@pytest.mark.parametrize("name", ["hw_les", "vv_les", "ab3_les"])
def test_turbulent_viscosity(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    u, _ = fields(sol)
    eng.set_field(tca.VEL, 2, u)
    eng.recalc_turb_visc((tca.VEL, 2))
    sol.recalc_turb_visc(u)
    ref = sol.get_field(tca.NUT, 0)
    got = eng.get_field(tca.NUT, 0)
    nz_el = sol.nz - 1 if sol.scheme == tca.SCHEME_AB3 else sol.nz  # AB3: elements only
    m = sol.nx * sol.ny * nz_el
    assert parity.rel(got[:m], ref[:m]) < TOL64


@pytest.mark.parametrize("name", ALL)
def test_gradient_divergence_solve(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    u, p = fields(sol)
    # ApplyGradient
    eng.set_field(tca.PRES, 1, p)
    eng.apply_gradient((tca.PRES, 1), (tca.FLUX, 2))
    assert parity.rel(eng.get_field(tca.FLUX, 2), sol.apply_gradient(p)) < TOL64
    # UnapplyGradient = solve(div v) incl. normalisation; v with a non-zero divergence
    v = u + 0.3 * np.stack([np.cos(np.arange(sol.n3d) * 0.37 + c) for c in range(3)], axis=1)
    if sol.scheme != tca.SCHEME_AB3:
        v[: sol.nx * sol.ny, 2] = 0.0
    eng.set_field(tca.FLUX, 1, v)
    eng.unapply_gradient((tca.FLUX, 1), (tca.PRES, 3))
    ref = sol.unapply_gradient(v)
    got = eng.get_field(tca.PRES, 3)
    assert parity.rel(got, ref) < TOL64
    # NormalizePressure on its own
    eng.set_field(tca.PRES, 0, p)
    eng.normalize_pressure((tca.PRES, 0))
    assert parity.rel(eng.get_field(tca.PRES, 0), sol.normalize_pressure(p)) < TOL64


@pytest.mark.parametrize("name", ALL)
def test_reductions(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    u, p = fields(sol)
    eng.set_field(tca.VEL, 3, u)
    eng.set_field(tca.PRES, 2, p)
    if sol.params.turb_visc:
        sol.recalc_turb_visc(u)
        eng.set_field(tca.NUT, 0, sol.get_field(tca.NUT, 0))
    np.testing.assert_allclose(eng.integral((tca.VEL, 3)), sol.integral(u), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(eng.kinetic_energy((tca.VEL, 3)), sol.kinetic_energy(u), rtol=1e-12)
    for mode in (1, 2, 3):
        np.testing.assert_allclose(eng.cfl((tca.VEL, 3), 1e-3, mode), sol.cfl(u, 1e-3, mode), rtol=1e-12)
    got = eng.profiles((tca.VEL, 3), 2)
    ref = sol.profiles(u, p)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("name", ALL)
def test_fused_diagnostics(name):
    """tc_diagnostics (one pass) against the oracle's CalcIntegral,
    CalcKineticEnergy, CalcCFL and CalcProfiles on the same fields."""
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    u, p = fields(sol)
    eng.set_field(tca.VEL, 1, u)
    eng.set_field(tca.PRES, 0, p)
    if sol.params.turb_visc:
        sol.recalc_turb_visc(u)
        eng.set_field(tca.NUT, 0, sol.get_field(tca.NUT, 0))
    for mode in (1, 2, 3):
        d = eng.diagnostics(1, 0, 1e-3, mode)
        np.testing.assert_allclose(d["cfl"], sol.cfl(u, 1e-3, mode), rtol=1e-12)
    np.testing.assert_allclose(d["integral"], sol.integral(u), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(d["kinetic_energy"], sol.kinetic_energy(u), rtol=1e-12)
    ref = sol.profiles(u, p)
    assert np.abs(d["profiles"] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_fused_diagnostics_fp32_matches_separate_calls():
    eng = parity.make_engine("vv_les", precision=tca.FP32_BULK)
    sol = parity.make_oracle("vv_les")
    u, p = fields(sol)
    eng.set_field(tca.VEL, 2, u)
    eng.set_field(tca.PRES, 2, p)
    eng.recalc_turb_visc((tca.VEL, 2))
    d = eng.diagnostics(2, 2, 2e-3, 3)
    np.testing.assert_allclose(d["integral"], eng.integral((tca.VEL, 2)), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d["cfl"], eng.cfl((tca.VEL, 2), 2e-3, 3), rtol=1e-12)
    np.testing.assert_allclose(d["profiles"], eng.profiles((tca.VEL, 2), 2), rtol=1e-12, atol=1e-12)


def load_history(sol, eng, cur, seed=11):
    """Fill vel/flux/pres slots cur, cur-1, cur-2 with consistent seeded data."""
    for k, slot in enumerate((cur, (cur + 3) % 4, (cur + 2) % 4)):
        u = parity.seeded_velocity(sol, seed=seed + k, noise=0.1 + 0.02 * k)
        f = sol.explicit_term(u)
        pz = np.cos(np.arange(sol.n3d) * 0.11 + k)
        for s_, e_ in ((tca.VEL, u), (tca.FLUX, f), (tca.PRES, pz)):
            sol.set_field(s_, slot, e_)
            eng.set_field(s_, slot, e_)


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_single_step(name, order):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    cur, tau = 1, 2e-3
    load_history(sol, eng, cur)
    sol.step(0.0, tau, cur, order)
    eng.step(0.0, tau, cur, order)
    nxt = (cur + 1) % 4
    for kind, slot in ((tca.VEL, nxt), (tca.PRES, nxt), (tca.FLUX, cur), (tca.FLUX, nxt)):
        e = parity.rel(eng.get_field(kind, slot), sol.get_field(kind, slot))
        assert e < TOL64, (kind, slot, e)
    if sol.params.turb_visc:
        m = sol.nx * sol.ny * (sol.nz - 1)
        assert parity.rel(eng.get_field(tca.NUT, 0)[:m], sol.get_field(tca.NUT, 0)[:m]) < TOL64


def run_trajectory(stepper, nsteps, tau):
    cur = 0
    for it in range(nsteps):
        stepper.step(it * tau, tau, cur, min(it + 1, 3))
        cur = (cur + 1) % 4
    return cur


@pytest.mark.parametrize("name", ["hw", "vv", "vv_les", "ab3", "ab3_les"])
@pytest.mark.parametrize("cfr", [0.0, 1.0])
def test_trajectory_ten_steps(name, cfr):
    sol = parity.make_oracle(name, cfr_ubulk=cfr)
    eng = parity.make_engine(name, cfr_ubulk=cfr)
    u0 = parity.seeded_velocity(sol, seed=42)
    sol.set_field(tca.VEL, 0, u0)
    eng.set_field(tca.VEL, 0, u0)
    tau = 2e-3
    cur = run_trajectory(sol, 10, tau)
    assert run_trajectory(eng, 10, tau) == cur
    for kind in (tca.VEL, tca.PRES):
        e = parity.rel(eng.get_field(kind, cur), sol.get_field(kind, cur))
        assert e < TOL_TRAJ, (kind, e)
    if cfr > 0:
        vol = sol.integral(sol.get_field(tca.VEL, cur))[0]
        assert abs(eng.integral((tca.VEL, cur))[0] - vol) < 1e-12 * abs(vol)


@pytest.mark.parametrize("name", ["vv_odd", "ab3"])
def test_initial_field_on_device(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    eng.initial_field(0, seed=42, ubulk=1.0, noise_rms=0.1)
    got = eng.get_field(tca.VEL, 0)
    ref = parity.seeded_velocity(sol, seed=42)
    assert parity.rel(got, ref) < TOL64


@pytest.mark.parametrize("name", ["vv", "vv_odd", "hw"])
def test_projection_is_divergence_free(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    eng.initial_field(0, seed=5)
    v = eng.get_field(tca.VEL, 0)
    assert np.abs(sol.apply_divergence(v)).max() < 1e-11


@pytest.mark.parametrize("name", ["hw", "vv", "vv_les", "ab3", "ab3_les", "hw_deep", "ab3_deep"])
def test_fp32_bulk_per_step_error(name):
    e64 = parity.make_engine(name, precision=tca.FP64)
    e32 = parity.make_engine(name, precision=tca.FP32_BULK)
    sol = parity.make_oracle(name)
    cur, tau = 1, 2e-3
    for k, slot in enumerate((cur, (cur + 3) % 4, (cur + 2) % 4)):
        u = parity.seeded_velocity(sol, seed=20 + k)
        f = sol.explicit_term(u)
        # the FP32 engine starts from the FP32-rounded state; the FP64 engine from the same values
        u32, f32 = u.astype(np.float32), f.astype(np.float32)
        pz = np.cos(np.arange(sol.n3d) * 0.11 + k).astype(np.float32)
        for kind, a in ((tca.VEL, u32), (tca.FLUX, f32), (tca.PRES, pz)):
            e32.set_field(kind, slot, a)
            e64.set_field(kind, slot, a.astype(np.float64))
    e64.step(0.0, tau, cur, 3)
    e32.step(0.0, tau, cur, 3)
    nxt = (cur + 1) % 4
    ev = parity.rel(e32.get_field(tca.VEL, nxt), e64.get_field(tca.VEL, nxt))
    assert ev < TOL32, ev


@pytest.mark.parametrize("name,precision,cfr", [("vv", tca.FP64, 1.0), ("hw_les", tca.FP64, 0.0),
                                                ("ab3_les", tca.FP32_BULK, 1.0), ("ab3", tca.FP64, 0.0)])
def test_graph_replay_matches_eager(name, precision, cfr):
    """Steps replayed from captured HIP graphs give the eager results bit for bit,
    on every replay, with other work (diagnostics, field reads) in between."""
    sol = parity.make_oracle(name, cfr_ubulk=cfr)
    u = parity.seeded_velocity(sol, seed=21)
    engs = [parity.make_engine(name, precision=precision, cfr_ubulk=cfr) for _ in range(2)]
    engs[1].graphs(True)
    for e in engs:
        e.set_field(tca.VEL, 0, u)
    cur = 0
    for it in range(12):
        tau = 1e-3 if it < 9 else 5e-4  # a change of step: new captures
        order = min(it + 1, 3) if it < 9 else min(it - 8, 3)
        for e in engs:
            e.step(it * tau, tau, cur, order)
        cur = (cur + 1) % 4
        a, b = (e.get_field(tca.VEL, cur) for e in engs)
        assert np.array_equal(a, b), it
        assert np.array_equal(engs[0].get_field(tca.PRES, cur), engs[1].get_field(tca.PRES, cur)), it
        d0, d1 = (e.diagnostics(cur, cur, tau) for e in engs)
        assert d0["cfl"] == d1["cfl"]


def test_timing_counts_launches():
    eng = parity.make_engine("vv")
    eng.initial_field(0)
    eng.timing(True)
    run_trajectory(eng, 3, 1e-3)
    t = eng.timing_read()
    assert t["flux"][1] == 3 and t["zsolve"][1] == 3 and t["flux"][0] > 0


def test_bad_slot_and_order_raise():
    eng = parity.make_engine("vv")
    with pytest.raises(tca.TCError):
        eng.step(0.0, 1e-3, 0, 4)
    with pytest.raises(tca.TCError):
        eng.get_field(tca.VEL, 7)
    short = np.zeros(5)
    rc = tca.lib().tc_set_field(eng._h, tca.PRES, 0, short.ctypes.data, short.nbytes)
    assert rc == 1 and "size mismatch" in tca.lib().tc_last_error().decode()


# Planes of 256 x 192 (FP64) and 384 x 256 (FP32) points: too large for an XCD's
# L2 to hold four of them, so the kernels run in their plain block order, and
# an odd number of line pairs per transform block (3 at N = 256 in FP64). A few
# planes keep the oracle fast.
LARGE = [(tca.SCHEME_HW, 256, 192, 8), (tca.SCHEME_VV, 256, 192, 8), (tca.SCHEME_AB3, 256, 192, 9)]


def divergent_field(u, scheme, nxy):
    """u plus a pattern with a divergence (the seeded field alone is a discrete
    curl, divergence-free in the HW discretisation); no flux through the bottom
    wall for the staggered schemes."""
    n = u.shape[0]
    v = u + 0.3 * np.stack([np.cos(np.arange(n) * 0.37 + c) for c in range(3)], axis=1)
    if scheme != tca.SCHEME_AB3:
        v[:nxy, 2] = 0.0
    return v


def large_problem(scheme, nx, ny, nz):
    if scheme == tca.SCHEME_AB3:
        zz = cases.ab3_mesh_tanh(nz, 1.8)
        kw = cases.bulk_kwargs(nx, ny, 5600, 180)
        kw.update(stab_type=1, fd_mode=1, visc_type=0)
    else:
        zz = cases.tanh_faces(nz, 1.8)
        kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    return zz, kw


@pytest.mark.parametrize("case", LARGE, ids=lambda c: f"{['hw', 'vv', 'ab3'][c[0]]}-{c[1]}x{c[2]}x{c[3]}")
def test_pressure_solve_large_planes(case):
    scheme, nx, ny, nz = case
    zz, kw = large_problem(scheme, nx, ny, nz)
    sol = oracle.Solver(scheme, nx, ny, nz, zz, **kw)
    eng = tca.Engine(scheme, nx, ny, nz, zz, **kw)
    u = divergent_field(sol.initial_field(seed=5, ubulk=1.0, noise_rms=0.3), scheme, nx * ny)
    eng.set_field(tca.VEL, 0, u)
    eng.unapply_gradient((tca.VEL, 0), (tca.PRES, 0))
    assert parity.rel(eng.get_field(tca.PRES, 0), sol.unapply_gradient(u)) < TOL64


@pytest.mark.parametrize("case", LARGE, ids=lambda c: f"{['hw', 'vv', 'ab3'][c[0]]}-{c[1]}x{c[2]}x{c[3]}")
def test_pressure_solve_large_planes_fp32(case):
    scheme, _, _, nz = case
    nx, ny = 384, 256
    zz, kw = large_problem(scheme, nx, ny, nz)
    e64 = tca.Engine(scheme, nx, ny, nz, zz, **kw)
    e32 = tca.Engine(scheme, nx, ny, nz, zz, precision=tca.FP32_BULK, **kw)
    u = divergent_field(tca.make_initial_field(e64.params, seed=6, ubulk=1.0, noise_rms=0.3), scheme,
                        nx * ny).astype(np.float32)
    e64.set_field(tca.VEL, 0, u.astype(np.float64))
    e32.set_field(tca.VEL, 0, u)
    for e in (e64, e32):
        e.unapply_gradient((tca.VEL, 0), (tca.PRES, 0))
    assert parity.rel(e32.get_field(tca.PRES, 0), e64.get_field(tca.PRES, 0)) < TOL32
