"""GPU parity of the collocated SRK stepper (tc_srk_step, collocated_SRK/func_global.h:317-408)
against the FP64 oracle's or_srk_step, through the C ABI, on the same seeded
velocity, pressure and dpdt. FP64 bar: max-norm relative error <= 1e-12 after one
step and <= 1e-10 after four; FP32-bulk: <= 1e-5 after one step."""
import numpy as np
import pytest

import oracle
import turbchannel_amd as tca
from tests import cases, parity

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL_TRAJ = 1e-10
TOL32 = 1e-5

# (method, StabType, alpha_tau, nx, ny, nz)
CASES = [
    (0, 0, 1.0, 16, 12, 18),
    (1, 1, 0.0, 16, 12, 18),   # alpha_tau <= 0: the reference's default 0.5 * alpha_tau_max
    (2, 1, 1.0, 16, 12, 18),   # ARS(3,4,3), StabType 1: the reference's own choice (main.cpp:342-352)
    (2, 0, 0.6, 30, 10, 14),
    (3, 0, 1.0, 16, 12, 18),
    (4, 1, 1.0, 16, 12, 18),
    (4, 0, 1.3, 12, 20, 14),
    (5, 1, 0.8, 16, 12, 18),
    (6, 0, 1.0, 16, 12, 18),
    (7, 1, 1.0, 30, 10, 14),
]


def ids(c):
    return f"{tca.IMEX_METHODS[c[0]]}-stab{c[1]}-a{c[2]}-{c[3]}x{c[4]}x{c[5]}"


def setup(method, stab, alpha_tau, nx, ny, nz, precision=tca.FP64):
    zz = cases.ab3_mesh_tanh(nz, 1.8)
    kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    kw.update(stab_type=stab, fd_mode=1, visc_type=0)
    sol = oracle.Solver(oracle.SCHEME_SRK, nx, ny, nz, zz, **kw)
    eng = tca.Engine(tca.SCHEME_SRK, nx, ny, nz, zz, precision=precision, imex=method, alpha_tau=alpha_tau, **kw)
    a = alpha_tau if alpha_tau > 0 else 0.5 * oracle.imex_table(method)["alpha_tau_max"][stab]
    u = parity.seeded_velocity(sol, seed=7 + method)
    rng = np.random.default_rng(method + 10 * stab)
    p = sol.normalize_pressure(0.05 * rng.standard_normal(sol.n3d))
    q = 0.05 * rng.standard_normal(sol.n3d)
    return sol, eng, a, u, p, q


def put(eng, u, p, q, dt=np.float64):
    eng.set_field(tca.VEL, 0, u.astype(dt))
    eng.set_field(tca.PRES, 0, p.astype(dt))
    eng.set_field(tca.PRES, 1, q.astype(dt))


def get(eng):
    return eng.get_field(tca.VEL, 0), eng.get_field(tca.PRES, 0), eng.get_field(tca.PRES, 1)


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_srk_step_parity(case):
    method, stab, alpha_tau, nx, ny, nz = case
    sol, eng, a, u, p, q = setup(*case)
    tau = 2e-3
    put(eng, u, p, q)
    q0 = q.copy()
    for k in range(4):
        u, p, q = sol.srk_step(k * tau, tau, u, p, q, method=method, stab_type=stab, alpha_tau=a)
        eng.srk_step(k * tau, tau, 0, 0, 1)
        gu, gp, gq = get(eng)
        tol = TOL64 if k == 0 else TOL_TRAJ
        assert parity.rel(gu, u) < tol, (k, parity.rel(gu, u))
        assert parity.rel(gp, p) < tol, (k, parity.rel(gp, p))
        if stab:
            assert parity.rel(gq, q) < tol, (k, parity.rel(gq, q))
        else:  # StabType 0 leaves dpdt alone
            assert np.array_equal(gq, q0)


@pytest.mark.parametrize("case", [CASES[2], CASES[5], CASES[9]], ids=ids)
def test_srk_step_fp32_bulk(case):
    method, stab, alpha_tau, nx, ny, nz = case
    sol, eng, a, u, p, q = setup(*case, precision=tca.FP32_BULK)
    tau = 2e-3
    put(eng, u, p, q, np.float32)
    u1, p1, q1 = sol.srk_step(0.0, tau, u, p, q, method=method, stab_type=stab, alpha_tau=a)
    eng.srk_step(0.0, tau, 0, 0, 1)
    gu, gp, gq = get(eng)
    assert parity.rel(gu, u1) < TOL32
    assert parity.rel(gp, p1) < 50 * TOL32  # pressure: difference quotient of FP32 stage values
    if stab:
        assert parity.rel(gq, q1) < 500 * TOL32


@pytest.mark.parametrize("case", [CASES[2], CASES[6]], ids=ids)
def test_srk_graph_replay_matches_eager(case):
    method, stab, alpha_tau, nx, ny, nz = case
    runs = []
    sol, eng, a, u, p, q = setup(*case)  # one initial state: the oracle's OpenMP sums are not ordered
    for graphs in (False, True):
        eng = tca.Engine(tca.SCHEME_SRK, nx, ny, nz, sol.zz, imex=method, alpha_tau=alpha_tau,
                         **{k: getattr(sol.params, k) for k in ("visc", "sour_dpdx", "inv_length_scale", "hx", "hy",
                                                                "stab_type", "fd_mode", "visc_type")})
        eng.graphs(graphs)
        put(eng, u, p, q)
        out = []
        for k, tau in enumerate([2e-3, 2e-3, 2e-3, 1e-3, 1e-3, 2e-3]):
            eng.srk_step(0.0, tau, 0, 0, 1)
            out.append(get(eng))
        runs.append(out)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_srk_slots_and_refusals():
    case = CASES[5]
    sol, eng, a, u, p, q = setup(*case)
    # any slots: velocity in 2, pressure in 3, dpdt in 0
    eng.set_field(tca.VEL, 2, u)
    eng.set_field(tca.PRES, 3, p)
    eng.set_field(tca.PRES, 0, q)
    eng.srk_step(0.0, 1e-3, 2, 3, 0)
    u1, p1, q1 = sol.srk_step(0.0, 1e-3, u, p, q, method=case[0], stab_type=case[1], alpha_tau=a)
    assert parity.rel(eng.get_field(tca.VEL, 2), u1) < TOL64
    assert parity.rel(eng.get_field(tca.PRES, 3), p1) < TOL64
    assert parity.rel(eng.get_field(tca.PRES, 0), q1) < TOL64
    with pytest.raises(tca.TCError):
        eng.step(0.0, 1e-3, 0, 1)
    with pytest.raises(tca.TCError):
        eng.srk_step(0.0, 1e-3, 0, 1, 1)  # dpdt in the pressure slot
    with pytest.raises(tca.TCError):
        eng.srk_step(0.0, -1.0, 0, 0, 1)
    ab3 = parity.make_engine("ab3")
    with pytest.raises(tca.TCError):
        ab3.srk_step(0.0, 1e-3, 0, 0, 1)


def test_srk_diagnostics_are_collocated():
    """Reductions on an SRK engine are the collocated ones (CalcIntegral,
    CalcVelocityProfile of collocated_SRK/func_global.h:145-231)."""
    sol, eng, a, u, p, q = setup(*CASES[2])
    put(eng, u, p, q)
    ref = sol.integral(u)
    np.testing.assert_allclose(eng.integral((tca.VEL, 0)), ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    got = eng.profiles((tca.VEL, 0), 0)
    ref = sol.profiles(u, p)
    assert parity.rel(got[:, :9], ref[:, :9]) < TOL64
    eng.normalize_pressure((tca.PRES, 0))
    assert parity.rel(eng.get_field(tca.PRES, 0), sol.normalize_pressure(p)) < TOL64
