"""Step diagnostics formed in the projection kernel (tc_step_diagnostics): the
flow rate, kinetic energy, CFL numbers and plane profiles of the step's new
fields (staggered/main.cpp:312-338, collocated_AB3/main.cpp:414-441) against the
oracle's CalcIntegral, CalcKineticEnergy, CalcCFL and CalcProfiles of the
oracle's own step, and against the engine's separate one-pass kernel. FP64 bar:
1e-11 relative (the step itself agrees to 1e-12); the fused and separate sums
of one engine differ only in summation order (1e-13)."""
import numpy as np
import pytest

import oracle
import turbchannel_amd as tca
from tests import cases, parity

pytestmark = pytest.mark.gpu

TOL = 1e-11
TOL_ORDER = 1e-13

# the parity cases, plus planes of several 256-cell chunks with a ragged last one,
# and of more than one 1024-cell block (4 cells per thread) with a ragged last block
NAMES = ["hw", "hw_les", "vv", "vv_les", "vv_odd", "ab3", "ab3_les", "ab3_stab0", "hw_wide", "ab3_wide", "vv_wide2",
         "ab3_wide2"]
WIDE = {
    "hw_wide": (tca.SCHEME_HW, 48, 20, 12, dict(turb_visc=1, visc_type=1)),
    "ab3_wide": (tca.SCHEME_AB3, 40, 24, 14, dict(stab_type=1, fd_mode=1, visc_type=0)),
    "vv_wide2": (tca.SCHEME_VV, 50, 24, 12, dict()),
    "ab3_wide2": (tca.SCHEME_AB3, 50, 24, 12, dict(stab_type=0, fd_mode=1, visc_type=1, turb_visc=1)),
}


def make(name, cfr=0.0, precision=tca.FP64):
    if name not in WIDE:
        return parity.make_oracle(name, cfr_ubulk=cfr), parity.make_engine(name, precision=precision, cfr_ubulk=cfr)
    scheme, nx, ny, nz, extra = WIDE[name]
    zz = cases.ab3_mesh_tanh(nz, 1.8) if scheme == tca.SCHEME_AB3 else cases.tanh_faces(nz, 1.8)
    kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    kw.update(extra, cfr_ubulk=cfr)
    return (oracle.Solver(scheme, nx, ny, nz, zz, **kw),
            tca.Engine(scheme, nx, ny, nz, zz, precision=precision, **kw))


def load_history(sol, engs, cur, seed=11, dt=np.float64):
    """Vel/flux/pres slots cur, cur-1, cur-2 with consistent seeded data."""
    for k, slot in enumerate((cur, (cur + 3) % 4, (cur + 2) % 4)):
        u = parity.seeded_velocity(sol, seed=seed + k, noise=0.1 + 0.02 * k)
        f = sol.explicit_term(u)
        pz = np.cos(np.arange(sol.n3d) * 0.11 + k)
        for kind, a in ((tca.VEL, u), (tca.FLUX, f), (tca.PRES, pz)):
            a = a.astype(dt).astype(np.float64)
            sol.set_field(kind, slot, a)
            for e in engs:
                e.set_field(kind, slot, a.astype(dt))


def expected(sol, slot, tau, mode):
    u = sol.get_field(tca.VEL, slot)
    p = sol.get_field(tca.PRES, slot)
    return dict(integral=sol.integral(u), kinetic_energy=sol.kinetic_energy(u), cfl=sol.cfl(u, tau, mode),
                profiles=sol.profiles(u, p))


def assert_close(d, ref, tol):
    s = np.abs(ref["integral"]).max()
    assert np.abs(d["integral"] - ref["integral"]).max() <= tol * s, (d["integral"], ref["integral"])
    assert abs(d["kinetic_energy"] - ref["kinetic_energy"]) <= tol * abs(ref["kinetic_energy"])
    assert abs(d["cfl"] - ref["cfl"]) <= tol * abs(ref["cfl"])
    pr = ref["profiles"]
    assert np.abs(d["profiles"] - pr).max() <= tol * max(1.0, np.abs(pr).max())


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("cfr", [0.0, 1.0])
def test_step_diagnostics_parity(name, cfr):
    sol, eng = make(name, cfr)
    _, sep = make(name, cfr)
    eng.step_diagnostics(True)
    cur, tau = 1, 2e-3
    load_history(sol, [eng, sep], cur)
    sol.step(0.0, tau, cur, 3)
    eng.step(0.0, tau, cur, 3)
    sep.step(0.0, tau, cur, 3)
    nxt = (cur + 1) % 4
    for mode in (1, 2, 3):  # every read of the step's record gives the step's values
        d = eng.diagnostics(nxt, nxt, tau, mode)
        assert_close(d, expected(sol, nxt, tau, mode), TOL)
        assert_close(d, sep.diagnostics(nxt, nxt, tau, mode), TOL_ORDER)
    assert np.array_equal(eng.get_field(tca.VEL, nxt), sep.get_field(tca.VEL, nxt))
    assert np.array_equal(eng.get_field(tca.PRES, nxt), sep.get_field(tca.PRES, nxt))


@pytest.mark.parametrize("name", ["vv_les", "ab3", "hw_wide"])
def test_step_diagnostics_fp32_bulk(name):
    """FP32 fields, FP64 sums: the fused record against the separate pass."""
    sol, eng = make(name, precision=tca.FP32_BULK)
    _, sep = make(name, precision=tca.FP32_BULK)
    eng.step_diagnostics(True)
    cur, tau = 2, 2e-3
    load_history(sol, [eng, sep], cur, dt=np.float32)
    eng.step(0.0, tau, cur, 2)
    sep.step(0.0, tau, cur, 2)
    nxt = (cur + 1) % 4
    assert np.array_equal(eng.get_field(tca.VEL, nxt), sep.get_field(tca.VEL, nxt))
    assert_close(eng.diagnostics(nxt, nxt, tau, 3), sep.diagnostics(nxt, nxt, tau, 3), TOL_ORDER)


@pytest.mark.parametrize("name", ["vv", "ab3_les"])
def test_step_record_is_dropped_when_fields_change(name):
    sol, eng = make(name)
    eng.step_diagnostics(True)
    cur, tau = 0, 1e-3
    load_history(sol, [eng], cur)
    sol.step(0.0, tau, cur, 1)
    eng.step(0.0, tau, cur, 1)
    nxt = 1
    assert_close(eng.diagnostics(nxt, nxt, tau, 3), expected(sol, nxt, tau, 3), TOL)
    # another slot: a pass of its own, which replaces the step's record
    assert_close(eng.diagnostics(cur, cur, tau, 3), expected(sol, cur, tau, 3), TOL)
    assert_close(eng.diagnostics(nxt, nxt, tau, 3), expected(sol, nxt, tau, 3), TOL)
    # a new field in the step's slot
    eng.step(0.0, tau, cur, 1)
    u2 = parity.seeded_velocity(sol, seed=99, noise=0.3)
    eng.set_field(tca.VEL, nxt, u2)
    sol.set_field(tca.VEL, nxt, u2)
    assert_close(eng.diagnostics(nxt, nxt, tau, 3), expected(sol, nxt, tau, 3), TOL)
    # nu_t recomputed from another field
    if sol.params.turb_visc:
        eng.step(0.0, tau, cur, 1)
        sol.step(0.0, tau, cur, 1)
        eng.recalc_turb_visc((tca.VEL, nxt))
        sol.recalc_turb_visc(sol.get_field(tca.VEL, nxt))
        assert_close(eng.diagnostics(nxt, nxt, tau, 3), expected(sol, nxt, tau, 3), TOL)
    # staggered: another pressure slot is a pass of its own
    if name == "vv":
        eng.step(0.0, tau, cur, 1)
        sol.step(0.0, tau, cur, 1)
        u = sol.get_field(tca.VEL, nxt)
        ref = sol.profiles(u, sol.get_field(tca.PRES, cur))
        got = eng.diagnostics(nxt, cur, tau, 3)["profiles"]
        assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("name,precision", [("hw_les", tca.FP64), ("ab3_wide", tca.FP64), ("vv", tca.FP32_BULK)])
def test_step_diagnostics_graph_replay(name, precision):
    """Captured steps with the fused record give the eager values bit for bit on every replay."""
    sol, e0 = make(name, cfr=1.0, precision=precision)
    _, e1 = make(name, cfr=1.0, precision=precision)
    dt = np.float32 if precision == tca.FP32_BULK else np.float64
    u = parity.seeded_velocity(sol, seed=21).astype(dt)
    for e in (e0, e1):
        e.step_diagnostics(True)
        e.set_field(tca.VEL, 0, u)
    e1.graphs(True)
    cur = 0
    for it in range(10):
        tau = 1e-3 if it < 7 else 5e-4
        for e in (e0, e1):
            e.step(it * tau, tau, cur, min(it + 1, 3))
        cur = (cur + 1) % 4
        a, b = (e.diagnostics(cur, cur, tau, 3) for e in (e0, e1))
        for k in a:
            assert np.array_equal(a[k], b[k]), (it, k)
        if it == 4:  # switched off and on again: new captures, same values
            for e in (e0, e1):
                e.step_diagnostics(False)
                e.step_diagnostics(True)


def test_step_diagnostics_refused_on_srk():
    zz = cases.ab3_mesh_tanh(14, 1.8)
    kw = cases.bulk_kwargs(16, 12, 5600, 180)
    kw.update(stab_type=1, fd_mode=1, visc_type=0)
    eng = tca.Engine(tca.SCHEME_SRK, 16, 12, 14, zz, imex=2, alpha_tau=1.0, **kw)
    with pytest.raises(tca.TCError):
        eng.step_diagnostics(True)


@pytest.mark.parametrize("name", ["vv", "hw_les", "ab3_les"])
@pytest.mark.parametrize("latest_first", [False, True])
def test_records_of_the_last_two_steps(name, latest_first):
    """Each of the last two steps keeps the record it formed (with the nu_t of its
    own step), so a caller may read a step's diagnostics after launching the next."""
    sol, eng = make(name)
    eng.step_diagnostics(True)
    cur, tau = 0, 1e-3
    load_history(sol, [eng], cur)
    sol.step(0.0, tau, 0, 1)
    exp1 = expected(sol, 1, tau, 3)
    sol.step(tau, tau, 1, 2)
    exp2 = expected(sol, 2, tau, 3)
    eng.step(0.0, tau, 0, 1)
    eng.step(tau, tau, 1, 2)
    order = [(2, exp2), (1, exp1)] if latest_first else [(1, exp1), (2, exp2)]
    for slot, exp in order * 2:  # each read leaves the records in place
        assert_close(eng.diagnostics(slot, slot, tau, 3), exp, TOL)
    eng.step(2 * tau, tau, 2, 3)  # the third step's record replaces the first's
    sol.step(2 * tau, tau, 2, 3)
    assert_close(eng.diagnostics(3, 3, tau, 3), expected(sol, 3, tau, 3), TOL)
    assert_close(eng.diagnostics(2, 2, tau, 3), exp2, TOL)
