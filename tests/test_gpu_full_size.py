"""GPU parity at the benchmark workloads' full sizes, against the FP64 oracle on
the same seeded input (the reference runs its own CPU-vs-GPU checks at its
production size, staggered/selftest.cpp:360-438).

Workloads (bench.py's table; reference axes x, span, wall-normal):
* hw_64x64x64      second-order staggered DNS, 200 steps (the CPU-path case);
* vv_192x160x128   fourth-order staggered DNS;
* ab3_384x384x192  collocated BDF3/EX3 DNS at bulk Re 13750;
* vv_384x384x192   fourth-order staggered DNS at the same size.
At 384 x 384 points per plane the node-grid kernels other than the projection
run in plain block order (four planes of a vector field exceed one XCD's L2), so
these cases cover that launch order; the smaller ones cover the XCD-contiguous
order.

Bars, each checked below:
* FP64: max-norm relative error of velocity and pressure <= 1e-12 after the
  first step and <= 1e-10 after the order ramp (1, 2, 3) and after 200 steps;
  for the staggered schemes (pressure matrix = D G) the discrete divergence of
  the engine's velocity, formed by the oracle's operator, at round-off: below
  1e-11 of one face-flux term (the oracle's own step reaches 3e-14 at 64^3,
  3e-13 at 192 x 160 x 128 and 1.1e-12 at 384 x 384 x 192); the collocated
  scheme's compact pressure Laplacian is not D G, so its divergence is not
  round-off and is not checked; the bulk velocity held at 1 to 1e-12 by the
  constant-flow-rate correction;
* FP32-bulk: one BDF3 step from the same (FP32-rounded) history as the FP64
  oracle, velocity within 1e-5 relative of the oracle's.
"""
import gc
import math
import os
import sys

import numpy as np
import pytest

import oracle
import turbchannel_amd as tca
from tests import parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL_TRAJ = 1e-10
TOL32 = 1e-5
FULL = ["hw_64x64x64", "vv_192x160x128", "ab3_384x384x192", "vv_384x384x192"]


def setup(name, precision):
    scheme, nx, ny, nz, zz, kw, dt = bench.problem(name, precision)
    okw = {k: v for k, v in kw.items() if k != "precision"}
    sol = oracle.Solver(scheme, nx, ny, nz, zz, **okw)
    eng = tca.Engine(scheme, nx, ny, nz, zz, **kw)
    return sol, eng, dt, kw


def divergence_scale(sol, kw, v):
    """Size of one face-flux term of the divergence: |v| times the largest face area."""
    dz = np.diff(sol.zz).max()
    return np.abs(v).max() * max(kw["hx"] * kw["hy"], kw["hx"] * dz, kw["hy"] * dz)


def bulk_velocity(sol, q):
    """Volume integral of u over the box volume (walls at z = 0 and 2)."""
    return q / (sol.params.hx * sol.nx * sol.params.hy * sol.ny * 2.0)


@pytest.mark.parametrize("name", FULL)
def test_full_size_fp64_ramp(name):
    oracle.set_threads(min(os.cpu_count() or 1, 16))
    sol, eng, dt, kw = setup(name, tca.FP64)
    u0 = sol.project(sol.initial_field(seed=42))
    sol.set_field(oracle.KIND_VEL, 0, u0)
    eng.set_field(tca.VEL, 0, u0)
    del u0
    nsteps = 200 if name == "hw_64x64x64" else 3
    checks = {1, 3, 50, 100, 200}
    cur = 0
    for it in range(nsteps):
        order = min(it + 1, 3)
        sol.step(it * dt, dt, cur, order)
        eng.step(it * dt, dt, cur, order)
        cur = (cur + 1) % 4
        if it + 1 in checks:
            ev = parity.rel(eng.get_field(tca.VEL, cur), sol.get_field(oracle.KIND_VEL, cur))
            ep = parity.rel(eng.get_field(tca.PRES, cur), sol.get_field(oracle.KIND_PRES, cur))
            print(f"{name} step {it + 1}: velocity {ev:.2e}, pressure {ep:.2e}", flush=True)
            tol = TOL64 if it == 0 else TOL_TRAJ
            assert ev < tol and ep < tol, (it + 1, ev, ep)
    ub = bulk_velocity(sol, eng.integral((tca.VEL, cur))[0])
    print(f"{name}: bulk velocity - 1 = {ub - 1:.2e}", flush=True)
    assert abs(ub - 1.0) < 1e-12
    if sol.scheme != tca.SCHEME_AB3:
        v = eng.get_field(tca.VEL, cur)
        div = np.abs(sol.apply_divergence(v)).max() / divergence_scale(sol, kw, v)
        print(f"{name}: divergence {div:.2e} of one face-flux term", flush=True)
        assert div < 1e-11
    eng.close()
    sol.close()
    gc.collect()


@pytest.mark.parametrize("name", ["vv_192x160x128", "ab3_384x384x192", "vv_384x384x192"])
def test_full_size_fp32_bulk_per_step(name):
    """One BDF3 step of the FP32-bulk engine against the FP64 oracle from the same
    history: the oracle's ramp (orders 1, 2, 3) fills every ring slot, all slots
    are rounded to FP32 and loaded into both, then both take one order-3 step."""
    oracle.set_threads(min(os.cpu_count() or 1, 16))
    sol, eng, dt, kw = setup(name, tca.FP32_BULK)
    sol.set_field(oracle.KIND_VEL, 0, sol.project(sol.initial_field(seed=7)))
    for it in range(3):
        sol.step(it * dt, dt, it, it + 1)
    for kind in (oracle.KIND_VEL, oracle.KIND_FLUX, oracle.KIND_PRES):
        for slot in range(4):
            a = sol.get_field(kind, slot).astype(np.float32)
            eng.set_field(kind, slot, a)
            sol.set_field(kind, slot, a.astype(np.float64))
            del a
    cur = 3
    sol.step(3 * dt, dt, cur, 3)
    eng.step(3 * dt, dt, cur, 3)
    nxt = (cur + 1) % 4
    ev = parity.rel(eng.get_field(tca.VEL, nxt), sol.get_field(oracle.KIND_VEL, nxt))
    ep = parity.rel(eng.get_field(tca.PRES, nxt), sol.get_field(oracle.KIND_PRES, nxt))
    ub = bulk_velocity(sol, eng.integral((tca.VEL, nxt))[0])
    print(f"{name} FP32-bulk one step vs FP64 oracle: velocity {ev:.2e}, pressure {ep:.2e}, "
          f"bulk - 1 = {ub - 1:.2e}", flush=True)
    assert ev < TOL32, ev
    assert abs(ub - 1.0) < 1e-6  # FP32 sums of an FP64-corrected field
    eng.close()
    sol.close()
    gc.collect()


def test_headline_grid_sizes():
    """The table above is the workloads' own: grid sizes as the benchmark sets them."""
    sizes = {n: bench.WORKLOADS[n][1:4] for n in FULL}
    assert sizes == {"hw_64x64x64": (64, 64, 64), "vv_192x160x128": (192, 160, 128),
                     "ab3_384x384x192": (384, 384, 192), "vv_384x384x192": (384, 384, 192)}
    assert math.isclose(bench.WORKLOADS["ab3_384x384x192"][5], 13750.0)
