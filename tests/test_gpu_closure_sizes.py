"""The learned closure and the LES steps at the sizes of the closure
configurations, against the CPU restatement (oracle/mlp.py) and the FP64
oracle:

* training samples from a 384 x 384 slab of the collocated DNS filtered 4x in x
  and y (the 96 x 96 coarse planes of the training set), FP64 and FP32-bulk
  engines;
* the loss gradient on one batch of 65536 samples;
* the closure stress at every node of the 64 x 64 x 96 LES grid, and one LES step
  with the learned closure there (the oracle steps with the GPU's tau);
* one LES step with the reference's Smagorinsky-van Driest model on the 64 x 64 x 96
  grid and on the 1.5x finer 96 x 96 x 144 grid.

Bars as in test_gpu_closure.py: samples 1e-6 (FP64 engine, FP32 outputs) and
2e-3 (FP32-bulk engine) max-norm relative; network outputs 2e-2 max-norm (shared
bfloat16 rounding model, FP32 against FP64 sums); gradient 2e-2 in 2-norm and
loss 1e-4; steps 1e-12 (FP64).
"""
import math

import numpy as np
import pytest

import oracle
import turbchannel_amd as tca
from oracle import mlp
from tests import cases, parity

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def empty(n):
    t = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    return t


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")
    torch.cuda.synchronize()
    return t


def les_problem(nx, ny, nz, **extra):
    """Collocated channel at Re_tau 590 (bulk Re 21900), box 2 pi x pi x 2, tanh 2.0."""
    zz = cases.ab3_mesh_tanh(nz, 2.0)
    kw = cases.bulk_kwargs(nx, ny, 21900.0, 590.0)
    kw.update(stab_type=1, fd_mode=1, visc_type=1, cfr_ubulk=1.0)
    kw.update(extra)
    return zz, kw


@pytest.mark.parametrize("precision,tol", [(tca.FP64, 1e-6), (tca.FP32_BULK, 2e-3)])
def test_samples_from_a_384_slab(precision, tol):
    nx, ny, nz = 384, 384, 10
    zz = cases.ab3_mesh_tanh(nz, 2.0)
    kw = cases.bulk_kwargs(nx, ny, 13750.0, 395.0)
    kw.update(stab_type=1, fd_mode=1, visc_type=0)
    sol = oracle.Solver(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    eng = tca.Engine(tca.SCHEME_AB3, nx, ny, nz, zz, precision=precision, **kw)
    u = sol.initial_field(seed=8, ubulk=1.0, noise_rms=0.3)
    eng.set_field(tca.VEL, 0, u)
    n = eng.sgs_sample_count(4)
    assert n == 96 * 96 * (nz - 2)
    x, t = empty(10 * n), empty(6 * n)
    assert eng.sgs_samples((tca.VEL, 0), 4, x.data_ptr(), t.data_ptr(), n) == n
    X = x.cpu().numpy().reshape(n, 10)
    T = t.cpu().numpy().reshape(n, 6)
    u_tau = kw["visc"] * kw["inv_length_scale"]
    uref = eng.get_field(tca.VEL, 0).astype(np.float64)  # the engine's (rounded) field
    Xr, Tr = mlp.samples(uref, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau, kw["inv_length_scale"], 4)
    ex, et = parity.rel(X, Xr), parity.rel(T, Tr)
    print(f"384x384x{nz} slab, factor 4: {n} samples, inputs {ex:.2e}, targets {et:.2e}", flush=True)
    assert ex < tol and et < tol
    assert np.abs(Tr).max() > 0


def test_gradient_on_a_batch_of_65536():
    n = 65536
    rng = np.random.default_rng(12)
    X = rng.normal(size=(n, 10)).astype(np.float32)
    X[:, 9] = rng.uniform(0, 2, n)
    T = (0.3 * rng.normal(size=(n, 6))).astype(np.float32)
    net = tca.Mlp(seed=9)
    p = net.get_params()
    x, t = dev(X), dev(T)
    loss, g = net.gradient(x.data_ptr(), t.data_ptr(), n)
    lref, gref = mlp.loss_and_grad(p, X, T, emulate=True)
    eg = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"batch 65536: loss rel {abs(loss - lref) / lref:.2e}, gradient 2-norm rel {eg:.2e}", flush=True)
    assert abs(loss - lref) < 1e-4 * lref
    assert eg < 2e-2


def closure_params(seed=21):
    net = tca.Mlp(seed=seed)
    p = net.get_params()
    net.close()
    p[-(6 * 64 + 6):] *= 0.05  # a small stress, as a trained closure gives
    return p


def load_history(sol, eng, cur, seed):
    """Consistent seeded history in slots cur, cur-1, cur-2 (velocity, flux, pressure)."""
    for k, slot in enumerate((cur, (cur + 3) % 4, (cur + 2) % 4)):
        u = sol.project(sol.initial_field(seed=seed + k, ubulk=1.0, noise_rms=0.15))
        f = sol.explicit_term(u)
        pz = np.cos(np.arange(sol.n3d) * 0.11 + k)
        for kind, a in ((tca.VEL, u), (tca.FLUX, f), (tca.PRES, pz)):
            sol.set_field(kind, slot, a)
            eng.set_field(kind, slot, a)


def test_learned_les_at_64x64x96():
    nx, ny, nz = 64, 64, 96
    zz, kw = les_problem(nx, ny, nz, sgs_model=1)
    sol = oracle.Solver(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    eng = tca.Engine(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    p = closure_params()
    eng.set_sgs_params(p)
    u_tau = kw["visc"] * kw["inv_length_scale"]
    # the closure stress at every node
    u = sol.project(sol.initial_field(seed=3, ubulk=1.0, noise_rms=0.2))
    eng.set_field(tca.VEL, 2, u)
    eng.recalc_turb_visc((tca.VEL, 2))
    got = eng.get_field(tca.TAU, 0)
    ref = mlp.node_tau(p, u, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau, kw["inv_length_scale"])
    e = parity.rel(got, ref)
    print(f"LES 64x64x96: node tau max-norm rel {e:.2e}", flush=True)
    assert e < 2e-2
    # one step: the oracle steps with the GPU's tau
    sol.set_field(oracle.KIND_TAU, 0, np.zeros((sol.n3d, 6)))
    cur, dt = 1, 2e-3
    load_history(sol, eng, cur, 60)
    eng.step(0.0, dt, cur, 3)
    sol.set_field(oracle.KIND_TAU, 0, eng.get_field(tca.TAU, 0))
    sol.step(0.0, dt, cur, 3)
    nxt = (cur + 1) % 4
    for kind in (tca.VEL, tca.PRES):
        e = parity.rel(eng.get_field(kind, nxt), sol.get_field(kind, nxt))
        print(f"LES 64x64x96 learned step: kind {kind} {e:.2e}", flush=True)
        assert e < 1e-12


@pytest.mark.parametrize("grid", [(64, 64, 96), (96, 96, 144)], ids=lambda g: "x".join(map(str, g)))
def test_smagorinsky_les_step(grid):
    nx, ny, nz = grid
    zz, kw = les_problem(nx, ny, nz, turb_visc=1)
    sol = oracle.Solver(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    eng = tca.Engine(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    cur, dt = 1, 2e-3
    load_history(sol, eng, cur, 70)
    sol.step(0.0, dt, cur, 3)
    eng.step(0.0, dt, cur, 3)
    nxt = (cur + 1) % 4
    for kind in (tca.VEL, tca.PRES):
        e = parity.rel(eng.get_field(kind, nxt), sol.get_field(kind, nxt))
        print(f"Smagorinsky LES {nx}x{ny}x{nz} step: kind {kind} {e:.2e}", flush=True)
        assert e < 1e-12
    m = nx * ny * (nz - 1)
    e = parity.rel(eng.get_field(tca.NUT, 0)[:m], sol.get_field(tca.NUT, 0)[:m])
    assert e < 1e-12
    assert math.isfinite(e)
