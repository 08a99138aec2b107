"""GPU parity of the learned subgrid closure (through the C ABI) against the CPU
restatement in oracle/mlp.py and the FP64 oracle.

Bars, each stated where it is checked:
* network forward / gradient: the GPU and the oracle share the bfloat16 rounding
  model but sum in FP32 vs FP64, so an activation can round to the neighbouring
  bfloat16 value; outputs agree to 2e-2 of their max norm (typically ~1e-3),
  gradients to 2e-2 in 2-norm, the loss to 1e-4 relative;
* Adam: 1e-6 absolute on the weights (lr 1e-3), fed the GPU's own gradient;
* training samples: 1e-9 relative (FP64 engine; output is FP32-rounded values
  of FP64 arithmetic: 1e-6) and 2e-3 (FP32-bulk engine);
* the divergence of tau inside the explicit term and the step: 1e-12, given the
  same tau field."""
import numpy as np
import pytest

import oracle
import turbchannel_amd as tca
from oracle import mlp
from tests import parity

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    """float32 device copy (torch used only for device memory)."""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")
    torch.cuda.synchronize()
    return t


def empty(n):
    t = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    return t


def rand_data(n, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 10)).astype(np.float32)
    X[:, 9] = rng.uniform(0, 2, n)
    T = (0.3 * rng.normal(size=(n, 6))).astype(np.float32)
    return X, T


def test_param_count_matches_oracle():
    assert tca.mlp_param_count() == mlp.n_params()


@pytest.mark.parametrize("n", [64, 1000, 70000])
def test_forward_matches_oracle(n):
    net = tca.Mlp(seed=3)
    p = net.get_params()
    X, _ = rand_data(n)
    x, y = dev(X), empty(6 * n)
    net.forward(x.data_ptr(), n, y.data_ptr())
    got = y.cpu().numpy().reshape(n, 6)
    ref, _ = mlp.forward(p, X, emulate=True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"forward n={n}: max-norm rel {err:.2e}, mean abs rel {np.abs(got - ref).mean() / np.abs(ref).mean():.2e}")
    assert err < 2e-2
    assert np.abs(got - ref).mean() < 2e-3 * np.abs(ref).mean()


@pytest.mark.parametrize("n", [200, 4096])
def test_gradient_matches_oracle(n):
    net = tca.Mlp(seed=5)
    p = net.get_params()
    X, T = rand_data(n, seed=2)
    x, t = dev(X), dev(T)  # keep the device buffers alive across the call
    loss, g = net.gradient(x.data_ptr(), t.data_ptr(), n)
    lref, gref = mlp.loss_and_grad(p, X, T, emulate=True)
    eg = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"gradient n={n}: loss rel {abs(loss - lref) / lref:.2e}, grad 2-norm rel {eg:.2e}")
    assert abs(loss - lref) < 1e-4 * lref
    assert eg < 2e-2
    # every layer's block, not only the largest
    off = 0
    for i, o in mlp.LAYERS:
        for sz in (o * i, o):
            blk = slice(off, off + sz)
            assert np.linalg.norm(g[blk] - gref[blk]) < 5e-2 * np.linalg.norm(gref[blk]) + 1e-12
            off += sz


def test_gradient_is_deterministic():
    net = tca.Mlp(seed=7)
    X, T = rand_data(5000, seed=4)
    x, t = dev(X), dev(T)
    l1, g1 = net.gradient(x.data_ptr(), t.data_ptr(), 5000)
    l2, g2 = net.gradient(x.data_ptr(), t.data_ptr(), 5000)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_adam_steps_match_oracle():
    net = tca.Mlp(seed=9)
    X, T = rand_data(3000, seed=6)
    x, t = dev(X), dev(T)
    p = net.get_params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in (1, 2, 3):
        _, g = net.gradient(x.data_ptr(), t.data_ptr(), 3000)
        net.train_step(x.data_ptr(), t.data_ptr(), 3000, lr=1e-3)
        p, m, v = mlp.adam(p, g.astype(np.float64), m, v, step, lr=1e-3)
        assert np.abs(net.get_params() - p).max() < 1e-6
    assert net.steps() == 3


def test_training_fits_a_teacher_network():
    teacher = tca.Mlp(seed=100)
    X, _ = rand_data(8192, seed=8)
    x = dev(X)
    y = empty(6 * 8192)
    teacher.forward(x.data_ptr(), 8192, y.data_ptr())
    student = tca.Mlp(seed=1)
    losses = [student.train_step(x.data_ptr(), y.data_ptr(), 8192, lr=3e-3) for _ in range(300)]
    print(f"teacher fit: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.1 * losses[0]


def test_set_params_round_trip_and_bad_count():
    net = tca.Mlp(seed=1)
    p = np.linspace(-1, 1, net.n_params).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    bad = np.zeros(10, dtype=np.float32)
    rc = tca.lib().tc_mlp_set_params(net._h, bad.ctypes.data_as(tca.ctypes.POINTER(tca.ctypes.c_float)), 10)
    assert rc == 1 and "count" in tca.lib().tc_last_error().decode()


# --- training samples --------------------------------------------------------
def sample_case(name="ab3_sgs"):
    scheme, nx, ny, nz, zz, kw = parity.problem(name)
    return nx, ny, nz, zz, kw


@pytest.mark.parametrize("precision,tol", [(tca.FP64, 1e-6), (tca.FP32_BULK, 2e-3)])
@pytest.mark.parametrize("factor", [2, 4])
def test_samples_match_oracle(precision, tol, factor):
    nx, ny, nz, zz, kw = sample_case()
    sol = parity.make_oracle("ab3_sgs")
    eng = parity.make_engine("ab3_sgs", precision=precision)
    u = parity.seeded_velocity(sol, seed=21, noise=0.3)
    eng.set_field(tca.VEL, 1, u)
    n = eng.sgs_sample_count(factor)
    x, t = empty(10 * n), empty(6 * n)
    assert eng.sgs_samples((tca.VEL, 1), factor, x.data_ptr(), t.data_ptr(), n) == n
    X = x.cpu().numpy().reshape(n, 10)
    T = t.cpu().numpy().reshape(n, 6)
    u_tau = kw["visc"] * kw["inv_length_scale"]
    Xr, Tr = mlp.samples(u, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau, kw["inv_length_scale"], factor)
    ex = parity.rel(X, Xr)
    et = parity.rel(T, Tr)
    print(f"samples factor {factor}: inputs {ex:.2e}, targets {et:.2e}")
    assert ex < tol and et < tol
    assert np.abs(Tr).max() > 0  # the filter leaves a sub-filter stress


def test_samples_reject_bad_factor():
    eng = parity.make_engine("ab3_sgs")
    x = empty(10)
    rc = tca.lib().tc_sgs_samples(eng._h, tca.VEL, 0, 3, x.data_ptr(), x.data_ptr(), 10, None)
    assert rc == 1 and "factor" in tca.lib().tc_last_error().decode()


# --- the closure inside the step ---------------------------------------------------
def closure_params(seed=11):
    return tca.Mlp(seed=seed).get_params()


@pytest.mark.parametrize("name", ["ab3_sgs", "ab3_sgs_gal"])
def test_node_tau_matches_oracle(name):
    nx, ny, nz, zz, kw = sample_case(name)
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    p = closure_params()
    eng.set_sgs_params(p)
    u = parity.seeded_velocity(sol, seed=31, noise=0.3)
    eng.set_field(tca.VEL, 2, u)
    eng.recalc_turb_visc((tca.VEL, 2))
    got = eng.get_field(tca.TAU, 0)
    u_tau = kw["visc"] * kw["inv_length_scale"]
    ref = mlp.node_tau(p, u, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau, kw["inv_length_scale"])
    e = parity.rel(got, ref)
    print(f"{name}: node tau max-norm rel {e:.2e}")
    assert e < 2e-2
    nxy = nx * ny
    assert np.all(got[:nxy] == 0) and np.all(got[-nxy:] == 0)


def test_node_tau_with_backscatter_clipping_matches_oracle():
    """sgs_model = 2: the closure stress is zero at the nodes where tau_ij S_ij > 0.
    Nodes whose sign of tau_ij S_ij is within the network's bfloat16 rounding of
    zero may go either way; every other node matches the oracle's clipped stress."""
    name = "ab3_sgs"
    nx, ny, nz, zz, kw = sample_case(name)
    sol = parity.make_oracle(name)
    scheme, *_, pkw = parity.problem(name)
    pkw = dict(pkw, sgs_model=2)
    eng = tca.Engine(scheme, nx, ny, nz, zz, **pkw)
    p = closure_params()
    eng.set_sgs_params(p)
    u = parity.seeded_velocity(sol, seed=31, noise=0.3)
    eng.set_field(tca.VEL, 2, u)
    eng.recalc_turb_visc((tca.VEL, 2))
    got = eng.get_field(tca.TAU, 0)
    u_tau = kw["visc"] * kw["inv_length_scale"]
    args = (p, u, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau, kw["inv_length_scale"])
    ref = mlp.node_tau(*args, clip=True)
    raw = mlp.node_tau(*args, clip=False)
    X = mlp.features([u.reshape(nz, ny, nx, 3)[..., c] for c in range(3)], zz, kw["hx"], kw["hy"], u_tau,
                     kw["inv_length_scale"]).reshape(-1, 10)
    inner = slice(nx * ny, -nx * ny)
    y = raw[inner] / (u_tau * u_tau)
    ts = mlp.stress_strain(y, X)
    bound = 2.0 * np.abs(y).sum(axis=1) * np.abs(X[:, :9]).max(axis=1)  # |tau_ij S_ij| <= bound
    firm = np.abs(ts) > 0.05 * bound
    clipped = (ts > 0) & firm
    print(f"clipped {clipped.mean():.2%} of the interior nodes, {firm.mean():.2%} firm")
    assert 0.05 < clipped.sum() / max(1, firm.sum()) < 0.95
    g, r = got[inner], ref[inner]
    e = np.abs(g[firm] - r[firm]).max() / np.abs(r).max()
    print(f"clipped node tau max-norm rel {e:.2e}")
    assert e < 2e-2
    assert np.all(g[clipped] == 0)
    assert np.all(got[:nx * ny] == 0) and np.all(got[-nx * ny:] == 0)


def test_node_tau_projected_on_the_strain_rate_matches_oracle():
    """sgs_model = 3: the closure stress is its dissipative part along the
    deviatoric strain rate. Nodes whose sign of tau:D is within the network's
    bfloat16 rounding of zero may go either way; every other node matches the
    oracle's projected stress, and the dropped ones are zero."""
    name = "ab3_sgs"
    nx, ny, nz, zz, kw = sample_case(name)
    sol = parity.make_oracle(name)
    scheme, *_, pkw = parity.problem(name)
    eng = tca.Engine(scheme, nx, ny, nz, zz, **dict(pkw, sgs_model=3))
    p = closure_params()
    eng.set_sgs_params(p)
    u = parity.seeded_velocity(sol, seed=32, noise=0.3)
    eng.set_field(tca.VEL, 2, u)
    eng.recalc_turb_visc((tca.VEL, 2))
    got = eng.get_field(tca.TAU, 0)
    u_tau = kw["visc"] * kw["inv_length_scale"]
    args = (p, u, nx, ny, nz, zz, kw["hx"], kw["hy"], u_tau, kw["inv_length_scale"])
    ref = mlp.node_tau(*args, project=True)
    raw = mlp.node_tau(*args)
    X = mlp.features([u.reshape(nz, ny, nx, 3)[..., c] for c in range(3)], zz, kw["hx"], kw["hy"], u_tau,
                     kw["inv_length_scale"]).reshape(-1, 10)
    inner = slice(nx * ny, -nx * ny)
    y = raw[inner] / (u_tau * u_tau)
    G = X[:, :9].reshape(-1, 3, 3)
    S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    D = S - (np.trace(S, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    d6 = np.stack([D[:, 0, 0], D[:, 1, 1], D[:, 2, 2], D[:, 0, 1], D[:, 0, 2], D[:, 1, 2]], -1)
    yd = np.sum(np.array([1, 1, 1, 2, 2, 2]) * y * d6, axis=1)  # tau:D
    bound = 2.0 * np.abs(y).sum(axis=1) * np.abs(d6).max(axis=1)  # |tau:D| <= bound
    firm = np.abs(yd) > 0.05 * bound
    kept = (yd < 0) & firm
    dropped = (yd > 0) & firm
    print(f"kept {kept.mean():.2%}, dropped {dropped.mean():.2%} of the interior nodes")
    assert kept.sum() > 0.05 * firm.sum() and dropped.sum() > 0.05 * firm.sum()
    g, r = got[inner], ref[inner]
    e = np.abs(g[firm] - r[firm]).max() / np.abs(r).max()
    print(f"projected node tau max-norm rel {e:.2e}")
    assert e < 2e-2
    assert np.all(g[dropped] == 0)
    assert np.all(got[:nx * ny] == 0) and np.all(got[-nx * ny:] == 0)


@pytest.mark.parametrize("name", ["ab3_sgs", "ab3_sgs_gal"])
def test_explicit_term_with_tau(name):
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    u = parity.seeded_velocity(sol, seed=41)
    rng = np.random.default_rng(3)
    tau = 1e-3 * rng.normal(size=(sol.n3d, 6))
    nxy = sol.nx * sol.ny
    tau[:nxy] = 0
    tau[-nxy:] = 0
    sol.set_field(oracle.KIND_TAU, 0, tau)
    eng.set_field(tca.TAU, 0, tau)
    eng.set_field(tca.VEL, 0, u)
    eng.explicit_term((tca.VEL, 0), (tca.FLUX, 0))
    ref = sol.explicit_term(u)
    assert parity.rel(eng.get_field(tca.FLUX, 0), ref) < 1e-12
    # the term is really there: without tau the answer changes
    sol.set_field(oracle.KIND_TAU, 0, np.zeros_like(tau))
    assert parity.rel(sol.explicit_term(u), ref) > 1e-6


@pytest.mark.parametrize("name", ["ab3_sgs", "ab3_sgs_gal"])
def test_step_with_learned_closure(name):
    """Step on the GPU (tau from its own network evaluation); the oracle steps
    with that same tau field."""
    sol = parity.make_oracle(name)
    eng = parity.make_engine(name)
    eng.set_sgs_params(closure_params(13))
    cur, tau_t = 1, 2e-3
    for k, slot in enumerate((cur, (cur + 3) % 4, (cur + 2) % 4)):
        u = parity.seeded_velocity(sol, seed=50 + k, noise=0.2)
        pz = np.cos(np.arange(sol.n3d) * 0.11 + k)
        for kind, a in ((tca.VEL, u), (tca.PRES, pz)):
            sol.set_field(kind, slot, a)
            eng.set_field(kind, slot, a)
        sol.set_field(oracle.KIND_TAU, 0, np.zeros((sol.n3d, 6)))
        f = sol.explicit_term(u)
        sol.set_field(tca.FLUX, slot, f)
        eng.set_field(tca.FLUX, slot, f)
    eng.step(0.0, tau_t, cur, 3)
    tau_gpu = eng.get_field(tca.TAU, 0)
    assert np.abs(tau_gpu).max() > 0
    sol.set_field(oracle.KIND_TAU, 0, tau_gpu)
    sol.step(0.0, tau_t, cur, 3)
    nxt = (cur + 1) % 4
    for kind in (tca.VEL, tca.PRES):
        e = parity.rel(eng.get_field(kind, nxt), sol.get_field(kind, nxt))
        assert e < 1e-12, (kind, e)


def test_step_without_weights_is_an_error():
    eng = parity.make_engine("ab3_sgs")
    with pytest.raises(tca.TCError, match="tc_set_sgs_params"):
        eng.step(0.0, 1e-3, 0, 1)


# --- what the trained closure gains ------------------------------------------------
def collect(eng, sol, seed, factor, snap_steps, dt=2e-3):
    """Samples of the filtered field at the given step counts of one run from the
    seeded initial field (noise 0.3)."""
    n = eng.sgs_sample_count(factor)
    X, T = [], []
    eng.set_field(tca.VEL, 0, parity.seeded_velocity(sol, seed=seed, noise=0.3))
    cur = 0
    for it in range(max(snap_steps)):
        eng.step(it * dt, dt, cur, min(it + 1, 3))
        cur = (cur + 1) % 4
        if it + 1 in snap_steps:
            x, t = empty(10 * n), empty(6 * n)
            eng.sgs_samples((tca.VEL, cur), factor, x.data_ptr(), t.data_ptr(), n)
            X.append(x.cpu().numpy().reshape(n, 10))
            T.append(t.cpu().numpy().reshape(n, 6))
    return np.concatenate(X), np.concatenate(T)


def test_trained_closure_beats_no_model_and_smagorinsky():
    """A priori, on a held-out run: mean-square error of the sub-filter stress
    predicted by the network trained on the GPU (standardised data, folded back
    into raw units), against tau = 0 and against the reference's
    Smagorinsky-van Driest model evaluated on the same filtered field."""
    from tests import cases
    from turbchannel_amd.closure import Standardizer, train

    nx, ny, nz, factor = 64, 32, 26, 4
    zz = cases.ab3_mesh_tanh(nz, 1.8)
    kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    kw.update(stab_type=1, fd_mode=1, visc_type=0, cfr_ubulk=1.0)
    sol = oracle.Solver(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    eng = tca.Engine(tca.SCHEME_AB3, nx, ny, nz, zz, **kw)
    snaps = [600, 900, 1200]
    parts = [collect(eng, sol, sd, factor, snaps) for sd in range(101, 107)]
    Xtr = np.concatenate([p[0] for p in parts])
    Ttr = np.concatenate([p[1] for p in parts])
    Xte, Tte = collect(eng, sol, 999, factor, [800, 1200])
    perm = np.random.default_rng(0).permutation(len(Xtr))
    Xtr, Ttr = Xtr[perm], Ttr[perm]
    st = Standardizer(Xtr, Ttr)
    net = tca.Mlp(seed=0)
    x, t = dev(st.inputs(Xtr)), dev(st.targets(Ttr))
    losses = train(net, x.data_ptr(), t.data_ptr(), len(Xtr), batch=4096, iters=1500, lr=3e-3, lr_final=3e-4)
    raw = tca.Mlp(seed=0)
    raw.set_params(st.fold(net.get_params()))
    xe, ye = dev(Xte), empty(6 * len(Xte))
    raw.forward(xe.data_ptr(), len(Xte), ye.data_ptr())
    pred = ye.cpu().numpy().reshape(-1, 6)
    dzc = np.diff(zz)
    cellz = 0.5 * (dzc[:-1] + dzc[1:])  # interior node cell heights
    plane = (nx // factor) * (ny // factor)
    hz = np.tile(np.repeat(cellz, plane), len(Xte) // (plane * (nz - 2)))
    cell_plus = (kw["hx"] * factor * kw["hy"] * factor * hz) ** (1.0 / 3.0) * kw["inv_length_scale"]
    smag = mlp.smagorinsky_tau(Xte.astype(np.float64), cell_plus)
    mse = lambda a: float(np.mean((a - Tte) ** 2))  # noqa: E731
    e_net, e_zero, e_smag = mse(pred), mse(0 * Tte), mse(smag)

    # sub-filter energy transfer per plane, <tau_ij S_ij> (wall units), the quantity a closure must carry
    def transfer(tau):
        G = Xte[:, :9].astype(np.float64).reshape(-1, 3, 3)
        S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
        ts = sum(tau[:, k] * S[:, a, b] * (1 if a == b else 2) for k, (a, b) in enumerate(mlp.PAIRS))
        return ts.reshape(-1, nz - 2, plane).mean(axis=(0, 2))

    P_true = transfer(Tte.astype(np.float64))
    perr = lambda tau: float(np.linalg.norm(transfer(tau) - P_true) / np.linalg.norm(P_true))  # noqa: E731
    p_net, p_smag = perr(pred.astype(np.float64)), perr(smag)
    print(f"training loss {losses[0]:.3f} -> {losses[-1]:.3f} (standardised); a priori MSE (wall units): "
          f"learned {e_net:.4e}, no model {e_zero:.4e}, Smagorinsky {e_smag:.4e}; "
          f"transfer-profile error: learned {p_net:.3f}, Smagorinsky {p_smag:.3f}")
    assert e_net < 0.97 * min(e_zero, e_smag)
    assert p_net < 0.5 * p_smag
