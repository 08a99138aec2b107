"""Pins of the learned-closure oracle (oracle/mlp.py) on the CPU.

The network, its gradient and Adam are checked against an independent
implementation (torch on the CPU, float64 autograd and torch.optim.Adam) and
central finite differences; bfloat16 rounding against torch's own conversion;
the training-sample filter against band-limited fields and a closed form."""
import numpy as np
import pytest
import torch

from oracle import mlp


def rand_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for i, o in mlp.LAYERS:
        b = scale / np.sqrt(i)
        out.append((rng.uniform(-b, b, (o, i)), rng.uniform(-b, b, o)))
    return mlp.pack(out)


def rand_data(n, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 10))
    X[:, 9] = rng.uniform(0, 2, n)
    T = 0.3 * rng.normal(size=(n, 6))
    return X, T


def torch_model(params):
    layers = mlp.unpack(params)
    mods = []
    for li, (W, b) in enumerate(layers):
        lin = torch.nn.Linear(W.shape[1], W.shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(W))
            lin.bias.copy_(torch.from_numpy(b))
        mods.append(lin)
        if li < len(layers) - 1:
            mods.append(torch.nn.GELU(approximate="tanh"))
    return torch.nn.Sequential(*mods)


def test_param_count():
    assert mlp.n_params() == 10 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 6 + 6 == 9414


def test_bf16_matches_torch_conversion():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-8, 8, 5000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0]])
    x32 = x.astype(np.float32)
    ref = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x32), ref)
    # ties go to even
    assert mlp.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert mlp.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_gelu_and_derivative_match_torch():
    z = torch.linspace(-8, 8, 2001, dtype=torch.float64, requires_grad=True)
    g = torch.nn.functional.gelu(z, approximate="tanh")
    (d,) = torch.autograd.grad(g.sum(), z)
    zn = z.detach().numpy()
    assert np.abs(mlp.gelu(zn) - g.detach().numpy()).max() < 1e-14
    assert np.abs(mlp.gelu_d(zn) - d.numpy()).max() < 1e-14


def test_forward_and_gradient_match_torch_autograd():
    params = rand_params()
    X, T = rand_data(300)
    net = torch_model(params)
    y = net(torch.from_numpy(X))
    loss = torch.mean((y - torch.from_numpy(T)) ** 2)
    loss.backward()
    ref_grad = np.concatenate(
        [np.concatenate([m.weight.grad.numpy().ravel(), m.bias.grad.numpy()]) for m in net
         if isinstance(m, torch.nn.Linear)])
    yo, _ = mlp.forward(params, X, emulate=False)
    assert np.abs(yo - y.detach().numpy()).max() < 1e-13
    lo, go = mlp.loss_and_grad(params, X, T, emulate=False)
    assert abs(lo - loss.item()) < 1e-14
    assert np.abs(go - ref_grad).max() < 1e-14 * max(1.0, np.abs(ref_grad).max()) * 100


def test_gradient_finite_differences():
    params = rand_params(3)
    X, T = rand_data(64, seed=4)
    _, g = mlp.loss_and_grad(params, X, T, emulate=False)
    rng = np.random.default_rng(9)
    idx = rng.choice(params.size, 40, replace=False)
    idx = np.concatenate([idx, [0, 639, 640, 703, 704, params.size - 1]])  # layer edges, biases
    h = 1e-6
    for k in idx:
        p1, p2 = params.copy(), params.copy()
        p1[k] += h
        p2[k] -= h
        fd = (mlp.loss_and_grad(p1, X, T, False)[0] - mlp.loss_and_grad(p2, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_emulated_rounding_is_bf16_sized():
    params = rand_params(7)
    X, T = rand_data(500, seed=8)
    ye, _ = mlp.forward(params, X, emulate=True)
    yx, _ = mlp.forward(params, X, emulate=False)
    err = np.abs(ye - yx).max() / np.abs(yx).max()
    assert 1e-4 < err < 3e-2  # rounding present, and of bfloat16 size
    _, ge = mlp.loss_and_grad(params, X, T, emulate=True)
    _, gx = mlp.loss_and_grad(params, X, T, emulate=False)
    assert np.linalg.norm(ge - gx) < 3e-2 * np.linalg.norm(gx)


def test_adam_matches_torch():
    params = rand_params(11)
    rng = np.random.default_rng(12)
    w = torch.nn.Parameter(torch.from_numpy(params.copy()))
    opt = torch.optim.Adam([w], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = params.copy(), np.zeros_like(params), np.zeros_like(params)
    for step in range(1, 6):
        g = rng.normal(size=params.size) * 10.0 ** rng.integers(-6, 1)
        w.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = mlp.adam(p, g, m, v, step)
        assert np.abs(p - w.detach().numpy()).max() < 1e-15 * 10


# --- training samples --------------------------------------------------------
def grid(nx=32, ny=16, nz=10):
    zz = 1.0 + np.tanh(1.5 * (2.0 * np.arange(nz) / (nz - 1) - 1.0)) / np.tanh(1.5)
    return nx, ny, nz, zz, 2 * np.pi / nx, np.pi / ny


def test_filter_keeps_resolved_modes_and_removes_the_rest():
    nx, ny, nz, zz, hx, hy = grid()
    x = np.arange(nx) * hx
    y = np.arange(ny) * hy
    X, Y = np.meshgrid(x, y)
    # the span is pi, so sin(2 k Y) is mode k; cut by 4: kept |kx| < 4, |ky| < 2
    low = 0.7 + np.cos(3 * X + 0.2) * np.sin(2 * Y) + 0.4 * np.sin(X - 2 * Y)
    high = np.cos(5 * X) + np.sin(6 * Y)  # kx = 5 and ky = 3 are cut
    q = np.broadcast_to((low + high)[None], (nz, ny, nx))
    f = mlp.spectral_filter(q, 4)
    assert f.shape == (nz, ny // 4, nx // 4)
    assert np.abs(f - low[None, ::4, ::4]).max() < 1e-13


def test_filter_matches_brute_force_dft():
    rng = np.random.default_rng(2)
    nx, ny, nz = 16, 8, 3
    q = rng.normal(size=(nz, ny, nx))
    f = mlp.spectral_filter(q, 2)
    nxc, nyc = nx // 2, ny // 2
    kx = np.arange(nx)
    kx = np.where(kx > nx // 2, kx - nx, kx)
    ky = np.arange(ny)
    ky = np.where(ky > ny // 2, ky - ny, ky)
    keep = (np.abs(kx)[None, :] < nxc // 2) & (np.abs(ky)[:, None] < nyc // 2)
    ref = np.zeros((nz, nyc, nxc))
    for iz in range(nz):
        for J in range(nyc):
            for I in range(nxc):
                acc = 0.0
                for b in range(ny):
                    for a in range(nx):
                        if not keep[b, a]:
                            continue
                        Qh = np.sum(q[iz] * np.exp(-2j * np.pi * (np.outer(np.arange(ny), np.full(nx, ky[b])) / ny
                                                                   + np.outer(np.ones(ny), np.arange(nx) * kx[a]) / nx)))
                        acc += (Qh * np.exp(2j * np.pi * (ky[b] * 2 * J / ny + kx[a] * 2 * I / nx))).real
                ref[iz, J, I] = acc / (nx * ny)
    assert np.abs(f - ref).max() < 1e-12


def test_sgs_stress_closed_form():
    """u = A cos(x) + B cos(6x), v = w = 0 on a 32-point x grid cut by 4 (kept |kx| < 4):
    filt(u) = A cos x, filt(u^2) - filt(u)^2 = B^2 / 2, deviatoric XX = B^2/3, YY = ZZ = -B^2/6."""
    nx, ny, nz, zz, hx, hy = grid()
    A, B = 0.8, 0.3
    x = np.arange(nx) * hx
    u = np.zeros((nz, ny, nx, 3))
    u[..., 0] = A * np.cos(x) + B * np.cos(6 * x)
    u_tau, inv_len = 0.05, 0.05 / 1e-3
    X, T = mlp.samples(u.reshape(-1, 3), nx, ny, nz, zz, hx, hy, u_tau, inv_len, 4)
    assert X.shape == ((nx // 4) * (ny // 4) * (nz - 2), 10) and T.shape[1] == 6
    ref = np.array([B * B / 3, -B * B / 6, -B * B / 6, 0, 0, 0]) / u_tau ** 2
    assert np.abs(T - ref).max() < 1e-11 * np.abs(ref).max()
    # du/dx of the filtered field by central differences on the coarse grid, in wall units
    xc = x[::4]
    dudx = A * (np.cos(xc + 4 * hx) - np.cos(xc - 4 * hx)) / (8 * hx) / (u_tau * inv_len)
    got = X[:, 0].reshape(nz - 2, ny // 4, nx // 4)
    assert np.abs(got - dudx[None, None, :]).max() < 1e-12 * np.abs(dudx).max()
    assert np.all(X[:, [1, 2, 3, 4, 5, 6, 7, 8]] == 0)


def test_z_derivative_exact_for_quadratics():
    nz = 9
    zz = np.sort(np.random.default_rng(0).uniform(0, 2, nz))
    f = (1.5 * zz * zz - 0.3 * zz + 2.0)[:, None, None]
    d = mlp.dz_nonuniform(f, zz)[:, 0, 0]
    assert np.abs(d - (3.0 * zz[1:-1] - 0.3)).max() < 1e-12


def test_smagorinsky_comparison_model_on_pure_shear():
    """du/dy+ = s only: |S+| = s, nu_t+ = f_VD Delta+^2 s, tau_xy+ = -nu_t+ s."""
    s, yp, cell = 0.2, 30.0, 40.0
    X = np.zeros((1, 10))
    X[0, 1] = s
    X[0, 9] = yp * mlp.YPLUS_SCALE
    t = mlp.smagorinsky_tau(X, cell)
    delta = min(0.1 * cell, 0.41 * yp)
    nut = (1 - np.exp(-(0.04 * yp) ** 3)) * delta * delta * s
    assert abs(t[0, 3] + nut * s) < 1e-15
    assert np.abs(t[0, [0, 1, 2, 4, 5]]).max() < 1e-15


@pytest.mark.parametrize("factor", [2, 4])
def test_band_limited_velocity_has_no_sgs_stress(factor):
    """Products of modes below a quarter of the cut-off stay below it: tau = 0."""
    nx, ny, nz, zz, hx, hy = grid()
    x = np.arange(nx) * hx
    y = np.arange(ny) * hy
    Xg, Yg = np.meshgrid(x, y)
    u = np.zeros((nz, ny, nx, 3))
    prof = zz * (2 - zz)
    u[..., 0] = prof[:, None, None] * (1 + 0.1 * np.cos(Xg))[None]
    u[..., 1] = 0.05 * prof[:, None, None] * np.cos(Xg)[None]
    u[..., 2] = 0.02 * prof[:, None, None] * np.sin(Xg)[None]
    _, T = mlp.samples(u.reshape(-1, 3), nx, ny, nz, zz, hx, hy, 0.05, 50.0, factor)
    assert np.abs(T).max() < 1e-11


def test_standardizer_fold_is_the_same_network():
    from turbchannel_amd.closure import Standardizer

    rng = np.random.default_rng(21)
    X = rng.normal(size=(400, 10)) * rng.uniform(0.1, 5, 10) + rng.uniform(-2, 2, 10)
    T = rng.normal(size=(400, 6)) * rng.uniform(0.5, 20, 6) + rng.uniform(-5, 5, 6)
    st = Standardizer(X, T)
    p = rand_params(5).astype(np.float32).astype(np.float64)
    y_std, _ = mlp.forward(p, st.inputs(X).astype(np.float64), emulate=False)
    want = y_std * st.st + st.mt
    got, _ = mlp.forward(st.fold(p), X, emulate=False)
    assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()


def test_output_scaling_and_filter_width_ratio():
    """closure.scale_outputs multiplies the network's outputs (the last layer is
    linear) and filter_width_ratio2 is the squared ratio of the filter widths in
    wall units, Re_tau L / modes."""
    from turbchannel_amd import closure
    p = rand_params(3)
    X, _ = rand_data(64, seed=4)
    y = mlp.forward(p, X, emulate=False)[0]
    y2 = mlp.forward(closure.scale_outputs(p, 2.5), X, emulate=False)[0]
    np.testing.assert_allclose(y2, 2.5 * y, rtol=1e-6, atol=1e-7)
    assert closure.filter_width_ratio2(395.0, (96, 96), 395.0, (96, 96)) == pytest.approx(1.0)
    assert closure.filter_width_ratio2(395.0, (96, 96), 790.0, (96, 96)) == pytest.approx(4.0)
    assert closure.filter_width_ratio2(395.0, (96, 96), 395.0, (48, 96)) == pytest.approx(2.0)
    assert closure.filter_width_ratio2(395.0, (96, 96), 590.0, (64, 64)) == pytest.approx((590 / 395 * 1.5) ** 2)


def _sym6(y):
    """[n, 6] (XX YY ZZ XY XZ YZ) -> [n, 3, 3]"""
    return np.stack([np.stack([y[:, 0], y[:, 3], y[:, 4]], -1), np.stack([y[:, 3], y[:, 1], y[:, 5]], -1),
                     np.stack([y[:, 4], y[:, 5], y[:, 2]], -1)], 1)


def test_strain_projection_is_the_dissipative_eddy_viscosity_part():
    """mlp.strain_projection keeps c D with c = (y:D)/(D:D) <= 0, D the deviatoric
    strain rate of the gradient inputs: an eddy-viscosity stress -2 nu D comes back
    unchanged for nu > 0 and is dropped for nu < 0; otherwise the result is
    traceless, parallel to D, carries the same dissipation y:D, leaves a remainder
    orthogonal to D, and projecting again changes nothing. The gradients have a
    trace and off-diagonal parts, so a strain that is not made deviatoric or an
    off-diagonal entry counted once fails."""
    rng = np.random.default_rng(12)
    X, _ = rand_data(400, seed=13)
    G = X[:, :9].reshape(-1, 3, 3)
    S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    D = S - (np.trace(S, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    d6 = np.stack([D[:, 0, 0], D[:, 1, 1], D[:, 2, 2], D[:, 0, 1], D[:, 0, 2], D[:, 1, 2]], -1)
    nu = rng.uniform(0.1, 2.0, 400)
    np.testing.assert_allclose(mlp.strain_projection(-2 * nu[:, None] * d6, X), -2 * nu[:, None] * d6, rtol=1e-12,
                               atol=1e-13)
    assert np.all(mlp.strain_projection(2 * nu[:, None] * d6, X) == 0)
    y = rng.normal(size=(400, 6))
    p = mlp.strain_projection(y, X)
    Y, P = _sym6(y), _sym6(p)
    yd = np.einsum("nij,nij->n", Y, D)
    dis = yd < 0
    assert 0.2 < dis.mean() < 0.8
    assert np.all(p[~dis] == 0)
    np.testing.assert_allclose(np.trace(P, axis1=1, axis2=2), 0, atol=1e-12)
    c = np.einsum("nij,nij->n", P, D) / np.einsum("nij,nij->n", D, D)
    np.testing.assert_allclose(P[dis], c[dis, None, None] * D[dis], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(np.einsum("nij,nij->n", P, D)[dis], yd[dis], rtol=1e-10)
    np.testing.assert_allclose(np.einsum("nij,nij->n", Y - P, D)[dis], 0, atol=1e-10)
    np.testing.assert_allclose(mlp.strain_projection(p, X), p, rtol=1e-12, atol=1e-13)


def test_training_target_projection_matches_oracle():
    """closure.strain_projection (the training targets of the eddy-viscosity
    network, chunked, float32 out) equals the oracle's projection."""
    from turbchannel_amd import closure
    X, T = rand_data(1000, seed=21)
    got = closure.strain_projection(T.astype(np.float32), X.astype(np.float32), chunk=300)
    ref = mlp.strain_projection(T.astype(np.float32).astype(np.float64), X.astype(np.float32).astype(np.float64))
    assert got.dtype == np.float32
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)
    assert 0.2 < np.mean(np.all(got == 0, axis=1)) < 0.8
