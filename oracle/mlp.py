"""CPU restatement of the learned subgrid closure and its training data (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP closure kernels. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.

The reference has no learned closure; its subgrid model is the Smagorinsky-van
Driest eddy viscosity (staggered/func_global.h:815-890, collocated_AB3/func_global.h:342-384),
restated in tc_oracle.c. What is restated here is the closure's own definition:

* Network: 10 -> 64 -> 64 -> 64 -> 6, GELU (tanh form) on the hidden layers,
  linear output. Flat FP32 parameter vector: for each layer W (out x in,
  row-major) then b (out).
* Rounding model of the GPU kernels (emulate=True): inputs, hidden activations,
  GELU derivatives and back-propagated gradients are rounded to bfloat16
  (round to nearest even, from FP32); weights are used as bfloat16 copies of the
  FP32 master weights; every product sums in FP32 (here in FP64, the difference
  being the checked tolerance); biases and the output stay FP32.
* Loss: mean over points and the 6 outputs of (y - t)^2.
* Adam on the FP32 master weights (same update as torch.optim.Adam).
* Training samples: a sharp spectral cut-off by `factor` in x and y of a
  collocated velocity field, sampled on the coarse grid; inputs are the 9
  velocity-gradient components in wall units (x, y by central differences on
  the coarse grid, z by the 3-point nonuniform formula) and y+/100; targets are
  the deviatoric tau_ij = filt(u_i u_j) - filt(u_i) filt(u_j) in wall units,
  ordered XX, YY, ZZ, XY, XZ, YZ (the order of the reference's uu[6],
  staggered/real3.h:42).

Pins: backward against torch.autograd and central finite differences
(emulate=False), Adam against torch.optim.Adam, the filter against band-limited
fields and single modes (tests/test_closure_oracle.py).
"""
import numpy as np

LAYERS = [(10, 64), (64, 64), (64, 64), (64, 6)]  # (in, out)
N_IN, N_OUT = 10, 6
YPLUS_SCALE = 0.01


def n_params():
    return sum(o * i + o for i, o in LAYERS)


def unpack(params):
    """Flat vector -> [(W, b), ...]."""
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for i, o in LAYERS:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    assert off == p.size
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def bf16(x):
    """Round to bfloat16 (nearest even) through FP32, returned as float64."""
    x32 = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


_C = np.sqrt(2.0 / np.pi)
_A = 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_C * (z + _A * z ** 3)))


def gelu_d(z):
    t = np.tanh(_C * (z + _A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _C * (1.0 + 3.0 * _A * z * z)


def forward(params, X, emulate=True):
    """y (P x 6) and the cache for backward."""
    rd = bf16 if emulate else (lambda v: v)
    L = unpack(params)
    a = rd(np.asarray(X, dtype=np.float64))
    acts, ders = [a], []
    for li, (W, b) in enumerate(L):
        Wq = rd(f32(W) if emulate else W)
        z = a @ Wq.T + (f32(b) if emulate else b)
        if li == len(L) - 1:
            return (f32(z) if emulate else z), (acts, ders)
        if emulate:
            z = f32(z)
        a = rd(gelu(z))
        ders.append(rd(gelu_d(z)))
        acts.append(a)


def loss_and_grad(params, X, T, emulate=True):
    """Mean-squared loss and its gradient (flat, like params)."""
    rd = bf16 if emulate else (lambda v: v)
    y, (acts, ders) = forward(params, X, emulate)
    P = y.shape[0]
    r = y - np.asarray(T, dtype=np.float64)
    loss = float(np.sum(r * r) / (P * N_OUT))
    L = unpack(params)
    g = rd(2.0 * r / (P * N_OUT))
    grads = [None] * len(L)
    for li in range(len(L) - 1, -1, -1):
        W, _ = L[li]
        dW = g.T @ acts[li]
        db = g.sum(axis=0)
        grads[li] = (dW, db)
        if li > 0:
            Wq = rd(f32(W) if emulate else W)
            ga = g @ Wq
            g = rd(ga * ders[li - 1])
    return loss, pack(grads)


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (params, m, v)."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------------------
# Training samples from a collocated field (nodes incl. walls)
# ---------------------------------------------------------------------------
def spectral_filter(q, factor):
    """Sharp cut-off by `factor` in x and y of q[nz, ny, nx], sampled on the
    coarse grid (every factor-th point): keeps |kx| < nxc/2 and |ky| < nyc/2."""
    nz, ny, nx = q.shape
    nxc, nyc = nx // factor, ny // factor
    Q = np.fft.rfft2(q, axes=(1, 2))
    C = np.zeros((nz, nyc, nxc // 2 + 1), dtype=complex)
    kxk = nxc // 2  # kx = 0 .. nxc/2 - 1 kept; nxc/2 zero
    hy = nyc // 2
    C[:, :hy, :kxk] = Q[:, :hy, :kxk]
    C[:, nyc - hy + 1:, :kxk] = Q[:, ny - hy + 1:, :kxk]
    return np.fft.irfft2(C, s=(nyc, nxc), axes=(1, 2)) * (nxc * nyc) / (nx * ny)


def dz_nonuniform(f, zz):
    """3-point derivative along axis 0 at interior nodes 1..nz-2."""
    hm = (zz[1:-1] - zz[:-2])[:, None, None]
    hp = (zz[2:] - zz[1:-1])[:, None, None]
    return (hm * hm * (f[2:] - f[1:-1]) + hp * hp * (f[1:-1] - f[:-2])) / (hp * hm * (hp + hm))


def gradients(U, zz, hx, hy):
    """g[i][j] = d u_i / d x_j at interior nodes; U: 3 arrays [nz, ny, nx]."""
    g = []
    for u in U:
        ux = (np.roll(u, -1, axis=2) - np.roll(u, 1, axis=2)) / (2.0 * hx)
        uy = (np.roll(u, -1, axis=1) - np.roll(u, 1, axis=1)) / (2.0 * hy)
        g.append([ux[1:-1], uy[1:-1], dz_nonuniform(u, zz)])
    return g


def features(U, zz, hx, hy, u_tau, inv_len):
    """Network inputs [nz-2, ny, nx, 10] at interior nodes."""
    g = gradients(U, zz, hx, hy)
    s = 1.0 / (u_tau * inv_len)  # nu / u_tau^2
    nz = len(zz)
    d = np.minimum(zz - zz[0], zz[-1] - zz)[1:-1]
    shape = g[0][0].shape
    F = np.empty(shape + (10,))
    for i in range(3):
        for j in range(3):
            F[..., 3 * i + j] = g[i][j] * s
    F[..., 9] = np.broadcast_to((d * inv_len * YPLUS_SCALE)[:, None, None], shape)
    assert F.shape[0] == nz - 2
    return F


PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def samples(u, nx, ny, nz, zz, hx, hy, u_tau, inv_len, factor):
    """(X [S, 10], T [S, 6]) at the coarse interior nodes, S = nxc*nyc*(nz-2),
    ordered (iz, iy, ix) with ix fastest."""
    u = np.asarray(u, dtype=np.float64).reshape(nz, ny, nx, 3)
    U = [u[..., c] for c in range(3)]
    Uf = [spectral_filter(q, factor) for q in U]
    tau = []
    for a, b in PAIRS:
        tau.append(spectral_filter(U[a] * U[b], factor) - Uf[a] * Uf[b])
    tr = (tau[0] + tau[1] + tau[2]) / 3.0
    for c in range(3):
        tau[c] = tau[c] - tr
    X = features(Uf, zz, hx * factor, hy * factor, u_tau, inv_len)
    T = np.stack([t[1:-1] / (u_tau * u_tau) for t in tau], axis=-1)
    return X.reshape(-1, 10), T.reshape(-1, 6)


def stress_strain(y, X):
    """tau_ij S_ij per sample from the outputs (XX YY ZZ XY XZ YZ) and the gradient
    inputs g_ij = X[3 i + j], S_ij = (g_ij + g_ji) / 2."""
    return (y[:, 0] * X[:, 0] + y[:, 1] * X[:, 4] + y[:, 2] * X[:, 8] + y[:, 3] * (X[:, 1] + X[:, 3]) +
            y[:, 4] * (X[:, 2] + X[:, 6]) + y[:, 5] * (X[:, 5] + X[:, 7]))


def strain_projection(y, X):
    """The dissipative part of the stress y (XX YY ZZ XY XZ YZ) along the
    deviatoric strain rate D of the gradient inputs: (y:D / D:D) D where
    y:D < 0, else 0, as six components."""
    G = X[:, :9].reshape(-1, 3, 3)
    S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    D = S - (np.trace(S, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    d6 = np.stack([D[:, 0, 0], D[:, 1, 1], D[:, 2, 2], D[:, 0, 1], D[:, 0, 2], D[:, 1, 2]], axis=-1)
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])  # off-diagonal entries appear twice in a:b
    yd = np.sum(w * y * d6, axis=1)
    dd = np.sum(w * d6 * d6, axis=1)
    c = np.where((yd < 0) & (dd > 0), yd / np.where(dd > 0, dd, 1.0), 0.0)
    return c[:, None] * d6


def node_tau(params, u, nx, ny, nz, zz, hx, hy, u_tau, inv_len, emulate=True, clip=False, project=False):
    """Closure stress at every node [N3D, 6] (zero on the wall planes), as the
    LES step computes it from its own velocity field; clip: zero where
    tau_ij S_ij > 0 (backscatter); project: the dissipative part along the
    deviatoric strain rate (strain_projection)."""
    u = np.asarray(u, dtype=np.float64).reshape(nz, ny, nx, 3)
    X = features([u[..., c] for c in range(3)], zz, hx, hy, u_tau, inv_len).reshape(-1, 10)
    y, _ = forward(params, X, emulate)
    if clip:
        y = np.where((stress_strain(y, X) > 0)[:, None], 0.0, y)
    if project:
        y = strain_projection(y, X)
    tau = np.zeros((nz, ny, nx, 6))
    tau[1:-1] = (y * u_tau * u_tau).reshape(nz - 2, ny, nx, 6)
    return tau.reshape(-1, 6)


def smagorinsky_tau(X, cell_plus, cs=0.1):
    """Deviatoric Smagorinsky-van Driest stress in wall units from the same
    inputs, as the reference's model defines nu_t (staggered/func_global.h:869-889):
    nu_t+ = f_VD Delta+^2 |S+|, f_VD = 1 - exp(-(0.04 y+)^3),
    Delta+ = min(C_s cell_plus, 0.41 y+) with cell_plus = (cell volume)^(1/3) in
    wall units; tau+ = -2 nu_t+ S+ (a comparison model for the closure)."""
    G = X[:, :9].reshape(-1, 3, 3)
    S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    absS = np.sqrt(2.0 * np.sum(S * S, axis=(1, 2)))
    yp = X[:, 9] / YPLUS_SCALE
    fvd = 1.0 - np.exp(-((0.04 * yp) ** 3))
    delta = np.minimum(cs * np.asarray(cell_plus, dtype=np.float64), 0.41 * yp)
    nut = fvd * delta * delta * absS
    tau = -2.0 * nut[:, None, None] * S
    tr = np.trace(tau, axis1=1, axis2=2) / 3.0
    for c in range(3):
        tau[:, c, c] -= tr
    return np.stack([tau[:, a, b] for a, b in PAIRS], axis=1)
