"""Host side of the learned subgrid closure: data standardisation, folding the
standardisation into the network's first and last layers, and the training loop
over device-resident samples (the arithmetic runs in the HIP kernels behind
``tc_mlp_*``; this module only orders the calls).

The closure has no counterpart in the reference; it replaces the role of
RecalcTurbVisc (staggered/func_global.h:815-890, collocated_AB3/func_global.h:342-384)
in a collocated LES when an engine is created with ``sgs_model=1``.
"""
import numpy as np

from . import Mlp, mlp_param_count

N_IN, N_HID, N_OUT = 10, 64, 6
_OFF_W1 = 0
_OFF_B1 = _OFF_W1 + N_HID * N_IN
_OFF_W4 = _OFF_B1 + N_HID + 2 * (N_HID * N_HID + N_HID)
_OFF_B4 = _OFF_W4 + N_OUT * N_HID


class Standardizer:
    """Per-column mean and scale of inputs and targets."""

    def __init__(self, X, T, eps=1e-12):
        X = np.asarray(X, dtype=np.float64)
        T = np.asarray(T, dtype=np.float64)
        self.mx, self.sx = X.mean(axis=0), np.maximum(X.std(axis=0), eps)
        self.mt, self.st = T.mean(axis=0), np.maximum(T.std(axis=0), eps)

    def inputs(self, X):
        return ((np.asarray(X, dtype=np.float64) - self.mx) / self.sx).astype(np.float32)

    def targets(self, T):
        return ((np.asarray(T, dtype=np.float64) - self.mt) / self.st).astype(np.float32)

    def fold(self, params):
        """Parameters of the network on standardised data -> the same network on
        raw inputs and outputs: W1 / sx, b1 - W1 (mx / sx); st W4, st b4 + mt."""
        p = np.asarray(params, dtype=np.float64).copy()
        assert p.size == mlp_param_count()
        W1 = p[_OFF_W1:_OFF_B1].reshape(N_HID, N_IN)
        b1 = p[_OFF_B1:_OFF_B1 + N_HID]
        p[_OFF_B1:_OFF_B1 + N_HID] = b1 - W1 @ (self.mx / self.sx)
        p[_OFF_W1:_OFF_B1] = (W1 / self.sx[None, :]).ravel()
        W4 = p[_OFF_W4:_OFF_B4].reshape(N_OUT, N_HID)
        b4 = p[_OFF_B4:_OFF_B4 + N_OUT]
        p[_OFF_W4:_OFF_B4] = (W4 * self.st[:, None]).ravel()
        p[_OFF_B4:_OFF_B4 + N_OUT] = b4 * self.st + self.mt
        return p.astype(np.float32)


def filter_width_ratio2(re_tau_train, modes_train, re_tau_les, modes_les):
    """(Delta+_les / Delta+_train)^2 for the sharp cut-offs of the training data
    and of the LES grid on the same box: modes_* are the cut-off mode counts in x
    and y, the filter width in wall units Re_tau L / modes (geometric mean of x
    and y)."""
    r = [(re_tau_les / re_tau_train) * (mt / ml) for mt, ml in zip(modes_train, modes_les)]
    return float(np.prod(r)) ** (2.0 / len(r))


def scale_outputs(params, s):
    """The network with its outputs multiplied by s (last layer's weights and biases)."""
    p = np.asarray(params, dtype=np.float64).copy()
    assert p.size == mlp_param_count()
    p[_OFF_W4:_OFF_B4 + N_OUT] *= s
    return p.astype(np.float32)


def strain_projection(T, X, chunk=1 << 22):
    """The dissipative part of each stress T[n] (XX YY ZZ XY XZ YZ) along the
    deviatoric strain rate D of the sample's gradient inputs X[n, :9]
    (g_ij = X[3 i + j], any positive scale): (T:D / D:D) D where T:D < 0, else
    0 -- the stress of an eddy viscosity -T:D / (2 D:D) >= 0, the part an LES
    engine with sgs_model 3 keeps of the network's output. float32 [n, 6]."""
    T = np.asarray(T)
    X = np.asarray(X)
    out = np.empty((T.shape[0], 6), dtype=np.float32)
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])  # off-diagonal entries appear twice in a:b
    for a in range(0, T.shape[0], chunk):
        g = X[a:a + chunk, :9].astype(np.float64)
        tr = (g[:, 0] + g[:, 4] + g[:, 8]) / 3.0
        d = np.stack([g[:, 0] - tr, g[:, 4] - tr, g[:, 8] - tr, 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]),
                      0.5 * (g[:, 5] + g[:, 7])], axis=-1)
        td = np.sum(w * T[a:a + chunk].astype(np.float64) * d, axis=1)
        dd = np.sum(w * d * d, axis=1)
        c = np.where((td < 0) & (dd > 0), td / np.where(dd > 0, dd, 1.0), 0.0)
        out[a:a + chunk] = c[:, None] * d
    return out


def train(net, dev_x, dev_t, n, *, batch=65536, iters=1000, lr=1e-3, lr_final=None):
    """Adam over contiguous mini-batches of device samples (addresses of float32
    [n][10] inputs and [n][6] targets, already shuffled); returns the losses.
    The learning rate decays geometrically to lr_final when given."""
    batch = min(batch, n)
    nb = max(1, n // batch)
    losses = []
    for it in range(iters):
        off = (it % nb) * batch
        a = lr if lr_final is None else lr * (lr_final / lr) ** (it / max(1, iters - 1))
        losses.append(net.train_step(dev_x + 4 * N_IN * off, dev_t + 4 * N_OUT * off, batch, lr=a))
    return np.array(losses)


__all__ = ["Standardizer", "train", "Mlp", "filter_width_ratio2", "scale_outputs"]
