"""Problem set-ups shared by the tests: the reference's own set-up
(staggered/main.cpp:200-241, collocated_AB3/main.cpp:186-267) and the
bulk-unit channel cases with a tanh wall-normal mesh."""
import math

import numpy as np

PI = math.pi


def sinh_mesh(nz, stretch=6.5, H=1.0):
    """Faces of the reference's sinh-stretched mesh (staggered/main.cpp:219-223)."""
    zz = np.zeros(nz + 1)
    for i in range(nz // 2 + 1):
        zz[i] = H * math.sinh(stretch * i / nz) / math.sinh(0.5 * stretch)
        zz[nz - i] = 2.0 * H - zz[i]
    return zz


def uniform_mesh(nz, Hz=2.0):
    return np.array([i * (Hz / nz) for i in range(nz + 1)])


def tanh_faces(n, gamma):
    """z_k = 1 + tanh(gamma (2k/n - 1)) / tanh(gamma), k = 0..n."""
    k = np.arange(n + 1)
    zz = 1.0 + np.tanh(gamma * (2.0 * k / n - 1.0)) / math.tanh(gamma)
    zz[0], zz[-1] = 0.0, 2.0
    return zz


def ref_staggered_kwargs(nx, ny, Re_tau=395.0):
    """Physical parameters of staggered/main.cpp:201-241."""
    visc = 1.0 / Re_tau
    u_tau = Re_tau * visc
    Hx, Hy = 4.0 * PI, 2.0 * PI
    return dict(visc=visc, sour_dpdx=u_tau * u_tau, inv_length_scale=u_tau / visc, hx=Hx / nx, hy=Hy / ny,
                xmin=0.0, ymin=-0.5 * Hy)


def bulk_kwargs(nx, ny, Re_b, Re_tau, Lx=2 * PI, Ly=PI):
    """Bulk units (H = 1, U_b = 1, nu = 2/Re_b), box Lx x 2 x Ly."""
    visc = 2.0 / Re_b
    u_tau = Re_tau * visc
    return dict(visc=visc, sour_dpdx=u_tau * u_tau, inv_length_scale=u_tau / visc, hx=Lx / nx, hy=Ly / ny)


def coords(nx, ny, nz, hx, hy, zz, ivar):
    """GetCoor (staggered/main.cpp:58-66) for all cells; returns (N3D, 3)."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = (ix + 0.5) * hx
    y = (iy + 0.5) * hy
    z = 0.5 * (zz[iz] + zz[iz + 1])
    if ivar == 0:
        x = ix * hx
    if ivar == 1:
        y = iy * hy
    if ivar == 2:
        z = zz[iz]
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def node_coords(nx, ny, nz, hx, hy, zz):
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([(ix * hx).ravel(), (iy * hy).ravel(), zz[iz].ravel()], axis=1)


def sin_sq(n, offset=0.0):
    """sin(in*in + offset) over cell indices, as the reference's pseudo-random fields."""
    i = np.arange(n, dtype=np.float64)
    return np.sin(i * i + offset)


def ab3_mesh_tanh(nz, gamma):
    """Collocated nodes z_k, k = 0..nz-1, with the walls at 0 and 2."""
    return tanh_faces(nz - 1, gamma)
