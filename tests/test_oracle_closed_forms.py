"""Closed-form pins of the oracle functions that the reference's self-checks
leave open: the NAIVE stress divergence with a varying turbulent viscosity
(including the averaging of staggered/func_global.h:598), the viscous wall rows,
the Galerkin viscous term with element viscosity
(collocated_AB3/func_fluxes.h:21-85, the weighting at :78-80), the collocated
integral, kinetic energy, profiles and CFL (collocated_AB3/func_global.h:173-289),
the viscous CFL modes of the staggered solver (staggered/func_global.h:692-704)
and the collocated divergence of a caller-given stress tensor.

Each pin is a function pin_*() that raises AssertionError when the oracle
disagrees with the mathematics; the test_* wrappers run them.
tests/test_oracle_mutants.py compiles copies of the oracle with one planted
mistake each and checks that the pin meant to catch it fails.
"""
import math

import numpy as np

import oracle
from oracle import SCHEME_AB3, SCHEME_HW, SCHEME_VV, KIND_NUT, KIND_TAU
from tests import cases


def _stag(scheme, nx, ny, nz, gamma=1.7, **kw):
    zz = cases.tanh_faces(nz, gamma)
    args = dict(visc=0.013, sour_dpdx=0.0, inv_length_scale=90.0, hx=0.37, hy=0.29)
    args.update(kw)
    return oracle.Solver(scheme, nx, ny, nz, zz, **args), zz, args


def _ab3(nx, ny, nz, gamma=1.6, **kw):
    zz = cases.ab3_mesh_tanh(nz, gamma)
    args = dict(visc=0.011, sour_dpdx=0.0, inv_length_scale=100.0, hx=0.31, hy=0.23)
    args.update(kw)
    return oracle.Solver(SCHEME_AB3, nx, ny, nz, zz, **args), zz, args


def _grid(nx, ny, nz):
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ix.ravel(), iy.ravel(), iz.ravel()


# ---------------------------------------------------------------------------
# Staggered NAIVE stress divergence with a linear turbulent viscosity.
# ---------------------------------------------------------------------------
def pin_naive_linear_viscosity():
    """u_i = A_ij x_j sampled at each component's staggered position, total
    viscosity mu = visc + nu_t with nu_t = n0 + b . x at the cell centres. Every
    difference quotient of a linear field is exact, so at interior cells (one
    cell from the periodic seams and the walls) the stress divergence
    d_j (mu (d_j u_i + d_i u_j)) = (A + A^T)_ij b_j holds term by term, up to two
    geometric facts of the staggered averaging:
    * the viscosity at an x-z or y-z edge is the mean of the four cell centres
      around it, i.e. mu at the mid-point of the two cell centres in z, so the
      z term of the u and v equations carries (zc[k+1] - zc[k-1]) / (2 DZ[k]);
    * sxy_190 averages mu(XR(YR(in))) (staggered/func_global.h:598) where the
      edge's fourth cell is XR(YL(in)): its average sits at (x + h_x/2, y), half a
      cell up in y, so the v equation's x term is (A01 + A10)(b_x + b_y h_y / (2 h_x)).
    """
    nx, ny, nz = 12, 10, 14
    s, zz, a = _stag(SCHEME_VV, nx, ny, nz, turb_visc=1, visc_type=1)
    hx, hy = a["hx"], a["hy"]
    rng = np.random.default_rng(17)
    A = rng.uniform(-1, 1, (3, 3))
    b = np.array([0.02, 0.035, -0.015])
    n0 = 0.05
    u = np.zeros((s.n3d, 3))
    for c in range(3):
        r = cases.coords(nx, ny, nz, hx, hy, zz, c)
        u[:, c] = r @ A[c]
    rc = cases.coords(nx, ny, nz, hx, hy, zz, -1)  # cell centres
    s.set_field(KIND_NUT, 0, n0 + rc @ b)
    f = s.visc_flux_naive(u)
    ix, iy, iz = _grid(nx, ny, nz)
    m = (ix >= 1) & (ix <= nx - 2) & (iy >= 1) & (iy <= ny - 2) & (iz >= 1) & (iz <= nz - 2)
    zc = 0.5 * (zz[:-1] + zz[1:])
    dz = np.diff(zz)
    k = np.clip(iz, 1, nz - 2)
    rz = (zc[k + 1] - zc[k - 1]) / (2.0 * dz[k])
    S2 = A + A.T
    fex = np.zeros((s.n3d, 3))
    fex[:, 0] = S2[0, 0] * b[0] + S2[0, 1] * b[1] + S2[0, 2] * b[2] * rz
    fex[:, 1] = S2[0, 1] * (b[0] + b[1] * hy / (2.0 * hx)) + S2[1, 1] * b[1] + S2[1, 2] * b[2] * rz
    fex[:, 2] = S2[0, 2] * b[0] + S2[1, 2] * b[1] + S2[2, 2] * b[2]
    err = np.abs(f - fex)[m].max() / np.abs(fex[m]).max()
    assert err < 1e-11, err


# ---------------------------------------------------------------------------
# Viscous wall rows: the volume sum of a conservative viscous term telescopes
# to the wall fluxes.
# ---------------------------------------------------------------------------
def _periodic_field(nx, ny, nz, seed):
    """Random field periodic in x and y: a random mean profile per component (so
    the wall fluxes do not vanish) plus plane waves; w = 0 on the bottom wall face."""
    rng = np.random.default_rng(seed)
    ix, iy, iz = _grid(nx, ny, nz)
    u = np.zeros((nx * ny * nz, 3))
    for c in range(3):
        u[:, c] = rng.uniform(0.5, 1.5, nz)[iz]
        for _ in range(3):
            kx, ky = rng.integers(0, 3, 2)
            ph = rng.uniform(0, 2 * math.pi, 2)
            amp = rng.uniform(0.5, 1.5, nz)
            u[:, c] += amp[iz] * np.cos(2 * math.pi * kx * ix / nx + ph[0]) * np.cos(2 * math.pi * ky * iy / ny + ph[1])
    u[iz == 0, 2] = 0.0
    return u


def pin_staggered_viscous_wall_rows():
    """Sum over cells of volume x viscous term, with the volumes of the
    staggered unknowns (DZ[k] hx hy for u and v, (DZ[k] + DZ[k-1])/2 hx hy for w at
    faces k >= 1): x and y differences cancel over the periodic directions and
    the z fluxes telescope, leaving the wall terms. With the turbulent viscosity
    varying in z only:
    * u, v (CROSS and NAIVE): -2 visc hx hy sum_xy (u_top / DZ[nz-1] + u_bottom / DZ[0])
      with the molecular viscosity only, as the wall stress of
      staggered/func_global.h:592-595,604-607 is written;
    * w, NAIVE: -2 hx hy sum_xy (mu_top w[nz-1] / DZ[nz-1] + mu_bottom w[1] / DZ[0]) with
      the total viscosity of the wall cells (szz uses mu);
    * w, CROSS: -visc hx hy sum_xy (w[nz-1] / DZ[nz-1] + w[1] / DZ[0]).
    """
    nx, ny, nz = 8, 6, 12
    for visc_type in (0, 1):
        for nut_on in ((0, 1) if visc_type else (0,)):
            s, zz, a = _stag(SCHEME_HW, nx, ny, nz, turb_visc=nut_on, visc_type=visc_type)
            hx, hy, visc = a["hx"], a["hy"], a["visc"]
            u = _periodic_field(nx, ny, nz, 5 + visc_type + 2 * nut_on)
            ix, iy, iz = _grid(nx, ny, nz)
            nut_plane = 0.02 + 0.01 * np.arange(nz) ** 1.5
            if nut_on:
                s.set_field(KIND_NUT, 0, nut_plane[iz])
            mu = visc + (nut_plane if nut_on else np.zeros(nz))
            f = s.flux_term(u, 0, 1)
            dz = np.diff(zz)
            vol = dz[iz] * hx * hy
            volw = np.where(iz > 0, 0.5 * (dz[iz] + dz[np.maximum(iz - 1, 0)]), 0.0) * hx * hy
            top, bot, one = iz == nz - 1, iz == 0, iz == 1
            for c in (0, 1):
                got = np.sum(f[:, c] * vol)
                exp = -2 * visc * hx * hy * (np.sum(u[top, c]) / dz[-1] + np.sum(u[bot, c]) / dz[0])
                assert abs(got - exp) < 1e-11 * np.abs(f[:, c] * vol).sum(), (visc_type, nut_on, c, got, exp)
            got = np.sum(f[:, 2] * volw)
            if visc_type:
                exp = -2 * hx * hy * (mu[-1] * np.sum(u[top, 2]) / dz[-1] + mu[0] * np.sum(u[one, 2]) / dz[0])
            else:
                exp = -visc * hx * hy * (np.sum(u[top, 2]) / dz[-1] + np.sum(u[one, 2]) / dz[0])
            assert abs(got - exp) < 1e-11 * np.abs(f[:, 2] * volw).sum(), (visc_type, nut_on, 2, got, exp)


# ---------------------------------------------------------------------------
# Collocated Galerkin viscous term with element viscosity.
# ---------------------------------------------------------------------------
def _element_nut(nx, ny, nz, zz, hx, hy, n0, b):
    """nu_t on elements (stored at their lower corner node; nz - 1 planes) from a
    linear function of the element centre; the last node plane holds zeros."""
    ix, iy, iz = _grid(nx, ny, nz)
    zc = 0.5 * (zz[np.minimum(iz, nz - 2)] + zz[np.minimum(iz + 1, nz - 1)])
    v = n0 + b[0] * (ix + 0.5) * hx + b[1] * (iy + 0.5) * hy + b[2] * zc
    return np.where(iz < nz - 1, v, 0.0)


def pin_galerkin_element_viscosity():
    """Q1 Galerkin stress form, mu_e = visc + nu_t of each element.
    (1) Linear velocity, nu_t linear over element centres: on each element the
        stress 2 mu_e S is constant, and summed over the 8 elements of a node the
        weak form gives mu's jump across the node times the face areas, so the
        term equals d_j (2 mu S_ij) = (A + A^T)_ij b_j exactly at interior nodes.
    (2) Random element nu_t, random fields vanishing on the walls: the operator
        times the node volume is the gradient of the symmetric bilinear form
        -sum_e mu_e int 2 S(u):S(v), so <v, vol L u> = <u, vol L v> and <u, vol L u> <= 0.
    """
    nx, ny, nz = 10, 8, 12
    s, zz, a = _ab3(nx, ny, nz, turb_visc=1, visc_type=1)
    hx, hy = a["hx"], a["hy"]
    X = cases.node_coords(nx, ny, nz, hx, hy, zz)
    rng = np.random.default_rng(23)
    A = rng.uniform(-1, 1, (3, 3))
    b = np.array([0.03, -0.02, 0.025])
    s.set_field(KIND_NUT, 0, _element_nut(nx, ny, nz, zz, hx, hy, 0.04, b))
    u = X @ A.T
    f = s.flux_term(u, 0, 1)
    ix, iy, iz = _grid(nx, ny, nz)
    m = (ix >= 1) & (ix <= nx - 2) & (iy >= 1) & (iy <= ny - 2) & (iz >= 1) & (iz <= nz - 2)
    fex = np.tile((A + A.T) @ b, (s.n3d, 1))
    err = np.abs(f - fex)[m].max() / np.abs(fex[m]).max()
    assert err < 1e-10, err
    # (2) symmetry with random element viscosity
    nut = np.where(iz < nz - 1, rng.uniform(0.0, 0.1, s.n3d), 0.0)
    s.set_field(KIND_NUT, 0, nut)
    wall = (iz == 0) | (iz == nz - 1)
    uu = rng.standard_normal((s.n3d, 3))
    vv = rng.standard_normal((s.n3d, 3))
    uu[wall] = 0.0
    vv[wall] = 0.0
    dz = np.diff(zz)
    h = np.where(iz == 0, 0.5 * dz[0], np.where(iz == nz - 1, 0.5 * dz[-1],
                                                 0.5 * (dz[np.maximum(iz - 1, 0)] + dz[np.minimum(iz, nz - 2)])))
    vol = (h * hx * hy)[:, None]
    Lu, Lv = s.flux_term(uu, 0, 1), s.flux_term(vv, 0, 1)
    a_uv, a_vu = np.sum(vv * Lu * vol), np.sum(uu * Lv * vol)
    scale = np.sqrt(np.sum(uu * Lu * vol) * np.sum(vv * Lv * vol))
    assert abs(a_uv - a_vu) < 1e-12 * scale, (a_uv, a_vu)
    assert np.sum(uu * Lu * vol) < 0 and np.sum(vv * Lv * vol) < 0


# ---------------------------------------------------------------------------
# Collocated reductions.
# ---------------------------------------------------------------------------
def pin_collocated_integrals():
    """CalcIntegral / CalcKineticEnergy weight each node by its cell height
    (half an interval at the walls), which is the trapezoidal rule in z: exact
    for fields linear in z, so uniform and linear-in-z fields integrate in
    closed form over the box Lx x Ly x 2."""
    nx, ny, nz = 10, 8, 13
    s, zz, a = _ab3(nx, ny, nz)
    box = a["hx"] * nx * a["hy"] * ny
    c = np.array([1.3, -0.7, 0.4])
    u = np.tile(c, (s.n3d, 1))
    np.testing.assert_allclose(s.integral(u), c * box * 2.0, rtol=1e-13)
    np.testing.assert_allclose(s.kinetic_energy(u), 0.5 * (c @ c) * box * 2.0, rtol=1e-13)
    X = cases.node_coords(nx, ny, nz, a["hx"], a["hy"], zz)
    lin = np.stack([X[:, 2], 2.0 - X[:, 2], 3.0 * X[:, 2] - 1.0], axis=1)
    # int_0^2 z dz = 2, int (2 - z) = 2, int (3z - 1) = 4
    np.testing.assert_allclose(s.integral(lin), np.array([2.0, 2.0, 4.0]) * box, rtol=1e-13)


def pin_collocated_profiles():
    """CalcProfiles (collocated): plane means of u (3), uu (XX YY ZZ XY XZ YZ) and
    nu_t. On plane-uniform fields the means are the values and products
    themselves; a cosine of a non-zero x wavenumber adds zero mean and half its
    square to the second moments."""
    nx, ny, nz = 12, 8, 9
    s, zz, a = _ab3(nx, ny, nz, turb_visc=1, visc_type=1)
    ix, iy, iz = _grid(nx, ny, nz)
    fa, fb, fc = 1.0 + iz, 2.0 - 0.3 * iz, 0.5 + 0.1 * iz ** 2
    amp = 0.7
    wave = amp * np.cos(2 * math.pi * 2 * ix / nx + 0.3)
    u = np.stack([fa + wave, fb, fc], axis=1)
    nut = 0.01 * (1 + iz)
    s.set_field(KIND_NUT, 0, nut)
    av = s.profiles(u, np.zeros(s.n3d))
    P = lambda x: x.reshape(nz, -1)[:, 0]  # noqa: E731  (plane value of a plane-uniform array)
    ea, eb, ec = P(fa), P(fb), P(fc)
    exp = np.stack([ea, eb, ec, ea ** 2 + amp ** 2 / 2, eb ** 2, ec ** 2, ea * eb, ea * ec, eb * ec,
                    P(nut)], axis=1)
    np.testing.assert_allclose(av, exp, rtol=1e-13, atol=1e-13)


def pin_collocated_cfl():
    """CalcCFL (collocated_AB3/func_global.h:188-205), per node
    conv = tau (|u|/hx + |v|/hy + |w|/h_z), visc = 4 (G + 1) tau (visc + nu_node)
    (1/hx^2 + 1/hy^2 + 1/h_z^2), G = 1 for the Galerkin term, h_z the node's cell
    height (half an interval at the walls), nu_node the mean of the 8 elements
    around an interior node (0 on the walls); mode 1, 2, 3 = conv, visc, both,
    maximised over nodes."""
    nx, ny, nz = 8, 8, 11
    tau = 1e-3
    for galerkin in (0, 1):
        s, zz, a = _ab3(nx, ny, nz, turb_visc=1, visc_type=galerkin)
        hx, hy, visc = a["hx"], a["hy"], a["visc"]
        dz = np.diff(zz)
        hnode = np.concatenate([[0.5 * dz[0]], 0.5 * (dz[:-1] + dz[1:]), [0.5 * dz[-1]]])
        fac = 4 * (galerkin + 1)
        ix, iy, iz = _grid(nx, ny, nz)
        c = np.array([0.8, -0.5, 0.3])
        u = np.tile(c, (s.n3d, 1))
        s.set_field(KIND_NUT, 0, np.zeros(s.n3d))
        conv = tau * (abs(c[0]) / hx + abs(c[1]) / hy + abs(c[2]) / hnode)
        vis = fac * tau * visc * (1 / hx ** 2 + 1 / hy ** 2 + 1 / hnode ** 2)
        np.testing.assert_allclose(s.cfl(u, tau, 1), conv.max(), rtol=1e-13)
        np.testing.assert_allclose(s.cfl(u, tau, 2), vis.max(), rtol=1e-13)
        np.testing.assert_allclose(s.cfl(u, tau, 3), (conv + vis).max(), rtol=1e-13)
        # one element of large nu_t: its 8 corner nodes see V / 8 (none on a wall)
        V, ke = 50.0, nz // 2
        e = (iz == ke) & (ix == 3) & (iy == 4)
        s.set_field(KIND_NUT, 0, np.where(e, V, 0.0))
        corner_planes = np.array([ke, ke + 1])
        expect = max(fac * tau * (visc + V / 8) * (1 / hx ** 2 + 1 / hy ** 2 + 1 / hnode[k] ** 2) for k in corner_planes)
        np.testing.assert_allclose(s.cfl(np.zeros_like(u), tau, 2), max(expect, vis.max()), rtol=1e-13)


# ---------------------------------------------------------------------------
# Staggered reductions: special cases.
# ---------------------------------------------------------------------------
def pin_staggered_integrals_and_cfl():
    """Staggered CalcIntegral / CalcKineticEnergy: u, v weighted by DZ[k] hx hy,
    w (faces k >= 1) by (DZ[k] + DZ[k-1])/2 hx hy, so a uniform field integrates to
    c Lx Ly 2 for u, v and c Lx Ly (2 - (DZ[0] + DZ[nz-1])/2) for w, and w = z to
    Lx Ly (2 - DZ[nz-1]).
    CalcCFL: mode 2 = 4 (NAIVE + 1) tau (visc + nu_t) (1/hx^2 + 1/hy^2 + 1/DZ^2), maximised:
    the thinnest cell without nu_t, the spiked cell with one large nu_t."""
    nx, ny, nz = 8, 6, 12
    tau = 2e-3
    for visc_type in (0, 1):
        s, zz, a = _stag(SCHEME_VV, nx, ny, nz, turb_visc=visc_type, visc_type=visc_type)
        hx, hy, visc = a["hx"], a["hy"], a["visc"]
        dz = np.diff(zz)
        box = hx * nx * hy * ny
        c = np.array([0.9, -1.1, 0.6])
        u = np.tile(c, (s.n3d, 1))
        wlen = 2.0 - 0.5 * (dz[0] + dz[-1])
        np.testing.assert_allclose(s.integral(u), [c[0] * box * 2, c[1] * box * 2, c[2] * box * wlen], rtol=1e-13)
        # w = z at the faces: the face weights are the trapezoidal rule on z_0..z_nz without
        # its two end terms, 0 at z = 0 and 2 DZ[nz-1] / 2 at z = 2
        lin = np.zeros((s.n3d, 3))
        lin[:, 2] = zz[np.arange(s.n3d) // (nx * ny)]
        np.testing.assert_allclose(s.integral(lin)[2], box * (2.0 - dz[-1]), rtol=1e-13)
        np.testing.assert_allclose(s.kinetic_energy(u), 0.5 * ((c[0] ** 2 + c[1] ** 2) * 2 + c[2] ** 2 * wlen) * box,
                                   rtol=1e-13)
        fac = 4 * (visc_type + 1)
        vis = fac * tau * visc * (1 / hx ** 2 + 1 / hy ** 2 + 1 / dz ** 2)
        conv = tau * (abs(c[0]) / hx + abs(c[1]) / hy + abs(c[2]) / dz)
        if visc_type:
            s.set_field(KIND_NUT, 0, np.zeros(s.n3d))
        np.testing.assert_allclose(s.cfl(u, tau, 2), vis.max(), rtol=1e-13)
        np.testing.assert_allclose(s.cfl(u, tau, 3), (vis + conv).max(), rtol=1e-13)
        if visc_type:
            ix, iy, iz = _grid(nx, ny, nz)
            V, k = 3.0, nz // 2
            s.set_field(KIND_NUT, 0, np.where((iz == k) & (ix == 2) & (iy == 3), V, 0.0))
            spike = fac * tau * (visc + V) * (1 / hx ** 2 + 1 / hy ** 2 + 1 / dz[k] ** 2)
            np.testing.assert_allclose(s.cfl(u, tau, 2), max(spike, vis.max()), rtol=1e-13)


def pin_staggered_profiles():
    """Staggered CalcProfiles: 12 columns per plane (u3, uu6, nu_t, p, pp), w taken
    at the cell centre as the mean of its two faces (0 above the top cell)."""
    nx, ny, nz = 8, 6, 7
    s, zz, a = _stag(SCHEME_HW, nx, ny, nz, turb_visc=1, visc_type=1)
    ix, iy, iz = _grid(nx, ny, nz)
    fw = np.where(iz > 0, 1.0 + 0.5 * iz, 0.0)
    u = np.stack([2.0 + iz, -1.0 + 0.2 * iz, fw], axis=1)
    p = 0.3 * iz - 1.0
    s.set_field(KIND_NUT, 0, 0.01 * iz)
    av = s.profiles(u, p)
    pw = fw.reshape(nz, -1)[:, 0]
    wc = 0.5 * (pw + np.concatenate([pw[1:], [0.0]]))
    ea, eb = 2.0 + np.arange(nz), -1.0 + 0.2 * np.arange(nz)
    pp = 0.3 * np.arange(nz) - 1.0
    exp = np.stack([ea, eb, wc, ea ** 2, eb ** 2, wc ** 2, ea * eb, ea * wc, eb * wc, 0.01 * np.arange(nz), pp,
                    pp ** 2], axis=1)
    np.testing.assert_allclose(av, exp, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------------------
# Collocated divergence of a given stress tensor (learned closure).
# ---------------------------------------------------------------------------
def pin_collocated_stress_divergence():
    """The explicit term with sgs_model = 1 and zero velocity is -d_j tau_ij
    (tau stored XX YY ZZ XY XZ YZ per node): central differences in x and y are
    exact for linear tau, the three-point formula on the nonuniform z nodes is
    exact for quadratics in z. So with tau_k = c_k + g_k . x + q_k z^2 the result
    at interior nodes is -(sum over j of d_j tau_ij) in closed form."""
    nx, ny, nz = 10, 8, 12
    s, zz, a = _ab3(nx, ny, nz, sgs_model=1)
    X = cases.node_coords(nx, ny, nz, a["hx"], a["hy"], zz)
    rng = np.random.default_rng(31)
    c = rng.standard_normal(6)
    G = rng.standard_normal((6, 3))
    q = rng.standard_normal(6)
    tau = c + X @ G.T + np.outer(X[:, 2] ** 2, q)
    s.set_field(KIND_TAU, 0, tau)
    f = s.flux_term(np.zeros((s.n3d, 3)), 0, 1)
    comp = [[0, 3, 4], [3, 1, 5], [4, 5, 2]]  # tau_ij -> stored component
    ix, iy, iz = _grid(nx, ny, nz)
    m = (ix >= 1) & (ix <= nx - 2) & (iy >= 1) & (iy <= ny - 2) & (iz >= 1) & (iz <= nz - 2)
    fex = np.zeros((s.n3d, 3))
    for i in range(3):
        fex[:, i] = -(G[comp[i][0], 0] + G[comp[i][1], 1] + G[comp[i][2], 2] + 2 * q[comp[i][2]] * X[:, 2])
    err = np.abs(f - fex)[m].max() / np.abs(fex[m]).max()
    assert err < 1e-11, err


PINS = {
    "naive_linear_viscosity": pin_naive_linear_viscosity,
    "staggered_viscous_wall_rows": pin_staggered_viscous_wall_rows,
    "galerkin_element_viscosity": pin_galerkin_element_viscosity,
    "collocated_integrals": pin_collocated_integrals,
    "collocated_profiles": pin_collocated_profiles,
    "collocated_cfl": pin_collocated_cfl,
    "staggered_integrals_and_cfl": pin_staggered_integrals_and_cfl,
    "staggered_profiles": pin_staggered_profiles,
    "collocated_stress_divergence": pin_collocated_stress_divergence,
}


def test_naive_linear_viscosity():
    pin_naive_linear_viscosity()


def test_staggered_viscous_wall_rows():
    pin_staggered_viscous_wall_rows()


def test_galerkin_element_viscosity():
    pin_galerkin_element_viscosity()


def test_collocated_integrals():
    pin_collocated_integrals()


def test_collocated_profiles():
    pin_collocated_profiles()


def test_collocated_cfl():
    pin_collocated_cfl()


def test_staggered_integrals_and_cfl():
    pin_staggered_integrals_and_cfl()


def test_staggered_profiles():
    pin_staggered_profiles()


def test_collocated_stress_divergence():
    pin_collocated_stress_divergence()


if __name__ == "__main__":  # python -m tests.test_oracle_closed_forms NAME...: run pins (exit 1 if one fails)
    import sys

    for name in sys.argv[1:]:
        PINS[name]()
