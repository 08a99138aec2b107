"""Pins of the FP64 CPU oracle.

The reference ships no golden vectors; its own known-answer tests are the
self-checks of staggered/selftest.cpp, which compute their own expected values.
They are restated here against the oracle, with thresholds (the reference
prints norms without thresholds). Further pins come from the mathematics:
the DFT (numpy), dense residuals, textbook BDF3/EX3 coefficients, closed forms
of the Smagorinsky model on linear shear, and exactness of the collocated
operators on low-order polynomials.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import SCHEME_AB3, SCHEME_HW, SCHEME_VV, KIND_FLUX, KIND_PRES, KIND_VEL
from tests import cases


def stag(scheme, nx=16, ny=16, nz=16, uniform=False, **kw):
    zz = cases.uniform_mesh(nz) if uniform else cases.sinh_mesh(nz)
    args = cases.ref_staggered_kwargs(nx, ny)
    args.update(kw)
    return oracle.Solver(scheme, nx, ny, nz, zz, **args), zz, args


def interior(nx, ny, nz, bb):
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = (ix >= bb) & (ix < nx - bb) & (iy >= bb) & (iy < ny - bb) & (iz >= bb) & (iz < nz - bb)
    return m.ravel()


# ---------------------------------------------------------------------------
# FFT (replaces FFTW/cuFFT): must be the unnormalised DFT.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 6, 8, 12, 30, 49, 96, 160, 192, 384])
def test_fft_is_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(oracle.fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(oracle.fft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("nx,ny", [(8, 8), (16, 4), (64, 32), (32, 128)])
def test_index_helpers_match_bitmask_at_pow2(nx, ny):
    assert oracle.index_selfcheck(nx, ny, 3) == 0


# ---------------------------------------------------------------------------
# staggered/selftest.cpp restated
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", [SCHEME_VV, SCHEME_HW])
def test_convection_consistency_linear_fields(scheme):
    """CheckConvectionConsistency (staggered/selftest.cpp:11-76), uniform mesh.
    hx = hy here (Hx = 4 pi over 32 cells, Hy = 2 pi over 16): see the next test
    for what the VV w-equation does when hx != hy."""
    nx, ny, nz = 32, 16, 16
    s, zz, a = stag(scheme, nx, ny, nz, uniform=True)
    C = [cases.coords(nx, ny, nz, a["hx"], a["hy"], zz, k) for k in range(3)]
    m = interior(nx, ny, nz, 6)
    err = 0.0
    for jtest in range(3):  # 1A
        for jcoor in range(3):
            u = np.zeros((s.n3d, 3))
            u[:, jtest] = C[jtest][:, jcoor]
            f = s.flux_term(u, 1, 0)
            fex = np.zeros((s.n3d, 3))
            if jcoor == jtest:
                fex[:, jcoor] -= 2 * C[jcoor][:, jcoor]
            err = max(err, np.abs(f - fex)[m].max())
    for it0 in range(3):  # 1B
        for it1 in range(3):
            for it2 in range(3):
                it = (it0, it1, it2)
                u = np.stack([C[0][:, it0], C[1][:, it1], C[2][:, it2]], axis=1)
                f = s.flux_term(u, 1, 0)
                divu = (it0 == 0) + (it1 == 1) + (it2 == 2)
                fex = np.zeros((s.n3d, 3))
                for c in range(3):
                    r = C[c]
                    fex[:, c] -= r[:, it[c]] * divu + r[:, it[it[c]]]
                err = max(err, np.abs(f - fex)[m].max())
    assert err < 1e-10


def test_vv_w_equation_y_flux_uses_hy():
    """The VV w-equation scales its y-face mass flux by hy where the face area is
    hx*dz (staggered/func_global.h:475,479,483,487). The oracle keeps this, so for
    u = x, v = x, w = y the w-flux is -(y + (hy/hx) x) instead of -(x + y); the
    u- and v-equations stay exact."""
    nx = ny = nz = 16
    s, zz, a = stag(SCHEME_VV, nx, ny, nz, uniform=True)
    assert a["hx"] != a["hy"]
    C = [cases.coords(nx, ny, nz, a["hx"], a["hy"], zz, k) for k in range(3)]
    m = interior(nx, ny, nz, 6)
    u = np.stack([C[0][:, 0], C[1][:, 0], C[2][:, 1]], axis=1)
    f = s.flux_term(u, 1, 0)
    fex_u = -(C[0][:, 0] * 1 + C[0][:, 0])
    fex_v = -(C[1][:, 0] * 1 + C[1][:, 0])
    fex_w = -(C[2][:, 1] + a["hy"] / a["hx"] * C[2][:, 0])
    assert np.abs(f[:, 0] - fex_u)[m].max() < 1e-10
    assert np.abs(f[:, 1] - fex_v)[m].max() < 1e-10
    assert np.abs(f[:, 2] - fex_w)[m].max() < 1e-10


def test_gradient_kernel_and_normalize_projection():
    """CheckGradientKernel (staggered/selftest.cpp:79-97)."""
    s, _, _ = stag(SCHEME_VV)
    kerg = s.mesh("KERG")
    p = np.repeat(kerg, s.nx * s.ny)
    g = s.apply_gradient(p)
    assert np.abs(g).max() < 1e-10 * np.abs(p).max() * s.mesh("GCZ").max()
    p = cases.sin_sq(s.n3d)
    p1 = s.normalize_pressure(p)
    p2 = s.normalize_pressure(p1)
    assert np.abs(p2 - p1).max() < 1e-14
    # and it removes exactly the KERG component
    proj = np.sum(p1 * np.repeat(kerg * s.mesh("DZ"), s.nx * s.ny))
    assert abs(proj) < 1e-12 * s.n3d


def test_pressure_matrix_symmetric_and_equals_DG():
    """CheckPressureMatrix (staggered/selftest.cpp:100-142)."""
    s, zz, _ = stag(SCHEME_VV, nz=24)
    M = s.fill_matrix()
    nz = s.nz
    for i in range(nz):
        for j in range(-3, 4):
            if 0 <= i + j < nz:
                assert abs(M[i, 3 + j] - M[i + j, 3 - j]) < 1e-9 * np.abs(M).max()
    pz = np.sin(np.arange(nz) ** 2 + 51.0)
    p = np.repeat(pz, s.nx * s.ny)
    dgp = s.apply_divergence(s.apply_gradient(p)).reshape(nz, -1)
    dgp_mat = np.array([sum(M[i, j + 3] * pz[i + j] for j in range(-3, 4) if 0 <= i + j < nz) for i in range(nz)])
    assert np.abs(dgp - dgp_mat[:, None]).max() < 1e-9 * np.abs(dgp_mat).max()
    # the lower bands / diagonals kept for the solver are the matrix's own
    A = s.mesh("MAT_A").reshape(nz, 3)
    for iz in range(nz):
        for i in range(3):
            exp = M[iz, i] if iz + i - 3 >= 0 else 0.0
            assert A[iz, i] == exp
    D = s.mesh("MAT_D")
    np.testing.assert_array_equal(D, [M[0, 3], M[1, 3], M[2, 3], M[nz - 3, 3], M[nz - 2, 3], M[nz - 1, 3]])
    # interior rows have zero row sum (staggered/func_global.h:138)
    assert np.abs(M[3:nz - 3].sum(axis=1)).max() < 1e-9 * np.abs(M).max()


@pytest.mark.parametrize("scheme", [SCHEME_VV, SCHEME_HW])
@pytest.mark.parametrize("nx,ny,nz", [(16, 16, 16), (12, 10, 14)])
def test_pressure_solver_round_trip(scheme, nx, ny, nz):
    """CheckPressureSolver (staggered/selftest.cpp:144-181), incl. non-power-of-2 sizes."""
    s, zz, _ = stag(scheme, nx, ny, nz)
    for itest in range(2):
        if itest == 0:
            iz = np.arange(s.n3d) // (nx * ny)
            p = np.sin(zz[iz])
        else:
            p = cases.sin_sq(s.n3d)
        p = s.normalize_pressure(p)
        dgp = s.apply_divergence(s.apply_gradient(p))
        pnew = s.normalize_pressure(s.solve_pressure(dgp))
        assert np.abs(p - pnew).max() < 1e-9, itest


@pytest.mark.parametrize("scheme", [SCHEME_VV, SCHEME_HW])
def test_projection_divergence_free_and_energy_conserving(scheme):
    """CheckConservation(1) (staggered/selftest.cpp:186-254), 3-D field on the stretched mesh.
    hx = hy as at the reference's default sizes (NX = 64 over 4 pi, NY = 32 over 2 pi);
    with hx != hy the VV w-equation's hy factor breaks the skew-symmetry."""
    s, zz, _ = stag(scheme, nx=32, ny=16, nz=16)
    n = s.n3d
    v = np.stack([np.sin(np.arange(n, dtype=float) ** 2 + math.cos(10.0 * c)) for c in range(3)], axis=1)
    v[: s.nx * s.ny, 2] = 0.0
    v = v - s.apply_gradient(s.unapply_gradient(v))
    p = s.unapply_gradient(v)
    assert np.abs(p).max() < 1e-9
    assert np.abs(s.apply_divergence(v)).max() < 1e-11
    f = s.flux_term(v, 1, 0)
    iz = np.arange(n) // (s.nx * s.ny)
    dz = s.mesh("DZ")
    if scheme == SCHEME_VV:
        hz = 1.125 * dz[iz] - 0.125 * s.mesh("DZW")[iz]
        gcz = s.mesh("GCZ")
        hzw = np.where(iz > 0, 1.0 / np.where(iz > 0, gcz[iz], 1.0), 0.0)
    else:
        hz = dz[iz]
        hzw = np.where(iz > 0, 0.5 * (dz[iz] + dz[np.maximum(iz - 1, 0)]), 0.0)
    K = [np.sum(f[:, 0] * v[:, 0] * hz), np.sum(f[:, 1] * v[:, 1] * hz), np.sum(f[:, 2] * v[:, 2] * hzw)]
    scale = np.sum(np.abs(f[:, 0] * v[:, 0] * hz))
    assert abs(sum(K)) < 1e-11 * scale


def test_conservation_2d_mode_uniform():
    """CheckConservation(0) (staggered/selftest.cpp:191-216): 2-D divergence-free mode, VV.
    The mode is solenoidal only for hx = hy (reference default NX = 64, NY = 32)."""
    nx, ny, nz = 32, 16, 8
    s, zz, _ = stag(SCHEME_VV, nx, ny, nz, uniform=True)
    kx, ky = 7, 9
    cx = math.sin(kx * math.pi / nx) - math.sin(kx * 3 * math.pi / nx) / 27.0
    cy = math.sin(ky * math.pi / ny) - math.sin(ky * 3 * math.pi / ny) / 27.0
    ix = np.arange(nx * ny) % nx
    iy = np.arange(nx * ny) // nx
    v2 = np.zeros((nx * ny, 3))
    v2[:, 0] = -cy * np.cos(2 * math.pi * (kx * (ix - 0.5) / nx + ky * iy / ny))
    v2[:, 1] = cx * np.cos(2 * math.pi * (kx * ix / nx + ky * (iy - 0.5) / ny))
    v = np.tile(v2, (nz, 1))
    assert np.abs(s.unapply_gradient(v)).max() < 1e-10
    f = s.flux_term(v, 1, 0)
    iz = np.arange(s.n3d) // (nx * ny)
    hz = 1.125 * s.mesh("DZ")[iz] - 0.125 * s.mesh("DZW")[iz]
    assert abs(np.sum(f[:, 0] * hz)) < 1e-10 and abs(np.sum(f[:, 1] * hz)) < 1e-10
    assert abs(np.sum((f[:, 0] * v[:, 0] + f[:, 1] * v[:, 1]) * hz)) < 1e-10


@pytest.mark.parametrize("uniform", [True, False])
def test_visc_naive_consistency(uniform):
    """CheckViscNaiveConsistency (staggered/selftest.cpp:256-340)."""
    nx = ny = nz = 16
    s, zz, a = stag(SCHEME_VV, nx, ny, nz, uniform=uniform, turb_visc=1)
    m = interior(nx, ny, nz, 3)
    err = 0.0
    for ic in range(3):
        r = cases.coords(nx, ny, nz, a["hx"], a["hy"], zz, ic)
        for px in range(3):
            for py in range(3):
                for pz in range(3):
                    if not uniform and pz == 2:
                        continue
                    v = np.zeros((s.n3d, 3))
                    v[:, ic] = r[:, 0] ** px * r[:, 1] ** py * r[:, 2] ** pz
                    v[: nx * ny, 2] = 0
                    f = s.visc_flux_naive(v) / a["visc"]
                    fex = np.zeros((s.n3d, 3))
                    for ivar in range(3):
                        qx = px - (ic == 0) - (ivar == 0)
                        qy = py - (ic == 1) - (ivar == 1)
                        qz = pz - (ic == 2) - (ivar == 2)
                        if qx < 0 or qy < 0 or qz < 0:
                            continue
                        rr = cases.coords(nx, ny, nz, a["hx"], a["hy"], zz, ivar)
                        val = rr[:, 0] ** qx * rr[:, 1] ** qy * rr[:, 2] ** qz
                        if qx < px and px == 2:
                            val = val * 2
                        if qy < py and py == 2:
                            val = val * 2
                        if qz < pz and pz == 2:
                            val = val * 2
                        fex[:, ivar] += val
                    if px == 2:
                        fex[:, ic] += 2 * r[:, 1] ** py * r[:, 2] ** pz
                    if py == 2:
                        fex[:, ic] += 2 * r[:, 0] ** px * r[:, 2] ** pz
                    if pz == 2:
                        fex[:, ic] += 2 * r[:, 0] ** px * r[:, 1] ** py
                    err = max(err, np.abs(f - fex)[m].max())
    assert err < 1e-8


def test_visc_cross_equals_naive_for_constant_viscosity_interior():
    """With nu_t = 0 the NAIVE stress divergence reduces to the CROSS Laplacian on
    divergence-free fields (nu (Lap u + grad div u)); interior cells."""
    nx = ny = nz = 16
    s0, zz, a = stag(SCHEME_HW, nx, ny, nz, visc_type=0)
    s1, _, _ = stag(SCHEME_HW, nx, ny, nz, visc_type=1)
    n = s0.n3d
    v = np.stack([np.sin(np.arange(n, dtype=float) ** 2 + c) for c in range(3)], axis=1)
    v[: nx * ny, 2] = 0
    v = v - s0.apply_gradient(s0.unapply_gradient(v))
    f0 = s0.flux_term(v, 0, 1)
    f1 = s1.flux_term(v, 0, 1)
    m = interior(nx, ny, nz, 2)
    assert np.abs(f0 - f1)[m].max() < 1e-9 * np.abs(f0).max()


# ---------------------------------------------------------------------------
# Smagorinsky-van Driest: closed forms on linear shear.
# ---------------------------------------------------------------------------
def test_abs_s_and_nut_on_linear_shear():
    nx = ny = nz = 16
    s, zz, a = stag(SCHEME_VV, nx, ny, nz, turb_visc=1)
    alpha = 0.7
    r = cases.coords(nx, ny, nz, a["hx"], a["hy"], zz, 0)
    u = np.zeros((s.n3d, 3))
    u[:, 0] = alpha * r[:, 2]
    S = s.calc_abs_s(u)
    iz = np.arange(s.n3d) // (nx * ny)
    m = (iz > 0) & (iz < nz - 1)
    np.testing.assert_allclose(S[m], abs(alpha), rtol=1e-12)
    s.recalc_turb_visc(u)
    nut = s.get_field(oracle.KIND_NUT, 0)
    zc = 0.5 * (zz[iz] + zz[iz + 1])
    d = np.minimum(zc, 2 - zc)
    yp = d * a["inv_length_scale"]
    delta = np.minimum(0.1 * (a["hx"] * a["hy"] * s.mesh("DZ")[iz]) ** (1 / 3), 0.41 * d)
    exp_nut = (1 - np.exp(-(0.04 ** 3) * yp ** 3)) * delta ** 2 * abs(alpha)
    np.testing.assert_allclose(nut[m], exp_nut[m], rtol=1e-12)


# ---------------------------------------------------------------------------
# Time step: BDF3 / EX3 weights are textbook values.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", [SCHEME_HW, SCHEME_VV, SCHEME_AB3])
def test_bdf_ex_weights(scheme):
    nx = ny = 8
    nz = 8
    if scheme == SCHEME_AB3:
        zz = cases.ab3_mesh_tanh(nz, 1.5)
    else:
        zz = cases.sinh_mesh(nz)
    s = oracle.Solver(scheme, nx, ny, nz, zz, visc=0.0, sour_dpdx=0.0, hx=0.3, hy=0.2, stab_type=0)
    n = s.n3d
    wall = np.zeros(n, bool)
    if scheme == SCHEME_AB3:
        wall[: nx * ny] = wall[-nx * ny:] = True
    tau = 0.01
    c0, c1, c2 = 1.3, 0.7, -0.4
    a1, a2 = 2.0, 5.0
    for order, ucoef, fcoef, bd0 in [
        (1, (1, 0, 0), (0, 0), 1.0),
        (2, (2, -0.5, 0), (-1, 0), 1.5),
        (3, (3, -1.5, 1 / 3), (-3, 1), 11 / 6),
    ]:
        cur = 2
        for slot, c in ((cur, c0), ((cur + 3) % 4, c1), ((cur + 2) % 4, c2)):
            u = np.zeros((n, 3))
            u[~wall, 0] = c
            s.set_field(KIND_VEL, slot, u)
        for slot, c in (((cur + 3) % 4, a1), ((cur + 2) % 4, a2)):
            f = np.zeros((n, 3))
            f[~wall, 0] = c
            s.set_field(KIND_FLUX, slot, f)
        s.step(0.0, tau, cur, order)
        un = s.get_field(KIND_VEL, (cur + 1) % 4)
        exp = (ucoef[0] * c0 + ucoef[1] * c1 + ucoef[2] * c2) / bd0 + tau * (fcoef[0] * a1 + fcoef[1] * a2) / bd0
        np.testing.assert_allclose(un[~wall, 0], exp, rtol=1e-13)
        assert np.abs(un[:, 1:]).max() < 1e-13


# ---------------------------------------------------------------------------
# Collocated-AB3
# ---------------------------------------------------------------------------
def ab3(nx=12, ny=10, nz=14, gamma=1.6, uniform=False, **kw):
    zz = np.linspace(0, 2, nz) if uniform else cases.ab3_mesh_tanh(nz, gamma)
    args = dict(visc=0.01, sour_dpdx=0.5, inv_length_scale=100.0, hx=0.31, hy=0.23)
    args.update(kw)
    return oracle.Solver(SCHEME_AB3, nx, ny, nz, zz, **args), zz, args


def test_ab3_convection_exact_on_linear_fields():
    nx, ny, nz = 16, 16, 16
    s, zz, a = ab3(nx, ny, nz, uniform=True, visc=0.0, sour_dpdx=0.0)
    X = cases.node_coords(nx, ny, nz, a["hx"], a["hy"], zz)
    m = interior(nx, ny, nz, 2)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((3, 3))
    c = rng.standard_normal(3)
    u = c + X @ B.T
    f = s.flux_term(u, 1, 0)
    grad = B  # du_i/dx_j = B[i, j]
    div = np.trace(B)
    fex = -(u @ grad.T + 0.5 * u * div)
    assert np.abs(f - fex)[m].max() < 1e-10


@pytest.mark.parametrize("galerkin", [0, 1])
def test_ab3_viscous_exact_on_quadratics_uniform(galerkin):
    """nu (Lap u + grad div u) for Galerkin Q1 (stress form) and nu Lap u for the FD form."""
    nx, ny, nz = 12, 12, 12
    s, zz, a = ab3(nx, ny, nz, uniform=True, visc=0.3, sour_dpdx=0.0, visc_type=galerkin)
    X = cases.node_coords(nx, ny, nz, a["hx"], a["hy"], zz)
    m = interior(nx, ny, nz, 2)
    for ic in range(3):
        for (px, py, pz) in [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]:
            u = np.zeros((s.n3d, 3))
            u[:, ic] = X[:, 0] ** px * X[:, 1] ** py * X[:, 2] ** pz
            f = s.flux_term(u, 0, 1) / 0.3
            fex = np.zeros((s.n3d, 3))
            p = (px, py, pz)
            lap = sum(2.0 * np.prod([X[:, k] ** p[k] for k in range(3) if k != j]) for j in range(3) if p[j] == 2)
            fex[:, ic] += lap
            if galerkin:
                for ivar in range(3):  # d/dx_ivar (d u_ic / dx_ic)
                    q = list(p)
                    coef = 1.0
                    for k in (ic, ivar):
                        coef *= q[k]
                        q[k] -= 1
                    if min(q) < 0 or coef == 0:
                        continue
                    fex[:, ivar] += coef * np.prod([X[:, k] ** q[k] for k in range(3)], axis=0)
            assert np.abs(f - fex)[m].max() < 1e-8, (ic, p)


def test_ab3_pressure_solver_inverts_compact_laplacian():
    """The collocated Poisson operator is the compact Laplacian with the MAT_A
    wall rows (collocated_AB3/func_global.h:8-61, FDmode=1); build it from its
    definition and check solve(L p) = p after normalisation."""
    nx, ny, nz = 12, 10, 14
    s, zz, a = ab3(nx, ny, nz)
    hx, hy = a["hx"], a["hy"]
    A = s.mesh("MAT_A")
    dz = np.diff(zz)
    p = cases.sin_sq(s.n3d).reshape(nz, ny, nx)
    Lp = np.zeros_like(p)
    for k in range(nz):
        hz = 0.5 * (dz[k - 1] if k > 0 else 0) + 0.5 * (dz[k] if k < nz - 1 else 0)
        if 0 < k < nz - 1:
            Lp[k] += 2 * hz * ((np.roll(p[k], -1, 1) - 2 * p[k] + np.roll(p[k], 1, 1)) * hy / hx / 2
                               + (np.roll(p[k], -1, 0) - 2 * p[k] + np.roll(p[k], 1, 0)) * hx / hy / 2)
        if k > 0:
            Lp[k] += A[k] * (p[k - 1] - p[k])
        if k < nz - 1:
            Lp[k] += A[k + 1] * (p[k + 1] - p[k])
    pn = s.normalize_pressure(p.ravel())
    Lp_flat = Lp.ravel()
    sol = s.normalize_pressure(s.solve_pressure(Lp_flat))
    # the operator is singular on constants; compare modulo a constant
    d = sol - pn
    assert np.abs(d - d.mean()).max() < 1e-9


def test_ab3_normalize_pressure_factor():
    """NormalizePressure with the hx*hy factor of collocated_AB3/func_global.h:167:
    a constant c maps to c (1 - hx hy)."""
    s, zz, a = ab3()
    out = s.normalize_pressure(np.full(s.n3d, 2.0))
    np.testing.assert_allclose(out, 2.0 * (1 - a["hx"] * a["hy"]), rtol=1e-12)


def test_ab3_element_strain_on_linear_field():
    nx, ny, nz = 10, 10, 12
    s, zz, a = ab3(nx, ny, nz, turb_visc=1, visc_type=1)
    X = cases.node_coords(nx, ny, nz, a["hx"], a["hy"], zz)
    alpha = 0.9
    u = np.zeros((s.n3d, 3))
    u[:, 0] = alpha * X[:, 2]
    s.recalc_turb_visc(u)
    nut = s.get_field(oracle.KIND_NUT, 0)
    iz = np.arange(s.n3d) // (nx * ny)
    m = iz < nz - 1
    Z = 0.5 * (zz[iz[m]] + zz[iz[m] + 1])
    d = np.minimum(Z, 2 - Z)
    yp = d * a["inv_length_scale"]
    hz = np.diff(zz)[iz[m]]
    delta = np.minimum(0.1 * (a["hx"] * a["hy"] * hz) ** (1 / 3), 0.41 * d)
    np.testing.assert_allclose(nut[m], (1 - np.exp(-(0.04 ** 3) * yp ** 3)) * delta ** 2 * alpha, rtol=1e-12)


# ---------------------------------------------------------------------------
# Reductions: integral, kinetic energy, profiles, CFL vs numpy definitions.
# ---------------------------------------------------------------------------
def test_reductions_staggered():
    nx, ny, nz = 12, 10, 14
    s, zz, a = stag(SCHEME_VV, nx, ny, nz, turb_visc=1)
    n = s.n3d
    u = np.stack([cases.sin_sq(n, 10 * c + 5) for c in range(3)], axis=1)
    u[: nx * ny, 2] = 0
    p = cases.sin_sq(n, 87)
    dz = np.diff(zz)
    iz = np.arange(n) // (nx * ny)
    vol = dz[iz] * a["hx"] * a["hy"]
    volw = np.where(iz > 0, 0.5 * (dz[iz] + dz[np.maximum(iz - 1, 0)]), 0.0) * a["hx"] * a["hy"]
    q = s.integral(u)
    np.testing.assert_allclose(q, [np.sum(u[:, 0] * vol), np.sum(u[:, 1] * vol), np.sum(u[:, 2] * volw)], rtol=1e-12)
    ke = s.kinetic_energy(u)
    np.testing.assert_allclose(ke, 0.5 * np.sum((u[:, 0] ** 2 + u[:, 1] ** 2) * vol + u[:, 2] ** 2 * volw), rtol=1e-12)
    av = s.profiles(u, p)
    U = u.reshape(nz, ny * nx, 3)
    wc = 0.5 * (U[:, :, 2] + np.concatenate([U[1:, :, 2], np.zeros((1, nx * ny))]))
    P = p.reshape(nz, -1)
    np.testing.assert_allclose(av[:, 0], U[:, :, 0].mean(1), atol=1e-14)
    np.testing.assert_allclose(av[:, 2], wc.mean(1), atol=1e-14)
    np.testing.assert_allclose(av[:, 7], (U[:, :, 0] * wc).mean(1), atol=1e-14)
    np.testing.assert_allclose(av[:, 11], (P * P).mean(1), atol=1e-14)
    tau = 1e-3
    cfl = s.cfl(u, tau, 1)
    exp = np.max(tau * (np.abs(u[:, 0]) / a["hx"] + np.abs(u[:, 1]) / a["hy"] + np.abs(u[:, 2]) / dz[iz]))
    np.testing.assert_allclose(cfl, exp, rtol=1e-13)


# ---------------------------------------------------------------------------
# New numerics: tanh mesh, seeded divergence-free initial field, CFR.
# ---------------------------------------------------------------------------
def test_tanh_mesh_closed_form():
    zz = oracle.tanh_mesh(64, 1.8)
    np.testing.assert_allclose(zz, cases.tanh_faces(64, 1.8), rtol=0, atol=1e-15)
    assert zz[0] == 0 and zz[-1] == 2 and np.all(np.diff(zz) > 0)
    np.testing.assert_allclose(zz + zz[::-1], 2.0, atol=1e-15)


@pytest.mark.parametrize("scheme", [SCHEME_HW, SCHEME_VV])
def test_initial_field_projected_is_divergence_free(scheme):
    nx, ny, nz = 16, 12, 16
    zz = cases.tanh_faces(nz, 1.8)
    args = cases.bulk_kwargs(nx, ny, 5600, 180)
    s = oracle.Solver(scheme, nx, ny, nz, zz, **args)
    u0 = s.initial_field(seed=42, ubulk=1.0, noise_rms=0.1)
    u1 = s.initial_field(seed=42, ubulk=1.0, noise_rms=0.1)
    np.testing.assert_array_equal(u0, u1)
    assert not np.array_equal(u0, s.initial_field(seed=43))
    pert = u0.copy()
    iz = np.arange(s.n3d) // (nx * ny)
    zc = 0.5 * (zz[iz] + zz[iz + 1])
    pert[:, 0] -= 1.5 * (1 - (zc - 1) ** 2)
    np.testing.assert_allclose(np.sqrt(np.mean(pert ** 2)), 0.1, rtol=1e-12)
    if scheme == SCHEME_HW:  # the discrete curl is exactly solenoidal for the 2nd-order MAC divergence
        assert np.abs(s.apply_divergence(u0)).max() < 1e-12
    u = s.project(u0)
    assert np.abs(s.apply_divergence(u)).max() < 1e-11


def test_cfr_correction_sets_bulk_and_keeps_divergence():
    nx, ny, nz = 12, 10, 14
    zz = cases.tanh_faces(nz, 1.8)
    args = cases.bulk_kwargs(nx, ny, 5600, 180)
    s = oracle.Solver(SCHEME_VV, nx, ny, nz, zz, cfr_ubulk=1.0, **args)
    u = s.project(s.initial_field(noise_rms=0.1)) * 0.93
    d0 = s.apply_divergence(u)
    u2, shift = s.cfr_correct(u)
    vol = (zz[-1] - zz[0]) * args["hx"] * nx * args["hy"] * ny
    np.testing.assert_allclose(s.integral(u2)[0] / vol, 1.0, rtol=1e-13)
    np.testing.assert_allclose(s.apply_divergence(u2), d0, atol=1e-12)
    assert shift > 0
