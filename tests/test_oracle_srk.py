"""Pins of the oracle's collocated SRK stepper (collocated_SRK/func_global.h:317-408)
and of its IMEX Butcher tables (collocated_SRK/imex.cpp).

The reference's SRK directory holds no tests or golden data; its main() only
compares CPU and GPU runs of ApplyGradient and UnapplyGradient, which are the
AB3 operators pinned elsewhere. So the pins here come from the mathematics:
  * the tables satisfy the order conditions of their published orders
    (explicit and implicit parts, and the coupling conditions of IMEX pairs);
  * for a velocity field u = (phi(z), 0, 0) with phi an eigenvector of the
    discrete viscous operator, the pressure stays zero and one step multiplies
    phi by the stability function R_E(lambda tau) of the explicit table;
  * with ARS(1,2,1) and alpha_tau = 1 a step is Chorin's projection method
    (StabType 0) or the incremental pressure-correction method (StabType 1);
  * the SRK pressure normalisation gives zero volume mean.
The bookkeeping of the multi-stage pressure sums (mujtilde, qj) and of the CK
terms (nu1tilde, d) beyond these cases is checked only by convergence in tau:
parity unpinned in that part.
"""
import numpy as np
import pytest

import oracle
from oracle import SCHEME_SRK

NOMINAL_ORDER = {"ARS_121": 1, "ARS_232": 2, "ARS_343": 3, "MARS_343": 3, "ARK3": 3, "ARK4": 4, "MARK3": 3,
                 "BHR_553": 3}
M = list(range(len(oracle.IMEX_METHODS)))


def srk(nx=8, ny=6, nz=9, **kw):
    zz = 1.0 - np.cos(np.pi * np.arange(nz) / (nz - 1))
    args = dict(visc=0.05, sour_dpdx=0.0, hx=0.4, hy=0.5, stab_type=0, fd_mode=1, visc_type=0)
    args.update(kw)
    return oracle.Solver(SCHEME_SRK, nx, ny, nz, zz, **args), zz


def rk_conditions(A, b, order):
    """Residuals of the Runge-Kutta order conditions up to `order` (Butcher trees)."""
    c = A.sum(1)
    r = [b.sum() - 1]
    if order >= 2:
        r.append(b @ c - 1 / 2)
    if order >= 3:
        r += [b @ c ** 2 - 1 / 3, b @ A @ c - 1 / 6]
    if order >= 4:
        r += [b @ c ** 3 - 1 / 4, (b * c) @ A @ c - 1 / 8, b @ A @ c ** 2 - 1 / 12, b @ A @ A @ c - 1 / 24]
    return np.abs(r)


@pytest.mark.parametrize("m", M)
def test_tables_meet_published_order(m):
    t = oracle.imex_table(m)
    s, E, I = t["s"], t["E"], t["I"]
    p = NOMINAL_ORDER[oracle.IMEX_METHODS[m]]
    tol = 1e-12 if m in (0, 1, 2, 4, 5) else 1e-11  # MARS/MARK3/BHR: coefficients given to ~14 digits
    assert rk_conditions(E[:s], E[s], p).max() < tol
    assert rk_conditions(I[:s], I[s], p).max() < tol
    # explicit part strictly lower triangular, implicit part lower triangular (DIRK)
    assert np.all(np.triu(E[:s]) == 0)
    assert np.all(np.triu(I[:s], 1) == 0)
    # the weights are not of a higher order than published (a wrong row would often pass p)
    if p < 4:
        assert rk_conditions(E[:s], E[s], p + 1).max() > 1e-6 or rk_conditions(I[:s], I[s], p + 1).max() > 1e-6


@pytest.mark.parametrize("m", M)
def test_imex_coupling_conditions(m):
    t = oracle.imex_table(m)
    s, E, I = t["s"], t["E"], t["I"]
    p = NOMINAL_ORDER[oracle.IMEX_METHODS[m]]
    AE, bE, AI, bI = E[:s], E[s], I[:s], I[s]
    cE, cI = AE.sum(1), AI.sum(1)
    tol = 1e-11
    np.testing.assert_allclose(cE, cI, atol=tol)  # common abscissae
    if p >= 2:
        assert abs(bE @ cI - 0.5) < tol and abs(bI @ cE - 0.5) < tol
    if p >= 3:
        for b in (bE, bI):
            for A1 in (AE, AI):
                for c in (cE, cI):
                    assert abs(b @ A1 @ c - 1 / 6) < tol


def test_type_and_stiff_accuracy():
    for m in M:
        t = oracle.imex_table(m)
        s, E, I = t["s"], t["E"], t["I"]
        np.testing.assert_array_equal(I[s], I[s - 1])  # stiffly accurate implicit part
        if t["type"] == oracle.IMEX_TYPE_ARS:
            assert np.all(I[:, 0] == 0) and E[s, 0] == 0  # ARS: first stage explicit only
        else:
            assert t["type"] == oracle.IMEX_TYPE_CK and np.all(I[1:s, 0] != 0) and I[0, 0] == 0
            np.testing.assert_array_equal(E[s], I[s])
        assert np.all(np.diag(I[1:s, 1:s]) == I[s - 1, s - 1])  # one diagonal value a_ss
    with pytest.raises(ValueError):
        oracle.imex_table(99)


def stability_function(A, b, z):
    s = len(b)
    return 1 + z * b @ np.linalg.solve(np.eye(s) - z * A, np.ones(s))


@pytest.mark.parametrize("stab", [0, 1])
@pytest.mark.parametrize("m", M)
def test_viscous_eigenmode_follows_explicit_stability_function(m, stab):
    """u = (phi(z), 0, 0): no convection and no divergence, so every pressure term
    vanishes and the step is the explicit method on u' = lambda u."""
    nx, ny, nz = 6, 4, 11
    s, zz = srk(nx, ny, nz, visc=0.3)
    n = s.n3d
    # the discrete viscous operator on the interior nodes, probed column by column
    A = np.zeros((nz - 2, nz - 2))
    for k in range(1, nz - 1):
        u = np.zeros((n, 3))
        u[k * nx * ny:(k + 1) * nx * ny, 0] = 1.0
        f = s.explicit_term(u)[:, 0].reshape(nz, ny, nx)
        A[:, k - 1] = f[1:-1, 0, 0]
    lam, V = np.linalg.eig(A)
    assert np.abs(lam.imag).max() < 1e-10
    t = oracle.imex_table(m)
    E = t["E"]
    for idx in (np.argmax(lam.real), np.argmin(lam.real)):
        phi = np.zeros(nz)
        phi[1:-1] = V[:, idx].real
        u = np.zeros((n, 3))
        u[:, 0] = np.repeat(phi, nx * ny)
        p = np.zeros(n)
        tau = 0.8 / abs(lam[idx].real) if idx == np.argmin(lam.real) else 0.05
        R = stability_function(E[:-1], E[-1], lam[idx].real * tau)
        v, pn, q = s.srk_step(0.0, tau, u, p, p, method=m, stab_type=stab, alpha_tau=0.7)
        np.testing.assert_allclose(v[:, 0], R * u[:, 0], rtol=0, atol=1e-13 * max(1, abs(R)))
        assert np.abs(v[:, 1:]).max() == 0 and np.abs(pn).max() < 1e-13 and np.abs(q).max() < 1e-13


def random_state(s, seed):
    rng = np.random.default_rng(seed)
    u = s.initial_field(seed=seed, ubulk=1.0, noise_rms=0.3)
    p = rng.standard_normal(s.n3d) * 0.1
    q = rng.standard_normal(s.n3d) * 0.1
    return u, s.normalize_pressure(p), q


def test_ars121_stab0_is_chorin_projection():
    s, _ = srk(sour_dpdx=0.2)
    u, p, q = random_state(s, 1)
    tau = 0.01
    v, pn, _ = s.srk_step(0.0, tau, u, p, q, method=0, stab_type=0, alpha_tau=1.0)
    F, G, Li = s.explicit_term, s.apply_gradient, s.unapply_gradient
    u1 = u + tau * (F(u) - G(p))      # explicit stage
    w = u + tau * F(u1)              # predictor
    phi = Li(w)                      # L^-1 D w
    np.testing.assert_allclose(v, w - G(phi), rtol=0, atol=1e-12)
    np.testing.assert_allclose(pn, phi / tau, rtol=0, atol=1e-10 * np.abs(phi / tau).max())


def test_ars121_stab1_is_incremental_projection():
    s, _ = srk(sour_dpdx=0.2, stab_type=1)
    u, p, q = random_state(s, 2)
    tau = 0.01
    v, pn, qn = s.srk_step(0.0, tau, u, p, q, method=0, stab_type=1, alpha_tau=1.0)
    F, G, Li = s.explicit_term, s.apply_gradient, s.unapply_gradient
    u1 = u + tau * (F(u) - G(p))
    w = u + tau * (F(u1) - G(p))     # predictor with the old pressure
    dp = Li(w) / tau                 # pressure increment
    np.testing.assert_allclose(v, w - tau * G(dp), rtol=0, atol=1e-12)
    np.testing.assert_allclose(pn, p + dp, rtol=0, atol=1e-10 * np.abs(dp).max())
    np.testing.assert_allclose(qn, dp / tau, rtol=0, atol=1e-10 * np.abs(dp / tau).max())


def test_srk_normalisation_is_zero_mean():
    s, zz = srk()
    rng = np.random.default_rng(5)
    p = s.normalize_pressure(rng.standard_normal(s.n3d) + 3.0)
    h = np.empty(len(zz))
    h[0], h[-1] = 0.5 * (zz[1] - zz[0]), 0.5 * (zz[-1] - zz[-2])
    h[1:-1] = 0.5 * (zz[2:] - zz[:-2])
    vol = np.repeat(h, s.nx * s.ny)
    assert abs(p @ vol) < 1e-12 * vol.sum()


@pytest.mark.parametrize("stab", [0, 1])
@pytest.mark.parametrize("m", [1, 2, 4, 5, 7])
def test_converges_in_tau(m, stab):
    """Errors against a fine-step run fall with tau. The pressure stabilisation
    of the scheme scales with tau, so the observed order is below the tables'
    (about 0.4 for StabType 0 and 1 for StabType 1 on this grid)."""
    s, _ = srk(8, 8, 11, visc=0.1, sour_dpdx=0.3, stab_type=stab)
    u0 = s.project(s.initial_field(seed=3, ubulk=1.0, noise_rms=0.2))
    p0 = np.zeros(s.n3d)
    T = 0.2

    def run(n):
        u, p, q = u0, p0, p0
        for k in range(n):
            u, p, q = s.srk_step(k * T / n, T / n, u, p, q, method=m, stab_type=stab, alpha_tau=1.0)
        return u

    ref = run(256)
    err = [np.abs(run(n) - ref).max() for n in (16, 32, 64)]
    assert err[0] > err[1] > err[2]
    assert err[2] < 0.2 * np.abs(u0).max() * (0.3 if stab else 1.0)
