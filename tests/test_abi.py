"""CPU checks of the C ABI library: it loads, exports every function of
include/turbchannel.h, and its host-side set-up (mesh data, initial field)
agrees with the oracle. No compute call needs a GPU here."""
import ctypes
import os

import numpy as np
import pytest

import turbchannel_amd as tca
from tests import parity


def test_library_exports_every_declared_function():
    L = tca.lib()
    names = tca.header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(tca, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(tca, "_lib", None)
    with pytest.raises(tca.TCError):
        tca.lib()


@pytest.mark.parametrize("name", ["hw", "vv", "vv_odd", "ab3"])
def test_mesh_data_matches_oracle(name):
    p = parity.make_params(name)
    o = parity.make_oracle(name)
    names = ["ZZ", "DZ", "MAT_A"]
    if p.scheme == tca.SCHEME_VV:
        names += ["DZW", "GC1", "GC2", "GCZ", "MAT_D", "KERG"]
    for n in names:
        np.testing.assert_allclose(tca.mesh_array(p, n), o.mesh(n), rtol=1e-14, atol=1e-300, err_msg=n)
    if p.scheme == tca.SCHEME_VV:
        np.testing.assert_allclose(tca.mesh_array(p, "KERG_norm2")[0], o.kerg_norm2(), rtol=1e-14)


@pytest.mark.parametrize("name", ["hw", "vv_odd", "ab3"])
def test_initial_field_matches_oracle(name):
    p = parity.make_params(name)
    o = parity.make_oracle(name)
    for seed in (42, 7):
        a = tca.make_initial_field(p, seed=seed, ubulk=1.0, noise_rms=0.1)
        b = o.initial_field(seed=seed, ubulk=1.0, noise_rms=0.1)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)


def test_create_rejects_bad_arguments():
    zz = np.linspace(0, 2, 9)
    p = tca.make_params(tca.SCHEME_VV, 15, 8, 8, zz, visc=0.01, hx=0.1, hy=0.1)  # odd nx
    h = ctypes.c_void_p()
    rc = tca.lib().tc_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert tca.lib().tc_last_error().decode()
    p = tca.make_params(tca.SCHEME_VV, 16, 8, 8, zz, visc=0.01, hx=0.1, hy=0.1, turb_visc=1, visc_type=0)
    rc = tca.lib().tc_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert rc == 1 and "CROSS" in tca.lib().tc_last_error().decode()


def test_mesh_array_rejects_bad_mesh():
    zz = np.array([0.0, 0.5, 0.4, 1.0, 1.5, 1.8, 2.0])  # not increasing
    p = tca.make_params(tca.SCHEME_HW, 8, 8, 6, zz, visc=0.01, hx=0.1, hy=0.1)
    assert tca.lib().tc_mesh_array(ctypes.byref(p), b"DZ", None, 0) == -1


@pytest.mark.parametrize("m", range(len(tca.IMEX_METHODS)))
def test_imex_tables_match_oracle(m):
    """tc_imex_table (host only) against the oracle's restatement of
    collocated_SRK/imex.cpp, whose order conditions tests/test_oracle_srk.py checks."""
    import oracle
    a, b = tca.imex_table(m), oracle.imex_table(m)
    assert a["s"] == b["s"] and a["type"] == b["type"]
    for k in ("E", "I", "d", "alpha_tau_max"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-15, atol=1e-16, err_msg=k)
    assert tca.lib().tc_imex_table(99, None, None, None, None, None, None) != 0


def test_srk_create_rejects_unsupported_options():
    zz = parity.problem("ab3")[4]
    h = ctypes.c_void_p()
    for extra, msg in ((dict(turb_visc=1), "turbulence"), (dict(visc_type=1), "FD"),
                       (dict(cfr_ubulk=1.0), "flow-rate"), (dict(imex=8), "IMEX"), (dict(stab_type=2), "StabType")):
        kw = dict(visc=0.01, hx=0.1, hy=0.1, stab_type=1)
        kw.update(extra)
        p = tca.make_params(tca.SCHEME_SRK, 16, 8, len(zz), zz, **kw)
        assert tca.lib().tc_create(ctypes.byref(p), 0, ctypes.byref(h)) == 1 and not h.value
        assert msg in tca.lib().tc_last_error().decode(), extra


def test_srk_mesh_data_is_collocated():
    zz = parity.problem("ab3")[4]
    p = tca.make_params(tca.SCHEME_SRK, 16, 12, len(zz), zz, visc=0.01, hx=0.1, hy=0.1)
    q = tca.make_params(tca.SCHEME_AB3, 16, 12, len(zz), zz, visc=0.01, hx=0.1, hy=0.1)
    for n in ("ZZ", "DZ", "MAT_A"):
        np.testing.assert_array_equal(tca.mesh_array(p, n), tca.mesh_array(q, n))


def _inlane_factors(W, R, KX, KY, nx, ny, nz):
    """The in-lane forward elimination of the FP32-bulk z-solve (ZInl::row in
    kernels_spectral.hip), restated in FP64 for every (kx, |ky|) column."""
    nxh, nyh = nx // 2 + 1, ny // 2 + 1
    ncols = nxh * nyh
    kx, ky = np.divmod(np.arange(ncols), nyh)
    M, D, U = np.zeros((W, nz, ncols)), np.zeros((nz, ncols)), np.zeros((W, nz, ncols))
    if W == 1:
        mm = KX[kx] + KY[ky]
        u = np.zeros(ncols)
        for i, (am, ap, c0, cq) in enumerate(R.reshape(nz, 4)):
            d = 1.0 / ((c0 + cq * mm) - am * u)
            u = ap * d
            M[0, i], D[i], U[0, i] = -am, d, u
        return M, D, U
    sx, s3x = KX.reshape(-1, 2)[kx].T
    sy, s3y = KY.reshape(-1, 2)[ky].T
    b1, b2, b3 = (np.zeros((3, ncols)) for _ in range(3))
    for i, r in enumerate(R.reshape(nz, 12)):
        cx, cy = r[7] * sx - r[8] * s3x, r[7] * sy - r[8] * s3y
        dg = r[3] - r[9] * cx * cx - r[10] * cy * cy
        q3 = -r[0] + 0 * dg
        l2 = r[1] + q3 * b3[0]
        l3 = r[2] + q3 * b3[1]
        dg = dg + q3 * b3[2]
        q2 = -l2
        l3 = l3 + q2 * b2[0]
        dg = dg + q2 * b2[1]
        u1 = r[4] + q2 * b2[2]
        q1 = -l3
        dg = dg + q1 * b1[0]
        u1 = u1 + q1 * b1[1]
        u2 = r[5] + q1 * b1[2]
        d = 1.0 / dg
        b3, b2, b1 = b2, b1, np.array([u1 * d, u2 * d, r[6] * d])
        M[0, i], M[1, i], M[2, i], D[i] = q1, q2, q3, d
        U[:, i] = b1
    return M, D, U


@pytest.mark.parametrize("name,extra", [("hw", {}), ("ab3", {}), ("ab3", {"fd_mode": 0}), ("ab3_stab0", {}),
                                        ("vv", {}), ("vv_odd", {})])
def test_zsolve_inlane_constants_reproduce_the_factors(name, extra):
    """The per-row constants and wavenumber terms from which the FP32-bulk
    z-solve forms its forward factors in-lane give the host's precomputed
    elimination (the factors the FP64 engine streams, which the GPU parity tests
    hold to the oracle's ThomasForPressure) for every column but the pinned (0,0)
    one, which the FP32-bulk engine solves from the precomputed FP64 factors."""
    scheme, nx, ny, nz, zz, kw = parity.problem(name)
    kw.update(extra)
    p = tca.make_params(scheme, nx, ny, nz, zz, **kw)
    W = 3 if scheme == tca.SCHEME_VV else 1
    ncols = (nx // 2 + 1) * (ny // 2 + 1)
    Mh = tca.mesh_array(p, "ZF_M").reshape(W, nz, ncols)
    Dh = tca.mesh_array(p, "ZF_D").reshape(nz, ncols)
    Uh = tca.mesh_array(p, "ZF_U").reshape(W, nz, ncols)
    with np.errstate(divide="ignore", invalid="ignore"):  # the unpinned (0,0) column may be singular
        M, D, U = _inlane_factors(W, tca.mesh_array(p, "ZF_R"), tca.mesh_array(p, "ZF_KX"),
                                  tca.mesh_array(p, "ZF_KY"), nx, ny, nz)
    for got, ref, what in ((M, Mh, "M"), (D, Dh, "D"), (U, Uh, "U")):
        g, r = got[..., 1:], ref[..., 1:]
        np.testing.assert_allclose(g, r, rtol=1e-11, atol=1e-12 * np.abs(r).max(), err_msg=what)
    # the (0,0) column differs: its first pivot is doubled by the pin
    assert not abs(D[0, 0] - Dh[0, 0]) <= 0.1 * abs(Dh[0, 0])
