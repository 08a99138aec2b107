"""Shared set-up for the parity tests: one oracle solver and one HIP engine on
the same problem, plus comparison helpers."""
import numpy as np

import oracle
import turbchannel_amd as tca
from tests import cases

# name -> (scheme, nx, ny, nz, mesh, extra keyword arguments)
CASES = {
    "hw": (tca.SCHEME_HW, 16, 12, 16, "sinh", dict()),
    "hw_les": (tca.SCHEME_HW, 16, 12, 16, "sinh", dict(turb_visc=1, visc_type=1)),
    "vv": (tca.SCHEME_VV, 16, 12, 16, "sinh", dict()),
    "vv_les": (tca.SCHEME_VV, 16, 12, 16, "sinh", dict(turb_visc=1, visc_type=1)),
    "vv_odd": (tca.SCHEME_VV, 30, 10, 12, "tanh", dict()),
    "ab3": (tca.SCHEME_AB3, 16, 12, 18, "tanh", dict(stab_type=1, fd_mode=1, visc_type=0)),
    "ab3_les": (tca.SCHEME_AB3, 16, 12, 18, "tanh", dict(stab_type=1, fd_mode=1, visc_type=1, turb_visc=1)),
    "ab3_stab0": (tca.SCHEME_AB3, 12, 20, 14, "tanh", dict(stab_type=0, fd_mode=1, visc_type=0)),
    # 192 planes: the FP64 z-solve keeps its forward intermediate over the
    # spectrum instead of in LDS, on the launch path of the 384 x 384 x 192 grids
    "hw_tall": (tca.SCHEME_HW, 8, 6, 192, "sinh", dict()),
    "vv_tall": (tca.SCHEME_VV, 8, 6, 192, "tanh", dict()),
    "ab3_tall": (tca.SCHEME_AB3, 8, 6, 192, "tanh", dict(stab_type=1, fd_mode=1, visc_type=0)),
    # 264 planes: the FP32-bulk z-solve forms its factors in-lane with the forward
    # intermediate over the spectrum (as the 512 x 512 x 256 grid does); FP32 tests only
    "hw_deep": (tca.SCHEME_HW, 8, 6, 264, "tanh", dict()),
    "ab3_deep": (tca.SCHEME_AB3, 8, 6, 264, "tanh", dict(stab_type=1, fd_mode=1, visc_type=0)),
    # learned closure (collocated only), FD and Galerkin viscous terms
    "ab3_sgs": (tca.SCHEME_AB3, 32, 16, 18, "tanh", dict(stab_type=1, fd_mode=1, visc_type=0, sgs_model=1)),
    "ab3_sgs_gal": (tca.SCHEME_AB3, 24, 16, 14, "tanh", dict(stab_type=1, fd_mode=1, visc_type=1, sgs_model=1)),
}


def problem(name, cfr_ubulk=0.0):
    scheme, nx, ny, nz, mesh, extra = CASES[name]
    if scheme == tca.SCHEME_AB3:
        zz = cases.ab3_mesh_tanh(nz, 1.8)
        kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    elif mesh == "sinh":
        zz = cases.sinh_mesh(nz)
        kw = cases.ref_staggered_kwargs(nx, ny)
        kw.pop("xmin"), kw.pop("ymin")
    else:
        zz = cases.tanh_faces(nz, 1.8)
        kw = cases.bulk_kwargs(nx, ny, 5600, 180)
    kw.update(extra)
    kw["cfr_ubulk"] = cfr_ubulk
    return scheme, nx, ny, nz, zz, kw


def make_oracle(name, cfr_ubulk=0.0):
    scheme, nx, ny, nz, zz, kw = problem(name, cfr_ubulk)
    return oracle.Solver(scheme, nx, ny, nz, zz, **kw)


def make_engine(name, precision=tca.FP64, cfr_ubulk=0.0):
    scheme, nx, ny, nz, zz, kw = problem(name, cfr_ubulk)
    return tca.Engine(scheme, nx, ny, nz, zz, precision=precision, **kw)


def make_params(name, precision=tca.FP64, cfr_ubulk=0.0):
    scheme, nx, ny, nz, zz, kw = problem(name, cfr_ubulk)
    return tca.make_params(scheme, nx, ny, nz, zz, precision=precision, **kw)


def seeded_velocity(sol, seed=42, ubulk=1.0, noise=0.1, project=True):
    u = sol.initial_field(seed=seed, ubulk=ubulk, noise_rms=noise)
    return sol.project(u) if project else u


def rel(a, b):
    """max |a - b| / max |b| (max-norm relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d
