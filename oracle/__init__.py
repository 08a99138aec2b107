"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this package; it is the checker, never the thing measured or shipped.
The C restatement it loads is ``oracle/tc_oracle.c`` (built by ``oracle/Makefile``).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TC_ORACLE_LIB names another build of the same source (the mutation tests load
# copies of the oracle with one planted mistake each this way).
_LIB_PATH = os.environ.get("TC_ORACLE_LIB") or os.path.join(_HERE, "libtc_oracle.so")

SCHEME_HW, SCHEME_VV, SCHEME_AB3, SCHEME_SRK = 0, 1, 2, 3
IMEX_METHODS = ("ARS_121", "ARS_232", "ARS_343", "MARS_343", "ARK3", "ARK4", "MARK3", "BHR_553")
IMEX_TYPE_A, IMEX_TYPE_ARS, IMEX_TYPE_CK = 0, 1, 2
KIND_VEL, KIND_FLUX, KIND_PRES, KIND_NUT, KIND_TAU = 0, 1, 2, 3, 4


class OrParams(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("turb_visc", ctypes.c_int), ("visc_type", ctypes.c_int), ("stab_type", ctypes.c_int),
        ("fd_mode", ctypes.c_int),
        ("visc", ctypes.c_double), ("sour_dpdx", ctypes.c_double), ("inv_length_scale", ctypes.c_double),
        ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("xmin", ctypes.c_double), ("ymin", ctypes.c_double),
        ("cfr_ubulk", ctypes.c_double), ("sgs_model", ctypes.c_int),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "libtc_oracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d = ctypes.c_double
        dp = P(d)
        vp = ctypes.c_void_p
        L.or_create.restype = vp
        L.or_create.argtypes = [P(OrParams), dp]
        L.or_destroy.argtypes = [vp]
        L.or_last_error.restype = ctypes.c_char_p
        L.or_mesh_array.argtypes = [vp, ctypes.c_char_p, dp, ctypes.c_int]
        L.or_kerg_norm2.argtypes = [vp]
        L.or_kerg_norm2.restype = d
        L.or_fill_matrix.argtypes = [vp, dp]
        L.or_solve_banded.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, d]
        for name in ("or_explicit_term", "or_apply_gradient", "or_apply_divergence", "or_unapply_gradient",
                     "or_calc_abs_s", "or_visc_flux_naive"):
            getattr(L, name).argtypes = [vp, dp, dp]
        L.or_flux_term.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int, dp]
        L.or_solve_pressure.argtypes = [vp, dp, dp]
        L.or_normalize_pressure.argtypes = [vp, dp]
        L.or_timers.argtypes = [vp, dp, ctypes.c_int]
        L.or_recalc_turb_visc.argtypes = [vp, dp]
        L.or_integral.argtypes = [vp, dp, dp]
        L.or_kinetic_energy.argtypes = [vp, dp]
        L.or_kinetic_energy.restype = d
        L.or_profiles.argtypes = [vp, dp, dp, dp]
        L.or_cfl.argtypes = [vp, dp, d, ctypes.c_int]
        L.or_cfl.restype = d
        L.or_set_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.or_get_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.or_step.argtypes = [vp, d, d, ctypes.c_int, ctypes.c_int]
        L.or_imex_table.argtypes = [ctypes.c_int, P(ctypes.c_int), dp, dp, dp, dp]
        L.or_srk_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, d, d, d, dp, dp, dp]
        L.or_cfr_correct.argtypes = [vp, dp]
        L.or_cfr_correct.restype = d
        L.or_tanh_mesh.argtypes = [ctypes.c_int, d, dp]
        L.or_initial_field.argtypes = [vp, ctypes.c_ulonglong, d, d, dp]
        L.or_fft.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.or_index_selfcheck.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tanh_mesh(n, gamma):
    zz = np.zeros(n + 1)
    lib().or_tanh_mesh(n, gamma, _p(zz))
    return zz


def fft(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().or_fft(len(x), _p(x.view(np.float64)), _p(out.view(np.float64)), sign)
    return out


def index_selfcheck(nx, ny, nz):
    return lib().or_index_selfcheck(nx, ny, nz)


def imex_table(method):
    """Butcher tables of one IMEX method (collocated_SRK/imex.cpp): dict with
    s, type, E and I ((s+1) x s; row s = weights), d (CalcD) and alpha_tau_max."""
    be, bi, d, am = np.zeros(42), np.zeros(42), np.zeros(6), np.zeros(2)
    t = ctypes.c_int(0)
    s = lib().or_imex_table(int(method), ctypes.byref(t), _p(be), _p(bi), _p(d), _p(am))
    if s == 0:
        raise ValueError(f"unknown IMEX method {method}")
    n = s * (s + 1)
    return {"s": s, "type": t.value, "E": be[:n].reshape(s + 1, s), "I": bi[:n].reshape(s + 1, s),
            "d": d[:s].copy(), "alpha_tau_max": am}


def set_threads(n):
    lib().or_set_threads(int(n))


class Solver:
    """Mirror of the reference's ``data`` struct (staggered/data_global_inc.h:102-167)."""

    def __init__(self, scheme, nx, ny, nz, zz, *, visc, sour_dpdx=0.0, inv_length_scale=1.0, hx, hy,
                 turb_visc=0, visc_type=None, stab_type=1, fd_mode=1, xmin=0.0, ymin=0.0, cfr_ubulk=0.0,
                 sgs_model=0):
        if visc_type is None:
            visc_type = 1 if turb_visc else 0
        self.params = OrParams(scheme, nx, ny, nz, turb_visc, visc_type, stab_type, fd_mode, visc, sour_dpdx,
                               inv_length_scale, hx, hy, xmin, ymin, cfr_ubulk, sgs_model)
        self.scheme, self.nx, self.ny, self.nz = scheme, nx, ny, nz
        self.n3d = nx * ny * nz
        self.zz = np.ascontiguousarray(zz, dtype=np.float64)
        self._h = lib().or_create(ctypes.byref(self.params), _p(self.zz))
        if not self._h:
            raise RuntimeError("oracle: " + lib().or_last_error().decode())

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- mesh -------------------------------------------------------------
    def mesh(self, name):
        n = lib().or_mesh_array(self._h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().or_mesh_array(self._h, name.encode(), _p(out), n)
        return out

    def kerg_norm2(self):
        return lib().or_kerg_norm2(self._h)

    def fill_matrix(self):
        M = np.zeros(7 * self.nz)
        lib().or_fill_matrix(self._h, _p(M))
        return M.reshape(self.nz, 7)

    # --- operators --------------------------------------------------------
    def vec(self):
        return np.zeros((self.n3d, 3))

    def scal(self):
        return np.zeros(self.n3d)

    def explicit_term(self, u):
        out = self.vec()
        lib().or_explicit_term(self._h, _p(np.ascontiguousarray(u)), _p(out))
        return out

    def flux_term(self, u, do_conv, do_visc):
        out = self.vec()
        lib().or_flux_term(self._h, _p(np.ascontiguousarray(u)), int(do_conv), int(do_visc), _p(out))
        return out

    def visc_flux_naive(self, u):
        out = self.vec()
        lib().or_visc_flux_naive(self._h, _p(np.ascontiguousarray(u)), _p(out))
        return out

    def apply_gradient(self, p):
        out = self.vec()
        lib().or_apply_gradient(self._h, _p(np.ascontiguousarray(p)), _p(out))
        return out

    def apply_divergence(self, v):
        out = self.scal()
        lib().or_apply_divergence(self._h, _p(np.ascontiguousarray(v)), _p(out))
        return out

    def solve_pressure(self, f):
        out = self.scal()
        lib().or_solve_pressure(self._h, _p(out), _p(np.ascontiguousarray(f)))
        return out

    def unapply_gradient(self, v):
        out = self.scal()
        lib().or_unapply_gradient(self._h, _p(np.ascontiguousarray(v)), _p(out))
        return out

    def normalize_pressure(self, p):
        p = np.array(p, dtype=np.float64, copy=True)
        lib().or_normalize_pressure(self._h, _p(p))
        return p

    TIMER_CHANNELS = ("explicit", "gradient", "divergence", "pressure_fft", "pressure_thomas",
                      "pressure_normalize", "other")

    def timers(self, reset=False):
        """Seconds per timer channel since the last reset (or_timers)."""
        out = np.zeros(len(self.TIMER_CHANNELS))
        lib().or_timers(self._h, _p(out), int(reset))
        return dict(zip(self.TIMER_CHANNELS, out.tolist()))

    def recalc_turb_visc(self, u):
        lib().or_recalc_turb_visc(self._h, _p(np.ascontiguousarray(u)))

    def calc_abs_s(self, u):
        out = self.scal()
        lib().or_calc_abs_s(self._h, _p(np.ascontiguousarray(u)), _p(out))
        return out

    def integral(self, u):
        out = np.zeros(3)
        lib().or_integral(self._h, _p(np.ascontiguousarray(u)), _p(out))
        return out

    def kinetic_energy(self, u):
        return lib().or_kinetic_energy(self._h, _p(np.ascontiguousarray(u)))

    def profiles(self, u, p):
        w = 10 if self.scheme in (SCHEME_AB3, SCHEME_SRK) else 12
        out = np.zeros((self.nz, w))
        lib().or_profiles(self._h, _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(p)), _p(out))
        return out

    def cfl(self, u, tau, mode=3):
        return lib().or_cfl(self._h, _p(np.ascontiguousarray(u)), tau, mode)

    def project(self, u):
        """u - G L^-1 D u (one UnapplyGradient + ApplyGradient)."""
        p = self.unapply_gradient(u)
        return u - self.apply_gradient(p)

    # --- state ------------------------------------------------------------
    def set_field(self, kind, slot, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if lib().or_set_field(self._h, kind, slot, _p(a)):
            raise ValueError("bad field")

    def get_field(self, kind, slot):
        out = self.scal() if kind in (KIND_PRES, KIND_NUT) else np.zeros((self.n3d, 6)) if kind == KIND_TAU \
            else self.vec()
        if lib().or_get_field(self._h, kind, slot, _p(out)):
            raise ValueError("bad field")
        return out

    def step(self, t, tau, cur, order):
        if lib().or_step(self._h, t, tau, cur, order):
            raise RuntimeError(lib().or_last_error().decode())

    def srk_step(self, t, tau, velocity, pressure, dpdt=None, *, method=2, stab_type=1, alpha_tau=1.0):
        """data::Step of the collocated SRK solver (collocated_SRK/func_global.h:317-408)
        on copies; returns (velocity, pressure, dpdt)."""
        v = np.array(velocity, dtype=np.float64, copy=True)
        p = np.array(pressure, dtype=np.float64, copy=True)
        q = np.zeros_like(p) if dpdt is None else np.array(dpdt, dtype=np.float64, copy=True)
        if lib().or_srk_step(self._h, int(method), int(stab_type), float(alpha_tau), float(t), float(tau),
                             _p(v), _p(p), _p(q)):
            raise RuntimeError(lib().or_last_error().decode())
        return v, p, q

    def cfr_correct(self, u):
        u = np.array(u, dtype=np.float64, copy=True)
        shift = lib().or_cfr_correct(self._h, _p(u))
        return u, shift

    def initial_field(self, seed=42, ubulk=1.0, noise_rms=0.1):
        out = self.vec()
        lib().or_initial_field(self._h, seed, ubulk, noise_rms, _p(out))
        return out
