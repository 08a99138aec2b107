"""MI355X time stepper for TurbChannel's staggered and collocated solvers.

Python mirror of the reference's solver interface (the ``data`` struct driven by
``main()``: staggered/data_global_inc.h:102-167) over the C ABI of
``include/turbchannel.h``, implemented by ``libturbchannel_amd.so`` (HIP, gfx950).
There is no fallback: if the library is missing or the device cannot be used,
every call raises.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libturbchannel_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "turbchannel.h")

SCHEME_HW, SCHEME_VV, SCHEME_AB3, SCHEME_SRK = 0, 1, 2, 3
# IMEX Runge-Kutta methods of the SRK solver, in tc_imex order (collocated_SRK/imex.h:26-42)
IMEX_METHODS = ("ARS_121", "ARS_232", "ARS_343", "MARS_343", "ARK3", "ARK4", "MARK3", "BHR_553")
IMEX_TYPE_A, IMEX_TYPE_ARS, IMEX_TYPE_CK = 0, 1, 2
FP64, FP32_BULK = 0, 1
VEL, FLUX, PRES, NUT, TAU = 0, 1, 2, 3, 4
KERNELS = ["turb_visc", "flux", "div", "fft_x", "fft_y", "zsolve", "ifft_y", "ifft_x", "project", "other"]


class TCError(RuntimeError):
    pass


class TcParams(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int), ("precision", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("turb_visc", ctypes.c_int), ("visc_type", ctypes.c_int), ("stab_type", ctypes.c_int),
        ("fd_mode", ctypes.c_int),
        ("visc", ctypes.c_double), ("sour_dpdx", ctypes.c_double), ("inv_length_scale", ctypes.c_double),
        ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("xmin", ctypes.c_double), ("ymin", ctypes.c_double),
        ("cfr_ubulk", ctypes.c_double),
        ("zz", ctypes.POINTER(ctypes.c_double)),
        ("sgs_model", ctypes.c_int),
        ("imex", ctypes.c_int), ("alpha_tau", ctypes.c_double),
    ]


_lib = None


def header_functions():
    """Names of the functions declared in include/turbchannel.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|long long|const char\*)\s+(tc_[a-z_0-9]+)\s*\(", text, re.M)))


def lib():
    """Load the HIP engine. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TCError(f"{LIB_PATH} is missing: build it with `python turbchannel_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    c_int, c_double, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    dp = P(c_double)
    L.tc_last_error.restype = ctypes.c_char_p
    L.tc_create.argtypes = [P(TcParams), c_int, P(vp)]
    L.tc_destroy.argtypes = [vp]
    L.tc_real_size.argtypes = [vp]
    L.tc_sync.argtypes = [vp]
    L.tc_set_field.argtypes = [vp, c_int, c_int, vp, ctypes.c_size_t]
    L.tc_get_field.argtypes = [vp, c_int, c_int, vp, ctypes.c_size_t]
    L.tc_field_device_ptr.argtypes = [vp, c_int, c_int, P(vp)]
    L.tc_step.argtypes = [vp, c_double, c_double, c_int, c_int]
    L.tc_set_graphs.argtypes = [vp, c_int]
    L.tc_step_diagnostics.argtypes = [vp, c_int]
    L.tc_srk_step.argtypes = [vp, c_double, c_double, c_int, c_int, c_int]
    L.tc_imex_table.argtypes = [c_int, P(c_int), P(c_int), dp, dp, dp, dp]
    for n in ("tc_explicit_term", "tc_apply_gradient", "tc_unapply_gradient"):
        getattr(L, n).argtypes = [vp, c_int, c_int, c_int, c_int]
    L.tc_normalize_pressure.argtypes = [vp, c_int, c_int]
    L.tc_recalc_turb_visc.argtypes = [vp, c_int, c_int]
    L.tc_integral.argtypes = [vp, c_int, c_int, dp]
    L.tc_kinetic_energy.argtypes = [vp, c_int, c_int, dp]
    L.tc_cfl.argtypes = [vp, c_int, c_int, c_double, c_int, dp]
    L.tc_profiles.argtypes = [vp, c_int, c_int, c_int, dp, ctypes.c_size_t]
    L.tc_diagnostics.argtypes = [vp, c_int, c_int, c_double, c_int, dp, dp, dp, dp, ctypes.c_size_t]
    L.tc_cfr_correct.argtypes = [vp, c_int, c_int, dp]
    L.tc_project.argtypes = [vp, c_int, c_int]
    L.tc_make_initial_field.argtypes = [P(TcParams), ctypes.c_ulonglong, c_double, c_double, dp]
    L.tc_initial_field.argtypes = [vp, c_int, ctypes.c_ulonglong, c_double, c_double]
    L.tc_mesh_array.argtypes = [P(TcParams), ctypes.c_char_p, dp, c_int]
    L.tc_timing_enable.argtypes = [vp, c_int]
    L.tc_timing_read.argtypes = [vp, c_int, dp, P(ctypes.c_longlong)]
    L.tc_timing_reset.argtypes = [vp]
    fp = P(ctypes.c_float)
    ll = ctypes.c_longlong
    L.tc_set_sgs_params.argtypes = [vp, fp, ctypes.c_size_t]
    L.tc_sgs_sample_count.argtypes = [vp, c_int]
    L.tc_sgs_sample_count.restype = ll
    L.tc_sgs_samples.argtypes = [vp, c_int, c_int, c_int, vp, vp, ll, P(ll)]
    L.tc_mlp_param_count.argtypes = []
    L.tc_mlp_create.argtypes = [c_int, ctypes.c_ulonglong, P(vp)]
    L.tc_mlp_destroy.argtypes = [vp]
    L.tc_mlp_set_params.argtypes = [vp, fp, ctypes.c_size_t]
    L.tc_mlp_get_params.argtypes = [vp, fp, ctypes.c_size_t]
    L.tc_mlp_forward.argtypes = [vp, vp, ll, vp]
    L.tc_mlp_gradient.argtypes = [vp, vp, vp, ll, fp, ctypes.c_size_t, dp]
    L.tc_mlp_train_step.argtypes = [vp, vp, vp, ll, ctypes.c_float, dp]
    L.tc_mlp_train_step_noalloc.argtypes = [vp, vp, vp, ll, ctypes.c_float, dp]
    L.tc_mlp_steps.argtypes = [vp, P(ll)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise TCError(f"turbchannel error {rc}: {lib().tc_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(scheme, nx, ny, nz, zz, *, precision=FP64, visc, sour_dpdx=0.0, inv_length_scale=1.0, hx, hy,
                turb_visc=0, visc_type=None, stab_type=1, fd_mode=1, xmin=0.0, ymin=0.0, cfr_ubulk=0.0, sgs_model=0,
                imex=2, alpha_tau=0.0):
    """tc_params for a problem; keeps the mesh array alive on the returned object.
    imex and alpha_tau apply to SCHEME_SRK (method index into IMEX_METHODS; alpha_tau
    <= 0 selects the reference's default 0.5 * alpha_tau_max[stab_type])."""
    if visc_type is None:
        visc_type = 1 if turb_visc else 0
    zz = np.ascontiguousarray(zz, dtype=np.float64)
    p = TcParams(scheme, precision, nx, ny, nz, turb_visc, visc_type, stab_type, fd_mode, visc, sour_dpdx,
                 inv_length_scale, hx, hy, xmin, ymin, cfr_ubulk, _dptr(zz), sgs_model, imex, alpha_tau)
    p._zz = zz
    return p


def imex_table(method):
    """Butcher tables of one IMEX method (tIMEXMethod, collocated_SRK/imex.cpp): dict
    with s, type, E and I ((s+1) x s, row s = weights), d (CalcD), alpha_tau_max."""
    n, t = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().tc_imex_table(int(method), ctypes.byref(n), ctypes.byref(t), None, None, None, None))
    s = n.value
    be, bi, d, am = np.zeros(s * (s + 1)), np.zeros(s * (s + 1)), np.zeros(s), np.zeros(2)
    _check(lib().tc_imex_table(int(method), None, None, _dptr(be), _dptr(bi), _dptr(d), _dptr(am)))
    return {"s": s, "type": t.value, "E": be.reshape(s + 1, s), "I": bi.reshape(s + 1, s), "d": d,
            "alpha_tau_max": am}


def mesh_array(params, name):
    """Host mesh data derived from the parameters (no device needed)."""
    n = lib().tc_mesh_array(ctypes.byref(params), name.encode(), None, 0)
    if n < 0:
        raise KeyError(name)
    out = np.zeros(n)
    lib().tc_mesh_array(ctypes.byref(params), name.encode(), _dptr(out), n)
    return out


def make_initial_field(params, seed=42, ubulk=1.0, noise_rms=0.1):
    out = np.zeros((params.nx * params.ny * params.nz, 3))
    _check(lib().tc_make_initial_field(ctypes.byref(params), seed, ubulk, noise_rms, _dptr(out)))
    return out


class Engine:
    """One engine on one GPU; methods follow the reference's ``data`` struct.

    Fields are addressed as (kind, slot) with kind in VEL, FLUX, PRES, NUT and
    slot in 0..3 (the ring buffers vel/flux/pres[4]; NUT has slot 0).
    Host arrays are float64 for FP64 engines and float32 for FP32-bulk engines,
    shaped (N3D, 3) for vectors and (N3D,) for scalars.
    """

    def __init__(self, scheme, nx, ny, nz, zz, device=0, **kw):
        self.params = make_params(scheme, nx, ny, nz, zz, **kw)
        self.scheme, self.nx, self.ny, self.nz = scheme, nx, ny, nz
        self.n3d = nx * ny * nz
        h = ctypes.c_void_p()
        _check(lib().tc_create(ctypes.byref(self.params), device, ctypes.byref(h)))
        self._h = h
        self.dtype = np.float64 if lib().tc_real_size(h) == 8 else np.float32

    @classmethod
    def from_params(cls, params, device=0):
        e = cls.__new__(cls)
        e.params = params
        e.scheme, e.nx, e.ny, e.nz = params.scheme, params.nx, params.ny, params.nz
        e.n3d = params.nx * params.ny * params.nz
        h = ctypes.c_void_p()
        _check(lib().tc_create(ctypes.byref(params), device, ctypes.byref(h)))
        e._h = h
        e.dtype = np.float64 if lib().tc_real_size(h) == 8 else np.float32
        return e

    def close(self):
        if getattr(self, "_h", None):
            lib().tc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- fields -----------------------------------------------------------
    def _shape(self, kind):
        return (self.n3d, 3) if kind in (VEL, FLUX) else (self.n3d, 6) if kind == TAU else (self.n3d,)

    def set_field(self, kind, slot, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype).reshape(self._shape(kind)))
        _check(lib().tc_set_field(self._h, kind, slot, a.ctypes.data, a.nbytes))

    def get_field(self, kind, slot):
        out = np.empty(self._shape(kind), dtype=self.dtype)
        _check(lib().tc_get_field(self._h, kind, slot, out.ctypes.data, out.nbytes))
        return out

    def device_ptr(self, kind, slot):
        p = ctypes.c_void_p()
        _check(lib().tc_field_device_ptr(self._h, kind, slot, ctypes.byref(p)))
        return p.value

    def sync(self):
        _check(lib().tc_sync(self._h))

    # --- solver -----------------------------------------------------------
    def step(self, t, tau, cur, bdf_order):
        """data::Step (staggered/func_global.h:897-942)."""
        _check(lib().tc_step(self._h, t, tau, cur, bdf_order))

    def srk_step(self, t, tau, vel_slot=0, p_slot=0, dpdt_slot=1):
        """data::Step of the collocated SRK solver (collocated_SRK/func_global.h:317-408),
        in place on the velocity slot and the pressure and dpdt slots (PRES kind)."""
        _check(lib().tc_srk_step(self._h, t, tau, vel_slot, p_slot, dpdt_slot))

    def graphs(self, on=True):
        """Replay repeated steps from captured HIP graphs (tc_set_graphs)."""
        _check(lib().tc_set_graphs(self._h, 1 if on else 0))

    def step_diagnostics(self, on=True):
        """Form the diagnostics of each step's new fields in its projection
        (tc_step_diagnostics); diagnostics() of that slot then reads them. The
        records of the last two steps are kept, and reading one waits only for
        its own step."""
        _check(lib().tc_step_diagnostics(self._h, 1 if on else 0))

    def explicit_term(self, src, dst):
        _check(lib().tc_explicit_term(self._h, src[0], src[1], dst[0], dst[1]))

    def apply_gradient(self, src, dst):
        _check(lib().tc_apply_gradient(self._h, src[0], src[1], dst[0], dst[1]))

    def unapply_gradient(self, src, dst):
        _check(lib().tc_unapply_gradient(self._h, src[0], src[1], dst[0], dst[1]))

    def normalize_pressure(self, f):
        _check(lib().tc_normalize_pressure(self._h, f[0], f[1]))

    def recalc_turb_visc(self, src):
        _check(lib().tc_recalc_turb_visc(self._h, src[0], src[1]))

    def integral(self, f):
        out = np.zeros(3)
        _check(lib().tc_integral(self._h, f[0], f[1], _dptr(out)))
        return out

    def kinetic_energy(self, f):
        out = ctypes.c_double()
        _check(lib().tc_kinetic_energy(self._h, f[0], f[1], ctypes.byref(out)))
        return out.value

    def cfl(self, f, tau, mode=3):
        out = ctypes.c_double()
        _check(lib().tc_cfl(self._h, f[0], f[1], tau, mode, ctypes.byref(out)))
        return out.value

    def profiles(self, vel, p_slot=0):
        nv = 10 if self.scheme in (SCHEME_AB3, SCHEME_SRK) else 12
        out = np.zeros((self.nz, nv))
        _check(lib().tc_profiles(self._h, vel[0], vel[1], p_slot, _dptr(out), out.size))
        return out

    def diagnostics(self, vel_slot, p_slot=None, tau=1.0, cfl_mode=3):
        """Integral, kinetic energy, CFL and profiles of velocity slot vel_slot in
        one pass (tc_diagnostics); returns a dict of the four."""
        nv = 10 if self.scheme in (SCHEME_AB3, SCHEME_SRK) else 12
        prof = np.zeros((self.nz, nv))
        q = np.zeros(3)
        ke, cfl = ctypes.c_double(), ctypes.c_double()
        _check(lib().tc_diagnostics(self._h, vel_slot, vel_slot if p_slot is None else p_slot, tau, cfl_mode, _dptr(q),
                                    ctypes.byref(ke), ctypes.byref(cfl), _dptr(prof), prof.size))
        return dict(integral=q, kinetic_energy=ke.value, cfl=cfl.value, profiles=prof)

    def cfr_correct(self, f):
        s = ctypes.c_double()
        _check(lib().tc_cfr_correct(self._h, f[0], f[1], ctypes.byref(s)))
        return s.value

    def project(self, f):
        _check(lib().tc_project(self._h, f[0], f[1]))

    def initial_field(self, slot, seed=42, ubulk=1.0, noise_rms=0.1):
        _check(lib().tc_initial_field(self._h, slot, seed, ubulk, noise_rms))

    # --- learned closure ---------------------------------------------------
    def set_sgs_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().tc_set_sgs_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))

    def sgs_sample_count(self, factor):
        return lib().tc_sgs_sample_count(self._h, factor)

    def sgs_samples(self, vel, factor, dev_x, dev_t, cap):
        """Writes training samples of field `vel` into device buffers (addresses);
        returns their number."""
        n = ctypes.c_longlong()
        _check(lib().tc_sgs_samples(self._h, vel[0], vel[1], factor, dev_x, dev_t, cap, ctypes.byref(n)))
        return n.value

    # --- timing -------------------------------------------------------------
    def timing(self, on=True):
        _check(lib().tc_timing_enable(self._h, 1 if on else 0))

    def timing_reset(self):
        _check(lib().tc_timing_reset(self._h))

    def timing_read(self):
        res = {}
        for k, name in enumerate(KERNELS):
            ms = ctypes.c_double()
            n = ctypes.c_longlong()
            _check(lib().tc_timing_read(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
            res[name] = (ms.value, n.value)
        return res


def mlp_param_count():
    return lib().tc_mlp_param_count()


class Mlp:
    """Trainer of the learned closure on one GPU (tc_mlp_*). Inputs and targets
    are device buffers given by address: x [n][10], t [n][6] float32."""

    def __init__(self, device=0, seed=0):
        h = ctypes.c_void_p()
        _check(lib().tc_mlp_create(device, seed, ctypes.byref(h)))
        self._h = h
        self.n_params = mlp_param_count()

    def close(self):
        if getattr(self, "_h", None):
            lib().tc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().tc_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().tc_mlp_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size))
        return out

    def forward(self, dev_x, n, dev_y):
        _check(lib().tc_mlp_forward(self._h, dev_x, n, dev_y))

    def gradient(self, dev_x, dev_t, n):
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().tc_mlp_gradient(self._h, dev_x, dev_t, n, g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     g.size, ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, dev_x, dev_t, n, lr=1e-3, no_alloc=False):
        """One Adam step; no_alloc uses tc_mlp_train_step_noalloc (no device allocation)."""
        loss = ctypes.c_double()
        fn = lib().tc_mlp_train_step_noalloc if no_alloc else lib().tc_mlp_train_step
        _check(fn(self._h, dev_x, dev_t, n, lr, ctypes.byref(loss)))
        return loss.value

    def steps(self):
        n = ctypes.c_longlong()
        _check(lib().tc_mlp_steps(self._h, ctypes.byref(n)))
        return n.value
