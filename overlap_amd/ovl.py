"""ctypes binding of the C ABI in include/ovl.h (libovl.so, built for gfx950).

There is no CPU fallback: if libovl.so is missing or no HIP device is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libovl.so")

OK, E_ARG, E_HIP, E_STATE, E_CONVERGE, E_NOMEM, E_SPECTRUM, E_NONFINITE = 0, -1, -2, -3, -4, -5, -6, -7
PSI, PSIMID, SCRATCH = 0, 1, 2
SIGN_CHEBYSHEV, SIGN_ZOLOTAREV = 0, 1
PREC_FP64, PREC_FP32, PREC_MIXED = 0, 1, 2
WILSON, OVERLAP = 0, 1
GAUGE_FROZEN, GAUGE_LINKS, GAUGE_FULL = 0, 1, 2

# exported symbols, in the order of include/ovl.h
SYMBOLS = [
    "ovl_create", "ovl_destroy", "ovl_error", "ovl_device_name", "ovl_sign_defaults", "ovl_sign_plan",
    "ovl_set_links", "ovl_get_links", "ovl_set_efield", "ovl_get_efield", "ovl_set_u1_flux",
    "ovl_set_modes", "ovl_get_modes", "ovl_copy_field",
    "ovl_apply_wilson", "ovl_apply_overlap", "ovl_apply_sign", "ovl_sign_residual", "ovl_ritz",
    "ovl_wilson_evo_compute", "ovl_overlap_operator_compute", "ovl_zolotarev_coefficients",
    "ovl_update_links", "ovl_update_efield", "ovl_fermion_update",
    "ovl_energy", "ovl_gauss_violation", "ovl_unitarity_violation", "ovl_ncs_dot", "ovl_currents",
    "ovl_fermion_update_observed", "ovl_mode_norms",
    "ovl_create_ensemble", "ovl_set_system_links", "ovl_get_system_links", "ovl_set_system_efield",
    "ovl_get_system_efield", "ovl_set_system_u1_flux", "ovl_set_system_modes", "ovl_get_system_modes",
    "ovl_system_observables", "ovl_observables_noalloc", "ovl_ensemble_balance", "ovl_gauss_restore", "ovl_qed_eigenbasis",
    "ovl_save_configuration", "ovl_load_configuration",
    "ovl_get_timing", "ovl_set_profiling", "ovl_autotune", "ovl_get_tuning", "ovl_set_tuning",
]
OBS_NAMES = ("e_el", "e_mag", "gauss_max", "gauss_rms", "unit_max", "unit_rms", "ncs_dot")


class Params(C.Structure):
    _fields_ = [("nc", C.c_int), ("n", C.c_int * 3), ("n_modes", C.c_int), ("r_wilson", C.c_double),
                ("m_fermion", C.c_double), ("m_dw", C.c_double), ("mode_block", C.c_int),
                ("reunitarize", C.c_int), ("traversal", C.c_int)]


class EnsembleParams(C.Structure):
    _fields_ = [("nc", C.c_int), ("n_systems", C.c_int), ("dims", C.POINTER(C.c_int)), ("n_modes", C.c_int),
                ("r_wilson", C.c_double), ("m_fermion", C.c_double), ("m_dw", C.c_double),
                ("mode_block", C.c_int), ("reunitarize", C.c_int)]


class SignOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("degree", C.c_int), ("cut", C.c_double),
                ("tol", C.c_double), ("head", C.c_int), ("lmin", C.c_double), ("lmax", C.c_double),
                ("mode_chunk", C.c_int), ("fp16_tail", C.c_int), ("head16", C.c_int),
                ("fp48", C.c_int), ("head48", C.c_int), ("guard", C.c_int), ("relax", C.c_int)]


class StepOpts(C.Structure):
    _fields_ = [("kind", C.c_int), ("dt", C.c_double), ("gauge", C.c_int), ("sign", SignOpts),
                ("use_graph", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("traversal", C.c_int), ("slabs", C.c_int), ("nt_store", C.c_int), ("waves_per_eu", C.c_int)]


class Observe(C.Structure):
    _fields_ = [("every", C.c_int), ("first_mode", C.c_int), ("count", C.c_int),
                ("axial_charge", C.POINTER(C.c_double)), ("ncs_dot", C.POINTER(C.c_double)),
                ("jv", C.POINTER(C.c_double)), ("ja", C.POINTER(C.c_double))]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("hop_ms", C.c_double), ("hop_launches", C.c_longlong),
                ("hop_bytes", C.c_double), ("fp64_steps", C.c_int), ("fp32_steps", C.c_int),
                ("fp16_steps", C.c_int), ("fp48_steps", C.c_int), ("var_ms", C.c_double * 5),
                ("var_launches", C.c_longlong * 5), ("var_bytes", C.c_double * 5), ("graph_captures", C.c_int),
                ("guard_checks", C.c_int), ("guard_level", C.c_int), ("guard_err", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.ovl_error.restype = C.c_char_p
        _lib.ovl_zolotarev_coefficients.restype = C.c_double
        _lib.ovl_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        _lib.ovl_create_ensemble.argtypes = [C.POINTER(EnsembleParams), C.POINTER(C.c_void_p)]
        for name in SYMBOLS:
            f = getattr(_lib, name)
            if name not in ("ovl_error", "ovl_zolotarev_coefficients", "ovl_sign_defaults"):
                f.restype = C.c_int
        _lib.ovl_sign_defaults.restype = None
        _lib.ovl_observables_noalloc.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]
    return _lib


class OvlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ovl error {code}: {msg}")
        self.code = code


def _d(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def sign_opts(**kw) -> SignOpts:
    o = SignOpts()
    lib().ovl_sign_defaults(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def sign_plan(opts: SignOpts):
    """(head48, head, head16): steps k < head48 run in FP64, head48 <= k < head in 48-bit storage,
    head <= k < head16 in FP32, the rest in FP16 storage."""
    h48, h, h16 = C.c_int(), C.c_int(), C.c_int()
    rc = lib().ovl_sign_plan(C.byref(opts), C.byref(h48), C.byref(h), C.byref(h16))
    if rc != OK:
        raise OvlError(rc, "invalid sign options")
    return h48.value, h.value, h16.value


def zolotarev_coefficients(amin, amax, errtol=1e-12, no=25):
    bz, dz = np.zeros(no), np.zeros(no)
    eps = lib().ovl_zolotarev_coefficients(C.c_double(amin), C.c_double(amax), C.c_double(errtol), C.c_int(no),
                                           _d(bz), _d(dz))
    return eps, bz, dz


def observables_noalloc(nc, n, links_dev: int, efield_dev: int, obs_dev: int, stream: int = 0) -> None:
    """ovl_observables_noalloc: device addresses of the links and E (reference layouts) and of
    OVL_OBS_COUNT output doubles; queued on `stream` (0 = default stream), no device allocation."""
    dims = (C.c_int * 3)(*n)
    rc = lib().ovl_observables_noalloc(int(nc), dims, C.c_void_p(links_dev), C.c_void_p(efield_dev),
                                       C.c_void_p(obs_dev), C.c_void_p(stream))
    if rc != OK:
        raise OvlError(rc, lib().ovl_error(None).decode())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    rc = lib().ovl_device_name(buf, 256)
    if rc != OK:
        raise OvlError(rc, "no HIP device")
    return buf.value.decode()


class Context:
    """One device context: gauge field, two mode-function fields (Psi, PsiMid) and scratch."""

    def __init__(self, nc, n, n_modes, r_wilson=1.0, m_fermion=0.0, m_dw=0.0, mode_block=0, reunitarize=0,
                 traversal=0):
        self.nc, self.n, self.n_modes = nc, tuple(n), n_modes
        self.V = int(np.prod(n))
        self.dims, self.Vs = [self.n], [self.V]
        p = Params(nc, (C.c_int * 3)(*n), n_modes, r_wilson, m_fermion, m_dw, mode_block, reunitarize, traversal)
        h = C.c_void_p()
        rc = lib().ovl_create(C.byref(p), C.byref(h))
        if rc != OK:
            raise OvlError(rc, "ovl_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ovl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, allow=()):
        if rc != OK and rc not in allow:
            raise OvlError(rc, lib().ovl_error(self.h).decode())
        return rc

    # gauge
    def set_links(self, links):
        self._chk(lib().ovl_set_links(self.h, _d(np.ascontiguousarray(links, np.float64))))

    def get_links(self):
        out = np.zeros(self.V * 3 * (4 if self.nc == 2 else 18))
        self._chk(lib().ovl_get_links(self.h, _d(out)))
        return out

    def set_efield(self, e):
        self._chk(lib().ovl_set_efield(self.h, _d(np.ascontiguousarray(e, np.float64))))

    def get_efield(self):
        out = np.zeros(self.V * 3 * (self.nc * self.nc - 1))
        self._chk(lib().ovl_get_efield(self.h, _d(out)))
        return out

    def set_u1_flux(self, nb, q=1.0):
        self._chk(lib().ovl_set_u1_flux(self.h, C.c_int(nb), C.c_double(q)))

    # modes: complex arrays (count, V*nc*4) in the reference AoS layout
    def set_modes(self, field, psi, first=0):
        psi = np.ascontiguousarray(np.atleast_2d(psi), np.complex128)
        self._chk(lib().ovl_set_modes(self.h, field, psi.ctypes.data_as(C.POINTER(C.c_double)), first, psi.shape[0]))

    def get_modes(self, field, first=0, count=None):
        count = self.n_modes - first if count is None else count
        out = np.zeros((count, self.V * self.nc * 4), np.complex128)
        self._chk(lib().ovl_get_modes(self.h, field, out.ctypes.data_as(C.POINTER(C.c_double)), first, count))
        return out

    def copy_field(self, src, dst):
        self._chk(lib().ovl_copy_field(self.h, src, dst))

    # operators
    def apply_wilson(self, fin, fout):
        self._chk(lib().ovl_apply_wilson(self.h, fin, fout))

    def apply_overlap(self, fin, fout, opts: SignOpts):
        return self._chk(lib().ovl_apply_overlap(self.h, fin, fout, C.byref(opts)), allow=(E_CONVERGE,))

    def apply_sign(self, fin, fout, opts: SignOpts):
        self._chk(lib().ovl_apply_sign(self.h, fin, fout, C.byref(opts)))

    def sign_residual(self, fin, opts: SignOpts) -> float:
        r = C.c_double()
        self._chk(lib().ovl_sign_residual(self.h, fin, C.byref(opts), C.byref(r)))
        return r.value

    def ritz(self, start_flat, sign, delitr=1e-16, nitlz=2000, nrestart=100) -> float:
        v = np.ascontiguousarray(start_flat, np.complex128)
        r = C.c_double()
        self._chk(lib().ovl_ritz(self.h, v.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(sign),
                                 C.c_double(delitr), nitlz, nrestart, C.byref(r)))
        return r.value

    def wilson_evo_compute(self, ucc, phi):
        ucc = np.ascontiguousarray(ucc, np.complex128)
        phi = np.ascontiguousarray(phi, np.complex128)
        out = np.zeros_like(phi)
        self._chk(lib().ovl_wilson_evo_compute(self.h, ucc.ctypes.data_as(C.POINTER(C.c_double)),
                                               phi.ctypes.data_as(C.POINTER(C.c_double)),
                                               out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def overlap_operator_compute(self, ucc, phi, lmin, lmax):
        ucc = np.ascontiguousarray(ucc, np.complex128)
        phi = np.ascontiguousarray(phi, np.complex128)
        out = np.zeros_like(phi)
        self._chk(lib().ovl_overlap_operator_compute(self.h, ucc.ctypes.data_as(C.POINTER(C.c_double)),
                                                     phi.ctypes.data_as(C.POINTER(C.c_double)),
                                                     out.ctypes.data_as(C.POINTER(C.c_double)),
                                                     C.c_double(lmin), C.c_double(lmax)), allow=(E_CONVERGE,))
        return out

    # dynamics
    def update_links(self, dt):
        self._chk(lib().ovl_update_links(self.h, C.c_double(dt)))

    def update_efield(self, dt):
        self._chk(lib().ovl_update_efield(self.h, C.c_double(dt)))

    def fermion_update(self, opts: StepOpts, nsteps):
        return self._chk(lib().ovl_fermion_update(self.h, C.byref(opts), nsteps), allow=(E_CONVERGE,))

    def fermion_update_observed(self, opts: StepOpts, nsteps, every, first_mode=0):
        """Returns (axial charge history, dN_CS/dt history, jv, ja of the last observation)."""
        n = nsteps // every if every > 0 else 0
        q5, ncs = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        jv, ja = np.zeros(4 * self.V), np.zeros(4 * self.V)
        ob = Observe(every, first_mode, 0, _d(q5), _d(ncs), _d(jv), _d(ja))
        self._chk(lib().ovl_fermion_update_observed(self.h, C.byref(opts), nsteps, C.byref(ob)), allow=(E_CONVERGE,))
        return q5[:ob.count], ncs[:ob.count], jv, ja

    # observables
    def _pair(self, fn):
        a, b = C.c_double(), C.c_double()
        self._chk(fn(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def energy(self):
        return self._pair(lib().ovl_energy)

    def gauss_violation(self):
        return self._pair(lib().ovl_gauss_violation)

    def unitarity_violation(self):
        return self._pair(lib().ovl_unitarity_violation)

    def ncs_dot(self):
        r = C.c_double()
        self._chk(lib().ovl_ncs_dot(self.h, C.byref(r)))
        return r.value

    def currents(self, field, opts: SignOpts, first_mode=0):
        jv, ja = np.zeros(4 * self.V), np.zeros(4 * self.V)
        q = C.c_double()
        self._chk(lib().ovl_currents(self.h, field, C.byref(opts), first_mode, _d(jv), _d(ja), C.byref(q)))
        return jv, ja, q.value

    def mode_norms(self, field):
        out = np.zeros(self.n_modes)
        self._chk(lib().ovl_mode_norms(self.h, field, _d(out)))
        return out

    def timing(self) -> Timing:
        t = Timing()
        self._chk(lib().ovl_get_timing(self.h, C.byref(t)))
        return t

    def set_profiling(self, on: bool):
        self._chk(lib().ovl_set_profiling(self.h, int(on)))

    def get_tuning(self, precision) -> Tuning:
        t = Tuning()
        self._chk(lib().ovl_get_tuning(self.h, precision, C.byref(t)))
        return t

    def set_tuning(self, precision, traversal, slabs=1, nt_store=1, waves_per_eu=1):
        self._chk(lib().ovl_set_tuning(self.h, precision, C.byref(Tuning(traversal, slabs, nt_store, waves_per_eu))))

    def autotune(self) -> str:
        buf = C.create_string_buffer(16384)
        self._chk(lib().ovl_autotune(self.h, buf, 16384))
        return buf.value.decode()

    # per-system calls (system 0 of an ordinary context)
    def _sys_volume(self, s):
        if not 0 <= s < len(self.Vs):
            raise OvlError(E_ARG, "system index out of range")
        return self.Vs[s]

    def set_system_links(self, s, links):
        self._chk(lib().ovl_set_system_links(self.h, s, _d(np.ascontiguousarray(links, np.float64))))

    def get_system_links(self, s):
        out = np.zeros(self._sys_volume(s) * 3 * (4 if self.nc == 2 else 18))
        self._chk(lib().ovl_get_system_links(self.h, s, _d(out)))
        return out

    def set_system_efield(self, s, e):
        self._chk(lib().ovl_set_system_efield(self.h, s, _d(np.ascontiguousarray(e, np.float64))))

    def get_system_efield(self, s):
        out = np.zeros(self._sys_volume(s) * 3 * (self.nc * self.nc - 1))
        self._chk(lib().ovl_get_system_efield(self.h, s, _d(out)))
        return out

    def set_system_u1_flux(self, s, nb, q=1.0):
        self._chk(lib().ovl_set_system_u1_flux(self.h, s, C.c_int(nb), C.c_double(q)))

    def set_system_modes(self, s, field, psi, first=0):
        psi = np.ascontiguousarray(np.atleast_2d(psi), np.complex128)
        self._chk(lib().ovl_set_system_modes(self.h, s, field, psi.ctypes.data_as(C.POINTER(C.c_double)), first,
                                             psi.shape[0]))

    def get_system_modes(self, s, field, first=0, count=None):
        count = self.n_modes - first if count is None else count
        out = np.zeros((count, self._sys_volume(s) * self.nc * 4), np.complex128)
        self._chk(lib().ovl_get_system_modes(self.h, s, field, out.ctypes.data_as(C.POINTER(C.c_double)), first,
                                             count))
        return out

    def qed_eigenbasis(self, opts: SignOpts, nb, q=1.0, dt=0.02, first_mode=0):
        """Sets Psi/PsiMid from the overlap eigenbasis of the U(1) background; returns (eigenvalues
        in descending order, max |H - H^+| of the assembled matrix)."""
        D = 4 * self.nc * self.V
        ev = np.zeros(D)
        herm = C.c_double()
        self._chk(lib().ovl_qed_eigenbasis(self.h, C.byref(opts), nb, C.c_double(q), C.c_double(dt), first_mode,
                                           _d(ev), C.byref(herm)))
        return ev, herm.value

    def gauss_restore(self, tol=1e-12, max_iter=10000, s=0):
        """Returns (iterations, max |G^a| of the final field)."""
        it, mx = C.c_int(), C.c_double()
        self._chk(lib().ovl_gauss_restore(self.h, s, C.c_double(tol), max_iter, C.byref(it), C.byref(mx)))
        return it.value, mx.value

    def save_configuration(self, ufile, efile=None, s=0):
        self._chk(lib().ovl_save_configuration(self.h, s, os.fsencode(ufile),
                                               None if efile is None else os.fsencode(efile)))

    def load_configuration(self, ufile, efile=None, s=0):
        self._chk(lib().ovl_load_configuration(self.h, s, os.fsencode(ufile),
                                               None if efile is None else os.fsencode(efile)))

    def system_observables(self, s) -> dict:
        out = np.zeros(len(OBS_NAMES))
        self._chk(lib().ovl_system_observables(self.h, s, _d(out)))
        return dict(zip(OBS_NAMES, out.tolist()))


class Ensemble(Context):
    """Many independent lattices (systems) of different sizes in one device context.

    Every stencil application runs as one kernel launch over all systems; links, E fields, U(1)
    flux and modes are set and read per system.
    """

    def __init__(self, nc, dims, n_modes, r_wilson=1.0, m_fermion=0.0, m_dw=0.0, mode_block=0, reunitarize=0):
        self.nc, self.n_modes = nc, n_modes
        self.dims = [tuple(int(v) for v in d) for d in dims]
        self.Vs = [int(np.prod(d)) for d in self.dims]
        self.V = sum(self.Vs)
        flat = (C.c_int * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        p = EnsembleParams(nc, len(self.dims), C.cast(flat, C.POINTER(C.c_int)), n_modes, r_wilson, m_fermion, m_dw,
                           mode_block, reunitarize)
        h = C.c_void_p()
        rc = lib().ovl_create_ensemble(C.byref(p), C.byref(h))
        if rc != OK:
            raise OvlError(rc, "ovl_create_ensemble failed")
        self.h = h

    @property
    def n_systems(self):
        return len(self.dims)

    def balance(self) -> float:
        r = C.c_double()
        self._chk(lib().ovl_ensemble_balance(self.h, C.byref(r)))
        return r.value

    def mode_norms(self, field):
        out = np.zeros(self.n_systems * self.n_modes)
        self._chk(lib().ovl_mode_norms(self.h, field, _d(out)))
        return out.reshape(self.n_systems, self.n_modes)
