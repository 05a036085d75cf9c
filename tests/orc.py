"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was built, the compiled
reference hot path (oracle/_ref/libovlref.so).  Test infrastructure: only tests, smoke() and
bench.py's cpu_baseline leg use this module."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libovlref.so")

_D = C.POINTER(C.c_double)


class Lat(C.Structure):
    _fields_ = [("nc", C.c_int), ("n", C.c_int * 3), ("r_wilson", C.c_double),
                ("m_fermion", C.c_double), ("m_dw", C.c_double)]


class StepOpts(C.Structure):
    _fields_ = [("kind", C.c_int), ("dt", C.c_double), ("update_links", C.c_int), ("nb", C.c_int),
                ("q", C.c_double), ("cheb_n", C.c_int), ("cheb_cut", C.c_double),
                ("zolo_lmin", C.c_double), ("zolo_lmax", C.c_double), ("threads", C.c_int)]


WILSON, OVERLAP_CHEB, OVERLAP_ZOLO = 0, 1, 2


def dp(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
        _lib = C.CDLL(ORACLE_SO)
        _lib.orc_hw_norm_bound.restype = C.c_double
        _lib.orc_zolotarev.restype = C.c_double
        _lib.orc_ritz.restype = C.c_double
        _lib.orc_edotb.restype = C.c_double
        _lib.orc_gauss_restore_step.restype = C.c_double
        _lib.orc_zolotrr.restype = C.c_int
        _lib.orc_overlap_evo_zolo.restype = C.c_int
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_zolotarev.restype = C.c_double
        _ref.ref_ritz.restype = C.c_double
        _ref.ref_su2_unitarity_norm.restype = C.c_double
    return _ref


def lat(nc, n, r_wilson=1.0, m_fermion=0.0, m_dw=0.0) -> Lat:
    return Lat(nc, (C.c_int * 3)(*n), r_wilson, m_fermion, m_dw)


def cz(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def u1_phases(n, nb, q):
    V = int(np.prod(n))
    out = np.zeros(3 * V, complex)
    lib().orc_u1_phases((C.c_int * 3)(*n), C.c_int(nb), C.c_double(q), dp(out))
    return out


def links_to_ucc(L: Lat, links, phase=None):
    V = L.n[0] * L.n[1] * L.n[2]
    out = np.zeros(3 * V * L.nc * L.nc, complex)
    lib().orc_links_to_ucc(C.byref(L), dp(np.ascontiguousarray(links, np.float64)),
                           dp(phase) if phase is not None else None, dp(out))
    return out


def wilson_evo(L: Lat, ucc, phi):
    out = np.zeros_like(cz(phi))
    lib().orc_wilson_evo(C.byref(L), dp(cz(ucc)), dp(cz(phi)), dp(out))
    return out


def hw(L: Lat, ucc, phi):
    out = np.zeros_like(cz(phi))
    lib().orc_hw(C.byref(L), dp(cz(ucc)), dp(cz(phi)), dp(out))
    return out


def norm_bound(L: Lat) -> float:
    return lib().orc_hw_norm_bound(C.byref(L))


def cheb_coeffs(n, cut):
    c = np.zeros(n + 1)
    lib().orc_cheb_coeffs(C.c_int(n), C.c_double(cut), dp(c))
    return c


def sign_cheb(L: Lat, ucc, lam, cut, n, coef, phi):
    out = np.zeros_like(cz(phi))
    lib().orc_sign_cheb(C.byref(L), dp(cz(ucc)), C.c_double(lam), C.c_double(cut), C.c_int(n),
                        dp(np.ascontiguousarray(coef, np.float64)), dp(cz(phi)), dp(out))
    return out


def overlap_evo_cheb(L: Lat, ucc, lam, cut, n, coef, phi):
    out = np.zeros_like(cz(phi))
    lib().orc_overlap_evo_cheb(C.byref(L), dp(cz(ucc)), C.c_double(lam), C.c_double(cut), C.c_int(n),
                               dp(np.ascontiguousarray(coef, np.float64)), dp(cz(phi)), dp(out))
    return out


def zolotarev(amin, amax, errtol=1e-12, no=25):
    bz, dz = np.zeros(no), np.zeros(no)
    eps = lib().orc_zolotarev(C.c_double(amin), C.c_double(amax), C.c_double(errtol), C.c_int(no), dp(bz), dp(dz))
    return eps, bz, dz


def overlap_evo_zolo(L: Lat, ucc, phi, lmin, lmax):
    out = np.zeros_like(cz(phi))
    it = lib().orc_overlap_evo_zolo(C.byref(L), dp(cz(ucc)), dp(cz(phi)), dp(out), C.c_double(lmin), C.c_double(lmax))
    return out, it


def ritz(L: Lat, ucc, vec_flat, delitr=1e-16, nitlz=2000, sign=1.0, nrestart=100):
    return lib().orc_ritz(C.byref(L), dp(cz(ucc)), dp(cz(vec_flat)), C.c_double(delitr), C.c_int(nitlz),
                          C.c_double(sign), C.c_int(nrestart))


def su2_iexp(c, alpha):
    out = np.zeros(4)
    lib().orc_su2_iexp(C.c_double(c), dp(np.ascontiguousarray(alpha, np.float64)), dp(out))
    return out


def su3_iexp(c, alpha):
    out = np.zeros(18)
    lib().orc_su3_iexp(C.c_double(c), dp(np.ascontiguousarray(alpha, np.float64)), dp(out))
    return out


def update_links(L: Lat, links, e, dt):
    links = np.array(links, np.float64)
    lib().orc_update_links(C.byref(L), dp(links), dp(np.ascontiguousarray(e, np.float64)), C.c_double(dt))
    return links


def reunitarize(L: Lat, links):
    links = np.array(links, np.float64)
    lib().orc_reunitarize(C.byref(L), dp(links))
    return links


def leapfrog(L: Lat, links, e, psi, psimid, opts: StepOpts, nsteps):
    """Returns (links, psi, psimid, efield) after nsteps."""
    links = np.array(links, np.float64)
    e = np.array(e, np.float64)
    psi = cz(psi).copy()
    psimid = cz(psimid).copy()
    nmodes = psi.shape[0] if psi.ndim > 1 else 1
    lib().orc_leapfrog(C.byref(L), dp(links), dp(e), C.c_int(nmodes),
                       dp(psi), dp(psimid), C.byref(opts), C.c_int(nsteps))
    return links, psi, psimid, e


def plaquette_energy(L: Lat, links, e):
    a, b = C.c_double(), C.c_double()
    lib().orc_plaquette_energy(C.byref(L), dp(np.ascontiguousarray(links, np.float64)),
                               dp(np.ascontiguousarray(e, np.float64)), C.byref(a), C.byref(b))
    return a.value, b.value


def gauss_violation(L: Lat, links, e):
    a, b = C.c_double(), C.c_double()
    lib().orc_gauss_violation(C.byref(L), dp(np.ascontiguousarray(links, np.float64)),
                              dp(np.ascontiguousarray(e, np.float64)), C.byref(a), C.byref(b))
    return a.value, b.value


def gauss_restore(L: Lat, links, e, tol, max_iter):
    """Returns (E after the restoration, iterations, final max |G^a|)."""
    e = np.array(e, np.float64)
    vmax = C.c_double()
    it = lib().orc_gauss_restore(C.byref(L), dp(np.ascontiguousarray(links, np.float64)), dp(e), C.c_double(tol),
                                 C.c_int(max_iter), C.byref(vmax))
    return e, it, vmax.value


def gauss_restore_step(L: Lat, links, e, gamma):
    """Returns (E after one relaxation step, max |G^a| before it)."""
    e = np.array(e, np.float64)
    m = lib().orc_gauss_restore_step(C.byref(L), dp(np.ascontiguousarray(links, np.float64)), dp(e),
                                     C.c_double(gamma))
    return e, m


def unitarity_violation(L: Lat, links):
    a, b = C.c_double(), C.c_double()
    lib().orc_unitarity_violation(C.byref(L), dp(np.ascontiguousarray(links, np.float64)), C.byref(a), C.byref(b))
    return a.value, b.value


def edotb(L: Lat, links, e):
    return lib().orc_edotb(C.byref(L), dp(np.ascontiguousarray(links, np.float64)),
                           dp(np.ascontiguousarray(e, np.float64)))


def currents(L: Lat, psi, hpsi, first_mode=0):
    V = L.n[0] * L.n[1] * L.n[2]
    jv, ja = np.zeros(4 * V), np.zeros(4 * V)
    psi, hpsi = cz(psi), cz(hpsi)
    nm = psi.shape[0] if psi.ndim > 1 else 1
    lib().orc_currents(C.byref(L), dp(psi), dp(hpsi), C.c_int(nm), C.c_int(first_mode), dp(jv), dp(ja))
    return jv, ja


# ---- compiled reference (SU(2) only) ----

def ref_set(L: Lat):
    ref().ref_set_params((C.c_int * 3)(*L.n), C.c_double(L.r_wilson), C.c_double(L.m_fermion), C.c_double(L.m_dw))


def ref_wilson_evo(L: Lat, ucc, phi):
    ref_set(L)
    out = np.zeros_like(cz(phi))
    ref().ref_wilson_evo(dp(cz(ucc)), dp(cz(phi)), dp(out))
    return out


def ref_zolotarev(amin, amax, errtol=1e-12, no=25):
    bz, dz = np.zeros(no), np.zeros(no)
    eps = ref().ref_zolotarev(C.c_double(amin), C.c_double(amax), C.c_double(errtol), C.c_int(no), dp(bz), dp(dz))
    return eps, bz, dz


def ref_overlap_evo(L: Lat, ucc, phi, lmin, lmax):
    ref_set(L)
    out = np.zeros_like(cz(phi))
    ref().ref_overlap_evo(dp(cz(ucc)), dp(cz(phi)), dp(out), C.c_double(lmin), C.c_double(lmax))
    return out


def ref_ritz(L: Lat, ucc, vec_flat, delitr=1e-16, nitlz=2000, sign=1.0, nrestart=100):
    ref_set(L)
    return ref().ref_ritz(dp(cz(ucc)), dp(cz(vec_flat)), C.c_double(delitr), C.c_int(nitlz), C.c_double(sign),
                          C.c_int(nrestart))


def ref_su2_iexp(c, alpha):
    out = np.zeros(4)
    ref().ref_su2_iexp(C.c_double(c), dp(np.ascontiguousarray(alpha, np.float64)), dp(out))
    return out


def ref_su2_uu(a, b):
    out = np.zeros(4)
    ref().ref_su2_uu(dp(np.ascontiguousarray(a, np.float64)), dp(np.ascontiguousarray(b, np.float64)), dp(out))
    return out


def aos_to_flat(nc, V, psi):
    """AoS [site][colour][spin] -> flat m = site + V(spin + 4 colour) (OverlapU.cpp:163)."""
    a = cz(psi).reshape(V, nc, 4)
    return np.ascontiguousarray(np.transpose(a, (1, 2, 0)).reshape(-1))


def efield_update(L: Lat, links, e, dt):
    e = np.array(e, np.float64)
    lib().orc_efield_update(C.byref(L), dp(np.ascontiguousarray(links, np.float64)), dp(e), C.c_double(dt))
    return e
