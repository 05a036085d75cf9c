"""C-ABI checks that need no GPU: the library loads, exports every entry point declared in
include/ovl.h, and its host-only arithmetic matches the reference's golden vectors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from overlap_amd import ovl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ovl.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ovl_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ovl.lib()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert sorted(ovl.SYMBOLS) == names


def test_library_is_gfx950_code_object():
    data = open(ovl.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_zolotarev_coefficients_bitwise_reference():
    f = np.load(os.path.join(ROOT, "tests", "golden", "zolotarev.npz"), allow_pickle=False)
    for (amin, amax), eps, bz, dz in zip(f["ranges"], f["eps"], f["bz"], f["dz"]):
        e, b, d = ovl.zolotarev_coefficients(amin, amax)
        assert e == eps
        assert np.array_equal(b, bz) and np.array_equal(d, dz)


def test_zolotarev_rejects_bad_interval():
    assert np.isnan(ovl.zolotarev_coefficients(0.0, 1.0)[0])
    assert np.isnan(ovl.zolotarev_coefficients(2.0, 1.0)[0])


def test_sign_defaults():
    o = ovl.sign_opts()
    assert o.method == ovl.SIGN_CHEBYSHEV and o.degree == 256 and abs(o.cut - 0.05) < 1e-15 and o.head == -1


def test_create_rejects_bad_params_without_touching_device():
    p = ovl.Params(4, (C.c_int * 3)(8, 8, 8), 16, 1.0, 0.0, 0.0, 0, 0, 0)
    h = C.c_void_p()
    assert ovl.lib().ovl_create(C.byref(p), C.byref(h)) == ovl.E_ARG
    p = ovl.Params(2, (C.c_int * 3)(8, 2, 8), 16, 1.0, 0.0, 0.0, 0, 0, 0)
    assert ovl.lib().ovl_create(C.byref(p), C.byref(h)) == ovl.E_ARG
    p = ovl.Params(2, (C.c_int * 3)(8, 8, 8), 16, 1.0, 0.0, 0.0, 48, 0, 0)
    assert ovl.lib().ovl_create(C.byref(p), C.byref(h)) == ovl.E_ARG


def test_create_ensemble_rejects_bad_params_without_touching_device():
    def create(nc, dims, nmodes=8, mb=0):
        flat = (C.c_int * max(1, 3 * len(dims)))(*[v for d in dims for v in d])
        p = ovl.EnsembleParams(nc, len(dims), C.cast(flat, C.POINTER(C.c_int)), nmodes, 1.0, 0.0, 1.0, mb, 0)
        h = C.c_void_p()
        return ovl.lib().ovl_create_ensemble(C.byref(p), C.byref(h))
    assert create(4, [(4, 4, 4)]) == ovl.E_ARG
    assert create(2, []) == ovl.E_ARG
    assert create(2, [(4, 4, 4), (4, 2, 4)]) == ovl.E_ARG
    assert create(2, [(4, 4, 4)], nmodes=0) == ovl.E_ARG
    assert create(2, [(4, 4, 4)], mb=24) == ovl.E_ARG
    p = ovl.Params(2, (C.c_int * 3)(8, 8, 8), 16, 1.0, 0.0, 0.0, 0, 0, 4)  # traversal 4 does not exist
    assert ovl.lib().ovl_create(C.byref(p), C.byref(C.c_void_p())) == ovl.E_ARG


@pytest.mark.skipif(bool(os.environ.get("HIP_VISIBLE_DEVICES")) or os.path.exists("/dev/kfd"),
                    reason="a GPU may be present")
def test_create_reports_missing_device():
    with pytest.raises(ovl.OvlError) as e:
        ovl.Context(2, (4, 4, 4), 4)
    assert e.value.code == ovl.E_HIP
