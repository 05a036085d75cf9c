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


def test_create_rejects_sizes_beyond_the_lane_limits():
    # at most 2^26 sites, and (mode blocks) x sites x (block size) lanes below 2^31, both checked
    # before any device call
    h = C.c_void_p()
    p = ovl.Params(2, (C.c_int * 3)(512, 512, 512), 16, 1.0, 0.0, 0.0, 16, 0, 0)  # 2^27 sites
    assert ovl.lib().ovl_create(C.byref(p), C.byref(h)) == ovl.E_ARG and not h.value
    p = ovl.Params(3, (C.c_int * 3)(256, 256, 256), 256, 1.0, 0.0, 0.0, 64, 0, 0)  # 2^24 sites x 256 modes
    assert ovl.lib().ovl_create(C.byref(p), C.byref(h)) == ovl.E_ARG and not h.value


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


def test_noalloc_observables_reject_bad_shapes_without_touching_device():
    # the argument checks come before any HIP call; the (never dereferenced) addresses are dummies
    dummy = 256

    def call(nc, n, ptr=dummy):
        return ovl.lib().ovl_observables_noalloc(nc, (C.c_int * 3)(*n), C.c_void_p(ptr), C.c_void_p(ptr),
                                                 C.c_void_p(ptr), None)
    assert call(4, (4, 4, 4)) == ovl.E_ARG
    assert b"nc must be 2 or 3" in ovl.lib().ovl_error(None)
    assert call(2, (4, 4, 4), ptr=0) == ovl.E_ARG
    assert call(3, (4, 2, 4)) == ovl.E_ARG
    assert b">= 3" in ovl.lib().ovl_error(None)
    assert call(2, (1024, 1024, 128)) == ovl.E_ARG  # 2^27 sites, beyond the 2^26 maximum
    assert b"exceed the maximum of 67108864" in ovl.lib().ovl_error(None)
    with pytest.raises(ovl.OvlError, match="maximum"):
        ovl.observables_noalloc(3, (512, 512, 512), dummy, dummy, dummy)


@pytest.mark.skipif(bool(os.environ.get("HIP_VISIBLE_DEVICES")) or os.path.exists("/dev/kfd"),
                    reason="a GPU may be present")
def test_create_reports_missing_device():
    with pytest.raises(ovl.OvlError) as e:
        ovl.Context(2, (4, 4, 4), 4)
    assert e.value.code == ovl.E_HIP


def _tier_starts_numpy(deg, cut, tol, fp16, fp48=True):
    # the precision ladder restated in numpy: B_k = max_t |b_k(t)| of the scalar Clenshaw
    # recurrence of the Chebyshev coefficients (as orc.cheb_coeffs computes them) on t in [-1, 1];
    # a tier with unit roundoff u starts at the smallest k whose whole tail has kappa u B_k <= its
    # share of tol: 1/4 each to the FP16 and 48-bit tiers when they are on, the rest to FP32;
    # kappa = 1 for FP32 and FP16, 0.1 for the 48-bit layout
    import orc
    c = orc.cheb_coeffs(deg, cut)
    t = np.cos(np.pi * np.arange(2049) / 2048)
    b1 = np.zeros_like(t)
    b2 = np.zeros_like(t)
    B = np.zeros(deg + 2)
    for k in range(deg, 0, -1):
        bk = c[k] + 2 * t * b1 - b2
        b2, b1 = b1, bk
        B[k] = np.abs(bk).max()

    def start(u, share, kappa=1.0):
        first = deg + 1
        for k in range(deg, 0, -1):
            if kappa * u * B[k] > share:
                break
            first = k
        return first
    s16, s48 = (0.25 if fp16 else 0.0), (0.25 if fp48 else 0.0)
    head = start(2.0 ** -24, (1 - s16 - s48) * tol)
    head16 = min(max(start(2.0 ** -11, s16 * tol), head + 1), deg + 1) if fp16 else deg + 1
    head48 = min(max(start(2.0 ** -37, s48 * tol, 0.1), 1), head) if fp48 else head
    return head48, head, head16


def test_sign_plan_host_only():
    o = ovl.sign_opts(degree=384, cut=0.05, precision=ovl.PREC_MIXED, tol=1e-10)
    assert ovl.sign_plan(o) == (12, 110, 204)  # the ladder of the 24^3 SU(3) bench workload at tol
    assert ovl.sign_plan(o) == _tier_starts_numpy(384, 0.05, 1e-10, True)
    # the accuracy guard's loosened level -1 plans for 4 tol with the head overrides dropped: the
    # ladder the bench runs (FP64 1 step, 48-bit 96, FP32 94, FP16 194 of 385)
    o4 = ovl.sign_opts(degree=384, cut=0.05, precision=ovl.PREC_MIXED, tol=4e-10)
    assert ovl.sign_plan(o4) == (1, 97, 191) == _tier_starts_numpy(384, 0.05, 4e-10, True)
    o.fp48 = 0
    assert ovl.sign_plan(o) == (106, 106, 204) == _tier_starts_numpy(384, 0.05, 1e-10, True, False)
    o.fp16_tail = 0
    assert ovl.sign_plan(o) == (103, 103, 385) == _tier_starts_numpy(384, 0.05, 1e-10, False, False)
    o.fp48 = 1
    assert ovl.sign_plan(o) == _tier_starts_numpy(384, 0.05, 1e-10, False, True)
    o = ovl.sign_opts(degree=256, cut=0.05, precision=ovl.PREC_MIXED, tol=1e-11)
    assert ovl.sign_plan(o) == _tier_starts_numpy(256, 0.05, 1e-11, True)
    assert ovl.sign_plan(ovl.sign_opts(degree=64, precision=ovl.PREC_FP64)) == (65, 65, 65)
    assert ovl.sign_plan(ovl.sign_opts(degree=64, precision=ovl.PREC_FP32)) == (0, 0, 65)
    o = ovl.sign_opts(degree=64, precision=ovl.PREC_MIXED, head=4, head16=9)
    assert ovl.sign_plan(o)[1:] == (4, 9)
    o = ovl.sign_opts(degree=64, precision=ovl.PREC_MIXED, head=6, head16=9, head48=2)
    assert ovl.sign_plan(o) == (2, 6, 9)
    o.head48 = 20  # a 48-bit start above the FP32 start leaves no 48-bit steps
    assert ovl.sign_plan(o) == (6, 6, 9)
    o = ovl.sign_opts(degree=64, precision=ovl.PREC_MIXED, head=65, head48=1)  # all but k = 0 in 48 bits
    assert ovl.sign_plan(o) == (1, 65, 65)
    with pytest.raises(ovl.OvlError):
        ovl.sign_plan(ovl.sign_opts(cut=1.5))


def test_bench_refuses_more_than_one_gpu():
    # the bench measures one GPU; asked for more it exits with a message before touching a device
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_python_structs_match_the_header(tmp_path):
    # the ctypes mirrors of the parameter, option and result structs have the size and field
    # offsets of include/ovl.h as a C compiler lays them out
    import ctypes as C
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pairs = [("ovl_params", ovl.Params), ("ovl_ensemble_params", ovl.EnsembleParams),
             ("ovl_sign_opts", ovl.SignOpts), ("ovl_step_opts", ovl.StepOpts), ("ovl_tuning", ovl.Tuning),
             ("ovl_observe", ovl.Observe), ("ovl_timing", ovl.Timing)]
    lines, want = [], []
    for cname, py in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(C.sizeof(py))
        for f in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f[0]}));')
            want.append(getattr(py, f[0]).offset)
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "ovl.h"\nint main(void)\n{\n    '
                   + "\n    ".join(lines) + "\n    return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want
