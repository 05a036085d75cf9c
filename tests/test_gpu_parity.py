"""GPU parity: the gfx950 path (through the C ABI) against the CPU oracle and the reference's
golden vectors.  Tolerances: FP64 kernels 1e-12 relative (integer-free floating point, FMA
contraction differs from the oracle), FP32 kernels 2e-5, mixed FP32/FP64 sign 1e-10 against FP64.
Tests that exercise one particular precision ladder (head overrides beyond what tol allows) turn
the mixed sign's accuracy guard off (guard=0), which would otherwise tighten it; the guard has its
own test."""
import os

import numpy as np
import pytest

import orc
import tiers
from overlap_amd import configs as cf
from overlap_amd import ovl

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make(nc, n, nmodes, seed, sigma=0.4, mb=0, r=1.0, m=0.0, M=0.0, nb=0, esig=0.2):
    g = cf.rng(seed)
    links = cf.su2_links(n, sigma, g) if nc == 2 else cf.su3_links(n, sigma, g)
    e = cf.efield(n, nc, esig, g)
    psi = cf.spinors(n, nc, nmodes, g)
    ctx = ovl.Context(nc, n, nmodes, r_wilson=r, m_fermion=m, m_dw=M, mode_block=mb)
    ctx.set_links(links)
    ctx.set_efield(e)
    if nb:
        ctx.set_u1_flux(nb, 1.0)
    ctx.set_modes(ovl.PSI, psi)
    L = orc.lat(nc, n, r, m, M)
    ucc = orc.links_to_ucc(L, links, orc.u1_phases(n, nb, 1.0) if nb else None)
    return ctx, L, ucc, links, e, psi


@pytest.mark.parametrize("nc,nb,mb", [(2, 0, 16), (2, 0, 64), (3, 3, 32), (3, 0, 64), (3, 2, 16)])
def test_wilson_batched_vs_oracle(gpu, nc, nb, mb):
    print(f"[case] wilson nc={nc} nb={nb} mb={mb}", flush=True)
    ctx, L, ucc, *_, psi = make(nc, (4, 3, 5), 40, 100 + nc + mb, mb=mb, m=0.1, M=1.3, nb=nb)
    ctx.apply_wilson(ovl.PSI, ovl.PSIMID)
    out = ctx.get_modes(ovl.PSIMID)
    for k in range(40):
        err = rel(out[k], orc.wilson_evo(L, ucc, psi[k]))
        assert err < 1e-13


def test_wilson_dropin_reference_golden(gpu):
    f = np.load(os.path.join(GOLD, "wilson_su2_4.npz"), allow_pickle=False)
    n = tuple(int(v) for v in f["n"])
    for k, (r, m, M) in enumerate(f["params"]):
        with ovl.Context(2, n, 1, r_wilson=r, m_fermion=m, m_dw=M) as ctx:
            out = ctx.wilson_evo_compute(f["ucc"], f["phi"][k])
            err = rel(out, f["psi"][k])
            assert err < 1e-14


def test_overlap_dropin_reference_golden(gpu):
    # OverlapOperator::Compute: 25-pole Zolotarev + multi-shift CG to ||r||^2 <= 1e-16 ||b||
    f = np.load(os.path.join(GOLD, "overlap_su2_4.npz"), allow_pickle=False)
    n = tuple(int(v) for v in f["n"])
    r, m, M = f["params"]
    lmin, lmax = f["interval"]
    with ovl.Context(2, n, 1, r_wilson=r, m_fermion=m, m_dw=M) as ctx:
        for k in range(2):
            out = ctx.overlap_operator_compute(f["ucc"], f["phi"][k], lmin, lmax)
            err = rel(out, f["psi"][k])
            assert err < 1e-9


def test_ritz_reference_golden(gpu):
    f = np.load(os.path.join(GOLD, "ritz_su2_4.npz"), allow_pickle=False)
    n = tuple(int(v) for v in f["n"])
    r, m, M = f["params"]
    ucc = f["ucc"]
    # master links from u_cc, inverting GetMatrix (SU2/GroupOperations.cpp:213-225); the reference's
    # own inverse GetGaugeLink (:227-239) returns -u2 in component 2, so it is not used here
    V = int(np.prod(n))
    u = ucc.reshape(3, V, 2, 2)
    links = np.empty((V, 3, 4))
    for mu in range(3):
        mm = u[mu]
        links[:, mu, 0] = mm[:, 0, 1].imag
        links[:, mu, 1] = mm[:, 0, 1].real
        links[:, mu, 2] = mm[:, 0, 0].imag
        links[:, mu, 3] = mm[:, 0, 0].real
    with ovl.Context(2, n, 1, r_wilson=r, m_fermion=m, m_dw=M) as ctx:
        ctx.set_links(links.reshape(-1))
        lo = ctx.ritz(f["start"], +1.0)
        hi = ctx.ritz(f["start"], -1.0)
    assert abs(lo - f["eig"][0]) < 1e-9 * abs(f["eig"][0]) + 1e-12
    assert abs(hi - f["eig"][1]) < 1e-9 * abs(f["eig"][1])


@pytest.mark.parametrize("nc,nb,prec,tol", [(2, 0, ovl.PREC_FP64, 1e-12), (3, 2, ovl.PREC_FP64, 1e-12),
                                            (3, 2, ovl.PREC_FP32, 2e-5), (2, 1, ovl.PREC_FP32, 2e-5)])
def test_chebyshev_sign_vs_oracle(gpu, nc, nb, prec, tol):
    print(f"[case] sign nc={nc} nb={nb} prec={prec}", flush=True)
    n = (4, 4, 3)
    ctx, L, ucc, *_, psi = make(nc, n, 6, 7 + nc, M=1.2, nb=nb)
    deg, cut = 96, 0.1
    o = ovl.sign_opts(degree=deg, cut=cut, precision=prec)
    ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
    out = ctx.get_modes(ovl.PSIMID)
    coef = orc.cheb_coeffs(deg, cut)
    lam = orc.norm_bound(L)
    for k in range(6):
        err = rel(out[k], orc.sign_cheb(L, ucc, lam, cut, deg, coef, psi[k]))
        assert err < tol


@pytest.mark.parametrize("nc,nb", [(2, 0), (3, 4)])
def test_overlap_chebyshev_vs_oracle(gpu, nc, nb):
    n = (4, 3, 4)
    ctx, L, ucc, *_, psi = make(nc, n, 5, 17 + nc, m=0.05, M=1.1, nb=nb)
    deg, cut = 80, 0.08
    ctx.apply_overlap(ovl.PSI, ovl.PSIMID, ovl.sign_opts(degree=deg, cut=cut))
    out = ctx.get_modes(ovl.PSIMID)
    coef = orc.cheb_coeffs(deg, cut)
    for k in range(5):
        err = rel(out[k], orc.overlap_evo_cheb(L, ucc, orc.norm_bound(L), cut, deg, coef, psi[k]))
        assert err < 1e-12


@pytest.mark.parametrize("nc,nb", [(2, 0), (3, 8)])
def test_mixed_precision_sign_reaches_1e10(gpu, nc, nb):
    # FP32 tail + FP64 head of the Clenshaw recurrence, head chosen for 1e-10
    n = (6, 6, 6)
    ctx, L, ucc, *_, psi = make(nc, n, 64, 31 + nc, sigma=0.25, M=1.2, nb=nb)
    deg = 384 if nc == 3 else 256
    o64 = ovl.sign_opts(degree=deg, cut=0.05, precision=ovl.PREC_FP64)
    ctx.apply_sign(ovl.PSI, ovl.PSIMID, o64)
    ref = ctx.get_modes(ovl.PSIMID)
    om = ovl.sign_opts(degree=deg, cut=0.05, precision=ovl.PREC_MIXED, tol=1e-10)
    ctx.apply_sign(ovl.PSI, ovl.SCRATCH, om)
    t = ctx.timing()
    mix = ctx.get_modes(ovl.SCRATCH)
    err = max(np.linalg.norm(mix[k] - ref[k]) / np.linalg.norm(psi[k]) for k in range(64))
    print(f"[mixed] nc={nc} steps fp64 {t.fp64_steps} 48-bit {t.fp48_steps} fp32 {t.fp32_steps} "
          f"fp16 {t.fp16_steps} err {err:.3e}", flush=True)
    assert t.fp32_steps > 0 and t.fp16_steps > 0 and t.fp48_steps > 0 and t.fp64_steps < deg + 1
    assert (t.fp64_steps, t.fp48_steps, t.fp32_steps, t.fp16_steps) == tuple(
        np.diff((0,) + ovl.sign_plan(om) + (deg + 1,)))
    assert err <= 1e-10
    # without the FP16 tail: FP64 head + FP32 rest
    om.fp16_tail = 0
    ctx.apply_sign(ovl.PSI, ovl.SCRATCH, om)
    t = ctx.timing()
    mix = ctx.get_modes(ovl.SCRATCH)
    err = max(np.linalg.norm(mix[k] - ref[k]) / np.linalg.norm(psi[k]) for k in range(64))
    assert t.fp16_steps == 0 and t.fp32_steps > 0 and err <= 1e-10


def _agg(a, b):
    return np.sqrt(sum(np.linalg.norm(x - y) ** 2 for x, y in zip(a, b)))


@pytest.mark.parametrize("nc,nb", [(2, 1), (3, 3)])
def test_fp16_tier_vs_oracle(gpu, nc, nb):
    # nearly the whole recurrence in FP16 storage (steps k >= 9 of 48): the result carries the FP16
    # rounding (u16 = 4.9e-4 times the size of the intermediates), the FP64 form does not.  The
    # host model of the same roundings (tiers.sign_tiers) sets the expected size of that error:
    # FP16 rounding under the FP32 arithmetic of the stencil is chaotic (a 3e-7 perturbation of
    # every h_W decorrelates the roundings), so the two agree in size, not pointwise; the size
    # moves by under 8% between perturbed host runs, by 1.4-3.7x with truncation instead of
    # rounding to nearest, by 1.3-2.4x with a scale 2^10 too small
    print(f"[case] fp16 tier nc={nc} nb={nb}", flush=True)
    n = (4, 3, 4)
    ctx, L, ucc, *_, psi = make(nc, n, 20, 60 + nc, M=1.2, nb=nb, mb=16)
    psi[3] *= 20.0  # modes of different norms share the scale of the largest
    ctx.set_modes(ovl.PSI, psi)
    deg, cut = 48, 0.1
    o = ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, head=4, head16=9, fp48=0, guard=0)
    ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
    t = ctx.timing()
    assert (t.fp64_steps, t.fp48_steps, t.fp32_steps, t.fp16_steps) == (4, 0, 5, deg + 1 - 9)
    out = ctx.get_modes(ovl.PSIMID)
    coef = orc.cheb_coeffs(deg, cut)
    lam = orc.norm_bound(L)
    errs = [np.linalg.norm(out[k] - orc.sign_cheb(L, ucc, lam, cut, deg, coef, psi[k])) / np.linalg.norm(psi[k])
            for k in range(20)]
    print(f"  fp16-tier error: max {max(errs):.2e} min {min(errs):.2e}", flush=True)
    assert max(errs) < 2e-2 and np.isfinite(out).all()
    ref = np.array([orc.sign_cheb(L, ucc, lam, cut, deg, coef, x) for x in psi])
    model = tiers.sign_tiers(L, ucc, lam, cut, coef, psi, ovl.sign_plan(o))
    ratio = _agg(out, ref) / _agg(model, ref)
    print(f"  fp16-tier error size GPU / host model: {ratio:.3f}", flush=True)
    assert 0.8 < ratio < 1.25
    # the FP16 variant is timed on its own
    assert t.var_launches[3] == 2 * (deg - 9) and t.var_launches[1] > 0  # b_n = c_n x needs no stencil


@pytest.mark.parametrize("nc,nb,mb", [(2, 2, 16), (3, 5, 64)])
def test_fp48_tier_vs_oracle(gpu, nc, nb, mb):
    # every step but the last (k = 0) in 48-bit storage: the result carries the 48-bit rounding
    # (u = 2^-37 times the size of the intermediates), the FP64 form does not; the two differ
    # but stay far inside 1e-9, and the oracle agrees with the FP64 form to 1e-12
    print(f"[case] 48-bit tier nc={nc} nb={nb} mb={mb}", flush=True)
    n = (4, 4, 3)
    ctx, L, ucc, *_, psi = make(nc, n, 16, 70 + nc, M=1.2, nb=nb, mb=mb)
    psi[5] *= 1e3  # the 48-bit layout keeps the FP64 exponent range: no scales
    ctx.set_modes(ovl.PSI, psi)
    deg, cut = 64, 0.1
    o = ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, head=deg + 1, head48=1, guard=0)
    assert ovl.sign_plan(o) == (1, deg + 1, deg + 1)
    ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
    t = ctx.timing()
    assert (t.fp64_steps, t.fp48_steps, t.fp32_steps, t.fp16_steps) == (1, deg, 0, 0)
    assert t.var_launches[4] == 2 * (deg - 1) and t.var_launches[0] == 3  # b_n = c_n x needs no stencil
    out = ctx.get_modes(ovl.PSIMID)
    ctx.apply_sign(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64))
    ref = ctx.get_modes(ovl.SCRATCH)
    coef = orc.cheb_coeffs(deg, cut)
    lam = orc.norm_bound(L)
    e48 = [np.linalg.norm(out[k] - ref[k]) / np.linalg.norm(psi[k]) for k in range(16)]
    print(f"  48-bit tier vs FP64: max {max(e48):.2e} min {min(e48):.2e}", flush=True)
    assert 0 < max(e48) < 1e-9
    for k in (0, 5, 15):
        assert rel(ref[k], orc.sign_cheb(L, ucc, lam, cut, deg, coef, psi[k])) < 1e-12
        assert np.linalg.norm(out[k] - orc.sign_cheb(L, ucc, lam, cut, deg, coef, psi[k])) / np.linalg.norm(psi[k]) < 1e-9


def test_sign_residual_and_mode_chunks(gpu):
    n = (6, 6, 6)
    ctx, L, ucc, *_, psi = make(3, n, 48, 77, sigma=0.25, M=1.2, nb=8, mb=16)
    o = ovl.sign_opts(degree=384, cut=0.05)
    r = ctx.sign_residual(ovl.PSI, o)
    print(f"[sign residual] {r:.3e}", flush=True)
    assert r < 1e-10
    ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
    a = ctx.get_modes(ovl.PSIMID)
    o.mode_chunk = 1
    ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
    b = ctx.get_modes(ovl.PSIMID)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nc,nb", [(2, 0), (3, 2)])
def test_sign_residual_vs_oracle(gpu, nc, nb):
    # OverlapSignTest (OverlapSignTest.cpp:155-165): max over modes of ||eps(eps x) - x|| / ||x||, on the
    # GPU and from the oracle's sign applied twice to the same inputs.  A low degree makes the
    # residual the polynomial's own error (1e-5 .. 1e-3 here), far above the rounding of either side
    print(f"[case] sign residual nc={nc} nb={nb}", flush=True)
    n = (4, 4, 3)
    ctx, L, ucc, *_, psi = make(nc, n, 8, 61 + nc, M=1.2, nb=nb)
    deg, cut = 40, 0.1
    r = ctx.sign_residual(ovl.PSI, ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64))
    coef, lam = orc.cheb_coeffs(deg, cut), orc.norm_bound(L)
    want = 0.0
    for k in range(8):
        e1 = orc.sign_cheb(L, ucc, lam, cut, deg, coef, psi[k])
        e2 = orc.sign_cheb(L, ucc, lam, cut, deg, coef, e1)
        want = max(want, np.linalg.norm(e2 - psi[k]) / np.linalg.norm(psi[k]))
    print(f"  gpu {r:.6e} oracle {want:.6e}", flush=True)
    assert 1e-7 < want < 1e-1
    assert abs(r - want) <= 1e-9 * want


@pytest.mark.parametrize("nc,n,nm,mb,nb", [(2, (4, 4, 4), 20, 16, 0), (3, (5, 3, 4), 13, 32, 2)])
def test_zolotarev_overlap_batched_vs_oracle(gpu, nc, n, nm, mb, nb):
    # the reference's rational sign (25-pole Zolotarev, multi-shift CG) batched over every mode,
    # including a ragged SU(3) block with flux, against the oracle's restatement mode by mode
    print(f"[case] zolotarev nc={nc} n={n} modes={nm} mb={mb} flux={nb}", flush=True)
    ctx, L, ucc, *_, psi = make(nc, n, nm, 91 + nc, m=0.1, M=1.0, mb=mb, nb=nb)
    o = ovl.sign_opts(method=ovl.SIGN_ZOLOTAREV, lmin=0.2, lmax=30.0)
    ctx.apply_overlap(ovl.PSI, ovl.PSIMID, o)
    out = ctx.get_modes(ovl.PSIMID)
    for k in range(nm):
        err = rel(out[k], orc.overlap_evo_zolo(L, ucc, psi[k], 0.2, 30.0)[0])
        assert err < 1e-9, (k, err)


def test_wilson_leapfrog_config0(gpu):
    # SU(2) 8^3, 16 modes, m = 0.1, r = 1, M = 0, dt = 0.05, links moving with E fixed
    c = cf.CONFIGS["su2_wilson_8"]
    links, e, psi = c.generate()
    steps = 40
    ctx = ovl.Context(2, c.n, c.nmodes, r_wilson=c.r_wilson, m_fermion=c.m_fermion, m_dw=c.m_dw)
    ctx.set_links(links)
    ctx.set_efield(e)
    ctx.set_modes(ovl.PSI, psi)
    ctx.set_modes(ovl.PSIMID, psi)
    so = ovl.StepOpts(kind=ovl.WILSON, dt=c.dt, gauge=ovl.GAUGE_LINKS, sign=ovl.sign_opts(), use_graph=0)
    ctx.fermion_update(so, steps)
    L = orc.lat(2, c.n, c.r_wilson, c.m_fermion, c.m_dw)
    oo = orc.StepOpts(orc.WILSON, c.dt, 1, 0, 1.0, 0, 0.0, 0.0, 0.0, 8)
    lk, p, pm, _ = orc.leapfrog(L, links, e, psi, psi, oo, steps)
    err = rel(ctx.get_links(), lk)
    assert err < 1e-13
    err = rel(ctx.get_modes(ovl.PSI), p)
    assert err < 1e-11
    err = rel(ctx.get_modes(ovl.PSIMID), pm)
    assert err < 1e-11


@pytest.mark.parametrize("prec", [ovl.PREC_FP64, ovl.PREC_MIXED])
def test_overlap_leapfrog_full_gauge_vs_oracle(gpu, prec):
    n = (4, 4, 4)
    nc, nb, deg, cut, dt, steps = 3, 2, 96, 0.1, 0.02, 3
    ctx, L, ucc, links, e, psi = make(nc, n, 8, 123, sigma=0.3, M=1.2, nb=nb)
    ctx.set_modes(ovl.PSIMID, psi)
    so = ovl.StepOpts(kind=ovl.OVERLAP, dt=dt, gauge=ovl.GAUGE_FULL,
                      sign=ovl.sign_opts(degree=deg, cut=cut, precision=prec, tol=1e-11), use_graph=0)
    ctx.fermion_update(so, steps)
    oo = orc.StepOpts(orc.OVERLAP_CHEB, dt, 2, nb, 1.0, deg, cut, 0.0, 0.0, 8)
    lk, p, pm, ee = orc.leapfrog(L, links, e, psi, psi, oo, steps)
    err = rel(ctx.get_links(), lk)
    assert err < 1e-13
    err = rel(ctx.get_efield(), ee)
    assert err < 1e-12
    tol = 1e-12 if prec == ovl.PREC_FP64 else 1e-10
    err = rel(ctx.get_modes(ovl.PSI), p)
    assert err < tol
    err = rel(ctx.get_modes(ovl.PSIMID), pm)
    assert err < tol


def _random_shape(case, seed=2026):
    # (nc, lattice, modes, mode block, flux quanta, field seed) of case number `case`, drawn from one
    # fixed stream: odd and uneven lattice sides, mode counts that leave the last block ragged
    g = np.random.default_rng(seed)
    for _ in range(case + 1):
        nc = int(g.choice([2, 3]))
        n = tuple(int(v) for v in g.integers(3, 8, size=3))
        nm = int(g.integers(1, 71))
        mb = int(g.choice([0, 16, 32, 64]))
        nb = int(g.integers(0, 4))
        fs = int(g.integers(1, 1 << 20))
    return nc, n, nm, mb, nb, fs


@pytest.mark.parametrize("case", range(8))
def test_random_shapes_vs_oracle(gpu, case):
    # the overlap leapfrog (links moving) on drawn shapes: FP64 against the oracle for the first and
    # last mode; the mixed ladder (every tier at degree 256, the guard allowed to loosen) against
    # FP64 for every mode; hipGraph replay bit-identical to stream launches
    nc, n, nm, mb, nb, fs = _random_shape(case)
    print(f"[case] random shape {case}: nc={nc} n={n} modes={nm} mb={mb} flux={nb}", flush=True)
    deg, cut, dt, steps = 256, 0.05, 0.02, 2
    ctx, L, ucc, links, e, psi = make(nc, n, nm, fs, sigma=0.3, mb=mb, M=1.2, nb=nb)
    with ctx:
        ctx.set_modes(ovl.PSIMID, psi)
        so = ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64)
        ctx.fermion_update(ovl.StepOpts(ovl.OVERLAP, dt, ovl.GAUGE_LINKS, so, 0), steps)
        p64, m64 = ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID)
        sample = sorted({0, nm - 1})
        oo = orc.StepOpts(orc.OVERLAP_CHEB, dt, 1, nb, 1.0, deg, cut, 0.0, 0.0, 8)
        lk, p, pm, _ = orc.leapfrog(L, links, e, psi[sample], psi[sample], oo, steps)
        assert rel(ctx.get_links(), lk) < 1e-13
        for i, s in enumerate(sample):
            assert rel(p64[s], p[i]) < 1e-12 and rel(m64[s], pm[i]) < 1e-12, (s, rel(p64[s], p[i]))
        runs = []
        for graph in (0, 1):
            ctx.set_links(links)
            ctx.set_modes(ovl.PSI, psi)
            ctx.set_modes(ovl.PSIMID, psi)
            so = ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, tol=1e-10, relax=2)
            ctx.fermion_update(ovl.StepOpts(ovl.OVERLAP, dt, ovl.GAUGE_LINKS, so, graph), steps)
            t = ctx.timing()
            runs.append((ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID), t.guard_level))
        assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
        err = max(max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(x, y))
                  for x, y in ((runs[0][0], p64), (runs[0][1], m64)))
        print(f"  guard level {runs[0][2]}, mixed vs fp64 {err:.2e}", flush=True)
        assert err <= 1e-10


@pytest.mark.parametrize("case", range(4))
def test_random_shapes_wilson_and_observables(gpu, case):
    # drawn shapes (another fixed stream) through the Wilson leapfrog with the full E update in FP64
    # and, links moving only, with FP32 spinor arithmetic, against the oracle for the first and last
    # mode; the FP64 observables of the evolved fields against the oracle's
    nc, n, nm, mb, nb, fs = _random_shape(case, seed=77)
    print(f"[case] random wilson shape {case}: nc={nc} n={n} modes={nm} mb={mb} flux={nb}", flush=True)
    dt, steps = 0.05, 4
    ctx, L, ucc, links, e, psi = make(nc, n, nm, fs, sigma=0.4, mb=mb, m=0.1, nb=nb, esig=0.3)
    sample = sorted({0, nm - 1})
    with ctx:
        ctx.set_modes(ovl.PSIMID, psi)
        so = ovl.StepOpts(ovl.WILSON, dt, ovl.GAUGE_FULL, ovl.sign_opts(precision=ovl.PREC_FP64), 1)
        ctx.fermion_update(so, steps)
        oo = orc.StepOpts(orc.WILSON, dt, 2, nb, 1.0, 0, 0.0, 0.0, 0.0, 8)
        lk, p, pm, ee = orc.leapfrog(L, links, e, psi[sample], psi[sample], oo, steps)
        assert rel(ctx.get_links(), lk) < 1e-13 and rel(ctx.get_efield(), ee) < 1e-12
        for i, s in enumerate(sample):
            assert rel(ctx.get_modes(ovl.PSI, s, 1)[0], p[i]) < 1e-12
            assert rel(ctx.get_modes(ovl.PSIMID, s, 1)[0], pm[i]) < 1e-12
        L2 = orc.lat(nc, n, 1.0, 0.1, 0.0)
        assert np.allclose(ctx.energy(), orc.plaquette_energy(L2, lk, ee), rtol=1e-12, atol=0)
        assert np.allclose(ctx.gauss_violation(), orc.gauss_violation(L2, lk, ee), rtol=1e-11, atol=1e-15)
        ref = orc.edotb(L2, lk, ee) / (8 * np.pi ** 2)
        assert abs(ctx.ncs_dot() - ref) <= 1e-12 * max(abs(ref), 1e-12)
        # FP32 spinor arithmetic (links moving, E fixed) from the same start
        ctx.set_links(links)
        ctx.set_efield(e)
        ctx.set_modes(ovl.PSI, psi)
        ctx.set_modes(ovl.PSIMID, psi)
        so = ovl.StepOpts(ovl.WILSON, dt, ovl.GAUGE_LINKS, ovl.sign_opts(precision=ovl.PREC_FP32), 1)
        ctx.fermion_update(so, steps)
        oo = orc.StepOpts(orc.WILSON, dt, 1, nb, 1.0, 0, 0.0, 0.0, 0.0, 8)
        _, p, pm, _ = orc.leapfrog(L, links, e, psi[sample], psi[sample], oo, steps)
        for i, s in enumerate(sample):
            assert rel(ctx.get_modes(ovl.PSI, s, 1)[0], p[i]) < 2e-5
            assert rel(ctx.get_modes(ovl.PSIMID, s, 1)[0], pm[i]) < 2e-5


def test_graph_replay_matches_stream(gpu):
    n = (6, 4, 4)
    outs = []
    for graph in (0, 1, 1):
        ctx, L, ucc, links, e, psi = make(3, n, 32, 55, sigma=0.3, M=1.2, nb=3)
        ctx.set_modes(ovl.PSIMID, psi)
        so = ovl.StepOpts(kind=ovl.OVERLAP, dt=0.02, gauge=ovl.GAUGE_LINKS,
                          sign=ovl.sign_opts(degree=64, cut=0.1, precision=ovl.PREC_MIXED), use_graph=graph)
        ctx.fermion_update(so, 3)
        outs.append((ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID), ctx.get_links()))
        ctx.close()
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


def test_kept_graph_follows_the_context(gpu):
    # a graph-mode update keeps its captured step: a later call with nothing changed replays it
    # without capturing, a change of options, of tuning or of the FP16 scales (larger modes)
    # captures anew, and every sequence gives the bits of the same calls on the stream
    print("[case] kept step graph", flush=True)
    n = (6, 4, 4)

    def opts(dt, graph):
        return ovl.StepOpts(kind=ovl.OVERLAP, dt=dt, gauge=ovl.GAUGE_LINKS, use_graph=graph,
                            sign=ovl.sign_opts(degree=64, cut=0.1, precision=ovl.PREC_MIXED, head48=4, head=12,
                                               head16=30, guard=0))
    runs = []
    for graph in (0, 1):
        ctx, L, ucc, links, e, psi = make(3, n, 32, 57, sigma=0.3, M=1.2, nb=3)
        ctx.set_modes(ovl.PSIMID, psi)
        caps = []

        def step(dt, k):
            ctx.fermion_update(opts(dt, graph), k)
            caps.append(ctx.timing().graph_captures)
        step(0.02, 2)  # first call: capture
        step(0.02, 3)  # unchanged: replay
        step(0.01, 1)  # other options
        t = ctx.get_tuning(3)
        ctx.set_tuning(3, t.traversal, t.slabs, t.nt_store, 3 if t.waves_per_eu != 3 else 2)
        step(0.01, 1)  # other tuning of the 48-bit stencil
        big = ctx.get_modes(ovl.PSI) * 1024.0
        ctx.set_modes(ovl.PSI, big)
        ctx.set_modes(ovl.PSIMID, big)
        step(0.01, 1)  # FP16 scales 2^-10 smaller
        step(0.01, 2)  # unchanged: replay
        runs.append((caps, ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID), ctx.get_links()))
        ctx.close()
    assert runs[0][0] == [0] * 6
    assert runs[1][0] == [1, 0, 1, 1, 1, 0]
    for a, b in zip(runs[1][1:], runs[0][1:]):
        assert np.isfinite(b).all() and np.array_equal(a, b)


@pytest.mark.parametrize("nc", [2, 3])
def test_observables_vs_oracle(gpu, nc):
    n = (5, 4, 4)
    ctx, L, ucc, links, e, psi = make(nc, n, 16, 200 + nc, sigma=0.6, M=1.0, esig=0.5)
    assert np.allclose(ctx.energy(), orc.plaquette_energy(L, links, e), rtol=1e-13, atol=0)
    g_gpu, g_orc = ctx.gauss_violation(), orc.gauss_violation(L, links, e)
    assert np.allclose(g_gpu, g_orc, rtol=1e-12, atol=0)
    # unitarity on links pushed off the group (each scaled by its own 1 + O(1e-2)): the oracle's
    # value is the closed form |lam^2 - 1| (times sqrt 3 for SU(3)), tests/test_oracle.py
    g = cf.rng(900 + nc)
    per = 4 if nc == 2 else 18
    lam = 1.0 + g.normal(0, 0.01, size=3 * int(np.prod(n)))
    off = (links.reshape(-1, per) * lam[:, None]).reshape(-1)
    ctx.set_links(off)
    u_gpu, u_orc = ctx.unitarity_violation(), orc.unitarity_violation(L, off)
    assert u_orc[0] > 1e-3
    assert np.allclose(u_gpu, u_orc, rtol=1e-12, atol=0)
    ctx.set_links(links)
    u_gpu, u_orc = ctx.unitarity_violation(), orc.unitarity_violation(L, links)
    assert abs(u_gpu[0] - u_orc[0]) < 1e-15 and abs(u_gpu[1] - u_orc[1]) < 1e-15
    assert abs(ctx.ncs_dot() - orc.edotb(L, links, e) / (8 * np.pi ** 2)) < 1e-12 * abs(orc.edotb(L, links, e))
    # one E update
    ctx.update_efield(0.03)
    err = rel(ctx.get_efield(), orc.efield_update(L, links, e, 0.03))
    assert err < 1e-13


@pytest.mark.parametrize("nc", [2, 3])
def test_link_update_reunitarizes(gpu, nc):
    # reunitarize = 1: every link update ends with the projection onto SU(Nc) (SU(2) V/|V|, SU(3)
    # polar factor with det 1); from links pushed off the group by 1e-3 the update must equal the
    # oracle's exponential update followed by its projection (pinned against numpy's SVD polar
    # factor) and land on the group to rounding
    print(f"[case] reunitarize nc={nc}", flush=True)
    n = (4, 3, 5)
    V = int(np.prod(n))
    g = cf.rng(910 + nc)
    links = cf.su2_links(n, 0.5, g) if nc == 2 else cf.su3_links(n, 0.5, g)
    e = cf.efield(n, nc, 0.4, g)
    if nc == 2:
        off = (links.reshape(-1, 4) * (1 + g.normal(0, 1e-3, size=(3 * V, 1)))).reshape(-1)
    else:
        U = cf.links_to_matrices(3, links)
        X = g.normal(size=U.shape) + 1j * g.normal(size=U.shape)
        cm = np.swapaxes(U @ (np.eye(3) + 1e-3 * X), -1, -2).reshape(V, 3, 9)
        o = np.empty((V, 3, 18))
        o[..., 0::2], o[..., 1::2] = cm.real, cm.imag
        off = o.reshape(-1)
    L = orc.lat(nc, n)
    for reunit in (0, 1):
        ctx = ovl.Context(nc, n, 16, reunitarize=reunit)
        ctx.set_links(off)
        ctx.set_efield(e)
        ctx.update_links(0.05)
        got = ctx.get_links()
        want = orc.update_links(L, off, e, 0.05)
        if reunit:
            want = orc.reunitarize(L, want)
            assert ctx.unitarity_violation()[0] < 1e-14
        else:
            assert ctx.unitarity_violation()[0] > 1e-4
        assert rel(got, want) < 1e-13
        ctx.close()


def test_currents_vs_oracle(gpu):
    n = (4, 4, 4)
    ctx, L, ucc, links, e, psi = make(2, n, 24, 300, sigma=0.3, M=1.0, mb=16)
    deg, cut = 64, 0.1
    o = ovl.sign_opts(degree=deg, cut=cut)
    jv, ja, q = ctx.currents(ovl.PSI, o, first_mode=3)
    coef = orc.cheb_coeffs(deg, cut)
    h = np.stack([orc.overlap_evo_cheb(L, ucc, orc.norm_bound(L), cut, deg, coef, p) for p in psi])
    jv0, ja0 = orc.currents(L, psi, h, first_mode=3)
    ev, ea = np.abs(jv - jv0).max(), np.abs(ja - ja0).max()
    assert ev < 1e-14 and ea < 1e-14
    assert abs(q - ja0[0::4].sum()) < 1e-13


def test_error_paths(gpu):
    ctx = ovl.Context(2, (4, 4, 4), 4)
    with pytest.raises(ovl.OvlError) as ex:
        ctx.apply_wilson(ovl.PSI, ovl.PSIMID)
    assert ex.value.code == ovl.E_STATE
    ctx.set_links(cf.su2_links((4, 4, 4), 0.1, cf.rng(1)))
    with pytest.raises(ovl.OvlError) as ex:
        ctx.apply_wilson(ovl.PSI, 7)
    assert ex.value.code == ovl.E_ARG
    with pytest.raises(ovl.OvlError) as ex:
        ctx.apply_sign(ovl.PSI, ovl.PSIMID, ovl.sign_opts(cut=1.5))
    assert ex.value.code == ovl.E_SPECTRUM
    bad = cf.spinors((4, 4, 4), 2, 1, cf.rng(2))
    bad[0, 5] = np.nan
    with pytest.raises(ovl.OvlError) as ex:
        ctx.set_modes(ovl.PSI, bad)
    assert ex.value.code == ovl.E_ARG
    with pytest.raises(ovl.OvlError) as ex:
        ctx.apply_overlap(ovl.PSI, ovl.PSIMID, ovl.sign_opts(method=ovl.SIGN_ZOLOTAREV, lmin=2.0, lmax=1.0))
    assert ex.value.code == ovl.E_SPECTRUM
    # the eigenbasis has D = 4 Nc V = 512 modes here: modes 510..513 run past it
    with pytest.raises(ovl.OvlError) as ex:
        ctx.qed_eigenbasis(ovl.sign_opts(degree=16, cut=0.1), 1, first_mode=510)
    assert ex.value.code == ovl.E_ARG
    # observed updates recover -i h_ov Psi from the dt-scaled step term: dt = 0 is refused up front
    so = ovl.StepOpts(ovl.OVERLAP, 0.0, ovl.GAUGE_FROZEN, ovl.sign_opts(degree=16, cut=0.1), 0)
    with pytest.raises(ovl.OvlError) as ex:
        ctx.fermion_update_observed(so, 2, 1)
    assert ex.value.code == ovl.E_ARG
    # the single-mode drop-ins check their host inputs for Inf/NaN (on the device, after the
    # upload): every such input is refused, and the next call is unaffected
    g = cf.rng(3)
    V = 64
    ucc = (g.standard_normal((3 * V, 2, 2)) + 1j * g.standard_normal((3 * V, 2, 2))) / 2
    phi = g.standard_normal(V * 8) + 1j * g.standard_normal(V * 8)
    good = ctx.wilson_evo_compute(ucc, phi)
    for which, idx in (("ucc", (3 * V - 1, 1, 0)), ("ucc", (0, 0, 1)), ("phi", (17,)), ("phi", (V * 8 - 1,))):
        for v in (np.nan, np.inf, complex(0.0, -np.inf), complex(0.0, np.nan)):
            u, p = ucc.copy(), phi.copy()
            (u if which == "ucc" else p)[idx] = v
            with pytest.raises(ovl.OvlError, match=f"non-finite input \\({'u_cc' if which == 'ucc' else 'phi'}\\)") as ex:
                ctx.wilson_evo_compute(u, p)
            assert ex.value.code == ovl.E_ARG
    with pytest.raises(ovl.OvlError) as ex:
        ctx.overlap_operator_compute(ucc, np.where(np.arange(V * 8) == 3, np.nan, phi), 0.01, 4.0)
    assert ex.value.code == ovl.E_ARG
    assert np.array_equal(ctx.wilson_evo_compute(ucc, phi), good)


def test_nonfinite_check_finds_every_position(gpu):
    # the device check streams 512 x 256 lanes over 16-byte units, four units per lane in flight, then
    # a remainder loop: on 28^3 SU(3) the u_cc input of the drop-in (592 704 units) runs both loops, so
    # one bad double is planted at the ends of each (and in both halves of a unit)
    print("[case] non-finite check positions", flush=True)
    n = (28, 28, 28)
    V = n[0] * n[1] * n[2]
    ctx = ovl.Context(3, n, 1)
    g = cf.rng(4)
    ucc = (g.standard_normal((3 * V, 3, 3)) + 1j * g.standard_normal((3 * V, 3, 3))) / 3
    phi = g.standard_normal(V * 12) + 1j * g.standard_normal(V * 12)
    good = ctx.wilson_evo_compute(ucc, phi)
    flat = ucc.reshape(-1)
    lanes = 512 * 256
    units = flat.size  # one complex = one 16-byte unit
    assert units > 4 * lanes
    for k in (0, 1, lanes - 1, lanes, 4 * lanes - 1, 4 * lanes, units - lanes, units - 1):
        for v in (complex(np.nan, 0.0), complex(0.0, np.inf)):
            u = flat.copy()
            u[k] = v
            with pytest.raises(ovl.OvlError, match="non-finite input") as ex:
                ctx.wilson_evo_compute(u.reshape(ucc.shape), phi)
            assert ex.value.code == ovl.E_ARG
    for k in (0, phi.size // 2, phi.size - 1):
        p = phi.copy()
        p[k] = complex(-np.inf, 1.0)
        with pytest.raises(ovl.OvlError, match="non-finite input") as ex:
            ctx.wilson_evo_compute(ucc, p)
        assert ex.value.code == ovl.E_ARG
    # the largest finite doubles are finite
    big = ucc.copy()
    big.reshape(-1)[7] = complex(np.finfo(float).max, -np.finfo(float).max)
    ctx.wilson_evo_compute(big, phi)
    assert np.array_equal(ctx.wilson_evo_compute(ucc, phi), good)
    ctx.close()


def test_leapfrog_beyond_its_stability_bound(gpu):
    # dt |omega| >> 2: the staggered leapfrog amplifies a mode by about dt |omega| per half step, so
    # after 40 steps at dt = 1e4 (|omega| up to 6) the norms pass the double range; the call reports
    # it instead of handing Inf back (every kernel still runs to its end: the host chose the inputs)
    print("[case] leapfrog beyond the stability bound", flush=True)
    ctx, L, ucc, links, e, psi = make(2, (4, 4, 4), 16, 808, M=0.0)
    with ctx:
        so = ovl.StepOpts(ovl.WILSON, 1e4, ovl.GAUGE_FROZEN, ovl.sign_opts(precision=ovl.PREC_FP64), 0)
        with pytest.raises(ovl.OvlError) as ex:
            ctx.fermion_update(so, 40)
        assert ex.value.code == ovl.E_NONFINITE
        ctx.set_modes(ovl.PSI, psi)
        ctx.set_modes(ovl.PSIMID, psi)
        so.dt = 0.05
        ctx.fermion_update(so, 2)  # a stable step afterwards works as before
    # the mixed overlap ladder stores its tail in FP16 scaled for norms within twice their start:
    # beyond dt ||h_ov|| = 2 (||h_ov|| <= 2 rho at m = 0) it stops at FP32
    ctx, *_ = make(3, (4, 4, 4), 16, 809, M=1.2, nb=1)
    with ctx:
        o = ovl.sign_opts(degree=64, cut=0.1, precision=ovl.PREC_MIXED, tol=1e-8)
        ctx.fermion_update(ovl.StepOpts(ovl.OVERLAP, 0.02, ovl.GAUGE_FROZEN, o, 0), 1)
        assert ctx.timing().fp16_steps > 0
        ctx.fermion_update(ovl.StepOpts(ovl.OVERLAP, 1.0, ovl.GAUGE_FROZEN, o, 0), 1)
        t = ctx.timing()
        assert t.fp16_steps == 0 and t.fp32_steps > 0


@pytest.mark.parametrize("nc,n,nm,mb", [(3, (4, 4, 4), 64, 64), (2, (8, 4, 6), 32, 16), (3, (6, 4, 4), 64, 32),
                                        (3, (6, 4, 4), 128, 16), (2, (8, 4, 6), 128, 16), (3, (4, 4, 8), 512, 64)])
def test_traversals_are_permutations(gpu, nc, n, nm, mb):
    # every lane traversal (lexicographic, XCD boxes, XCD mode split with y slabs, and whatever the
    # autotuner picks) only reorders lanes: results equal bit for bit
    print(f"[case] traversals nc={nc} n={n} modes={nm} mb={mb}", flush=True)
    outs = []
    g = cf.rng(404)
    links = cf.su2_links(n, 0.4, g) if nc == 2 else cf.su3_links(n, 0.4, g)
    psi = cf.spinors(n, nc, nm, g)
    for trav in (1, 2, 3, 0):
        with ovl.Context(nc, n, nm, m_fermion=0.1, m_dw=1.2, mode_block=mb, traversal=trav) as ctx:
            ctx.set_links(links)
            ctx.set_modes(ovl.PSI, psi)
            if trav == 0:
                print("  autotune:", ctx.autotune(), flush=True)
            ctx.apply_sign(ovl.PSI, ovl.PSIMID,
                           ovl.sign_opts(degree=32, cut=0.1, precision=ovl.PREC_MIXED, head=8, head16=20, guard=0))
            outs.append(ctx.get_modes(ovl.PSIMID))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    L = orc.lat(nc, n, 1.0, 0.1, 1.2)
    ucc = orc.links_to_ucc(L, links)
    coef = orc.cheb_coeffs(32, 0.1)
    err = rel(outs[0][nm - 1], orc.sign_cheb(L, ucc, orc.norm_bound(L), 0.1, 32, coef, psi[nm - 1]))
    assert err < 1e-5


def test_qed_eigenbasis_zolotarev_vs_dense_oracle(gpu):
    # the reference's own sign (25-pole Zolotarev, multi-shift CG) on a fixed interval of h_W^2:
    # the block spectrum against numpy's eigh of the oracle's dense matrix of the same operator
    n, nc, nb, M, dt, lmin, lmax = (3, 3, 4), 2, 1, 1.0, 0.02, 0.2, 30.0
    print(f"[case] qed eigenbasis zolotarev n={n}", flush=True)
    V = int(np.prod(n))
    D = 4 * nc * V
    ctx = ovl.Context(nc, n, 16, m_dw=M, mode_block=16)
    ev, herm = ctx.qed_eigenbasis(ovl.sign_opts(method=ovl.SIGN_ZOLOTAREV, lmin=lmin, lmax=lmax), nb, 1.0, dt)
    L = orc.lat(nc, n, 1.0, 0.0, M)
    ph = orc.u1_phases(n, nb, 1.0)
    ucc = np.zeros((3, V, nc, nc), complex)
    ucc[:, :, 0, 0] = np.asarray(ph).reshape(3, V)
    H = np.empty((D, D), complex)
    for k in range(D):
        e = np.zeros(D, complex)
        e[k] = 1.0
        H[:, k] = 1j * orc.overlap_evo_zolo(L, ucc.reshape(-1), e, lmin, lmax)[0]
    Hh = 0.5 * (H + H.conj().T)
    w = np.linalg.eigvalsh(Hh)
    assert herm < 1e-8
    assert np.abs(ev - w[::-1]).max() < 1e-8


def test_qed_eigenbasis_beyond_dense_size(gpu):
    # SU(3) 12 x 10 x 12: D = 4 Nc V = 17280, more than a dense 16384 problem; the block solve gives
    # the whole spectrum, the returned modes are orthogonal eigenvectors of the operator the oracle
    # applies (U(1) phase on colour 0, no other hopping)
    n, nc, nb, deg, cut, M, dt = (12, 10, 12), 3, 3, 48, 0.1, 1.2, 0.02
    print(f"[case] qed eigenbasis n={n} nc={nc}", flush=True)
    V = int(np.prod(n))
    D = 4 * nc * V
    nm = 8
    ctx = ovl.Context(nc, n, nm, m_dw=M, mode_block=16)
    ev, herm = ctx.qed_eigenbasis(ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64), nb, 1.0, dt,
                                  first_mode=5)
    assert ev.shape == (D,) and np.all(np.diff(ev) <= 0) and herm < 1e-12 and np.isfinite(ev).all()
    L = orc.lat(nc, n, 1.0, 0.0, M)
    ph = orc.u1_phases(n, nb, 1.0)
    ucc = np.zeros((3, V, nc, nc), complex)
    ucc[:, :, 0, 0] = np.asarray(ph).reshape(3, V)
    coef = orc.cheb_coeffs(deg, cut)
    lam = orc.norm_bound(L)
    psi = ctx.get_modes(ovl.PSI)
    gram = psi.conj() @ psi.T
    assert np.abs(gram - V * np.eye(nm)).max() < 1e-9 * V
    for s in range(nm):
        g = 5 + s
        particle, level = g % 2, g // 2
        b = D // 2 - 1 - level if particle == 0 else D // 2 + level
        hv = 1j * orc.overlap_evo_cheb(L, ucc.reshape(-1), lam, cut, deg, coef, psi[s])
        assert np.linalg.norm(hv - ev[b] * psi[s]) < 1e-9 * np.sqrt(V)


def test_sign_polynomial_follows_degree_and_cut(gpu):
    # the context keeps the last sign polynomial (coefficients and Clenshaw bounds): a change of
    # degree or cut between applies on one context must give the bits of a fresh context
    print("[case] sign polynomial per degree and cut", flush=True)
    n = (4, 4, 6)
    opts = [ovl.sign_opts(degree=32, cut=0.1, precision=ovl.PREC_MIXED, tol=1e-10),
            ovl.sign_opts(degree=48, cut=0.1, precision=ovl.PREC_MIXED, tol=1e-10),
            ovl.sign_opts(degree=48, cut=0.05, precision=ovl.PREC_MIXED, tol=1e-10),
            ovl.sign_opts(degree=48, cut=0.05, precision=ovl.PREC_FP64)]
    ctx, *_ = make(3, n, 16, 611, M=1.2, nb=2, mb=16)
    seq = []
    for o in opts + opts[:1]:
        ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
        seq.append(ctx.get_modes(ovl.PSIMID))
    for o, got in zip(opts, seq):
        fresh, *_ = make(3, n, 16, 611, M=1.2, nb=2, mb=16)
        fresh.apply_sign(ovl.PSI, ovl.PSIMID, o)
        assert np.array_equal(fresh.get_modes(ovl.PSIMID), got)
    assert np.array_equal(seq[-1], seq[0])
    assert rel(seq[1], seq[2]) > 1e-6  # the cut changes the polynomial


@pytest.mark.parametrize("nc,mb", [(3, 64), (2, 16)])
def test_occupancy_variants_are_bit_identical(gpu, nc, mb):
    # occupancy variants 1-5 (5: one workgroup per CU through an LDS reservation) run the same
    # per-lane code: every tier of the recurrence gives the same bits under each
    print(f"[case] occupancy variants nc={nc} mb={mb}", flush=True)
    n = (4, 4, 8)
    ctx, L, ucc, *_, psi = make(nc, n, 64, 505 + nc, M=1.2, nb=2, mb=mb)
    o = ovl.sign_opts(degree=32, cut=0.1, precision=ovl.PREC_MIXED, head48=2, head=8, head16=20, guard=0)
    assert ovl.sign_plan(o) == (2, 8, 20)
    outs = []
    for v in (1, 2, 3, 4, 5):
        for prec in range(4):
            t = ctx.get_tuning(prec)
            ctx.set_tuning(prec, t.traversal, t.slabs, t.nt_store, v)
            assert ctx.get_tuning(prec).waves_per_eu == v
        ctx.apply_sign(ovl.PSI, ovl.PSIMID, o)
        outs.append(ctx.get_modes(ovl.PSIMID))
    for x in outs[1:]:
        assert np.array_equal(x, outs[0])
    with pytest.raises(ovl.OvlError):
        ctx.set_tuning(0, 1, 1, 1, 6)


@pytest.mark.parametrize("nc", [2, 3])
def test_gauss_restore_vs_oracle(gpu, nc):
    print(f"[case] gauss restore nc={nc}", flush=True)
    n = (5, 4, 3)
    ctx, L, ucc, links, e, psi = make(nc, n, 4, 450 + nc, sigma=0.5, esig=0.3)
    it, vmax = ctx.gauss_restore(tol=0.0, max_iter=7)
    e_ref, it_ref, vmax_ref = orc.gauss_restore(L, links, e, 0.0, 7)
    assert it == it_ref == 7
    err = rel(ctx.get_efield(), e_ref)
    assert err < 1e-12
    assert abs(vmax - vmax_ref) < 1e-12 * max(vmax_ref, 1e-3)
    # to convergence; links untouched
    it, vmax = ctx.gauss_restore(tol=1e-11, max_iter=5000)
    assert vmax <= 1e-11 and it < 5000
    assert np.array_equal(ctx.get_links(), links)
    assert ctx.gauss_violation()[0] == vmax


def test_observed_update_currents(gpu):
    # currents every 2 steps from the leapfrog's own overlap apply: the run is bit-identical to a
    # plain one, and the currents are those of -i h_ov[U(t + dt/2)] Psi(t + dt) (the operator the
    # reference's ComputeCurrents applies after Update)
    print("[case] observed update", flush=True)
    n, nc, nb, deg, cut, dt = (4, 4, 4), 3, 2, 64, 0.1, 0.02
    so = ovl.StepOpts(kind=ovl.OVERLAP, dt=dt, gauge=ovl.GAUGE_LINKS,
                      sign=ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64), use_graph=1)
    runs = []
    for observed in (False, True):
        ctx, L, ucc, links, e, psi = make(nc, n, 16, 700, sigma=0.3, M=1.2, nb=nb, mb=16)
        ctx.set_modes(ovl.PSIMID, psi)
        if observed:
            q5, ncs, jv, ja = ctx.fermion_update_observed(so, 5, every=2, first_mode=3)
        else:
            ctx.fermion_update(so, 5)
        runs.append((ctx, ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID), ctx.get_links()))
    for a, b in zip(runs[0][1:], runs[1][1:]):
        assert np.array_equal(a, b)
    assert len(q5) == 2 and len(ncs) == 2
    # the last observation came after step 4: undo step 5 on the CPU for the state of step 4
    ctx4, L, ucc, links, e, psi = make(nc, n, 16, 700, sigma=0.3, M=1.2, nb=nb, mb=16)
    ctx4.set_modes(ovl.PSIMID, psi)
    ctx4.fermion_update(so, 4)
    psi4, links4 = ctx4.get_modes(ovl.PSI), ctx4.get_links()
    half = orc.update_links(L, links4, e, -0.5 * dt)  # U(t + dt/2) = exp(+i dt/2 E) U(t + dt)
    uh = orc.links_to_ucc(L, half, orc.u1_phases(n, nb, 1.0))
    coef = orc.cheb_coeffs(deg, cut)
    h = np.stack([orc.overlap_evo_cheb(L, uh, orc.norm_bound(L), cut, deg, coef, p) for p in psi4])
    jv0, ja0 = orc.currents(L, psi4, h, first_mode=3)
    assert abs(q5[-1] - ja0[0::4].sum()) < 1e-12
    assert abs(ncs[-1] - ctx4.ncs_dot()) <= 1e-15 * max(1.0, abs(ncs[-1]))
    # jv/ja returned by the call are those of the last observation (after step 4)
    assert np.abs(jv - jv0).max() < 1e-13 and np.abs(ja - ja0).max() < 1e-13


@pytest.mark.parametrize("nc,n,nb,nm,mb,first", [(2, (4, 4, 3), 2, 16, 16, 2), (3, (3, 3, 4), 1, 16, 16, 2),
                                                 (2, (4, 4, 3), 2, 48, 16, 2), (3, (3, 3, 4), 1, 64, 32, 2),
                                                 (3, (3, 3, 4), 1, 16, 16, 352)])
def test_qed_eigenbasis_vs_dense_oracle(gpu, nc, n, nb, nm, mb, first):
    # the overlap Hamiltonian of the U(1) background alone (colour 0 carries the phase), assembled
    # on the CPU from the oracle's overlap apply and diagonalised by numpy, against the GPU's.  The
    # cases cover one and several mode blocks per batch of unit sources (the z-momentum transform
    # is one batched product over the blocks), several batches, and the colour-1 spin sources in
    # a batch of their own, after the last colour-0 sources in the same batch, and in a later block.
    # Modes from 352 on (levels 176-183 of the 432 levels) lie in the +-2.4 eigenspaces, where the
    # colours without hopping have their eigenvectors: the single-site spin modes.
    print(f"[case] qed eigenbasis nc={nc} n={n} nb={nb} modes={nm} mb={mb} first={first}", flush=True)
    V = int(np.prod(n))
    D = 4 * nc * V
    deg, cut, M, dt = 64, 0.1, 1.2, 0.02
    ctx = ovl.Context(nc, n, nm, m_dw=M, mode_block=mb)
    ev, herm = ctx.qed_eigenbasis(ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64), nb, 1.0, dt,
                                  first_mode=first)
    L = orc.lat(nc, n, 1.0, 0.0, M)
    ph = orc.u1_phases(n, nb, 1.0)  # complex [mu][site]
    ucc = np.zeros((3, V, nc, nc), complex)
    ucc[:, :, 0, 0] = np.asarray(ph).reshape(3, V)
    coef = orc.cheb_coeffs(deg, cut)
    lam = orc.norm_bound(L)
    H = np.empty((D, D), complex)
    for k in range(D):
        e = np.zeros(D, complex)
        e[k] = 1.0
        H[:, k] = 1j * orc.overlap_evo_cheb(L, ucc.reshape(-1), lam, cut, deg, coef, e)
    assert np.abs(H - H.conj().T).max() < 1e-12 and herm < 1e-12
    w, vecs = np.linalg.eigh(H)
    assert np.abs(ev - w[::-1]).max() < 1e-11
    # modes: eigenvectors of H with norm^2 V; PsiMid = Psi exp(-i omega dt / 2)
    psi, mid = ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID)
    on_colour0 = (np.arange(D) // 4) % nc == 0  # AoS entry (site, colour, spin)
    single_site = 0
    for s in range(nm):
        g = first + s
        particle, level = g % 2, g // 2
        b = D // 2 - 1 - level if particle == 0 else D // 2 + level
        omega = ev[b]
        v = psi[s]
        assert abs(np.vdot(v, v).real - V) < 1e-9 * V
        assert np.linalg.norm(H @ v - omega * v) < 1e-9 * np.sqrt(V)
        assert np.allclose(mid[s], v * np.exp(-0.5j * omega * dt), rtol=0, atol=1e-14)
        if not np.any(v[on_colour0]):
            assert np.count_nonzero(v) <= 4  # one site, one colour >= 1
            single_site += 1
    assert single_site == (nm if first == 352 else 0)
    # the SU(Nc) stencil is back afterwards
    ctx2, L2, ucc2, links, e2, psi2 = make(nc, n, nm, 5, M=M, mb=16)
    ctx2.qed_eigenbasis(ovl.sign_opts(degree=16, cut=0.2, precision=ovl.PREC_FP64), nb, 1.0, dt)
    ctx2.set_modes(ovl.PSI, psi2)
    ctx2.apply_wilson(ovl.PSI, ovl.PSIMID)
    assert rel(ctx2.get_modes(ovl.PSIMID)[0], orc.wilson_evo(L2, ucc2, psi2[0])) < 1e-13


def test_guard_tightens_an_aggressive_ladder(gpu):
    # a mixed ladder forced far too low (FP32 from step 6, FP16 from step 12 of degree 240): with the
    # guard off it misses 1e-10 by orders of magnitude; with it on, the first check of the sampled
    # mode block fails, the context drops the overrides (ladder from tol), the recheck passes, and the
    # apply and a leapfrog run end within 1e-10 of FP64 and of the oracle.  (6^3 SU(3), 8 flux
    # quanta, M = 1.2: the smallest |h_W| is 0.22, above the cut 0.04 Lambda = 0.19; the host model
    # of the storage roundings puts the tol ladder at 1.2e-12 and the forced one at 1.1e-3.)
    print("[case] guard", flush=True)
    n, nc, nm = (6, 6, 6), 3, 128
    ctx, L, ucc, links, e, psi = make(nc, n, nm, 777, sigma=0.3, M=1.2, nb=8, mb=64)
    deg, cut = 240, 0.04
    bad = dict(degree=deg, cut=cut, precision=ovl.PREC_MIXED, tol=1e-10, head=6, head16=12, head48=2)
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, cut=cut))
    ref = ctx.get_modes(ovl.SCRATCH)
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(guard=0, **bad))
    unguarded = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(ctx.get_modes(ovl.SCRATCH), ref))
    assert unguarded > 1e-8
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(**bad))
    t = ctx.timing()
    guarded = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(ctx.get_modes(ovl.SCRATCH), ref))
    print(f"  unguarded {unguarded:.2e}, guarded {guarded:.2e}, checks {t.guard_checks}, level {t.guard_level}, "
          f"last check {t.guard_err:.2e}", flush=True)
    assert t.guard_checks == 2 and t.guard_level == 1 and t.guard_err <= 1e-10
    assert guarded < 1e-10
    # a leapfrog run with the same bad options: the context keeps the tightened ladder (no new check
    # is due: same links, same options, one apply since), and the run matches the oracle's FP64
    # leapfrog; asked to check every 2 steps, the next call checks again
    ctx.set_modes(ovl.PSIMID, psi)
    so = ovl.StepOpts(ovl.OVERLAP, 0.02, ovl.GAUGE_LINKS, ovl.sign_opts(**bad), 1)
    ctx.fermion_update(so, 3)
    t = ctx.timing()
    assert t.guard_level >= 1 and t.guard_checks == 0
    o = orc.StepOpts(orc.OVERLAP_CHEB, 0.02, 1, 8, 1.0, deg, cut, 0.0, 0.0, 8)
    _, p, pm, _ = orc.leapfrog(L, links, e, psi[:4], psi[:4], o, 3)
    assert rel(ctx.get_modes(ovl.PSI, 0, 4), p) < 1e-10 and rel(ctx.get_modes(ovl.PSIMID, 0, 4), pm) < 1e-10
    so.sign.guard = 2
    ctx.fermion_update(so, 1)
    t = ctx.timing()
    assert t.guard_checks == 1 and t.guard_level >= 1 and t.guard_err <= 1e-10
    # new links: the guard starts again from the requested ladder
    ctx.set_links(links)
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, tol=1e-10))
    t = ctx.timing()
    assert t.guard_level == 0 and t.guard_checks == 1 and t.guard_err <= 1e-10


def test_guard_falls_back_to_fp64_below_the_cut(gpu):
    # a rough lattice whose h_W spectrum reaches below the cut (smallest |h_W| 0.18 against
    # 0.08 Lambda = 0.38): the Clenshaw intermediates of those modes grow beyond the bounds the
    # ladder is planned from, and no mixed ladder meets 1e-10 (the host model: 3.7e-7 at the tol
    # ladder).  The guard tightens level by level and ends in plain FP64, whose result the apply
    # then returns bit for bit.
    print("[case] guard below the cut", flush=True)
    n, nc, nm = (6, 6, 6), 3, 64
    ctx, L, ucc, links, e, psi = make(nc, n, nm, 777, sigma=0.6, M=1.2, nb=8, mb=64)
    deg, cut = 160, 0.08
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, cut=cut))
    ref = ctx.get_modes(ovl.SCRATCH)
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, tol=1e-10))
    t = ctx.timing()
    print(f"  checks {t.guard_checks}, level {t.guard_level}, last check {t.guard_err:.2e}", flush=True)
    assert t.guard_level == 6 and t.guard_checks == 5 and t.guard_err > 1e-10
    assert np.array_equal(ctx.get_modes(ovl.SCRATCH), ref)
    # allowed to loosen, the guard still ends in FP64 here: loosening starts only from a level that
    # passes
    ctx.set_links(links)
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH,
                      ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, tol=1e-10, relax=2))
    t = ctx.timing()
    assert t.guard_level == 6
    assert np.array_equal(ctx.get_modes(ovl.SCRATCH), ref)


def test_guard_loosens_within_half_tol(gpu):
    # relax > 0: a check whose error leaves 4x to spare below tol / 2 tries the ladder of 4 tol on
    # every mode block and keeps it while that check stays within tol / 2.  On 6^3 SU(2) at degree 256
    # the tol ladder's error is 4e-12, the looser one's 2e-11 (a level further would need 8e-11):
    # one level.  Every mode of the apply, and of a 3-step leapfrog after it, stays within tol of
    # the FP64 form; without relax the ladder stays at the tol level.
    print("[case] guard loosening", flush=True)
    n, nc, nm, deg, tol = (6, 6, 6), 2, 64, 256, 1e-10
    ctx, L, ucc, links, e, psi = make(nc, n, nm, 32, sigma=0.25, M=1.2)
    with ctx:
        ctx.apply_sign(ovl.PSI, ovl.PSIMID, ovl.sign_opts(degree=deg, precision=ovl.PREC_FP64))
        ref = ctx.get_modes(ovl.PSIMID)
        o0 = ovl.sign_opts(degree=deg, precision=ovl.PREC_MIXED, tol=tol)
        ctx.apply_sign(ovl.PSI, ovl.SCRATCH, o0)
        t0 = ctx.timing()
        assert t0.guard_level == 0 and t0.guard_checks == 1
        o2 = ovl.sign_opts(degree=deg, precision=ovl.PREC_MIXED, tol=tol, relax=2)
        ctx.apply_sign(ovl.PSI, ovl.SCRATCH, o2)  # new options: the check is due again
        t = ctx.timing()
        mix = ctx.get_modes(ovl.SCRATCH)
        err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref))
        print(f"  level {t.guard_level}, checks {t.guard_checks}, last check {t.guard_err:.2e}, all modes "
              f"{err:.2e}, fp16 steps {t0.fp16_steps} -> {t.fp16_steps}", flush=True)
        assert -2 <= t.guard_level <= -1 and t.guard_err <= 0.5 * tol and err <= 0.5 * tol
        assert t.fp16_steps > t0.fp16_steps and t.fp64_steps < t0.fp64_steps
        # the leapfrog with the loosened ladder against the FP64 leapfrog from the same state
        runs = []
        for prec, relax in ((ovl.PREC_FP64, 0), (ovl.PREC_MIXED, 2)):
            ctx.set_links(links)
            ctx.set_modes(ovl.PSI, psi)
            ctx.set_modes(ovl.PSIMID, psi)
            so = ovl.sign_opts(degree=deg, precision=prec, tol=tol, relax=relax)
            ctx.fermion_update(ovl.StepOpts(ovl.OVERLAP, 0.02, ovl.GAUGE_LINKS, so, 1), 3)
            runs.append((ctx.get_modes(ovl.PSI), ctx.timing().guard_level))
        err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(runs[1][0], runs[0][0]))
        print(f"  leapfrog: level {runs[1][1]}, mixed vs fp64 after 3 steps {err:.2e}", flush=True)
        assert runs[1][1] < 0 and err <= tol
    # four mode blocks: the try of a looser ladder is checked on all of them, so the check's error is
    # the apply's error over every mode (the periodic checks take one block in turn)
    ctx, L, ucc, links, e, psi = make(nc, n, 128, 33, sigma=0.25, M=1.2, mb=32)
    with ctx:
        ctx.apply_sign(ovl.PSI, ovl.PSIMID, ovl.sign_opts(degree=deg, precision=ovl.PREC_FP64))
        ref = ctx.get_modes(ovl.PSIMID)
        ctx.apply_sign(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, precision=ovl.PREC_MIXED, tol=tol, relax=2))
        t = ctx.timing()
        mix = ctx.get_modes(ovl.SCRATCH)
        err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref))
        print(f"  4 blocks: level {t.guard_level}, checks {t.guard_checks}, check {t.guard_err:.4e}, "
              f"all modes {err:.4e}", flush=True)
        assert t.guard_level < 0 and err <= 0.5 * tol
        assert abs(t.guard_err - err) <= 1e-3 * err
    with pytest.raises(ovl.OvlError):
        with ovl.Context(nc, n, nm) as c2:
            c2.set_links(links)
            c2.apply_sign(ovl.PSI, ovl.PSIMID, ovl.sign_opts(precision=ovl.PREC_MIXED, relax=9))


def test_timed_mode_block(gpu):
    # mode_block = -1: every block size is timed on the lattice and the fastest kept; the context
    # then computes exactly what a context created with that block size computes
    import re
    print("[case] timed mode block", flush=True)
    n, nc, nm = (8, 8, 6), 3, 96
    g = cf.rng(880)
    links = cf.su3_links(n, 0.3, g)
    psi = cf.spinors(n, nc, nm, g)
    with ovl.Context(nc, n, nm, m_dw=1.2, mode_block=-1) as ctx:
        ctx.set_links(links)
        ctx.set_modes(ovl.PSI, psi)
        rep = ctx.autotune()
        m = re.search(r"mode block:((?: \d+ [0-9.]+ us)+), chosen (\d+)", rep)
        assert m, rep[:200]
        sizes = [int(v) for v in m.group(1).split()[0::3]]
        chosen = int(m.group(2))
        print(f"  {m.group(0)}", flush=True)
        assert sizes == [16, 32, 64] and chosen in sizes
        ctx.apply_wilson(ovl.PSI, ovl.PSIMID)
        a = ctx.get_modes(ovl.PSIMID)
    with ovl.Context(nc, n, nm, m_dw=1.2, mode_block=chosen) as ref:
        ref.set_links(links)
        ref.set_modes(ovl.PSI, psi)
        ref.apply_wilson(ovl.PSI, ovl.PSIMID)
        assert np.array_equal(a, ref.get_modes(ovl.PSIMID))
    # 20 modes: a 64-mode block would more than double the lanes and is not a candidate
    with ovl.Ensemble(2, [(4, 4, 4), (5, 4, 3)], 20, mode_block=-1) as ens:
        for k, d in enumerate([(4, 4, 4), (5, 4, 3)]):
            ens.set_system_links(k, cf.su2_links(d, 0.3, g))
        m = re.search(r"mode block:((?: \d+ [0-9.]+ us)+), chosen (\d+)", ens.autotune())
        assert m and [int(v) for v in m.group(1).split()[0::3]] == [16, 32]
