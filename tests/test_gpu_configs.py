"""The five benchmark workloads at their full sizes on the GPU.  Sampled modes are checked against
the CPU oracle run on the same inputs (FP64: 1e-12 relative; a few modes, few steps, so the oracle
finishes in seconds); the whole state is checked through size-independent properties: the mixed
FP64/FP32/FP16 form against the FP64 form (1e-10 relative per mode), the leapfrog invariant and
the link unitarity.

Leapfrog invariant: a step is Psi' = Psi + dt A1 Mid, Mid' = Mid + dt A2 Psi' with A1, A2 = -i h
anti-Hermitian (OverlapUpdate.cpp:9-58; A2 uses the links half a step later), so
<Psi', Mid'> = <Psi, Mid> + dt <A1 Mid, Mid> + dt <Psi', A2 Psi'> changes only by imaginary terms:
Re<Psi, Mid> is conserved exactly, for every mode, whatever the links do.  |Psi|^2 itself is
not: with Mid = Psi at t = 0 it swings by O((dt h)^2)."""
import numpy as np
import pytest

import orc
from overlap_amd import configs as cf
from overlap_amd import ovl

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(c: cf.Config, nmodes=None):
    links, e, psi = c.generate(nmodes)
    ctx = ovl.Context(c.nc, c.n, psi.shape[0], r_wilson=c.r_wilson, m_fermion=c.m_fermion, m_dw=c.m_dw)
    ctx.set_links(links)
    ctx.set_efield(e)
    if c.nb:
        ctx.set_u1_flux(c.nb, c.q)
    ctx.set_modes(ovl.PSI, psi)
    ctx.set_modes(ovl.PSIMID, psi)
    return ctx, links, e, psi


def stagger_invariant(ctx, chunk=64):
    """Re<Psi_i, PsiMid_i> of every mode, fetched in chunks."""
    out = []
    for f in range(0, ctx.n_modes, chunk):
        n = min(chunk, ctx.n_modes - f)
        p, m = ctx.get_modes(ovl.PSI, f, n), ctx.get_modes(ovl.PSIMID, f, n)
        out.append(np.einsum("ij,ij->i", p.conj(), m).real)
    return np.concatenate(out)


def step_opts(c: cf.Config, prec, graph=1, relax=0):
    kind = ovl.OVERLAP if c.kind == "overlap" else ovl.WILSON
    so = ovl.sign_opts(degree=max(c.cheb_n, 1), cut=c.cheb_cut, precision=prec, tol=1e-10, relax=relax)
    return ovl.StepOpts(kind=kind, dt=c.dt, gauge=ovl.GAUGE_LINKS, sign=so, use_graph=graph)


def oracle_steps(c: cf.Config, links, e, psi, steps, threads=8):
    L = orc.lat(c.nc, c.n, c.r_wilson, c.m_fermion, c.m_dw)
    kind = orc.OVERLAP_CHEB if c.kind == "overlap" else orc.WILSON
    o = orc.StepOpts(kind, c.dt, 1, c.nb, c.q, c.cheb_n, c.cheb_cut, 0.0, 0.0, threads)
    return orc.leapfrog(L, links, e, psi, psi, o, steps)


def test_config0_su2_wilson_8_full_run(gpu):
    # SU(2) 8^3, 16 modes, 100 steps at dt 0.05: every mode against the oracle
    c = cf.CONFIGS["su2_wilson_8"]
    print(f"[case] {c.name} {c.steps} steps", flush=True)
    ctx, links, e, psi = setup(c)
    g0 = ctx.gauss_violation()
    ctx.fermion_update(step_opts(c, ovl.PREC_FP64), c.steps)
    lk, p, pm, _ = oracle_steps(c, links, e, psi, c.steps)
    assert rel(ctx.get_links(), lk) < 1e-12
    assert rel(ctx.get_modes(ovl.PSI), p) < 1e-11 and rel(ctx.get_modes(ovl.PSIMID), pm) < 1e-11
    assert np.abs(stagger_invariant(ctx) - 1).max() < 1e-12
    # the links move with E fixed; they stay unitary
    assert ctx.unitarity_violation()[0] < 1e-13
    assert np.isfinite(g0).all()


def test_config1_su2_overlap_16_full_size(gpu):
    # SU(2) 16^3, 64 modes, degree 256: two modes against the oracle (FP64, 1 step); all modes
    # mixed against FP64 (2 steps); sign residual of the mixed form
    c = cf.CONFIGS["su2_overlap_16"]
    print(f"[case] {c.name}", flush=True)
    ctx, links, e, psi = setup(c)
    ctx.fermion_update(step_opts(c, ovl.PREC_FP64), 1)
    _, p, pm, _ = oracle_steps(c, links, e, psi[:2], 1)
    assert rel(ctx.get_modes(ovl.PSI, 0, 2), p) < 1e-12 and rel(ctx.get_modes(ovl.PSIMID, 0, 2), pm) < 1e-12
    f64 = [ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID)]
    ctx2, *_ = setup(c)
    ctx2.fermion_update(step_opts(c, ovl.PREC_MIXED), 1)
    for f, ref in zip((ovl.PSI, ovl.PSIMID), f64):
        mix = ctx2.get_modes(f)
        err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref))
        print(f"  mixed vs fp64 field {f}: {err:.2e}", flush=True)
        assert err < 1e-10
    assert np.abs(stagger_invariant(ctx2) - 1).max() < 1e-10
    r = ctx2.sign_residual(ovl.PSI, step_opts(c, ovl.PREC_MIXED).sign)
    print(f"  sign residual {r:.2e}", flush=True)
    assert r < 1e-9


def test_config2_su3_wilson_32_full_size(gpu):
    # SU(3) 32^3 with 4 flux quanta, 512 modes in one batched launch, 3 steps: two modes against
    # the oracle, norms of all 512
    c = cf.CONFIGS["su3_wilson_32"]
    print(f"[case] {c.name}", flush=True)
    ctx, links, e, psi = setup(c)
    steps = 3
    ctx.fermion_update(step_opts(c, ovl.PREC_FP64), steps)
    sample = [0, 511]
    lk, p, pm, _ = oracle_steps(c, links, e, psi[sample], steps)
    assert rel(ctx.get_links(), lk) < 1e-12
    for i, s in enumerate(sample):
        assert rel(ctx.get_modes(ovl.PSI, s, 1)[0], p[i]) < 1e-12
        assert rel(ctx.get_modes(ovl.PSIMID, s, 1)[0], pm[i]) < 1e-12
    inv = stagger_invariant(ctx)
    print(f"  leapfrog invariant deviation over 512 modes: {np.abs(inv - 1).max():.2e}", flush=True)
    assert inv.shape == (512,) and np.abs(inv - 1).max() < 1e-12


def test_config3_su3_overlap_24_full_size(gpu):
    # the bench workload: SU(3) 24^3, 8 flux quanta, 256 modes, degree 384: one overlap apply of
    # one mode against the oracle (FP64), the mixed form against FP64 for every mode
    c = cf.CONFIGS["su3_overlap_24"]
    print(f"[case] {c.name}", flush=True)
    ctx, links, e, psi = setup(c)
    o64 = step_opts(c, ovl.PREC_FP64).sign
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, o64)
    L = orc.lat(c.nc, c.n, c.r_wilson, c.m_fermion, c.m_dw)
    ucc = orc.links_to_ucc(L, links, orc.u1_phases(c.n, c.nb, c.q))
    coef = orc.cheb_coeffs(c.cheb_n, c.cheb_cut)
    want = orc.overlap_evo_cheb(L, ucc, orc.norm_bound(L), c.cheb_cut, c.cheb_n, coef, psi[17])
    got = ctx.get_modes(ovl.SCRATCH, 17, 1)[0]
    assert rel(got, want) < 1e-12
    ref = ctx.get_modes(ovl.SCRATCH)
    ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, step_opts(c, ovl.PREC_MIXED).sign)
    t = ctx.timing()
    mix = ctx.get_modes(ovl.SCRATCH)
    err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref))
    print(f"  mixed vs fp64 over 256 modes: {err:.2e} (steps {t.fp64_steps}/{t.fp32_steps}/{t.fp16_steps})",
          flush=True)
    assert err < 1e-10 and t.fp16_steps > 0


def test_config4_ensemble_full_size(gpu):
    # 32 SU(2) systems of uneven sizes: one overlap apply of every system, mixed against FP64 per
    # system, and the smallest system's first mode against the oracle
    ec = cf.ENSEMBLES["su2_overlap_ensemble"]
    print(f"[case] {ec.name}", flush=True)
    sy = [ec.system(k) for k in range(ec.n_systems)]
    ens = ovl.Ensemble(ec.nc, [s[0] for s in sy], ec.nmodes, m_dw=ec.m_dw)
    for k, (n, nb, links, e, psi) in enumerate(sy):
        ens.set_system_links(k, links)
        ens.set_system_efield(k, e)
        ens.set_system_u1_flux(k, nb, ec.q)
        ens.set_system_modes(k, ovl.PSI, psi)
    so = lambda prec: ovl.sign_opts(degree=ec.cheb_n, cut=ec.cheb_cut, precision=prec, tol=1e-10)  # noqa: E731
    ens.apply_overlap(ovl.PSI, ovl.SCRATCH, so(ovl.PREC_FP64))
    ref = [ens.get_system_modes(k, ovl.SCRATCH) for k in range(ec.n_systems)]
    ens.apply_overlap(ovl.PSI, ovl.SCRATCH, so(ovl.PREC_MIXED))
    worst = 0.0
    for k in range(ec.n_systems):
        mix = ens.get_system_modes(k, ovl.SCRATCH)
        worst = max(worst, max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref[k])))
    print(f"  worst system mixed vs fp64: {worst:.2e}", flush=True)
    assert worst < 1e-10
    k = min(range(ec.n_systems), key=lambda i: np.prod(sy[i][0]))
    n, nb, links, e, psi = sy[k]
    L = orc.lat(ec.nc, n, ec.r_wilson, ec.m_fermion, ec.m_dw)
    ucc = orc.links_to_ucc(L, links, orc.u1_phases(n, nb, ec.q) if nb else None)
    coef = orc.cheb_coeffs(ec.cheb_n, ec.cheb_cut)
    want = orc.overlap_evo_cheb(L, ucc, orc.norm_bound(L), ec.cheb_cut, ec.cheb_n, coef, psi[0])
    assert rel(ref[k][0], want) < 1e-12


def test_config1_su2_overlap_16_all_50_steps(gpu):
    # the workload's full run: 50 leapfrog steps of 64 modes, the mixed form (the guard only
    # tightening, and allowed to loosen by two levels) against FP64 for every mode (the links move
    # with E fixed and do not depend on the modes, so the runs see the same operators and differ
    # only by the mixed form's rounding, accumulated over 100 overlap applies); the sign residual
    # ||eps(eps psi) - psi|| of the final modes
    c = cf.CONFIGS["su2_overlap_16"]
    print(f"[case] {c.name} {c.steps} steps", flush=True)
    runs = []
    for prec, relax in ((ovl.PREC_FP64, 0), (ovl.PREC_MIXED, 0), (ovl.PREC_MIXED, 2)):
        ctx, links, e, psi = setup(c)
        ctx.fermion_update(step_opts(c, prec, relax=relax), c.steps)
        runs.append((ctx, ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID), int(ctx.timing().guard_level)))
    c64, p64, m64, _ = runs[0]
    assert np.abs(stagger_invariant(c64) - 1).max() < 1e-12
    for relax, (cmx, pmx, mmx, level) in zip((0, 2), runs[1:]):
        err = max(max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(x, y))
                  for x, y in ((pmx, p64), (mmx, m64)))
        so = step_opts(c, ovl.PREC_MIXED, relax=relax).sign
        r = cmx.sign_residual(ovl.PSI, so)
        print(f"  relax {relax} (guard level {level}): mixed vs fp64 after {c.steps} steps: {err:.2e}; "
              f"sign residual {r:.2e}", flush=True)
        assert err < 1e-10
        assert r < 1e-9
        assert np.array_equal(cmx.get_links(), c64.get_links())
        assert np.abs(stagger_invariant(cmx) - 1).max() < 1e-10
        assert level <= 0 if relax == 0 else level >= -relax


def test_config2_su3_wilson_32_all_200_steps(gpu):
    # the workload's full run: 200 steps of 512 modes on 32^3 with 4 flux quanta: the leapfrog
    # invariant of every mode, the links' unitarity, and the first mode against a 3-step-shorter
    # run continued by 3 oracle steps (the oracle continues the GPU's state exactly)
    c = cf.CONFIGS["su3_wilson_32"]
    print(f"[case] {c.name} {c.steps} steps", flush=True)
    ctx, links, e, psi = setup(c)
    ctx.fermion_update(step_opts(c, ovl.PREC_FP64), c.steps - 3)
    lk, p, pm = ctx.get_links(), ctx.get_modes(ovl.PSI, 0, 1), ctx.get_modes(ovl.PSIMID, 0, 1)
    ctx.fermion_update(step_opts(c, ovl.PREC_FP64), 3)
    L = orc.lat(c.nc, c.n, c.r_wilson, c.m_fermion, c.m_dw)
    o = orc.StepOpts(orc.WILSON, c.dt, 1, c.nb, c.q, 0, 0.0, 0.0, 0.0, 8)
    lk3, p3, pm3, _ = orc.leapfrog(L, lk, e, p, pm, o, 3)
    assert rel(ctx.get_links(), lk3) < 1e-12
    assert rel(ctx.get_modes(ovl.PSI, 0, 1), p3) < 1e-12 and rel(ctx.get_modes(ovl.PSIMID, 0, 1), pm3) < 1e-12
    inv = stagger_invariant(ctx)
    print(f"  leapfrog invariant deviation after {c.steps} steps: {np.abs(inv - 1).max():.2e}", flush=True)
    assert np.abs(inv - 1).max() < 1e-11
    assert ctx.unitarity_violation()[0] < 1e-12


def test_config3_su3_overlap_24_all_100_steps(gpu):
    # the bench workload's full run (256 modes, degree 384, 8 flux quanta, 100 steps at dt 0.02) with
    # its observables after every step: the mixed form, with the guard only tightening and with it
    # allowed to loosen by two levels, against FP64 for every mode at the end, the leapfrog
    # invariant, the dN_CS/dt history (gauge fields only: the same bits in every run) and the axial
    # charge history every 5 steps
    c = cf.CONFIGS["su3_overlap_24"]
    print(f"[case] {c.name} {c.steps} steps", flush=True)
    runs = []
    for prec, relax in ((ovl.PREC_FP64, 0), (ovl.PREC_MIXED, 0), (ovl.PREC_MIXED, 2)):
        ctx, links, e, psi = setup(c)
        q5, rate, _, _ = ctx.fermion_update_observed(step_opts(c, prec, relax=relax), c.steps, every=1)
        runs.append((ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID), stagger_invariant(ctx), q5, rate,
                     int(ctx.timing().guard_level)))
        ctx.close()
    p64, m64, i64, q64, r64, _ = runs[0]
    assert np.abs(i64 - 1).max() < 1e-12
    for relax, (pmx, mmx, imx, qmx, rmx, level) in zip((0, 2), runs[1:]):
        err = max(max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(x, y))
                  for x, y in ((pmx, p64), (mmx, m64)))
        dq = np.abs(qmx - q64).max()
        print(f"  relax {relax} (guard level {level}): mixed vs fp64 after {c.steps} steps over 256 modes: "
              f"{err:.2e}; axial charge {q64[4::5]} (largest mixed-fp64 difference {dq:.2e}); "
              f"dN_CS/dt {r64[4::5]}", flush=True)
        assert err < 1e-10
        assert np.abs(imx - 1).max() < 1e-10
        assert len(r64) == c.steps and np.array_equal(r64, rmx) and np.isfinite(r64).all()
        # each mode adds a bilinear of norm <= (1 + |h_ov| / 2) |psi|^2 / 2V <= 4 / 2V to the charge
        # (unit modes, |h_ov| <= 2 rho): a relative error of 1e-10 per mode moves it by at most
        # 4e-10 M / 2V
        assert dq <= 1e-10 * 4 * c.nmodes / (2 * c.volume)
        assert level <= 0 if relax == 0 else level >= -relax


def test_config2_su3_wilson_32_fp32_all_200_steps(gpu):
    # the workload as specified: FP32 spinor arithmetic (FP32 copies of Psi and PsiMid stepped by the
    # FP32 stencil with the axpy in its epilogue), links and observables FP64.  All 512 modes after
    # 200 steps against the FP64 run of the same inputs (itself pinned to the oracle above), the
    # FP32 rounding allowed 2e-5 relative per mode; two modes over 3 steps against the oracle; the
    # leapfrog invariant to FP32 accuracy.
    c = cf.CONFIGS["su3_wilson_32"]
    print(f"[case] {c.name} fp32 {c.steps} steps", flush=True)
    ctx, links, e, psi = setup(c)
    ctx.fermion_update(step_opts(c, ovl.PREC_FP32), 3)
    sample = [0, 511]
    lk, p, pm, _ = oracle_steps(c, links, e, psi[sample], 3)
    for i, s in enumerate(sample):
        assert rel(ctx.get_modes(ovl.PSI, s, 1)[0], p[i]) < 2e-5
        assert rel(ctx.get_modes(ovl.PSIMID, s, 1)[0], pm[i]) < 2e-5
    ctx.fermion_update(step_opts(c, ovl.PREC_FP32), c.steps - 3)
    c64, *_ = setup(c)
    c64.fermion_update(step_opts(c, ovl.PREC_FP64), c.steps)
    assert np.array_equal(ctx.get_links(), c64.get_links())
    worst = 0.0
    for f in (ovl.PSI, ovl.PSIMID):
        for m0 in range(0, 512, 64):
            a, b = ctx.get_modes(f, m0, 64), c64.get_modes(f, m0, 64)
            worst = max(worst, max(np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(a, b)))
    inv = stagger_invariant(ctx)
    print(f"  fp32 vs fp64 after {c.steps} steps over 512 modes: {worst:.2e}; invariant {np.abs(inv - 1).max():.2e}",
          flush=True)
    assert worst < 2e-5
    assert np.abs(inv - 1).max() < 1e-5
    t = ctx.timing()
    assert t.var_launches[1] > 0 and t.var_launches[0] == 0  # FP32 stencil only


def test_config4_ensemble_all_50_steps(gpu):
    # the ensemble workload at its full length: 32 SU(2) overlap systems of uneven sizes, 50 leapfrog
    # steps of the mixed form (the guard only tightening, and allowed to loosen by two levels) and of
    # the FP64 form from the same state, every mode of every system compared (1e-10 relative per
    # mode), and each system's links against the oracle's link update
    ec = cf.ENSEMBLES["su2_overlap_ensemble"]
    print(f"[case] {ec.name} {ec.steps} steps", flush=True)
    sy = [ec.system(k) for k in range(ec.n_systems)]
    ens = ovl.Ensemble(ec.nc, [s[0] for s in sy], ec.nmodes, m_dw=ec.m_dw)
    runs, levels = [], []
    for prec, relax in ((ovl.PREC_FP64, 0), (ovl.PREC_MIXED, 0), (ovl.PREC_MIXED, 2)):
        for k, (n, nb, links, e, psi) in enumerate(sy):
            ens.set_system_links(k, links)
            ens.set_system_efield(k, e)
            ens.set_system_u1_flux(k, nb, ec.q)
            ens.set_system_modes(k, ovl.PSI, psi)
            ens.set_system_modes(k, ovl.PSIMID, psi)
        so = ovl.sign_opts(degree=ec.cheb_n, cut=ec.cheb_cut, precision=prec, tol=1e-10, relax=relax)
        ens.fermion_update(ovl.StepOpts(ovl.OVERLAP, ec.dt, ovl.GAUGE_LINKS, so, 1), ec.steps)
        runs.append([(ens.get_system_modes(k, ovl.PSI), ens.get_system_modes(k, ovl.PSIMID))
                     for k in range(ec.n_systems)])
        levels.append(int(ens.timing().guard_level))
    for relax, mixed, level in ((0, runs[1], levels[1]), (2, runs[2], levels[2])):
        worst = []
        for mix, ref in zip(mixed, runs[0]):
            worst.append(max(max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(x, y))
                             for x, y in zip(mix, ref)))
        print(f"  relax {relax} (guard level {level}): per-system mixed vs fp64 after {ec.steps} steps: "
              f"max {max(worst):.2e} min {min(worst):.2e}", flush=True)
        assert max(worst) < 1e-10
        assert level <= 0 if relax == 0 else level >= -relax
    for k in (0, ec.n_systems - 1):
        n, nb, links, e, psi = sy[k]
        L = orc.lat(ec.nc, n)
        want = links
        for _ in range(2 * ec.steps):
            want = orc.update_links(L, want, e, 0.5 * ec.dt)
        assert rel(ens.get_system_links(k), want) < 1e-12
