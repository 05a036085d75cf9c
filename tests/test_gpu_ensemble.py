"""GPU parity of ensemble contexts: many lattices of different sizes, each stencil application one
launch over all of them.  Every system is checked against the CPU oracle run on that system alone
(FP64 1e-12 relative, mixed FP32/FP64 1e-10 against the FP64 GPU run) and against an ordinary
one-lattice context holding the same system."""
import numpy as np
import pytest

import orc
from overlap_amd import configs as cf
from overlap_amd import ovl

pytestmark = pytest.mark.gpu

DIMS = [(4, 4, 4), (6, 5, 4), (3, 4, 5), (5, 5, 3)]
FLUX = [0, 2, 1, 3]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def systems(nc, nmodes, seed, dims=DIMS, sigma=0.4, esig=0.2):
    out = []
    for k, n in enumerate(dims):
        g = cf.rng(seed + k)
        links = cf.su2_links(n, sigma, g) if nc == 2 else cf.su3_links(n, sigma, g)
        e = cf.efield(n, nc, esig, g)
        psi = cf.spinors(n, nc, nmodes, g)
        out.append((n, links, e, psi))
    return out


def make_ensemble(nc, nmodes, sy, mb=0, r=1.0, m=0.0, M=0.0, flux=FLUX):
    ens = ovl.Ensemble(nc, [s[0] for s in sy], nmodes, r_wilson=r, m_fermion=m, m_dw=M, mode_block=mb)
    for k, (n, links, e, psi) in enumerate(sy):
        ens.set_system_links(k, links)
        ens.set_system_efield(k, e)
        if flux[k]:
            ens.set_system_u1_flux(k, flux[k], 1.0)
        ens.set_system_modes(k, ovl.PSI, psi)
    return ens


def oracle_lat(nc, n, nb, r=1.0, m=0.0, M=0.0, links=None):
    L = orc.lat(nc, n, r, m, M)
    ucc = orc.links_to_ucc(L, links, orc.u1_phases(n, nb, 1.0) if nb else None)
    return L, ucc


@pytest.mark.parametrize("nc,mb,nm", [(2, 16, 20), (2, 64, 64), (3, 32, 40), (3, 64, 8)])
def test_ensemble_wilson_vs_oracle_and_single(gpu, nc, mb, nm):
    print(f"[case] ensemble wilson nc={nc} mb={mb} modes={nm}", flush=True)
    sy = systems(nc, nm, 500 + nc + mb)
    ens = make_ensemble(nc, nm, sy, mb=mb, m=0.1, M=1.3)
    assert ens.balance() >= 1.0
    ens.apply_wilson(ovl.PSI, ovl.PSIMID)
    for k, (n, links, e, psi) in enumerate(sy):
        out = ens.get_system_modes(k, ovl.PSIMID)
        L, ucc = oracle_lat(nc, n, FLUX[k], 1.0, 0.1, 1.3, links)
        for j in range(0, nm, max(1, nm // 8)):
            err = rel(out[j], orc.wilson_evo(L, ucc, psi[j]))
            assert err < 1e-13, (k, j, err)
        with ovl.Context(nc, n, nm, m_fermion=0.1, m_dw=1.3, mode_block=mb) as ctx:
            ctx.set_links(links)
            if FLUX[k]:
                ctx.set_u1_flux(FLUX[k], 1.0)
            ctx.set_modes(ovl.PSI, psi)
            ctx.apply_wilson(ovl.PSI, ovl.PSIMID)
            single = ctx.get_modes(ovl.PSIMID)
        err = rel(out, single)
        print(f"  system {k} {n}: ensemble vs single context {err:.2e}", flush=True)
        assert err < 1e-14


@pytest.mark.parametrize("nc", [2, 3])
def test_ensemble_sign_fp64_and_mixed(gpu, nc):
    print(f"[case] ensemble sign nc={nc}", flush=True)
    nm, deg, cut = 16, 96, 0.1
    sy = systems(nc, nm, 610 + nc, sigma=0.3)
    ens = make_ensemble(nc, nm, sy, M=1.2)
    ens.apply_sign(ovl.PSI, ovl.PSIMID, ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_FP64))
    ens.apply_sign(ovl.PSI, ovl.SCRATCH, ovl.sign_opts(degree=deg, cut=cut, precision=ovl.PREC_MIXED, tol=1e-11))
    t = ens.timing()
    assert t.fp32_steps > 0
    coef = orc.cheb_coeffs(deg, cut)
    for k, (n, links, e, psi) in enumerate(sy):
        L, ucc = oracle_lat(nc, n, FLUX[k], 1.0, 0.0, 1.2, links)
        lam = orc.norm_bound(L)
        f64 = ens.get_system_modes(k, ovl.PSIMID)
        mix = ens.get_system_modes(k, ovl.SCRATCH)
        for j in (0, nm - 1):
            err = rel(f64[j], orc.sign_cheb(L, ucc, lam, cut, deg, coef, psi[j]))
            assert err < 1e-12, (k, j, err)
        emix = max(np.linalg.norm(mix[j] - f64[j]) / np.linalg.norm(psi[j]) for j in range(nm))
        print(f"  system {k} {n}: mixed vs fp64 {emix:.2e}", flush=True)
        assert emix < 1e-10
    # the accuracy guard's check ran on every mode of every system, each mode against its own FP64
    # norm within its own system: its error is the apply's largest per-system, per-mode error
    worst = max(max(np.linalg.norm(a - b) / np.linalg.norm(b)
                    for a, b in zip(ens.get_system_modes(k, ovl.SCRATCH), ens.get_system_modes(k, ovl.PSIMID)))
                for k in range(len(sy)))
    print(f"  guard: checks {t.guard_checks}, level {t.guard_level}, error {t.guard_err:.4e} (apply {worst:.4e})",
          flush=True)
    assert t.guard_checks >= 1 and abs(t.guard_err - worst) <= 1e-3 * worst
    r = ens.sign_residual(ovl.PSI, ovl.sign_opts(degree=256, cut=0.05, precision=ovl.PREC_FP64))
    assert r < 1e-8


@pytest.mark.parametrize("prec,graph", [(ovl.PREC_FP64, 0), (ovl.PREC_MIXED, 1)])
def test_ensemble_overlap_leapfrog_vs_oracle(gpu, prec, graph):
    print(f"[case] ensemble leapfrog prec={prec} graph={graph}", flush=True)
    nc, nm, deg, cut, dt, steps = 2, 8, 64, 0.1, 0.02, 3
    sy = systems(nc, nm, 720, sigma=0.3)
    ens = make_ensemble(nc, nm, sy, M=1.0)
    for k, s in enumerate(sy):
        ens.set_system_modes(k, ovl.PSIMID, s[3])
    so = ovl.StepOpts(kind=ovl.OVERLAP, dt=dt, gauge=ovl.GAUGE_LINKS,
                      sign=ovl.sign_opts(degree=deg, cut=cut, precision=prec, tol=1e-11), use_graph=graph)
    ens.fermion_update(so, steps)
    tol = 1e-12 if prec == ovl.PREC_FP64 else 1e-10
    for k, (n, links, e, psi) in enumerate(sy):
        L = orc.lat(nc, n, 1.0, 0.0, 1.0)
        oo = orc.StepOpts(orc.OVERLAP_CHEB, dt, 1, FLUX[k], 1.0, deg, cut, 0.0, 0.0, 8)
        lk, p, pm, _ = orc.leapfrog(L, links, e, psi, psi, oo, steps)
        err = rel(ens.get_system_links(k), lk)
        assert err < 1e-13, (k, err)
        err = rel(ens.get_system_modes(k, ovl.PSI), p)
        assert err < tol, (k, err)
        err = rel(ens.get_system_modes(k, ovl.PSIMID), pm)
        assert err < tol, (k, err)


def test_ensemble_wilson_leapfrog_full_gauge(gpu):
    nc, nm, dt, steps = 3, 16, 0.05, 4
    sy = systems(nc, nm, 800, sigma=0.3, esig=0.3)
    ens = make_ensemble(nc, nm, sy, m=0.1)
    for k, s in enumerate(sy):
        ens.set_system_modes(k, ovl.PSIMID, s[3])
    so = ovl.StepOpts(kind=ovl.WILSON, dt=dt, gauge=ovl.GAUGE_FULL, sign=ovl.sign_opts(), use_graph=1)
    ens.fermion_update(so, 1)
    assert ens.timing().graph_captures == 1
    ens.fermion_update(so, steps - 1)
    assert ens.timing().graph_captures == 0  # the kept graph of the first call replays
    for k, (n, links, e, psi) in enumerate(sy):
        L = orc.lat(nc, n, 1.0, 0.1, 0.0)
        oo = orc.StepOpts(orc.WILSON, dt, 2, FLUX[k], 1.0, 0, 0.0, 0.0, 0.0, 8)
        lk, p, pm, ee = orc.leapfrog(L, links, e, psi, psi, oo, steps)
        assert rel(ens.get_system_links(k), lk) < 1e-13
        assert rel(ens.get_system_efield(k), ee) < 1e-12
        assert rel(ens.get_system_modes(k, ovl.PSI), p) < 1e-11
        assert rel(ens.get_system_modes(k, ovl.PSIMID), pm) < 1e-11


@pytest.mark.parametrize("nc", [2, 3])
def test_ensemble_observables_and_norms(gpu, nc):
    nm = 12
    sy = systems(nc, nm, 900 + nc, sigma=0.6, esig=0.5)
    ens = make_ensemble(nc, nm, sy, M=1.0)
    norms = ens.mode_norms(ovl.PSI)
    assert norms.shape == (len(sy), nm)
    for k, (n, links, e, psi) in enumerate(sy):
        assert np.allclose(norms[k], (np.abs(psi) ** 2).sum(axis=1), rtol=1e-13, atol=0)
        L = orc.lat(nc, n, 1.0, 0.0, 1.0)
        ob = ens.system_observables(k)
        assert np.allclose((ob["e_el"], ob["e_mag"]), orc.plaquette_energy(L, links, e), rtol=1e-13, atol=0)
        assert np.allclose((ob["gauss_max"], ob["gauss_rms"]), orc.gauss_violation(L, links, e), rtol=1e-12, atol=0)
        u = orc.unitarity_violation(L, links)
        assert abs(ob["unit_max"] - u[0]) < 1e-15 and abs(ob["unit_rms"] - u[1]) < 1e-15
        eb = orc.edotb(L, links, e) / (8 * np.pi ** 2)
        assert abs(ob["ncs_dot"] - eb) < 1e-12 * abs(eb)


def test_ensemble_rejects_single_lattice_calls(gpu):
    sy = systems(2, 4, 950)
    ens = make_ensemble(2, 4, sy)
    for call in (lambda: ens.set_modes(ovl.PSI, sy[0][3]), lambda: ens.get_modes(ovl.PSI), ens.energy,
                 ens.ncs_dot,
                 lambda: ens.apply_overlap(ovl.PSI, ovl.PSIMID,
                                           ovl.sign_opts(method=ovl.SIGN_ZOLOTAREV, lmin=0.1, lmax=30.0))):
        with pytest.raises(ovl.OvlError) as ex:
            call()
        assert ex.value.code == ovl.E_STATE
    with pytest.raises(ovl.OvlError) as ex:
        ens.set_system_links(len(sy), sy[0][1])
    assert ex.value.code == ovl.E_ARG
    # the per-system calls also address system 0 of an ordinary context
    with ovl.Context(2, DIMS[0], 4) as ctx:
        ctx.set_system_modes(0, ovl.PSI, sy[0][3])
        assert np.array_equal(ctx.get_modes(ovl.PSI), sy[0][3])
        with pytest.raises(ovl.OvlError) as ex:
            ctx.get_system_modes(1, ovl.PSI)
        assert ex.value.code == ovl.E_ARG


@pytest.mark.parametrize("nc,mb,nm", [(2, 32, 32), (3, 64, 64), (2, 16, 20)])
def test_ensemble_work_queues_are_permutations(gpu, nc, mb, nm):
    # the persistent work-queue kernel (tuning traversal 4, 1 / 4 / 16 tiles per claim) runs the same
    # tiles as the static table in another order: a Wilson apply, a mixed sign (pure and epilogue
    # launches of every storage tier) and a graph-replayed leapfrog are bit-identical, and repeated
    # launches find the queues reset
    print(f"[case] ensemble queues nc={nc} mb={mb} modes={nm}", flush=True)
    dims = [(12, 12, 12), (6, 5, 4), (8, 8, 8), (3, 4, 5), (10, 10, 10)]
    sy = systems(nc, nm, 700 + nc, dims=dims)
    so = ovl.sign_opts(degree=40, cut=0.1, precision=ovl.PREC_MIXED, head48=3, head=10, head16=24, guard=0)
    step = ovl.StepOpts(ovl.OVERLAP, 0.02, ovl.GAUGE_LINKS, so, 1)
    outs = []
    for trav, per in ((1, 1), (4, 1), (4, 4), (4, 16)):
        ens = make_ensemble(nc, nm, sy, mb=mb, M=1.2, flux=[1, 0, 2, 0, 3])
        for p in range(4):
            t = ens.get_tuning(p)
            ens.set_tuning(p, trav, per, t.nt_store, t.waves_per_eu)
        ens.apply_wilson(ovl.PSI, ovl.PSIMID)
        r = [ens.get_system_modes(k, ovl.PSIMID) for k in range(len(dims))]
        ens.apply_sign(ovl.PSI, ovl.SCRATCH, so)
        r += [ens.get_system_modes(k, ovl.SCRATCH) for k in range(len(dims))]
        for k, s_ in enumerate(sy):
            ens.set_system_modes(k, ovl.PSIMID, s_[3])
        ens.fermion_update(step, 3)
        r += [ens.get_system_modes(k, ovl.PSI) for k in range(len(dims))]
        assert ens.get_tuning(0).traversal == trav
        outs.append(r)
        ens.close()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.isfinite(a).all() and np.array_equal(a, b)


def _random_ensemble(case, seed=4051):
    # (nc, system sides, flux per system, modes, mode block) of case number `case`, from one fixed
    # stream: 2-6 systems of uneven, odd sides, ragged mode blocks
    g = np.random.default_rng(seed)
    for _ in range(case + 1):
        nc = int(g.choice([2, 3]))
        ns = int(g.integers(2, 7))
        dims = [tuple(int(v) for v in g.integers(3, 8, size=3)) for _ in range(ns)]
        flux = [int(v) for v in g.integers(0, 4, size=ns)]
        nm = int(g.integers(1, 41))
        mb = int(g.choice([0, 16, 32, 64]))
    return nc, dims, flux, nm, mb


@pytest.mark.parametrize("case", range(4))
def test_ensemble_random_shapes(gpu, case):
    # drawn ensembles through the overlap leapfrog with moving links: each system in FP64 against an
    # ordinary context holding that system alone, the mixed ladder (every tier, the guard allowed
    # to loosen) against FP64 per system, and hipGraph replay bit-identical to stream launches
    nc, dims, flux, nm, mb = _random_ensemble(case)
    print(f"[case] random ensemble {case}: nc={nc} systems={dims} flux={flux} modes={nm} mb={mb}", flush=True)
    deg, cut, dt, steps, M = 256, 0.05, 0.02, 2, 1.2
    sy = systems(nc, nm, 700 + 10 * case, dims=dims, sigma=0.3)

    def opts(prec, graph):
        so = ovl.sign_opts(degree=deg, cut=cut, precision=prec, tol=1e-10, relax=2)
        return ovl.StepOpts(ovl.OVERLAP, dt, ovl.GAUGE_LINKS, so, graph)

    ens = make_ensemble(nc, nm, sy, mb=mb, M=M, flux=flux)
    runs = {}
    for key, prec, graph in (("fp64", ovl.PREC_FP64, 0), ("mixed", ovl.PREC_MIXED, 0), ("mixed_graph", ovl.PREC_MIXED, 1)):
        for k, (n, links, e, psi) in enumerate(sy):
            ens.set_system_links(k, links)
            ens.set_system_modes(k, ovl.PSI, psi)
            ens.set_system_modes(k, ovl.PSIMID, psi)
        ens.fermion_update(opts(prec, graph), steps)
        runs[key] = [(ens.get_system_modes(k, ovl.PSI), ens.get_system_modes(k, ovl.PSIMID)) for k in range(len(sy))]
    level = ens.timing().guard_level
    ens.close()
    worst = 0.0
    for k, (n, links, e, psi) in enumerate(sy):
        with ovl.Context(nc, n, nm, m_dw=M, mode_block=mb) as ctx:
            ctx.set_links(links)
            ctx.set_efield(e)
            if flux[k]:
                ctx.set_u1_flux(flux[k], 1.0)
            ctx.set_modes(ovl.PSI, psi)
            ctx.set_modes(ovl.PSIMID, psi)
            ctx.fermion_update(opts(ovl.PREC_FP64, 0), steps)
            single = (ctx.get_modes(ovl.PSI), ctx.get_modes(ovl.PSIMID))
        for f in range(2):
            assert rel(runs["fp64"][k][f], single[f]) < 1e-13, (k, f)
            assert np.array_equal(runs["mixed"][k][f], runs["mixed_graph"][k][f]), (k, f)
            worst = max(worst, max(np.linalg.norm(a - b) / np.linalg.norm(b)
                                   for a, b in zip(runs["mixed"][k][f], runs["fp64"][k][f])))
    print(f"  guard level {level}, mixed vs fp64 per system: max {worst:.2e}", flush=True)
    assert worst <= 1e-10
