"""Benchmark of the OVERLAP fermion hot path on one MI355X.

Workload (default): SU(3) overlap fermions on a 24^3 lattice, U(1) background with 8 flux
quanta, 256 mode functions, Chebyshev sign of degree 384 (cut 0.05), rho = 1.2, dt = 0.02.
A step is one full leapfrog step of all modes: two overlap half steps (each an overlap
apply of every mode plus the axpy into the other field) and two gauge-link half steps.

value = site*mode updates/s of the mixed form (the Clenshaw recurrence stored in FP64, 48-bit,
FP32 and FP16 tiers, its accuracy guard allowed to loosen the ladder by --relax levels), inputs
resident in HBM, timed as hipGraph replay of the captured step; the plain FP64 form and the CPU
oracle are timed beside it, with the roofline of the dominant stencil variant (HIP events around
every launch of an eager step) and its HBM traffic (two rocprofv3 counter passes run as child
processes).  Also timed: the 32-system ensemble, the three smaller workloads over their own run
lengths, and the calls either side of the leapfrog.  Prints one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import csv
import platform
import re
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from overlap_amd import configs as cf  # noqa: E402
from overlap_amd import ovl  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "site·mode updates/s + %HBM roofline, mixed vs FP64 GPU at ≤10⁻¹⁰ overlap residual"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def step_opts(c: cf.Config, prec, graph, tol=1e-10, chunk=0, relax=0):
    kind = ovl.OVERLAP if c.kind == "overlap" else ovl.WILSON
    so = ovl.sign_opts(degree=max(c.cheb_n, 1), cut=c.cheb_cut, precision=prec, tol=tol, mode_chunk=chunk,
                       relax=relax)
    return ovl.StepOpts(kind=kind, dt=c.dt, gauge=ovl.GAUGE_LINKS, sign=so, use_graph=int(graph))


def host_description() -> dict:
    """nproc, the CPUs this process may run on, OMP_NUM_THREADS and the CPU model name."""
    model = ""
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        for line in r.stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if not model and os.path.exists("/proc/cpuinfo"):
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "model": model or platform.processor() or platform.machine()}


def default_cpu_threads() -> int:
    """The CPU share of this process: OMP_NUM_THREADS when set (the GPU box sets it to its 16-core
    share per GPU), else the CPUs it may run on."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    if n > 0:
        return n
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def cpu_baseline(c: cf.Config, links, e, psi, threads: int, modes: int):
    """The oracle (plain C FP64 restatement, OpenMP over modes) on a bounded sample:
    `modes` mode functions, one full leapfrog step, on `threads` host threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    L = orc.lat(c.nc, c.n, c.r_wilson, c.m_fermion, c.m_dw)
    kind = orc.OVERLAP_CHEB if c.kind == "overlap" else orc.WILSON
    o = orc.StepOpts(kind, c.dt, 1, c.nb, c.q, c.cheb_n, c.cheb_cut, 0.0, 0.0, threads)
    sample = psi[:modes]
    t0 = time.perf_counter()
    orc.leapfrog(L, links, e, sample, sample, o, 1)
    dt = time.perf_counter() - t0
    return c.volume * modes / dt, dt


def _pmc_pass(counter: str, child_args: list, outdir: str, timeout: float):
    """One rocprofv3 counter pass over a short run of this script (a child process); returns the
    dispatch rows of its counter CSV, or None."""
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    d = os.path.join(outdir, counter.lower())
    cmd = [rp, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "c", "--", sys.executable,
           os.path.join(ROOT, "bench.py")] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    with open(os.path.join(outdir, counter.lower() + ".log"), "w") as lg:
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=lg, stderr=subprocess.STDOUT, start_new_session=True)
        try:
            p.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            log(f"{counter} pass timed out")
            return None
    f = os.path.join(d, "c_counter_collection.csv")
    if not os.path.exists(f):  # the profiler may end badly after the program; use complete results only
        log(f"{counter} pass: no counter file (rc {p.returncode})")
        return None
    with open(f) as fh:
        return list(csv.DictReader(fh))


def _kernel_name(n: str) -> str:
    """Kernel name as in the source (rocprofv3 leaves the _Float16 instantiations mangled)."""
    m = re.match(r"_ZN3ovl12_GLOBAL__N_1\d+(k_hop(?:_ens)?)ILi(\d)EDF16_DF16_Li(\d+)ELi(\d)E(?:Lb(\d)E)?", n)
    if m:
        g = m.groups()
        epi = "" if g[4] is None else (", true" if g[4] == "1" else ", false")
        return "%s<%s, _Float16, _Float16, %s, %s%s>" % (g[0], g[1], g[2], g[3], epi)
    return n.replace("void ", "").replace("ovl::(anonymous namespace)::", "")


def tuning_string(ctx) -> str:
    """The context's stencil tuning in the --tuning format."""
    parts = []
    for p, name in ((0, "64"), (1, "32"), (2, "16"), (3, "48")):
        t = ctx.get_tuning(p)
        parts.append(f"{name}:{t.traversal}/{t.slabs}/{t.nt_store}/{t.waves_per_eu}")
    return ",".join(parts)


def measure_traffic(args, c: cf.Config, prefix: str, n_last: int, tuning: str):
    """HBM bytes per launch of the dominant stencil variant from FETCH_SIZE and WRITE_SIZE, each
    scaled by what the same counter reads for the calibration copy of one FP64 field (known bytes
    read and written).  The counted run is `bench.py --quick --steps 1` of the same workload; the
    last `n_last` dispatches whose name starts with `prefix` are its timed step."""
    child = ["--quick", "--steps", "1", "--warmup", "0", "--workload", c.name, "--mb", str(args.mb),
             "--traversal", str(args.traversal), "--tol", str(args.tol), "--relax", str(args.relax), "--chunk", str(args.chunk),
             "--tuning", tuning]
    if args.modes:
        child += ["--modes", str(args.modes)]
    outdir = tempfile.mkdtemp(prefix="ovl_pmc_", dir="/tmp")
    field_bytes = c.volume * c.nmodes * 4 * c.nc * 16
    field_cplx = c.volume * c.nmodes * 4 * c.nc
    res = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            log(f"traffic: {counter} pass")
            rows = _pmc_pass(counter, child, outdir, 300)
            if not rows:
                return None
            name = lambda r: _kernel_name(r["Kernel_Name"])  # noqa: E731
            calib = [r for r in rows if name(r).startswith("k_lincomb<double>") and int(r["Grid_Size"]) == field_cplx]
            hops = [r for r in rows if name(r).startswith(prefix)]
            if not calib or len(hops) < n_last:
                return None
            factor = field_bytes / (float(calib[0]["Counter_Value"]) * 1024.0)
            per = [float(r["Counter_Value"]) * 1024.0 * factor for r in hops[-n_last:]]
            # the variant's launches of the step: pure h_W and epilogue instantiations, averaged
            kinds = sorted({name(r).split("(")[0] for r in hops[-n_last:]})
            res[counter] = {"factor": factor, "bytes_per_launch": sum(per) / len(per), "kernel": " + ".join(kinds)}
    finally:
        shutil.rmtree(outdir, ignore_errors=True)
    return res


def bench_ensemble(ec: cf.EnsembleConfig, steps: int, warmup: int, tol: float, relax: int = 0, graph: bool = True):
    """Many SU(2) overlap systems of uneven sizes in one context (one stencil launch per Clenshaw
    step over all systems): mixed and FP64 timing from the same start (hipGraph replay of the
    step, captured in an untimed warmup step), per-system error of the mixed form against the FP64
    form after the warmup and `steps` steps; the stencil rate from one eager step with HIP events
    around every launch."""
    t0 = time.perf_counter()
    sy = [ec.system(k) for k in range(ec.n_systems)]
    dims = [s[0] for s in sy]
    sites = sum(int(np.prod(d)) for d in dims)
    log(f"ensemble {ec.name}: {ec.n_systems} systems, {sites} sites, inputs in {time.perf_counter() - t0:.1f}s")
    ens = ovl.Ensemble(ec.nc, dims, ec.nmodes, r_wilson=ec.r_wilson, m_fermion=ec.m_fermion, m_dw=ec.m_dw)

    def load():
        for k, (n, nb, links, e, psi) in enumerate(sy):
            ens.set_system_links(k, links)
            ens.set_system_efield(k, e)
            ens.set_system_u1_flux(k, nb, ec.q)
            ens.set_system_modes(k, ovl.PSI, psi)
            ens.set_system_modes(k, ovl.PSIMID, psi)

    def opts(prec, graph=False):
        so = ovl.sign_opts(degree=ec.cheb_n, cut=ec.cheb_cut, precision=prec, tol=tol, relax=relax)
        return ovl.StepOpts(kind=ovl.OVERLAP, dt=ec.dt, gauge=ovl.GAUGE_LINKS, sign=so, use_graph=int(graph))

    load()
    tune = ens.autotune()
    # per-system error of one overlap apply (-i h_ov psi), mixed against FP64
    apply_err = []
    ens.apply_overlap(ovl.PSI, ovl.SCRATCH, opts(ovl.PREC_FP64).sign)
    ref = [ens.get_system_modes(k, ovl.SCRATCH) for k in range(ec.n_systems)]
    ens.apply_overlap(ovl.PSI, ovl.SCRATCH, opts(ovl.PREC_MIXED).sign)
    for k in range(ec.n_systems):
        mix = ens.get_system_modes(k, ovl.SCRATCH)
        apply_err.append(float(max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref[k]))))
    del ref
    ens.fermion_update(opts(ovl.PREC_MIXED), warmup)
    out = {}
    for name, prec in (("mixed", ovl.PREC_MIXED), ("fp64", ovl.PREC_FP64)):
        load()
        # new links make the mixed sign's accuracy guard due: its check runs here, untimed (in a
        # run it recurs every 100 steps)
        ens.apply_overlap(ovl.PSI, ovl.SCRATCH, opts(prec).sign)
        ens.fermion_update(opts(prec, graph), 1)  # untimed warmup step: captures the step graph
        t0 = time.perf_counter()
        ens.fermion_update(opts(prec, graph), steps)
        wall = time.perf_counter() - t0
        psi = [ens.get_system_modes(k, ovl.PSI) for k in range(ec.n_systems)]
        # the stencil's rate: one eager step with HIP events around every stencil launch
        ens.set_profiling(True)
        ens.fermion_update(opts(prec), 1)
        ens.set_profiling(False)
        t = ens.timing()
        var = int(np.argmax([t.var_ms[i] for i in range(5)]))
        out[name] = {"ms_per_step": 1e3 * wall / steps, "value": sites * ec.nmodes * steps / wall,
                     "kernel": ["k_hop_ens FP64", "k_hop_ens FP32", "k_hop_ens FP32->FP64", "k_hop_ens FP16",
                                "k_hop_ens 48-bit"][var],
                     "kernel_GBps": t.var_bytes[var] / max(t.var_ms[var], 1e-9) / 1e6,
                     "kernel_share_of_eager_step": t.var_ms[var] / max(t.total_ms, 1e-9),
                     "clenshaw_steps_of_one_step": {"fp64": t.fp64_steps, "fp48": t.fp48_steps,
                                                    "fp32": t.fp32_steps, "fp16": t.fp16_steps},
                     "guard_level": int(t.guard_level)}
        out[name]["psi"] = psi
        log(f"ensemble {name}: {out[name]['ms_per_step']:.1f} ms/step {out[name]['value']:.4e} site*mode/s, "
            f"{out[name]['kernel']} {out[name]['kernel_GBps']:.0f} GB/s")
    errs = [float(max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(m, f)))
            for m, f in zip(out["mixed"].pop("psi"), out["fp64"].pop("psi"))]
    res = {"workload": ec.name, "systems": ec.n_systems, "sites": sites, "modes_per_system": ec.nmodes,
           "chebyshev_degree": ec.cheb_n, "steps": steps, "xcd_imbalance": round(ens.balance(), 4),
           "autotune": tune.split(";")[-1].strip(), "mixed": out["mixed"], "fp64": out["fp64"],
           "speedup_vs_fp64": out["mixed"]["value"] / out["fp64"]["value"],
           "per_system_rel_err_vs_fp64": [float(f"{e:.3e}") for e in errs],
           "per_system_overlap_apply_rel_err_vs_fp64": [float(f"{e:.3e}") for e in apply_err],
           "systems_layout": [[d[0], nb] for (d, nb, *_r) in sy]}
    ens.close()
    log(f"ensemble per-system err vs fp64: evolved modes max {max(errs):.3e}, one overlap apply max "
        f"{max(apply_err):.3e}")
    return res


VARIANT_NAMES = ["k_hop FP64", "k_hop FP32", "k_hop FP32->FP64", "k_hop FP16", "k_hop 48-bit"]
# storage bytes per real of the spinor fields and of the links each stencil variant reads (FP16 runs
# on the FP32 links, the 48-bit layout on the FP64 links)
DEGREE_BYTES_PER_REAL = {0: (8, 8), 1: (4, 4), 3: (2, 4), 4: (6, 8)}


def wilson_step_alg_bytes(c: cf.Config, real_bytes: int) -> float:
    """Algorithmic bytes of one Wilson leapfrog step of all modes: two half steps of 3S + L/M per
    site-mode (read the stencil input and the target, write the target; links once)."""
    S = 4 * c.nc * 2 * real_bytes
    Lb = 3 * (4 if c.nc == 2 else 18) * real_bytes
    return 2 * (3 * S + Lb / c.nmodes) * c.volume * c.nmodes


def bench_other_workloads(tol: float, relax: int = 0, cpu: bool = True, graph: bool = True):
    """The other three single-lattice workloads, timed the same way as the main one (inputs
    resident, autotuned stencil, HIP events around every stencil launch): the SU(2) Wilson
    leapfrog in FP64, the SU(2) overlap in the mixed form, the SU(3) 32^3 Wilson leapfrog with FP32
    spinor arithmetic (as that workload specifies; FP64 beside it).  Each is timed over its own run
    length (100, 50 and 200 steps) in one call replaying the captured step graph, as the workload
    runs; the stencil rate and its share come from a short eager call with HIP events around every
    stencil launch.  Their cost does not depend on the mode values, so only the first 64 modes are
    drawn and repeated to fill the mode count."""
    out = {}
    for name in ("su2_wilson_8", "su2_overlap_16", "su3_wilson_32"):
        c = cf.CONFIGS[name]
        steps = c.steps
        nm = min(c.nmodes, 64)
        links, e, psi = c.generate(nm)
        psi0 = psi[:4].copy()
        ctx = ovl.Context(c.nc, c.n, c.nmodes, r_wilson=c.r_wilson, m_fermion=c.m_fermion, m_dw=c.m_dw)
        ctx.set_links(links)
        ctx.set_efield(e)
        if c.nb:
            ctx.set_u1_flux(c.nb, c.q)
        for f in range(0, c.nmodes, nm):
            k = min(nm, c.nmodes - f)
            ctx.set_modes(ovl.PSI, psi[:k], first=f)
            ctx.set_modes(ovl.PSIMID, psi[:k], first=f)
        del psi
        ctx.autotune()
        if c.kind == "overlap":
            precs = [ovl.PREC_MIXED]
        elif name == "su3_wilson_32":
            precs = [ovl.PREC_FP32, ovl.PREC_FP64]
        else:
            precs = [ovl.PREC_FP64]
        pname = {ovl.PREC_MIXED: "mixed", ovl.PREC_FP64: "fp64", ovl.PREC_FP32: "fp32"}
        for prec in precs:
            # the run: hipGraph replay of the step, captured in an untimed warmup step (eager stream
            # launches with graph=False)
            so = step_opts(c, prec, graph, tol, relax=relax)
            ctx.fermion_update(so, 1)
            t0 = time.perf_counter()
            ctx.fermion_update(so, steps)
            wall = time.perf_counter() - t0
            # the stencil's rate: a short eager call with HIP events around every stencil launch
            ctx.set_profiling(True)
            ctx.fermion_update(step_opts(c, prec, False, tol, relax=relax), 2 if c.kind == "overlap" else 10)
            ctx.set_profiling(False)
            t = ctx.timing()
            var = int(np.argmax([t.var_ms[i] for i in range(5)]))
            gbps = t.var_bytes[var] / max(t.var_ms[var], 1e-9) / 1e6
            r = {"lattice": list(c.n), "nc": c.nc, "modes": c.nmodes, "kind": c.kind, "steps": steps,
                 "precision": pname[prec], "ms_per_step": round(1e3 * wall / steps, 3),
                 "value": c.volume * c.nmodes * steps / wall, "unit": "site*mode updates/s",
                 "kernel": VARIANT_NAMES[var], "kernel_GBps": round(gbps, 1),
                 "kernel_frac_of_8TBps": round(gbps / HBM_PEAK_GBS, 4),
                 "kernel_share_of_eager_step": round(t.var_ms[var] / max(t.total_ms, 1e-9), 3)}
            if c.kind == "wilson":
                # the whole call: the FP32 form converts Psi and PsiMid to FP32 when it starts and back
                # when it ends (in the timed call, amortised over its steps)
                ab = wilson_step_alg_bytes(c, 4 if prec == ovl.PREC_FP32 else 8)
                r["step_alg_bytes_per_site_mode"] = round(ab / (c.volume * c.nmodes), 2)
                r["step_alg_frac"] = round(ab / (1e-3 * r["ms_per_step"]) / 1e9 / HBM_PEAK_GBS, 4)
            key = name if prec == precs[0] else f"{name}_{pname[prec]}"
            out[key] = r
            log(f"{key}: {r['ms_per_step']:.2f} ms/step {r['value']:.4e} site*mode/s, "
                f"{VARIANT_NAMES[var]} {gbps:.0f} GB/s")
        ctx.close()
        if c.nc == 2 and cpu:
            out[name]["cpu_reference"] = reference_cpu(c, links, psi0, default_cpu_threads())
    return out


def reference_cpu(c: cf.Config, links, psi, threads: int):
    """The reference's own code on the host (oracle/_ref, built from the reference's SU(2) sources):
    WilsonEvoOperator::Compute (WilsonOperator.cpp:30) or OverlapU with its Zolotarev coefficients
    (OverlapOperator.cpp:25-64 without the unused random vector) on the interval the Chebyshev sign
    uses.  Modes are independent (the reference's MPI decomposition, ModePartition.cpp:12-25), so
    the sample of modes runs on `threads` host threads at once (the functions keep no state between
    calls); a leapfrog step applies the operator twice per mode (OverlapUpdate.cpp:9-58).  The
    overlap step also makes four Ritz solves (OverlapUpdate.cpp:61-97, 99-135), serial, on the whole
    lattice: each iteration one WilsonUU; the MaxEValue solve (sign -1) runs all 2000 iterations, its
    stopping test never firing (RitzSolver.cpp:250), the MinEValue one at least one.  Their cost is
    the measured time of 20 iterations scaled to 2 x 2000 per step (a lower bound of the four); the
    reference's solver prints its own note that the iteration limit was reached."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    from concurrent.futures import ThreadPoolExecutor
    if not orc.ref_available():
        return None
    L = orc.lat(c.nc, c.n, c.r_wilson, c.m_fermion, c.m_dw)
    ucc = orc.links_to_ucc(L, links, orc.u1_phases(c.n, c.nb, c.q) if c.nb else None)
    lam = orc.norm_bound(L)
    orc.ref_set(L)  # the reference's globals (lattice size, masses), set once before the threads start
    R, uc = orc.ref(), orc.cz(ucc)
    ins = [orc.cz(p) for p in psi]
    outs = [np.zeros_like(ins[0]) for _ in range(2 * threads)]  # distinct per concurrent call

    def apply(k):
        i = k % len(ins)
        o = outs[k % len(outs)]
        if c.kind == "overlap":
            R.ref_overlap_evo(orc.dp(uc), orc.dp(ins[i]), orc.dp(o), C.c_double((c.cheb_cut * lam) ** 2),
                              C.c_double(lam ** 2))
        else:
            R.ref_wilson_evo(orc.dp(uc), orc.dp(ins[i]), orc.dp(o))

    apply(0)  # first call outside the clock
    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while done == 0 or (time.perf_counter() - t0 < 6.0 and done < 64 * threads):
            list(ex.map(apply, range(done, done + threads)))
            done += threads
    wall = time.perf_counter() - t0
    apply_wall = wall / done  # wall seconds per mode apply with `threads` running at once
    out = {"kind": "reference", "cores": threads, "modes_timed": done, "unit": "site*mode updates/s",
           "function": "OverlapU (Zolotarev, 25 poles, multi-shift CG)" if c.kind == "overlap"
           else "WilsonEvoOperator::Compute",
           "wall_seconds_per_mode_apply": round(apply_wall, 6)}
    step = 2 * c.nmodes * apply_wall
    if c.kind == "overlap":
        g = cf.rng(7)
        v = (g.normal(size=psi[0].size) + 1j * g.normal(size=psi[0].size))
        v = v / np.linalg.norm(v)
        t1 = time.perf_counter()
        orc.ref_ritz(L, ucc, v.reshape(-1), delitr=1e-16, nitlz=20, sign=-1.0, nrestart=100)
        per_it = (time.perf_counter() - t1) / 20
        out["ritz_seconds_per_iteration"] = round(per_it, 6)
        out["ritz_seconds_per_step_lower_bound"] = round(2 * 2000 * per_it, 3)
        out["value_without_ritz"] = c.volume * c.nmodes / step
        step += 2 * 2000 * per_it
    out["seconds_per_step"] = round(step, 3)
    out["value"] = c.volume * c.nmodes / step
    return out


def bench_next_rows(c: cf.Config, links, e):
    """Wall time of the calls either side of the leapfrog on the bench lattice: Gauss-law
    restoration (100 sweeps, E relaxed on the fixed links), the configuration text files (save and
    load of links and E), and the overlap eigenbasis of the U(1) background of the same lattice and
    flux (FP64 Chebyshev sign of the workload's degree; 256 modes set)."""
    out = {}
    with ovl.Context(c.nc, c.n, 64) as ctx:
        ctx.set_links(links)
        ctx.set_efield(e)
        ctx.gauss_restore(tol=0.0, max_iter=2)
        t0 = time.perf_counter()
        it, vmax = ctx.gauss_restore(tol=0.0, max_iter=100)
        w = time.perf_counter() - t0
        out["gauss_restore"] = {"lattice": list(c.n), "nc": c.nc, "sweeps": it, "ms_per_sweep": round(1e3 * w / it, 4),
                                "max_violation_after": vmax}
        d = tempfile.mkdtemp(prefix="ovl_io_", dir="/tmp")
        try:
            uf, ef = os.path.join(d, "u.txt"), os.path.join(d, "e.txt")
            t0 = time.perf_counter()
            ctx.save_configuration(uf, ef)
            ws = time.perf_counter() - t0
            nbytes = os.path.getsize(uf) + os.path.getsize(ef)
            t0 = time.perf_counter()
            ctx.load_configuration(uf, ef)
            wl = time.perf_counter() - t0
        finally:
            shutil.rmtree(d, ignore_errors=True)
        out["configuration_files"] = {"bytes": nbytes, "save_s": round(ws, 3), "load_s": round(wl, 3),
                                      "save_MBps": round(nbytes / ws / 1e6, 1), "load_MBps": round(nbytes / wl / 1e6, 1)}
        # the single-mode drop-in with the reference's host arguments (WilsonEvoOperator::Compute:
        # u_cc and phi in, psi out, pageable host arrays): every call copies them over PCIe
        g = np.random.default_rng(11)
        shp = (3 * c.volume, c.nc, c.nc)
        ucc = (g.standard_normal(shp) + 1j * g.standard_normal(shp)) / c.nc
        phi = g.standard_normal(c.volume * 4 * c.nc) + 1j * g.standard_normal(c.volume * 4 * c.nc)
        ctx.wilson_evo_compute(ucc, phi)
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.wilson_evo_compute(ucc, phi)
        w = (time.perf_counter() - t0) / reps
        hb = ucc.nbytes + 2 * phi.nbytes
        out["host_buffer_wilson"] = {"lattice": list(c.n), "nc": c.nc, "modes_per_call": 1,
                                     "ms_per_call": round(1e3 * w, 3), "site_updates_per_s": c.volume / w,
                                     "host_bytes_per_call": hb, "host_GBps": round(hb / w / 1e9, 2)}
    with ovl.Context(c.nc, c.n, 256, m_fermion=c.m_fermion, m_dw=c.m_dw) as ctx:
        o = ovl.sign_opts(degree=c.cheb_n, cut=c.cheb_cut, precision=ovl.PREC_FP64)
        t0 = time.perf_counter()
        ev, herm = ctx.qed_eigenbasis(o, c.nb, c.q, c.dt)
        w = time.perf_counter() - t0
        out["qed_eigenbasis"] = {"lattice": list(c.n), "nc": c.nc, "flux_quanta": c.nb, "dimension": int(ev.size),
                                 "blocks": f"{c.n[2]} x {4 * c.n[0] * c.n[1]} + 4", "modes_set": 256,
                                 "chebyshev_degree": c.cheb_n, "seconds": round(w, 3), "hermiticity": herm,
                                 "spectrum": [float(ev.min()), float(ev.max())]}
    log("next rows", json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="su3_overlap_24", choices=sorted(cf.CONFIGS))
    ap.add_argument("--modes", type=int, default=0, help="override the mode count (test runs only)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--chunk", type=int, default=0, help="mode blocks per recurrence pass (0 = all)")
    ap.add_argument("--relax", type=int, default=2,
                    help="levels the accuracy guard may loosen the mixed ladder by (each x4 on tol) while its "
                         "check stays within tol/2; 0 = the ladder is only ever tightened")
    ap.add_argument("--mb", type=int, default=-1,
                    help="modes per SoA block (16/32/64; 0 = by the mode count; -1 = timed on this lattice)")
    ap.add_argument("--traversal", type=int, default=0, help="0 = XCD boxes, 1 = lexicographic")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="timed region only (layout / tile sweeps)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = OMP_NUM_THREADS, else the CPUs allowed")
    ap.add_argument("--cpu-modes", type=int, default=0, help="0 = two per thread")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble workload")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--tuning", default="", help="fixed stencil tuning instead of the autotune sweep, e.g. "
                    "'64:2/1/1/2,32:2/1/1/1,16:2/1/0/1' = precision:traversal/slabs/nt_store/waves_per_eu")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the other three single-lattice workloads")
    ap.add_argument("--no-next", action="store_true", help="skip the Gauss restore / files / eigenbasis timings")
    ap.add_argument("--next-only", action="store_true", help="only the Gauss restore / files / eigenbasis timings")
    ap.add_argument("--other-only", action="store_true", help="only the other three single-lattice workloads")
    ap.add_argument("--no-graph", action="store_true",
                    help="the ensemble's and the other workloads' runs as eager stream launches (for kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    if args.ensemble_only:
        r = bench_ensemble(cf.ENSEMBLES["su2_overlap_ensemble"], args.steps, args.warmup, args.tol, args.relax,
                           not args.no_graph)
        print(json.dumps(r))
        return
    if args.other_only:
        print(json.dumps({"other_workloads": bench_other_workloads(args.tol, args.relax, not args.no_cpu,
                                                                   not args.no_graph)}))
        return

    c = cf.CONFIGS[args.workload]
    if args.modes:
        c = cf.Config(**{**c.__dict__, "nmodes": args.modes})
    log("device", ovl.device_name())
    log("workload", c.name, "n", c.n, "modes", c.nmodes)
    t0 = time.perf_counter()
    links, e, psi = c.generate()
    log(f"inputs generated in {time.perf_counter() - t0:.1f}s")
    if args.next_only:
        del psi
        print(json.dumps({"workload": c.name, "next_rows": bench_next_rows(c, links, e)}))
        return
    # the calls either side of the leapfrog, first: the set-up of a run (its initial eigenbasis, the
    # Gauss restore, the files) comes before its steps, and the single-mode host-array drop-in is
    # timed in a process that has not yet run the other workloads (after them its pageable copies
    # ran at a third of the rate in one bench sequence, which no single workload reproduced)
    nxt = None
    if not args.no_next and not args.quick:
        nxt = bench_next_rows(c, links, e)
    ctx = ovl.Context(c.nc, c.n, c.nmodes, r_wilson=c.r_wilson, m_fermion=c.m_fermion, m_dw=c.m_dw,
                      mode_block=args.mb, traversal=args.traversal)
    ctx.set_links(links)
    ctx.set_efield(e)
    if c.nb:
        ctx.set_u1_flux(c.nb, c.q)
    ctx.set_modes(ovl.PSI, psi)
    ctx.set_modes(ovl.PSIMID, psi)
    # a plain copy of one FP64 field (k_lincomb<double>, known bytes): the calibration point of the
    # FETCH_SIZE / WRITE_SIZE counters in profiles/
    ctx.copy_field(ovl.PSI, ovl.SCRATCH)
    log(f"calibration copy: {c.volume * c.nmodes * 4 * c.nc * 16} bytes read and written")

    # --- accuracy of the mixed form against the plain FP64 form (one overlap apply, all modes)
    prec_mixed = ovl.PREC_MIXED
    so64 = step_opts(c, ovl.PREC_FP64, False).sign
    som = step_opts(c, prec_mixed, False, args.tol, args.chunk, args.relax).sign
    res = {}
    if args.quick:
        args.no_fp64 = args.no_cpu = args.no_ensemble = args.no_traffic = args.no_other = args.no_next = True
    if c.kind == "overlap" and not args.quick:
        # every mode of one overlap apply, mixed against FP64 (the FP64 result held on the host)
        ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, so64)
        ref = ctx.get_modes(ovl.SCRATCH)
        ctx.apply_overlap(ovl.PSI, ovl.SCRATCH, som)
        tm = ctx.timing()
        worst = 0.0
        for f in range(0, c.nmodes, 32):
            mix = ctx.get_modes(ovl.SCRATCH, f, min(32, c.nmodes - f))
            worst = max(worst, max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mix, ref[f:f + 32])))
        res["rel_err_mixed_vs_fp64"] = float(worst)
        res["rel_err_modes"] = c.nmodes
        res["fp64_steps"] = int(tm.fp64_steps)
        res["fp48_steps"] = int(tm.fp48_steps)
        res["fp32_steps"] = int(tm.fp32_steps)
        res["fp16_steps"] = int(tm.fp16_steps)
        del ref, mix
        log(f"mixed vs fp64 rel err (all {c.nmodes} modes) {res['rel_err_mixed_vs_fp64']:.3e}  steps fp64 {tm.fp64_steps} "
            f"48-bit {tm.fp48_steps} fp32 {tm.fp32_steps} fp16 {tm.fp16_steps}")

    # --- stencil autotune for this shape (or a fixed tuning)
    if args.tuning:
        for item in args.tuning.split(","):
            p, v = item.split(":")
            prec = {"64": 0, "32": 1, "16": 2, "48": 3}[p.strip()]
            ctx.set_tuning(prec, *[int(x) for x in v.split("/")])
        tune = "fixed " + args.tuning
    else:
        tune = ctx.autotune()
    log("autotune", tune)
    tuning = tuning_string(ctx)
    res["autotune"] = tune.split(";")[-1].strip()  # the choice; every candidate's time is in the log
    res["tuning"] = tuning
    mbm = re.search(r"mode block:((?: \d+ [0-9.]+ us)+), chosen (\d+)", tune)
    if mbm:  # the timed mode block (--mb -1): each size's Clenshaw step time, and the choice
        res["mode_block_timing"] = mbm.group(0)
        args.mb = int(mbm.group(2))  # the counter passes below run the chosen block size

    # --- warmup + timed region: mixed form, per-launch HIP events on the library stream
    so = step_opts(c, prec_mixed, False, args.tol, args.chunk, args.relax)
    for _ in range(args.warmup):
        ctx.fermion_update(so, 1)
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    ctx.fermion_update(so, args.steps)
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    t = ctx.timing()
    ms_per_step = 1e3 * wall / args.steps
    value = c.volume * c.nmodes * args.steps / wall
    log(f"mixed: {ms_per_step:.1f} ms/step  {value:.4e} site*mode/s  (events {t.total_ms / args.steps:.1f} ms/step)")

    # dominant stencil variant of the timed region
    var = int(np.argmax([t.var_ms[i] for i in range(5)]))
    names = VARIANT_NAMES
    avg_ms = t.var_ms[var] / max(1, t.var_launches[var])
    unfused_per_launch = t.var_bytes[var] / max(1, t.var_launches[var])
    bytes_per_launch = unfused_per_launch
    deg = None
    if c.kind == "overlap" and var in DEGREE_BYTES_PER_REAL:
        # the fused Clenshaw degree of the survey's roofline, 4S + 2L/M bytes per site-mode (read
        # b_{k+1}, b_{k+2}, x, write b_k; links once per launch over all modes), is two launches here
        # (w = h b_{k+1}, then the epilogue): per launch, half of it
        rb, lb = DEGREE_BYTES_PER_REAL[var]
        S = 4 * c.nc * 2 * rb
        Lb = 3 * (4 if c.nc == 2 else 18) * lb
        per_site_mode = 4 * S + 2 * Lb / c.nmodes
        bytes_per_launch = per_site_mode * c.volume * c.nmodes / 2
        deg = {"bytes_per_site_mode": per_site_mode, "S": S, "L": Lb, "launches_per_degree": 2,
               "us_per_degree": round(2 * avg_ms * 1e3, 2), "bytes_per_degree": 2 * bytes_per_launch}
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    alg_step_bytes = None
    if c.kind == "overlap":
        S = 4 * c.nc * 2 * 4
        Lb = 3 * (4 if c.nc == 2 else 18) * 4
        M = c.nmodes
        n = c.cheb_n
        per_half = n * (4 * S + 2 * Lb / M) + (2 * S + Lb / M) + 3 * S
        alg_step_bytes = 2 * per_half * c.volume * M
    roofline = {"bound": "hbm", "kernel": names[var], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(avg_ms * 1e3, 2), "alg_bytes_per_launch": bytes_per_launch,
                "launches": int(t.var_launches[var]),
                "kernel_share_of_step": round(t.var_ms[var] / max(t.total_ms, 1e-9), 3),
                "algorithmic_bytes": ("fused Clenshaw degree 4S + 2L/M per site-mode, two launches per degree"
                                      if deg else "stencil input once, pointwise operands and output once"),
                "degree": deg,
                # the same launches counted as they run (unfused h_W^2: stencil input, each pointwise
                # operand and the output once, 7S per degree)
                "unfused_alg_bytes_per_launch": unfused_per_launch,
                "unfused_achieved": round(unfused_per_launch / (avg_ms * 1e-3) / 1e9, 1),
                "unfused_frac": round(unfused_per_launch / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if alg_step_bytes:
        roofline["step_alg_bytes"] = alg_step_bytes
        roofline["step_alg_frac"] = round(alg_step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    roofline["variants"] = {
        names[i]: {"launches": int(t.var_launches[i]), "ms": round(t.var_ms[i], 3),
                   "GBps": round(t.var_bytes[i] / max(t.var_ms[i], 1e-9) / 1e6, 1)}
        for i in range(5) if t.var_launches[i]}
    log("variants", json.dumps(roofline["variants"]))

    # --- graph-captured replay of the same step (capture and instantiation inside the timed
    # region): the headline value; the eager evented run above gives the per-kernel roofline
    graph_ms = None
    eager = {"ms_per_step": ms_per_step, "value": value}
    timing = "eager stream, HIP events around every stencil launch"
    if not args.quick:
        sog = step_opts(c, prec_mixed, True, args.tol, args.chunk, args.relax)
        ctx.fermion_update(sog, 1)  # untimed warmup step: captures the step graph, which the context keeps
        t0 = time.perf_counter()
        ctx.fermion_update(sog, args.steps)
        wg = time.perf_counter() - t0
        graph_ms = 1e3 * wg / args.steps
        captures = int(ctx.timing().graph_captures)
        log(f"mixed (hipGraph): {graph_ms:.1f} ms/step, {captures} captures in the timed call")
        if wg < wall:
            ms_per_step = graph_ms
            value = c.volume * c.nmodes * args.steps / wg
            timing = ("hipGraph replay of the step graph captured in an untimed warmup step (the timed call "
                      f"bounds the mode norms, checks the step's key and replays; captures in it: {captures}); "
                      "the roofline is from the eager run with HIP events around every stencil launch")
            if alg_step_bytes:
                roofline["step_alg_frac"] = round(alg_step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)

    # --- cost of one accuracy-guard check of the mixed ladder (due every 100 steps and whenever new
    # links are set): the links are set again to their current values, which makes a check due, and
    # a one-step graph call with the check is timed against the next one without it
    guard = None
    if c.kind == "overlap" and not args.quick:
        sog1 = step_opts(c, prec_mixed, True, args.tol, args.chunk, args.relax)
        ctx.set_links(ctx.get_links())
        t0 = time.perf_counter()
        ctx.fermion_update(sog1, 1)
        w_check = time.perf_counter() - t0
        tg = ctx.timing()
        t0 = time.perf_counter()
        ctx.fermion_update(sog1, 1)
        w_plain = time.perf_counter() - t0
        check_ms = 1e3 * (w_check - w_plain)
        every = int(sog1.sign.guard)
        guard = {"checks_in_call": int(tg.guard_checks), "level_after": int(tg.guard_level),
                 "max_rel_err_of_check": float(tg.guard_err), "check_ms": round(check_ms, 2),
                 "every_steps": every, "amortised_ms_per_step": round(check_ms / max(every, 1), 3),
                 "ms_per_step_with_guard": round(ms_per_step + check_ms / max(every, 1), 2)}
        log(f"guard: {json.dumps(guard)}")

    # --- the workload's full run from its initial state (config steps, mixed form, graph replay):
    # dN_CS/dt after every step, integrated by the trapezoid rule into the Chern-Simons number
    # history N_CS(t) - N_CS(0), and the axial charge every 5 steps (currents from the leapfrog's own
    # overlap apply)
    observables = None
    if c.kind == "overlap" and not args.quick:
        ctx.set_links(links)
        ctx.set_modes(ovl.PSI, psi)
        ctx.set_modes(ovl.PSIMID, psi)
        so_obs = step_opts(c, prec_mixed, True, args.tol, args.chunk, args.relax)
        rate0 = ctx.ncs_dot()
        t0 = time.perf_counter()
        q5, rate, _, _ = ctx.fermion_update_observed(so_obs, c.steps, every=1)
        secs = time.perf_counter() - t0
        rates = np.concatenate([[rate0], rate])
        ncs = np.concatenate([[0.0], np.cumsum(0.5 * c.dt * (rates[1:] + rates[:-1]))])
        every = 5
        observables = {"steps": c.steps, "dt": c.dt, "every": every, "t": [round(k * c.dt, 6) for k in range(0, c.steps + 1, every)],
                       "ncs": [float(v) for v in ncs[::every]], "ncs_dot": [float(v) for v in rates[::every]],
                       "axial_charge": [float(v) for v in q5[every - 1::every]],
                       "axial_charge_t": [round(k * c.dt, 6) for k in range(every, c.steps + 1, every)],
                       "seconds": round(secs, 3), "ms_per_step": round(1e3 * secs / c.steps, 2)}
        log(f"observed run ({c.steps} steps, {observables['ms_per_step']} ms/step): N_CS(t) - N_CS(0) "
            f"{observables['ncs'][-1]:.4e}, axial charge {observables['axial_charge'][-1]:.4e}")

    # --- plain FP64 form
    fp64 = None
    if not args.no_fp64:
        so64s = step_opts(c, ovl.PREC_FP64, True)  # as the headline: graph replay
        ctx.fermion_update(so64s, 1)
        k64 = args.steps
        t0 = time.perf_counter()
        ctx.fermion_update(so64s, k64)
        w64 = time.perf_counter() - t0
        fp64 = {"value": c.volume * c.nmodes * k64 / w64, "ms_per_step": 1e3 * w64 / k64, "steps": k64}
        log(f"fp64: {fp64['ms_per_step']:.1f} ms/step")

    # --- sign-function residual of the mixed form on the current state
    if c.kind == "overlap" and not args.quick:
        res["sign_residual_mixed"] = ctx.sign_residual(ovl.PSI, som)
        log(f"sign residual (mixed) {res['sign_residual_mixed']:.3e}")
    ctx.close()
    del psi

    # HBM traffic of the dominant stencil variant, from two counter passes (child processes)
    if not args.no_traffic:
        prefix = ["k_hop<%d, double, double" % c.nc, "k_hop<%d, float, float" % c.nc,
                  "k_hop<%d, float, double" % c.nc, "k_hop<%d, _Float16, _Float16" % c.nc,
                  "k_hop<%d, f48, f48" % c.nc][var]
        tr = measure_traffic(args, c, prefix, int(round(t.var_launches[var] / args.steps)), tuning)
        if tr:
            traffic = tr["FETCH_SIZE"]["bytes_per_launch"] + tr["WRITE_SIZE"]["bytes_per_launch"]
            roofline["traffic"] = traffic
            roofline["traffic_detail"] = {
                "read_bytes_per_launch": tr["FETCH_SIZE"]["bytes_per_launch"],
                "write_bytes_per_launch": tr["WRITE_SIZE"]["bytes_per_launch"],
                "fetch_calibration": round(tr["FETCH_SIZE"]["factor"], 4),
                "write_calibration": round(tr["WRITE_SIZE"]["factor"], 4),
                "kernel": tr["FETCH_SIZE"]["kernel"],
                "traffic_over_algorithmic": round(traffic / bytes_per_launch, 3),
                "traffic_over_unfused": round(traffic / unfused_per_launch, 3),
                "traffic_GBps": round(traffic / (avg_ms * 1e-3) / 1e9, 1)}
            log("traffic", json.dumps(roofline["traffic_detail"]))

    ens = None
    if not args.no_ensemble:
        ens = bench_ensemble(cf.ENSEMBLES["su2_overlap_ensemble"], 10, 1, args.tol, args.relax, not args.no_graph)

    other = None
    if not args.no_other:
        other = bench_other_workloads(args.tol, args.relax, not args.no_cpu)


    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or default_cpu_threads()
        nm = args.cpu_modes or 2 * threads
        links, e, psi = c.generate(nmodes=nm)
        v, secs = cpu_baseline(c, links, e, psi, threads, nm)
        cpu = {"value": v, "unit": "site*mode updates/s", "cores": threads, "kind": "port",
               "sample": f"{nm} modes x 1 full step of {c.name} ({secs:.1f} s; two modes per thread), "
                         f"oracle/liboracle.so FP64 C restatement (the reference has no SU(3) path), "
                         f"OpenMP over modes, {threads} threads",
               "host": host_description()}
        log(f"cpu baseline {v:.4e} site*mode/s in {secs:.1f}s on {threads} threads")

    out = {"metric": METRIC, "value": value, "unit": "site*mode updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp64+48-bit+fp32+fp16-storage (mixed)", "data": "synthetic",
           "config": {"workload": c.name, "nc": c.nc, "lattice": list(c.n), "modes": c.nmodes,
                      "chebyshev_degree": c.cheb_n, "cut": c.cheb_cut, "rho": c.m_dw, "flux_quanta": c.nb,
                      "dt": c.dt, "tol": args.tol, "guard_relax_levels": args.relax, "mode_block": args.mb, "mode_chunk": args.chunk,
                      "traversal": args.traversal, "parallelism": "single GPU"},
           "roofline": roofline, "cpu_baseline": cpu, "fp64_gpu": fp64,
           "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
           # a leapfrog step applies the overlap operator twice to every mode (each a full-lattice apply)
           "overlap_applies_per_s": (2.0 * value / c.volume) if c.kind == "overlap" else None,
           "fp64_overlap_applies_per_s": (2.0 * fp64["value"] / c.volume) if (fp64 and c.kind == "overlap") else None,
           "graph_ms_per_step": graph_ms, "eager": eager, "timing": timing, "accuracy": res, "guard": guard,
           "observables": observables, "ensemble": ens,
           "other_workloads": other, "next_rows": nxt}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
