#!/usr/bin/env python3
"""Benchmark of the explicit unstructured finite-volume Euler residual on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c4|c2|c3|c3mixed|c2rk|c2out]

c5 (default): 1,024 wedge instances of uneven size, Mach number and angle in one
batch, FP32-mixed, with the same batch in FP64 beside it: both throughputs, their
ratio and the per-instance density error of mixed against FP64 after the quoted
2,000 steps; the C2 (4M-triangle Riemann, FP64) and C3 (16M-triangle ramp, FP64 and
mixed) single meshes and the C4 batch (256 wedges, FP64) as secondary lines.  c4 is the
256-wedge FP64 batch alone, c2/c3/c3mixed the single meshes alone, c2rk the SSP-RK
integrators' cost, c2out the output path (files written while stepping).

A step is one residual evaluation plus the forward-Euler update (src/Base/Euler.cpp:85-97),
with the CFL time step.  The state is resident in HBM before the timed region starts;
exactly --steps steps are timed after --warmup untimed ones.  Prints ONE JSON line with
cell-updates/s (interior cells x steps / s), the HBM roofline of the dominant kernel
and the reference's CPU solver timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "cell-updates/s per residual eval + % HBM peak; FP32-mixed vs FP64 at stated err"


def alg_bytes(Nc: int, Nf: int, mixed: bool):
    """Algorithmic HBM bytes per step of the minimal two-pass dataflow that the
    method's dependencies force (cells incl. ghosts, every array touched once per
    pass, face geometry read once per pass):
    FP64:  pass A 104 Nc + 64 Nf (U 32 + 1/V 8 read, limited gradient 64 written;
                                  face record 64)
           pass B 136 Nc + 64 Nf (U 32 + limited gradient 64 + 1/V 8 read, U 32
                                  written; face record 64)
    mixed: pass A  68 Nc + 36 Nf;  pass B 100 Nc + 36 Nf.
    The kernels' own intermediate (face states instead of the limited gradient)
    is a design choice; the figure stays the method's minimum."""
    if mixed:
        a, b = 68 * Nc + 36 * Nf, 100 * Nc + 36 * Nf
    else:
        a, b = 104 * Nc + 64 * Nf, 136 * Nc + 64 * Nf
    return a, b, a + b


PMC_SUMMARY = {"c2": "pmc_riemann_4m_fp64.md", "c3": "pmc_ramp_16m_fp64.md", "c3mixed": "pmc_ramp_16m_mixed.md",
               "c4": "pmc_wedges_256_fp64.md", "c5": "pmc_wedges_1024_mixed.md"}


def read_traffic(workload: str):
    """HBM bytes per launch of each pass from the committed rocprofv3 --pmc summary
    of the workload (profiles/pmc_*.md, lines 'traffic_bytes <kernel> <bytes>':
    FETCH_SIZE and WRITE_SIZE scaled by what they read for a copy of known size)."""
    path = os.path.join(ROOT, "profiles", PMC_SUMMARY.get(workload, "none"))
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        m = re.match(r"\s*traffic_bytes\s+(\S+)\s+([0-9.eE+]+)", line)
        if m:
            out[m.group(1)] = float(m.group(2))
    return out or None


def roofline_record(kernel: str, alg_bytes: float, ms: float, traffic):
    """The HBM roofline of one kernel: achieved = algorithmic bytes per launch (the
    minimal two-pass dataflow, see alg_bytes) / the live mean launch time.  The
    single-kernel step keeps its face states on chip and moves fewer bytes than
    that minimum, so its fraction can exceed 1; traffic (measured HBM bytes per
    launch from the committed rocprofv3 summary) gives the bytes it really moves."""
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    rec = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "alg_bytes": alg_bytes, "traffic": traffic,
           "timing": "HIP event pairs around eager launches right after the timed region"}
    if traffic:
        rec["traffic_gbs"] = traffic / (ms * 1e-3) / 1e9
        rec["traffic_frac"] = rec["traffic_gbs"] / HBM_PEAK_GBS
    return rec


def cpu_baseline(case, ctx, dt: float, target_s: float):
    """The reference's own solver (oracle/_ref, built from its sources) on one core,
    on the same mesh in the GPU's interior cell order, for a bounded number of
    steps of the dt the GPU run ended with: the VEC layout (AVX2 build) and the Base
    layout, the faster of the two as the baseline (they are within a few percent
    of each other; SoA and AoSoA do not give the reference's results)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as ol
    from cfd_optimisations_amd.formats import renumber_cells, write_hmsh
    hm = case.hmesh()
    perm = ctx.permutation()
    rm = renumber_cells(hm, perm)
    nint = hm.Nc - (hm.Nf - hm.F_BC)
    U0 = np.zeros((4, hm.Nc))
    U0[:, perm] = ctx.get_state()
    bcmap = [0] + [int(b) for b in case.bctypes]

    def timed(run):
        t0 = time.perf_counter()
        U = run(U0, 1)
        one = time.perf_counter() - t0
        n = max(1, min(200, int(target_s / 2 / max(one, 1e-6))))
        t0 = time.perf_counter()
        run(U, n)
        return n, time.perf_counter() - t0

    out = {}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "bench.hmsh")
        write_hmsh(path, rm)
        if ol.ref_available(ol.REF_VEC_SO):
            vc = ol.VecRefCase(path, case.gamma, bcmap, case.qbc)
            out["vec"] = timed(lambda U, n: vc.run(U, n, dt))
            vc.close()
        if ol.ref_available(ol.REF_SO):
            rc = ol.RefCase(path, case.gamma, bcmap, case.qbc, ol.REF_SO)
            rc.zero_bc_iv()
            out["base"] = timed(lambda U, n: rc.run(U, n, dt))
            rc.close()
        if not out:
            od = ol.OracleDomain(rm, case.gamma, case.bctypes, case.qbc)
            od.zero_bc_iv()
            out["port"] = timed(lambda U, n: od.run(U, n, dt=dt)[0])
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    rate = {k: nint * n / sec for k, (n, sec) in out.items()}
    best = max(rate, key=rate.get)  # the faster of the reference layouts measured
    n, sec = out[best]
    label = {"vec": "reference VEC layout (AVX2 build, oracle/_ref)",
             "base": "reference Base layout (oracle/_ref)", "port": "oracle port"}[best]
    rec = {"value": rate[best], "unit": "cell-updates/s", "cores": 1, "kind": "port" if best == "port" else "reference",
           "sample": f"{label}, the faster of the reference's valid CPU layouts timed here, same mesh in the "
                     f"GPU's cell order, {n} steps of fixed dt after 1 warm-up step, 1 thread on {cpu}",
           "seconds": round(sec, 3), "layouts": rate}
    return rec


def build_case(name: str):
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    W = C.WORKLOADS
    if name == "c2":
        w = W["c2_riemann_fp64"]
        return C.riemann_case(w["nx"], w["ny"], seed=0), w, L.FP64, \
            f"c2: four-quadrant Riemann problem on [0,1]^2, {w['nx']}x{w['ny']} quads -> {2 * w['nx'] * w['ny']:,} " \
            f"triangles, nodes jittered +-20% (seed 0), cells shuffled then renumbered on the device, FP64, " \
            f"CFL {w['cfl']}"
    if name in ("c3", "c3mixed"):
        w = W["c3_ramp_16m"]
        prec = L.MIXED if name == "c3mixed" else L.FP64
        return C.ramp_case(w["nx"], w["ny"], w["mach"], w["angle"], seed=0), w, prec, \
            f"{name}: Mach {w['mach']} flow over a {w['angle']} deg ramp, {w['nx']}x{w['ny']} quads -> " \
            f"{2 * w['nx'] * w['ny']:,} triangles, {'FP32-mixed' if prec else 'FP64'}, CFL 0.5"
    raise SystemExit(f"unknown workload {name}")


def run_single(name: str, steps: int, warmup: int, cpu_s: float, graph: bool, secondary: bool, tune: bool,
               order: str = "sfc", step_path: str = "default"):
    import numpy as np
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    case, w, prec, desc = build_case(name)
    steps = steps or w["steps"]
    cfl = w["cfl"]
    flags = (L.FLAG_GRAPH if graph else 0) | (L.FLAG_ORDER_SFC if order == "sfc" else 0)
    t0 = time.perf_counter()
    ctx = C.make_context(case, prec, flags)
    setup_s = time.perf_counter() - t0
    inf = ctx.info()
    if tune:
        ctx.autotune(steps=10, cfl=cfl, cache=os.path.join(tempfile.gettempdir(), "hcfd_tune.txt"))
    if step_path != "default":
        ctx.set_schedule(fused=1 if step_path == "single" else 0)
    # warm-up steps, then the state is put back: the timed steps are steps
    # 1..steps of the workload's run from its initial state (the Riemann problem
    # is quoted on 500 steps and leaves the physical range some steps later)
    U_init = ctx.get_state()
    ctx.step(warmup, cfl=cfl)
    ctx.set_state(U_init)
    ctx.fault()  # clears the flag
    # timed region: exactly `steps` steps (hipGraph replays with --graph), the
    # stream drained on both sides
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps, cfl=cfl)
    wall = time.perf_counter() - t0
    fault = ctx.fault()  # any non-physical state among the timed steps
    # per-kernel durations: HIP event pairs around eager launches right after
    # it, again from the initial state
    ctx.set_state(U_init)
    tm = ctx.time_steps(min(steps, 100), cfl=cfl)
    ms_step = 1e3 * wall / steps
    value = inf.Nint * steps / wall
    a_b, b_b, step_b = alg_bytes(inf.Nc, inf.Nf, prec == L.MIXED)
    fused = bool(ctx.schedule()["fused"])
    if fused:   # one launch performs the whole step
        kern = {"fused_step": (tm["pass_a_ms"], step_b)}
    else:
        kern = {"pass_a": (tm["pass_a_ms"], a_b), "pass_b": (tm["pass_b_ms"], b_b)}
    dom = max(kern, key=lambda k: kern[k][0])
    dms, dbytes = kern[dom]
    achieved = dbytes / (dms * 1e-3) / 1e9
    traffic = read_traffic(name)
    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == L.MIXED else "f64", "data": "synthetic",
        "config": {"workload": desc, "Nc": inf.Nc, "Nint": inf.Nint, "Nf": inf.Nf, "steps": steps, "cfl": cfl,
                   "schedule": ctx.schedule(), "graph": graph, "setup_s": round(setup_s, 2), "cell_order": order,
                   "bandwidth": {"what": "max |CL - CR| over interior faces", "caller_order": inf.bandwidth_in,
                                 f"{'hilbert' if order == 'sfc' else 'rcm'}_order": inf.bandwidth_out}},
        "hbm_pct_step": 100.0 * step_b / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "alg_bytes_per_step": step_b,
        "kernel_ms": ({"fused_step": tm["pass_a_ms"]} if fused else
                      {"pass_a": tm["pass_a_ms"], "pass_b": tm["pass_b_ms"]}) |
                     {"dt_reduce": tm["dt_reduce_ms"], "event_wall_per_step": tm["wall_ms"] / min(steps, 100)},
        "roofline": roofline_record(dom, dbytes, dms, (traffic or {}).get(dom)),
        "fault": fault,
        "copy_ceiling_gbs": L.copy_bandwidth(1 << 30, 20),
    }
    line["roofline"]["frac_of_copy_ceiling"] = achieved / line["copy_ceiling_gbs"]
    if cpu_s > 0:
        dt = ctx.last_dt()
        line["cpu_baseline"] = cpu_baseline(case, ctx, dt, cpu_s)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    ctx.close()
    if secondary and name == "c2":
        line["secondary"] = [r for r in secondary_lines() if not r["workload"].startswith("c2")]
    return line


def secondary_lines():
    """The other workloads beside the headline: C2 (4M-triangle Riemann, FP64, its
    quoted 500 steps from the initial state) and C3 (16M-triangle ramp, 1000 steps,
    FP64 and FP32-mixed), each timed over its quoted steps with its own roofline
    (event-timed dominant kernel, traffic from the committed rocprofv3 --pmc
    summary), the C3 mixed run's error per conserved variable against the C3
    FP64 run after the same steps, and the C4 FP64 batch (batch_secondary)."""
    import numpy as np
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    out = []
    states = {}
    for name in ("c2", "c3", "c3mixed"):
        case, w, prec, desc = build_case(name)
        steps, cfl = w["steps"], w["cfl"]
        ctx = C.make_context(case, prec, L.FLAG_ORDER_SFC)
        inf = ctx.info()
        U_init = ctx.get_state()
        ctx.step(3, cfl=cfl)            # warm-up, then the initial state again
        ctx.set_state(U_init)
        ctx.fault()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps, cfl=cfl)
        wall = time.perf_counter() - t0
        fault = ctx.fault()
        if name != "c2":
            states[prec] = ctx.get_state()
        ctx.set_state(U_init)
        nev = min(steps, 100)
        tm = ctx.time_steps(nev, cfl=cfl)
        a_b, b_b, step_b = alg_bytes(inf.Nc, inf.Nf, prec == L.MIXED)
        fused = bool(ctx.schedule()["fused"])
        kern = {"fused_step": (tm["pass_a_ms"], step_b)} if fused else \
            {"pass_a": (tm["pass_a_ms"], a_b), "pass_b": (tm["pass_b_ms"], b_b)}
        dom = max(kern, key=lambda k: kern[k][0])
        achieved = kern[dom][1] / (kern[dom][0] * 1e-3) / 1e9
        ms = 1e3 * wall / steps
        out.append({"workload": desc + f", {steps} steps from the initial state",
                    "value": inf.Nint * steps / wall, "unit": "cell-updates/s", "ms_per_step": ms,
                    "dtype": "f32+f64" if prec == L.MIXED else "f64", "schedule": ctx.schedule(),
                    "kernel_ms": {dom: kern[dom][0], "dt_reduce": tm["dt_reduce_ms"]},
                    "hbm_pct_step": 100.0 * step_b / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
                    "roofline": roofline_record(dom, kern[dom][1], kern[dom][0], (read_traffic(name) or {}).get(dom)),
                    "fault": fault})
        log(f"{name}: {ms:.4f} ms per step, {dom} {kern[dom][0]:.4f} ms ({achieved / HBM_PEAK_GBS:.3f} of peak)")
        ctx.close()
    Ua, Um = states[L.FP64], states[L.MIXED]
    names = ("rho", "rho_u", "rho_v", "rho_E")
    out[2]["error_vs_fp64"] = {
        names[v]: {"max_abs": float(np.max(np.abs(Um[v] - Ua[v]))),
                   "rel_l2": float(np.linalg.norm(Um[v] - Ua[v]) / np.linalg.norm(Ua[v]))} for v in range(4)}
    out[2]["speedup_vs_fp64"] = out[2]["value"] / out[1]["value"]
    out.append(batch_secondary("c4", 200, 20))
    return out


def batch_secondary(name: str, steps: int, warmup: int):
    """A batch as a secondary line: the 256-wedge FP64 batch (C4), autotuned,
    hipGraph replay, timed over `steps` steps after `warmup` from the initial
    state, with its roofline (per-instance agreement with the CPU solver is in
    the GPU tests)."""
    b = run_batch(name, steps, warmup, True, False, 0, False, 0.0, False)
    return {"workload": b["config"]["workload"].replace(
                "; timed steps follow 0 untimed steps from the initial state",
                f", {steps} steps after {warmup} from the initial state"),
            "value": b["value"], "unit": b["unit"], "ms_per_step": b["ms_per_step"], "dtype": b["dtype"],
            "instances": b["config"]["instances"], "Nint": b["config"]["Nint"],
            "schedule": b["config"]["schedule"], "kernel_ms": b["kernel_ms"], "hbm_pct_step": b["hbm_pct_step"],
            "roofline": b["roofline"], "faulted_instances": b["faulted_instances"]}


def run_integrators(steps: int):
    """C2's mesh and state stepped with forward Euler, SSP-RK2 and SSP-RK3 (fixed dt,
    the CFL step of the initial state / stages so the simulated time per stage
    matches): time per step and per stage, and cell-updates/s counted per stage."""
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    w = C.WORKLOADS["c2_riemann_fp64"]
    case = C.riemann_case(w["nx"], w["ny"], seed=0)
    ctx = C.make_context(case, L.FP64, L.FLAG_ORDER_SFC)
    inf = ctx.info()
    U0 = ctx.get_state()
    ctx.step(1, cfl=w["cfl"])
    dt = ctx.last_dt()
    rows = []
    for stages in (1, 2, 3):
        ctx.set_state(U0)
        ctx.set_integrator(stages)
        ctx.step(3, dt=dt)
        tm = ctx.time_steps(steps, dt=dt)
        ms = tm["wall_ms"] / steps
        rows.append({"stages": stages, "ms_per_step": ms, "ms_per_stage": ms / stages,
                     "stage_updates_per_s": inf.Nint * stages / (ms * 1e-3), "fault": ctx.fault()})
        log(f"stages {stages}: {ms:.4f} ms per step")
    sched = ctx.schedule()
    ctx.close()
    return {"metric": METRIC + " (integrator cost)", "value": rows[2]["stage_updates_per_s"],
            "unit": "cell-updates/s (per stage, SSP-RK3)", "n_gpus": 1, "steps": steps, "warmup": 3,
            "ms_per_step": rows[2]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"c2 mesh ({w['nx']}x{w['ny']} quads, 4M triangles), fixed dt, FP64, "
                                   f"forward Euler / SSP-RK2 / SSP-RK3", "schedule": sched, "dt": dt},
            "integrators": rows}


def run_output_overlap(steps: int, every: int):
    """C2 with primitives written to an .hsln file every `every` steps: stepping
    alone, the reference's pattern (stop, compute primitives, copy, write), and
    snapshots taken on the device and copied on a second stream while stepping
    continues, the file written by the host meanwhile."""
    import numpy as np
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    w = C.WORKLOADS["c2_riemann_fp64"]
    case = C.riemann_case(w["nx"], w["ny"], seed=0)
    ctx = C.make_context(case, L.FP64, L.FLAG_ORDER_SFC)
    inf = ctx.info()
    cfl = w["cfl"]
    out_dir = tempfile.mkdtemp(prefix="hcfd_out_")
    U0 = ctx.get_state()
    nout = steps // every
    res = {}
    ctx.step(5, cfl=cfl)
    for mode in ("none", "sync", "overlap"):
        ctx.set_state(U0)
        ctx.sync()
        t0 = time.perf_counter()
        if mode == "none":
            ctx.step(steps, cfl=cfl)
        elif mode == "sync":
            for k in range(nout):
                ctx.step(every, cfl=cfl)
                L.write_hsln(os.path.join(out_dir, f"sync_{k}.hsln"), ctx.get_prims())
        else:
            pending = None
            for k in range(nout):
                ctx.step_async(every, cfl=cfl)
                slot = k & 1
                ctx.output_begin(slot)
                if pending is not None:
                    Q, _ = ctx.output_wait(pending)
                    L.write_hsln(os.path.join(out_dir, f"overlap_{k - 1}.hsln"), Q)
                pending = slot
            Q, _ = ctx.output_wait(pending)
            L.write_hsln(os.path.join(out_dir, f"overlap_{nout - 1}.hsln"), Q)
            ctx.sync()
        res[mode] = time.perf_counter() - t0
        log(f"output mode {mode}: {res[mode]:.3f} s")
    same = all(np.array_equal(L.read_hsln(os.path.join(out_dir, f"sync_{k}.hsln")),
                              L.read_hsln(os.path.join(out_dir, f"overlap_{k}.hsln"))) for k in range(nout))
    for f in os.listdir(out_dir):
        os.remove(os.path.join(out_dir, f))
    os.rmdir(out_dir)
    ctx.close()
    return {"metric": METRIC + " (output path)", "value": inf.Nint * steps / res["overlap"],
            "unit": "cell-updates/s (with output)", "n_gpus": 1, "steps": steps, "warmup": 5,
            "ms_per_step": 1e3 * res["overlap"] / steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"c2 ({w['nx']}x{w['ny']} quads, 4M triangles), FP64, CFL {cfl}, "
                                   f"{steps} steps, .hsln of all cells every {every} steps ({nout} files of "
                                   f"{32 * inf.Nc / 1e6:.0f} MB)"},
            "wall_s": res, "output_overhead_sync": res["sync"] / res["none"] - 1.0,
            "output_overhead_overlap": res["overlap"] / res["none"] - 1.0, "files_identical": same}


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def batch_cpu_baseline(cases, specs, order, dts, target_s: float, threads: int):
    """The reference's own VEC layout (oracle/_ref, AVX2 build), one instance per
    thread (the ctypes calls release the interpreter lock, so the threads run on
    separate cores): the `threads` instances nearest the batch's median size,
    each stepped with the fixed dt its GPU run ended with for about target_s
    seconds after one warm-up step.  value = interior cells x steps summed over
    the instances / wall time of the parallel section."""
    import concurrent.futures as cf
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from cfd_optimisations_amd.formats import write_hmsh
    if not ol.ref_available(ol.REF_VEC_SO):
        return None
    sizes = np.array([c.n_interior() for c in cases])
    pick = [int(k) for k in np.argsort(np.abs(sizes - np.median(sizes)))[:threads]]
    td = tempfile.mkdtemp()
    jobs = []
    for k in pick:
        c = cases[k]
        hm = c.hmesh()
        path = os.path.join(td, f"inst{k}.hmsh")
        write_hmsh(path, hm)
        vc = ol.VecRefCase(path, c.gamma, [0] + [int(b) for b in c.bctypes], c.qbc)
        q = c.qic
        U0 = np.zeros((4, hm.Nc))
        U0[0], U0[1], U0[2] = q[0], q[0] * q[1], q[0] * q[2]
        U0[3] = q[3] / (c.gamma - 1) + 0.5 * q[0] * (q[1] ** 2 + q[2] ** 2)
        jobs.append((k, vc, U0, float(dts[k])))

    def probe(job):
        k, vc, U0, dt = job
        t0 = time.perf_counter()
        U = vc.run(U0, 1, dt)
        return k, U, time.perf_counter() - t0

    with cf.ThreadPoolExecutor(len(jobs)) as ex:
        first = {k: (U, sec) for k, U, sec in ex.map(probe, jobs)}
        nsteps = {k: max(1, int(target_s / max(first[k][1], 1e-6))) for k, _, _, _ in jobs}

        def timed(job):
            k, vc, _, dt = job
            vc.run(first[k][0], nsteps[k], dt)
            return k

        t0 = time.perf_counter()
        list(ex.map(timed, jobs))
        wall = time.perf_counter() - t0
    for _, vc, _, _ in jobs:
        vc.close()
    for k in pick:
        os.remove(os.path.join(td, f"inst{k}.hmsh"))
    os.rmdir(td)
    work = sum(int(sizes[k]) * nsteps[k] for k in pick)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": work / wall, "unit": "cell-updates/s", "cores": len(jobs), "kind": "reference",
            "sample": f"reference VEC layout (AVX2 build, oracle/_ref, FP64: the reference has no other precision), "
                      f"one instance per thread on {len(jobs)} threads of {cpu}: the {len(jobs)} instances nearest "
                      f"the batch's median size ({int(np.median(sizes)):,} interior cells), each stepped "
                      f"{min(nsteps.values())}-{max(nsteps.values())} steps of its GPU run's last dt after one "
                      f"warm-up step from its free stream",
            "seconds": round(wall, 3), "instances": pick}


def run_batch(name: str, steps: int, warmup: int, tune: bool, compare_fp64: bool, err_steps: int,
              sensitivity: bool, cpu_s: float, secondary: bool, force_width: int = 0):
    """Batched wedges (one launch per step for all instances, hipGraph replay,
    autotuned schedule).  value = interior cells of all instances x steps / wall
    over exactly `steps` steps after `warmup`.  Before the timed steps each batch
    runs the workload's quoted step count (err_steps) from its initial state, so
    the FP32-mixed batch can be compared with the FP64 batch after the same steps;
    those runs are outside the timed region."""
    import numpy as np
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    key = {"c4": "c4_batch_fp64", "c5": "c5_batch_mixed"}[name]
    w = C.WORKLOADS[key]
    steps = steps or w["steps"]
    err_steps = w["steps"] if err_steps is None else err_steps
    cfl = w["cfl"]
    prec = w["precision"]
    t0 = time.perf_counter()
    specs = C.wedge_batch_specs(w["n"], w["seed"], w["mach"], w["angle"], w["tris"])
    cases = [C.ramp_case(sp["nx"], sp["ny"], sp["mach"], sp["angle"], seed=sp["seed"]) for sp in specs]
    gen_s = time.perf_counter() - t0
    log(f"{name}: generated {len(cases)} meshes in {gen_s:.1f} s")
    flags = L.FLAG_GRAPH | L.FLAG_ORDER_SFC

    def one(precision, perturb=0.0, sched=None, timed=True, keep=None):
        """Builds the batch, runs err_steps steps from the initial state (untimed),
        then `warmup` + exactly `steps` timed steps; returns the context, its record,
        the per-instance states after the err_steps steps (variables in `keep`)."""
        t1 = time.perf_counter()
        ctx = C.make_batch(cases, precision, flags)
        setup = time.perf_counter() - t1
        inf = ctx.info()
        log(f"{'mixed' if precision else 'fp64'} batch: {inf.Nint:,} interior cells, built in {setup:.1f} s")
        if perturb:
            # every conserved value of every cell times (1 + perturb * N(0,1)), seeded
            rng = np.random.default_rng(12345)
            for k in range(ctx.ninst):
                U = ctx.inst_state(k)
                ctx.set_inst_state(k, U * (1.0 + perturb * rng.standard_normal(U.shape)))
            log(f"initial states perturbed by {perturb:g} (relative)")
        if sched is None and force_width:
            sched = ctx.schedule() | {"aosoa_width": force_width}
        if sched is not None:
            ctx.set_schedule(**sched)
        else:
            sched = ctx.autotune(steps=4, cfl=cfl) if tune else ctx.schedule()
        log(f"schedule {sched}")
        rec = {"setup_s": round(setup, 2), "schedule": sched, "Nc": inf.Nc, "Nint": inf.Nint, "Nf": inf.Nf,
               "device_bytes": inf.device_bytes}
        states = None
        if err_steps:
            t1 = time.perf_counter()
            ctx.step(err_steps, cfl=cfl)
            log(f"{err_steps} steps from the initial state in {time.perf_counter() - t1:.1f} s (untimed)")
            tb, dtb, fb = ctx.status()
            rec["err_run"] = {"steps": err_steps, "faulted_instances": int(fb.sum()),
                              "faulted": [{"instance": int(k), "mach": round(specs[k]["mach"], 3),
                                           "angle": round(specs[k]["angle"], 2),
                                           "triangles": 2 * specs[k]["nx"] * specs[k]["ny"]}
                                          for k in np.nonzero(fb)[0]]}
            rec["last_dt"] = dtb
            if keep is not None:
                states = [ctx.inst_state(k)[keep].copy() for k in range(ctx.ninst)]
        if timed:
            ctx.step(warmup, cfl=cfl)
            ctx.sync()
            t1 = time.perf_counter()
            ctx.step(steps, cfl=cfl)          # graph replays; returns after the stream drained
            wall = time.perf_counter() - t1
            log(f"{steps} timed steps in {wall:.3f} s")
            nev = min(steps, 20)
            tm = ctx.time_steps(nev, cfl=cfl)   # per-kernel event timing, eager launches
            tb, dtb, fb = ctx.status()
            rec |= {"value": inf.Nint * steps / wall, "ms_per_step": 1e3 * wall / steps,
                    "kernel_ms": ({"fused_step": tm["pass_a_ms"]} if sched["fused"] else
                                  {"pass_a": tm["pass_a_ms"], "pass_b": tm["pass_b_ms"]}) |
                                 {"dt_reduce": tm["dt_reduce_ms"]},
                    "faulted_instances": int(fb.sum()),
                    "steps_total": err_steps + warmup + steps + nev}
            rec.setdefault("last_dt", dtb)
        return ctx, rec, states

    keep_all = slice(0, 4) if (compare_fp64 and prec == L.MIXED) else None
    ctx, rec, Um = one(prec, keep=keep_all)
    a_b, b_b, step_b = alg_bytes(rec["Nc"], rec["Nf"], prec == L.MIXED)
    if rec["schedule"]["fused"]:
        kern = {"fused_step": (rec["kernel_ms"]["fused_step"], step_b)}
    else:
        kern = {"pass_a": (rec["kernel_ms"]["pass_a"], a_b), "pass_b": (rec["kernel_ms"]["pass_b"], b_b)}
    dom = max(kern, key=lambda k: kern[k][0])
    achieved = kern[dom][1] / (kern[dom][0] * 1e-3) / 1e9
    traffic = read_traffic(name)
    line = {
        "metric": METRIC, "value": rec["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": rec["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if prec == L.MIXED else "f64", "data": "synthetic",
        "config": {"workload": f"{name}: {w['n']} wedges, Mach U{list(w['mach'])}, angle U{list(w['angle'])} deg, "
                               f"log-U{[int(x) for x in w['tris']]} triangles (seed {w['seed']}), "
                               f"{'FP32-mixed (FP32 differences, gradients, limiter and flux; residual accumulated in FP64; FP64 update and dt)' if prec == L.MIXED else 'FP64'}, "
                               f"one launch per step for all instances, hipGraph replay, "
                               f"{'autotuned' if tune else 'default'} schedule, CFL {cfl}; timed steps follow "
                               f"{err_steps} untimed steps from the initial state",
                   "instances": w["n"], "generate_s": round(gen_s, 1), **{k: rec[k] for k in
                   ("Nc", "Nint", "Nf", "setup_s", "schedule", "device_bytes")}},
        "hbm_pct_step": 100.0 * step_b / (rec["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "alg_bytes_per_step": step_b,
        "kernel_ms": rec["kernel_ms"],
        "roofline": roofline_record(dom, kern[dom][1], kern[dom][0], (traffic or {}).get(dom)),
        "faulted_instances": rec["faulted_instances"],
        "err_run": rec.get("err_run"),
        # plain 16-byte-per-lane copy of 1 GiB each way: the achievable HBM ceiling,
        # and the known-size kernel the rocprofv3 byte counters are calibrated with
        "copy_ceiling_gbs": L.copy_bandwidth(1 << 30, 20),
    }
    line["roofline"]["frac_of_copy_ceiling"] = achieved / line["copy_ceiling_gbs"]
    last_dt = rec["last_dt"]
    ctx.close()
    if compare_fp64 and prec == L.MIXED:
        ctx64, rec64, S64 = one(L.FP64, timed=True, keep=slice(0, 4))
        sched64 = rec64["schedule"]
        line["fp64_batch"] = {"value": rec64["value"], "ms_per_step": rec64["ms_per_step"],
                              "kernel_ms": rec64["kernel_ms"], "schedule": sched64,
                              "faulted_instances": rec64["faulted_instances"], "err_run": rec64.get("err_run")}
        line["speedup_vs_fp64_batch"] = rec["value"] / rec64["value"]
        if S64 is None:  # --err-steps 0: timings only
            ctx64.close()
            compare_fp64 = False
    if compare_fp64 and prec == L.MIXED:
        names = ("rho", "rho_u", "rho_v", "rho_E")
        rel = np.full((4, ctx64.ninst), np.nan)
        mabs = np.zeros(4)
        Ua = []
        for k in range(ctx64.ninst):
            Ua_k = S64[k]
            S64[k] = None
            Ua.append(Ua_k[0].copy())
            d = Um[k] - Ua_k
            rel[:, k] = np.linalg.norm(d, axis=1) / np.linalg.norm(Ua_k, axis=1)
            fin = np.isfinite(d).all(axis=1)
            mabs = np.where(fin, np.maximum(mabs, np.nanmax(np.abs(d), axis=1)), mabs)
        ctx64.close()
        del Um
        e = rel[0]
        fin = np.isfinite(e)
        line["error_vs_fp64"] = {
            "what": f"per instance, relative L2 of each conserved variable of the FP32-mixed batch against the "
                    f"FP64 batch after the same {err_steps} CFL steps from the same initial state; max_abs over all "
                    f"cells of the instances that stayed finite",
            **{names[v]: {"rel_l2_median": float(np.median(rel[v][np.isfinite(rel[v])])) if fin.any() else None,
                          "rel_l2_max": float(np.max(rel[v][np.isfinite(rel[v])])) if fin.any() else None,
                          "max_abs": float(mabs[v])} for v in range(4)}}
        line["rho_rel_l2_vs_fp64"] = {"median": float(np.median(e[fin])) if fin.any() else None,
                                      "max": float(e[fin].max()) if fin.any() else None,
                                      "non_finite_instances": int((~fin).sum()),
                                      "per_instance": [float(x) for x in e]}
        line["rho_error_bound"] = {"bound": "rho rel L2 <= 1e-5 per instance",
                                   "instances_within": int((e[fin] <= 1e-5).sum()),
                                   "finite_instances": int(fin.sum()),
                                   "median_within": bool(fin.any() and np.median(e[fin]) <= 1e-5)}
        if sensitivity:
            # the FP64 path's own sensitivity: the same run from initial states
            # perturbed by 1e-7 (relative), about one FP32 rounding, once at t = 0
            ctxp, _, Sp = one(L.FP64, perturb=1e-7, sched=sched64, timed=False, keep=0)
            fl = np.array([float(np.linalg.norm(Sp[k] - Ua[k]) / np.linalg.norm(Ua[k])) for k in range(ctxp.ninst)])
            ctxp.close()
            ok = fin & np.isfinite(fl)
            line["fp64_sensitivity"] = {
                "what": "rho rel L2 of the FP64 batch run from initial states perturbed by 1e-7 (relative, seeded) "
                        "against the unperturbed FP64 run, same steps",
                "median": float(np.median(fl[ok])) if ok.any() else None,
                "max": float(fl[ok].max()) if ok.any() else None,
                "instances_above_1e-5": int((fl[ok] > 1e-5).sum()),
                "per_instance": [float(x) for x in fl]}
    if cpu_s > 0:
        line["cpu_baseline"] = batch_cpu_baseline(cases, specs, None, last_dt, cpu_s, 8)
        if line["cpu_baseline"]:
            line["gpu_over_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    if secondary:
        line["secondary"] = secondary_lines()
    return line


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the workload's)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c5", choices=["c2", "c3", "c3mixed", "c4", "c5", "c2rk", "c2out"],
                    help="c5 (default): the 1,024-wedge FP32-mixed batch against the FP64 batch")
    ap.add_argument("--no-compare", action="store_true", help="c5: skip the FP64 batched comparison run")
    ap.add_argument("--err-steps", type=int, default=None,
                    help="batches: untimed steps from the initial state before the timed ones, after which the "
                         "mixed and FP64 batches are compared (default: the workload's quoted steps)")
    ap.add_argument("--no-sensitivity", action="store_true", help="c5: skip the perturbed FP64 run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length (0 = skip)")
    ap.add_argument("--graph", action="store_true", help="single meshes: replay a captured step graph in the "
                    "timed region (measured no faster than eager launches on C2)")
    ap.add_argument("--order", default="sfc", choices=["sfc", "rcm"],
                    help="interior cell order: Hilbert curve of centroids (sfc) or reverse Cuthill-McKee (rcm)")
    ap.add_argument("--tune", action="store_true", help="run the autotuner before timing")
    ap.add_argument("--no-tune", action="store_true", help="batches: keep the default schedule")
    ap.add_argument("--aosoa-width", type=int, default=0, choices=[0, 1, 8, 64],
                    help="batches: the default schedule with this state tile width instead of the autotuner's choice "
                         "(the profiles' commands use the width the autotuner picks, to leave its trial runs out)")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--step-path", default="default", choices=["default", "single", "two-pass"],
                    help="single meshes: force the single-kernel step or the two passes")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        return 2
    if a.workload == "c2rk":
        line = run_integrators(a.steps or 100)
    elif a.workload == "c2out":
        line = run_output_overlap(a.steps or 500, 50)
    elif a.workload in ("c4", "c5"):
        line = run_batch(a.workload, a.steps, a.warmup, not a.no_tune, not a.no_compare, a.err_steps,
                         not a.no_sensitivity, a.cpu_seconds, not a.no_secondary, a.aosoa_width)
    else:
        line = run_single(a.workload, a.steps, a.warmup, a.cpu_seconds, a.graph, not a.no_secondary, a.tune,
                          a.order, a.step_path)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
