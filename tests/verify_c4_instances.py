"""Per-instance agreement of the 256-wedge FP64 batch (Mach U[1.5, 4.0], angle
U[5, 25] deg, 20k-500k triangles log-uniform, seed 1) with the reference's own
CPU solver over the quoted 2,000 CFL steps at CFL 0.5 (--batch c5: the same for
the 1,024 wedges of the mixed-precision sweep run in FP64, the comparator of its
error figures).

    python tests/verify_c4_instances.py [--count N --max-triangles T | --range A B]
                                        [--jobs J] [--log-dir D]

Phase 1 (this process, the GPU): the whole batch in one context (default
schedule, hipGraph replay), 2,000 steps.  For the instances checked (--range:
instances A..B-1; otherwise --count instances, default 16, chosen evenly over
those of --min-triangles to --max-triangles triangles, default 0 to 200,000) it writes the
instance's mesh as .hmsh and its initial state, final state and clock to a
temporary directory.  Instances that go non-physical on the GPU are skipped,
unless named with --instances: those are compared with their NaN values in
place.

Phase 2 (one child process per instance, at most --jobs at once, default 16,
the largest first; no GPU): the reference's
Base solver compiled from its sources with -ffp-contract=off
(oracle/_ref/libhcfd_ref_nofma.so) on the instance's mesh from the same initial
state, each step with the CFL dt of its own state (the oracle's orc_cfl_dt: the
reference has no CFL step), compared bit for bit with the GPU's state and clock.
Each child writes its progress to <log_dir>/<batch>_ref_<instance>.log.

A script rather than a collected test: the reference needs several minutes of
one CPU core per instance.  Exits 1 if any instance differs.
"""
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402


def say(msg):
    print(msg, flush=True)


def child(npz_path: str, log_path: str) -> int:
    """Phase 2 for one instance (CPU only: the oracle and the reference build)."""
    import oracle_lib as ol
    from cfd_optimisations_amd.formats import read_hmsh
    d = np.load(npz_path)
    k, steps, cfl, gamma = int(d["inst"]), int(d["steps"]), float(d["cfl"]), float(d["gamma"])
    hm = read_hmsh(str(d["mesh"]))
    bct, qbc = d["bctypes"], d["qbc"]
    with open(log_path, "w") as log:
        def note(msg):
            log.write(msg + "\n")
            log.flush()
        od = ol.OracleDomain(hm, gamma, bct, qbc)
        od.zero_bc_iv()
        bymap = np.zeros(int(hm.coltag.max()) + 1, np.int32)
        for b in range(hm.NBCcol):
            bymap[hm.coltag[hm.Ncol - hm.NBCcol + b]] = bct[b]
        rc = ol.RefCase(str(d["mesh"]), gamma, bymap, qbc, ol.REF_NOFMA_SO)
        rc.zero_bc_iv()
        U, t = d["U0"], 0.0
        t0 = time.time()
        for s in range(1, steps + 1):
            dt = od.cfl_dt(U, cfl)
            U = rc.run(U, 1, dt)
            t += dt
            if s % 200 == 0:
                note(f"instance {k}: reference step {s}/{steps}, {(time.time() - t0) / s:.3f} s per step")
        rc.close()
        # an instance that goes non-physical carries NaN in both runs, in the same
        # places when the arithmetic is the same (NaN rates drop out of the CFL
        # step in both: the GPU's maximum rate and the oracle's minimum bound)
        tg = float(d["t"])
        same_u = np.array_equal(U, d["U"], equal_nan=True)
        same_t = t == tg or (np.isnan(t) and np.isnan(tg))
        fin = np.isfinite(U) & np.isfinite(d["U"])
        diff = np.abs(U - d["U"])[fin].max() / np.abs(U[fin]).max() if fin.any() else 0.0
        if not np.isfinite(U).all():
            note(f"instance {k}: non-physical state, {int((~np.isfinite(U)).sum()):,} non-finite values "
                 f"(GPU: {int((~np.isfinite(d['U'])).sum()):,})")
        line = (f"instance {k} ({hm.Nc:,} cells, Mach {float(d['mach']):.3f}, {float(d['angle']):.2f} deg): "
                f"state {'==' if same_u else '!='} reference (max |dU| / max |U| = {diff:.3e}), "
                f"clock {'==' if same_t else '!='} ({float(d['t'])!r} vs {t!r})")
        note(line)
    print(line, flush=True)
    return 0 if same_u and same_t else 1


def main(argv=None) -> int:
    import argparse
    import oracle_lib as ol
    from cfd_optimisations_amd import cases as C
    from cfd_optimisations_amd import lib as L
    from cfd_optimisations_amd.formats import write_hmsh
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=16)
    ap.add_argument("--max-triangles", type=int, default=200000)
    ap.add_argument("--min-triangles", type=int, default=0)
    ap.add_argument("--range", type=int, nargs=2)
    ap.add_argument("--jobs", type=int, default=16)
    ap.add_argument("--log-dir", default=os.path.join(ROOT, "gpurun_out"))
    ap.add_argument("--instances", type=int, nargs="+", help="check these instances (faulted ones included)")
    ap.add_argument("--batch", choices=("c4", "c5"), default="c4",
                    help="c5: the 1,024 wedges of the mixed-precision sweep, run here in FP64 (the "
                         "comparator of its error figures); instances that fault are skipped")
    a = ap.parse_args(argv)
    if not ol.ref_available(ol.REF_NOFMA_SO):
        say("the reference build oracle/_ref/libhcfd_ref_nofma.so is missing")
        return 2
    w = C.WORKLOADS["c4_batch_fp64" if a.batch == "c4" else "c5_batch_mixed"]
    steps, cfl = w["steps"], w["cfl"]
    specs = C.wedge_batch_specs(w["n"], w["seed"], w["mach"], w["angle"], w["tris"])
    cases = [C.ramp_case(sp["nx"], sp["ny"], sp["mach"], sp["angle"], seed=sp["seed"]) for sp in specs]
    tris = np.array([2 * sp["nx"] * sp["ny"] for sp in specs])
    if a.instances:
        pick = sorted(set(k for k in a.instances if 0 <= k < len(cases)))
        say(f"{len(cases)} instances generated; checking instances {pick}")
    elif a.range:
        pick = list(range(max(a.range[0], 0), min(a.range[1], len(cases))))
        say(f"{len(cases)} instances generated; checking instances {pick[0]}..{pick[-1]}")
    else:
        small = [k for k in np.argsort(tris, kind="stable") if a.min_triangles <= tris[k] <= a.max_triangles]
        n = a.count
        pick = sorted({int(small[int(round(j * (len(small) - 1) / max(n - 1, 1)))]) for j in range(n)})
        say(f"{len(cases)} instances generated; checking {len(pick)} of the {len(small)} with "
            f"{a.min_triangles:,} to {a.max_triangles:,} triangles: {pick}")
    os.makedirs(a.log_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as td:
        t0 = time.time()
        with C.make_batch(cases, L.FP64, L.FLAG_GRAPH | L.FLAG_ORDER_SFC) as ctx:
            U0 = {k: ctx.inst_state(k) for k in pick}
            ctx.step(steps, cfl=cfl)
            tb, _, fb = ctx.status()
            Ug = {k: ctx.inst_state(k) for k in pick}
            sched = ctx.schedule()
        say(f"GPU: {steps} steps of the batch, schedule {sched}, {int(fb.sum())} faulted instances, "
            f"{time.time() - t0:.1f} s")
        skipped = [] if a.instances else [k for k in pick if fb[k]]
        if skipped:
            say(f"skipped (faulted on the GPU, as under the reference method): {skipped}")
            pick = [k for k in pick if not fb[k]]
        jobs = []
        for k in sorted(pick, key=lambda k: -tris[k]):
            mesh = os.path.join(td, f"w{k}.hmsh")
            write_hmsh(mesh, cases[k].hmesh())
            npz = os.path.join(td, f"w{k}.npz")
            np.savez(npz, inst=k, steps=steps, cfl=cfl, gamma=C.GAMMA, mesh=mesh, bctypes=cases[k].bctypes,
                     qbc=cases[k].qbc, U0=U0[k], U=Ug[k], t=tb[k], mach=specs[k]["mach"], angle=specs[k]["angle"])
            jobs.append((k, npz, os.path.join(a.log_dir, f"{a.batch}_ref_{k}.log")))
        del U0, Ug
        running, bad, done = [], 0, 0
        while jobs or running:
            while jobs and len(running) < a.jobs:
                k, npz, log = jobs.pop(0)
                running.append((k, subprocess.Popen([sys.executable, "-u", os.path.abspath(__file__), "--ref", npz,
                                                     log], stdout=subprocess.DEVNULL)))
            time.sleep(0.5)
            for k, p in list(running):
                rc = p.poll()
                if rc is None:
                    continue
                running.remove((k, p))
                done += 1
                bad += rc != 0
                say(f"instance {k} ({tris[k]:,} triangles): {'PASS' if rc == 0 else 'FAIL'} [{done}/{len(pick)}]")
    say(f"PASS: all {len(pick)} instances checked are bit-identical to the reference, state and clock" if not bad
        else f"FAIL: {bad} instance(s) differ")
    return 1 if bad else 0


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--ref":
        sys.exit(child(sys.argv[2], sys.argv[3]))
    sys.exit(main())
