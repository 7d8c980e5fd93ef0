"""The C++ drop-in classes (namespace MI355X) run the reference's own case
(src/Test.cpp:50-74) through the C ABI; their .hsln output must match the
oracle bit for bit and the shipped golden file to FMA noise."""
import os
import subprocess

import numpy as np
import pytest

from cfd_optimisations_amd import cases as C
from cfd_optimisations_amd.formats import read_hsln
import oracle_lib as ol
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

RUN = os.path.join(ROOT, "cfd_optimisations_amd", "hcfd_run")


def oracle_prims(wedge_mesh, steps):
    case = C.reference_wedge_case(wedge_mesh)
    od = ol.OracleDomain(wedge_mesh, case.gamma, case.bctypes, case.qbc)
    od.zero_bc_iv()
    U0 = od.initialise(case.qic)
    U, _ = od.run(U0, steps, dt=1e-3)
    return od.prims(U0), od.prims(U)


@pytest.mark.parametrize("mode", ["fused", "stages"])
def test_dropin_driver_matches_oracle(tmp_path, wedge_mesh, mode):
    steps = 1000 if mode == "fused" else 100
    stem = str(tmp_path / "MI355X")
    args = [RUN, os.path.join(GOLDEN, "wedge.hmsh"), stem, str(steps)] + (["--stages"] if mode == "stages" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "MI355X" in r.stdout
    Q0, Q1 = oracle_prims(wedge_mesh, steps)
    _, S1 = read_hsln(stem + "_1.hsln")
    assert np.array_equal(S1, Q1)
    if mode == "fused":
        _, S0 = read_hsln(stem + "_0.hsln")
        assert np.array_equal(S0, Q0)
        _, B1 = read_hsln(os.path.join(GOLDEN, "Base_1.hsln"))
        err = [np.linalg.norm(S1[v] - B1[v]) / np.linalg.norm(B1[v]) for v in range(4)]
        assert max(err) < 1e-13
        assert os.path.getsize(stem + "_1.hsln") == os.path.getsize(os.path.join(GOLDEN, "Base_1.hsln"))


def test_dropin_driver_reports_bad_mesh(tmp_path):
    bad = tmp_path / "bad.hmsh"
    bad.write_bytes(b"\x02\x00\x00\x00" * 3)
    r = subprocess.run([RUN, str(bad), str(tmp_path / "x"), "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "truncated" in r.stderr


def _run(args):
    r = subprocess.run([RUN] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def _output_steps(T, dt, dt_out):
    """Steps after which Euler::run writes a file (src/Base/Euler.cpp:83-104)."""
    t, t_out, n, out = 0.0, dt_out, 0, []
    while t < T:
        t += dt
        n += 1
        if t >= t_out:
            out.append(n)
            t_out += dt_out
    return out


def test_dropin_outputs_written_while_stepping(tmp_path, wedge_mesh):
    """Four outputs over 400 steps: each file, written by the host while the device
    steps on, holds the oracle's primitives of its own step."""
    stem = str(tmp_path / "out")
    _run([os.path.join(GOLDEN, "wedge.hmsh"), stem, 400, "--outputs", 4])
    case = C.reference_wedge_case(wedge_mesh)
    od = ol.OracleDomain(wedge_mesh, case.gamma, case.bctypes, case.qbc)
    od.zero_bc_iv()
    U = od.initialise(case.qic)
    _, S0 = read_hsln(stem + "_0.hsln")
    assert np.array_equal(S0, od.prims(U))
    done = 0
    steps = _output_steps(400 * 1e-3, 1e-3, 400 * 1e-3 / 4)
    assert len(steps) == 4
    for k, n in enumerate(steps, start=1):
        U, _ = od.run(U, n - done, dt=1e-3)
        done = n
        _, Sk = read_hsln(f"{stem}_{k}.hsln")
        assert np.array_equal(Sk, od.prims(U)), k
    assert not os.path.exists(f"{stem}_5.hsln")


def test_dropin_ssprk_and_restart(tmp_path, wedge_mesh):
    """SSP-RK3 through the drop-in Euler, then a restart from its output file:
    both equal the oracle (the restart starts from the file's primitives)."""
    stem = str(tmp_path / "rk")
    _run([os.path.join(GOLDEN, "wedge.hmsh"), stem, 150, "--rk", 3])
    case = C.reference_wedge_case(wedge_mesh)
    od = ol.OracleDomain(wedge_mesh, case.gamma, case.bctypes, case.qbc)
    od.zero_bc_iv()
    U, _ = od.run_ssprk(od.initialise(case.qic), 150, 3, dt=1e-3)
    _, S1 = read_hsln(stem + "_1.hsln")
    assert np.array_equal(S1, od.prims(U))
    stem2 = str(tmp_path / "again")
    _run([os.path.join(GOLDEN, "wedge.hmsh"), stem2, 100, "--restart", stem + "_1.hsln"])
    _, R0 = read_hsln(stem2 + "_0.hsln")
    U2 = od.initialise_prims(S1)
    assert np.array_equal(R0, od.prims(U2))
    U3, _ = od.run(U2, 100, dt=1e-3)
    _, R1 = read_hsln(stem2 + "_1.hsln")
    assert np.array_equal(R1, od.prims(U3))
    r = subprocess.run([RUN, os.path.join(GOLDEN, "wedge.hmsh"), str(tmp_path / "x"), "1", "--restart",
                        str(tmp_path / "missing.hsln")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_dropin_checkpoint_resume(tmp_path):
    """hcfd_run: 100 steps straight through, against 60 steps that end with
    --checkpoint and 40 more that start with --resume: the final files are equal bit
    for bit (the reference's t += dt loop takes 60, 40 and 100 steps for these ends)."""
    w = os.path.join(GOLDEN, "wedge.hmsh")
    ck = tmp_path / "a.ckp"
    _run([w, tmp_path / "full", 100])
    _run([w, tmp_path / "a", 60, "--checkpoint", ck])
    _run([w, tmp_path / "b", 40, "--resume", ck])
    _, F = read_hsln(str(tmp_path / "full_1.hsln"))
    _, B = read_hsln(str(tmp_path / "b_1.hsln"))
    assert np.array_equal(F, B)
    r = subprocess.run([RUN, w, str(tmp_path / "x"), "1", "--resume", str(tmp_path / "missing.ckp")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_dropin_mixed_domain_within_bound(tmp_path, wedge_mesh):
    """MI355X::Domain(mesh, HCFD_MIXED) through the drop-in Euler: src/Test.cpp's
    1000 steps stay within the mixed path's stated bound of the oracle (relative
    L2 of the density 1e-5), and the per-stage classes refuse the mixed Domain."""
    stem = str(tmp_path / "mixed")
    _run([os.path.join(GOLDEN, "wedge.hmsh"), stem, 1000, "--mixed"])
    _, Q1 = oracle_prims(wedge_mesh, 1000)
    _, S1 = read_hsln(stem + "_1.hsln")
    err = [np.linalg.norm(S1[v] - Q1[v]) / np.linalg.norm(Q1[v]) for v in range(4)]
    assert err[0] < 1e-5 and max(err) < 1e-4, err
    assert not np.array_equal(S1, Q1)  # the mixed arithmetic did run
    r = subprocess.run([RUN, os.path.join(GOLDEN, "wedge.hmsh"), str(tmp_path / "x"), "1", "--stages", "--mixed"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "FP64" in r.stderr


def test_case_file_wedge_matches_oracle(tmp_path, wedge_mesh):
    """src/Test.cpp's case written as a case file (hcfd_run --case): the same
    files as the oracle's run, bit for bit."""
    from test_host import WEDGE_CASE
    p = tmp_path / "wedge.ini"
    p.write_text(WEDGE_CASE.format(mesh=os.path.join(GOLDEN, "wedge.hmsh"), T=1.0, stem="w"))
    r = subprocess.run([RUN, "--case", str(p)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    Q0, Q1 = oracle_prims(wedge_mesh, 1000)
    assert np.array_equal(read_hsln(str(tmp_path / "w_0.hsln"))[1], Q0)
    assert np.array_equal(read_hsln(str(tmp_path / "w_1.hsln"))[1], Q1)


def test_case_file_cfl_ssprk3_matches_oracle(tmp_path):
    """A case file with expression initial states, SSP-RK3, a CFL number and an
    output cadence: every file equals the oracle's state at the same step."""
    from cfd_optimisations_amd import lib as L
    from cfd_optimisations_amd.formats import read_hmsh
    from test_host import RIEMANN_CASE, riemann_ic_np
    nm = L.NativeMesh.riemann(40, 30, 0.2, 6, True)
    mp = str(tmp_path / "r.hmsh")
    nm.save(mp)
    hm = read_hmsh(mp)
    T, dtout = 0.06, 0.02
    (tmp_path / "r.ini").write_text(RIEMANN_CASE.format(mesh=mp, prec="fp64", scheme="ssprk3", T=T, dtout=dtout))
    r = subprocess.run([RUN, "--case", str(tmp_path / "r.ini")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    od = ol.OracleDomain(hm, 1.4, np.full(4, 4), np.tile([1.5, 0.0, 0.0, 1.5], (4, 1)))
    od.zero_bc_iv()
    U = od.initialise_prims(riemann_ic_np(od.geometry()["xc"]))
    files = [od.prims(U)]
    t, t_out = 0.0, dtout
    while t < T:
        U, h = od.run_ssprk(U, 1, 3, cfl=0.4)
        t += h
        if t >= t_out:
            files.append(od.prims(U))
            t_out += dtout
    assert len(files) >= 4 and np.isfinite(files[-1]).all()
    for k, Qo in enumerate(files):
        _, Qg = read_hsln(str(tmp_path / f"r_{k}.hsln"))
        assert np.array_equal(Qg, Qo), k
    assert not os.path.exists(tmp_path / f"r_{len(files)}.hsln")
