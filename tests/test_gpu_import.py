"""GPU parity on meshes imported from gmsh (.msh v4 ASCII, interface/Mesh.py:21-127,
159-378): the importer's mesh goes through hcfd_create like a .hmsh file and the
HIP path must equal the oracle on the same imported mesh."""
import numpy as np
import pytest

from cfd_optimisations_amd import cases as C
from cfd_optimisations_amd import lib as L

import oracle_lib as ol
from test_host import write_square_msh

pytestmark = pytest.mark.gpu

TWO_PASS = dict(block=256, lanes_per_cell=1, aosoa_width=1, persistent=0, xcd_remap=1, fused=0)
FUSED = dict(TWO_PASS, fused=1)


def imported_case(tmp_path, nx=48, ny=30, quads_left=16):
    """[0,1]^2 with quadrilaterals on the left third and triangles elsewhere; a
    Mach-2 stream entering at 10 degrees: SupInlet left, SupOutlet right, Wall at
    the bottom, Symmetry at the top (colours bottom, right, top, left)."""
    p = str(tmp_path / "mixed.msh")
    write_square_msh(p, nx, ny, quads_left=quads_left)
    mesh = L.NativeMesh.import_msh(p)
    a = np.radians(10.0)
    q = np.array([1.0, 2.0 * np.cos(a), 2.0 * np.sin(a), 1.0 / 1.4])
    bct = np.array([L.WALL, L.SUPOUTLET, L.SYMMETRY, L.SUPINLET], np.int32)
    return C.Case(mesh, bct, np.tile(q, (4, 1)), qic=q, meta={"name": "imported mixed cells"})


@pytest.mark.parametrize("sched", [TWO_PASS, FUSED], ids=["two_pass", "fused"])
@pytest.mark.parametrize("order", [0, L.FLAG_ORDER_SFC], ids=["rcm", "hilbert"])
def test_imported_mixed_cells_exact(tmp_path, sched, order):
    case = imported_case(tmp_path)
    hm = case.hmesh()
    assert set(np.diff(hm.cn_off)) == {2, 3, 4}   # ghost lines, triangles, quadrilaterals
    od = ol.OracleDomain(hm, case.gamma, case.bctypes, case.qbc)
    od.zero_bc_iv()
    Uo, to = od.run(od.initialise(case.qic), 80, cfl=0.5)
    with C.make_context(case, L.FP64, order) as ctx:
        inf = ctx.info()
        assert inf.maxdeg == 4
        if sched["fused"]:
            assert inf.fused_available
        ctx.set_schedule(**sched)
        t = ctx.step(80, cfl=0.5)
        assert t == to
        assert np.array_equal(ctx.get_state(), Uo)
        assert ctx.fault() == 0


def test_imported_mixed_cells_fp32_mixed(tmp_path):
    """The FP32-mixed path on the imported mesh: its single kernel equals its two
    passes bit for bit, and it stays within its bound of the FP64 result."""
    case = imported_case(tmp_path)
    out = {}
    for name, prec, sched in (("fp64", L.FP64, FUSED), ("mixed_fused", L.MIXED, FUSED),
                              ("mixed_two_pass", L.MIXED, TWO_PASS)):
        with C.make_context(case, prec, L.FLAG_ORDER_SFC) as ctx:
            ctx.set_schedule(**sched)
            ctx.step(200, cfl=0.5)
            out[name] = ctx.get_state()
    assert np.array_equal(out["mixed_fused"], out["mixed_two_pass"])
    e = np.linalg.norm(out["mixed_fused"][0] - out["fp64"][0]) / np.linalg.norm(out["fp64"][0])
    assert e <= 1e-5, e
