"""ctypes binding of libhcfd.so (the C ABI in include/hcfd.h).

There is no fallback: if the HIP library is missing or no GPU is visible, the
calls raise.  Arrays cross the boundary as numpy float64/int32, per variable
``(4, Nc)`` in the caller's cell order.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from .formats import HMesh

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhcfd.so")

FP64, MIXED = 0, 1
FLAG_EXACT, FLAG_NO_RENUMBER, FLAG_GRAPH, FLAG_ORDER_SFC = 1, 2, 4, 8
SYMMETRY, WALL, SUPINLET, SUBINLET, SUPOUTLET, SUBOUTLET = range(6)
ARR_U, ARR_G, ARR_PHI, ARR_R = range(4)

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)


class HcfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hcfd error {code}: {msg}")
        self.code = code


class MeshView(ctypes.Structure):
    _fields_ = [("dim", I), ("Nn", I), ("Ncol", I), ("NBCcol", I), ("Nc", I), ("Nf", I),
                ("coltag", IP), ("coloff", IP), ("nodes", DP), ("fn_con", IP), ("fc_con", IP),
                ("cn_off", IP), ("cn_con", IP)]


class Case(ctypes.Structure):
    _fields_ = [("bctype", IP), ("qbc", DP), ("qic", D * 4)]


class Schedule(ctypes.Structure):
    _fields_ = [("block", I), ("lanes_per_cell", I), ("aosoa_width", I), ("persistent", I), ("xcd_remap", I),
                ("fused", I)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Info(ctypes.Structure):
    _fields_ = [("Nc", I), ("Nf", I), ("Nint", I), ("Nghost", I), ("ninst", I), ("maxdeg", I),
                ("precision", I), ("flags", I), ("bandwidth_in", I), ("bandwidth_out", I), ("ncolours", I),
                ("device_bytes", ctypes.c_longlong), ("sched", Schedule), ("fused_available", I),
                ("fused_lds_bytes", I), ("halo_max_cells", I)]


_lib = None


def load() -> ctypes.CDLL:
    """Loads libhcfd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cfd_optimisations_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.hcfd_last_error.restype = ctypes.c_char_p
    lib.hcfd_mesh_pname.restype = ctypes.c_char_p
    lib.hcfd_mesh_pname.argtypes = [P, I]
    sig = {
        "hcfd_device_count": [P],
        "hcfd_create": [P, I, I, P],
        "hcfd_batch_create": [I, P, P, D, I, I, P],
        "hcfd_destroy": [P],
        "hcfd_get_info": [P, P],
        "hcfd_get_geometry": [P] * 8,
        "hcfd_get_cf": [P, P, P],
        "hcfd_get_permutation": [P, P],
        "hcfd_get_face_colours": [P, P, P],
        "hcfd_set_bcs": [P, I, P, P, D],
        "hcfd_initialise": [P, P],
        "hcfd_initialise_prims": [P, P],
        "hcfd_set_state": [P, P],
        "hcfd_get_state": [P, P],
        "hcfd_get_prims": [P, P],
        "hcfd_stage_set": [P, I, P],
        "hcfd_stage_get": [P, I, P],
        "hcfd_stage_gradient": [P],
        "hcfd_stage_limiter": [P],
        "hcfd_stage_bc_grads": [P],
        "hcfd_stage_bc_limiter": [P],
        "hcfd_stage_flux": [P],
        "hcfd_stage_update": [P, D],
        "hcfd_stage_bc_cells": [P],
        "hcfd_step": [P, I, D, D, P],
        "hcfd_get_last_dt": [P, P],
        "hcfd_checkpoint_save": [P, ctypes.c_char_p],
        "hcfd_checkpoint_load": [P, ctypes.c_char_p],
        "hcfd_get_fault": [P, P],
        "hcfd_time_steps": [P, I, D, D, P],
        "hcfd_set_schedule": [P, P],
        "hcfd_set_integrator": [P, I],
        "hcfd_get_integrator": [P, P],
        "hcfd_get_schedule": [P, P],
        "hcfd_autotune": [P, I, D, ctypes.c_char_p, P],
        "hcfd_batch_get_state": [P, I, P],
        "hcfd_batch_get_prims": [P, I, P],
        "hcfd_batch_set_state": [P, I, P],
        "hcfd_batch_get_status": [P, P, P, P],
        "hcfd_batch_instance_size": [P, I, P, P],
        "hcfd_mesh_load": [ctypes.c_char_p, P],
        "hcfd_mesh_save": [P, ctypes.c_char_p],
        "hcfd_mesh_view_of": [P, P],
        "hcfd_mesh_free": [P],
        "hcfd_meshgen_riemann": [I, I, D, ctypes.c_ulonglong, I, P],
        "hcfd_meshgen_ramp": [I, I, D, D, D, D, D, ctypes.c_ulonglong, I, P],
        "hcfd_write_hsln": [ctypes.c_char_p, I, P, P],
        "hcfd_read_hsln": [ctypes.c_char_p, P, P],
        "hcfd_write_vtu": [ctypes.c_char_p, P, P],
        "hcfd_step_async": [P, I, D, D],
        "hcfd_step_async_noalloc": [P, I, D, D],
        "hcfd_rusanov3d": [I, D, P, P, P, P],
        "hcfd_prims3d": [I, D, P, P],
        "hcfd_cons3d": [I, D, P, P],
        "hcfd_sync": [P, P],
        "hcfd_output_begin": [P, I],
        "hcfd_output_wait": [P, I, P, P],
        "hcfd_mesh_import_msh": [ctypes.c_char_p, P],
        "hcfd_stage_bc": [P, I, I],
        "hcfd_set_gamma": [P, D],
        "hcfd_copy_bandwidth": [ctypes.c_longlong, I, P],
    }
    for name, args in sig.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = I
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        raise HcfdError(rc, load().hcfd_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return P(a.ctypes.data)


def _rows(arr: np.ndarray):
    """(4, n) contiguous float64 -> double*[4] array for the C side."""
    assert arr.dtype == np.float64 and arr.flags.c_contiguous and arr.shape[0] == 4
    return (DP * 4)(*[arr[v].ctypes.data_as(DP) for v in range(4)])


def device_count() -> int:
    n = I(0)
    check(load().hcfd_device_count(ctypes.byref(n)))
    return n.value


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """GB/s of the library's plain copy kernel (read + write bytes / time)."""
    g = D(0.0)
    check(load().hcfd_copy_bandwidth(nbytes, reps, ctypes.byref(g)))
    return g.value


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
class NativeMesh:
    """A mesh owned by the library (loaded, generated or imported)."""

    def __init__(self, handle):
        self.h = handle
        self.view = MeshView()
        check(load().hcfd_mesh_view_of(self.h, ctypes.byref(self.view)))

    @classmethod
    def load(cls, path: str) -> "NativeMesh":
        h = P()
        check(load().hcfd_mesh_load(path.encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def riemann(cls, nx: int, ny: int, jitter: float = 0.2, seed: int = 0, shuffle: bool = True) -> "NativeMesh":
        h = P()
        check(load().hcfd_meshgen_riemann(nx, ny, jitter, seed, int(shuffle), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def ramp(cls, nx: int, ny: int, length: float = 1.0, height: float = 0.5, x_ramp: float = 0.2,
             angle_deg: float = 15.0, jitter: float = 0.2, seed: int = 0, shuffle: bool = False) -> "NativeMesh":
        h = P()
        check(load().hcfd_meshgen_ramp(nx, ny, length, height, x_ramp, angle_deg, jitter, seed, int(shuffle),
                                       ctypes.byref(h)))
        return cls(h)

    @classmethod
    def import_msh(cls, path: str) -> "NativeMesh":
        h = P()
        check(load().hcfd_mesh_import_msh(path.encode(), ctypes.byref(h)))
        return cls(h)

    def save(self, path: str) -> None:
        check(load().hcfd_mesh_save(ctypes.byref(self.view), path.encode()))

    def pnames(self):
        out, k = [], 0
        while True:
            s = load().hcfd_mesh_pname(self.h, k)
            if s is None:
                return out
            out.append(s.decode())
            k += 1

    def to_hmesh(self) -> HMesh:
        v = self.view

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        cn_off = arr(v.cn_off, v.Nc + 1, np.int32)
        return HMesh(2, v.Nn, v.Ncol, v.NBCcol, v.Nc, v.Nf, self.pnames(), arr(v.coltag, v.Ncol, np.int32),
                     arr(v.coloff, v.Ncol + 1, np.int32), arr(v.nodes, 2 * v.Nn, np.float64).reshape(v.Nn, 2),
                     arr(v.fn_con, 2 * v.Nf, np.int32).reshape(v.Nf, 2), arr(v.fc_con, 2 * v.Nf, np.int32).reshape(v.Nf, 2),
                     cn_off, arr(v.cn_con, int(cn_off[-1]), np.int32))

    def close(self):
        if self.h:
            load().hcfd_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _HMeshView:
    """Keeps numpy arrays of an HMesh alive behind a MeshView."""

    def __init__(self, m: HMesh):
        self.keep = [np.ascontiguousarray(x, dtype=t) for x, t in (
            (m.coltag, np.int32), (m.coloff, np.int32), (m.nodes, np.float64), (m.fn_con, np.int32),
            (m.fc_con, np.int32), (m.cn_off, np.int32), (m.cn_con, np.int32))]
        ct, co, nd, fn, fc, cno, cnc = self.keep
        self.view = MeshView(2, m.Nn, m.Ncol, m.NBCcol, m.Nc, m.Nf, ct.ctypes.data_as(IP), co.ctypes.data_as(IP),
                             nd.ctypes.data_as(DP), fn.ctypes.data_as(IP), fc.ctypes.data_as(IP),
                             cno.ctypes.data_as(IP), cnc.ctypes.data_as(IP))


def as_view(mesh):
    """NativeMesh, HMesh or MeshView -> (MeshView, keep-alive object)."""
    if isinstance(mesh, NativeMesh):
        return mesh.view, mesh
    if isinstance(mesh, HMesh):
        hv = _HMeshView(mesh)
        return hv.view, hv
    if isinstance(mesh, MeshView):
        return mesh, None
    raise TypeError(type(mesh))


# ---------------------------------------------------------------------------
# contexts
# ---------------------------------------------------------------------------
class Context:
    """One GPU context: a mesh (or a batch of meshes) and its solver state."""

    def __init__(self, handle, keep, Nc: int, inst_sizes: Optional[Sequence[int]] = None):
        self.h = handle
        self._keep = keep
        self.Nc = Nc
        self.inst_sizes = list(inst_sizes) if inst_sizes is not None else [Nc]

    @classmethod
    def create(cls, mesh, precision: int = FP64, flags: int = 0) -> "Context":
        view, keep = as_view(mesh)
        h = P()
        check(load().hcfd_create(ctypes.byref(view), precision, flags, ctypes.byref(h)))
        return cls(h, keep, view.Nc)

    @classmethod
    def batch(cls, meshes, cases, gamma: float = 1.4, precision: int = FP64, flags: int = 0) -> "Context":
        """cases: list of (bctypes[NBC], qbc[NBC][4], qic[4])."""
        n = len(meshes)
        views, keep = [], []
        for m in meshes:
            v, k = as_view(m)
            views.append(v)
            keep.append(k)
        varr = (MeshView * n)(*views)
        carr = (Case * n)()
        for k, (bct, qbc, qic) in enumerate(cases):
            b = np.ascontiguousarray(bct, dtype=np.int32)
            q = np.ascontiguousarray(qbc, dtype=np.float64).reshape(-1)
            keep += [b, q]
            carr[k].bctype = b.ctypes.data_as(IP)
            carr[k].qbc = q.ctypes.data_as(DP)
            for v in range(4):
                carr[k].qic[v] = float(qic[v])
        h = P()
        check(load().hcfd_batch_create(n, varr, carr, gamma, precision, flags, ctypes.byref(h)))
        keep += [varr, carr]
        return cls(h, keep, sum(v.Nc for v in views), [v.Nc for v in views])

    def close(self):
        if self.h:
            load().hcfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- info -------------------------------------------------------------
    def info(self) -> Info:
        i = Info()
        check(load().hcfd_get_info(self.h, ctypes.byref(i)))
        return i

    def geometry(self):
        inf = self.info()
        Nf, Nc = inf.Nf, inf.Nc
        g = dict(xf=np.zeros((Nf, 2)), xc=np.zeros((Nc, 2)), rf=np.zeros((Nf, 2, 2)), nf=np.zeros((Nf, 2)),
                 Af=np.zeros(Nf), If=np.zeros(Nf), iv=np.zeros(Nc))
        check(load().hcfd_get_geometry(self.h, *(_ptr(g[k]) for k in ("xf", "xc", "rf", "nf", "Af", "If", "iv"))))
        return g

    def cf(self):
        inf = self.info()
        off = np.zeros(inf.Nc + 1, np.int32)
        con = np.zeros(2 * inf.Nf, np.int32)
        check(load().hcfd_get_cf(self.h, _ptr(off), _ptr(con)))
        return off, con

    def permutation(self):
        p = np.zeros(self.Nc, np.int32)
        check(load().hcfd_get_permutation(self.h, _ptr(p)))
        return p

    def face_colours(self):
        inf = self.info()
        c = np.zeros(inf.Nf, np.int32)
        n = I(0)
        check(load().hcfd_get_face_colours(self.h, _ptr(c), ctypes.byref(n)))
        return c, n.value

    # ---- setup ------------------------------------------------------------
    def set_bcs(self, bctypes, qbc, gamma: float = 1.4):
        b = np.ascontiguousarray(bctypes, dtype=np.int32)
        q = np.ascontiguousarray(qbc, dtype=np.float64).reshape(-1)
        check(load().hcfd_set_bcs(self.h, len(b), _ptr(b), _ptr(q), gamma))

    def initialise(self, qic):
        q = np.ascontiguousarray(qic, dtype=np.float64)
        check(load().hcfd_initialise(self.h, _ptr(q)))

    def initialise_prims(self, Q: np.ndarray):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        check(load().hcfd_initialise_prims(self.h, _rows(Q)))

    def set_state(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(load().hcfd_set_state(self.h, _rows(U)))

    def get_state(self) -> np.ndarray:
        U = np.zeros((4, self.Nc))
        check(load().hcfd_get_state(self.h, _rows(U)))
        return U

    def get_prims(self) -> np.ndarray:
        Q = np.zeros((4, self.Nc))
        check(load().hcfd_get_prims(self.h, _rows(Q)))
        return Q

    # ---- stages -----------------------------------------------------------
    def stage_set(self, which: int, X: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(4, -1)
        check(load().hcfd_stage_set(self.h, which, _rows(X)))

    def stage_get(self, which: int) -> np.ndarray:
        n = 2 * self.Nc if which == ARR_G else self.Nc
        X = np.zeros((4, n))
        check(load().hcfd_stage_get(self.h, which, _rows(X)))
        return X.reshape(4, 2, self.Nc) if which == ARR_G else X

    def stage(self, name: str, dt: float = 0.0):
        lib = load()
        if name == "update":
            check(lib.hcfd_stage_update(self.h, dt))
        else:
            check(getattr(lib, "hcfd_stage_" + name)(self.h))

    # ---- stepping ----------------------------------------------------------
    def step(self, nsteps: int, dt: float = 0.0, cfl: float = 0.0) -> float:
        t = D(0.0)
        check(load().hcfd_step(self.h, nsteps, dt, cfl, ctypes.byref(t)))
        return t.value

    def step_async(self, nsteps: int, dt: float = 0.0, cfl: float = 0.0) -> None:
        check(load().hcfd_step_async(self.h, nsteps, dt, cfl))

    def step_async_noalloc(self, nsteps: int, dt: float = 0.0, cfl: float = 0.0) -> None:
        """hcfd_step_async with the CFL scratch in the library's static device
        array: allocates nothing (eager launches; call sync() to wait)."""
        check(load().hcfd_step_async_noalloc(self.h, nsteps, dt, cfl))

    def sync(self) -> float:
        t = D(0.0)
        check(load().hcfd_sync(self.h, ctypes.byref(t)))
        return t.value

    def output_begin(self, slot: int) -> None:
        check(load().hcfd_output_begin(self.h, slot))

    def output_wait(self, slot: int):
        """(primitives (4, Nc) in caller order, simulated time) of the slot's snapshot."""
        Q = np.zeros((4, self.Nc))
        t = D(0.0)
        check(load().hcfd_output_wait(self.h, slot, _rows(Q), ctypes.byref(t)))
        return Q, t.value

    def checkpoint_save(self, path: str) -> None:
        """Writes the state, clocks, last dt and fault flags (hcfd_checkpoint_save)."""
        check(load().hcfd_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path: str) -> None:
        """Restores a checkpoint of the same mesh(es) (hcfd_checkpoint_load)."""
        check(load().hcfd_checkpoint_load(self.h, os.fsencode(path)))

    def last_dt(self) -> float:
        d = D(0.0)
        check(load().hcfd_get_last_dt(self.h, ctypes.byref(d)))
        return d.value

    def fault(self) -> int:
        f = I(0)
        check(load().hcfd_get_fault(self.h, ctypes.byref(f)))
        return f.value

    def time_steps(self, nsteps: int, dt: float = 0.0, cfl: float = 0.0):
        ms = (ctypes.c_float * 4)()
        check(load().hcfd_time_steps(self.h, nsteps, dt, cfl, ms))
        return {"wall_ms": ms[0], "pass_a_ms": ms[1], "pass_b_ms": ms[2], "dt_reduce_ms": ms[3]}

    def set_integrator(self, stages: int) -> None:
        """1 = forward Euler (the reference's), 2 = SSP-RK2, 3 = SSP-RK3."""
        check(load().hcfd_set_integrator(self.h, stages))

    def integrator(self) -> int:
        n = I(0)
        check(load().hcfd_get_integrator(self.h, ctypes.byref(n)))
        return n.value

    def schedule(self) -> dict:
        s = Schedule()
        check(load().hcfd_get_schedule(self.h, ctypes.byref(s)))
        return s.as_dict()

    def set_schedule(self, block=256, lanes_per_cell=1, aosoa_width=1, persistent=0, xcd_remap=1, fused=0):
        s = Schedule(block, lanes_per_cell, aosoa_width, persistent, xcd_remap, fused)
        check(load().hcfd_set_schedule(self.h, ctypes.byref(s)))

    def autotune(self, steps: int = 10, cfl: float = 0.0, cache: Optional[str] = None) -> dict:
        s = Schedule()
        check(load().hcfd_autotune(self.h, steps, cfl, cache.encode() if cache else None, ctypes.byref(s)))
        return s.as_dict()

    # ---- batches --------------------------------------------------------------
    @property
    def ninst(self) -> int:
        return len(self.inst_sizes)

    def inst_state(self, k: int) -> np.ndarray:
        U = np.zeros((4, self.inst_sizes[k]))
        check(load().hcfd_batch_get_state(self.h, k, _rows(U)))
        return U

    def inst_prims(self, k: int) -> np.ndarray:
        Q = np.zeros((4, self.inst_sizes[k]))
        check(load().hcfd_batch_get_prims(self.h, k, _rows(Q)))
        return Q

    def set_inst_state(self, k: int, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(load().hcfd_batch_set_state(self.h, k, _rows(U)))

    def status(self):
        n = self.ninst
        t, d, f = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
        check(load().hcfd_batch_get_status(self.h, _ptr(t), _ptr(d), _ptr(f)))
        return t, d, f


def write_hsln(path: str, Q: np.ndarray, names: Optional[Sequence[str]] = None) -> None:
    """Writer::write_soln's file (src/Writer.cpp:22-39); names: the four variable
    names (default Density, Velocity_X, Velocity_Y, Pressure)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    nm = None
    if names is not None:
        if len(names) != 4:
            raise ValueError("four variable names")
        nm = (ctypes.c_char_p * 4)(*[n.encode() for n in names])
    check(load().hcfd_write_hsln(path.encode(), Q.shape[1], _rows(Q), nm))


def read_hsln(path: str) -> np.ndarray:
    """Primitives (4, Nc) of a .hsln file, through the library's reader."""
    n = I(0)
    check(load().hcfd_read_hsln(path.encode(), ctypes.byref(n), None))
    Q = np.zeros((4, n.value))
    check(load().hcfd_read_hsln(path.encode(), ctypes.byref(n), _rows(Q)))
    return Q


def write_vtu(path: str, mesh, Q: np.ndarray) -> None:
    """VTU file of a mesh (NativeMesh, HMesh or MeshView) with primitives Q (4, Nc)."""
    v, keep = as_view(mesh)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    check(load().hcfd_write_vtu(path.encode(), ctypes.byref(v), _rows(Q)))
    del keep


# ---- the reference's NDIMS == 3 branches (arrays of states; include/hcfd.h) ------
def rusanov3d(UL: np.ndarray, UR: np.ndarray, nrm: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    """Flux::Rusanov with the 3-D EulerFlux on the GPU: UL, UR [n][5], nrm [n][3]."""
    UL, UR, nrm = (np.ascontiguousarray(x, dtype=np.float64) for x in (UL, UR, nrm))
    n = UL.shape[0]
    if UL.shape != (n, 5) or UR.shape != (n, 5) or nrm.shape != (n, 3):
        raise ValueError("UL, UR must be [n][5] and nrm [n][3]")
    F = np.zeros((n, 5))
    check(load().hcfd_rusanov3d(n, gamma, _ptr(UL), _ptr(UR), _ptr(nrm), _ptr(F)))
    return F


def prims3d(U: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64)
    Q = np.zeros_like(U)
    check(load().hcfd_prims3d(U.shape[0], gamma, _ptr(U), _ptr(Q)))
    return Q


def cons3d(Q: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    U = np.zeros_like(Q)
    check(load().hcfd_cons3d(Q.shape[0], gamma, _ptr(Q), _ptr(U)))
    return U
