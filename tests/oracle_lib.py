"""ctypes bindings of the CPU references: the C oracle (oracle/hcfd_oracle.c) and,
where it was built, the reference's own sources (oracle/_ref).

Test infrastructure: used only by tests, the smoke entry point and the CPU
baseline leg of the benchmark.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libhcfd_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libhcfd_ref.so")
REF_NOFMA_SO = os.path.join(ROOT, "oracle", "_ref", "libhcfd_ref_nofma.so")
REF_VEC_SO = os.path.join(ROOT, "oracle", "_ref", "libhcfd_ref_vec.so")

P = ctypes.c_void_p
D = ctypes.c_double
I = ctypes.c_int


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else P(a.ctypes.data)


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_dom_create.restype = P
        lib.orc_dom_create.argtypes = [I, I, I, I, I, P, P, P, P, P, P]
        lib.orc_dom_free.argtypes = [P]
        lib.orc_zero_bc_iv.argtypes = [P]
        lib.orc_geometry.argtypes = [P] * 9
        lib.orc_cf.argtypes = [P, P, P]
        lib.orc_gradient.argtypes = [P, P, P]
        lib.orc_limiter.argtypes = [P, P, P, P]
        lib.orc_bc_grads.argtypes = [P, P]
        lib.orc_bc_limiter.argtypes = [P, P]
        lib.orc_flux.argtypes = [P, D, P, P, P, P]
        lib.orc_update.argtypes = [P, D, P, P]
        lib.orc_bc_cells.argtypes = [P, P, P, D, P]
        lib.orc_initialise.argtypes = [P, P, P, P, D, P]
        lib.orc_initialise_prims.argtypes = [P, P, P, P, D, P]
        lib.orc_prims.argtypes = [P, D, P, P]
        lib.orc_cfl_dt.restype = D
        lib.orc_cfl_dt.argtypes = [P, D, D, P]
        lib.orc_run.restype = D
        lib.orc_run.argtypes = [P, P, P, D, P, I, D, D]
        lib.orc_run_ssprk.restype = D
        lib.orc_run_ssprk.argtypes = [P, P, P, D, P, I, D, D, I]
        lib.orc_rusanov.argtypes = [D, P, P, P, P]
        lib.orc_euler_flux.argtypes = [D, P, P, P]
        lib.orc_rusanov3d.argtypes = [I, D, P, P, P, P]
        lib.orc_euler_flux3d.argtypes = [I, D, P, P, P]
        lib.orc_prims3d.argtypes = [I, D, P, P]
        lib.orc_cons3d.argtypes = [I, D, P, P]
        lib.orc_cons_of_prims.argtypes = [D, P, P]
        lib.orc_count_steps.restype = I
        lib.orc_count_steps.argtypes = [D, D]
        _oracle = lib
    return _oracle


class OracleDomain:
    """The oracle's Domain over raw mesh arrays (formats.HMesh)."""

    def __init__(self, m, gamma: float = 1.4, bctypes=None, qbc=None):
        lib = oracle_lib()
        self.lib, self.m, self.gamma = lib, m, gamma
        self._keep = [np.ascontiguousarray(x, dtype=t) for x, t in (
            (m.coloff, np.int32), (m.nodes, np.float64), (m.fn_con, np.int32), (m.fc_con, np.int32),
            (m.cn_off, np.int32), (m.cn_con, np.int32))]
        co, nd, fn, fc, cno, cnc = self._keep
        self.h = lib.orc_dom_create(m.Nn, m.Nc, m.Nf, m.Ncol, m.NBCcol, _ptr(co), _ptr(nd), _ptr(fn), _ptr(fc),
                                    _ptr(cno), _ptr(cnc))
        self.Nc, self.Nf = m.Nc, m.Nf
        nbc = m.NBCcol
        self.bctypes = np.ascontiguousarray(bctypes if bctypes is not None else np.full(nbc, 4), dtype=np.int32)
        self.qbc = np.ascontiguousarray(qbc if qbc is not None else np.tile([1.0, 0, 0, 1.0], (nbc, 1)),
                                        dtype=np.float64).reshape(nbc, 4)

    def __del__(self):
        try:
            self.lib.orc_dom_free(self.h)
        except Exception:
            pass

    def zero_bc_iv(self):
        self.lib.orc_zero_bc_iv(self.h)

    def geometry(self):
        Nf, Nc = self.Nf, self.Nc
        g = dict(C=np.zeros((Nf, 2), np.int32), xf=np.zeros((Nf, 2)), xc=np.zeros((Nc, 2)),
                 rf=np.zeros((Nf, 2, 2)), nf=np.zeros((Nf, 2)), Af=np.zeros(Nf), If=np.zeros(Nf), iv=np.zeros(Nc))
        self.lib.orc_geometry(self.h, *(_ptr(g[k]) for k in ("C", "xf", "xc", "rf", "nf", "Af", "If", "iv")))
        return g

    def cf(self):
        off = np.zeros(self.Nc + 1, np.int32)
        con = np.zeros(2 * self.Nf, np.int32)
        self.lib.orc_cf(self.h, _ptr(off), _ptr(con))
        return off, con

    def gradient(self, U):
        G = np.zeros((4, 2, self.Nc))
        self.lib.orc_gradient(self.h, _ptr(np.ascontiguousarray(U)), _ptr(G))
        return G

    def limiter(self, U, G):
        phi = np.zeros((4, self.Nc))
        self.lib.orc_limiter(self.h, _ptr(np.ascontiguousarray(U)), _ptr(np.ascontiguousarray(G)), _ptr(phi))
        return phi

    def bc_grads(self, G):
        G = np.array(G, copy=True)
        self.lib.orc_bc_grads(self.h, _ptr(G))
        return G

    def bc_limiter(self, phi):
        phi = np.array(phi, copy=True)
        self.lib.orc_bc_limiter(self.h, _ptr(phi))
        return phi

    def flux(self, U, G, phi, R=None):
        R = np.zeros((4, self.Nc)) if R is None else np.array(R, copy=True)
        self.lib.orc_flux(self.h, self.gamma, _ptr(np.ascontiguousarray(U)), _ptr(np.ascontiguousarray(G)),
                          _ptr(np.ascontiguousarray(phi)), _ptr(R))
        return R

    def bc_cells(self, U):
        U = np.array(U, copy=True)
        self.lib.orc_bc_cells(self.h, _ptr(self.bctypes), _ptr(self.qbc), self.gamma, _ptr(U))
        return U

    def initialise(self, QIC):
        U = np.zeros((4, self.Nc))
        q = np.ascontiguousarray(QIC, dtype=np.float64)
        self.lib.orc_initialise(self.h, _ptr(q), _ptr(self.bctypes), _ptr(self.qbc), self.gamma, _ptr(U))
        return U

    def initialise_prims(self, Q):
        U = np.zeros((4, self.Nc))
        q = np.ascontiguousarray(Q, dtype=np.float64)
        self.lib.orc_initialise_prims(self.h, _ptr(q), _ptr(self.bctypes), _ptr(self.qbc), self.gamma, _ptr(U))
        return U

    def prims(self, U):
        Q = np.zeros((4, self.Nc))
        self.lib.orc_prims(self.h, self.gamma, _ptr(np.ascontiguousarray(U)), _ptr(Q))
        return Q

    def cfl_dt(self, U, cfl):
        return self.lib.orc_cfl_dt(self.h, self.gamma, cfl, _ptr(np.ascontiguousarray(U)))

    def run(self, U, nsteps, dt=0.0, cfl=0.0):
        U = np.array(U, copy=True)
        t = self.lib.orc_run(self.h, _ptr(self.bctypes), _ptr(self.qbc), self.gamma, _ptr(U), nsteps, dt, cfl)
        return U, t

    def run_ssprk(self, U, nsteps, stages, dt=0.0, cfl=0.0):
        """nsteps SSP-RK steps of 1, 2 or 3 stages (1 = forward Euler)."""
        U = np.array(U, copy=True)
        t = self.lib.orc_run_ssprk(self.h, _ptr(self.bctypes), _ptr(self.qbc), self.gamma, _ptr(U), nsteps, dt,
                                   cfl, stages)
        if t < 0:
            raise ValueError("unsupported stage count %d" % stages)
        return U, t


def ref_available(path: str = REF_SO) -> bool:
    return os.path.exists(path)


class VecRefCase:
    """The reference's VEC layout (oracle/_ref, AVX2 build), driven through
    ref_vec_driver.cpp: its fastest valid CPU layout, used as the CPU baseline.
    The state is [4][Nc]; ghost 1/V is zeroed at creation as Euler does."""

    def __init__(self, hmsh_path: str, gamma: float, bcmap_by_coltag, qbc, so: str = REF_VEC_SO):
        lib = ctypes.CDLL(so)
        lib.vref_create.restype = P
        lib.vref_create.argtypes = [ctypes.c_char_p, D, P, P]
        lib.vref_destroy.argtypes = [P]
        lib.vref_run.argtypes = [P, P, I, D]
        lib.vref_veclen.restype = I
        lib.vref_prims.argtypes = [P, D, P, P]
        self.lib, self.gamma = lib, gamma
        self._bcmap = np.ascontiguousarray(bcmap_by_coltag, dtype=np.int32)
        self._qbc = np.ascontiguousarray(qbc, dtype=np.float64)
        self.h = lib.vref_create(hmsh_path.encode(), gamma, _ptr(self._bcmap), _ptr(self._qbc))
        self.veclen = lib.vref_veclen()

    def run(self, U, nsteps: int, dt: float):
        U = np.array(U, dtype=np.float64, copy=True, order="C")
        self.lib.vref_run(self.h, _ptr(U), nsteps, dt)
        return U

    def prims(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        Q = np.zeros_like(U)
        self.lib.vref_prims(self.h, self.gamma, _ptr(U), _ptr(Q))
        return Q

    def close(self):
        if self.h:
            self.lib.vref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefCase:
    """The reference's own Base classes (oracle/_ref), driven through ref_driver.cpp."""

    def __init__(self, hmsh_path: str, gamma: float, bcmap_by_coltag, qbc, so: str = REF_SO):
        lib = ctypes.CDLL(so)
        lib.ref_create.restype = P
        lib.ref_create.argtypes = [ctypes.c_char_p, D, P, P]
        for name, nargs in (("ref_destroy", 1), ("ref_sizes", 2), ("ref_geometry", 9), ("ref_cf", 3),
                            ("ref_zero_bc_iv", 1), ("ref_gradient", 3), ("ref_limiter", 4), ("ref_bc_grads", 2),
                            ("ref_bc_limiter", 2), ("ref_flux", 5), ("ref_bc_cells", 2)):
            getattr(lib, name).argtypes = [P] * nargs
        lib.ref_initialise.argtypes = [P, P, D, P]
        lib.ref_run.argtypes = [P, P, I, D]
        lib.ref_prims.argtypes = [P, D, P, P]
        self.lib, self.gamma = lib, gamma
        self._bcmap = np.ascontiguousarray(bcmap_by_coltag, dtype=np.int32)
        self._qbc = np.ascontiguousarray(qbc, dtype=np.float64)
        self.h = lib.ref_create(hmsh_path.encode(), gamma, _ptr(self._bcmap), _ptr(self._qbc))
        sz = np.zeros(4, np.int32)
        lib.ref_sizes(self.h, _ptr(sz))
        self.Nc, self.Nf = int(sz[0]), int(sz[1])

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        Nf, Nc = self.Nf, self.Nc
        g = dict(C=np.zeros((Nf, 2), np.int32), xf=np.zeros((Nf, 2)), xc=np.zeros((Nc, 2)),
                 rf=np.zeros((Nf, 2, 2)), nf=np.zeros((Nf, 2)), Af=np.zeros(Nf), If=np.zeros(Nf), iv=np.zeros(Nc))
        self.lib.ref_geometry(self.h, *(_ptr(g[k]) for k in ("C", "xf", "xc", "rf", "nf", "Af", "If", "iv")))
        return g

    def cf(self):
        off = np.zeros(self.Nc + 1, np.int32)
        con = np.zeros(2 * self.Nf, np.int32)
        self.lib.ref_cf(self.h, _ptr(off), _ptr(con))
        return off, con

    def zero_bc_iv(self):
        self.lib.ref_zero_bc_iv(self.h)

    def gradient(self, U):
        G = np.zeros((4, 2, self.Nc))
        self.lib.ref_gradient(self.h, _ptr(np.ascontiguousarray(U)), _ptr(G))
        return G

    def limiter(self, U, G):
        phi = np.zeros((4, self.Nc))
        self.lib.ref_limiter(self.h, _ptr(np.ascontiguousarray(U)), _ptr(np.ascontiguousarray(G)), _ptr(phi))
        return phi

    def bc_grads(self, G):
        G = np.array(G, copy=True)
        self.lib.ref_bc_grads(self.h, _ptr(G))
        return G

    def bc_limiter(self, phi):
        phi = np.array(phi, copy=True)
        self.lib.ref_bc_limiter(self.h, _ptr(phi))
        return phi

    def flux(self, U, G, phi, R=None):
        R = np.zeros((4, self.Nc)) if R is None else np.array(R, copy=True)
        self.lib.ref_flux(self.h, _ptr(np.ascontiguousarray(U)), _ptr(np.ascontiguousarray(G)),
                          _ptr(np.ascontiguousarray(phi)), _ptr(R))
        return R

    def bc_cells(self, U):
        U = np.array(U, copy=True)
        self.lib.ref_bc_cells(self.h, _ptr(U))
        return U

    def initialise(self, QIC):
        U = np.zeros((4, self.Nc))
        q = np.ascontiguousarray(QIC, dtype=np.float64)
        self.lib.ref_initialise(self.h, _ptr(q), self.gamma, _ptr(U))
        return U

    def run(self, U, nsteps, dt):
        U = np.array(U, copy=True)
        self.lib.ref_run(self.h, _ptr(U), nsteps, dt)
        return U

    def prims(self, U):
        Q = np.zeros((4, self.Nc))
        self.lib.ref_prims(self.h, self.gamma, _ptr(np.ascontiguousarray(U)), _ptr(Q))
        return Q
