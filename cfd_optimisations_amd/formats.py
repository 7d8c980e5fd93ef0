"""Readers and writers for the solver's on-disk formats, in numpy.

``.hmsh`` (mesh) follows the layout the reference's C++ reader consumes
(src/Mesh.cpp:11-45) and its Python writer produces (interface/Mesh.py:441-463):

    int32[8]   dim, Nn, Ncol, NBCcol, Nc, Nf, N_pnames, Nchars
    char[Nchars]              NUL-separated physical names
    int32[Ncol]               coltag
    int32[Ncol+1]             coloff  (face offsets per colour; colour 0 interior)
    float64[Nn*dim]           nodes (x, y interleaved)
    int32[2*Nf]               fn_con  (2 nodes per face)
    int32[2*Nf]               fc_con  (CL, CR per face; BC faces have the ghost as CR)
    int32[Nc+1], int32[...]   cn_off, cn_con
    int32[Nc+1], int32[...]   cc_off, cc_con
    int32[Nc]                 c_ctypes (gmsh element types; optional trailer)

``.hsln`` (solution) follows Writer::write_soln (src/Writer.cpp:22-39):

    int32 Nc, int32 nvars, int32 nchars; char[nchars] names; float64[nvars][Nc]
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional

import numpy as np

VARNAMES = ("Density", "Velocity_X", "Velocity_Y", "Pressure")


@dataclasses.dataclass
class HMesh:
    """Raw mesh arrays, as in a ``.hmsh`` file (Mesh members, includes/Mesh.hpp:4-22)."""
    dim: int
    Nn: int
    Ncol: int
    NBCcol: int
    Nc: int
    Nf: int
    pnames: List[str]
    coltag: np.ndarray
    coloff: np.ndarray
    nodes: np.ndarray      # (Nn, 2) float64
    fn_con: np.ndarray     # (Nf, 2) int32
    fc_con: np.ndarray     # (Nf, 2) int32
    cn_off: np.ndarray
    cn_con: np.ndarray
    cc_off: Optional[np.ndarray] = None
    cc_con: Optional[np.ndarray] = None
    c_ctypes: Optional[np.ndarray] = None

    @property
    def F_BC(self) -> int:
        """First boundary face (Domain::compute_rf, src/Base/Domain.cpp:74)."""
        return int(self.coloff[self.Ncol - self.NBCcol])

    def ghost_mask(self) -> np.ndarray:
        m = np.zeros(self.Nc, dtype=bool)
        m[self.fc_con[self.F_BC:, 1]] = True
        return m


def read_hmsh(path: str) -> HMesh:
    with open(path, "rb") as fh:
        raw = fh.read()
    pos = 0

    def take(dtype, count):
        nonlocal pos
        dt = np.dtype(dtype)
        nbytes = dt.itemsize * int(count)
        if pos + nbytes > len(raw):
            raise ValueError(f"{path}: truncated .hmsh (need {nbytes} bytes at offset {pos}, have {len(raw) - pos})")
        arr = np.frombuffer(raw, dtype=dt, count=int(count), offset=pos).copy()
        pos += nbytes
        return arr

    hdr = take("<i4", 8)
    dim, Nn, Ncol, NBCcol, Nc, Nf, Npn, Nch = (int(x) for x in hdr)
    if dim != 2 or min(Nn, Ncol, Nc, Nf) <= 0 or NBCcol < 0 or NBCcol > Ncol or Nch < 0:
        raise ValueError(f"{path}: bad .hmsh header {hdr.tolist()}")
    names = take("u1", Nch).tobytes().split(b"\0")
    pnames = [n.decode("utf-8", "replace") for n in names if n]
    coltag = take("<i4", Ncol)
    coloff = take("<i4", Ncol + 1)
    nodes = take("<f8", Nn * dim).reshape(Nn, dim)
    fn_con = take("<i4", 2 * Nf).reshape(Nf, 2)
    fc_con = take("<i4", 2 * Nf).reshape(Nf, 2)
    cn_off = take("<i4", Nc + 1)
    cn_con = take("<i4", cn_off[-1])
    cc_off = take("<i4", Nc + 1)
    cc_con = take("<i4", cc_off[-1])
    c_ctypes = take("<i4", Nc) if pos + 4 * Nc <= len(raw) else None
    return HMesh(dim, Nn, Ncol, NBCcol, Nc, Nf, pnames, coltag, coloff, nodes, fn_con, fc_con,
                 cn_off, cn_con, cc_off, cc_con, c_ctypes)


def build_cc(m: HMesh):
    """Cell->cell CSR of face neighbours (cells sharing a face)."""
    cl, cr = m.fc_con[:, 0].astype(np.int64), m.fc_con[:, 1].astype(np.int64)
    rows = np.concatenate([cl, cr])
    cols = np.concatenate([cr, cl])
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    off = np.zeros(m.Nc + 1, dtype=np.int32)
    np.add.at(off, rows + 1, 1)
    return np.cumsum(off, dtype=np.int64).astype(np.int32), cols.astype(np.int32)


def write_hmsh(path: str, m: HMesh) -> None:
    if m.cc_off is None or m.cc_con is None:
        m.cc_off, m.cc_con = build_cc(m)
    names = b"".join(s.encode() + b"\0" for s in m.pnames)
    hdr = np.array([2, m.Nn, m.Ncol, m.NBCcol, m.Nc, m.Nf, len(m.pnames), len(names)], dtype="<i4")
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(names)
        for arr, dt in ((m.coltag, "<i4"), (m.coloff, "<i4"), (m.nodes, "<f8"), (m.fn_con, "<i4"),
                        (m.fc_con, "<i4"), (m.cn_off, "<i4"), (m.cn_con, "<i4"), (m.cc_off, "<i4"),
                        (m.cc_con, "<i4")):
            fh.write(np.ascontiguousarray(arr, dtype=dt).tobytes())
        ct = m.c_ctypes if m.c_ctypes is not None else np.zeros(m.Nc, dtype="<i4")
        fh.write(np.ascontiguousarray(ct, dtype="<i4").tobytes())


def renumber_cells(m: HMesh, new_of_old: np.ndarray) -> HMesh:
    """The same mesh with cell i renamed new_of_old[i]; faces keep their
    orientation and colour and are sorted, within each colour, by their lower
    new cell id (the importer's face order, interface/Mesh.py:321-378)."""
    p = np.asarray(new_of_old, dtype=np.int64)
    old_of_new = np.empty_like(p)
    old_of_new[p] = np.arange(m.Nc)
    nn = np.diff(m.cn_off)[old_of_new]
    cn_off = np.zeros(m.Nc + 1, np.int64)
    np.cumsum(nn, out=cn_off[1:])
    starts = m.cn_off[:-1][old_of_new]
    idx = np.repeat(starts - cn_off[:-1], nn) + np.arange(cn_off[-1])
    cn_con = m.cn_con[idx]
    fc = p[m.fc_con]
    order = []
    for c in range(m.Ncol):
        f0, f1 = int(m.coloff[c]), int(m.coloff[c + 1])
        lo = fc[f0:f1].min(axis=1)
        order.append(f0 + np.lexsort((np.arange(f1 - f0), lo)))
    order = np.concatenate(order) if order else np.zeros(0, np.int64)
    ct = m.c_ctypes[old_of_new] if m.c_ctypes is not None else None
    return HMesh(m.dim, m.Nn, m.Ncol, m.NBCcol, m.Nc, m.Nf, list(m.pnames), m.coltag.copy(), m.coloff.copy(),
                 m.nodes.copy(), m.fn_con[order].copy(), fc[order].astype(np.int32), cn_off.astype(np.int32),
                 cn_con.astype(np.int32), None, None, ct)


def read_hsln(path: str):
    """Returns (names, Q[nvars][Nc])."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if len(raw) < 12:
        raise ValueError(f"{path}: truncated .hsln")
    Nc, nv, nch = (int(x) for x in np.frombuffer(raw, "<i4", 3))
    names = [n.decode() for n in raw[12:12 + nch].split(b"\0") if n]
    need = 12 + nch + 8 * nv * Nc
    if len(raw) != need:
        raise ValueError(f"{path}: size {len(raw)} != expected {need}")
    Q = np.frombuffer(raw, "<f8", nv * Nc, 12 + nch).reshape(nv, Nc).copy()
    return names, Q


def write_hsln(path: str, Q: np.ndarray, names=VARNAMES) -> None:
    """Byte layout of Writer::write_soln (src/Writer.cpp:27-35)."""
    Q = np.ascontiguousarray(Q, dtype="<f8")
    nv, Nc = Q.shape
    nm = b"".join(s.encode() + b"\0" for s in names[:nv])
    with open(path, "wb") as fh:
        fh.write(np.array([Nc, nv, len(nm)], "<i4").tobytes())
        fh.write(nm)
        fh.write(Q.tobytes())


def hsln_name(fname: str, runcount: int) -> str:
    """``<fname>_<runcount>.hsln`` as Writer::write_soln names it (src/Writer.cpp:23)."""
    return f"{fname}_{runcount}.hsln"


def file_size(path: str) -> int:
    return os.stat(path).st_size
