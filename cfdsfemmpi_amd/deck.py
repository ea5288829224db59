"""Mesh, control and solver-option decks in the reference's text formats.

Formats follow cfdsfemmpi's readers:

* mesh ``<prefix>.dat`` -- ``StabFEM::readInputData`` (src/base/StabFEM.cpp:80-264):
  seven ``key value`` header lines (ndim, npElem, nNode, nElem, nDBC, nFBC,
  nOutNodes), then ``nodes`` (id x y [z]), ``elements`` (id t1 t2 t3 n1..nk,
  1-based node ids), ``prescribed boundary conditions`` (node dof value,
  1-based), ``outputnodes`` (node, 1-based).  Whitespace-separated tokens are
  read positionally; the section keywords are skipped as single tokens (the
  DBC header is three tokens, :209).
* control file -- ``StabFEM::readControlParameters`` (StabFEM.cpp:268-322):
  eleven positional ``key value`` lines; keys are ignored.
* solver options -- the PETSc options files the reference passes to
  ``PetscInitialize`` (src/cfdsfemmpi.cpp:40, input/petsc_options*.dat).
"""
from __future__ import annotations

import dataclasses
import gzip
import io
import os
from typing import Optional

import warnings

import numpy as np


@dataclasses.dataclass
class Mesh:
    ndim: int
    npElem: int
    coords: np.ndarray            # (nNode, ndim) float64
    conn: np.ndarray              # (nElem, npElem) int32, 0-based
    dbc_node: np.ndarray          # (nDBC,) int32, 0-based, in file order
    dbc_dof: np.ndarray           # (nDBC,) int32, 0-based
    dbc_value: np.ndarray         # (nDBC,) float64
    out_nodes: np.ndarray         # (nOut,) int32, 0-based
    nFBC: int = 0

    @property
    def nNode(self) -> int:
        return int(self.coords.shape[0])

    @property
    def nElem(self) -> int:
        return int(self.conn.shape[0])

    @property
    def nDBC(self) -> int:
        return int(self.dbc_node.shape[0])

    @property
    def ndof(self) -> int:
        return self.ndim + 1


@dataclasses.dataclass
class Control:
    elemData: np.ndarray          # 11 entries: 0 rho, 1 mu, 2-4 body force, 8-10 SUPG/PSPG/LSIC
    tis: int
    rhoInf: float
    dt: float
    timeFinal: float
    stepsMax: int
    outputFreq: int
    conv_tol: float


@dataclasses.dataclass
class SolverOptions:
    """Krylov / preconditioner choice in the vocabulary of the PETSc options file."""
    ksp_type: str = "bcgs"        # bcgs | gmres | preonly
    pc_type: str = "ilu"          # none | jacobi | bjacobi4 | ilu | lu
    rtol: float = 1e-8
    max_it: int = 1000
    restart: int = 30


def _open_text(path: str):
    if path.endswith(".gz"):
        return io.TextIOWrapper(gzip.open(path, "rb"), encoding="ascii")
    return open(path, "r", encoding="ascii")


def read_mesh(path: str) -> Mesh:
    """Read a reference mesh deck.  ``path`` may omit ``.dat`` (the reference
    appends it, StabFEM.cpp:84) and may be gzip-compressed."""
    if not os.path.exists(path):
        for cand in (path + ".dat", path + ".dat.gz"):
            if os.path.exists(cand):
                path = cand
                break
    with _open_text(path) as f:
        tok = f.read().split()
    pos = 0

    def take(n=1):
        nonlocal pos
        v = tok[pos:pos + n]
        pos += n
        return v

    hdr = {}
    for _ in range(7):
        k, v = take(2)
        hdr[len(hdr)] = int(v)
    ndim, npElem, nNode, nElem, nDBC, nFBC, nOut = (hdr[i] for i in range(7))
    take(1)  # "nodes"
    rec = np.array(take(nNode * (1 + ndim)), dtype=np.float64).reshape(nNode, 1 + ndim)
    coords = np.ascontiguousarray(rec[:, 1:])
    take(1)  # "elements"
    if npElem not in (3, 4, 8):
        raise ValueError(f"invalid npElem {npElem}")
    rec = np.array(take(nElem * (4 + npElem)), dtype=np.int64).reshape(nElem, 4 + npElem)
    conn = np.ascontiguousarray(rec[:, 4:] - 1).astype(np.int32)
    take(3)  # "prescribed boundary conditions"
    rec = take(nDBC * 3)
    dnode = np.array(rec[0::3], dtype=np.int64) - 1
    ddof = np.array(rec[1::3], dtype=np.int64) - 1
    dval = np.array(rec[2::3], dtype=np.float64)
    take(1)  # "outputnodes"
    outn = np.array(take(nOut), dtype=np.int64) - 1
    return Mesh(ndim, npElem, coords, conn, dnode.astype(np.int32), ddof.astype(np.int32),
                dval, outn.astype(np.int32), nFBC)


def write_mesh(path: str, m: Mesh) -> None:
    """Write a mesh deck in the reference layout (inverse of :func:`read_mesh`)."""
    with open(path, "w", encoding="ascii") as f:
        f.write(f"ndim       {m.ndim}\nnpElem     {m.npElem}\nnNode      {m.nNode}\n"
                f"nElem      {m.nElem}\nnDBC       {m.nDBC}\nnFBC       {m.nFBC}\n"
                f"nOutNodes  {len(m.out_nodes)}\nnodes\n")
        for i in range(m.nNode):
            f.write(f"{i + 1}\t" + "\t".join(f"{x:.17g}" for x in m.coords[i]) + "\n")
        f.write("elements\n")
        for e in range(m.nElem):
            f.write(f"{e + 1}\t1\t1\t1\t" + "\t".join(str(int(n) + 1) for n in m.conn[e]) + "\n")
        f.write("prescribed boundary conditions\n")
        for k in range(m.nDBC):
            f.write(f"{int(m.dbc_node[k]) + 1}\t{int(m.dbc_dof[k]) + 1}\t{m.dbc_value[k]:.17g}\n")
        f.write("outputnodes\n")
        for n in m.out_nodes:
            f.write(f"{int(n) + 1}\n")


def read_control(path: str) -> Control:
    with _open_text(path) as f:
        tok = f.read().split()
    ed = np.zeros(11)
    p = 0
    ed[0] = float(tok[p + 1]); p += 2                       # density
    ed[1] = float(tok[p + 1]); p += 2                       # viscosity
    ed[2:5] = [float(x) for x in tok[p + 1:p + 4]]; p += 4  # body force
    ed[8:11] = [float(x) for x in tok[p + 1:p + 4]]; p += 4  # SUPG PSPG LSIC
    tis = int(tok[p + 1]); p += 2
    rhoInf = float(tok[p + 1]); p += 2
    dt = float(tok[p + 1]); p += 2
    timeFinal = float(tok[p + 1]); p += 2
    stepsMax = int(tok[p + 1]); p += 2
    outputFreq = int(tok[p + 1]); p += 2
    conv_tol = float(tok[p + 1]); p += 2
    return Control(ed, tis, rhoInf, dt, timeFinal, stepsMax, outputFreq, conv_tol)


def write_control(path: str, c: Control) -> None:
    ed = [repr(float(x)) for x in c.elemData]
    r = lambda x: repr(float(x))
    with open(path, "w", encoding="ascii") as f:
        f.write(f"density           {ed[0]}\nviscosity         {ed[1]}\n"
                f"bodyforce         {ed[2]} {ed[3]} {ed[4]}\n"
                f"stabilisation     {ed[8]} {ed[9]} {ed[10]}\n"
                f"tis               {c.tis}\nrhoInf            {r(c.rhoInf)}\n"
                f"dt                {r(c.dt)}\nFinalTime         {r(c.timeFinal)}\n"
                f"MaximumSteps      {c.stepsMax}\nOutputFrequency   {c.outputFreq}\n"
                f"tolerance         {r(c.conv_tol)}\n")


def read_solver_options(path: str) -> SolverOptions:
    """Map a PETSc options file onto :class:`SolverOptions`.

    ``-pc_type bjacobi`` with ``-sub_pc_type ilu`` (input/petsc_options_bcgs.dat)
    becomes ILU(0) on the single block a one-GPU run has; fill levels other than 0
    (``-sub_pc_factor_levels`` / ``-pc_factor_levels``) are not built, and a warning says
    that ILU(0) runs instead.  ``-pc_type lu`` (the test decks) is kept as "lu"; the
    driver maps it onto the device direct solver.
    """
    opts = {}
    with _open_text(path) as f:
        tok = f.read().split()
    i = 0
    while i < len(tok):
        if tok[i].startswith("-"):
            key = tok[i][1:]
            if i + 1 < len(tok) and not tok[i + 1].startswith("-"):
                opts[key] = tok[i + 1]
                i += 2
                continue
            opts[key] = ""
        i += 1
    so = SolverOptions()
    ksp = opts.get("ksp_type", "gmres")
    so.ksp_type = {"bcgs": "bcgs", "bicg": "bcgs", "gmres": "gmres", "preonly": "preonly",
                   "fgmres": "gmres"}.get(ksp, ksp)
    pc = opts.get("pc_type", "ilu")
    if pc == "bjacobi":
        sub = opts.get("sub_pc_type", "ilu")
        pc = {"ilu": "ilu", "jacobi": "jacobi", "none": "none"}.get(sub, sub)
    so.pc_type = pc
    for key in ("sub_pc_factor_levels", "pc_factor_levels"):
        if key in opts and opts[key] not in ("", "0"):
            warnings.warn(f"{path}: -{key} {opts[key]} asks for ILU({opts[key]}); ILU(0) runs instead",
                          stacklevel=2)
    if "ksp_rtol" in opts:
        so.rtol = float(opts["ksp_rtol"])
    if "ksp_max_it" in opts:
        so.max_it = int(opts["ksp_max_it"])
    if "ksp_gmres_restart" in opts:
        so.restart = int(opts["ksp_gmres_restart"])
    return so


def read_result(path: str) -> np.ndarray:
    """Reference result file (``StabFEM::writeResult``, StabFEM.cpp:457-494):
    node count, then one row of ndof values per node."""
    with _open_text(path) as f:
        tok = f.read().split()
    n = int(tok[0])
    vals = np.array(tok[1:], dtype=np.float64)
    return vals.reshape(n, -1)


def write_result(path: str, soln: np.ndarray, ndof: int) -> None:
    s = np.asarray(soln).reshape(-1, ndof)
    with open(path, "w", encoding="ascii") as f:
        f.write(f"{s.shape[0]}\n")
        for row in s:
            f.write("\t".join(f"{v:.8f}" for v in row) + "\n")


def check_result(soln: np.ndarray, ref: np.ndarray, ndim: int, tol: float = 1.0e-4) -> int:
    """``StabFEM::checkResult`` (StabFEM.cpp:501-554): 0 on match, -1 on a node
    count mismatch, -2 if any node's ||(du, dv, dp)||_2 exceeds ``tol`` (2D).
    Values are compared in 3D as well (the reference reads but skips them)."""
    ndof = ndim + 1
    s = np.asarray(soln).reshape(-1, ndof)
    if ref.shape[0] != s.shape[0]:
        return -1
    d = np.linalg.norm(ref[:, :ndof] - s, axis=1)
    return 0 if np.all(d <= tol) else -2
