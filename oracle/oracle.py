"""ctypes front end of liboracle.so and the restated time/Newton loop.

TEST INFRASTRUCTURE ONLY (see the package docstring).

``OracleFEM`` mirrors ``StabFEM`` (src/base/StabFEM.h:35-154) on one rank:
``prepare`` = prepareInputData + setSolver/prepareMatrixPattern
(StabFEM.cpp:328-386, StabFEM2.cpp:10-319), ``solve_fully_implicit`` =
``solveFullyImplicit`` (StabFEM2.cpp:593-832).  The direct solve ("lu") stands
in for MUMPS (SuperLU through scipy: same answer to rounding); "bcgs"/"gmres"
use the restated Krylov solvers in liboracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
TRIA3, QUAD4, TETRA4, HEXA8 = 0, 1, 2, 3


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.orc_elem_type.argtypes = [C.c_int, C.c_int]
    L.orc_elem_ngp.argtypes = [C.c_int]
    L.orc_time_params.argtypes = [C.c_int, C.c_double, C.c_double, f64p]
    L.orc_gauss.argtypes = [C.c_int, f64p, f64p]
    L.orc_shape.argtypes = [C.c_int, f64p, f64p, f64p, f64p, f64p]
    L.orc_shape.restype = None
    L.orc_prepare.argtypes = [C.c_int, C.c_int64, f64p, i32p, f64p, f64p, f64p, f64p, f64p, f64p]
    L.orc_prepare.restype = C.c_int64
    L.orc_element_KF.argtypes = [C.c_int, C.c_int64, i32p, f64p, f64p, f64p, f64p, f64p, f64p,
                                 f64p, f64p, f64p, f64p, f64p, f64p, f64p]
    L.orc_element_KF.restype = None
    L.orc_dof_numbering.argtypes = [C.c_int64, C.c_int, C.c_int64, i32p, i32p, i32p, i32p]
    L.orc_dof_numbering.restype = C.c_int64
    L.orc_pattern.argtypes = [C.c_int64, C.c_int, C.c_int, i32p, i32p, C.c_int64, i64p, C.c_void_p]
    L.orc_pattern.restype = C.c_int64
    L.orc_time_update.argtypes = [C.c_int64] + [f64p] * 7
    L.orc_time_update.restype = None
    L.orc_update_iter_step.argtypes = [C.c_int64] + [f64p] * 10
    L.orc_update_iter_step.restype = None
    L.orc_assemble.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int, i32p, f64p, f64p, f64p, f64p,
                               f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, i32p, i64p, i32p,
                               C.c_int64, C.c_void_p, f64p, C.c_int]
    L.orc_assemble.restype = None
    L.orc_spmv.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p]
    L.orc_spmv.restype = None
    L.orc_pc_create.argtypes = [C.c_int, C.c_int64, i64p, i32p, f64p, C.c_int64, C.c_void_p,
                                C.POINTER(C.c_int)]
    L.orc_pc_create.restype = C.c_void_p
    L.orc_pc_apply.argtypes = [C.c_void_p, f64p, f64p]
    L.orc_pc_apply.restype = None
    L.orc_pc_lu.argtypes = [C.c_void_p, f64p]
    L.orc_pc_lu.restype = C.c_int64
    L.orc_pc_destroy.argtypes = [C.c_void_p]
    L.orc_pc_destroy.restype = None
    L.orc_bicgstab.argtypes = [C.c_int64, i64p, i32p, f64p, C.c_void_p, f64p, f64p, C.c_double,
                               C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.orc_gmres.argtypes = [C.c_int64, i64p, i32p, f64p, C.c_void_p, f64p, f64p, C.c_int,
                            C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = None
    L.orc_get_threads.argtypes = []
    L.orc_get_threads.restype = C.c_int
    L.orc_flops_per_element.argtypes = [C.c_int, C.c_int]
    L.orc_flops_per_element.restype = C.c_int64
    _LIB = L
    return L


def set_threads(n: int) -> None:
    """Host threads of the oracle's parallel paths (1 = sequential, the default)."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def time_params(tis: int, rhoInf: float, dt: float) -> np.ndarray:
    td = np.zeros(16)
    lib().orc_time_params(tis, rhoInf, dt, td)
    return td


def gauss(etype: int):
    gp = np.zeros(24)
    gw = np.zeros(8)
    n = lib().orc_gauss(etype, gp, gw)
    return gp[:3 * n].reshape(n, 3), gw[:n]


def shape(etype: int, xi):
    x = np.zeros(3)
    x[:len(xi)] = xi
    N, d1, d2, d3 = (np.zeros(8) for _ in range(4))
    lib().orc_shape(etype, x, N, d1, d2, d3)
    return N, d1, d2, d3


def prepare(etype: int, coords: np.ndarray, conn: np.ndarray):
    ngp = lib().orc_elem_ngp(etype)
    ne, npe = conn.shape
    nl = npe
    N = np.zeros(ne * ngp * nl); dx = np.zeros_like(N); dy = np.zeros_like(N); dz = np.zeros_like(N)
    dv = np.zeros(ne * ngp); cl = np.zeros(ne)
    rc = lib().orc_prepare(etype, ne, np.ascontiguousarray(coords, np.float64),
                           np.ascontiguousarray(conn, np.int32), N, dx, dy, dz, dv, cl)
    if rc != 0:
        raise ValueError(f"negative Jacobian in element {-rc - 1}")
    return dict(N=N, dNdx=dx, dNdy=dy, dNdz=dz, dvol=dv, charlen=cl)


def element_KF(etype, e, conn, prep, elemData, td, solnPrev, solnCur, solnDotCur):
    npe = conn.shape[1]
    ndof = 3 if etype in (TRIA3, QUAD4) else 4
    ns = npe * ndof
    K = np.zeros(ns * ns)
    F = np.zeros(ns)
    lib().orc_element_KF(etype, e, np.ascontiguousarray(conn, np.int32), prep["N"], prep["dNdx"],
                         prep["dNdy"], prep["dNdz"], prep["dvol"], prep["charlen"],
                         np.ascontiguousarray(elemData, np.float64), td,
                         np.ascontiguousarray(solnPrev), np.ascontiguousarray(solnCur),
                         np.ascontiguousarray(solnDotCur), K, F)
    return K.reshape(ns, ns), F


def dof_numbering(nNode, ndof, dbc_node, dbc_dof):
    nodeDof = np.zeros(nNode * ndof, np.int32)
    assy = np.zeros(nNode * ndof, np.int32)
    nf = lib().orc_dof_numbering(nNode, ndof, len(dbc_node), np.ascontiguousarray(dbc_node, np.int32),
                                 np.ascontiguousarray(dbc_dof, np.int32), nodeDof, assy)
    return nodeDof, assy[:nf].copy(), int(nf)


def pattern(conn, ndof, nodeDof, nfree):
    conn = np.ascontiguousarray(conn, np.int32)
    rowptr = np.zeros(nfree + 1, np.int64)
    nnz = lib().orc_pattern(conn.shape[0], conn.shape[1], ndof, conn, nodeDof, nfree, rowptr, None)
    colidx = np.zeros(max(nnz, 1), np.int32)
    lib().orc_pattern(conn.shape[0], conn.shape[1], ndof, conn, nodeDof, nfree, rowptr,
                      colidx.ctypes.data_as(C.c_void_p))
    return rowptr, colidx[:nnz]


def flops_per_element(etype: int, want_K: bool = True) -> int:
    return int(lib().orc_flops_per_element(etype, 1 if want_K else 0))


PC_TYPES = {"none": 0, "jacobi": 1, "bjacobi4": 2, "ilu": 3}


class OraclePC:
    def __init__(self, pc_type, rowptr, colidx, vals, blk_ptr=None):
        st = C.c_int(0)
        bp = None
        nblk = 0
        if blk_ptr is not None:
            self._bp = np.ascontiguousarray(blk_ptr, np.int64)
            bp = self._bp.ctypes.data_as(C.c_void_p)
            nblk = len(blk_ptr) - 1
        self.h = lib().orc_pc_create(PC_TYPES[pc_type], len(rowptr) - 1, rowptr, colidx, vals, nblk,
                                     bp, C.byref(st))
        self.status = st.value

    def lu(self, nnz):
        """ILU(0) factor values on the CSR pattern (L strictly lower, U upper)."""
        out = np.zeros(nnz)
        if lib().orc_pc_lu(self.h, out) != nnz:
            raise ValueError("not an ILU(0) preconditioner")
        return out

    def apply(self, r):
        z = np.zeros_like(r)
        lib().orc_pc_apply(self.h, np.ascontiguousarray(r), z)
        return z

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pc_destroy(self.h)
            self.h = None


def krylov(ksp, pc, rowptr, colidx, vals, b, rtol, maxit, restart=30):
    x = np.zeros_like(b)
    it = C.c_int(0)
    rr = C.c_double(0.0)
    if ksp == "bcgs":
        rc = lib().orc_bicgstab(len(b), rowptr, colidx, vals, pc.h, b, x, rtol, maxit, C.byref(it),
                                C.byref(rr))
    else:
        rc = lib().orc_gmres(len(b), rowptr, colidx, vals, pc.h, b, x, restart, rtol, maxit,
                             C.byref(it), C.byref(rr))
    return x, rc, it.value, rr.value


def node_blocks(nodeDof: np.ndarray, ndof: int) -> np.ndarray:
    """Row ranges of the free DOFs of each node that has any (nodal blocks)."""
    nd = nodeDof.reshape(-1, ndof)
    cnt = (nd >= 0).sum(axis=1)
    cnt = cnt[cnt > 0]
    return np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)


class OracleFEM:
    """Restated ``StabFEM`` on one rank (see module docstring)."""

    def __init__(self, mesh, control, solver=None):
        self.mesh = mesh
        self.ctl = control
        self.etype = lib().orc_elem_type(mesh.ndim, mesh.npElem)
        if self.etype < 0:
            raise ValueError("unsupported element")
        self.ndof = mesh.ndim + 1
        self.solver = solver
        self.conn = np.ascontiguousarray(mesh.conn, np.int32)
        n = mesh.nNode * self.ndof
        z = lambda: np.zeros(n)
        self.soln, self.solnPrev, self.solnPrev2, self.solnPrev3, self.solnPrev4 = z(), z(), z(), z(), z()
        self.solnCur, self.solnDot, self.solnDotPrev, self.solnDotCur, self.solnApplied = z(), z(), z(), z(), z()
        self.td = np.zeros(16)
        self.log = []

    def prepare(self):
        m = self.mesh
        self.prep = prepare(self.etype, m.coords, self.conn)
        self.nodeDof, self.assyForSoln, self.nfree = dof_numbering(m.nNode, self.ndof, m.dbc_node, m.dbc_dof)
        self.rowptr, self.colidx = pattern(self.conn, self.ndof, self.nodeDof, self.nfree)
        self.vals = np.zeros(len(self.colidx))
        self.rhs = np.zeros(self.nfree)

    def assign_boundary_conditions(self, timeFact):
        """StabFEM::assignBoundaryConditions (StabFEM.cpp:391-410)"""
        m = self.mesh
        for k in range(m.nDBC):
            ind = int(m.dbc_node[k]) * self.ndof + int(m.dbc_dof[k])
            self.solnApplied[ind] = m.dbc_value[k] * timeFact - self.soln[ind]

    def update_iter_step(self):
        lib().orc_update_iter_step(len(self.soln), self.td, self.soln, self.solnPrev, self.solnPrev2,
                                   self.solnPrev3, self.solnPrev4, self.solnDot, self.solnDotPrev,
                                   self.solnCur, self.solnDotCur)

    def time_update(self):
        lib().orc_time_update(len(self.soln), self.soln, self.solnPrev, self.solnPrev2, self.solnPrev3,
                              self.solnPrev4, self.solnDot, self.solnDotPrev)

    def assemble(self, it, want_K=True, e0=0, e1=None):
        e1 = self.mesh.nElem if e1 is None else e1
        p = self.prep
        vals_ptr = self.vals.ctypes.data_as(C.c_void_p) if want_K else None
        lib().orc_assemble(self.etype, e0, e1, it, self.conn, p["N"], p["dNdx"], p["dNdy"], p["dNdz"],
                           p["dvol"], p["charlen"], np.ascontiguousarray(self.ctl.elemData, np.float64),
                           self.td, self.solnPrev, self.solnCur, self.solnDotCur, self.solnApplied,
                           self.nodeDof, self.rowptr, self.colidx, self.nfree, vals_ptr, self.rhs, 1)

    def linear_solve(self):
        so = self.solver
        if so is None or so.pc_type == "lu" or so.ksp_type == "preonly":
            import scipy.sparse as sp
            import scipy.sparse.linalg as spla
            A = sp.csr_matrix((self.vals, self.colidx, self.rowptr), shape=(self.nfree, self.nfree))
            du = spla.spsolve(A.tocsc(), self.rhs)
            return du, 0, 1
        blk = node_blocks(self.nodeDof, self.ndof) if so.pc_type == "bjacobi4" else None
        pc = OraclePC(so.pc_type, self.rowptr, self.colidx, self.vals, blk)
        du, rc, its, rr = krylov(so.ksp_type, pc, self.rowptr, self.colidx, self.vals, self.rhs,
                                 so.rtol, so.max_it, so.restart)
        if rc != 0:
            raise RuntimeError(f"Krylov solver did not converge (rc={rc}, its={its}, relres={rr})")
        return du, rc, its

    def solve_fully_implicit(self, newton_tol=1.0e-8, max_iter=10, steps=None, reuse_K=False, timers=None):
        """StabFEM::solveFullyImplicit (StabFEM2.cpp:593-832) on one rank.

        reuse_K: assemble the tangent at iteration 0 only and the residual alone after
        it.  The residual is affine in the state (the tangent depends on the previous
        step only), so the reference's re-assembled tangent is the same matrix; the HIP
        path does the same.  timers: a dict that receives the wall time of the element
        loops ("assembly") and of the linear solves ("solver"), the split of the
        reference's MPI_Wtime phase timers (StabFEM.cpp:711-727)."""
        import time as _time
        c = self.ctl
        m = self.mesh
        stepsMax = c.stepsMax if steps is None else steps
        timeNow = c.dt
        if timers is not None:
            timers.setdefault("assembly", 0.0)
            timers.setdefault("solver", 0.0)
        for tstep in range(stepsMax):
            self.td = time_params(c.tis, c.rhoInf, c.dt)
            self.time_update()
            timeFact = timeNow if timeNow <= 1.0 else 1.0
            self.assign_boundary_conditions(timeFact)
            for it in range(max_iter):
                self.update_iter_step()
                t0 = _time.perf_counter()
                self.assemble(it, want_K=(it == 0 or not reuse_K))
                if timers is not None:
                    timers["assembly"] += _time.perf_counter() - t0
                if it == 0:
                    for k in range(m.nDBC):
                        jj = int(m.dbc_node[k]) * self.ndof + int(m.dbc_dof[k])
                        self.soln[jj] += self.solnApplied[jj]
                norm_rhs = float(np.sqrt(np.dot(self.rhs, self.rhs)))
                self.log.append((tstep, it, norm_rhs))
                if norm_rhs < newton_tol:
                    break
                t0 = _time.perf_counter()
                du, rc, its = self.linear_solve()
                if timers is not None:
                    timers["solver"] += _time.perf_counter() - t0
                    timers["krylov_its"] = timers.get("krylov_its", 0) + its
                self.soln[self.assyForSoln] += du
            timeNow = timeNow + c.dt
            if timeNow > c.timeFinal:
                break
        return self.soln
