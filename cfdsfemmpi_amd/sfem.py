"""ctypes binding of libsfem_hip.so and the host driver mirroring ``StabFEM``.

``StabFEM`` keeps the reference driver's method names and call order
(src/base/StabFEM.h:101-153, src/cfdsfemmpi.cpp:22-89):
``readInputData`` -> ``readControlParameters`` -> ``prepareInputData`` ->
``setSolver`` -> ``solveFullyImplicit`` -> ``checkResult`` / ``writeResult``.
Every numerical step runs in the HIP library through the C ABI of
``include/sfem.h``; there is no CPU fallback -- if the library cannot be
loaded or no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import deck as _deck

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfem_hip.so")

SFEM_OK = 0
ERRORS = {-1: "SFEM_ERR_ARG", -2: "SFEM_ERR_STATE", -3: "SFEM_ERR_NEGJAC", -4: "SFEM_ERR_HIP",
          -5: "SFEM_ERR_DIVERGED", -6: "SFEM_ERR_MAXIT", -7: "SFEM_ERR_NOMEM", -8: "SFEM_ERR_NOTCONV",
          -9: "SFEM_ERR_UNSUPPORTED"}
PREC_FP64, PREC_MIXED, PREC_STORE32 = 0, 1, 2
PRED_PREVIOUS, PRED_EXTRAPOLATE, PRED_LEARNED, PRED_QUADRATIC = 0, 1, 2, 3
LEARNER_NPARAM = 5190
KT_ASSEMBLY, KT_SPMV, KT_PC = 0, 1, 2
TUNE_TIMED, TUNE_FIXED = 0, 1
KSP = {"bcgs": 0, "gmres": 1, "preonly": 2}
PC = {"none": 0, "jacobi": 1, "bjacobi4": 2, "ilu": 3, "lu": 4}

# (function name, restype, argtypes) of every entry point declared in include/sfem.h
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


class SolverOpts(C.Structure):
    _fields_ = [("ksp_type", C.c_int), ("pc_type", C.c_int), ("rtol", C.c_double),
                ("max_it", C.c_int), ("restart", C.c_int), ("reorder", C.c_int), ("factor_bf16", C.c_int),
                ("atol", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("krylov_iters", C.c_int), ("rhs_norm0", C.c_double),
                ("rhs_norm", C.c_double), ("t_assembly_ms", C.c_double), ("t_solver_ms", C.c_double)]


ABI = [
    ("sfem_version", C.c_char_p, []),
    ("sfem_last_error", C.c_char_p, []),
    ("sfem_create", _vp, [C.c_int, C.c_int, C.c_int64, C.c_int64, _f64p, _i32p, C.c_int]),
    ("sfem_destroy", None, [_vp]),
    ("sfem_set_dirichlet", C.c_int, [_vp, C.c_int64, _i32p, _i32p, _f64p]),
    ("sfem_prepare_elements", C.c_int, [_vp]),
    ("sfem_build_pattern", C.c_int, [_vp, _i64p, _i64p]),
    ("sfem_get_pattern", C.c_int, [_vp, _i64p, _i32p, _i32p]),
    ("sfem_set_params", C.c_int, [_vp, _f64p, C.c_int, C.c_double, C.c_double]),
    ("sfem_set_precision", C.c_int, [_vp, C.c_int]),
    ("sfem_set_predictor", C.c_int, [_vp, C.c_int]),
    ("sfem_set_pc_lag", C.c_int, [_vp, C.c_int]),
    ("sfem_set_coarse_correction", C.c_int, [_vp, C.c_int, C.c_int]),
    ("sfem_coarse_info", C.c_int, [_vp, _i64p]),
    ("sfem_coarse_inverse", C.c_int, [_vp, C.c_int, _f64p, C.c_int]),
    ("sfem_coarse_step", C.c_int, [_vp, C.POINTER(SolverOpts)]),
    ("sfem_set_inner_tolerance", C.c_int, [_vp, C.c_double, C.c_double]),
    ("sfem_time_update", C.c_int, [_vp, C.c_double]),
    ("sfem_assemble", C.c_int, [_vp, C.c_int, C.c_int, _f64p]),
    ("sfem_element_KF", C.c_int, [_vp, C.c_int64, _f64p, _f64p]),
    ("sfem_get_system", C.c_int, [_vp, _f64p, _f64p]),
    ("sfem_solve", C.c_int, [_vp, C.POINTER(SolverOpts), C.POINTER(C.c_int), _f64p]),
    ("sfem_setup_solver", C.c_int, [_vp, C.POINTER(SolverOpts)]),
    ("sfem_update_solution", C.c_int, [_vp]),
    ("sfem_step", C.c_int, [_vp, C.c_double, C.POINTER(SolverOpts), C.c_double, C.c_int,
                            C.POINTER(StepStats)]),
    ("sfem_get_solution", C.c_int, [_vp, _f64p]),
    ("sfem_set_solution", C.c_int, [_vp, _f64p]),
    ("sfem_get_increment", C.c_int, [_vp, _f64p]),
    ("sfem_set_state", C.c_int, [_vp, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
    ("sfem_reset_state", C.c_int, [_vp]),
    ("sfem_enable_kernel_timing", C.c_int, [_vp, C.c_int]),
    ("sfem_kernel_timing", C.c_int, [_vp, C.c_int, _i64p, _f64p]),
    ("sfem_get_ilu", C.c_int, [_vp, _i64p, _i32p, _i32p, _i32p, _f64p]),
    ("sfem_pc_apply", C.c_int, [_vp, _f64p, _f64p]),
    ("sfem_pc_apply_noalloc", C.c_int, [_vp, _f64p, _f64p]),
    ("sfem_solver_stats", C.c_int, [_vp, _i64p, _i64p]),
    ("sfem_set_batch", C.c_int, [_vp, C.c_int32, _i32p, _f64p]),
    ("sfem_step_batch", C.c_int, [_vp, C.c_double, C.POINTER(SolverOpts), C.c_double, C.c_int, _i32p, _i32p, _f64p]),
    ("sfem_batch_use_graph", C.c_int, [_vp, C.c_int]),
    ("sfem_batch_narrowing", C.c_int, [_vp, C.c_int]),
    ("sfem_block_info", C.c_int, [_vp, _i64p, _i64p, _i64p]),
    ("sfem_tile_info", C.c_int, [_vp, _i64p, _i64p, _i64p]),
    ("sfem_tuning", C.c_int, [_vp, C.POINTER(C.c_int)]),
    ("sfem_tuning_levels", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    ("sfem_element_tuning", C.c_int, [_vp, C.POINTER(C.c_int)]),
    ("sfem_set_tuning_mode", C.c_int, [_vp, C.c_int]),
    ("sfem_get_tuning_table", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    ("sfem_set_tuning_table", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int]),
    ("sfem_direct_info", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("sfem_info", C.c_int, [_vp, _i64p, _i64p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("sfem_last_timings", C.c_int, [_vp, _f64p, _f64p]),
    ("sfem_learner_create", _vp, [_vp, C.c_int64, C.c_uint64]),
    ("sfem_learner_destroy", None, [_vp]),
    ("sfem_learner_record", C.c_int, [_vp, _vp, C.c_int]),
    ("sfem_learner_info", C.c_int, [_vp, _i64p, _i64p, _i64p]),
    ("sfem_learner_train", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, _f64p]),
    ("sfem_learner_gradient", C.c_int, [_vp, C.c_int, _i32p, _i32p, C.POINTER(C.c_float), _f64p]),
    ("sfem_learner_weights", C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("sfem_learner_set_components", C.c_int, [_vp, C.c_int]),
    ("sfem_set_learner", C.c_int, [_vp, _vp]),
    ("sfem_learned_candidate", C.c_int, [_vp, _f64p]),
    ("sfem_predictor_stats", C.c_int, [_vp, _i64p, _f64p]),
]

_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libsfem_hip.so (raises if it is missing -- build it first)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python -m cfdsfemmpi_amd.build` first")
    L = C.CDLL(path)
    for name, res, args in ABI:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


class SfemError(RuntimeError):
    def __init__(self, code: int, where: str):
        lib = load_library()
        msg = lib.sfem_last_error().decode()
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


def _p(a: Optional[np.ndarray], ct):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def _check(code: int, where: str):
    if code != SFEM_OK:
        raise SfemError(code, where)


class Context:
    """Thin owner of one ``sfem_ctx``."""

    def __init__(self, mesh: _deck.Mesh, device: int = -1):
        self.lib = load_library()
        self.mesh = mesh
        self.ndof = mesh.ndim + 1
        self._coords = np.ascontiguousarray(mesh.coords, np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, np.int32)
        h = self.lib.sfem_create(mesh.ndim, mesh.npElem, mesh.nNode, mesh.nElem,
                                 _p(self._coords, C.c_double), _p(self._conn, C.c_int32), device)
        if not h:
            raise RuntimeError("sfem_create failed: " + self.lib.sfem_last_error().decode())
        self.h = C.c_void_p(h)
        self.nfree = 0
        self.nnz = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.sfem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -------------------------------------------------------------
    def set_dirichlet(self, node, dof, value):
        node = np.ascontiguousarray(node, np.int32)
        dof = np.ascontiguousarray(dof, np.int32)
        value = np.ascontiguousarray(value, np.float64)
        _check(self.lib.sfem_set_dirichlet(self.h, len(node), _p(node, C.c_int32), _p(dof, C.c_int32),
                                           _p(value, C.c_double)), "sfem_set_dirichlet")

    def prepare_elements(self):
        _check(self.lib.sfem_prepare_elements(self.h), "sfem_prepare_elements")

    def build_pattern(self):
        nf = C.c_int64(0)
        nz = C.c_int64(0)
        _check(self.lib.sfem_build_pattern(self.h, C.byref(nf), C.byref(nz)), "sfem_build_pattern")
        self.nfree, self.nnz = nf.value, nz.value
        return self.nfree, self.nnz

    def get_pattern(self):
        rp = np.zeros(self.nfree + 1, np.int64)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        af = np.zeros(max(self.nfree, 1), np.int32)
        _check(self.lib.sfem_get_pattern(self.h, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(af, C.c_int32)),
               "sfem_get_pattern")
        return rp, ci[:self.nnz], af[:self.nfree]

    def set_params(self, elemData, tis: int, rhoInf: float, dt: float):
        ed = np.ascontiguousarray(elemData, np.float64)
        assert ed.size == 11
        _check(self.lib.sfem_set_params(self.h, _p(ed, C.c_double), int(tis), float(rhoInf), float(dt)),
               "sfem_set_params")

    def set_precision(self, prec: int):
        _check(self.lib.sfem_set_precision(self.h, int(prec)), "sfem_set_precision")

    def set_predictor(self, kind: int):
        _check(self.lib.sfem_set_predictor(self.h, int(kind)), "sfem_set_predictor")

    def set_learner(self, learner: Optional["Learner"]):
        _check(self.lib.sfem_set_learner(self.h, learner.h if learner is not None else None), "sfem_set_learner")
        self._learner = learner  # keeps the learner alive while attached

    def learned_candidate(self) -> np.ndarray:
        out = np.zeros(self.mesh.nNode * self.ndof)
        _check(self.lib.sfem_learned_candidate(self.h, _p(out, C.c_double)), "sfem_learned_candidate")
        return out

    def predictor_stats(self):
        """(steps started from the [linear, quadratic, learned] start, their ||F|| at the
        last step) under PRED_LEARNED."""
        cnt = np.zeros(3, np.int64)
        nrm = np.zeros(3)
        _check(self.lib.sfem_predictor_stats(self.h, _p(cnt, C.c_int64), _p(nrm, C.c_double)),
               "sfem_predictor_stats")
        return cnt, nrm

    # -- hot path ----------------------------------------------------------
    def time_update(self, timeNow: float):
        """Start a time step at the accumulated time timeNow (the library applies the lid
        ramp min(timeNow, 1))."""
        _check(self.lib.sfem_time_update(self.h, float(timeNow)), "sfem_time_update")

    def assemble(self, it: int, want_K: bool = True) -> float:
        nrm = C.c_double(0.0)
        _check(self.lib.sfem_assemble(self.h, int(it), 1 if want_K else 0, C.byref(nrm)), "sfem_assemble")
        return nrm.value

    def element_KF(self, e: int):
        ns = self.mesh.npElem * self.ndof
        K = np.zeros(ns * ns)
        F = np.zeros(ns)
        _check(self.lib.sfem_element_KF(self.h, int(e), _p(K, C.c_double), _p(F, C.c_double)),
               "sfem_element_KF")
        return K.reshape(ns, ns), F

    def get_system(self):
        v = np.zeros(max(self.nnz, 1))
        r = np.zeros(max(self.nfree, 1))
        _check(self.lib.sfem_get_system(self.h, _p(v, C.c_double), _p(r, C.c_double)), "sfem_get_system")
        return v[:self.nnz], r[:self.nfree]

    def solve(self, opts: SolverOpts):
        it = C.c_int(0)
        rr = C.c_double(0.0)
        rc = self.lib.sfem_solve(self.h, C.byref(opts), C.byref(it), C.byref(rr))
        _check(rc, "sfem_solve")
        return it.value, rr.value

    def setup_solver(self, opts: SolverOpts):
        _check(self.lib.sfem_setup_solver(self.h, C.byref(opts)), "sfem_setup_solver")

    def update_solution(self):
        _check(self.lib.sfem_update_solution(self.h), "sfem_update_solution")

    def step(self, timeNow: float, opts: SolverOpts, newton_tol: float = 1e-8, max_iter: int = 10,
             allow_notconv: bool = True) -> StepStats:
        st = StepStats()
        rc = self.lib.sfem_step(self.h, float(timeNow), C.byref(opts), float(newton_tol), int(max_iter),
                                C.byref(st))
        if rc == -8 and allow_notconv:  # the reference moves on after 10 iterations
            return st
        _check(rc, "sfem_step")
        return st

    def get_solution(self) -> np.ndarray:
        s = np.zeros(self.mesh.nNode * self.ndof)
        _check(self.lib.sfem_get_solution(self.h, _p(s, C.c_double)), "sfem_get_solution")
        return s

    def set_solution(self, s: np.ndarray):
        s = np.ascontiguousarray(s, np.float64)
        _check(self.lib.sfem_set_solution(self.h, _p(s, C.c_double)), "sfem_set_solution")

    def get_increment(self) -> np.ndarray:
        d = np.zeros(max(self.nfree, 1))
        _check(self.lib.sfem_get_increment(self.h, _p(d, C.c_double)), "sfem_get_increment")
        return d[:self.nfree]

    def set_state(self, soln=None, solnPrev=None, solnCur=None, solnDotCur=None, solnDot=None,
                  solnDotPrev=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64)
                for a in (soln, solnPrev, solnCur, solnDotCur, solnDot, solnDotPrev)]
        _check(self.lib.sfem_set_state(self.h, *[_p(a, C.c_double) for a in arrs]), "sfem_set_state")

    def reset_state(self):
        _check(self.lib.sfem_reset_state(self.h), "sfem_reset_state")

    def enable_kernel_timing(self, on: bool = True):
        _check(self.lib.sfem_enable_kernel_timing(self.h, 1 if on else 0), "sfem_enable_kernel_timing")

    def kernel_timing(self, kind: int):
        n, ms = C.c_int64(0), C.c_double(0.0)
        _check(self.lib.sfem_kernel_timing(self.h, int(kind), C.byref(n), C.byref(ms)), "sfem_kernel_timing")
        return n.value, ms.value

    def block_info(self):
        """Sizes of the ILU(0) node-block form: block rows, neighbour blocks, values."""
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        _check(self.lib.sfem_block_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "sfem_block_info")
        return dict(nblk=a.value, npairs=b.value, nvals=c.value)

    def tile_info(self):
        """Tiles of a tiled ILU(0) plan: count, colours, factor values inside tiles."""
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        _check(self.lib.sfem_tile_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "sfem_tile_info")
        return dict(ntiles=a.value, ncolors=b.value, tile_vals=c.value)

    def set_batch(self, elem_mesh, mu):
        """Declare the mesh as a disjoint union of len(mu) meshes (elem_mesh: mesh of each element)."""
        em = np.ascontiguousarray(elem_mesh, np.int32)
        m = np.ascontiguousarray(mu, np.float64)
        _check(self.lib.sfem_set_batch(self.h, len(m), _p(em, C.c_int32), _p(m, C.c_double)), "sfem_set_batch")
        self.nmesh = len(m)

    def batch_use_graph(self, enable: bool):
        _check(self.lib.sfem_batch_use_graph(self.h, 1 if enable else 0), "sfem_batch_use_graph")

    def batch_narrowing(self, enable: bool):
        _check(self.lib.sfem_batch_narrowing(self.h, 1 if enable else 0), "sfem_batch_narrowing")

    def step_batch(self, timeNow: float, opts: SolverOpts, newton_tol: float = 1e-8, max_newton: int = 10):
        """One time step of every mesh: (newton_iters, krylov_iters, rhs_norm) per mesh."""
        ni = np.zeros(self.nmesh, np.int32)
        ki = np.zeros(self.nmesh, np.int32)
        rn = np.zeros(self.nmesh)
        _check(self.lib.sfem_step_batch(self.h, float(timeNow), C.byref(opts), float(newton_tol), int(max_newton),
                                        _p(ni, C.c_int32), _p(ki, C.c_int32), _p(rn, C.c_double)), "sfem_step_batch")
        return ni, ki, rn

    def fp32_fallbacks(self) -> int:
        """Mixed-precision solves finished by the FP32 GMRES fallback."""
        return self.solver_stats()[0]

    def solver_stats(self):
        """(FP32 GMRES fallbacks, FP32 solves that handed a partial iterate to Newton)."""
        v, w = C.c_int64(0), C.c_int64(0)
        _check(self.lib.sfem_solver_stats(self.h, C.byref(v), C.byref(w)), "sfem_solver_stats")
        return v.value, w.value

    def set_inner_tolerance(self, theta: float, rtol_max: float = 0.5):
        """Mixed paths: each Newton iteration's inner rtol becomes
        min(rtol_max, max(opts.rtol, theta * newton_tol / ||F||)) (theta 0: opts.rtol)."""
        _check(self.lib.sfem_set_inner_tolerance(self.h, float(theta), float(rtol_max)),
               "sfem_set_inner_tolerance")

    def set_pc_lag(self, lag: int):
        """Rebuild the preconditioner on every lag-th new tangent only (1 = always)."""
        _check(self.lib.sfem_set_pc_lag(self.h, int(lag)), "sfem_set_pc_lag")

    def set_coarse_correction(self, boxes_per_dim: int, multiplicative: bool = False):
        """Coarse pressure correction of every ILU(0) application: piecewise-constant
        pressure on boxes_per_dim boxes per direction, E = Z^T A Z (0 switches it off);
        additive z = M^-1 r + Z E^-1 Z^T r, or multiplicative z0 + Z E^-1 Z^T (r - A z0)."""
        _check(self.lib.sfem_set_coarse_correction(self.h, int(boxes_per_dim), int(bool(multiplicative))),
               "sfem_set_coarse_correction")

    def coarse_info(self) -> dict:
        """State of the coarse pressure correction."""
        v = np.zeros(6, np.int64)
        _check(self.lib.sfem_coarse_info(self.h, _p(v, C.c_int64)), "sfem_coarse_info")
        return {"k": int(v[0]), "unknowns": int(v[1]), "pressure_dofs": int(v[2]), "chunks": int(v[3]),
                "singular": bool(v[4]), "built": bool(v[5])}

    def coarse_inverse(self, which: int = 0) -> np.ndarray:
        """E^-1 (which 0) or the mixed-precision inverse B without the global constant
        (which 1) of the last preconditioner rebuild (unknowns x unknowns)."""
        nc = self.coarse_info()["unknowns"]
        out = np.zeros(max(nc * nc, 1))
        _check(self.lib.sfem_coarse_inverse(self.h, int(which), _p(out, C.c_double), out.size),
               "sfem_coarse_inverse")
        return out[:nc * nc].reshape(nc, nc)

    def coarse_step(self, opts: SolverOpts):
        """Increment = Z E^-1 Z^T rhs (FP64); apply it with update_solution()."""
        _check(self.lib.sfem_coarse_step(self.h, C.byref(opts)), "sfem_coarse_step")

    def get_ilu(self):
        """(perm, rowptr, colidx, vals) of the current ILU(0) factor in its ordering."""
        nnz = C.c_int64(0)
        _check(self.lib.sfem_get_ilu(self.h, C.byref(nnz), None, None, None, None), "sfem_get_ilu")
        n = self.nfree
        perm = np.zeros(max(n, 1), np.int32)
        rowptr = np.zeros(n + 1, np.int32)
        colidx = np.zeros(max(nnz.value, 1), np.int32)
        vals = np.zeros(max(nnz.value, 1))
        _check(self.lib.sfem_get_ilu(self.h, C.byref(nnz), _p(perm, C.c_int32), _p(rowptr, C.c_int32),
                                     _p(colidx, C.c_int32), _p(vals, C.c_double)), "sfem_get_ilu")
        return perm[:n], rowptr, colidx[:nnz.value], vals[:nnz.value]

    def pc_apply(self, r):
        """z = M^{-1} r with the preconditioner of the last solve (DOF order)."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros(max(self.nfree, 1))
        _check(self.lib.sfem_pc_apply(self.h, _p(r, C.c_double), _p(z, C.c_double)), "sfem_pc_apply")
        return z[:self.nfree]

    def element_tuning(self):
        """Element-loop variant per assembly mode {"K64": .., "F": .., "K32": ..} (-1: not
        chosen yet): 0-2 = 256/128/64 lanes per workgroup, 3-5 = the same with workgroups
        walking their colour class."""
        out = (C.c_int * 3)()
        _check(self.lib.sfem_element_tuning(self.h, out), "sfem_element_tuning")
        return dict(zip(("K64", "F", "K32"), list(out)))

    def tuning(self):
        """Autotuned kernel shapes per precision: SpMV (G, U, XCD order) and the
        per-level ILU(0) apply shapes, forward levels then backward levels."""
        out = (C.c_int * 8)()
        _check(self.lib.sfem_tuning(self.h, out), "sfem_tuning")
        v = list(out)
        res = {}
        for k, name in ((0, "fp64"), (1, "fp32")):
            w = v[4 * k:4 * k + 4]
            levels = []
            if w[3]:
                buf = (C.c_int * (3 * w[3]))()
                n = C.c_int(0)
                _check(self.lib.sfem_tuning_levels(self.h, k, buf, w[3], C.byref(n)), "sfem_tuning_levels")
                b = list(buf)
                levels = [tuple(b[3 * i:3 * i + 3]) for i in range(n.value)]
            res[name] = {"spmv_GUX": w[0:3], "ilu_levels_GUX": levels}
        return res

    def set_tuning_mode(self, mode: int):
        """TUNE_TIMED (default): time every kernel shape on the first solve; TUNE_FIXED:
        default shapes, so FP32 results repeat exactly from run to run."""
        _check(self.lib.sfem_set_tuning_mode(self.h, int(mode)), "sfem_set_tuning_mode")

    def get_tuning_table(self, fp32: bool) -> list:
        n = C.c_int(0)
        _check(self.lib.sfem_get_tuning_table(self.h, 1 if fp32 else 0, None, 0, C.byref(n)), "sfem_get_tuning_table")
        buf = (C.c_int * max(n.value, 1))()
        _check(self.lib.sfem_get_tuning_table(self.h, 1 if fp32 else 0, buf, n.value, C.byref(n)),
               "sfem_get_tuning_table")
        return list(buf)[:n.value]

    def set_tuning_table(self, fp32: bool, table):
        buf = (C.c_int * max(len(table), 1))(*[int(v) for v in table])
        _check(self.lib.sfem_set_tuning_table(self.h, 1 if fp32 else 0, buf, len(table)), "sfem_set_tuning_table")

    def direct_info(self):
        """(kl, ku) of the direct solver's reordered tangent (0, 0 before an LU solve)."""
        kl, ku = C.c_int(0), C.c_int(0)
        _check(self.lib.sfem_direct_info(self.h, C.byref(kl), C.byref(ku)), "sfem_direct_info")
        return kl.value, ku.value

    def info(self):
        nf, nz = C.c_int64(0), C.c_int64(0)
        nc, nd = C.c_int(0), C.c_int(0)
        _check(self.lib.sfem_info(self.h, C.byref(nf), C.byref(nz), C.byref(nc), C.byref(nd)), "sfem_info")
        return dict(nfree=nf.value, nnz=nz.value, ncolors=nc.value, ndof=nd.value)

    def last_timings(self):
        a, s = C.c_double(0.0), C.c_double(0.0)
        _check(self.lib.sfem_last_timings(self.h, C.byref(a), C.byref(s)), "sfem_last_timings")
        return a.value, s.value


class Learner:
    """Owner of one ``sfem_learner``: the training store and the MLP of the learned
    initial iterate for the mesh of ``ctx`` (2D meshes)."""

    def __init__(self, ctx: "Context", capacity: int, seed: int = 0):
        self.lib = load_library()
        h = self.lib.sfem_learner_create(ctx.h, int(capacity), int(seed) & (2 ** 64 - 1))
        if not h:
            raise RuntimeError("sfem_learner_create failed: " + self.lib.sfem_last_error().decode())
        self.h = C.c_void_p(h)

    def record(self, ctx: "Context", new_trajectory: bool = False):
        _check(self.lib.sfem_learner_record(self.h, ctx.h, 1 if new_trajectory else 0), "sfem_learner_record")

    def info(self):
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        _check(self.lib.sfem_learner_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "sfem_learner_info")
        return a.value, b.value, c.value

    def train(self, iters: int, batch: int, lr: float) -> np.ndarray:
        loss = np.zeros(max(int(iters), 1))
        _check(self.lib.sfem_learner_train(self.h, int(iters), int(batch), float(lr), _p(loss, C.c_double)),
               "sfem_learner_train")
        return loss[:int(iters)]

    def gradient(self, window: np.ndarray, node: np.ndarray):
        w = np.ascontiguousarray(window, np.int32)
        n = np.ascontiguousarray(node, np.int32)
        g = np.zeros(LEARNER_NPARAM, np.float32)
        loss = C.c_double(0.0)
        _check(self.lib.sfem_learner_gradient(self.h, len(w), _p(w, C.c_int32), _p(n, C.c_int32),
                                              _p(g, C.c_float), C.byref(loss)), "sfem_learner_gradient")
        return g, loss.value

    def get_weights(self) -> np.ndarray:
        w = np.zeros(LEARNER_NPARAM, np.float32)
        _check(self.lib.sfem_learner_weights(self.h, _p(w, C.c_float), None), "sfem_learner_weights")
        return w

    def set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == LEARNER_NPARAM
        _check(self.lib.sfem_learner_weights(self.h, None, _p(w, C.c_float)), "sfem_learner_weights")

    def set_components(self, n: int):
        """Correct (and train) u, v (n = 2, the default) or u, v, p (n = 3)."""
        _check(self.lib.sfem_learner_set_components(self.h, int(n)), "sfem_learner_set_components")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sfem_learner_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solver_opts(ksp="bcgs", pc="ilu", rtol=1e-8, max_it=1000, restart=30, reorder=0,
                factor_bf16=0, atol=0.0) -> SolverOpts:
    if pc == "lu":
        # the reference's test decks (-ksp_type preonly -pc_type lu, MUMPS): the device
        # banded LU direct solve, whatever -ksp_type the options file names
        ksp = "preonly"
    return SolverOpts(KSP[ksp], PC[pc], float(rtol), int(max_it), int(restart), int(reorder), int(factor_bf16),
                      float(atol))


class StabFEM:
    """Host driver with the reference's orchestration (one GPU, one rank)."""

    def __init__(self, device: int = -1):
        self.device = device
        self.mesh: Optional[_deck.Mesh] = None
        self.ctl: Optional[_deck.Control] = None
        self.ctx: Optional[Context] = None
        self.opts = solver_opts()
        self.newton_tol = 1.0e-8   # hard-coded absolute tolerance (StabFEM2.cpp:749)
        self.forces = []           # (timeNow, Fx, Fy) per step, the forces file's lines
        # postProcess after every step (StabFEM2.cpp:802-806) writes <infilename>-<n>.vtu;
        # off unless asked for, the files go through a background writer thread
        self.postprocess = False
        self.infilename = "sfem"
        self.fileCount = 0
        self._vtu = None
        self.max_newton = 10       # hard-coded iteration cap (StabFEM2.cpp:654)
        self.precision = PREC_FP64
        self.tuning = TUNE_TIMED   # kernel shapes of the solver: timed on first use, or fixed
        self.coarse_boxes = 0      # coarse pressure correction of ILU(0): boxes per direction (0: off)
        self.coarse_mult = False   # its multiplicative form
        self.history = []          # (tstep, iter, rhs_norm, krylov_its)

    # StabFEM::readInputData (StabFEM.cpp:80-264)
    def readInputData(self, fname: str):
        self.mesh = _deck.read_mesh(fname)
        base = os.path.basename(fname)
        for suffix in (".gz", ".dat"):
            if base.endswith(suffix):
                base = base[:-len(suffix)]
        self.infilename = os.path.join(os.path.dirname(fname), base)

    # StabFEM::readControlParameters (StabFEM.cpp:268-322)
    def readControlParameters(self, fname: str):
        self.ctl = _deck.read_control(fname)

    def setMesh(self, mesh: _deck.Mesh):
        self.mesh = mesh

    def setControl(self, ctl: _deck.Control):
        self.ctl = ctl

    # StabFEM::prepareInputData (StabFEM.cpp:328-386)
    def prepareInputData(self):
        m = self.mesh
        self.ctx = Context(m, self.device)
        self.ctx.set_dirichlet(m.dbc_node, m.dbc_dof, m.dbc_value)

    # StabFEM::setSolver -> prepareMatrixPattern (StabFEM2.cpp:10-319); the options
    # file stands in for PetscInitialize(optfile) (src/cfdsfemmpi.cpp:40)
    def setSolver(self, options: Optional[object] = None):
        if isinstance(options, str):
            so = _deck.read_solver_options(options)
            self.opts = solver_opts(so.ksp_type, so.pc_type, so.rtol, so.max_it, so.restart)
        elif isinstance(options, SolverOpts):
            self.opts = options
        self.ctx.prepare_elements()
        self.ctx.build_pattern()
        c = self.ctl
        self.ctx.set_params(c.elemData, c.tis, c.rhoInf, c.dt)
        self.ctx.set_precision(self.precision)
        self.ctx.set_tuning_mode(self.tuning)
        if self.coarse_boxes:
            self.ctx.set_coarse_correction(self.coarse_boxes, self.coarse_mult)
        self.ctx.setup_solver(self.opts)

    # StabFEM::solveFullyImplicit (StabFEM2.cpp:593-832)
    def solveFullyImplicit(self, steps: Optional[int] = None):
        c = self.ctl
        ctx = self.ctx
        stepsMax = c.stepsMax if steps is None else steps
        timeNow = c.dt
        for tstep in range(stepsMax):
            ctx.time_update(timeNow)  # the lid ramp timeFact = min(timeNow, 1) is applied inside
            for it in range(self.max_newton):
                nrm = ctx.assemble(it, want_K=(it == 0))
                if nrm < self.newton_tol:
                    self.history.append((tstep, it, nrm, 0))
                    break
                if self.coarse_boxes and self.precision == PREC_MIXED:
                    # FP64 coarse step before the FP32 inner solve (as sfem_step does)
                    ctx.coarse_step(self.opts)
                    ctx.update_solution()
                    nrm = ctx.assemble(it + 1, want_K=False)
                    if nrm < self.newton_tol:
                        self.history.append((tstep, it, nrm, 0))
                        break
                its, _ = ctx.solve(self.opts)
                self.history.append((tstep, it, nrm, its))
                ctx.update_solution()
            # StabFEM2.cpp:808-815: the drag/lift line of the forces file, summed from the
            # reaction vector at the output nodes; the reference zeroes that vector every
            # iteration (StabFEM2.cpp:660) and never adds to it, so both sums are 0
            ndof = self.mesh.ndim + 1
            reacVec = np.zeros(self.mesh.nNode * ndof)
            out = np.asarray(self.mesh.out_nodes, np.int64)
            self.forces.append((timeNow, float(reacVec[out * ndof].sum()), float(reacVec[out * ndof + 1].sum())))
            if self.postprocess:
                self.postProcess()
            timeNow = timeNow + c.dt
            if timeNow > c.timeFinal:
                break
        if self._vtu is not None:
            self._vtu.close()
        return ctx.get_solution()

    # StabFEM::postProcess (StabFEM3.cpp:6-160): velocity and pressure of the current
    # solution on the mesh as <infilename>-<fileCount %06d>.vtu, fileCount from 0
    def postProcess(self):
        from .vtu import VTUWriter
        m = self.mesh
        if self._vtu is None:
            self._vtu = VTUWriter(m.coords, m.conn, m.ndim, asynchronous=True)
        self._vtu.submit(f"{self.infilename}-{self.fileCount:06d}.vtu", self.ctx.get_solution())
        self.fileCount += 1

    # the "<mesh>-forces.dat" file opened in StabFEM::readInputData (StabFEM.cpp:249-259)
    # and written once per step (StabFEM2.cpp:815): fixed notation, 14 decimals
    def writeForces(self, fname: str):
        with open(fname, "w") as f:
            for t, fx, fy in self.forces:
                f.write(f"{t:.14f}\t{fx:.14f}\t{fy:.14f}\n")

    def getSolution(self) -> np.ndarray:
        return self.ctx.get_solution()

    # StabFEM::checkResult (StabFEM.cpp:501-554)
    def checkResult(self, fname: str) -> int:
        ref = _deck.read_result(fname)
        return _deck.check_result(self.getSolution(), ref, self.mesh.ndim)

    # StabFEM::writeResult (StabFEM.cpp:457-494)
    def writeResult(self, fname: str):
        _deck.write_result(fname, self.getSolution(), self.mesh.ndim + 1)
