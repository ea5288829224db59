"""GPU tests of the additive coarse pressure correction of ILU(0), through the C ABI.

The correction adds Z E^-1 Z^T r to every ILU(0) application, with Z the
piecewise-constant pressure on k boxes per direction of the bounding box and
E = Z^T A Z.  It has no counterpart in the reference (PETSc ILU(0) alone,
SolverPetsc.cpp:284-298), so the checks are against what the definition fixes:
E^-1 against numpy's inverse of Z^T A Z formed from the exported CSR tangent, the
change of M^-1 r against Z E^-1 Z^T r, and converged Newton solutions against the
uncorrected solver at the same tolerances (a preconditioner changes iteration
counts, never the answer).
"""
import numpy as np
import pytest

import oracle
from cfdsfemmpi_amd import deck, meshgen, sfem

pytestmark = pytest.mark.gpu

ED = np.array([1.0, 1e-2, 0, 0, 0, 0, 0, 0, 1, 1, 0])


def _ctx(mesh, prec=sfem.PREC_FP64):
    ctx = sfem.Context(mesh)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    ctx.set_params(ED, 2, 0.0, 0.1)
    ctx.set_precision(prec)
    ctx.set_tuning_mode(sfem.TUNE_FIXED)
    return ctx


def _coarse_space(mesh, k):
    """Z (free DOFs x coarse unknowns) by the documented rule: node box
    b_d = min(k - 1, floor(k (x_d - min_d) / extent_d)), one column per box holding a free
    pressure DOF, boxes in x-fastest order; and per free pressure row whether its node
    shares an element with a node whose pressure is prescribed."""
    nd = mesh.ndim
    ndof = nd + 1
    _, assy, nfree = oracle.dof_numbering(mesh.nNode, ndof, mesh.dbc_node, mesh.dbc_dof)
    assy = np.asarray(assy, np.int64)
    x = np.asarray(mesh.coords, np.float64).reshape(-1, nd)
    lo, hi = x.min(0), x.max(0)
    b = np.minimum(np.floor((x - lo) / (hi - lo) * k).astype(np.int64), k - 1)
    lin = b[:, 0] + (k * b[:, 1] if nd > 1 else 0) + (k * k * b[:, 2] if nd > 2 else 0)
    rows = np.nonzero(assy % ndof == ndof - 1)[0]
    boxes = lin[assy[rows] // ndof]
    uniq, col = np.unique(boxes, return_inverse=True)
    Z = np.zeros((nfree, len(uniq)))
    Z[rows, col] = 1.0
    conn = np.asarray(mesh.conn, np.int64).reshape(mesh.nElem, -1)
    fixed = np.zeros(mesh.nNode, bool)
    dn, dd = np.asarray(mesh.dbc_node), np.asarray(mesh.dbc_dof)
    fixed[dn[dd == ndof - 1]] = True
    near = np.zeros(mesh.nNode, bool)
    near[np.unique(conn[fixed[conn].any(1)])] = True
    return Z, rows, col, near[assy[rows] // ndof]


def _coarse_matrices(mesh, A, k):
    """E = Z^T A Z summed directly, and E formed as the library documents it: the
    couplings between different boxes summed directly, each diagonal entry as the row
    sums of the rows next to a prescribed pressure minus the box's couplings to the other
    boxes."""
    Z, rows, col, near = _coarse_space(mesh, k)
    E = Z.T @ (A @ Z)
    App = A[rows][:, rows].tocoo()
    nc = Z.shape[1]
    ci, cj = col[App.row], col[App.col]
    off = ci != cj
    Eid = np.zeros((nc, nc))
    np.add.at(Eid, (ci[off], cj[off]), App.data[off])
    rs = np.asarray(A[rows][:, rows].sum(axis=1)).ravel()
    diag = np.zeros(nc)
    np.add.at(diag, col[near], rs[near])
    Eid[np.arange(nc), np.arange(nc)] = diag - Eid.sum(axis=1)
    return Z, E, Eid


def _system(ctx):
    import scipy.sparse as sp
    rp, ci, _ = ctx.get_pattern()
    vals, rhs = ctx.get_system()
    n = ctx.nfree
    return sp.csr_matrix((vals, ci, rp), shape=(n, n)), rhs


CASES = [("hex", 6, 2), ("hex", 8, 3), ("quad", 16, 4), ("tria", 12, 3)]


def _mesh(kind, n):
    if kind == "hex":
        return meshgen.ldc3d_hex(n, graded=True)
    if kind == "quad":
        return meshgen.ldc2d_quad(n, graded=True)
    return meshgen.ldc2d_tria(n, jitter=0.2, seed=5)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("kind,n,k", CASES)
def test_coarse_inverse_and_apply(kind, n, k, prec):
    """E^-1 from the device (E assembled from the plan's block tangent, inverted by
    Gauss-Jordan with partial pivoting) against numpy on the exported CSR tangent (FP64,
    or the FP32 tangent widened).  FP64: the documented E equals Z^T A Z summed directly
    within 1e-10 (the pressure rows of the stabilised tangent sum to zero away from
    prescribed pressures) and E^-1 E = I within 1e-9.  Both: E^-1 within 1e-8 relative of
    numpy's inverse of the documented E; M_c^-1 r - M^-1 r = Z E^-1 Z^T r within 1e-10 of
    ||M_c^-1 r||, the same twice in a row bit for bit."""
    mesh = _mesh(kind, n)
    print(f"[case] coarse inverse/apply {kind} n={n} k={k} {prec}", flush=True)
    ctx = _ctx(mesh, sfem.PREC_MIXED if prec == "mixed" else sfem.PREC_FP64)
    ctx.time_update(1.0)
    ctx.assemble(0)
    opts = sfem.solver_opts("bcgs", "ilu", 1e-8 if prec == "fp64" else 1e-4, 2000, reorder=3)
    ctx.solve(opts)
    r = np.random.default_rng(3).standard_normal(ctx.nfree)
    z0 = ctx.pc_apply(r)
    ctx.set_coarse_correction(k)
    ctx.solve(opts)
    info = ctx.coarse_info()
    A, _ = _system(ctx)
    Z, E, Eid = _coarse_matrices(mesh, A, k)
    nc = Z.shape[1]
    assert info["built"] and not info["singular"]
    assert info["unknowns"] == nc and info["pressure_dofs"] == int(Z.sum())
    Einv = ctx.coarse_inverse()
    Eref = np.linalg.inv(Eid)
    if prec == "fp64":
        eid = np.max(np.abs(Eid - E)) / np.max(np.abs(E))
        e1 = np.max(np.abs(Einv @ Eid - np.eye(nc)))
        assert eid <= 1e-10 and e1 <= 1e-9, (eid, e1)
    e2 = np.max(np.abs(Einv - Eref)) / np.max(np.abs(Eref))
    print(f"[coarse] nc={nc} cond={np.linalg.cond(Eid):.2e} rel={e2:.2e}", flush=True)
    assert e2 <= 1e-8
    if prec == "mixed":
        # the FP32 Krylov vectors' inverse: leading block of [E b1; b1^T 0]^-1
        bb = np.max(np.abs(np.diag(Eid)))
        Bord = np.zeros((nc + 1, nc + 1))
        Bord[:nc, :nc] = Eid
        Bord[:nc, nc] = bb
        Bord[nc, :nc] = bb
        Bref = np.linalg.inv(Bord)[:nc, :nc]
        B = ctx.coarse_inverse(1)
        eb = np.max(np.abs(B - Bref)) / np.max(np.abs(Bref))
        y = B @ np.random.default_rng(5).standard_normal(nc)
        print(f"[coarse] mean-free inverse rel={eb:.2e}", flush=True)
        assert eb <= 1e-8 and abs(y.sum()) <= 1e-9 * np.abs(y).sum()
    else:
        with pytest.raises(sfem.SfemError):
            ctx.coarse_inverse(1)
    z1 = ctx.pc_apply(r)
    z2 = ctx.pc_apply(r)
    assert np.array_equal(z1, z2)
    corr = Z @ (Eref @ (Z.T @ r))
    d = np.linalg.norm((z1 - z0) - corr) / np.linalg.norm(z1)
    print(f"[coarse] apply diff {d:.2e}, |corr|/|z| {np.linalg.norm(corr) / np.linalg.norm(z1):.2e}", flush=True)
    assert np.linalg.norm(corr) > 1e-6 * np.linalg.norm(z1)
    assert d <= 1e-10
    ctx.close()


@pytest.mark.parametrize("prec", [sfem.PREC_FP64, sfem.PREC_MIXED])
def test_coarse_step_increment(prec):
    """sfem_coarse_step: the increment equals Z E^-1 Z^T rhs (numpy, from the exported
    tangent and residual) within 1e-10 relative."""
    mesh = meshgen.ldc3d_hex(7, graded=True)
    ctx = _ctx(mesh, prec)
    ctx.time_update(1.0)
    ctx.assemble(0)
    opts = sfem.solver_opts("bcgs", "ilu", 1e-6, 2000, reorder=3)
    ctx.set_coarse_correction(2)
    ctx.coarse_step(opts)
    du = ctx.get_increment()
    A, rhs = _system(ctx)
    Z, _, Eid = _coarse_matrices(mesh, A, 2)
    ref = Z @ np.linalg.solve(Eid, Z.T @ rhs)
    d = np.linalg.norm(du - ref) / np.linalg.norm(ref)
    print(f"[coarse step] prec={prec} rel diff {d:.2e}", flush=True)
    assert d <= 1e-10
    ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_coarse_multiplicative_apply(prec):
    """Multiplicative form, standalone application (host vectors): with z0 = M_ILU^-1 r
    from the same factor, M_c^-1 r - z0 = Z E^-1 Z^T (r - A z0) (numpy, exported tangent)
    within 1e-10 of ||M_c^-1 r||."""
    mesh = meshgen.ldc3d_hex(8, graded=True)
    ctx = _ctx(mesh, sfem.PREC_MIXED if prec == "mixed" else sfem.PREC_FP64)
    ctx.time_update(1.0)
    ctx.assemble(0)
    opts = sfem.solver_opts("bcgs", "ilu", 1e-8 if prec == "fp64" else 1e-4, 2000, reorder=3)
    ctx.solve(opts)
    r = np.random.default_rng(4).standard_normal(ctx.nfree)
    z0 = ctx.pc_apply(r)
    ctx.set_coarse_correction(3, True)
    ctx.solve(opts)
    z1 = ctx.pc_apply(r)
    A, _ = _system(ctx)
    Z, _, Eid = _coarse_matrices(mesh, A, 3)
    corr = Z @ np.linalg.solve(Eid, Z.T @ (r - A @ z0))
    d = np.linalg.norm((z1 - z0) - corr) / np.linalg.norm(z1)
    print(f"[coarse mult] {prec} apply diff {d:.2e}, |corr|/|z| {np.linalg.norm(corr) / np.linalg.norm(z1):.2e}",
          flush=True)
    assert np.linalg.norm(corr) > 1e-6 * np.linalg.norm(z1)
    assert d <= 1e-10
    ctx.close()


def _cavity(n, steps, prec, k, rtol, newton_tol, mult=False):
    mesh = meshgen.ldc3d_hex(n, graded=True)
    ctl = deck.Control(np.array([1.0, 1e-3, 0, 0, 0, 0, 0, 0, 1, 1, 0]), 2, 0.0, 0.1, 1e9, steps, 1, 1e-6)
    fem = sfem.StabFEM()
    fem.setMesh(mesh)
    fem.setControl(ctl)
    fem.precision = prec
    fem.newton_tol = newton_tol
    fem.tuning = sfem.TUNE_FIXED
    fem.coarse_boxes = k
    fem.coarse_mult = mult
    fem.prepareInputData()
    fem.setSolver(sfem.solver_opts("bcgs", "ilu", rtol, 5000, reorder=3))
    return fem


def _fields(s):
    s = s.reshape(-1, 4)
    return s[:, :3], s[:, 3]


@pytest.mark.parametrize("mult", [False, True])
@pytest.mark.parametrize("prec", [sfem.PREC_FP64, sfem.PREC_MIXED])
def test_coarse_cavity_same_answer_fewer_iterations(prec, mult):
    """Graded 12^3 cavity at Re 1000, 4 steps, against the FP64 solver without the
    correction, both driven to ||F|| < 1e-11: velocity within 1e-8 and pressure within
    1e-6 (relative L2; pressure is the weakly observed field of this PSPG system), and
    no more Krylov iterations in total than the same precision without it (the mixed
    path adds the FP64 coarse step before each FP32 inner solve)."""
    n, steps = 12, 4
    ref = _cavity(n, steps, sfem.PREC_FP64, 0, 1e-12, 1e-11)
    s_ref = ref.solveFullyImplicit()
    base = _cavity(n, steps, prec, 0, 1e-12 if prec == sfem.PREC_FP64 else 1e-4, 1e-11)
    base.solveFullyImplicit()
    co = _cavity(n, steps, prec, 4, 1e-12 if prec == sfem.PREC_FP64 else 1e-4, 1e-11, mult)
    s_co = co.solveFullyImplicit()
    its_base = sum(h[3] for h in base.history)
    its_co = sum(h[3] for h in co.history)
    u0, p0 = _fields(s_ref)
    u1, p1 = _fields(s_co)
    eu = np.linalg.norm(u1 - u0) / np.linalg.norm(u0)
    ep = np.linalg.norm(p1 - p0) / np.linalg.norm(p0)
    print(f"[coarse cavity] prec={prec} mult={mult} its {its_base} -> {its_co}; rel L2 u {eu:.2e} p {ep:.2e}",
          flush=True)
    assert all(h[2] < 1e-11 for h in co.history if h[3] == 0)
    assert eu <= 1e-8 and ep <= 1e-6
    assert its_co <= its_base
    for f in (ref, base, co):
        f.ctx.close()


def _channel():
    return meshgen.cylinder_channel(nr=8, nth=32, ratio=1.2, n_left=6, n_right=16, n_bottom=6, n_top=6,
                                    perturb=0.01, seed=3)


def _pinv_sym(E):
    """Pseudo-inverse of the symmetric part of E without its eigenvalue nearest zero
    (numpy's symmetric eigensolver)."""
    w, V = np.linalg.eigh(0.5 * (E + E.T))
    keep = np.argsort(np.abs(w))[1:]
    return (V[:, keep] / w[keep]) @ V[:, keep].T


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_coarse_without_prescribed_pressure(prec):
    """Channel past a cylinder with a traction-free outflow and no prescribed pressure
    (Quad4): every row of the documented E sums to zero, so E is singular with the
    constant as null vector, and the device's correction is its pseudo-inverse: both
    inverses the library exposes within 1e-8 relative of numpy's (symmetric
    eigendecomposition without the null eigenvalue).  FP64: the multiplicative
    application's change M_c^-1 r - M^-1 r equals Z E^+ Z^T (r - A M^-1 r) within 1e-10
    of ||M_c^-1 r||."""
    mesh = _channel()
    print(f"[case] coarse without prescribed pressure {prec}", flush=True)
    ctx = _ctx(mesh, sfem.PREC_MIXED if prec == "mixed" else sfem.PREC_FP64)
    ctx.time_update(1.0)
    ctx.assemble(0)
    opts = sfem.solver_opts("bcgs", "ilu", 1e-8 if prec == "fp64" else 0.5, 3000, reorder=3)
    ctx.solve(opts)
    r = np.random.default_rng(6).standard_normal(ctx.nfree)
    z0 = ctx.pc_apply(r)
    k = 6
    ctx.set_coarse_correction(k, True)
    ctx.solve(opts)
    info = ctx.coarse_info()
    assert info["built"] and not info["singular"]
    A, _ = _system(ctx)
    Z, E, Eid = _coarse_matrices(mesh, A, k)
    nc = Z.shape[1]
    assert info["unknowns"] == nc
    dmax = np.max(np.abs(np.diag(Eid)))
    assert np.max(np.abs(Eid.sum(axis=1))) <= 1e-12 * dmax
    ref = _pinv_sym(Eid)
    for which in ((0, 1) if prec == "mixed" else (0,)):
        P = ctx.coarse_inverse(which)
        e = np.max(np.abs(P - ref)) / np.max(np.abs(ref))
        print(f"[coarse float] {prec} inverse {which}: nc={nc} rel={e:.2e}", flush=True)
        assert e <= 1e-8
    if prec == "fp64":
        z1 = ctx.pc_apply(r)
        corr = Z @ (ref @ (Z.T @ (r - A @ z0)))
        d = np.linalg.norm((z1 - z0) - corr) / np.linalg.norm(z1)
        print(f"[coarse float] apply diff {d:.2e}, |corr|/|z| {np.linalg.norm(corr) / np.linalg.norm(z1):.2e}",
              flush=True)
        assert np.linalg.norm(corr) > 1e-6 * np.linalg.norm(z1)
        assert d <= 1e-10
    ctx.close()


def test_coarse_batch_of_channels_pseudo_inverse():
    """A batch of two channels without a prescribed pressure and one pinned cavity: each
    channel's block of the correction is the pseudo-inverse of its own E (numpy, on the
    exported tangent of that mesh alone, at the batch's first tangent), the cavity's block
    the inverse of its E, and nothing couples the blocks."""
    ch = [meshgen.cylinder_channel(nr=8, nth=32, ratio=1.2, n_left=6, n_right=16, n_bottom=6, n_top=6,
                                   perturb=0.01, seed=s) for s in (3, 4)]
    cav = meshgen.ldc2d_quad(12, graded=True)
    meshes = [ch[0], cav, ch[1]]
    mesh, em = meshgen.disjoint_union(meshes)
    mu = np.array([0.01, 0.01, 0.01])
    k = 4
    opts = sfem.solver_opts("bcgs", "ilu", 1e-8, 3000, reorder=2)
    ctx = sfem.Context(mesh)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    ctx.set_params(ED, 2, 0.0, 0.1)
    ctx.set_tuning_mode(sfem.TUNE_FIXED)
    ctx.set_batch(em, mu)
    ctx.set_coarse_correction(k, True)
    ctx.setup_solver(opts)
    ctx.step_batch(0.1, opts, 1e-8, 10)
    info = ctx.coarse_info()
    assert info["built"] and not info["singular"]
    P = ctx.coarse_inverse()
    ctx.close()
    off = 0
    for j, m in enumerate(meshes):
        c1 = _ctx(m)
        c1.time_update(0.1)
        c1.assemble(0)
        A, _ = _system(c1)
        Z, _, Eid = _coarse_matrices(m, A, k)
        nc = Z.shape[1]
        ref = _pinv_sym(Eid) if j != 1 else np.linalg.inv(Eid)
        blk = P[off:off + nc, off:off + nc]
        e = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        print(f"[coarse batch float] mesh {j}: {nc} unknowns, rel diff {e:.2e}", flush=True)
        assert e <= 1e-8
        assert np.all(P[off:off + nc, :off] == 0) and np.all(P[off:off + nc, off + nc:] == 0)
        off += nc
        c1.close()


def test_coarse_channel_same_answer_fewer_iterations():
    """The channel at Re 100, 4 steps from rest (inflow ramped), FP64 BiCGStab + ILU(0)
    stopped at ||r|| <= max(1e-12 ||b||, 5e-11), Newton to ||F|| < 1e-10, with and without
    the multiplicative correction on 6^2 boxes: velocity within 1e-8 and pressure within
    1e-6 (relative L2), and fewer Krylov iterations with it."""
    mesh = _channel()
    ctl = deck.Control(np.array([1.0, 0.01, 0, 0, 0, 0, 0, 0, 1, 1, 0]), 2, 0.0, 0.1, 1e9, 4, 1, 1e-6)
    sols, its = [], []
    for k in (0, 6):
        print(f"[case] channel, coarse boxes {k}", flush=True)
        fem = sfem.StabFEM()
        fem.setMesh(mesh)
        fem.setControl(ctl)
        fem.newton_tol = 1e-10
        fem.tuning = sfem.TUNE_FIXED
        fem.coarse_boxes = k
        fem.coarse_mult = True
        fem.prepareInputData()
        fem.setSolver(sfem.solver_opts("bcgs", "ilu", 1e-12, 5000, reorder=3, atol=5e-11))
        sols.append(fem.solveFullyImplicit().reshape(-1, 3))
        its.append(sum(h[3] for h in fem.history))
        assert all(h[2] < 1e-10 for h in fem.history if h[3] == 0)
        fem.ctx.close()
    (a, b) = sols
    eu = np.linalg.norm(b[:, :2] - a[:, :2]) / np.linalg.norm(a[:, :2])
    ep = np.linalg.norm(b[:, 2] - a[:, 2]) / np.linalg.norm(a[:, 2])
    print(f"[coarse channel] its {its[0]} -> {its[1]}; rel L2 u {eu:.2e} p {ep:.2e}", flush=True)
    assert eu <= 1e-8 and ep <= 1e-6
    assert its[1] < its[0]


def test_coarse_rejections():
    """Too many boxes and boxes narrower than an element are refused."""
    mesh = meshgen.ldc3d_hex(4, graded=True)
    ctx = _ctx(mesh)
    with pytest.raises(sfem.SfemError):
        ctx.set_coarse_correction(11)  # 1331 boxes > 1024
    with pytest.raises(sfem.SfemError):
        ctx.set_coarse_correction(8)   # boxes of 1/8 < the 4^3 mesh's elements
    ctx.set_coarse_correction(2)
    ctx.set_coarse_correction(0)
    assert ctx.coarse_info()["k"] == 0
    ctx.close()


def _batch(count, seed, n_range, prec, opts, dt, k, mult):
    mesh, em, mu, ns, meshes = meshgen.c4_batch(count, seed=seed, n_range=n_range)
    ed = np.array([1.0, 0.01, 0, 0, 0, 0, 0, 0, 1, 1, 0])
    ctx = sfem.Context(mesh)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    ctx.set_params(ed, 2, 0.0, dt)
    ctx.set_precision(prec)
    ctx.set_tuning_mode(sfem.TUNE_FIXED)
    ctx.set_batch(em, mu)
    ctx.set_coarse_correction(k, mult)
    ctx.setup_solver(opts)
    return ctx, mesh, em, mu, ns, meshes


def test_coarse_batch_block_inverse():
    """A batch gets k^2 boxes per mesh on the mesh's own bounding box: E is block diagonal,
    one block per mesh, each equal to that mesh's own E (numpy on the exported tangent of
    the mesh alone) within 1e-8 relative in its inverse."""
    opts = sfem.solver_opts("bcgs", "ilu", 1e-8, 3000, reorder=2)
    ctx, mesh, em, mu, ns, meshes = _batch(3, 7, (6, 14), sfem.PREC_FP64, opts, 0.05, 3, True)
    ctx.step_batch(0.05, opts, 1e-8, 10)
    Einv = ctx.coarse_inverse()
    info = ctx.coarse_info()
    print(f"[coarse batch] sizes {list(ns)} unknowns {info['unknowns']}", flush=True)
    assert info["built"] and not info["singular"]
    ctx.close()
    off = 0
    for k, m in enumerate(meshes):
        ed = np.array([1.0, mu[k], 0, 0, 0, 0, 0, 0, 1, 1, 0])
        c1 = sfem.Context(m)
        c1.set_dirichlet(m.dbc_node, m.dbc_dof, m.dbc_value)
        c1.prepare_elements()
        c1.build_pattern()
        c1.set_params(ed, 2, 0.0, 0.05)
        c1.time_update(0.05)
        c1.assemble(0)
        A, _ = _system(c1)
        Z, _, Eid = _coarse_matrices(m, A, 3)
        nc = Z.shape[1]
        blk = Einv[off:off + nc, off:off + nc]
        # the batch's block comes from the batch's first tangent (zero state, as here)
        ref = np.linalg.inv(Eid)
        e = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        print(f"[coarse batch] mesh {k}: {nc} unknowns, rel diff {e:.2e}", flush=True)
        assert e <= 1e-8
        assert np.all(Einv[off:off + nc, :off] == 0) and np.all(Einv[off:off + nc, off + nc:] == 0)
        off += nc
        c1.close()


@pytest.mark.parametrize("prec", [sfem.PREC_FP64, sfem.PREC_STORE32, sfem.PREC_MIXED])
def test_coarse_batch_matches_single_meshes(prec):
    """A batch of jittered Tria3 cavities with the multiplicative coarse correction (4^2
    boxes per mesh), stepped together, against each mesh stepped alone through the FP64
    single-mesh path without it, both to ||F|| < 1e-10: within 1e-7 (FP64 Krylov) or 5e-7
    (FP32), the bars of the batch test without the correction."""
    dt, steps, ntol = 0.05, 3, 1e-10
    rtol = 1e-10 if prec == sfem.PREC_FP64 else 1e-4
    opts = sfem.solver_opts("bcgs", "ilu", rtol, 3000, reorder=2)
    ctx, mesh, em, mu, ns, meshes = _batch(6, 3, (8, 48), prec, opts, dt, 4, True)
    t = dt
    tot = 0
    for s in range(steps):
        ni, ki, rn = ctx.step_batch(min(t, 1.0), opts, ntol, 10)
        print(f"[coarse batch] prec={prec} step {s}: newton {ni.tolist()} krylov {ki.tolist()}", flush=True)
        assert np.all(rn < ntol)
        tot += int(np.sum(ki))
        t += dt
    sb = ctx.get_solution().reshape(-1, 3)
    ctx.close()
    off = 0
    o64 = sfem.solver_opts("bcgs", "ilu", 1e-10, 3000, reorder=2)
    for k, m in enumerate(meshes):
        ed = np.array([1.0, mu[k], 0, 0, 0, 0, 0, 0, 1, 1, 0])
        c1 = sfem.Context(m)
        c1.set_dirichlet(m.dbc_node, m.dbc_dof, m.dbc_value)
        c1.prepare_elements()
        c1.build_pattern()
        c1.set_params(ed, 2, 0.0, dt)
        t = dt
        for s in range(steps):
            st = c1.step(min(t, 1.0), o64, ntol, 10)
            assert st.rhs_norm < ntol
            t += dt
        s1 = c1.get_solution().reshape(-1, 3)
        c1.close()
        d = np.linalg.norm(sb[off:off + m.nNode] - s1) / np.linalg.norm(s1)
        print(f"[coarse batch] mesh {k}: rel diff vs FP64 alone {d:.2e}", flush=True)
        assert d <= (1e-7 if prec == sfem.PREC_FP64 else 5e-7)
        off += m.nNode
