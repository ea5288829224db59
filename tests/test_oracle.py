"""Pins the oracle (CPU restatement) before it is trusted as the GPU checker.

Each check is chosen so that a plausible transcription mistake fails it:
closed forms of the time coefficients and quadrature, finite-difference and
partition-of-unity properties of the shape functions (including the reference's
Hexa8 derivative table), closed-form geometry of affine elements, the affine
invariant Fe(u + d) - Fe(u) = -Ke d, special states with closed-form
residuals, brute-force pattern counts, a direct-solve check of the Krylov
solvers, and the reference's own end-to-end fixtures.
"""
import os

import numpy as np
import pytest

import oracle
from cfdsfemmpi_amd import deck, meshgen

ALL = [oracle.TRIA3, oracle.QUAD4, oracle.TETRA4, oracle.HEXA8]
NPE = {oracle.TRIA3: 3, oracle.QUAD4: 4, oracle.TETRA4: 4, oracle.HEXA8: 8}
NDIM = {oracle.TRIA3: 2, oracle.QUAD4: 2, oracle.TETRA4: 3, oracle.HEXA8: 3}


# ---------------------------------------------------------------- time / quadrature
def test_time_params_closed_forms():
    dt = 0.1
    td = oracle.time_params(2, 0.0, dt)      # generalised alpha, rho_inf = 0
    af, am, g = 1.0, 1.5, 1.0
    assert np.allclose(td[[0, 1, 2, 3, 4]], [dt, am, af, am, g], rtol=0, atol=1e-15)
    assert np.isclose(td[5], af * dt) and np.isclose(td[8], am / g / dt)
    assert np.isclose(td[9], 1 / g / dt) and np.isclose(td[10], -1 / g / dt) and td[15] == 0.0
    td = oracle.time_params(2, 0.5, dt)      # rho_inf = 0.5: af = 2/3, am = 5/6, gamma = 2/3
    af, am = 1 / 1.5, 0.5 * 2.5 / 1.5
    g = 0.5 + am - af
    assert np.allclose(td[[1, 2, 4, 8, 9, 15]], [am, af, g, am / g / dt, 1 / g / dt, 1 - 1 / g])
    td = oracle.time_params(0, 0.0, dt)      # quasi-static: only td0..td4
    assert np.array_equal(td, np.r_[dt, 1, 1, 1, 1, np.zeros(11)])
    td = oracle.time_params(5, 0.0, dt)      # BDF2 weights 3/2, -2, 1/2
    assert np.allclose(td[[9, 10, 11]] * dt, [1.5, -2.0, 0.5])
    td = oracle.time_params(7, 0.0, dt)      # BDF4 weights sum to zero
    assert abs(td[9:14].sum()) < 1e-12


def _iter_step(tis, rho_inf, dt, soln, prev, prev2, prev3, prev4, dot_prev):
    td = oracle.time_params(tis, rho_inf, dt)
    n = soln.size
    dot, cur, dot_cur = np.zeros(n), np.zeros(n), np.zeros(n)
    oracle.lib().orc_update_iter_step(n, td, soln, prev, prev2, prev3, prev4, dot, dot_prev, cur, dot_cur)
    return dot, cur, dot_cur


def test_update_iter_step_generalised_alpha():
    """SolutionData::updateIterStep (SolutionData.cpp:83-91) for gen-alpha against the
    Jansen-Whiting-Hulbert update written in its trapezoidal form,
    u_{n+1} = u_n + dt ((1 - gamma) udot_n + gamma udot_{n+1}),
    u_{n+af} = u_n + af (u_{n+1} - u_n), udot_{n+am} = udot_n + am (udot_{n+1} - udot_n)."""
    rng = np.random.default_rng(7)
    n, dt = 64, 0.1
    soln, prev, dot_prev = (rng.standard_normal(n) for _ in range(3))
    junk = [rng.standard_normal(n) for _ in range(3)]  # older history must not enter
    for rho in (0.0, 0.5, 1.0):
        am, af = 0.5 * (3 - rho) / (1 + rho), 1 / (1 + rho)
        gam = 0.5 + am - af
        dot, cur, dot_cur = _iter_step(2, rho, dt, soln, prev, *junk, dot_prev)
        dot_ref = ((soln - prev) / dt - (1 - gam) * dot_prev) / gam
        assert np.allclose(dot, dot_ref, rtol=1e-13, atol=1e-13)
        assert np.allclose(prev + dt * ((1 - gam) * dot_prev + gam * dot), soln, rtol=1e-13, atol=1e-13)
        assert np.allclose(cur, prev + af * (soln - prev), rtol=1e-14, atol=1e-14)
        assert np.allclose(dot_cur, dot_prev + am * (dot_ref - dot_prev), rtol=1e-13, atol=1e-13)
    # rho_inf = 0 (the configs' setting): af = 1, am = 3/2, gamma = 1
    dot, cur, dot_cur = _iter_step(2, 0.0, dt, soln, prev, *junk, dot_prev)
    assert np.allclose(dot, (soln - prev) / dt, rtol=1e-14, atol=1e-14)
    assert np.array_equal(cur, soln)
    assert np.allclose(dot_cur, 1.5 * dot - 0.5 * dot_prev, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("tis,order", [(3, 1), (5, 2), (6, 3), (7, 4)])
def test_update_iter_step_bdf_exact_on_polynomials(tis, order):
    """BDF-k differentiates a polynomial of degree k in time exactly: with the history
    u(t_{n+1}), u(t_n), ..., the updated velocity is u'(t_{n+1}) for every node; the stage
    values are the new state (af = am = 1)."""
    rng = np.random.default_rng(tis)
    n, dt, t1 = 32, 0.05, 1.3
    c = rng.standard_normal((order + 1, n))
    u = lambda t: sum(c[k] * t ** k for k in range(order + 1))
    du = lambda t: sum(k * c[k] * t ** (k - 1) for k in range(1, order + 1))
    hist = [u(t1 - m * dt) for m in range(5)]  # soln, prev, prev2, prev3, prev4
    dot_prev = rng.standard_normal(n)          # BDF ignores the old velocity
    dot, cur, dot_cur = _iter_step(tis, 0.0, dt, *hist, dot_prev)
    assert np.allclose(dot, du(t1), rtol=1e-10, atol=1e-10)
    assert np.array_equal(cur, hist[0]) and np.array_equal(dot_cur, dot)
    # one degree higher is no longer exact (the history weights are the order-k ones)
    c2 = np.vstack([c, rng.standard_normal((1, n))])
    u2 = lambda t: sum(c2[k] * t ** k for k in range(order + 2))
    du2 = lambda t: sum(k * c2[k] * t ** (k - 1) for k in range(1, order + 2))
    dot2, _, _ = _iter_step(tis, 0.0, dt, *[u2(t1 - m * dt) for m in range(5)], dot_prev)
    assert np.max(np.abs(dot2 - du2(t1))) > 1e-6


def test_time_update_shifts_history():
    """SolutionData::timeUpdate (SolutionData.cpp:64-78): prev4 <- prev3 <- prev2 <- prev <- soln,
    dotPrev <- dot; soln and dot are left as they are."""
    rng = np.random.default_rng(3)
    n = 16
    a = [rng.standard_normal(n) for _ in range(7)]  # soln, prev, prev2, prev3, prev4, dot, dotPrev
    b = [x.copy() for x in a]
    oracle.lib().orc_time_update(n, *b)
    soln, prev, prev2, prev3, prev4, dot, dot_prev = b
    assert np.array_equal(soln, a[0]) and np.array_equal(dot, a[5])
    assert np.array_equal(prev, a[0]) and np.array_equal(prev2, a[1])
    assert np.array_equal(prev3, a[2]) and np.array_equal(prev4, a[3])
    assert np.array_equal(dot_prev, a[5])


def test_gauss_tables():
    gp, gw = oracle.gauss(oracle.HEXA8)
    g = 0.577350269189626
    assert gp.shape == (8, 3) and np.all(gw == 1.0)
    assert np.array_equal(gp[:, 0], [-g, g] * 4)             # x fastest
    assert np.array_equal(gp[:, 1], [-g, -g, g, g] * 2)
    assert np.array_equal(gp[:, 2], [-g] * 4 + [g] * 4)
    gp, gw = oracle.gauss(oracle.QUAD4)
    assert np.array_equal(gp[:, :2], [[-g, -g], [g, -g], [-g, g], [g, g]]) and np.all(gw == 1.0)
    gp, gw = oracle.gauss(oracle.TRIA3)                        # quad 1-point rule: vertex 0
    assert np.array_equal(gp, [[0, 0, 0]]) and np.array_equal(gw, [4.0])
    gp, gw = oracle.gauss(oracle.TETRA4)
    assert np.array_equal(gp, [[0.25, 0.25, 0.25]]) and np.isclose(gw[0], 1 / 6)


# ---------------------------------------------------------------- shape functions
@pytest.mark.parametrize("et", ALL)
def test_shape_partition_and_fd(et):
    rng = np.random.default_rng(1)
    n = NPE[et]
    for _ in range(5):
        xi = rng.uniform(-0.8, 0.8, 3) if et in (oracle.QUAD4, oracle.HEXA8) else rng.uniform(0.05, 0.3, 3)
        N, d1, d2, d3 = oracle.shape(et, xi)
        assert abs(N[:n].sum() - 1.0) < 1e-14
        for d in (d1, d2, d3)[:NDIM[et]]:
            assert abs(d[:n].sum()) < 1e-14
        h = 1e-6
        fd = []
        for k in range(NDIM[et]):
            e = np.zeros(3); e[k] = h
            fd.append((oracle.shape(et, xi + e)[0][:n] - oracle.shape(et, xi - e)[0][:n]) / (2 * h))
        if et != oracle.HEXA8:
            for k, d in enumerate((d1, d2, d3)[:NDIM[et]]):
                assert np.allclose(d[:n], fd[k], atol=1e-9)
        else:
            # dN/dxi1 is the derivative of N; the xi2 and xi3 columns are the
            # reference's table (BasisFunctionsLagrange.cpp:1056-1074), which
            # takes node 3's derivative for node 2 and so on.
            swap = [0, 1, 3, 2, 4, 5, 7, 6]
            assert np.allclose(d1[:8], fd[0], atol=1e-9)
            assert np.allclose(d2[:8], fd[1][swap], atol=1e-9)
            assert np.allclose(d3[:8], fd[2][swap], atol=1e-9)
            assert not np.allclose(d2[:8], fd[1], atol=1e-6)


# ---------------------------------------------------------------- geometry
def _single(et, scale=(1.0, 2.0, 0.5), shear=0.3):
    """One affine element: reference element mapped by x = A xi + b."""
    A = np.diag(scale[:NDIM[et]]).astype(float)
    A[0, 1] = shear
    if et == oracle.HEXA8:
        ref = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    elif et == oracle.QUAD4:
        ref = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float)
    elif et == oracle.TETRA4:
        ref = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    else:
        ref = np.array([[0, 0], [1, 0], [0, 1]], float)
    coords = ref @ A.T + 0.1
    return coords, np.arange(len(ref), dtype=np.int32)[None, :], A


@pytest.mark.parametrize("et", ALL)
def test_prepare_affine_closed_forms(et):
    coords, conn, A = _single(et)
    p = oracle.prepare(et, coords, conn)
    detA = np.linalg.det(A)
    vol_ref = {oracle.HEXA8: 8.0, oracle.QUAD4: 4.0, oracle.TETRA4: 1 / 6, oracle.TRIA3: 0.5}[et]
    vol = p["dvol"].sum()
    if et == oracle.TRIA3:
        assert np.isclose(vol, 8 * (0.5 * detA))            # weight 4 x Jac = 2A -> 8A
    else:
        assert np.isclose(vol, vol_ref * detA)
    if NDIM[et] == 2:
        assert np.isclose(p["charlen"][0], np.sqrt(4 * abs(vol) / np.pi))
    else:
        assert np.isclose(p["charlen"][0], (6 * abs(vol) / np.pi) ** (1 / 3))
    # gradients reproduce linear fields exactly (also with the Hexa8 table)
    ng = oracle.lib().orc_elem_ngp(et)
    n = NPE[et]
    grads = [p["dNdx"], p["dNdy"], p["dNdz"]][:NDIM[et]]
    for g in range(ng):
        for c in range(NDIM[et]):
            for k in range(NDIM[et]):
                val = sum(coords[i, k] * grads[c][g * n + i] for i in range(n))
                assert abs(val - (1.0 if c == k else 0.0)) < 1e-12


def test_negative_jacobian_detected():
    coords, conn, _ = _single(oracle.QUAD4)
    with pytest.raises(ValueError):
        oracle.prepare(oracle.QUAD4, coords, conn[:, [0, 3, 2, 1]])


# ---------------------------------------------------------------- element kernels
def _elem_setup(et, seed=0, tis=2):
    mesh = {oracle.HEXA8: meshgen.ldc3d_hex(2, graded=True), oracle.QUAD4: meshgen.ldc2d_quad(3, graded=True),
            oracle.TETRA4: meshgen.ldc3d_tet(2, graded=True),
            oracle.TRIA3: meshgen.ldc2d_tria(3, jitter=0.2, seed=1)}[et]
    prep = oracle.prepare(et, mesh.coords, mesh.conn)
    td = oracle.time_params(tis, 0.0, 0.1)
    ed = np.array([1.3, 0.02, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0])
    n = mesh.nNode * (mesh.ndim + 1)
    rng = np.random.default_rng(seed)
    return mesh, prep, td, ed, rng, n


@pytest.mark.parametrize("et", ALL)
@pytest.mark.parametrize("tis", [0, 2])
def test_affine_invariant(et, tis):
    """F is affine in soln: Fe(soln + d) - Fe(soln) = -Ke d with solnCur += af d and
    solnDotCur += am td9 d (SolutionData.cpp:83-91), at any fixed solnPrev."""
    mesh, prep, td, ed, rng, n = _elem_setup(et, tis=tis)
    prev, cur, dot = (rng.uniform(-1, 1, n) for _ in range(3))
    d = rng.uniform(-1, 1, n)
    af, am, td9 = td[2], td[1], td[9]
    ndof = mesh.ndim + 1
    for e in range(mesh.nElem):
        K, F0 = oracle.element_KF(et, e, mesh.conn, prep, ed, td, prev, cur, dot)
        _, F1 = oracle.element_KF(et, e, mesh.conn, prep, ed, td, prev, cur + af * d, dot + am * td9 * d)
        gd = (mesh.conn[e][:, None] * ndof + np.arange(ndof)[None, :]).ravel()
        lhs = F1 - F0
        rhs = -K @ d[gd]
        assert np.max(np.abs(lhs - rhs)) <= 1e-12 * max(1.0, np.max(np.abs(rhs))), e


@pytest.mark.parametrize("et", ALL)
def test_zero_state_zero_residual(et):
    mesh, prep, td, ed, rng, n = _elem_setup(et)
    z = np.zeros(n)
    for e in range(mesh.nElem):
        _, F = oracle.element_KF(et, e, mesh.conn, prep, ed, td, z, z, z)
        assert np.all(F == 0.0)


@pytest.mark.parametrize("et", [oracle.TETRA4, oracle.TRIA3])
def test_constant_pressure_closed_form(et):
    """u = 0, p = c: Fe_u(i) = c * sum_gp w dN_i (pressure in the pseudo-stress),
    Fe_p = 0; on a simplex dN_i is the constant textbook gradient."""
    coords, conn, A = _single(et)
    p = oracle.prepare(et, coords, conn)
    td = oracle.time_params(2, 0.0, 0.1)
    ed = np.array([1.0, 0.01, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0])
    nd = NDIM[et] + 1
    n = NPE[et]
    cur = np.zeros(n * nd)
    cur[nd - 1::nd] = 2.5
    prev = np.zeros(n * nd)
    _, F = oracle.element_KF(et, 0, conn, p, ed, td, prev, cur, np.zeros(n * nd))
    # textbook simplex gradients: grad N = inverse-transpose of the edge matrix
    E = (coords[1:] - coords[0]).T
    G = np.linalg.inv(E).T
    gradN = np.vstack([-G.sum(axis=1), G.T])
    vol = p["dvol"].sum()
    F = F.reshape(n, nd)
    assert np.allclose(F[:, :nd - 1], 2.5 * vol * gradN, atol=1e-13)
    assert np.allclose(F[:, nd - 1], 0.0, atol=1e-14)


def _weak_form_numpy(et, e, mesh, prep, ed, td, prev, cur, dot):
    """Second, independently written (vectorised) evaluation of the residual:
    F_u = -sum_gp w [N rho (udot + Gp u + G up - Gp up) + dN . (mu G - p I)] (- SUPG in 2D)
    F_p = s sum_gp w [N div u + tauP dN . (res + grad p)],  s = -1 (Quad4/Tria3/Tetra4), +1 (Hexa8)."""
    D = mesh.ndim
    nd = D + 1
    npe = mesh.npElem
    ng = oracle.lib().orc_elem_ngp(et)
    nodes = mesh.conn[e]
    U = lambda v: v.reshape(-1, nd)[nodes]
    Pv, Cv, Dv = U(prev), U(cur), U(dot)
    rho, mu = ed[0], ed[1]
    h = prep["charlen"][e]
    Gf = 4.0 / h ** 2
    CI = 4.0 if D == 2 else 10.0
    s = 1.0 if et == oracle.HEXA8 else -1.0
    F = np.zeros((npe, nd))
    for g in range(ng):
        sl = slice((e * ng + g) * npe, (e * ng + g + 1) * npe)
        N = prep["N"][sl]
        dN = np.stack([prep["dNdx"][sl], prep["dNdy"][sl], prep["dNdz"][sl]][:D], axis=1)  # (npe, D)
        w = prep["dvol"][e * ng + g]
        up = N @ Pv[:, :D]
        Gp = Pv[:, :D].T @ dN            # Gp[a, b] = d up_a / dx_b
        u = N @ Cv[:, :D]
        G = Cv[:, :D].T @ dN
        p = N @ Cv[:, D]
        gp_ = dN.T @ Cv[:, D]
        ud = N @ Dv[:, :D]
        res = rho * (ud + Gp @ u + G @ up - Gp @ up)
        tau = 1.0 / np.sqrt(Gf * (up @ up) + CI * (mu / rho) ** 2 * D * Gf ** 2) / rho
        sigma = mu * G - p * np.eye(D)
        F[:, :D] -= w * (np.outer(N, res) + dN @ sigma.T)
        if D == 2:
            F[:, :D] -= w * ed[8] * tau * rho * np.outer(dN @ up, res + gp_)
        F[:, D] += s * w * (N * np.trace(G) + ed[9] * tau * dN @ (res + gp_))
    return F.ravel()


@pytest.mark.parametrize("et", ALL)
def test_residual_against_independent_evaluation(et):
    mesh, prep, td, ed, rng, n = _elem_setup(et, seed=4)
    prev, cur, dot = (rng.uniform(-1, 1, n) for _ in range(3))
    for e in range(mesh.nElem):
        _, F = oracle.element_KF(et, e, mesh.conn, prep, ed, td, prev, cur, dot)
        Fn = _weak_form_numpy(et, e, mesh, prep, ed, td, prev, cur, dot)
        assert np.max(np.abs(F - Fn)) <= 1e-12 * np.max(np.abs(Fn)), e


def test_flop_counts_frozen():
    """Counts of the reference arithmetic as written (frozen for the roofline)."""
    counts = {et: oracle.flops_per_element(et, True) for et in ALL}
    fonly = {et: oracle.flops_per_element(et, False) for et in ALL}
    assert counts == {oracle.TRIA3: 930, oracle.QUAD4: 6027, oracle.TETRA4: 2267, oracle.HEXA8: 62723}
    assert fonly == {oracle.TRIA3: 273, oracle.QUAD4: 1355, oracle.TETRA4: 555, oracle.HEXA8: 7939}
    assert fonly[oracle.HEXA8] < counts[oracle.HEXA8] // 5


# ---------------------------------------------------------------- pattern
def _pairs_numpy(mesh, nodeDof, nfree):
    nd = mesh.ndim + 1
    fav = nodeDof.reshape(-1, nd)[mesh.conn].reshape(mesh.nElem, -1)
    r = np.repeat(fav, fav.shape[1], axis=1)
    c = np.tile(fav, (1, fav.shape[1]))
    ok = (r >= 0) & (c >= 0)
    key = np.unique(r[ok].astype(np.int64) * nfree + c[ok])
    return key // nfree, key % nfree


@pytest.mark.parametrize("mk", ["quad10", "quad10_pin_corner", "hex4", "tet3", "tri6"])
def test_pattern_vs_numpy_bruteforce(mk, golden):
    mesh = {"quad10": deck.read_mesh(os.path.join(golden, "ldc2dquad10", "LDC2Dquad10.dat")),
            "quad10_pin_corner": meshgen.ldc2d_quad(10), "hex4": meshgen.ldc3d_hex(4),
            "tet3": meshgen.ldc3d_tet(3), "tri6": meshgen.ldc2d_tria(6, 0.2, 2)}[mk]
    nd = mesh.ndim + 1
    nodeDof, assy, nfree = oracle.dof_numbering(mesh.nNode, nd, mesh.dbc_node, mesh.dbc_dof)
    rp, ci = oracle.pattern(mesh.conn, nd, nodeDof, nfree)
    r, c = _pairs_numpy(mesh, nodeDof, nfree)
    assert rp[-1] == len(r)
    assert np.array_equal(np.repeat(np.arange(nfree), np.diff(rp)), r)
    assert np.array_equal(ci, c)
    if mk == "quad10":
        assert (nfree, rp[-1]) == (282, 6354)   # the reference deck pins p at node 6
    if mk == "quad10_pin_corner":
        assert (nfree, rp[-1]) == (282, 6366)


def test_pattern_counts_32cube():
    """C2 sizes: 32^3 graded hex cavity, 125,309 free DOFs and 12,520,933 nonzeros."""
    mesh = meshgen.ldc3d_hex(32)
    nodeDof, assy, nfree = oracle.dof_numbering(mesh.nNode, 4, mesh.dbc_node, mesh.dbc_dof)
    assert nfree == 125309
    rp = np.zeros(nfree + 1, np.int64)
    nnz = oracle.lib().orc_pattern(mesh.nElem, 8, 4, np.ascontiguousarray(mesh.conn, np.int32), nodeDof,
                                   nfree, rp, None)
    assert nnz == 12520933


# ---------------------------------------------------------------- Krylov
@pytest.mark.parametrize("ksp,pc", [("bcgs", "ilu"), ("bcgs", "bjacobi4"), ("bcgs", "jacobi"),
                                    ("gmres", "ilu"), ("gmres", "bjacobi4")])
def test_oracle_krylov_vs_direct(ksp, pc):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    mesh = meshgen.ldc3d_hex(4)
    ctl = deck.Control(np.array([1.0, 1e-2, 0, 0, 0, 0, 0, 0, 1, 1, 0]), 2, 0.0, 0.1, 1e9, 1, 1, 1e-6)
    f = oracle.OracleFEM(mesh, ctl)
    f.prepare()
    f.td = oracle.time_params(2, 0.0, 0.1)
    f.time_update()
    f.assign_boundary_conditions(1.0)
    f.update_iter_step()
    f.assemble(0)
    A = sp.csr_matrix((f.vals, f.colidx, f.rowptr), shape=(f.nfree, f.nfree))
    xd = spla.spsolve(A.tocsc(), f.rhs)
    blk = oracle.node_blocks(f.nodeDof, 4) if pc == "bjacobi4" else None
    P = oracle.OraclePC(pc, f.rowptr, f.colidx, f.vals, blk)
    assert P.status == 0
    x, rc, its, rr = oracle.krylov(ksp, P, f.rowptr, f.colidx, f.vals, f.rhs, 1e-10, 5000, 40)
    assert rc == 0
    assert np.linalg.norm(f.rhs - A @ x) <= 1.01e-10 * np.linalg.norm(f.rhs)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


def test_ilu0_exact_on_tridiagonal():
    """ILU(0) of a tridiagonal matrix is its exact LU: one application solves."""
    n = 50
    rng = np.random.default_rng(0)
    main = rng.uniform(4, 5, n)
    off = rng.uniform(-1, 1, n - 1)
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, off[i - 1] if i else 0), (i, main[i]), (i + 1, off[i] if i < n - 1 else 0)):
            if 0 <= j < n:
                rows.append(i); cols.append(j); vals.append(v)
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    rp = A.indptr.astype(np.int64); ci = A.indices.astype(np.int32); va = A.data.astype(np.float64)
    P = oracle.OraclePC("ilu", rp, ci, va)
    b = rng.uniform(-1, 1, n)
    x = P.apply(b)
    assert np.allclose(A @ x, b, atol=1e-13)


# ---------------------------------------------------------------- end-to-end fixtures
def test_reference_ldc2dquad10(golden):
    d = os.path.join(golden, "ldc2dquad10")
    m = deck.read_mesh(os.path.join(d, "LDC2Dquad10"))
    c = deck.read_control(os.path.join(d, "control-parameters-ldc-Re100.dat"))
    f = oracle.OracleFEM(m, c)
    f.prepare()
    s = f.solve_fully_implicit()
    ref = deck.read_result(os.path.join(d, "test-ldc-stru-nelem10-ref.dat"))
    assert deck.check_result(s, ref, 2) == 0
    assert np.max(np.linalg.norm(ref - s.reshape(-1, 3), axis=1)) < 2e-8   # 8-decimal fixture


def test_reference_ldc2dquad_mesh2(golden):
    d = os.path.join(golden, "ldc2dquad-mesh2")
    m = deck.read_mesh(os.path.join(d, "LDC2Dquad-mesh2.dat.gz"))
    c = deck.read_control(os.path.join(d, "control-parameters-Re1000.dat"))
    f = oracle.OracleFEM(m, c)
    f.prepare()
    s = f.solve_fully_implicit()
    ref = deck.read_result(os.path.join(d, "test-ldc-stru-mesh2-ref.dat.gz"))
    assert deck.check_result(s, ref, 2) == 0
    assert np.max(np.linalg.norm(ref - s.reshape(-1, 3), axis=1)) < 2e-8


def test_threaded_oracle_matches_sequential():
    """The oracle's parallel paths (element colour classes, level-scheduled ILU(0), the
    node-by-node pattern, chunked dot products) against its sequential paths: identical
    pattern and prepare data, assembly and ILU(0) factor equal to rounding, and the
    level-scheduled sweeps equal to the sequential substitution for the same factor."""
    mesh = meshgen.ldc3d_hex(7, graded=True)
    ctl = deck.Control(np.array([1.0, 1e-3, 0, 0, 0, 0, 0, 0, 1, 1, 0]), 2, 0.0, 0.1, 1e9, 5, 1, 1e-6)
    res = {}
    try:
        for th in (1, 4):
            oracle.set_threads(th)
            f = oracle.OracleFEM(mesh, ctl)
            f.prepare()
            rng = np.random.default_rng(3)
            n = mesh.nNode * 4
            f.soln[:] = rng.uniform(-1, 1, n)
            f.solnPrev[:] = rng.uniform(-1, 1, n)
            f.td = oracle.time_params(2, 0.0, 0.1)
            f.time_update()
            f.assign_boundary_conditions(0.5)
            f.update_iter_step()
            f.assemble(0)
            pc = oracle.OraclePC("ilu", f.rowptr, f.colidx, f.vals)
            r = rng.standard_normal(f.nfree)
            res[th] = (f, pc.lu(len(f.vals)), pc.apply(r))
    finally:
        oracle.set_threads(1)
    (f1, lu1, z1), (f4, lu4, z4) = res[1], res[4]
    assert np.array_equal(f1.rowptr, f4.rowptr) and np.array_equal(f1.colidx, f4.colidx)
    for k in f1.prep:
        assert np.array_equal(f1.prep[k], f4.prep[k])
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))
    assert rel(f4.vals, f1.vals) <= 1e-14 and rel(f4.rhs, f1.rhs) <= 1e-14
    assert rel(lu4, lu1) <= 1e-12
    # the level-scheduled sweeps apply exactly the sequential substitution of their factor
    pc_seq = oracle.OraclePC("ilu", f4.rowptr, f4.colidx, f4.vals)
    assert np.array_equal(pc_seq.lu(len(f4.vals)), lu4)
    assert rel(z4, pc_seq.apply(r)) <= 1e-15 or np.array_equal(z4, pc_seq.apply(r))
