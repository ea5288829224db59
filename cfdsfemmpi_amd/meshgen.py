"""Generators of the lid-driven cavity meshes in the reference's deck layout.

Node ids run x-fastest from the origin and elements are counter-clockwise, as in
the reference's ``testdata/ldc2dquad10/LDC2Dquad10.dat`` (nodes :9-129, Quad4
connectivity ``n, n+1, n+1+(nx+1), n+(nx+1)`` at :131).  Hexa8 uses the same
four nodes at z and then at z + 1, which is the node order of the reference's
trilinear basis (BasisFunctionsLagrange.cpp:1044-1051).

Boundary conditions are listed node by node in increasing node id, dof by dof
(u, v[, w]), which is how ``LDC2Dquad10.dat`` lists them; the pressure pin
comes last.  The lid (y = 1 in 2D, z = 1 in 3D) includes its edges and
corners, and every other boundary node is no-slip.
"""
from __future__ import annotations

import numpy as np

from .deck import Mesh


def _cos_grading(n: int) -> np.ndarray:
    i = np.arange(n + 1, dtype=np.float64)
    return 0.5 * (1.0 - np.cos(np.pi * i / n))


def ldc2d_quad(nx: int, ny: int | None = None, graded: bool = False,
               p_pin: tuple = (0.0, 0.0), lid_u: float = 1.0) -> Mesh:
    """Structured Quad4 unit-square cavity with ``nx`` x ``ny`` cells.

    ``p_pin`` is the (x, y) of the node whose pressure is prescribed to 0
    (the reference deck ``LDC2Dquad10.dat`` pins node 6 = (0.5, 0)).
    """
    ny = nx if ny is None else ny
    xs = _cos_grading(nx) if graded else np.linspace(0.0, 1.0, nx + 1)
    ys = _cos_grading(ny) if graded else np.linspace(0.0, 1.0, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n = (j * (nx + 1) + i).ravel()
    conn = np.stack([n, n + 1, n + 1 + (nx + 1), n + (nx + 1)], axis=1).astype(np.int32)
    ii = np.arange(coords.shape[0]) % (nx + 1)
    jj = np.arange(coords.shape[0]) // (nx + 1)
    wall = (ii == 0) | (ii == nx) | (jj == 0) | (jj == ny)
    lid = jj == ny
    dn, dd, dv = [], [], []
    for node in np.nonzero(wall)[0]:
        for d in range(2):
            dn.append(node); dd.append(d)
            dv.append(lid_u if (lid[node] and d == 0) else 0.0)
    pin = int(np.argmin(np.abs(coords[:, 0] - p_pin[0]) + np.abs(coords[:, 1] - p_pin[1])))
    dn.append(pin); dd.append(2); dv.append(0.0)
    return Mesh(2, 4, coords, conn, np.array(dn, np.int32), np.array(dd, np.int32),
                np.array(dv, np.float64), np.zeros(0, np.int32))


def ldc2d_tria(nx: int, jitter: float = 0.0, seed: int = 0, p_pin=(0.0, 0.0),
               lid_u: float = 1.0) -> Mesh:
    """Unit-square cavity of ``nx`` x ``nx`` quads, each split into two Tria3
    (diagonal from node n to n+1+(nx+1)), interior nodes jittered uniformly by
    +-``jitter``*h with a seeded generator."""
    q = ldc2d_quad(nx, nx, p_pin=p_pin, lid_u=lid_u)
    coords = q.coords.copy()
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        h = 1.0 / nx
        ii = np.arange(coords.shape[0]) % (nx + 1)
        jj = np.arange(coords.shape[0]) // (nx + 1)
        interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < nx)
        coords[interior] += rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), 2))
    c = q.conn
    t1 = c[:, [0, 1, 2]]
    t2 = c[:, [0, 2, 3]]
    conn = np.stack([t1, t2], axis=1).reshape(-1, 3).astype(np.int32)
    return Mesh(2, 3, coords, conn, q.dbc_node, q.dbc_dof, q.dbc_value, q.out_nodes)


def ldc3d_hex(n: int, graded: bool = True, p_pin=(0.0, 0.0, 0.0), lid_u: float = 1.0) -> Mesh:
    """Unit-cube cavity of ``n``^3 Hexa8, nodes graded toward every wall by
    x_i = 0.5 (1 - cos(pi i / n)) when ``graded``; lid u = ``lid_u`` on z = 1."""
    xs = _cos_grading(n) if graded else np.linspace(0.0, 1.0, n + 1)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    m = n + 1
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    b = (k * m * m + j * m + i).ravel()
    conn = np.stack([b, b + 1, b + 1 + m, b + m,
                     b + m * m, b + m * m + 1, b + m * m + 1 + m, b + m * m + m],
                    axis=1).astype(np.int32)
    nn = coords.shape[0]
    ii = np.arange(nn) % m
    jj = (np.arange(nn) // m) % m
    kk = np.arange(nn) // (m * m)
    wall = (ii == 0) | (ii == n) | (jj == 0) | (jj == n) | (kk == 0) | (kk == n)
    lid = kk == n
    wn = np.nonzero(wall)[0]
    dn = np.repeat(wn, 3)
    dd = np.tile(np.arange(3), wn.size)
    dv = np.where((dd == 0) & lid[dn], lid_u, 0.0)
    pin = int(np.argmin(np.abs(coords - np.asarray(p_pin)).sum(axis=1)))
    dn = np.append(dn, pin)
    dd = np.append(dd, 3)
    dv = np.append(dv, 0.0)
    return Mesh(3, 8, coords, conn, dn.astype(np.int32), dd.astype(np.int32),
                dv.astype(np.float64), np.zeros(0, np.int32))


def ldc3d_tet(n: int, graded: bool = False, p_pin=(0.0, 0.0, 0.0), lid_u: float = 1.0) -> Mesh:
    """Unit-cube cavity with each hex of :func:`ldc3d_hex` split into 6 Tetra4
    (Kuhn split along the 0-6 diagonal), every tet positively oriented."""
    h = ldc3d_hex(n, graded=graded, p_pin=p_pin, lid_u=lid_u)
    c = h.conn
    # Kuhn subdivision: paths 0 -> 6 through the cube's vertices
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    conn = np.concatenate([c[:, list(t)][:, None, :] for t in tets], axis=1).reshape(-1, 4)
    x = h.coords[conn]
    vol = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
    flip = vol < 0
    conn[flip] = conn[flip][:, [0, 2, 1, 3]]
    return Mesh(3, 4, h.coords, conn.astype(np.int32), h.dbc_node, h.dbc_dof, h.dbc_value,
                h.out_nodes)


def c4_batch(count: int = 256, seed: int = 0, n_range=(16, 160), jitter: float = 0.2, re_range=(50.0, 2000.0),
             lid_u: float = 1.0):
    """Batch of ``count`` independent unit-square cavities: mesh k has n_k x n_k quads
    with n_k drawn uniformly from ``n_range`` (inclusive), each quad split into two
    Tria3, interior nodes jittered by +-``jitter``*h, and Re_k drawn log-uniformly from
    ``re_range`` (mu_k = 1/Re_k with rho = 1).  All draws come from one generator seeded
    with ``seed``.

    Returns ``(mesh, elem_mesh, mu, ns, meshes)``: the disjoint union (nodes, elements
    and boundary conditions mesh by mesh), the mesh of every element, the viscosities,
    the sizes and the individual meshes.
    """
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    res = np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]), size=count))
    seeds = rng.integers(0, 2**31 - 1, size=count)
    meshes = [ldc2d_tria(int(n), jitter=jitter, seed=int(s), lid_u=lid_u) for n, s in zip(ns, seeds)]
    mesh, em = disjoint_union(meshes)
    return mesh, em, 1.0 / res, ns, meshes


def disjoint_union(meshes):
    """One mesh holding ``meshes`` side by side (nodes, elements and boundary conditions
    mesh by mesh, node numbers offset), and the mesh of every element: the input of a
    batched context (``Context.set_batch``)."""
    coords, conn, dn, dd, dv, em = [], [], [], [], [], []
    off = 0
    for k, m in enumerate(meshes):
        assert m.ndim == meshes[0].ndim and m.npElem == meshes[0].npElem
        coords.append(m.coords)
        conn.append(m.conn + off)
        dn.append(m.dbc_node + off)
        dd.append(m.dbc_dof)
        dv.append(m.dbc_value)
        em.append(np.full(m.nElem, k, np.int32))
        off += m.nNode
    m0 = meshes[0]
    mesh = Mesh(m0.ndim, m0.npElem, np.concatenate(coords),
                np.concatenate(conn).astype(np.int32), np.concatenate(dn).astype(np.int32),
                np.concatenate(dd).astype(np.int32), np.concatenate(dv), np.zeros(0, np.int32))
    return mesh, np.concatenate(em)


def _stretch(h0: float, length: float, n: int) -> np.ndarray:
    """n cells over ``length`` growing geometrically from a first cell of ``h0``
    (uniform when n * h0 >= length): offsets 0..length."""
    if n * h0 >= length:
        return np.linspace(0.0, length, n + 1)
    lo, hi = 1.0, 2.0
    while h0 * (hi ** n - 1.0) / (hi - 1.0) < length:
        hi *= 2.0
    for _ in range(200):
        q = 0.5 * (lo + hi)
        if h0 * (q ** n - 1.0) / (q - 1.0) < length:
            lo = q
        else:
            hi = q
    q = 0.5 * (lo + hi)
    x = np.concatenate([[0.0], np.cumsum(h0 * q ** np.arange(n))])
    return x * (length / x[-1])


def cylinder_channel(nr: int = 64, nth: int = 256, ratio: float = 1.05, radius: float = 0.5,
                     core: float = 2.5, box=(-10.0, 30.0, -10.0, 10.0), n_left: int = 24,
                     n_right: int = 150, n_bottom: int = 24, n_top: int = 24, u_in: float = 1.0,
                     perturb: float = 0.0, seed: int = 0) -> Mesh:
    """Flow past a cylinder of radius ``radius`` at the origin in the channel ``box``,
    Quad4.  Near field: an O-grid of ``nr`` radial x ``nth`` circumferential cells
    between the circle and the square [-core, core]^2, radial spacing growing
    geometrically by ``ratio``; each ring point lies on the ray of its square-boundary
    point.  Far field: eight structured blocks around the square, uniform along the
    square's sides (``nth`` / 4 cells each) and stretched geometrically away from it.

    Boundary conditions (all values ramped by the reference's min(t, 1)):
    inflow x = box[0]: u = u_in (1 + perturb xi(y)), v = 0, xi a seeded smooth function
    with max |xi| = 1; top and bottom walls: slip, v = 0; cylinder: u = v = 0; outflow
    x = box[1]: traction free (no entry).  No pressure is pinned.
    """
    assert nth % 4 == 0
    m = nth // 4
    x0, x1, y0, y1 = box
    h0 = 2.0 * core / m
    xl = -core - _stretch(h0, -core - x0, n_left)[::-1]
    xr = core + _stretch(h0, x1 - core, n_right)
    yb = -core - _stretch(h0, -core - y0, n_bottom)[::-1]
    yt = core + _stretch(h0, y1 - core, n_top)
    xm = np.linspace(-core, core, m + 1)
    X = np.concatenate([xl[:-1], xm, xr[1:]])
    Y = np.concatenate([yb[:-1], xm, yt[1:]])
    NX, NY = len(X) - 1, len(Y) - 1
    i0, j0 = n_left, n_bottom  # square = grid [i0, i0+m] x [j0, j0+m]
    gi, gj = np.meshgrid(np.arange(NX + 1), np.arange(NY + 1), indexing="xy")
    gi, gj = gi.ravel(), gj.ravel()
    keep = ~((gi > i0) & (gi < i0 + m) & (gj > j0) & (gj < j0 + m))
    gid = -np.ones((NY + 1) * (NX + 1), np.int64)
    gid[keep] = np.arange(int(keep.sum()))
    grid = lambda i, j: int(gid[j * (NX + 1) + i])
    coords = [np.stack([X[gi[keep]], Y[gj[keep]]], axis=1)]
    nb = int(keep.sum())
    # far-field cells
    ci, cj = np.meshgrid(np.arange(NX), np.arange(NY), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    outside = ~((ci >= i0) & (ci < i0 + m) & (cj >= j0) & (cj < j0 + m))
    ci, cj = ci[outside], cj[outside]
    n00 = gid[cj * (NX + 1) + ci]
    conn = [np.stack([n00, gid[cj * (NX + 1) + ci + 1], gid[(cj + 1) * (NX + 1) + ci + 1],
                      gid[(cj + 1) * (NX + 1) + ci]], axis=1)]
    # square boundary, counter-clockwise from the corner (core, -core)
    sq = []
    for k in range(m):
        sq.append(grid(i0 + m, j0 + k))
    for k in range(m):
        sq.append(grid(i0 + m - k, j0 + m))
    for k in range(m):
        sq.append(grid(i0, j0 + m - k))
    for k in range(m):
        sq.append(grid(i0 + k, j0))
    sq = np.array(sq, np.int64)
    S = coords[0][sq]
    th = np.arctan2(S[:, 1], S[:, 0])
    C = radius * np.stack([np.cos(th), np.sin(th)], axis=1)
    s = (ratio ** np.arange(nr + 1) - 1.0) / (ratio ** nr - 1.0)
    ring = np.empty((nr + 1, nth), np.int64)
    ring[nr] = sq
    for k in range(nr):
        ring[k] = nb + k * nth + np.arange(nth)
        coords.append((1.0 - s[k]) * C + s[k] * S)
    jj = np.arange(nth)
    for k in range(nr):
        conn.append(np.stack([ring[k, jj], ring[k + 1, jj], ring[k + 1, (jj + 1) % nth],
                              ring[k, (jj + 1) % nth]], axis=1))
    coords = np.concatenate(coords)
    conn = np.concatenate(conn).astype(np.int32)
    # boundary conditions
    xy = coords
    tol = 1e-9
    inflow = np.abs(xy[:, 0] - x0) < tol
    walls = (np.abs(xy[:, 1] - y0) < tol) | (np.abs(xy[:, 1] - y1) < tol)
    cyl = np.zeros(len(xy), bool)
    cyl[ring[0]] = True
    if perturb:
        rng = np.random.default_rng(seed)
        modes = np.arange(1, 6)
        amp = rng.standard_normal(len(modes)) / modes
        ph = rng.uniform(0.0, 2.0 * np.pi, len(modes))
        yy = np.linspace(y0, y1, 2001)
        f = lambda y: (amp[None, :] * np.sin(np.pi * modes[None, :] * (y[:, None] - y0) / (y1 - y0)
                                             + ph[None, :])).sum(axis=1)
        scale = 1.0 / np.max(np.abs(f(yy)))
        xi = lambda y: scale * f(y)
    else:
        xi = lambda y: np.zeros_like(y)
    dn, dd, dv = [], [], []
    uin = u_in * (1.0 + perturb * xi(xy[:, 1]))
    for node in range(len(xy)):
        if inflow[node]:
            dn += [node, node]; dd += [0, 1]; dv += [float(uin[node]), 0.0]
        elif cyl[node]:
            dn += [node, node]; dd += [0, 1]; dv += [0.0, 0.0]
        elif walls[node]:
            dn.append(node); dd.append(1); dv.append(0.0)
    return Mesh(2, 4, coords, conn, np.array(dn, np.int32), np.array(dd, np.int32), np.array(dv, np.float64),
                np.zeros(0, np.int32))
