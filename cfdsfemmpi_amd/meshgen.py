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
    coords, conn, dn, dd, dv, em = [], [], [], [], [], []
    off = 0
    for k, m in enumerate(meshes):
        coords.append(m.coords)
        conn.append(m.conn + off)
        dn.append(m.dbc_node + off)
        dd.append(m.dbc_dof)
        dv.append(m.dbc_value)
        em.append(np.full(m.nElem, k, np.int32))
        off += m.nNode
    mesh = Mesh(2, 3, np.concatenate(coords), np.concatenate(conn).astype(np.int32),
                np.concatenate(dn).astype(np.int32), np.concatenate(dd).astype(np.int32),
                np.concatenate(dv), np.zeros(0, np.int32))
    return mesh, np.concatenate(em), 1.0 / res, ns, meshes
