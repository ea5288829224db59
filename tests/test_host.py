"""Host-side logic on the CPU: the C-ABI library loads and exports every entry
point of include/sfem.h; deck readers/writers; mesh generators against the
reference deck; solver-option mapping."""
import ctypes
import os
import re

import numpy as np
import pytest

from cfdsfemmpi_amd import deck, meshgen, sfem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "sfem.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sfem_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(sfem.LIB_PATH)
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_binding_table_matches_header():
    assert sorted(n for n, _, _ in sfem.ABI) == header_functions()


def test_version_and_error_strings():
    lib = sfem.load_library()
    assert lib.sfem_version().decode().startswith("sfem-mi355x")
    assert isinstance(lib.sfem_last_error(), bytes)


def test_create_rejects_bad_mesh_without_touching_device():
    lib = sfem.load_library()
    coords = np.zeros((4, 2))
    conn = np.array([[0, 1, 2, 7]], np.int32)   # node 7 does not exist
    h = lib.sfem_create(2, 4, 4, 1, coords.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        conn.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), -1)
    assert not h
    assert "connectivity" in lib.sfem_last_error().decode()
    h = lib.sfem_create(2, 5, 4, 1, coords.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        conn.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), -1)
    assert not h and "unsupported" in lib.sfem_last_error().decode()


def test_mesh_roundtrip(tmp_path):
    m = meshgen.ldc3d_hex(3)
    p = str(tmp_path / "m.dat")
    deck.write_mesh(p, m)
    r = deck.read_mesh(p)
    assert np.array_equal(r.coords, m.coords) and np.array_equal(r.conn, m.conn)
    assert np.array_equal(r.dbc_node, m.dbc_node) and np.array_equal(r.dbc_dof, m.dbc_dof)
    assert np.array_equal(r.dbc_value, m.dbc_value)


def test_control_roundtrip(tmp_path, golden):
    c = deck.read_control(os.path.join(golden, "ldc2dquad10", "control-parameters-ldc-Re100.dat"))
    assert c.elemData[0] == 1.0 and c.elemData[1] == 0.01 and list(c.elemData[8:11]) == [1.0, 1.0, 0.0]
    assert (c.tis, c.rhoInf, c.dt, c.timeFinal, c.stepsMax) == (0, 0.0, 0.1, 1.0, 10)
    p = str(tmp_path / "c.dat")
    deck.write_control(p, c)
    c2 = deck.read_control(p)
    assert np.array_equal(c.elemData, c2.elemData) and c.tis == c2.tis and c.dt == c2.dt


def test_generator_matches_reference_deck(golden):
    """The 10x10 generator with the deck's pressure pin reproduces LDC2Dquad10.dat."""
    ref = deck.read_mesh(os.path.join(golden, "ldc2dquad10", "LDC2Dquad10.dat"))
    g = meshgen.ldc2d_quad(10, p_pin=(0.5, 0.0))
    assert np.max(np.abs(g.coords - ref.coords)) < 1e-6    # the deck prints 6 decimals
    assert np.array_equal(g.conn, ref.conn)
    key = lambda m: sorted(zip(m.dbc_node.tolist(), m.dbc_dof.tolist(), np.abs(m.dbc_value).tolist()))
    assert key(g) == key(ref)


def test_solver_options_mapping(golden):
    so = deck.read_solver_options(os.path.join(golden, "ldc2dquad10", "petsc_options.dat"))
    assert (so.ksp_type, so.pc_type) == ("preonly", "lu")
    o = sfem.solver_opts(so.ksp_type, so.pc_type, so.rtol, so.max_it)
    assert o.ksp_type == sfem.KSP["preonly"] and o.pc_type == sfem.PC["lu"]


def test_duplicate_dirichlet_entries_in_mesh2(golden):
    """LDC2Dquad-mesh2 lists 8 DOFs twice (all with value 0); the reader keeps them."""
    m = deck.read_mesh(os.path.join(golden, "ldc2dquad-mesh2", "LDC2Dquad-mesh2.dat.gz"))
    assert m.nDBC == 1033
    keys = m.dbc_node.astype(np.int64) * 3 + m.dbc_dof
    assert len(np.unique(keys)) == 1025


def test_check_result_rule():
    ref = np.zeros((3, 3))
    s = np.zeros(9)
    assert deck.check_result(s, ref, 2) == 0
    s[4] = 2e-4
    assert deck.check_result(s, ref, 2) == -2
    assert deck.check_result(np.zeros(6), ref, 2) == -1


def test_generators_geometry():
    h = meshgen.ldc3d_hex(4)
    assert h.nNode == 125 and h.nElem == 64
    x = h.coords[:5, 0]
    assert np.allclose(x, 0.5 * (1 - np.cos(np.pi * np.arange(5) / 4)))
    t = meshgen.ldc3d_tet(2)
    xx = t.coords[t.conn]
    vol = np.einsum("ij,ij->i", np.cross(xx[:, 1] - xx[:, 0], xx[:, 2] - xx[:, 0]), xx[:, 3] - xx[:, 0]) / 6
    assert np.all(vol > 0) and np.isclose(vol.sum(), 1.0)
    tr = meshgen.ldc2d_tria(5, jitter=0.2, seed=4)
    xx = tr.coords[tr.conn]
    area = 0.5 * ((xx[:, 1, 0] - xx[:, 0, 0]) * (xx[:, 2, 1] - xx[:, 0, 1]) -
                  (xx[:, 2, 0] - xx[:, 0, 0]) * (xx[:, 1, 1] - xx[:, 0, 1]))
    assert np.all(area > 0) and np.isclose(area.sum(), 1.0)


def test_cylinder_channel_mesh():
    """The cylinder-in-channel Quad4 mesh: counts of the default 64 x 256 O-grid, every
    quad convex and counter-clockwise, the area equal to the channel minus the polygonal
    cylinder, each node used, and the boundary conditions on the right node sets."""
    m = meshgen.cylinder_channel()
    assert m.nElem == 64 * 256 + (24 + 64 + 150) * (24 + 64 + 24) - 64 * 64
    assert 38000 <= m.nElem <= 42000
    assert np.array_equal(np.unique(m.conn), np.arange(m.nNode))
    xx = m.coords[m.conn]
    cross = [(xx[:, (k + 1) % 4, 0] - xx[:, k, 0]) * (xx[:, (k + 2) % 4, 1] - xx[:, (k + 1) % 4, 1])
             - (xx[:, (k + 1) % 4, 1] - xx[:, k, 1]) * (xx[:, (k + 2) % 4, 0] - xx[:, (k + 1) % 4, 0])
             for k in range(4)]
    assert np.all(np.array(cross) > 0)
    area = 0.5 * np.sum(xx[:, :, 0] * np.roll(xx[:, :, 1], -1, axis=1) - np.roll(xx[:, :, 0], -1, axis=1) * xx[:, :, 1])
    c = m.coords
    r = np.hypot(c[:, 0], c[:, 1])
    cyl = np.isclose(r, 0.5)
    ring = c[cyl][np.argsort(np.arctan2(c[cyl, 1], c[cyl, 0]))]  # the inscribed polygon of the cylinder
    polygon = 0.5 * np.sum(ring[:, 0] * np.roll(ring[:, 1], -1) - np.roll(ring[:, 0], -1) * ring[:, 1])
    assert 0.785 < polygon < np.pi / 4
    assert np.isclose(area, 40.0 * 20.0 - polygon, rtol=1e-12)
    inflow = np.isclose(c[:, 0], -10.0)
    walls = (np.isclose(c[:, 1], -10.0) | np.isclose(c[:, 1], 10.0)) & ~inflow
    assert cyl.sum() == 256 and inflow.sum() == 24 + 64 + 24 + 1
    fixed = {(int(n), int(d)): v for n, d, v in zip(m.dbc_node, m.dbc_dof, m.dbc_value)}
    assert len(fixed) == m.nDBC == 2 * (cyl.sum() + inflow.sum()) + walls.sum()
    assert all(fixed[(n, 0)] == 1.0 and fixed[(n, 1)] == 0.0 for n in np.where(inflow)[0])
    assert all(fixed[(n, 0)] == 0.0 and fixed[(n, 1)] == 0.0 for n in np.where(cyl)[0])
    assert all((n, 0) not in fixed and fixed[(n, 1)] == 0.0 for n in np.where(walls)[0])
    assert not any(d == 2 for _, d in fixed)
    # the seeded inlet perturbation: relative amplitude exactly 1% at its peak, same seed same profile
    p1 = meshgen.cylinder_channel(perturb=0.01, seed=3)
    u_in = np.array([v for n, d, v in zip(p1.dbc_node, p1.dbc_dof, p1.dbc_value) if d == 0 and inflow[n]])
    assert 0.99 - 1e-12 <= u_in.min() and u_in.max() <= 1.01 + 1e-12 and np.ptp(u_in) > 0.005
    assert np.array_equal(p1.dbc_value, meshgen.cylinder_channel(perturb=0.01, seed=3).dbc_value)


def test_disjoint_union_of_meshes():
    """Meshes side by side: node numbers of mesh k offset by the nodes before it, its
    elements and boundary conditions carried over in order, one mesh id per element."""
    ms = [meshgen.ldc2d_tria(n, jitter=0.2, seed=n) for n in (3, 5, 4)]
    u, em = meshgen.disjoint_union(ms)
    off = np.cumsum([0] + [m.nNode for m in ms])
    assert u.nNode == off[-1] and u.nElem == sum(m.nElem for m in ms) and u.npElem == 3
    e0 = 0
    for k, m in enumerate(ms):
        assert np.array_equal(u.coords[off[k]:off[k + 1]], m.coords)
        assert np.array_equal(u.conn[e0:e0 + m.nElem], m.conn + off[k])
        assert np.all(em[e0:e0 + m.nElem] == k)
        e0 += m.nElem
    assert np.array_equal(u.dbc_node, np.concatenate([m.dbc_node + off[k] for k, m in enumerate(ms)]))
    assert np.array_equal(u.dbc_value, np.concatenate([m.dbc_value for m in ms]))
    b, bem, _, ns, bms = meshgen.c4_batch(3, seed=2, n_range=(4, 6))
    b2, bem2 = meshgen.disjoint_union(bms)
    assert np.array_equal(b.conn, b2.conn) and np.array_equal(bem, bem2)


def test_forces_file_format(tmp_path):
    """The forces file: one 'timeNow Fx Fy' line per step in fixed notation with 14
    decimals (StabFEM.cpp:257-259, StabFEM2.cpp:815)."""
    fem = sfem.StabFEM()
    fem.forces = [(0.1, 0.0, 0.0), (0.2, 0.0, 0.0)]
    fn = tmp_path / "m-forces.dat"
    fem.writeForces(str(fn))
    assert fn.read_text() == "0.10000000000000\t0.00000000000000\t0.00000000000000\n" \
                             "0.20000000000000\t0.00000000000000\t0.00000000000000\n"


def test_vtu_writer_round_trip(tmp_path):
    """VTU content of StabFEM::postProcess (StabFEM3.cpp:6-160): points with z = 0 in 2D,
    one cell per element with its VTK type, pressure and velocity point arrays."""
    from cfdsfemmpi_amd import vtu
    for mesh, ctype in ((meshgen.ldc2d_quad(3), 9), (meshgen.ldc2d_tria(3), 5), (meshgen.ldc3d_hex(2), 12),
                        (meshgen.ldc3d_tet(2), 10)):
        nd = mesh.ndim + 1
        rng = np.random.default_rng(3)
        soln = rng.uniform(-1, 1, mesh.nNode * nd)
        w = vtu.VTUWriter(mesh.coords, mesh.conn, mesh.ndim, asynchronous=True)
        fn = tmp_path / f"m{mesh.ndim}{mesh.npElem}-000000.vtu"
        w.submit(str(fn), soln)
        w.close()
        a = vtu.read_vtu_arrays(str(fn))
        s = soln.reshape(-1, nd)
        assert np.allclose(a["Points/points"][:, :mesh.ndim], mesh.coords, rtol=1e-6, atol=1e-7)
        assert np.all(a["Points/points"][:, mesh.ndim:] == 0)
        assert np.array_equal(a["Cells/connectivity"], mesh.conn.reshape(-1))
        assert np.array_equal(a["Cells/offsets"], np.arange(1, mesh.nElem + 1) * mesh.npElem)
        assert np.all(a["Cells/types"] == ctype)
        assert np.allclose(a["PointData/pressure"], s[:, mesh.ndim], rtol=1e-6, atol=1e-7)
        assert np.allclose(a["PointData/velocity"][:, :mesh.ndim], s[:, :mesh.ndim], rtol=1e-6, atol=1e-7)
        assert np.all(a["PointData/procId"] == 0) and np.all(a["CellData/procId"] == 0)
