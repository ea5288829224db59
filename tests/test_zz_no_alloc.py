"""The allocation-free preconditioner apply (sfem_pc_apply_noalloc) against the
allocating one (sfem_pc_apply), called through the C ABI: bit-identical results on two
calls in a row, the second on a smaller system, and sizes beyond the fixed scratch
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from cfdsfemmpi_amd import meshgen, sfem

pytestmark = pytest.mark.gpu


def _ctx(mesh, reorder):
    ctx = sfem.Context(mesh)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    ctx.set_params(np.array([1.0, 1e-2, 0, 0, 0, 0, 0, 0, 1, 1, 0]), 2, 0.0, 0.1)
    ctx.time_update(1.0)
    ctx.assemble(0)
    ctx.solve(sfem.solver_opts("bcgs", "ilu", 1e-6, 5000, 30, reorder))
    return ctx


def _noalloc(ctx, r):
    r = np.ascontiguousarray(r, np.float64)
    z = np.full(ctx.nfree, np.nan)
    rc = ctx.lib.sfem_pc_apply_noalloc(ctx.h, r.ctypes.data_as(C.POINTER(C.c_double)),
                                       z.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, z


@pytest.mark.parametrize("reorder", [0, 3])
def test_noalloc_pc_apply_matches_allocating_apply(reorder):
    rng = np.random.default_rng(17)
    for n in (7, 4):  # the second call on a smaller system
        mesh = meshgen.ldc3d_hex(n, graded=True)
        ctx = _ctx(mesh, reorder)
        print(f"[case] noalloc pc apply reorder={reorder} n={n}^3 nfree={ctx.nfree}", flush=True)
        r = rng.standard_normal(ctx.nfree)
        z_ref = ctx.pc_apply(r)
        rc, z = _noalloc(ctx, r)
        assert rc == 0, ctx.lib.sfem_last_error().decode()
        assert np.array_equal(z, z_ref)
        ctx.close()


def test_noalloc_pc_apply_without_preconditioner_is_rejected():
    mesh = meshgen.ldc3d_hex(3, graded=True)
    ctx = sfem.Context(mesh)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    rc, z = _noalloc(ctx, np.ones(ctx.nfree))
    assert rc == -2 and np.all(np.isnan(z))
    ctx.close()
