"""The pressure projection on non-cubic and tiny grids (partial column tiles, chunks longer than
the grid, one-cell-thick boxes) under each solver form: Jacobi CG as two sweeps per iteration and
as one (forced), and multigrid-preconditioned CG; against the oracle's FP64 PCG at 1e-8.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(17, 9, 30), (4, 3, 5), (1, 5, 6), (6, 6, 1), (2, 2, 2)]
FORMS = [("jacobi", 0), ("jacobi", 1), ("mg", -1)]


def _velocity(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(nx + 1, ny, nz)).astype(np.float32); u[0] = u[-1] = 0
    v = rng.normal(size=(nx, ny + 1, nz)).astype(np.float32); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(nx, ny, nz + 1)).astype(np.float32); w[:, :, 0] = w[:, :, -1] = 0
    return u, v, w


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precond,fused", FORMS)
def test_projection_shapes_match_oracle(fia, oracle, shape, precond, fused):
    print("case projection", shape, precond, fused, flush=True)
    nx, ny, nz = shape
    h = 1.0 / 16
    u, v, w = _velocity(shape, sum(shape))
    c = fia.Context(0)
    try:
        c.set_pressure_preconditioner(precond)
        if fused >= 0:
            c.set_tuning(shape, pcg_fused=fused)
        grid = fia.Grid3D(nx, ny, nz, h, None, u.copy(), v.copy(), w.copy())
        it, res = fia.calc_pressure_gradient_term_3d(c, grid, rel_tol=1e-8)
    finally:
        c.close()
    assert res <= 1e-8
    ou, ov, ow, p, oit, ores, rc = oracle.pressure_project(u, v, w, h, tol=1e-8)
    assert rc == 0
    for got, ref in ((grid.u, ou), (grid.v, ov), (grid.w, ow)):
        assert np.isfinite(got).all()
        assert np.abs(got - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())
