"""Non-cubic grids on the GPU (the reference CLI takes grid_num_x/y/z separately, src/main.cpp:55-57):
the density step and velocity self-advection on boxes of unequal edges, with a smooth
divergence-free curl-noise velocity that has no flow through the walls, against the FP64 oracle.
The same cases run through the kernels' host bodies in tests/test_pass_host.py."""
import numpy as np
import pytest

from conftest import rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _box_scene(scenes, shape, seed, cfl, h):
    nx, ny, nz = shape
    u, v, w = [np.ascontiguousarray(a, dtype=np.float32) for a in scenes.curl_noise(shape, seed=seed)]
    x = (np.arange(nx) + 0.5) * h; y = (np.arange(ny) + 0.5) * h; z = (np.arange(nz) + 0.5) * h
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    rho = np.exp(-((X - 0.4 * nx * h) ** 2 + (Y - 0.45 * ny * h) ** 2 + (Z - 0.5 * nz * h) ** 2) / (2 * 0.15 ** 2))
    dt = cfl * h / max(abs(u).max(), abs(v).max(), abs(w).max())
    return rho.astype(np.float32), u, v, w, dt


@pytest.mark.parametrize("shape,cfl", [((20, 12, 28), 2.0), ((9, 24, 14), 3.0)])
def test_move_substances_box_matches_oracle(fia, oracle, ctx, scenes, shape, cfl):
    print("case move box", shape, cfl, flush=True)
    nx, ny, nz = shape
    h = 1.0 / 32
    rho, u, v, w, dt = _box_scene(scenes, shape, 7, cfl, h)
    grid = fia.Grid3D(nx, ny, nz, h, rho.copy(), u, v, w)
    fia.move_substances_3d(ctx, grid, 0, dt, split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    ref = oracle.move_substances(rho.astype(np.float64), u, v, w, h, dt, split="xz-CellFaceAligned")
    assert rel_linf(grid.density, ref) < TOL


@pytest.mark.parametrize("shape", [(20, 12, 28), (9, 24, 14)])
def test_velocity_box_matches_oracle(fia, oracle, ctx, scenes, shape):
    # every cell must agree with the oracle run in one of the reference's two cutter arithmetics
    # (FP32 as src/facecutter.h, or FP64): corners on the wall rows of pinned vertices lie exactly
    # on cut lines, where the pieces depend on that rounding (tests/test_pass_host.py)
    print("case velocity box", shape, flush=True)
    nx, ny, nz = shape
    h = 1.0 / 32
    _, u, v, w, dt = _box_scene(scenes, shape, 9, 2.0, h)
    grid = fia.Grid3D(nx, ny, nz, h, None, u.copy(), v.copy(), w.copy())
    fia.advect_velocity_flux_advection_3d(ctx, grid, -1, dt, interpolation_method="WENO6")
    refs = [oracle.advect_velocity(u, v, w, h, dt, method="WENO6", cprec=cp) for cp in (0, 1)]
    for a, got in enumerate((grid.u, grid.v, grid.w)):
        got = np.asarray(got, dtype=np.float64)
        err = np.min([np.abs(got - r[a]) for r in refs], axis=0).max() / np.abs(refs[0][a]).max()
        assert err < TOL, (a, err)


@pytest.mark.parametrize("shape,seed,cfl", [((4, 3, 5), 3, 1.5), ((2, 7, 3), 5, 2.0), ((6, 6, 1), 7, 1.0),
                                             ((1, 5, 6), 9, 1.0), ((3, 3, 3), 4, 3.0)])
def test_move_substances_tiny_grid_matches_oracle(fia, oracle, ctx, shape, seed, cfl):
    # grids narrower than the six-sample WENO6 stencil (every stencil clamped) and one cell thick;
    # the same cases run through the kernels' host bodies in tests/test_pass_host.py
    print("case move tiny", shape, cfl, flush=True)
    nx, ny, nz = shape
    h = 1.0 / 16
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0, 1, size=shape).astype(np.float32)
    u = rng.normal(size=(nx + 1, ny, nz)); u[0] = u[-1] = 0
    v = rng.normal(size=(nx, ny + 1, nz)); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(nx, ny, nz + 1)); w[:, :, 0] = w[:, :, -1] = 0
    u, v, w = [np.ascontiguousarray(a, dtype=np.float32) for a in (u, v, w)]
    dt = cfl * h / max(abs(u).max(), abs(v).max(), abs(w).max())
    grid = fia.Grid3D(nx, ny, nz, h, rho.copy(), u, v, w)
    fia.move_substances_3d(ctx, grid, 0, dt, split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    ref = oracle.move_substances(rho.astype(np.float64), u, v, w, h, dt, split="xz-CellFaceAligned")
    assert np.isfinite(grid.density).all()
    assert rel_linf(grid.density, ref) < TOL
