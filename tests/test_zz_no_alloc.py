"""fia_grid_totals_noalloc against fia_grid_totals, through the C ABI: the same sums with the same
kernels (bit-identical), two calls in a row with the second on a smaller grid, no growth of the
context's device scratch, host fields rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device_grid(fia, torch, shape, seed, with_density=True):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    T = lambda s: torch.from_numpy(rng.normal(size=s).astype(np.float32)).to(dev)
    grid = fia.Grid3D(nx, ny, nz, 1.0 / 32, T((nx, ny, nz)), T((nx + 1, ny, nz)), T((nx, ny + 1, nz)),
                      T((nx, ny, nz + 1)))
    if not with_density:
        grid.density = None   # the totals of a grid without density: mass 0
    return grid


def test_noalloc_totals_match_the_allocating_entry_point(fia):
    import torch
    print("case grid totals without allocation", flush=True)
    if fia.device_count() < 1:
        pytest.skip("no GPU visible")
    ctx = fia.Context(0)
    try:
        bytes0 = ctx.device_bytes()
        for shape, seed, dens in (((48, 33, 40), 1, True), ((17, 9, 12), 2, True), ((5, 7, 3), 3, False)):
            grid = _device_grid(fia, torch, shape, seed, dens)
            torch.cuda.synchronize()
            got = fia.grid_totals_noalloc(ctx, grid)
            # the call allocated nothing: the context's scratch did not grow
            assert ctx.device_bytes() == bytes0
            want = fia.grid_totals(ctx, grid)
            assert got == want, (shape, got, want)
            if not dens:
                assert got[0] == 0.0
            # against an FP64 sum on the host
            vol = grid.dx ** 3
            ref = [0.0 if not dens else grid.density.double().sum().item() * vol] + \
                  [a.double().sum().item() * vol for a in (grid.u, grid.v, grid.w)]
            assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
            bytes0 = ctx.device_bytes()
        host = fia.Grid3D(4, 4, 4, 0.25)
        with pytest.raises(fia.FiaError):
            fia.grid_totals_noalloc(ctx, host)
    finally:
        ctx.close()
