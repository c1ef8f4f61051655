"""The one-sweep Jacobi CG iteration (pcg_fused) against the two-sweep iteration and the oracle.

The fused sweep finishes iteration k and starts k + 1 in one pass (beta from the previous sweep's
dot products), so its rounding differs from the two-sweep iteration's: both must reach the
tolerance, project the velocity to the same field within the tolerance's reach, and take nearly
the same number of iterations.  Cases: obstacles, a non-cubic box with odd edges (partial tiles
in y and z), chunk lengths of 1, 3 and the sized default (tile halo planes at every chunk edge),
and the oracle's PCG on a 16^3 box.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(nx + 1, ny, nz)).astype(np.float32); u[0] = u[-1] = 0
    v = rng.normal(size=(nx, ny + 1, nz)).astype(np.float32); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(nx, ny, nz + 1)).astype(np.float32); w[:, :, 0] = w[:, :, -1] = 0
    return u, v, w


def _mask(shape, seed):
    # a few solid balls in a box of any shape (cell centres in units of the largest edge)
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    h = 1.0 / max(shape)
    X, Y, Z = np.meshgrid((np.arange(nx) + 0.5) * h, (np.arange(ny) + 0.5) * h, (np.arange(nz) + 0.5) * h,
                          indexing="ij")
    m = np.zeros(shape, dtype=np.uint8)
    for _ in range(3):
        c = rng.uniform(0.2, 0.8, 3) * np.array(shape) * h
        r = rng.uniform(0.05, 0.12)
        m |= (((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) <= r * r).astype(np.uint8)
    return m


def _solve(fia, c, shape, fields, solid, tol):
    import torch
    nx, ny, nz = shape
    u, v, w = [torch.tensor(a, device="cuda:0") for a in fields]
    g = fia.Grid3D(nx, ny, nz, 1.0 / max(shape), None, u, v, w)
    g.density = None
    sd = torch.tensor(solid, device="cuda:0") if solid is not None else None
    it, res = fia.calc_pressure_gradient_term_3d(c, g, solid=sd, rel_tol=tol)
    torch.cuda.synchronize()
    return it, res, [t.double().cpu().numpy() for t in (g.u, g.v, g.w)]


@pytest.mark.parametrize("shape,chunk,with_solid", [((40, 40, 40), 0, True), ((40, 40, 40), 1, True),
                                                    ((17, 23, 31), 3, True), ((17, 23, 31), 0, False),
                                                    ((64, 64, 64), 0, False), ((33, 70, 19), 5, True)])
def test_fused_sweep_matches_two_sweeps(fia, shape, chunk, with_solid):
    print("case fused CG", shape, "chunk", chunk, "solid", with_solid, flush=True)
    fields = _field(shape, 11)
    solid = _mask(shape, 4) if with_solid else None
    c = fia.Context(0)
    try:
        c.set_pressure_preconditioner("jacobi")
        out = {}
        for fused in (0, 1):
            c.set_tuning(shape, pcg_chunk_x=chunk, pcg_fused=fused)
            out[fused] = _solve(fia, c, shape, fields, solid, 1e-8)
        (i0, r0, f0), (i1, r1, f1) = out[0], out[1]
        print("iterations two-sweep %d fused %d, residuals %.2e %.2e" % (i0, i1, r0, r1), flush=True)
        assert r0 <= 1e-8 and r1 <= 1e-8
        assert abs(i1 - i0) <= max(2, i0 // 50)
        umax = max(np.abs(a).max() for a in f0)
        for a, b in zip(f0, f1):
            assert np.abs(a - b).max() <= 1e-5 * umax
    finally:
        c.close()


def test_fused_sweep_matches_the_oracle(fia, oracle, scenes):
    print("case fused CG vs oracle 16^3", flush=True)
    n = 16
    shape = (n, n, n)
    fields = _field(shape, 7)
    solid = scenes.spherical_obstacles(n, count=3, seed=2)
    c = fia.Context(0)
    try:
        c.set_pressure_preconditioner("jacobi")
        c.set_tuning(shape, pcg_fused=1)
        it, res, got = _solve(fia, c, shape, fields, solid, 1e-8)
    finally:
        c.close()
    ou, ov, ow, p, oit, ores, rc = oracle.pressure_project(*fields, 1.0 / n, solid=solid, tol=1e-8)
    assert rc == 0 and res <= 1e-8
    for a, ref in zip(got, (ou, ov, ow)):
        assert np.abs(a - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())
