"""GPU checks at the benchmark's full size (256^3) through properties that hold at any size.

- Integer translation: with a uniform velocity that moves every vertex by a whole number of cells
  (dt u / h = (2, 1, 3)), a density step maps the field onto itself shifted by those cells, away
  from the walls.  It is exact in FP64 (the oracle gives 0 difference at 24^3 for both cases below,
  zero-velocity correction on): every sample coordinate falls on a node, where WENO6 and linear
  interpolation return the sample, and the density term along y vanishes.  The cases are WENO6
  with non-integral psi on uncut faces (one piece per face), and 1Dy_linear with cell-face-aligned
  colours, whose y faces are cut into 2 x 2 pieces by the one-pass segment cutter and whose piece
  values the cutter stores itself.  The GPU keeps the fields in FP32, so the bar is 1e-6 relative.
- Fractional translation: a density linear in y moved by a uniform velocity of (1.37, 0.61, 2.29)
  cells per step with the WENO6 integral rule (the dominant quadrature kernel, one Gauss point per
  interval) gives the exact average of the shifted profile in the interior: the psi prefix, the
  psi point on the face level and the midpoint of the residual integral are exact for a linear
  profile (the oracle holds it to 1e-12 at 32^3, tests/test_oracle_closed_forms.py).  A velocity
  component linear in y (v = V0 + b y, u and w uniform) likewise gives closed forms: every control
  volume is moved along x and z and stretched in y.
- Uniform velocity: self-advection of a uniform velocity moving a fraction of cells per step
  keeps the interior uniform (exact in FP64 up to rounding).
- Projection: the multigrid-preconditioned solve with obstacles reaches its residual, and the
  divergence it leaves matches that residual.
- Determinism: the benchmark's density step (plume blob, curl-noise velocity, CFL 4, WENO6
  integral, xz-CellFaceAligned) run twice from the same input gives bit-identical density (face
  fluxes are summed in a fixed order and gathered without atomics).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 256
SHIFT = (2, 1, 3)
MARGIN = 8


@pytest.mark.parametrize("interp,split", [("WENO6", "none"), ("1Dy_linear", "xz-CellFaceAligned")])
def test_full_size_integer_translation(fia, ctx, interp, split):
    print("case full-size translation", interp, split, flush=True)
    n, h = N, 1.0 / N
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.5, 1.5, size=(n, n, n)).astype(np.float32)
    u = np.full((n + 1, n, n), SHIFT[0] * h, np.float32)
    v = np.full((n, n + 1, n), SHIFT[1] * h, np.float32)
    w = np.full((n, n, n + 1), SHIFT[2] * h, np.float32)
    grid = fia.Grid3D(n, n, n, h, rho.copy(), u, v, w)
    fia.move_substances_3d(ctx, grid, 0, 1.0, split_method=split, interpolation_method=interp, use_integral=False)
    m, (a, b, c) = MARGIN, SHIFT
    got = grid.density[m:n - m, m:n - m, m:n - m].astype(np.float64)
    want = rho[m - a:n - m - a, m - b:n - m - b, m - c:n - m - c].astype(np.float64)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("relative L-inf against the shifted field %.3e" % err, flush=True)
    assert err <= 1e-6


def test_full_size_fractional_translation_is_exact(fia, ctx):
    print("case full-size fractional translation, WENO6 integral", flush=True)
    n, h = N, 1.0 / N
    s = (1.37, 0.61, 2.29)
    a, b = 0.8, 0.3
    yc = (np.arange(n) + 0.5) * h
    rho = np.ascontiguousarray(np.broadcast_to((a + b * yc)[None, :, None], (n, n, n)), dtype=np.float32)
    u = np.full((n + 1, n, n), s[0] * h, np.float32)
    v = np.full((n, n + 1, n), s[1] * h, np.float32)
    w = np.full((n, n, n + 1), s[2] * h, np.float32)
    grid = fia.Grid3D(n, n, n, h, rho.copy(), u, v, w)
    fia.move_substances_3d(ctx, grid, 0, 1.0, split_method="xz-CellFaceAligned", interpolation_method="WENO6",
                           use_integral=True)
    # the velocity arrays are FP32, so the displacement is s[1] to FP32 rounding of s[1] * h
    dy = float(np.float32(s[1] * h))
    want = a + b * (yc - dy)
    m = MARGIN
    got = grid.density[m:n - m, m:n - m, m:n - m].astype(np.float64)
    err = float(np.abs(got - want[None, m:n - m, None]).max() / np.abs(want).max())
    print("relative L-inf against the shifted linear profile %.3e" % err, flush=True)
    assert err <= 1e-6


def test_full_size_velocity_linear_in_y_is_exact(fia, ctx):
    print("case full-size velocity linear in y, WENO6 integral", flush=True)
    n, h = N, 1.0 / N
    V0, b, U0, W0 = 0.61, 0.5, 1.37, 2.29
    dt = h
    u = np.full((n + 1, n, n), U0, np.float32)
    w = np.full((n, n, n + 1), W0, np.float32)
    yf = np.arange(n + 1) * h
    v = np.ascontiguousarray(np.broadcast_to((V0 + b * yf)[None, :, None], (n, n + 1, n)), dtype=np.float32)
    grid = fia.Grid3D(n, n, n, h, None, u, v, w)
    fia.advect_velocity_flux_advection_3d(ctx, grid, -1, dt, interpolation_method="WENO6")
    back = lambda y: y - dt * (V0 + b * y)
    lo, hi = back((np.arange(n + 1) - 0.5) * h), back((np.arange(n + 1) + 0.5) * h)
    v_want = (V0 * (hi - lo) + 0.5 * b * (hi ** 2 - lo ** 2)) / h
    m = MARGIN
    err_v = float(np.abs(grid.v[m:-m, m:-m, m:-m].astype(np.float64) - v_want[None, m:-m, None]).max() / np.abs(v_want).max())
    stretch = (back(h) - back(0.0)) / h
    err_u = float(np.abs(grid.u[m:-m, m:-m, m:-m].astype(np.float64) - U0 * stretch).max() / U0)
    err_w = float(np.abs(grid.w[m:-m, m:-m, m:-m].astype(np.float64) - W0 * stretch).max() / W0)
    print("relative L-inf: v %.3e, u %.3e, w %.3e (stretch %.6f)" % (err_v, err_u, err_w, stretch), flush=True)
    assert max(err_v, err_u, err_w) <= 1e-5


def test_full_size_density_step_is_deterministic(fia, ctx, scenes):
    print("case full-size determinism", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import bench
    h, rho, u, v, w, dt = bench.make_scene(scenes, N, 4.0)
    out = []
    for _ in range(2):
        grid = fia.Grid3D(N, N, N, h, rho.copy(), u, v, w)
        fia.move_substances_3d(ctx, grid, 0, dt, split_method="xz-CellFaceAligned", interpolation_method="WENO6")
        out.append(grid.density.copy())
    assert np.isfinite(out[0]).all()
    assert np.array_equal(out[0], out[1])


def test_full_size_uniform_velocity_keeps_itself(fia, ctx):
    # velocity self-advection (WENO6 integral, xz-CellFaceAligned, every component) of a uniform
    # field moving (1.37, 0.61, 2.29) cells per step: the interior stays uniform (the oracle keeps
    # it to 5.5e-15 at 24^3); the faces move rigidly and are cut along both colours (2 x 2 pieces
    # on the y faces of every component grid)
    print("case full-size uniform velocity", flush=True)
    n, h = N, 1.0 / N
    s = (1.37, 0.61, 2.29)
    c = [np.float32(x * h) for x in s]
    u = np.full((n + 1, n, n), c[0], np.float32)
    v = np.full((n, n + 1, n), c[1], np.float32)
    w = np.full((n, n, n + 1), c[2], np.float32)
    grid = fia.Grid3D(n, n, n, h, None, u, v, w)
    fia.advect_velocity_flux_advection_3d(ctx, grid, -1, 1.0, interpolation_method="WENO6")
    m = MARGIN
    for name, got, ref in (("u", grid.u, c[0]), ("v", grid.v, c[1]), ("w", grid.w, c[2])):
        err = float(np.abs(got[m:-m, m:-m, m:-m].astype(np.float64) - float(ref)).max() / abs(float(ref)))
        print("%s: relative L-inf against the uniform value %.3e" % (name, err), flush=True)
        assert err <= 1e-5


def test_full_size_projection_removes_divergence(fia, ctx, scenes):
    # the pressure projection at 256^3 with 8 obstacles under the library's default (multigrid)
    # preconditioner: the relative residual reaches 1e-6, and the divergence left in the fluid cells
    # (relative to the divergence before) matches that residual
    print("case full-size projection", flush=True)
    n, h = N, 1.0 / N
    rng = np.random.default_rng(23)
    u = rng.normal(size=(n + 1, n, n)).astype(np.float32); u[0] = u[-1] = 0
    v = rng.normal(size=(n, n + 1, n)).astype(np.float32); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(n, n, n + 1)).astype(np.float32); w[:, :, 0] = w[:, :, -1] = 0
    solid = scenes.spherical_obstacles(n, count=8, seed=5)
    # the solve zeroes the faces that touch an obstacle before it forms the right-hand side
    sb = np.pad(solid.astype(bool), 1, constant_values=False)
    u0, v0, w0 = u.astype(np.float64), v.astype(np.float64), w.astype(np.float64)
    u0[sb[:-1, 1:-1, 1:-1] | sb[1:, 1:-1, 1:-1]] = 0.0
    v0[sb[1:-1, :-1, 1:-1] | sb[1:-1, 1:, 1:-1]] = 0.0
    w0[sb[1:-1, 1:-1, :-1] | sb[1:-1, 1:-1, 1:]] = 0.0
    d0 = scenes.divergence(u0, v0, w0, h) * h
    del u0, v0, w0
    grid = fia.Grid3D(n, n, n, h, None, u, v, w)
    it, res = fia.calc_pressure_gradient_term_3d(ctx, grid, solid=solid, rel_tol=1e-6)
    print("iterations %d, relative residual %.3e" % (it, res), flush=True)
    assert res <= 1e-6 and it < 100
    d1 = scenes.divergence(grid.u.astype(np.float64), grid.v.astype(np.float64), grid.w.astype(np.float64), h) * h
    fl = solid == 0
    d0 = np.where(fl, d0, 0.0); d1 = np.where(fl, d1, 0.0)
    last = np.flatnonzero(fl.reshape(-1))[-1]   # the pinned cell: the last fluid cell
    d0.reshape(-1)[last] = d1.reshape(-1)[last] = 0.0
    ratio = np.linalg.norm(d1) / np.linalg.norm(d0)
    print("divergence norm after / before %.3e" % ratio, flush=True)
    assert abs(ratio - res) < 0.5 * res + 2e-7
