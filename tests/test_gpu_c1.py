"""The whole simulation loop on the GPU against the oracle's FP64 run (tests/golden/c1_run.npz,
written by tests/golden/make_c1_run.py): 32^3 plain method, 10 steps of CFL-1 time step, buoyancy,
velocity self-advection (WENO6), projection to 1e-8 (the library's default multigrid-preconditioned
CG; the oracle solves the same system with its own CG) and density advection (WENO6,
xz-CellFaceAligned).  The GPU computes the interpolation and quadrature in FP32 and carries the
FP32 fields from step to step, so its difference from the FP64 run grows with the step count.

Bars after 10 steps: density relative L-inf <= 1e-5 and relative L2 <= 1e-5 (the single-step
bar of the density step), total-mass difference <= 1e-6 relative, velocity relative L-inf <= 5e-5
(the single-step velocity tests hold 1e-5; FP32 velocity error adds up over the steps), and the
final velocity divergence-free on every cell but the pinned one.  Velocity self-advection moves the
wall-normal face velocities (as the reference's scheme does; the oracle shows the same), so the
closed box's compatibility condition fails by the net wall flux, which the pinned last cell takes.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_plain_method_32_ten_steps(fia, ctx, scenes):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_c1_run as C1
    ref = np.load(os.path.join(HERE, "golden", "c1_run.npz"))
    rho, u, v, w = [np.ascontiguousarray(a, dtype=np.float32) for a in C1.initial(scenes)]
    n, h = C1.N, 1.0 / C1.N
    grid = fia.Grid3D(n, n, n, h, rho, u, v, w)
    t, frame = 0.0, 0
    for s in range(C1.STEPS):
        print("case c1 step", s, flush=True)
        dt, write = fia.calc_time_step_length_from_cfl_number_3d(ctx, grid, t, frame, 1, 60.0)
        assert abs(dt - ref["dts"][s]) <= 1e-6 * ref["dts"][s]
        grid.v = np.ascontiguousarray(C1.buoyancy(grid.v, grid.density, np.float32(dt)), dtype=np.float32)
        fia.advect_velocity_xyz_flux_advection_3d(ctx, grid, dt, interpolation_method="WENO6")
        it, res = fia.calc_pressure_gradient_term_3d(ctx, grid, rel_tol=C1.TOL)
        assert res <= C1.TOL
        fia.move_substances_3d(ctx, grid, frame, dt, split_method="xz-CellFaceAligned", interpolation_method="WENO6")
        t += dt
        frame += int(write)
    r = ref["rho"]
    d = grid.density.astype(np.float64)
    linf = np.abs(d - r).max() / np.abs(r).max()
    l2 = np.linalg.norm(d - r) / np.linalg.norm(r)
    mass = abs(d.sum() - r.sum()) / r.sum()
    vel = max(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()
              for a, b in ((grid.u, ref["u"]), (grid.v, ref["v"]), (grid.w, ref["w"])))
    div = scenes.divergence(grid.u.astype(np.float64), grid.v.astype(np.float64), grid.w.astype(np.float64), h)
    print("c1 after %d steps: density rel L-inf %.3e, rel L2 %.3e, mass difference %.3e, velocity rel L-inf %.3e, "
          "max |div u| h %.3e" % (C1.STEPS, linf, l2, mass, vel, np.abs(div).max() * h), flush=True)
    assert linf <= 1e-5 and l2 <= 1e-5
    assert mass <= 1e-6
    assert vel <= 5e-5
    # the last projection came after the last velocity advection: divergence-free but for the
    # pinned cell, which holds the net flux through the walls
    free = div.copy().reshape(-1)
    pinned = free[-1]
    free[-1] = 0.0
    assert np.abs(free).max() * h <= 1e-6
    wall = ((grid.u[-1].astype(np.float64).sum() - grid.u[0].astype(np.float64).sum())
            + (grid.v[:, -1].astype(np.float64).sum() - grid.v[:, 0].astype(np.float64).sum())
            + (grid.w[:, :, -1].astype(np.float64).sum() - grid.w[:, :, 0].astype(np.float64).sum()))
    assert abs(pinned * h - wall) <= 1e-6 * max(1.0, abs(wall))
