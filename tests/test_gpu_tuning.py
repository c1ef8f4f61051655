"""Launch shapes per grid shape (fia_ctx_set_tuning) and their autotuner (fia_ctx_autotune).

The face, piece and cell block sizes must leave the density step and the velocity
self-advection bit-identical (every face, piece and cell is one thread's work, summed in a fixed
order).  The CG chunking regroups the FP64 partial sums, so a solve under another chunking must
reach the same tolerance and agree with the default one to it.  The autotuner must leave its grid
untouched, store what it picked for the grid's shape, and pick only valid shapes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plume(fia, scenes, n, seed=3, cfl=3.0):
    import torch
    u, v, w = scenes.curl_noise(n, seed=seed)
    h = 1.0 / n
    rho = np.maximum(scenes.gaussian_blob(n), scenes.sphere_source(n))
    dt = scenes.cfl_dt(u, v, w, h, cfl)
    t = [torch.tensor(a.astype(np.float32), device="cuda:0") for a in (rho, u, v, w)]
    return t, h, dt


def _grid(fia, t, n, h):
    return fia.Grid3D(n, n, n, h, *[x.clone() for x in t])


def _density_step(fia, c, t, n, h, dt, od):
    import ctypes
    import torch
    g = _grid(fia, t, n, h)
    gc = g._c(with_density=True)
    fia._check(fia.lib().fia_move_substances_3d(c._h, ctypes.byref(gc), 0, dt, ctypes.byref(od)))
    torch.cuda.synchronize()
    return g


def test_block_sizes_leave_the_advection_bit_identical(fia, scenes, ctx):
    print("case block sizes, 40^3 CFL 3 density and velocity", flush=True)
    import ctypes
    import torch
    n = 40
    t, h, dt = _plume(fia, scenes, n)
    od = fia.make_opts(split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    ov = fia.make_opts(interpolation_method="WENO6")
    c = fia.Context(0)
    try:
        ref = _density_step(fia, c, t, n, h, dt, od)
        gv = _grid(fia, t, n, h)
        gvc = gv._c(with_density=False)
        fia._check(fia.lib().fia_advect_velocity_3d(c._h, ctypes.byref(gvc), -1, dt, ctypes.byref(ov)))
        torch.cuda.synchronize()
        for shapes in ({"face_block": 64}, {"face_block": 128, "piece_block": 64}, {"piece_block": 128},
                       {"cell_block": 64}, {"face_block": 64, "piece_block": 64, "cell_block": 128}):
            c.set_tuning((n, n, n), **shapes)
            assert c.get_tuning((n, n, n))["face_block"] == shapes.get("face_block", 0)
            got = _density_step(fia, c, t, n, h, dt, od)
            assert got.density.equal(ref.density), shapes
            g2 = _grid(fia, t, n, h)
            g2c = g2._c(with_density=False)
            fia._check(fia.lib().fia_advect_velocity_3d(c._h, ctypes.byref(g2c), -1, dt, ctypes.byref(ov)))
            torch.cuda.synchronize()
            for a, b in ((g2.u, gv.u), (g2.v, gv.v), (g2.w, gv.w)):
                assert a.equal(b), shapes
        # another shape's entry and the default do not apply to this one
        c.set_tuning((n + 1, n, n), face_block=64)
        assert c.get_tuning((n, n, n))["face_block"] == 64   # still the last entry set for (n, n, n)
        c.set_tuning(None, piece_block=128)
        assert c.get_tuning((7, 7, 7))["piece_block"] == 128
    finally:
        c.close()


def test_invalid_launch_shapes_are_rejected(fia, ctx):
    print("case invalid launch shapes", flush=True)
    for bad in ({"face_block": 100}, {"piece_block": 512}, {"pcg_chunk_x": -1}, {"pcg_recompute": 2},
                {"pcg_fused": 2}):
        with pytest.raises(fia.FiaError):
            ctx.set_tuning((16, 16, 16), **bad)
    with pytest.raises(fia.FiaError):
        ctx.set_tuning((16, 0, 16), face_block=64)


@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_cg_chunking_reaches_the_same_solution(fia, scenes, precond):
    print("case CG chunking", precond, flush=True)
    import torch
    n = 64
    rng = np.random.default_rng(5)
    u = rng.normal(size=(n + 1, n, n)).astype(np.float32); u[0] = u[-1] = 0
    v = rng.normal(size=(n, n + 1, n)).astype(np.float32); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(n, n, n + 1)).astype(np.float32); w[:, :, 0] = w[:, :, -1] = 0
    solid = torch.tensor(scenes.spherical_obstacles(n, count=3, seed=9), dtype=torch.uint8, device="cuda:0")
    base = [torch.tensor(a, device="cuda:0") for a in (u, v, w)]
    c = fia.Context(0)
    try:
        c.set_pressure_preconditioner(precond)
        out = []
        for shapes in ({}, {"pcg_chunk_x": 2, "pcg_recompute": 1, "pcg_fused": 0},
                       {"pcg_chunk_x": 16, "pcg_recompute": 0, "pcg_fused": 0},
                       {"pcg_chunk_x": 64, "pcg_recompute": 1, "pcg_fused": 0}, {"pcg_chunk_x": 3, "pcg_fused": 1},
                       {"pcg_chunk_x": 1, "pcg_fused": 1}):
            c.set_tuning((n, n, n), **shapes)
            g = fia.Grid3D(n, n, n, 1.0 / n, None, *[x.clone() for x in base])
            g.density = None
            it, res = fia.calc_pressure_gradient_term_3d(c, g, solid=solid, rel_tol=1e-6)
            torch.cuda.synchronize()
            print(shapes, "iterations", it, "residual %.3e" % res, flush=True)
            assert res <= 1e-6
            out.append((g.u.double(), g.v.double(), g.w.double(), it))
        ref = out[0]
        umax = max(float(x.abs().max()) for x in ref[:3])
        for o in out[1:]:
            for a, b in zip(o[:3], ref[:3]):
                assert float((a - b).abs().max()) < 1e-4 * umax
            assert abs(o[3] - ref[3]) <= max(3, ref[3] // 20)
    finally:
        c.close()


def test_autotune_keeps_the_grid_and_stores_its_choice(fia, scenes):
    print("case autotune 48^3", flush=True)
    import torch
    n = 48
    t, h, dt = _plume(fia, scenes, n, seed=4)
    od = fia.make_opts(split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    solid = torch.tensor(scenes.spherical_obstacles(n, count=2, seed=1), dtype=torch.uint8, device="cuda:0")
    c = fia.Context(0)
    try:
        ref = _density_step(fia, c, t, n, h, dt, od)
        g = _grid(fia, t, n, h)
        before = [x.clone() for x in (g.density, g.u, g.v, g.w)]
        best = c.autotune(g, od, dt, solid=solid, reps=2)
        torch.cuda.synchronize()
        print("picked", best, flush=True)
        for a, b in zip(before, (g.density, g.u, g.v, g.w)):
            assert a.equal(b)
        assert c.get_tuning((n, n, n)) == best
        for k in ("face_block", "piece_block", "cell_block"):
            assert best[k] in (0, 64, 128, 256)
        assert best["pcg_chunk_x"] >= 0 and best["pcg_recompute"] in (-1, 0, 1) and best["pcg_fused"] in (-1, 0, 1)
        # the step under the picked shapes is the same step
        got = _density_step(fia, c, t, n, h, dt, od)
        assert got.density.equal(ref.density)
    finally:
        c.close()
