"""CG iteration batches replayed from captured HIP graphs (fia_ctx_set_graphs) against the same
solves launched kernel by kernel.  The projected velocities, the pressure, the iteration counts
and the residuals must be bit-identical: under Jacobi and multigrid, after a learned start (its
probe of 5 iterations makes every later batch start at an odd index), with an iteration cap that
is not a multiple of the batch length (the solve stops unconverged in a shorter batch), and for
the scenes step with its pressure solves run per scene and batched.  The second solve of a system
of the same shape must replay the graphs captured by the first.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _velocity(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n + 1, n, n)).astype(np.float32); u[0] = u[-1] = 0
    v = rng.normal(size=(n, n + 1, n)).astype(np.float32); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(n, n, n + 1)).astype(np.float32); w[:, :, 0] = w[:, :, -1] = 0
    dev = torch.device("cuda:0")
    return [torch.tensor(a, device=dev) for a in (u, v, w)]


def _project(fia, c, n, seed, solid, graphs, rel_tol=1e-6, max_iters=20000, net=None):
    import ctypes
    import torch
    u, v, w = _velocity(n, seed)
    grid = fia.Grid3D(n, n, n, 1.0 / n, None, u, v, w)
    grid.density = None
    c.set_graphs(graphs)
    if net is None:
        # through the C ABI: a capped solve reports FIA_ERR_NOT_CONVERGED with its count and residual
        g = grid._c(with_density=False)
        grid.pressure = torch.zeros((n, n, n), dtype=torch.float32, device="cuda:0")
        it, res = ctypes.c_int(0), ctypes.c_double(0.0)
        st = fia.lib().fia_pressure_project(c._h, ctypes.byref(g), solid.data_ptr(), grid.pressure.data_ptr(), 0,
                                            float(rel_tol), int(max_iters), ctypes.byref(it), ctypes.byref(res))
        assert fia.STATUS[st] == ("FIA_OK" if max_iters == 20000 else "FIA_ERR_NOT_CONVERGED"), fia.STATUS[st]
        it, res, info = it.value, res.value, None
    else:
        it, res, info = fia.calc_pressure_gradient_term_3d_learned(c, grid, net, solid=solid, rel_tol=rel_tol,
                                                                   max_iters=max_iters)
    torch.cuda.synchronize()
    return (grid.u.clone(), grid.v.clone(), grid.w.clone(), grid.pressure.clone()), it, res, info


def _solid(scenes, n, seed):
    import torch
    s = scenes.spherical_obstacles(n, count=3, seed=seed)
    return torch.tensor(s, dtype=torch.uint8, device="cuda:0")


def _same(a, b):
    return all(x.equal(y) for x, y in zip(a, b))


@pytest.mark.parametrize("precond,n,max_iters", [("jacobi", 48, 20000), ("mg", 64, 20000), ("jacobi", 40, 77),
                                                  ("mg", 40, 11)])
def test_graph_replay_matches_direct_launches(fia, scenes, precond, n, max_iters):
    print("case graphs", precond, n, max_iters, flush=True)
    c = fia.Context(0)
    try:
        c.set_pressure_preconditioner(precond)
        solid = _solid(scenes, n, 4)
        tol = 1e-6 if max_iters == 20000 else 1e-13   # capped: the solve stops unconverged
        ref, it0, res0, _ = _project(fia, c, n, 3, solid, graphs=False, rel_tol=tol, max_iters=max_iters)
        c.reset_stats()
        got, it1, res1, _ = _project(fia, c, n, 3, solid, graphs=True, rel_tol=tol, max_iters=max_iters)
        s1 = c.stats()
        again, it2, res2, _ = _project(fia, c, n, 3, solid, graphs=True, rel_tol=tol, max_iters=max_iters)
        s2 = c.stats()
        print("iterations %d, residual %.3e, captures %d, replays %d then %d" %
              (it0, res0, s1["graph_captures"], s1["graph_replays"], s2["graph_replays"]), flush=True)
        assert (it1, res1) == (it0, res0) and _same(got, ref)
        assert (it2, res2) == (it0, res0) and _same(again, ref)
        if max_iters < 20000:
            assert it0 == max_iters and res0 > tol
        else:
            assert res0 <= 1e-6
        poll = 8 if precond == "mg" else 32
        if it0 > poll:   # batches after the first one came from graphs
            assert s1["graph_captures"] >= 1 and s1["graph_replays"] >= 1
            assert s2["graph_captures"] == s1["graph_captures"]            # the second solve captured nothing
            assert s2["graph_replays"] > s1["graph_replays"]
    finally:
        c.close()


def test_graph_replay_after_learned_start(fia, scenes):
    print("case graphs after the learned start (odd batch starts)", flush=True)
    n = 48
    c = fia.Context(0)
    try:
        c.set_pressure_preconditioner("jacobi")
        net = fia.UNet(c, seed=2)
        solid = _solid(scenes, n, 6)
        ref, it0, res0, info0 = _project(fia, c, n, 8, solid, graphs=False, net=net)
        c.reset_stats()
        got, it1, res1, info1 = _project(fia, c, n, 8, solid, graphs=True, net=net)
        s = c.stats()
        print("iterations %d, residual %.3e, network used %s, replays %d" %
              (it0, res0, info0["used_network"], s["graph_replays"]), flush=True)
        assert (it1, res1) == (it0, res0) and info1 == info0 and _same(got, ref)
        assert res0 <= 1e-6 and s["graph_replays"] >= 1
    finally:
        c.close()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replay_in_scene_steps(fia, batched):
    import importlib
    import torch
    batch = importlib.import_module("flux-interpolated-advection_amd.batch")
    print("case graphs in scene steps, batched", batched, flush=True)
    od = fia.make_opts(split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    ov = fia.make_opts(interpolation_method="WENO6")
    runs = []
    for graphs in (False, True):
        ctxs = [fia.Context(0) for _ in range(3)]
        try:
            for c in ctxs:
                c.set_graphs(graphs)
            sset = batch.SceneSet(6, seed=12, edges=(24, 32, 40))
            for _ in range(2):
                sset.step(ctxs, od, ov, batched=batched)
            torch.cuda.synchronize()
            runs.append((sset, sum(c.stats()["graph_replays"] for c in ctxs)))
        finally:
            for c in ctxs:
                c.close()
    (a, ra), (b, rb) = runs
    print("replays without graphs %d, with %d" % (ra, rb), flush=True)
    assert ra == 0 and rb >= 1
    for k, (x, y) in enumerate(zip(a.tensors, b.tensors)):
        for f in ("rho", "u", "v", "w", "p"):
            assert x[f].equal(y[f]), (k, f)
    for k, (x, y) in enumerate(zip(a.array, b.array)):
        assert (x.steps, x.status, x.pcg_iters, x.pcg_residual) == (y.steps, y.status, y.pcg_iters, y.pcg_residual), k
        assert y.status == 0 and y.pcg_residual <= 1e-6, k
