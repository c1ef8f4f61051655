"""Many independent smoke scenes of uneven sizes advanced together (fia_scenes_step, or
fia_scenes_step_batched with the pressure solves of all scenes batched), with autotuners for the
number of concurrent workers and for the launch shapes of each grid size (fia_ctx_autotune).  No
counterpart in the reference, which runs one grid per process.

Each scene is seeded: cubic grid edge drawn from `edges`, a density source sphere (radius 0.06)
in the lower part of the box, buoyancy in [0.5, 2] on the vertical velocity, 0 - max_obstacles
spherical obstacles kept clear of the source, and an integer CFL number in [cfl_min, cfl_max].
"""
import ctypes
import importlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.basename(HERE)


def _pkg():
    root = os.path.dirname(HERE)
    if root not in sys.path:
        sys.path.insert(0, root)
    return importlib.import_module(PKG), importlib.import_module(PKG + ".scenes")


class SceneSet:
    def __init__(self, count, seed=0, edges=(32, 48, 64, 80, 96), max_obstacles=4, cfl_range=(1, 3),
                 with_pressure=True):
        import torch
        fia, scenes = _pkg()
        self.fia = fia
        rng = np.random.default_rng(seed)
        dev = torch.device("cuda:0")
        self.tensors = []
        self.array = (fia.SceneC * count)()
        self.edges = []
        for k in range(count):
            n = int(rng.choice(edges))
            centre = (rng.uniform(0.25, 0.75), rng.uniform(0.08, 0.2), rng.uniform(0.25, 0.75))
            nob = int(rng.integers(0, max_obstacles + 1))
            clear = scenes.sphere_source(n, centre, 0.14)
            solid = scenes.spherical_obstacles(n, count=nob, seed=seed * 1000 + k, keep_clear=clear) if nob else None
            z = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
            t = {"rho": z(n, n, n), "u": z(n + 1, n, n), "v": z(n, n + 1, n), "w": z(n, n, n + 1),
                 "solid": torch.tensor(solid, dtype=torch.uint8, device=dev) if solid is not None else None,
                 "p": z(n, n, n) if with_pressure else None}
            self.tensors.append(t)
            sc = self.array[k]
            sc.grid = fia.GridC(n, n, n, 1.0 / n, t["rho"].data_ptr(), t["u"].data_ptr(), t["v"].data_ptr(),
                                t["w"].data_ptr(), 1)
            sc.solid = t["solid"].data_ptr() if t["solid"] is not None else None
            sc.pressure = t["p"].data_ptr() if t["p"] is not None else None
            sc.source_x, sc.source_y, sc.source_z = centre
            sc.source_radius, sc.source_value = 0.06, 1.0
            sc.buoyancy = float(rng.uniform(0.5, 2.0))
            sc.cfl = int(rng.integers(cfl_range[0], cfl_range[1] + 1))
            sc.movie_fps = 60.0
            self.edges.append(n)
        self.cells = int(sum(e ** 3 for e in self.edges))

    def step(self, ctxs, density_opts, velocity_opts, nets=None, rel_tol=1e-6, batched=False):
        self.fia.scenes_step(ctxs, self.array, density_opts, velocity_opts, rel_tol=rel_tol, nets=nets,
                             batched=batched)

    def stats(self):
        a = self.array
        return {"pcg_iters_mean": float(np.mean([s.pcg_iters for s in a])),
                "pcg_residual_max": float(max(s.pcg_residual for s in a)),
                "network_used": int(sum(s.used_network for s in a)), "steps": int(min(s.steps for s in a))}


def autotune_shapes(sset, ctxs, density_opts, reps=2, log=print):
    """fia_ctx_autotune on one scene of each grid edge (on copies: the scenes do not move), with dt
    from the scene's CFL number; the picked launch shapes go to every context.  Returns
    {edge: shapes}."""
    import torch
    fia, scenes = _pkg()
    picked = {}
    for k, n in enumerate(sset.edges):
        if n in picked:
            continue
        t = sset.tensors[k]
        g = fia.Grid3D(n, n, n, 1.0 / n, t["rho"], t["u"], t["v"], t["w"])
        vmax = max(float(t[f].abs().max()) for f in ("u", "v", "w"))
        dt = sset.array[k].cfl * (1.0 / n) / max(vmax, 1e-3)
        picked[n] = ctxs[0].autotune(g, density_opts, dt, solid=t["solid"], reps=reps)
        for c in ctxs[1:]:
            c.set_tuning((n, n, n), **picked[n])
        log("autotune: %d^3 %s" % (n, picked[n]))
    torch.cuda.synchronize()
    return picked


def autotune_workers(sset, ctxs, density_opts, velocity_opts, candidates, nets=None, batched=False, log=print):
    """Times one step of the whole set with each worker count (these steps advance the scenes) and
    returns the fastest count with the timings."""
    import torch
    timings = {}
    for w in candidates:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sset.step(ctxs[:w], density_opts, velocity_opts, nets=nets[:w] if nets else None, batched=batched)
        torch.cuda.synchronize()
        timings[w] = time.perf_counter() - t0
        log("autotune: %d workers, %.1f ms per step" % (w, timings[w] * 1e3))
    return min(timings, key=timings.get), timings
