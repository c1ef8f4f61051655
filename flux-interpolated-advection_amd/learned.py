"""Training data, training and evaluation of the learned pressure warm start (the U-Net of
libfia.so, fia_unet_*).  It has no counterpart in the reference, whose pressure solver is amgcl
(src/calc_pressure_3d.cpp:141-166).

Data: buoyant plume scenes simulated by this library on the GPU.  Each scene has a seeded source
sphere (radius 0.06, centre in the lower part of the box), buoyancy in [0.5, 2] on the vertical
(y) velocity and 0-8 spherical obstacles kept clear of the source.  A step adds the source and
the buoyancy, self-advects the velocity (flux-interpolated, WENO6), projects with PCG to 1e-6
(Jacobi-preconditioned by default: `--precond`) and advects the density.  Every step after a short warm-up gives one sample: the right-hand
side of the pressure system just before the projection, and the obstacle mask.

    python flux-interpolated-advection_amd/learned.py train --n 128 --scenes 12 --steps 40 \\
        --seconds 420 --out weights.npy
    python flux-interpolated-advection_amd/learned.py evaluate --n 128 --weights weights.npy

Training minimises the residual-norm loss |L y - b|^2 / |b|^2 of the network output (Adam,
cosine-decayed learning rate); evaluation compares the projection from a zero guess with the
learned one (same system, same tolerance) on scenes with seeds not used in training.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.basename(HERE)


def _pkg():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    return importlib.import_module(PKG), importlib.import_module(PKG + ".scenes")


class PlumeSim:
    """One seeded plume scene on the GPU (torch device tensors, library kernels)."""

    def __init__(self, fia, scenes, ctx, n, seed, cfl=2, velocity_interpolation="WENO6"):
        import torch
        self.fia, self.ctx, self.n, self.cfl = fia, ctx, n, cfl
        self.vi = velocity_interpolation
        rng = np.random.default_rng(seed)
        centre = (rng.uniform(0.25, 0.75), rng.uniform(0.08, 0.2), rng.uniform(0.25, 0.75))
        self.beta = float(rng.uniform(0.5, 2.0))
        self.obstacles = int(rng.integers(0, 9))
        src = scenes.sphere_source(n, centre, 0.06)
        clear = scenes.sphere_source(n, centre, 0.14)
        dev = torch.device("cuda:0")
        self.src = torch.tensor(src, dtype=torch.float32, device=dev)
        self.solid = None
        if self.obstacles:
            solid = scenes.spherical_obstacles(n, count=self.obstacles, seed=seed + 7919, keep_clear=clear)
            self.solid = torch.tensor(solid, dtype=torch.uint8, device=dev)
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
        self.grid = fia.Grid3D(n, n, n, 1.0 / n, z(n, n, n), z(n + 1, n, n), z(n, n + 1, n), z(n, n, n + 1))
        self.t, self.frame, self.dt = 0.0, 0, 0.0

    def before_projection(self):
        """Source, buoyancy and velocity self-advection of one step."""
        import torch
        fia, g = self.fia, self.grid
        dt, write = fia.calc_time_step_length_from_cfl_number_3d(self.ctx, g, self.t, self.frame, self.cfl, 60.0)
        torch.maximum(g.density, self.src, out=g.density)
        g.v[:, 1:-1, :] += dt * self.beta * 0.5 * (g.density[:, 1:, :] + g.density[:, :-1, :])
        fia.advect_velocity_xyz_flux_advection_3d(self.ctx, g, dt, interpolation_method=self.vi)
        self.dt = dt
        if write:
            self.frame += 1

    def rhs(self):
        return self.fia.pressure_rhs(self.ctx, self.grid, self.solid)

    def project(self, rel_tol=1e-6):
        return self.fia.calc_pressure_gradient_term_3d(self.ctx, self.grid, solid=self.solid, rel_tol=rel_tol)

    def after_projection(self):
        self.fia.move_substances_3d(self.ctx, self.grid, self.frame, self.dt, split_method="xz-CellFaceAligned")
        self.t += self.dt


def generate(fia, scenes, ctx, n, scene_seeds, steps, warmup=3, log=print):
    """Samples (rhs FP64, obstacle mask uint8) of the given scenes, stacked on the GPU."""
    import torch
    rhs, solid = [], []
    t0 = time.time()
    for k, seed in enumerate(scene_seeds):
        sim = PlumeSim(fia, scenes, ctx, n, seed)
        for step in range(warmup + steps):
            sim.before_projection()
            if step >= warmup:
                rhs.append(sim.rhs())
                solid.append(sim.solid if sim.solid is not None else torch.zeros((n, n, n), dtype=torch.uint8,
                                                                                   device=rhs[-1].device))
            sim.project()
            sim.after_projection()
        log("data: scene %d/%d (seed %d, %d obstacles, buoyancy %.2f) %d samples, %.1f s" %
            (k + 1, len(scene_seeds), seed, sim.obstacles, sim.beta, len(rhs), time.time() - t0))
    return torch.stack(rhs), torch.stack(solid)


def train(fia, net, rhs, solid, seconds, batch=4, lr=1e-3, lr_final=1e-4, seed=0, loss="residual", log=print):
    """Adam on random batches for `seconds` of wall time, cosine-decayed learning rate."""
    import torch
    rng = np.random.default_rng(seed)
    S = rhs.shape[0]
    t0 = time.time()
    step, window, wres, history = 0, [], [], []
    last = t0
    while True:
        el = time.time() - t0
        if el >= seconds:
            break
        f = el / seconds
        lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * f))
        idx = torch.tensor(rng.choice(S, size=batch, replace=False), device=rhs.device)
        lv, rv = net.train_step(rhs.index_select(0, idx).contiguous(), solid.index_select(0, idx).contiguous(), lr=lr_t,
                                loss=loss)
        window.append(lv)
        wres.append(rv)
        step += 1
        if time.time() - last >= 10.0:
            history.append((step, float(np.mean(window)), float(np.mean(wres))))
            log("train: step %d, %.0f s, lr %.2e, mean %s loss %.4e, residual norm %.4e" %
                (step, time.time() - t0, lr_t, loss, np.mean(window), np.mean(wres)))
            window, wres, last = [], [], time.time()
    if window:
        history.append((step, float(np.mean(window)), float(np.mean(wres))))
    return step, history


def evaluate(fia, scenes, ctx, net, n, scene_seeds, steps, warmup=3, rel_tol=1e-6, log=print):
    """Per projection: iterations and milliseconds from a zero guess and with the warm start, on
    the same velocity state; the scene continues from the zero-guess projection."""
    import torch
    rows = []
    for seed in scene_seeds:
        sim = PlumeSim(fia, scenes, ctx, n, seed)
        for step in range(warmup + steps):
            sim.before_projection()
            if step >= warmup:
                g = sim.grid
                copy = fia.Grid3D(n, n, n, g.dx, None, g.u.clone(), g.v.clone(), g.w.clone())
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                it0, r0 = sim.project(rel_tol)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                it1, r1, info = fia.calc_pressure_gradient_term_3d_learned(ctx, copy, net, solid=sim.solid,
                                                                           rel_tol=rel_tol)
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                du = float((copy.u - g.u).abs().max())
                rows.append({"seed": seed, "step": step, "plain_iters": it0, "plain_ms": (t1 - t0) * 1e3,
                             "plain_res": r0, "learned_iters": it1, "learned_ms": (t2 - t1) * 1e3, "learned_res": r1,
                             "used_network": info["used_network"], "guess_rel_residual": info["guess_rel_residual"],
                             "guess_scale": info["guess_scale"], "probe_res_network": info["probe_res_network"],
                             "probe_res_zero": info["probe_res_zero"],
                             "max_velocity_difference": du})
            else:
                sim.project(rel_tol)
            sim.after_projection()
        last = [r for r in rows if r["seed"] == seed]
        log("eval: seed %d: plain %.0f it %.1f ms, learned %.0f it %.1f ms, network used %d/%d" %
            (seed, np.mean([r["plain_iters"] for r in last]), np.mean([r["plain_ms"] for r in last]),
             np.mean([r["learned_iters"] for r in last]), np.mean([r["learned_ms"] for r in last]),
             sum(r["used_network"] for r in last), len(last)))
    return rows


def guess_quality(fia, scenes, ctx, n, seed, eps_list=(1e-1, 1e-2, 1e-3), warmup=3, rel_tol=1e-6):
    """How good an initial guess must be to cut the CG's iterations, on the velocity state of one
    plume step: p* from a solve to 1e-10, then guesses (1 - eps) p*, whose error is a scaled copy
    of the solution (the slow modes CG spends its iterations on) and whose residual is eps |b| up
    to p*'s FP32 rounding, and the iterations to rel_tol from each, beside the zero guess.  Each
    row: eps, the guess's measured relative residual, iterations, their ratio to the zero guess's."""
    import ctypes
    sim = PlumeSim(fia, scenes, ctx, n, seed)
    for _ in range(warmup):
        sim.before_projection()
        sim.project(rel_tol)
        sim.after_projection()
    sim.before_projection()
    g = sim.grid
    sp = sim.solid.data_ptr() if sim.solid is not None else None

    def fresh():
        return fia.Grid3D(n, n, n, g.dx, None, g.u.clone(), g.v.clone(), g.w.clone())

    def solve(guess, tol, max_iters):
        # through the C ABI: a capped solve reports FIA_ERR_NOT_CONVERGED with its count and residual
        gr = fresh()
        gc = gr._c(with_density=False)
        gr.pressure = guess.clone()
        it, res = ctypes.c_int(0), ctypes.c_double(0.0)
        st = fia.lib().fia_pressure_project(ctx._h, ctypes.byref(gc), sp, gr.pressure.data_ptr(), 1, float(tol),
                                            int(max_iters), ctypes.byref(it), ctypes.byref(res))
        if fia.STATUS[st] not in ("FIA_OK", "FIA_ERR_NOT_CONVERGED"):
            raise fia.FiaError(st, fia.lib().fia_last_error().decode())
        return it.value, res.value, gr.pressure

    zero = g.u.new_zeros((n, n, n))
    _, _, p_star = solve(zero, 1e-10, 50000)
    it0, _, _ = solve(zero, rel_tol, 20000)
    rows = [{"eps": 1.0, "guess_rel_residual": 1.0, "iters": it0, "iters_ratio": 1.0}]
    for eps in eps_list:
        guess = (1.0 - eps) * p_star
        _, r0, _ = solve(guess, rel_tol, 0)
        it, _, _ = solve(guess, rel_tol, 20000)
        rows.append({"eps": eps, "guess_rel_residual": r0, "iters": it, "iters_ratio": it / max(it0, 1)})
    return rows


def summary(rows):
    p_it = float(np.mean([r["plain_iters"] for r in rows]))
    l_it = float(np.mean([r["learned_iters"] for r in rows]))
    p_ms = float(np.mean([r["plain_ms"] for r in rows]))
    l_ms = float(np.mean([r["learned_ms"] for r in rows]))
    return {"projections": len(rows), "plain_iters": p_it, "learned_iters": l_it, "plain_ms": p_ms,
            "learned_ms": l_ms, "speedup": p_ms / l_ms if l_ms > 0 else None,
            "network_used": int(sum(r["used_network"] for r in rows)),
            "guess_rel_residual_mean": float(np.mean([r["guess_rel_residual"] for r in rows])),
            "max_rel_residual": float(max(max(r["plain_res"], r["learned_res"]) for r in rows))}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--precond", choices=("jacobi", "mg"), default="jacobi",
                    help="preconditioner of the scene and evaluation projections")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--n", type=int, default=128)
    t.add_argument("--scenes", type=int, default=12)
    t.add_argument("--steps", type=int, default=40)
    t.add_argument("--seconds", type=float, default=420.0)
    t.add_argument("--batch", type=int, default=4)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--init", default=None, help="start from these parameters (.npy)")
    t.add_argument("--loss", choices=["residual", "energy"], default="residual")
    t.add_argument("--out", required=True)
    t.add_argument("--eval-scenes", type=int, default=2)
    t.add_argument("--eval-steps", type=int, default=6)
    e = sub.add_parser("evaluate")
    e.add_argument("--n", type=int, default=128)
    e.add_argument("--weights", required=True)
    e.add_argument("--scenes", type=int, default=2)
    e.add_argument("--steps", type=int, default=8)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    fia, scenes = _pkg()
    ctx = fia.Context(0)
    ctx.set_pressure_preconditioner(a.precond)
    net = fia.UNet(ctx, seed=1)
    report = {"n": a.n}
    if a.cmd == "train":
        if a.init:
            net.load(a.init)
        rhs, solid = generate(fia, scenes, ctx, a.n, range(1000, 1000 + a.scenes), a.steps, log=log)
        report.update({"samples": int(rhs.shape[0]), "scene_seeds": [1000, 1000 + a.scenes - 1]})
        steps, hist = train(fia, net, rhs, solid, a.seconds, batch=a.batch, lr=a.lr, loss=a.loss, log=log)
        net.save(a.out)
        report.update({"loss": a.loss, "train_steps": steps, "batch": a.batch, "seconds": a.seconds,
                       "loss_history": hist})
        del rhs, solid
        rows = evaluate(fia, scenes, ctx, net, a.n, range(5000, 5000 + a.eval_scenes), a.eval_steps, log=log)
    else:
        net.load(a.weights)
        rows = evaluate(fia, scenes, ctx, net, a.n, range(5000, 5000 + a.scenes), a.steps, log=log)
    report["evaluation"] = summary(rows)
    report["rows"] = rows
    print(json.dumps({k: v for k, v in report.items() if k != "rows"}), flush=True)
    if a.cmd == "train":
        with open(os.path.splitext(a.out)[0] + "_report.json", "w") as f:
            json.dump(report, f, indent=1)
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
