"""Benchmark of the flux-interpolated advection path on one MI355X.

Workloads (--workload), each on the 256^3 plume of the measured configuration: 256^3 MAC grid in
the unit cube, density = a Gaussian plume blob plus a sphere source, divergence-free seeded
curl-noise velocity (wavenumbers 1-4) scaled so that dt = 4 dx / max|u| (CFL 4), WENO6, integral
psi (Gauss n=1), zero-velocity correction on.  Fields stay resident on the GPU.

  plume       (default) one step = one step of the simulation loop of the reference's main loop
              (src/main.cpp:596-740) on that scene with 8 seeded spherical obstacles (radius
              0.03-0.08) as a solid mask, run natively by fia_scenes_step on one context: dt from
              CFL 4 (no movie-frame clipping), density source and buoyancy 1.0 on v, velocity
              self-advection (u, v, w; WENO6 integral, 3 passes each), pressure projection with
              the obstacles to 1e-6 relative residual from a zero guess by Jacobi-PCG (the plain
              PCG of the configuration), density advection (3 passes).  The simulation runs on:
              each timed step continues the last.  Beside it (same line): the same steps under
              the multigrid-preconditioned CG, the learned warm start against plain PCG at
              128^3 (step ms of NN-warm-PCG vs PCG at 1e-6), and the mass change of one density
              step on the final state.
  density     one step = one move_substances_3D (xz-CellFaceAligned): 3 flux passes
  velocity    one step = advect_velocity_xyz_flux_advection_3D (u, v, w, 3 passes each)
  projection  one step = one pressure projection (Jacobi-PCG to 1e-6 relative residual, zero
              initial guess) of the curl-noise velocity plus a buoyancy impulse, 8 spherical
              obstacles as a solid mask
  scenes      one step = one step of --scenes (64) independent plume scenes (edges drawn from
              --edges, {32,48,64,80,96}; seeded sources, buoyancy, 0-4 obstacles, CFL 1-3):
              forcing, velocity advection,
              projection (U-Net warm start unless --plain) and density advection, on the number
              of concurrent workers the autotuner picks; value = all cells x steps / time;
              --batched: the pressure solves of all scenes batched (fia_scenes_step_batched)
  warmstart   one step = the pressure projection of one plume step (128^3 unless --n; seeded
              source, buoyancy, 0-8 obstacles; scenes not used in training) twice on the same
              velocity: from a zero guess and with the learned U-Net warm start (committed
              weights), both to 1e-6; value and ms_per_step are the warm-started projection's

Prints one JSON line: value = cell-steps per second / 1e6 over the timed steps; a roofline entry
for the dominant kernel measured with HIP events on the library's stream (plume: the CG iteration
of the pressure solve, HBM; density, velocity: the WENO6 quadrature kernel, FP32 VALU), and
(plume and density, unless --no-cpu-baseline) a CPU baseline: the oracle's restatement of the
reference path, multi-threaded, on a bounded sample of the same workload (plume: one whole step at
--cpu-n, 64^3 by default; density: one density step at 128^3).
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "flux-advection Mcell-steps/s at 256^3; step ms NN-warm-PCG vs PCG at 1e-6 res"
PEAK_FP32_TFLOPS = 157.3          # published FP32 vector rate (packed FP32 instructions)
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # one FP32 FMA per lane and cycle: this build emits no packed FP32
# Work the quadrature kernel executes: its 1D WENO6 evaluations, counted on the device, at 73 flop
# each (an FMA counts 2, the reciprocal 1): the arithmetic of one evaluation in the compiled kernel
# (five sample differences, second and third differences, three candidates, the common squared
# first difference and three indicator forms over 7 products, weights, blend).
FLOP_PER_WENO6_EXECUTED = 73
# The reference rule's count: every psi point and density point is one 3D WENO6 interpolation,
# 43 1D evaluations of 200 flop as written in the reference's weno.h:34-68 (8,600 flop per point).
FLOP_PER_WENO3D_POINT = 43 * 200
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of piece_kernel<1> at the default workload (256^3, CFL 4), from two
# rocprofv3 counter passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a 1 GiB device copy);
# recorded in profiles/density_step_256_cfl4_hbm_counters.md
TRAFFIC_BYTES_256_CFL4 = 12.124e9
# Jacobi-PCG algorithmic traffic per cell and iteration (FP64 x and r, FP32 search direction p,
# two kernels): direction kernel reads r (8), p_old (4) and the flag byte, writes p (4); update
# kernel reads x, r (2 x 8), p (4) and the flag byte, writes x, r (2 x 8); q = L p is recomputed,
# not stored
PCG_BYTES_PER_CELL_ITER = 54
# the one-sweep Jacobi iteration (pcg_fused_kernel), x stepped on every other sweep: odd sweeps
# flags 1, p 4 + 4, r 8 + 8 (25); even sweeps also x 8 + 8 and p_(k-2) 4 (45); 35 on average
PCG_FUSED_BYTES_PER_CELL_ITER = 35
# HBM bytes per pcg_fused_kernel launch (mean of an odd and an even sweep) on the 256^3 plume
# (8 obstacles), from two rocprofv3 passes (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, scaled by the
# counters' reading of a 1 GiB copy): odd 512.2 MB, even 865.8 MB; recorded in
# profiles/plume_step_256_cfl4_hbm_counters.md
PCG_FUSED_TRAFFIC_BYTES_256 = 689.0e6
# one multigrid V-cycle per iteration, per fine cell: the down kernel reads r (8) and the flag
# byte, writes x (4) and the aggregate's b (0.5); the up kernel reads x (4), r (8), the flag byte
# and the correction (0.5), writes z (8); the coarser levels add about a seventh of a fine level's
# 28 bytes; the direction kernel reads z in place of r (no change)
MG_BYTES_PER_CELL_VCYCLE = 39


def make_scene(scenes, n, cfl, seed=1):
    h = 1.0 / n
    rho = scenes.gaussian_blob(n, centre=(0.5, 0.3, 0.5), sigma=0.1) + scenes.sphere_source(n)
    u, v, w = scenes.curl_noise(n, seed=seed)
    dt = scenes.cfl_dt(u, v, w, h, cfl)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return h, f32(rho), f32(u), f32(v), f32(w), dt


def cpu_baseline(scenes, n, cfl, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.lib()
    h, rho, u, v, w, dt = make_scene(scenes, n, cfl)
    t0 = time.perf_counter()
    O.move_substances(rho.astype(np.float64), u, v, w, h, dt, split="xz-CellFaceAligned", method="WENO6",
                      prec=1, cprec=1, nthreads=threads)
    sec = time.perf_counter() - t0
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
    except OSError:
        pass
    return {"value": round(n ** 3 / sec / 1e6, 5), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "host": {"cpu_model": model, "logical_cpus_visible": os.cpu_count(), "threads_used": threads},
            "sample": "one move_substances_3D step at %d^3 (same workload, CFL %g), FP32 oracle restatement of the "
                      "reference path (per-cell faces, std::thread over x-slabs), %.1f s" % (n, cfl, sec)}


def quad_roofline(st, steps, n, cfl, density, optimized=False):
    """Roofline entry of the WENO6 integral quadrature kernel (piece_kernel<1>) from the
    library's HIP-event spans and device work counters of the timed steps.  achieved counts the
    flops the kernel executes (its 1D WENO6 evaluations x 73); the reference rule's count
    (8,600 flop per 3D point) is reported beside it: the kernel shares the z pass between the
    points of a piece and needs no y pass for psi on a face level, so that count over the
    kernel's time is a throughput of the reference's work, not a fraction of the hardware."""
    points = st["psi_points"] + st["density_points"]
    quad_s = st["quad_ms"] / 1e3
    qlaunches = max(st["quad_launches"], 1)
    executed = st["weno_evals"] * FLOP_PER_WENO6_EXECUTED / quad_s / 1e12 if quad_s > 0 else 0.0
    rule = points * FLOP_PER_WENO3D_POINT / quad_s / 1e12 if quad_s > 0 else 0.0
    traffic = TRAFFIC_BYTES_256_CFL4 if (density and not optimized and n == 256 and cfl == 4.0) else None
    avg_quad_s = quad_s / qlaunches
    return {"bound": "valu", "achieved": round(executed, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(executed / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
            "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS,
            "frac_nonpacked": round(executed / PEAK_FP32_NONPACKED_TFLOPS, 4),
            "kernel": ("piece_kernel<1, true> (WENO6-optimized integral quadrature; flops priced as the WENO6 "
                       "form's, without its two extra reciprocals)" if optimized else
                       "piece_kernel<1> (WENO6 integral quadrature)"), "compute": "FP32 VALU, no MFMA",
            "avg_launch_ms": round(st["quad_ms"] / qlaunches, 3),
            "weno6_evals_per_launch": st["weno_evals"] / qlaunches, "flop_per_weno6": FLOP_PER_WENO6_EXECUTED,
            "reference_rule": {"TFLOPs": round(rule, 3), "weno3d_points_per_launch": points / qlaunches,
                               "flop_per_point": FLOP_PER_WENO3D_POINT},
            "hbm": None if traffic is None else {
                "achieved_GBs": round(traffic / avg_quad_s / 1e9, 1), "peak_GBs": PEAK_HBM_GBS,
                "frac": round(traffic / avg_quad_s / 1e9 / PEAK_HBM_GBS, 4)}}


def run_advection(args, fia, scenes, torch, velocity):
    n = args.n
    h, rho, u, v, w, dt = make_scene(scenes, n, args.cfl)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(a).to(dev)
    rho_d, u_d, v_d, w_d = T(rho), T(u), T(v), T(w)
    grid = fia.Grid3D(n, n, n, h, rho_d, u_d, v_d, w_d)
    ctx = fia.Context(0)
    if velocity:
        opts = dict(interpolation_method="WENO6", use_integral=True,
                    num_gauss_quadrature_point_for_integrate_density=1,
                    enable_eliminate_zero_velocity_false_diffusion=True)
        # every step starts from the same curl-noise field (a device copy of 0.2 GB, ~0.1 ms, inside
        # the timed region): repeated self-advection without projection drifts into fields no
        # simulation step sees
        u0, v0, w0 = u_d.clone(), v_d.clone(), w_d.clone()

        def step(i):
            u_d.copy_(u0); v_d.copy_(v0); w_d.copy_(w0)
            torch.cuda.synchronize()
            fia.advect_velocity_xyz_flux_advection_3d(ctx, grid, dt, **opts)
    else:
        opts = dict(split_method="xz-CellFaceAligned", interpolation_method=args.interp, use_integral=True,
                    num_gauss_quadrature_point_for_integrate_density=1,
                    enable_eliminate_zero_velocity_false_diffusion=True)
        step = lambda i: fia.move_substances_3d(ctx, grid, i, dt, **opts)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    st = ctx.stats()
    launches = max(st["face_launches"], 1)
    if velocity:
        workload = ("advect_velocity_xyz_flux_advection_3D step (u, v, w; 3 passes each), %d^3 curl-noise velocity, "
                    "CFL %g, WENO6, integral psi (Gauss 1), per-component splits, zero-velocity correction on; "
                    "each step from the same initial field" % (n, args.cfl))
        extra = {"max_abs_velocity_after": float(max(u_d.abs().max().item(), v_d.abs().max().item(),
                                                     w_d.abs().max().item()))}
    else:
        workload = ("move_substances_3D density step, %d^3 plume blob + sphere source, curl-noise velocity, CFL %g, "
                    "%s, integral psi (Gauss 1), xz-CellFaceAligned, zero-velocity correction on" %
                    (n, args.cfl, args.interp))
        extra = {"mass_after": float(rho_d.double().sum().item())}
    out = {
        "metric": METRIC, "value": round(n ** 3 * args.steps / sec / 1e6, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "grid": [n, n, n], "cfl": args.cfl,
                   "passes_per_step": 9 if velocity else 3},
        "roofline": quad_roofline(st, args.steps, n, args.cfl, not velocity,
                                  optimized=not velocity and args.interp == "WENO6-optimized"),
        "kernels_ms_per_step": {k: round(st[k] / args.steps, 3)
                                for k in ("face_ms", "quad_ms", "vertex_ms", "psi_ms", "gather_ms")},
        "face_work": {"launches": st["face_launches"], "weno_pieces_per_step": st["pieces"] / args.steps,
                      "density_points_per_step": st["density_points"] / args.steps,
                      "avg_face_launch_ms": round(st["face_ms"] / launches, 3)},
    }
    out.update(extra)
    if not velocity and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scenes, args.cpu_n, args.cfl, min(os.cpu_count() or 1, 16))
    ctx.close()
    return out


def plume_scene(scenes, n, seed=1):
    """The 256^3 plume of the measured configuration: blob + source density, curl-noise velocity
    (max |u| = 1), 8 seeded spherical obstacles kept clear of the source."""
    h = 1.0 / n
    source = (0.5, 0.1, 0.5)
    rho = scenes.gaussian_blob(n, centre=(0.5, 0.3, 0.5), sigma=0.1) + scenes.sphere_source(n, source, 0.06)
    u, v, w = scenes.curl_noise(n, seed=seed)
    clear = scenes.sphere_source(n, source, 0.14)
    solid = scenes.spherical_obstacles(n, count=8, rmin=0.03, rmax=0.08, seed=3, keep_clear=clear)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return h, source, f32(rho), f32(u), f32(v), f32(w), np.ascontiguousarray(solid, dtype=np.uint8)


class PlumeRun:
    """One 256^3 plume scene stepped by fia_scenes_step on one context."""

    def __init__(self, fia, torch, n, state, precond, cfl):
        h, source, rho, u, v, w, solid = state
        dev = torch.device("cuda:0")
        T = lambda a: torch.from_numpy(a).to(dev)
        self.fia, self.torch, self.n = fia, torch, n
        self.t = {"rho": T(rho), "u": T(u), "v": T(v), "w": T(w), "solid": T(solid)}
        self.ctx = fia.Context(0)
        self.ctx.set_pressure_preconditioner(precond)
        self.scenes = (fia.SceneC * 1)()
        sc = self.scenes[0]
        t = self.t
        sc.grid = fia.GridC(n, n, n, h, t["rho"].data_ptr(), t["u"].data_ptr(), t["v"].data_ptr(), t["w"].data_ptr(), 1)
        sc.solid = t["solid"].data_ptr()
        sc.pressure = None
        sc.source_x, sc.source_y, sc.source_z = source
        sc.source_radius, sc.source_value, sc.buoyancy = 0.06, 1.0, 1.0
        sc.cfl = int(cfl)
        sc.movie_fps = 1e-3   # frames 1000 s apart: dt is the CFL step every time
        self.od = fia.make_opts(split_method="xz-CellFaceAligned", interpolation_method="WENO6", use_integral=True,
                                num_gauss_quadrature_point_for_integrate_density=1)
        self.ov = fia.make_opts(interpolation_method="WENO6", use_integral=True,
                                num_gauss_quadrature_point_for_integrate_density=1)
        self.iters, self.res, self.dts = [], [], []

    def step(self):
        self.fia.scenes_step([self.ctx], self.scenes, self.od, self.ov, rel_tol=1e-6, max_iters=20000)
        sc = self.scenes[0]
        self.iters.append(int(sc.pcg_iters)); self.res.append(float(sc.pcg_residual)); self.dts.append(float(sc.dt))

    def timed(self, steps, warmup, log):
        torch = self.torch
        for i in range(warmup):
            self.step()
            log("warm-up step %d: %d CG iterations" % (i, self.iters[-1]))
        # launch shapes for this grid, timed on copies of the warmed-up state (fia_ctx_autotune)
        self.shapes = None
        if self.dts:
            self.shapes = self.ctx.autotune(self.grid(), self.od, self.dts[-1], solid=self.t["solid"], reps=2)
            log("autotuned launch shapes: %s" % self.shapes)
        self.iters, self.res, self.dts = [], [], []
        torch.cuda.synchronize()
        self.ctx.set_timing(True)
        self.ctx.reset_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            self.step()
            log("timed step %d: %d CG iterations, residual %.2e" % (i, self.iters[-1], self.res[-1]))
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        st = self.ctx.stats()
        self.ctx.set_timing(False)
        return sec, st

    def grid(self):
        t = self.t
        return self.fia.Grid3D(self.n, self.n, self.n, 1.0 / self.n, t["rho"], t["u"], t["v"], t["w"])

    def close(self):
        self.ctx.close()


def pcg_is_fused(ctx, n, precond):
    """Whether the context's solves of an n^3 grid run the one-sweep Jacobi iteration (its launch
    shapes for that grid: fused on request, or by default from 4M cells)."""
    if precond != "jacobi":
        return False
    f = ctx.get_tuning((n, n, n))["pcg_fused"]
    return f == 1 or (f == -1 and n ** 3 >= 1 << 22)


def pcg_roofline(st, n, precond, fused=False):
    """Roofline entry of the CG iteration from the library's HIP-event spans around its launches:
    algorithmic bytes of one iteration over all cells, divided by the measured time per iteration."""
    per_cell = (PCG_FUSED_BYTES_PER_CELL_ITER if fused else PCG_BYTES_PER_CELL_ITER) + \
        (MG_BYTES_PER_CELL_VCYCLE if precond == "mg" else 0)
    iters = max(st["pcg_iters"], 1)
    ms_iter = st["pcg_ms"] / iters
    gbs = per_cell * n ** 3 / (ms_iter / 1e3) / 1e9 if ms_iter > 0 else 0.0
    kernels = ("pcg_fused_kernel<false|true> (one-sweep Jacobi-PCG iteration, FP64 r, FP32 p, FP64 x stepped "
               "on every other sweep)" if fused else
               "pcg_direction_kernel + pcg_update_kernel<true> (Jacobi-PCG iteration, FP64 vectors)"
               if precond == "jacobi" else "multigrid-preconditioned CG iteration (direction, update, V-cycle)")
    traffic = PCG_FUSED_TRAFFIC_BYTES_256 if (fused and n == 256) else None
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel": kernels,
            "bytes_per_cell_iteration": per_cell, "avg_iteration_ms": round(ms_iter, 4),
            "iterations": st["pcg_iters"], "solves": st["pcg_solves"]}


def plume_cpu_baseline(scenes, n, cfl, threads):
    """The oracle's restatement of one whole plume step at n^3 (velocity self-advection, pressure
    projection with the obstacles, density advection; FP32 fields, per-cell faces as the reference
    has them, std::thread over x-slabs; the CG of the oracle runs on one thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.lib()
    h, source, rho, u, v, w, solid = plume_scene(scenes, n)
    dt = cfl * h / max(abs(u).max(), abs(v).max(), abs(w).max())
    t0 = time.perf_counter()
    ua, va, wa = O.advect_velocity(u, v, w, h, dt, method="WENO6", prec=1, cprec=1, nthreads=threads)
    ub, vb, wb, p, it, res, rc = O.pressure_project(ua, va, wa, h, solid=solid, tol=1e-6, max_iters=20000)
    O.move_substances(rho.astype(np.float64), ub, vb, wb, h, dt, split="xz-CellFaceAligned", method="WENO6",
                      prec=1, cprec=1, nthreads=threads)
    sec = time.perf_counter() - t0
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
    except OSError:
        pass
    return {"value": round(n ** 3 / sec / 1e6, 5), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "host": {"cpu_model": model, "logical_cpus_visible": os.cpu_count(), "threads_used": threads},
            "sample": "one whole plume step at %d^3 (velocity self-advection, CG projection to 1e-6 with the 8 "
                      "obstacles: %d iterations, density advection; CFL %g), the oracle's FP32 restatement of the "
                      "reference path (per-cell faces, std::thread over x-slabs), %.1f s" % (n, it, cfl, sec)}


def run_plume(args, fia, scenes, torch):
    n = args.n
    log = (lambda s: print(s, file=sys.stderr, flush=True)) if args.verbose else (lambda s: None)
    log("plume %d^3: building the scene" % n)
    state = plume_scene(scenes, n)
    run = PlumeRun(fia, torch, n, state, "jacobi", args.cfl)
    sec, st = run.timed(args.steps, args.warmup, log)
    fused = pcg_is_fused(run.ctx, n, "jacobi")
    mass0 = fia.grid_totals(run.ctx, run.grid())[0]
    main = {"iters": list(run.iters), "res": list(run.res), "dts": list(run.dts), "shapes": run.shapes}
    # the mass change of one density step on the final state (no source, no forcing)
    g = run.grid()
    before = float(g.density.double().sum().item())
    fia.move_substances_3d(run.ctx, g, 0, run.dts[-1], split_method="xz-CellFaceAligned",
                           interpolation_method="WENO6")
    torch.cuda.synchronize()
    after = float(g.density.double().sum().item())
    run.close()
    del run
    torch.cuda.empty_cache()
    if not args.no_multigrid:
        log("multigrid variant")
        mg = PlumeRun(fia, torch, n, state, "mg", args.cfl)
        sec_mg, st_mg = mg.timed(args.steps, args.warmup, log)
        mg_iters = list(mg.iters)
        mg.close()
        del mg
        torch.cuda.empty_cache()
    out = {
        "metric": METRIC, "value": round(n ** 3 * args.steps / sec / 1e6, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 advection, f64 solve",
        "data": "synthetic",
        "config": {"workload": "plume simulation step, %d^3 (blob + sphere source, curl-noise start, buoyancy 1.0, "
                               "8 spherical obstacles r 0.03-0.08), CFL %g: velocity self-advection and density "
                               "advection (WENO6, integral psi, xz-CellFaceAligned density split, zero-velocity "
                               "correction on), Jacobi-PCG projection to 1e-6 from a zero guess" % (n, args.cfl),
                   "grid": [n, n, n], "cfl": args.cfl, "preconditioner": "jacobi",
                   "launch_shapes": main["shapes"],
                   "piece_rule": "the reference's: one psi point per piece plus 1-point Gauss along y for the "
                                 "density term; the 4x4 per-piece Gauss rule the configuration names is not built"},
        "roofline": pcg_roofline(st, n, "jacobi", fused),
        "pcg": {"iterations": main["iters"], "max_rel_residual": max(main["res"]),
                "dt": [round(d, 7) for d in main["dts"]]},
        "kernels_ms_per_step": {"pcg_ms": round(st["pcg_ms"] / args.steps, 3),
                                **{k: round(st[k] / args.steps, 3)
                                   for k in ("face_ms", "quad_ms", "vertex_ms", "psi_ms", "gather_ms")}},
        "roofline_advection": quad_roofline(st, args.steps, n, args.cfl, False),
        "multigrid": None if args.no_multigrid else {
            "ms_per_step": round(sec_mg / args.steps * 1e3, 3), "value": round(n ** 3 * args.steps / sec_mg / 1e6, 3),
            "pcg_iterations": mg_iters, "pcg_ms_per_step": round(st_mg["pcg_ms"] / args.steps, 3),
            "roofline": pcg_roofline(st_mg, n, "mg")},
        "mass": {"total_after_timed_steps": mass0,
                 "density_step_rel_change": (after - before) / before,
                 "note": "one density step on the final state, no source: the change is the scheme's wall "
                         "terms (moving wall faces sample the top-level psi at shifted points; the FP64 oracle "
                         "shows the same change, tests/test_gpu_c1.py), interior faces cancel exactly"},
    }
    if not args.no_warmstart:
        log("warm start at 128^3")
        out["warm_start"] = warmstart_summary(args, fia, scenes, torch)
    if not args.no_cpu_baseline:
        log("CPU baseline")
        out["cpu_baseline"] = plume_cpu_baseline(scenes, args.cpu_n, args.cfl, min(os.cpu_count() or 1, 16))
    return out


def warmstart_summary(args, fia, scenes, torch, n=128, steps=4):
    """NN-warm-PCG vs PCG step ms at 1e-6 on held-out plume scene 5000 (128^3, Jacobi-PCG, the
    solve the network was trained to start)."""
    learned = importlib.import_module("flux-interpolated-advection_amd.learned")
    if not os.path.exists(args.weights):
        return None
    ctx = fia.Context(0)
    ctx.set_pressure_preconditioner("jacobi")
    net = fia.UNet(ctx, seed=0)
    net.load(args.weights)
    rows = learned.evaluate(fia, scenes, ctx, net, n, [5000], 1 + steps, log=lambda s: None)[1:]
    s = learned.summary(rows)
    net.close()
    ctx.close()
    return {"grid": [n, n, n], "solver": "Jacobi-PCG to 1e-6", "plain_ms_per_step": round(s["plain_ms"], 3),
            "learned_ms_per_step": round(s["learned_ms"], 3), "speedup": round(s["plain_ms"] / s["learned_ms"], 4),
            "plain_iters": s["plain_iters"], "learned_iters": s["learned_iters"],
            "network_used": s["network_used"], "projections": s["projections"],
            "guess_rel_residual_mean": s["guess_rel_residual_mean"]}


def run_projection(args, fia, scenes, torch):
    n = args.n
    h = 1.0 / n
    u, v, w = scenes.curl_noise(n, seed=1)
    rho = scenes.gaussian_blob(n, centre=(0.5, 0.3, 0.5), sigma=0.1) + scenes.sphere_source(n)
    # buoyancy impulse on the y faces (y is up): v += dt_b * rho averaged to the face
    vb = np.zeros_like(v)
    vb[:, 1:-1, :] = 0.5 * (rho[:, 1:, :] + rho[:, :-1, :])
    v = v + 0.5 * vb
    solid = scenes.spherical_obstacles(n, count=8, seed=3)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
    u0, v0, w0 = T(u), T(v), T(w)
    solid_d = torch.from_numpy(np.ascontiguousarray(solid, dtype=np.uint8)).to(dev)
    ctx = fia.Context(0)
    ctx.set_pressure_preconditioner(args.precond)
    grid = fia.Grid3D(n, n, n, h, None, u0.clone(), v0.clone(), w0.clone())
    times, iters, res = [], [], []
    for i in range(args.warmup + args.steps):
        grid.u.copy_(u0); grid.v.copy_(v0); grid.w.copy_(w0)
        grid.pressure = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, rr = fia.calc_pressure_gradient_term_3d(ctx, grid, solid=solid_d, rel_tol=1e-6, max_iters=20000)
        torch.cuda.synchronize()
        if i >= args.warmup:
            times.append(time.perf_counter() - t0); iters.append(it); res.append(rr)
    sec = sum(times)
    it_mean = float(np.mean(iters))
    fused = pcg_is_fused(ctx, n, args.precond)
    per_iter = (PCG_FUSED_BYTES_PER_CELL_ITER if fused else PCG_BYTES_PER_CELL_ITER) + \
        (MG_BYTES_PER_CELL_VCYCLE if args.precond == "mg" else 0)
    gbs = per_iter * n ** 3 * it_mean / (sec / args.steps) / 1e9
    ctx.close()
    solver = "Jacobi-PCG" if args.precond == "jacobi" else "multigrid-preconditioned CG (V(1,1) per iteration)"
    return {
        "metric": METRIC, "value": round(n ** 3 * args.steps / sec / 1e6, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "pressure projection, %d^3, %s (FP64 CG) to 1e-6 relative residual from a zero "
                               "initial guess, curl-noise velocity + buoyancy impulse, 8 spherical obstacles" % (n, solver),
                   "grid": [n, n, n], "preconditioner": args.precond},
        "pcg": {"iterations": iters, "rel_residual": res, "warm_start": "none (zero initial guess; --workload warmstart measures the learned warm start)"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "whole CG iteration (%d B per cell and iteration)" % per_iter},
    }


def run_warmstart(args, fia, scenes, torch):
    learned = importlib.import_module("flux-interpolated-advection_amd.learned")
    if not os.path.exists(args.weights):
        print("no trained weights at %s (flux-interpolated-advection_amd/learned.py train)" % args.weights,
              file=sys.stderr)
        sys.exit(2)
    n = args.n
    ctx = fia.Context(0)
    ctx.set_pressure_preconditioner(args.precond)
    net = fia.UNet(ctx, seed=0)
    net.load(args.weights)
    # warm-up projections untimed, then the timed ones, all on held-out scene 5000 (one plume run)
    rows = learned.evaluate(fia, scenes, ctx, net, n, [5000], args.warmup + args.steps, log=lambda s: None)
    timed = rows[args.warmup:]
    s = learned.summary(timed)
    fused = pcg_is_fused(ctx, n, args.precond)
    # how good a guess must be: iterations from guesses (1 - eps) p* of the same system
    gq = learned.guess_quality(fia, scenes, ctx, n, 5000)
    net.close()
    ctx.close()
    l_ms, p_ms = s["learned_ms"], s["plain_ms"]
    per_iter = (PCG_FUSED_BYTES_PER_CELL_ITER if fused else PCG_BYTES_PER_CELL_ITER) + \
        (MG_BYTES_PER_CELL_VCYCLE if args.precond == "mg" else 0)
    solver = "Jacobi-PCG" if args.precond == "jacobi" else "multigrid-preconditioned CG"
    gbs = per_iter * n ** 3 * s["learned_iters"] / (l_ms / 1e3) / 1e9
    return {
        "metric": METRIC, "value": round(n ** 3 / (l_ms / 1e3) / 1e6, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(l_ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solve, bf16 network", "data": "synthetic",
        "config": {"workload": "pressure projection of a %d^3 buoyant plume step with the learned U-Net warm "
                               "start (energy-trained, 16/32/48/64 channels) vs from a zero guess, %s to 1e-6"
                               % (n, solver), "grid": [n, n, n], "preconditioner": args.precond,
                   "weights": os.path.relpath(args.weights, ROOT)},
        "warm_start": {"plain_ms_per_step": round(p_ms, 3), "learned_ms_per_step": round(l_ms, 3),
                       "speedup": round(p_ms / l_ms, 4), "plain_iters": s["plain_iters"],
                       "learned_iters": s["learned_iters"], "network_used": s["network_used"],
                       "projections": s["projections"], "guess_rel_residual_mean": s["guess_rel_residual_mean"],
                       "max_rel_residual": s["max_rel_residual"],
                       "guess_quality": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
                                         for r in gq]},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "warm-started projection: PCG iterations at %d B per cell over its whole time "
                               "(network and probe included)" % per_iter},
    }


def run_scenes(args, fia, torch):
    batch = importlib.import_module("flux-interpolated-advection_amd.batch")
    sset = batch.SceneSet(args.scenes, seed=5, edges=tuple(int(e) for e in args.edges.split(",")))
    candidates = (1, 2, 4, 8)
    ctxs = [fia.Context(0) for _ in range(max(candidates))]
    for c in ctxs:
        c.set_pressure_preconditioner(args.precond)
    nets = None
    if not args.plain:
        nets = []
        for _ in range(max(candidates)):
            u = fia.UNet(ctxs[0], seed=0)
            u.load(args.weights)
            nets.append(u)
    od = fia.make_opts(split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    ov = fia.make_opts(interpolation_method="WENO6")
    # the autotuner's steps are the warm-up; the launch shapes are then tuned per grid edge on the
    # moving scenes (on copies)
    best, tune = batch.autotune_workers(sset, ctxs, od, ov, candidates, nets=nets, batched=args.batched,
                                        log=lambda s: None)
    shapes = batch.autotune_shapes(sset, ctxs, od, log=lambda s: None)
    for _ in range(args.warmup):
        sset.step(ctxs[:best], od, ov, nets=nets[:best] if nets else None, batched=args.batched)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sset.step(ctxs[:best], od, ov, nets=nets[:best] if nets else None, batched=args.batched)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    st = sset.stats()
    for u in nets or []:
        u.close()
    for c in ctxs:
        c.close()
    return {
        "metric": METRIC, "value": round(sset.cells * args.steps / sec / 1e6, 3), "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup + len(candidates),
        "ms_per_step": round(sec / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 advection, f64 solve", "data": "synthetic",
        "config": {"workload": "%d independent plume scenes, edges %s (%d cells in all), CFL 1-3, 0-4 obstacles; "
                               "one step = forcing, velocity self-advection, projection to 1e-6 (%s), density "
                               "advection; WENO6%s" % (args.scenes, sorted(set(sset.edges)), sset.cells,
                                                       "plain PCG" if args.plain else "U-Net warm start",
                                                       "; pressure solves batched" if args.batched else ""),
                   "scenes": args.scenes, "cells": sset.cells},
        "scheduler": {"workers": best, "autotune_ms_per_step": {str(k): round(v * 1e3, 1) for k, v in tune.items()},
                      "launch_shapes": {str(k): v for k, v in sorted(shapes.items())}},
        "pcg": st,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None, help="grid edge (256; warmstart: 128)")
    ap.add_argument("--cfl", type=float, default=4.0)
    ap.add_argument("--interp", choices=("WENO6", "WENO6-optimized"), default="WENO6",
                    help="density workload: interpolation method of the step")
    ap.add_argument("--workload", choices=("plume", "density", "velocity", "projection", "warmstart", "scenes"),
                    default="plume")
    ap.add_argument("--verbose", action="store_true", help="progress lines on stderr")
    ap.add_argument("--no-warmstart", action="store_true", help="plume: skip the 128^3 warm-start comparison")
    ap.add_argument("--no-multigrid", action="store_true", help="plume: skip the multigrid-preconditioned variant")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--plain", action="store_true", help="scenes: plain PCG instead of the warm start")
    ap.add_argument("--batched", action="store_true", help="scenes: the pressure solves of all scenes batched")
    ap.add_argument("--edges", default="32,48,64,80,96", help="scenes: grid edges drawn from this list")
    ap.add_argument("--precond", choices=("jacobi", "mg"), default=None,
                    help="projection, warmstart, scenes: preconditioner of the pressure CG (default: mg, "
                         "the library's default; warmstart: jacobi, the solve the network was trained to start)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "flux-interpolated-advection_amd", "weights",
                                                      "unet_plume128.npy"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=None, help="CPU baseline grid (plume: 64, density: 128)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain device copy of 1 GiB (a known-size kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.n is None:
        args.n = 128 if args.workload == "warmstart" else 256
    if args.precond is None:
        args.precond = "jacobi" if args.workload == "warmstart" else "mg"
    if args.cpu_n is None:
        args.cpu_n = 64 if args.workload == "plume" else 128
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (one GPU's share of a "
              "sharded grid is the single-GPU workload)", file=sys.stderr)
        sys.exit(2)

    import torch
    fia = importlib.import_module("flux-interpolated-advection_amd")
    scenes = importlib.import_module("flux-interpolated-advection_amd.scenes")
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")   # 1 GiB read + 1 GiB written
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "projection":
        out = run_projection(args, fia, scenes, torch)
    elif args.workload == "warmstart":
        out = run_warmstart(args, fia, scenes, torch)
    elif args.workload == "scenes":
        out = run_scenes(args, fia, torch)
    elif args.workload == "plume":
        out = run_plume(args, fia, scenes, torch)
    else:
        out = run_advection(args, fia, scenes, torch, args.workload == "velocity")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
