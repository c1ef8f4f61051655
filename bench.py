"""Benchmark: flux-interpolated density step (move_substances_3D) at 256^3, CFL 4, on one MI355X.

Workload (BASELINE configuration "256^3 plume"): 256^3 MAC grid in the unit cube, density = a
Gaussian plume blob plus a sphere source, divergence-free seeded curl-noise velocity (wavenumbers
1-4) scaled so that dt = 4 dx / max|u| (CFL 4).  One step = one move_substances_3D: WENO6 flux
pass with integral psi (Gauss n=1), xz-CellFaceAligned split, plus the zero-velocity correction
passes -- the reference's default step.  Fields stay resident on the GPU.

Prints one JSON line: value = cell-steps per second / 1e6 over the timed steps, a roofline entry
for the face kernel measured with HIP events on the library's stream, and a CPU baseline: the
oracle's FP32 restatement of the reference path, multi-threaded, on a bounded 64^3 sample.
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
PEAK_FP32_TFLOPS = 157.3
# Algorithmic work of the quadrature: every psi point and density point is one 3D WENO6
# interpolation of the reference's rule, 43 1D evaluations of 200 flop as written in the
# reference's weno.h:34-68 (8,600 flop per point).
FLOP_PER_WENO3D_POINT = 43 * 200
# Work the kernel executes instead: its 1D evaluations (counted; the z pass is shared by the points
# of a piece and psi on a face level needs no y pass) at 100 flop each (sample differences,
# three Newton candidates, three smoothness indicators over 12 shared products, weights, blend;
# an FMA counts 2, the reciprocal 1).
FLOP_PER_WENO6_EXECUTED = 100
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of piece_kernel<1> at the default workload (256^3, CFL 4), from two
# rocprofv3 counter passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a 1 GiB device copy);
# recorded in profiles/density_step_256_cfl4_hbm_counters.md
TRAFFIC_BYTES_256_CFL4 = 11.561e9


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
    return {"value": round(n ** 3 / sec / 1e6, 5), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": "one move_substances_3D step at %d^3 (same workload, CFL %g), FP32 oracle restatement of the "
                      "reference path (per-cell faces, std::thread over x-slabs), %.1f s" % (n, cfl, sec)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cfl", type=float, default=4.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain device copy of 1 GiB (a known-size kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    fia = importlib.import_module("flux-interpolated-advection_amd")
    scenes = importlib.import_module("flux-interpolated-advection_amd.scenes")
    n = args.n
    h, rho, u, v, w, dt = make_scene(scenes, n, args.cfl)
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)   # 1 GiB read + 1 GiB written
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    T = lambda a: torch.from_numpy(a).to(dev)
    rho_d, u_d, v_d, w_d = T(rho), T(u), T(v), T(w)
    grid = fia.Grid3D(n, n, n, h, rho_d, u_d, v_d, w_d)
    ctx = fia.Context(0)
    opts = dict(split_method="xz-CellFaceAligned", interpolation_method="WENO6", use_integral=True,
                num_gauss_quadrature_point_for_integrate_density=1,
                enable_eliminate_zero_velocity_false_diffusion=True)
    for _ in range(args.warmup):
        fia.move_substances_3d(ctx, grid, 0, dt, **opts)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    ctx.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        fia.move_substances_3d(ctx, grid, i, dt, **opts)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    st = ctx.stats()
    cells = n ** 3
    value = cells * args.steps / sec / 1e6
    # roofline of the dominant kernel: the WENO6 integral quadrature (piece_kernel<1>), timed with
    # HIP events on the library's stream around each of its launches
    points = st["psi_points"] + st["density_points"]
    flop = points * FLOP_PER_WENO3D_POINT
    flop_exec = st["weno_evals"] * FLOP_PER_WENO6_EXECUTED
    quad_s = st["quad_ms"] / 1e3
    achieved = flop / quad_s / 1e12 if quad_s > 0 else 0.0
    executed = flop_exec / quad_s / 1e12 if quad_s > 0 else 0.0
    launches = max(st["face_launches"], 1)
    qlaunches = max(st["quad_launches"], 1)
    default_workload = n == 256 and args.cfl == 4.0
    traffic = TRAFFIC_BYTES_256_CFL4 if default_workload else None
    avg_quad_s = quad_s / qlaunches
    mass = float(rho_d.double().sum().item())
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "move_substances_3D density step, %d^3 plume blob + sphere source, curl-noise "
                               "velocity, CFL %g, WENO6, integral psi (Gauss 1), xz-CellFaceAligned, "
                               "zero-velocity correction on" % (n, args.cfl),
                   "grid": [n, n, n], "cfl": args.cfl, "passes_per_step": 3},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": "piece_kernel<1> (WENO6 integral quadrature)", "compute": "FP32 VALU, no MFMA",
                     "avg_launch_ms": round(st["quad_ms"] / qlaunches, 3),
                     "weno3d_points_per_launch": points / qlaunches, "flop_per_point": FLOP_PER_WENO3D_POINT,
                     "executed": {"TFLOPs": round(executed, 3), "frac": round(executed / PEAK_FP32_TFLOPS, 4),
                                  "weno6_evals_per_launch": st["weno_evals"] / qlaunches,
                                  "flop_per_weno6": FLOP_PER_WENO6_EXECUTED},
                     "hbm": None if traffic is None else {
                         "achieved_GBs": round(traffic / avg_quad_s / 1e9, 1), "peak_GBs": PEAK_HBM_GBS,
                         "frac": round(traffic / avg_quad_s / 1e9 / PEAK_HBM_GBS, 4)}},
        "kernels_ms_per_step": {k: round(st[k] / args.steps, 3)
                                for k in ("face_ms", "quad_ms", "vertex_ms", "psi_ms", "gather_ms")},
        "face_work": {"launches": st["face_launches"], "weno_pieces_per_step": st["pieces"] / args.steps,
                      "density_points_per_step": st["density_points"] / args.steps,
                      "avg_face_launch_ms": round(st["face_ms"] / launches, 3)},
        "mass_after": mass,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scenes, args.cpu_n, args.cfl, min(os.cpu_count() or 1, 16))
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
