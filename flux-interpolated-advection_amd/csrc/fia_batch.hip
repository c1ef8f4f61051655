// fia_batch.hip - many independent scenes of uneven sizes advanced together (no counterpart in
// the reference, which runs one grid per process).
//
// A step of a scene is the order of the reference's main loop (src/main.cpp:596-740): dt from
// the CFL number, forcing (a density source and buoyancy on the vertical velocity: the scene's
// own terms, the reference has none), velocity self-advection, pressure projection, density
// advection.  Scenes are independent, so they go to worker threads, each owning one library
// context (one HIP stream and its scratch): small grids are launch- and sync-bound, and the
// streams of different workers overlap on the GPU.  Scenes are taken largest first from a shared
// counter (longest-processing-time order), so the step ends with small scenes.
//
// fia_scenes_step_batched splits the step into three phases with the pressure solves of all
// scenes in the middle one: a CG iteration of a small grid is two short launches and its
// convergence polls are host round trips, so one launch per kernel for every scene and one poll
// for all of them replace many.
#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fia.h"
#include "fia_internal.hpp"
#include "fia_math.hpp"

namespace fia {

// rho = max(rho, value) inside the source sphere (cell centres)
__global__ void forcing_kernel(int nx, int ny, int nz, double h, float *rho, double cx, double cy, double cz, double r,
                               double value) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (c >= n) return;
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((size_t)ny * nz));
    const double x = (i + 0.5) * h - cx, y = (j + 0.5) * h - cy, z = (k + 0.5) * h - cz;
    if (r > 0.0 && x * x + y * y + z * z <= r * r) rho[c] = fmaxf(rho[c], (float)value);
}

// v += dt * beta * (rho averaged to the y face) on interior y faces (y is up)
__global__ void buoyancy_kernel(int nx, int ny, int nz, const float *rho, float *v, double dt_beta) {
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * (ny + 1) * nz;
    if (f >= n) return;
    const int k = (int)(f % nz), j = (int)((f / nz) % (ny + 1)), i = (int)(f / ((size_t)(ny + 1) * nz));
    if (j == 0 || j == ny) return;
    const double avg = 0.5 * ((double)rho[idx3(i, j, k, ny, nz)] + (double)rho[idx3(i, j - 1, k, ny, nz)]);
    v[f] = (float)((double)v[f] + dt_beta * avg);
}

hipError_t launch_forcing(const fia_grid3d &g, double cx, double cy, double cz, double r, double value, double dt_beta,
                          hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.nz, nv = (size_t)g.nx * (g.ny + 1) * g.nz;
    hipLaunchKernelGGL(forcing_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.nx, g.ny, g.nz, g.dx,
                       g.density, cx, cy, cz, r, value);
    if (dt_beta != 0.0)
        hipLaunchKernelGGL(buoyancy_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, g.nx, g.ny, g.nz,
                           g.density, g.v, dt_beta);
    return hipGetLastError();
}

}  // namespace fia

namespace {

// f(worker, scene index) over scenes[order[..]] on `workers` threads taking scenes from a shared
// counter; scene->status records each failure.  Returns the first failure, its message handed to
// the caller's thread (the messages are thread-local).
template <class F>
fia_status for_scenes(int workers, fia_scene *scenes, const std::vector<int> &order, F f) {
    const int n = (int)order.size();
    std::atomic<int> next{0};
    std::vector<fia_status> status(workers, FIA_OK);
    std::vector<std::string> errors(workers);
    auto run = [&](int w) {
        while (true) {
            const int k = next.fetch_add(1);
            if (k >= n) return;
            const fia_status st = f(w, order[k]);
            scenes[order[k]].status = st;
            if (st != FIA_OK && status[w] == FIA_OK) {
                status[w] = st;
                errors[w] = fia_last_error();
            }
        }
    };
    const int nthreads = std::min(workers, std::max(n, 1));
    if (nthreads == 1) {
        run(0);
    } else {
        std::vector<std::thread> pool;
        for (int w = 0; w < nthreads; ++w) pool.emplace_back(run, w);
        for (auto &t : pool) t.join();
    }
    for (int w = 0; w < nthreads; ++w)
        if (status[w] != FIA_OK) return fia::set_error(status[w], errors[w]);
    return FIA_OK;
}

// scene indices, largest grid first (stable)
std::vector<int> largest_first(const fia_scene *scenes, int n) {
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const fia_grid3d &ga = scenes[a].grid, &gb = scenes[b].grid;
        return (long long)ga.nx * ga.ny * ga.nz > (long long)gb.nx * gb.ny * gb.nz;
    });
    return order;
}

// Marks the contexts of a step of several scenes as sharing the device for the duration of the step.
struct SharedDevice {
    fia_ctx *const *ctxs;
    int workers;
    SharedDevice(fia_ctx *const *c, int w, bool shared) : ctxs(c), workers(w) {
        for (int i = 0; i < workers; ++i) fia::ctx_set_shared(ctxs[i], shared);
    }
    ~SharedDevice() {
        for (int i = 0; i < workers; ++i) fia::ctx_set_shared(ctxs[i], false);
    }
};

fia_status check_step_args(fia_ctx *const *ctxs, int workers, const fia_scene *scenes, int n_scenes,
                           const fia_advect_opts *od, const fia_advect_opts *ov) {
    if (!ctxs || workers <= 0 || (!scenes && n_scenes > 0) || n_scenes < 0 || !od || !ov)
        return fia::set_error(FIA_ERR_INVALID_ARG, "bad argument");
    for (int w = 0; w < workers; ++w)
        if (!ctxs[w]) return fia::set_error(FIA_ERR_INVALID_ARG, "NULL context");
    return FIA_OK;
}

// one scene step on one context; returns the first failing status
fia_status scene_step(fia_ctx *c, fia_scene *sc, const fia_advect_opts *od, const fia_advect_opts *ov, double rel_tol,
                      int max_iters, fia_unet *net) {
    fia_grid3d *g = &sc->grid;
    if (!g->on_device || !g->density) return fia::set_error(FIA_ERR_INVALID_ARG, "a scene needs a device grid with density");
    fia_status st;
    double dt = 0.0;
    int write = 0;
    if ((st = fia_cfl_time_step(c, g, sc->time, sc->frame, sc->cfl, sc->movie_fps, &dt, &write)) != FIA_OK) return st;
    if ((st = fia_scene_forcing(c, sc, dt)) != FIA_OK) return st;
    if ((st = fia_advect_velocity_3d(c, g, -1, dt, ov)) != FIA_OK) return st;
    int iters = 0;
    double res = 0.0;
    if (net) {
        fia_learned_info info;
        st = fia_pressure_project_learned(c, g, sc->solid, sc->pressure, net, rel_tol, max_iters, &iters, &res, &info);
        sc->used_network = info.used_network;
    } else {
        st = fia_pressure_project(c, g, sc->solid, sc->pressure, 0, rel_tol, max_iters, &iters, &res);
        sc->used_network = 0;
    }
    if (st != FIA_OK) return st;
    if ((st = fia_move_substances_3d(c, g, sc->frame, dt, od)) != FIA_OK) return st;
    sc->dt = dt;
    sc->pcg_iters = iters;
    sc->pcg_residual = res;
    sc->time += dt;
    if (write) sc->frame += 1;
    sc->steps += 1;
    return FIA_OK;
}

}  // namespace

extern "C" {

fia_status fia_scene_forcing(fia_ctx *c, fia_scene *sc, double dt) {
    if (!c || !sc) return fia::set_error(FIA_ERR_INVALID_ARG, "ctx or scene is NULL");
    const fia_grid3d &g = sc->grid;
    if (!g.on_device || !g.density || !g.v || g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
        return fia::set_error(FIA_ERR_INVALID_ARG, "a scene needs a device grid with density and velocity");
    if (hipSetDevice(fia::ctx_device(c)) != hipSuccess) return fia::set_error(FIA_ERR_DEVICE, "hipSetDevice failed");
    const hipError_t e = fia::launch_forcing(g, sc->source_x, sc->source_y, sc->source_z, sc->source_radius,
                                             sc->source_value, dt * sc->buoyancy, fia::ctx_stream(c));
    if (e != hipSuccess) return fia::set_error(FIA_ERR_DEVICE, std::string("forcing: ") + hipGetErrorString(e));
    return FIA_OK;
}

fia_status fia_scenes_step(fia_ctx *const *ctxs, int workers, fia_scene *scenes, int n_scenes,
                           const fia_advect_opts *density_opts, const fia_advect_opts *velocity_opts, double rel_tol,
                           int max_iters, fia_unet *const *nets) {
    fia_status st;
    if ((st = check_step_args(ctxs, workers, scenes, n_scenes, density_opts, velocity_opts)) != FIA_OK) return st;
    const SharedDevice shared(ctxs, workers, n_scenes > 1);
    return for_scenes(workers, scenes, largest_first(scenes, n_scenes), [&](int w, int i) {
        return scene_step(ctxs[w], &scenes[i], density_opts, velocity_opts, rel_tol, max_iters,
                          nets ? nets[w] : nullptr);
    });
}

fia_status fia_scenes_step_batched(fia_ctx *const *ctxs, int workers, fia_scene *scenes, int n_scenes,
                                   const fia_advect_opts *density_opts, const fia_advect_opts *velocity_opts,
                                   double rel_tol, int max_iters, fia_unet *const *nets) {
    fia_status st;
    if ((st = check_step_args(ctxs, workers, scenes, n_scenes, density_opts, velocity_opts)) != FIA_OK) return st;
    if (!(rel_tol > 0) || max_iters < 0) return fia::set_error(FIA_ERR_INVALID_ARG, "bad tolerance or iteration limit");
    if (nets)
        for (int w = 0; w < workers; ++w)
            if (!nets[w]) return fia::set_error(FIA_ERR_INVALID_ARG, "NULL network");
    for (int i = 0; i < n_scenes; ++i) {
        const fia_grid3d &g = scenes[i].grid;
        if (!g.on_device || !g.density || g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
            return fia::set_error(FIA_ERR_INVALID_ARG, "a scene needs a device grid with density");
    }
    fia_ctx *main = ctxs[0];
    const SharedDevice shared(ctxs, workers, n_scenes > 1);
    if ((st = fia::scenes_solve_alloc(main, scenes, n_scenes)) != FIA_OK) return st;
    const std::vector<int> order = largest_first(scenes, n_scenes);
    std::vector<double> dt(n_scenes, 0.0);
    std::vector<int> write(n_scenes, 0);
    // A: dt, forcing, velocity self-advection, pressure system and the start of its solve
    const fia_status sa = for_scenes(workers, scenes, order, [&](int w, int i) {
        fia_scene *sc = &scenes[i];
        fia_status e;
        if ((e = fia_cfl_time_step(ctxs[w], &sc->grid, sc->time, sc->frame, sc->cfl, sc->movie_fps, &dt[i],
                                   &write[i])) != FIA_OK)
            return e;
        if ((e = fia_scene_forcing(ctxs[w], sc, dt[i])) != FIA_OK) return e;
        if ((e = fia_advect_velocity_3d(ctxs[w], &sc->grid, -1, dt[i], velocity_opts)) != FIA_OK) return e;
        return fia::scene_solve_setup(ctxs[w], main, i, sc, nets ? nets[w] : nullptr, rel_tol, max_iters);
    });
    const std::string msg_a = sa != FIA_OK ? fia_last_error() : "";
    // B: the solves of the scenes that got this far, together, and the projections
    std::vector<int> live;
    for (int i = 0; i < n_scenes; ++i)
        if (scenes[i].status == FIA_OK) live.push_back(i);
    if ((st = fia::scenes_solve_run(main, scenes, live.data(), (int)live.size(), max_iters)) != FIA_OK) return st;
    fia_status sb = FIA_OK;
    std::vector<int> order_c;
    for (int i : order) {
        if (scenes[i].status == FIA_OK) order_c.push_back(i);
        else if (scenes[i].status == FIA_ERR_NOT_CONVERGED)
            sb = FIA_ERR_NOT_CONVERGED;
    }
    // C: density advection, the clock
    const fia_status sc_ = for_scenes(workers, scenes, order_c, [&](int w, int i) {
        fia_scene *sc = &scenes[i];
        fia_status e;
        if ((e = fia_move_substances_3d(ctxs[w], &sc->grid, sc->frame, dt[i], density_opts)) != FIA_OK) return e;
        sc->dt = dt[i];
        sc->time += dt[i];
        if (write[i]) sc->frame += 1;
        sc->steps += 1;
        return FIA_OK;
    });
    // the first failure: A's, then a solve that did not converge, then C's
    if (sa != FIA_OK) return fia::set_error(sa, msg_a);
    if (sb != FIA_OK) return fia::set_error(sb, "a scene's PCG reached max_iters before rel_tol");
    return sc_;
}

}  // extern "C"
