/*
 * fia.h - C ABI of the MI355X (gfx950) flux-interpolated advection library (libfia.so).
 *
 * Every entry point is a plain C function over plain pointers and sizes.  Arrays use the
 * reference's Grid_3D MAC layout: [x][y][z] with z fastest (reference include/utils.h:92-165):
 *   density    nx     * ny     * nz
 *   u (x-face) (nx+1) * ny     * nz
 *   v (y-face) nx     * (ny+1) * nz
 *   w (z-face) nx     * ny     * (nz+1)
 * Values are FP32 (the reference's MY_FLOAT_TYPE, reference include/define_float_type.h:5).
 *
 * Pointers in fia_grid3d are either all host pointers (on_device = 0: the call copies them to the
 * GPU, runs, and copies the results back, matching the reference's in-place update of Grid_3D's
 * std::vectors) or all device pointers (on_device = 1: the call runs in place on the GPU).
 *
 * Errors: every function returns fia_status.  The reference returns void and calls exit(0) or
 * assert() on failures (reference src/facecutter.h:188-204, src/split_face_3D.cpp:54,73); here the
 * same conditions come back as status codes and fia_last_error() gives a message.
 */
#ifndef FIA_H
#define FIA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fia_status {
    FIA_OK = 0,
    FIA_ERR_INVALID_ARG = 1,   /* bad size, null pointer, unknown option value            */
    FIA_ERR_UNSUPPORTED = 2,   /* option the flux path does not implement (MacCormack...)  */
    FIA_ERR_CAPACITY = 3,      /* a face produced more polygon vertices or pieces than fit */
    FIA_ERR_NONFINITE = 4,     /* NaN/Inf in an input field or in the time step            */
    FIA_ERR_DEVICE = 5,        /* HIP runtime error                                        */
    FIA_ERR_NO_DEVICE = 6,     /* no GPU visible                                           */
    FIA_ERR_NOT_CONVERGED = 7, /* pressure solve hit its iteration limit                   */
    FIA_ERR_CUT = 8            /* a cut line crossed a piece's boundary an odd number of
                                  times (the reference's cutter exits there,
                                  src/facecutter.h:190-203)                                */
} fia_status;

/* Interpolation methods (the reference's interpolation_method strings). */
typedef enum fia_interp {
    FIA_INTERP_LINEAR = 0,      /* "linear"     : trilinear, clamp-to-edge                    */
    FIA_INTERP_WENO6 = 1,       /* "WENO6"      : 6^3 WENO6, z then y then x passes           */
    FIA_INTERP_1DY_LINEAR = 2,  /* "1Dy_linear" : linear in y inside the containing column    */
    FIA_INTERP_1DY_WENO6 = 3,   /* "1Dy_WENO6"  : WENO6 in y inside the containing column     */
    FIA_INTERP_WENO6_OPTIMIZED = 4  /* "WENO6-optimized": the density step's quadrature points
                                       interpolated by the reference's ISPC WENO6
                                       (src/sginterp3.cpp:203-383, src/weno.ispc:10-51):
                                       alpha = C / (1e-30 + S^2), coordinates clamped to the
                                       samples' range; vertices and psi as "WENO6".  Density only:
                                       the reference's velocity path has no branch for it
                                       (src/calc_velocity_from_quadrature_point_list_use_integral.cpp:85-111) */
} fia_interp;

/* Face split methods (reference src/split_face_3D.cpp:26-43).  Any other string of the
 * reference (including its default "y-CellFaceAligned") leaves faces uncut: FIA_SPLIT_NONE. */
typedef enum fia_split {
    FIA_SPLIT_NONE = 0,
    FIA_SPLIT_XZ_CELLFACE = 1,           /* "xz-CellFaceAligned"              */
    FIA_SPLIT_XZ_AXIS = 2,               /* "xz-AxisAligned"                  */
    FIA_SPLIT_X_CELLFACE_Z_AXIS = 3,     /* "x-CellFaceAligned_z-AxisAligned" */
    FIA_SPLIT_X_AXIS_Z_CELLFACE = 4      /* "x-AxisAligned_z-CellFaceAligned" */
} fia_split;

/* How the y-prefix psi is built (interpolation_method_in_calc_psi of the reference). */
typedef enum fia_psi_mode {
    FIA_PSI_CONST = 0,   /* psi(j) = psi(j-1) + q(j-1) dx                                   */
    FIA_PSI_LINEAR = 1,  /* trilinear q at y = j-3/4, j-1/4, half-cell weights               */
    FIA_PSI_WENO6 = 2    /* Gauss-Legendre over 1D-y WENO6 of q on both half cells           */
} fia_psi_mode;

/* Preconditioner of the pressure solve's conjugate gradients (the reference uses amgcl: CG with a
 * smoothed-aggregation AMG preconditioner, src/calc_pressure_3d.cpp:141-166). */
typedef enum fia_precond {
    FIA_PRECOND_JACOBI = 0,   /* diagonal                                                        */
    FIA_PRECOND_MG = 1        /* geometric multigrid (default): 2x2x2 aggregation, Galerkin coarse */
                              /* operators, one V(1,1) cycle with damped Jacobi per CG iteration  */
} fia_precond;

/* Options of move_substances_3D / advect_velocity_*_flux_advection_3D_std_thread, one field per
 * argument of the reference signature (reference include/move_substances_3d.h:48-67,
 * include/advect_velocity_flux_advection_3d.h:14-124).  Strings are parsed once by the caller
 * (the C++ shim in fia_smoke_simulation.hpp does it for the reference strings). */
typedef struct fia_advect_opts {
    int use_flux_advection;                 /* 0 selects semi-Lagrangian: FIA_ERR_UNSUPPORTED    */
    int use_maccormack;                     /* nonzero: FIA_ERR_UNSUPPORTED                      */
    int use_clamping_in_maccormack;         /* ignored (MacCormack only)                         */
    int num_gauss_quad_boundary;            /* ignored by the 3D flux path                       */
    int num_gauss_quad_bulk;                /* ignored by the 3D flux path                       */
    int enable_zero_velocity_correction;    /* enable_eliminate_zero_velocity_false_diffusion    */
    fia_split split_method;
    int num_gauss_quadrature_points;        /* ignored by the optimized split                    */
    fia_interp interpolation_method;
    int use_integral;
    int num_gauss_points_integrate_density; /* 1,3,5,7,9,11,13,25 or 49                          */
    int enable_cell_volume_correction;      /* nonzero: FIA_ERR_UNSUPPORTED                      */
    double minimal_area;                    /* triangle subdivision threshold in xi-eta space    */
} fia_advect_opts;

/* Grid_3D (reference include/grid_3d.h:14-88): sizes, cell length and the four fields. */
typedef struct fia_grid3d {
    int nx, ny, nz;
    double dx;        /* _cell_length; the domain is [0, n*dx] on every axis (min_pos = 0) */
    float *density;   /* substance_density                                                  */
    float *u;         /* velocity_in_voxel_face_x                                           */
    float *v;         /* velocity_in_voxel_face_y                                           */
    float *w;         /* velocity_in_voxel_face_z                                           */
    int on_device;    /* 1: device pointers, 0: host pointers                                */
} fia_grid3d;

typedef struct fia_ctx fia_ctx;

/* Fills opts with the reference's command-line defaults (reference src/main.cpp:48-83) and
 * interpolation_method = WENO6, split_method = the reference's default string (FIA_SPLIT_NONE). */
void fia_default_opts(fia_advect_opts *opts);

/* Context: owns the device scratch (psi, backtraced vertices, face fluxes) for the last grid
 * size it saw, and one HIP stream of default priority.  device < 0 picks the current device. */
fia_status fia_ctx_create(int device, fia_ctx **out);
fia_status fia_ctx_destroy(fia_ctx *ctx);
/* A context that enqueues its work on the caller's HIP stream (a hipStream_t of the device),
 * so that library work orders with the caller's own kernels (torch's current stream, the streams
 * of a caller's scene scheduler).  Calls still return when their results are complete: the flux path reads
 * piece-slot counts back between its phases.  The context does not destroy the stream. */
fia_status fia_ctx_create_on_stream(int device, void *stream, fia_ctx **out);
/* The stream a context enqueues on (a hipStream_t). */
void *fia_ctx_stream(fia_ctx *ctx);
/* Bytes of device memory the context holds now. */
size_t fia_ctx_device_bytes(const fia_ctx *ctx);
/* Message of the last error of this thread (never NULL). */
const char *fia_last_error(void);
/* Number of visible GPUs (0 when none); does not create a context. */
int fia_device_count(void);

/* One density step: replaces smoke_simulation::move_substances_3D
 * (reference include/move_substances_3d.h:48-67, src/move_substances_3d.cpp:38-240).
 * Main flux pass, then, if opts->enable_zero_velocity_correction, the zero-velocity pass and
 * the error advection pass (reference src/move_substances_3d.cpp:115-226).  i_frame is unused
 * by the reference's 3D path and kept for signature parity. */
fia_status fia_move_substances_3d(fia_ctx *ctx, fia_grid3d *grid, int i_frame, double dt,
                                  const fia_advect_opts *opts);

/* One flux pass over a cell-centred field q (nx*ny*nz), in place: replaces
 * advect_density_flux_advection_3D_std_thread (reference src/move_substances_3d.cpp:585-841).
 * interp drives velocity and q interpolation, psi_mode the psi build, zero_velocity the
 * use_zero_velocity_for_backtrace flag.  q and the grid pointers follow grid->on_device. */
fia_status fia_advect_flux_pass(fia_ctx *ctx, const fia_grid3d *grid, float *q, double dt,
                                fia_split split, fia_interp interp, fia_psi_mode psi_mode,
                                int zero_velocity, int use_integral, int num_gauss_density,
                                double minimal_area);

/* Velocity self-advection of one component (dim 0,1,2) or of all three (dim = -1, every
 * component backtraced through the pre-step velocity): replaces
 * advect_velocity_{x,y,z,xyz}_flux_advection_3D_std_thread
 * (reference include/advect_velocity_flux_advection_3d.h:38-124,
 * src/advect_velocity_flux_advection_3d.cpp:580-1272). */
fia_status fia_advect_velocity_3d(fia_ctx *ctx, fia_grid3d *grid, int dim, double dt,
                                  const fia_advect_opts *opts);

/* psi_y of a cell-centred field of size nx*ny*nz into psi (nx*(ny+1)*nz, FP64, in units of
 * q*dx): replaces calc_psi_on_cell_face_from_density_on_cell_center_3D
 * (reference src/calc_psi_3D.cpp:10-308).  Pointers follow on_device. */
fia_status fia_calc_psi_y(fia_ctx *ctx, int nx, int ny, int nz, double dx, const float *q,
                          double *psi, fia_psi_mode mode, int use_integral, int num_gauss,
                          int on_device);

/* Interpolates one staggered field at n points: replaces WENO6_interpolation_3D_*_values,
 * linear_interpolation_3D_*, linear_interpolation_1Dy_* and WENO6_interpolation_1Dy_*
 * (reference src/WENO6_interpolation_3d.cpp:7-1020, src/linear_interpolation_3d.cpp:7-567,
 * src/linear_interpolation_1d.cpp:314-702, src/WENO6_interpolation_1d.cpp:315-779).
 * The field has dims (mx,my,mz) and sample point (i,j,k) sits at ((i+ox)dx,(j+oy)dx,(k+oz)dx);
 * points are xyz triples in physical units.  Pointers follow on_device. */
fia_status fia_interpolate(fia_ctx *ctx, const float *field, int mx, int my, int mz,
                           double ox, double oy, double oz, double dx, fia_interp method,
                           const float *points_xyz, size_t n, float *out, int on_device);

/* Backtraced positions of every vertex of the (component) lattice: replaces
 * calc_backtraced_face_3D + clamp_vertex_position_by_grid_3D applied to all faces
 * (reference src/calc_backtraced_face_3D.cpp:9-159).  dim = -1 for density, 0..2 for the
 * velocity component lattices.  out holds 3 arrays of (nx+1+ex)(ny+1+ey)(nz+1+ez) floats:
 * the backtraced positions minus the original ones, in units of dx. */
fia_status fia_backtrace_vertices(fia_ctx *ctx, const fia_grid3d *grid, int dim, double dt,
                                  fia_interp interp, float *disp_xyz, int on_device);

/* Splits n faces given by 4 vertex positions each (n*12 floats, physical units, reference
 * vertex order) with the face cutter (reference src/split_face_3D.cpp:17-93,
 * src/facecutter.h:59-324).  For each face writes at most max_pieces pieces:
 * piece_count[n], centroid_xyz[n*max_pieces*3], area_xyz[n*max_pieces*3] (vector areas).
 * FIA_ERR_CAPACITY when a face has more pieces than max_pieces.  Pointers follow on_device. */
fia_status fia_split_faces(fia_ctx *ctx, const float *quads_xyz, size_t n, double dx,
                           fia_split split, double minimal_area, int max_pieces,
                           int *piece_count, double *centroid_xyz, double *area_xyz,
                           int on_device);

/* CFL time step: replaces calc_time_step_length_from_CFL_number_3D
 * (reference src/calc_time_step_length_from_CFL_number.cpp:81-132).  cfl_number is an int as
 * in the reference (include/calc_time_step_length_from_CFL_number.h:38).  When now_time+dt
 * passes the next movie frame the step is shortened to it and *write_frame is set to 1. */
fia_status fia_cfl_time_step(fia_ctx *ctx, const fia_grid3d *grid, double now_time,
                             int i_movie_frame, int cfl_number, double movie_fps,
                             double *dt_out, int *write_frame);

/* Pressure projection: replaces calc_pressure_gradient_term_3D
 * (reference src/calc_pressure_3d.cpp:18-269).  Solves the reference's 7-point Neumann system
 * (last cell pinned to p = 0) with preconditioned CG (the context's preconditioner, geometric
 * multigrid unless fia_ctx_set_pressure_preconditioner chose Jacobi), FP64 dots and residuals, to
 * ||b - A p|| <= rel_tol * ||b||, then u -= grad p / dx on interior faces.  solid (optional,
 * nx*ny*nz bytes, nonzero = obstacle cell) turns faces next to obstacles into Neumann walls and
 * zeroes the velocity on them.  pressure (optional, nx*ny*nz floats) receives p and, when
 * warm_start is nonzero, is read as the initial guess.  iters/rel_residual are outputs. */
fia_status fia_pressure_project(fia_ctx *ctx, fia_grid3d *grid, const uint8_t *solid,
                                float *pressure, int warm_start, double rel_tol, int max_iters,
                                int *iters, double *rel_residual);

/* Instrumentation (no counterpart in the reference, which only prints wall-clock phases at
 * src/move_substances_3d.cpp:607-840): when enabled, the library records HIP events around every
 * face-kernel launch on its own stream and counts the quadrature work those launches did. */
typedef struct fia_kernel_stats {
    double face_ms;              /* summed face-phase time (HIP events on the context's stream)   */
    long long face_launches;     /* number of face phases (one per flux pass)                     */
    long long pieces;            /* face pieces integrated by the WENO6 integral quadrature       */
    long long psi_points;        /* its psi points (one per piece)                                */
    long long density_points;    /* its density quadrature points                                 */
    long long vertices;          /* backtraced lattice vertices                                   */
    double vertex_ms;            /* summed vertex-kernel time                                     */
    double psi_ms;               /* summed psi-kernel time                                        */
    double gather_ms;            /* summed gather-kernel time                                     */
    long long gather_cells;      /* cells written by gather kernels                               */
    long long weno_evals;        /* 1D WENO6 evaluations of the WENO6 integral quadrature kernel   */
    double quad_ms;              /* summed time of the WENO6 integral quadrature launches alone    */
    long long quad_launches;     /* number of those launches                                       */
    double pcg_ms;               /* summed time of the pressure solves' CG iteration launches       */
    long long pcg_iters;         /* CG iterations those solves ran                                  */
    long long pcg_solves;        /* pressure solves                                                 */
    long long twisted_faces;     /* faces cut by the recursive (general) cutter, both orientations  */
    long long general_pieces;    /* the pieces it made of them                                      */
    long long graph_captures;    /* CG iteration batches captured into HIP graphs (fia_ctx_set_graphs) */
    long long graph_replays;     /* CG iteration batches launched from captured graphs              */
} fia_kernel_stats;

fia_status fia_ctx_set_timing(fia_ctx *ctx, int enable);
/* The preconditioner of every later pressure solve on this context (fia_pressure_project,
 * fia_pressure_project_learned, fia_scenes_step); a new context uses FIA_PRECOND_MG.  No
 * counterpart in the reference, whose solver is fixed; fia_scenes_step_batched takes it from
 * ctxs[0]. */
fia_status fia_ctx_set_pressure_preconditioner(fia_ctx *ctx, fia_precond kind);
/* Whether the pressure solves of this context replay their CG iteration batches from captured HIP
 * graphs (on by default; a new capture only when the solve's buffers, grid or multigrid levels
 * change, up to 16 kept per context).  The results are bit-identical either way.  No counterpart
 * in the reference (amgcl on the CPU, src/calc_pressure_3d.cpp:141-166).  Synchronises; disabling
 * frees the captured graphs. */
fia_status fia_ctx_set_graphs(fia_ctx *ctx, int enable);

/* ---- Launch shapes and their autotuner ------------------------------------------------------
 * No counterpart in the reference (std::thread fan-out on the CPU).  Thread-block shapes of the
 * advection's face kernels and of the CG sweep, kept per grid shape in the context.  The face,
 * piece and cell block sizes leave every result bit-identical (each face, piece and cell is one
 * thread's work, summed in a fixed order); the CG chunking changes the grouping of the FP64
 * partial sums, so solves agree to their tolerance, not bit for bit. */
typedef struct fia_tuning {
    int face_block;      /* faces per block of the count, cutter, sum and zero-velocity kernels:
                            64, 128 or 256 (0: 256)                                              */
    int piece_block;     /* pieces per block of the quadrature kernel: 64, 128 or 256 (0: 256)   */
    int cell_block;      /* cells per block of the gather kernel: 64, 128 or 256 (0: 256)        */
    int pcg_chunk_x;     /* x planes one 64 x 16 CG tile sweeps (0: one block per resident slot) */
    int pcg_recompute;   /* the CG update recomputes q = L p (1) or streams it (0); -1: by size  */
    int pcg_fused;       /* Jacobi: one sweep per CG iteration (1) or the direction and update
                            sweeps (0); -1: fused from 4M cells.  Multigrid and batched solves
                            use two sweeps. */
} fia_tuning;
/* The shapes used on grids of nx x ny x nz cells (their staggered velocity components and
 * pressure systems included); nx = ny = nz = 0 sets them for shapes without an entry of their
 * own.  At most 64 shapes per context. */
fia_status fia_ctx_set_tuning(fia_ctx *ctx, int nx, int ny, int nz, const fia_tuning *t);
fia_status fia_ctx_get_tuning(fia_ctx *ctx, int nx, int ny, int nz, fia_tuning *t);
/* Times candidate shapes on copies of grid g (g is left as it was): density steps with opts and
 * time step dt for the block sizes (64, 128, 256 for each in turn), then iterations of g's
 * pressure system (obstacle mask solid, may be NULL; the context's preconditioner) for the CG
 * chunk length (2 .. 128 planes) and q policy.  The fastest of reps timed runs counts; a
 * candidate must beat the current choice by 2%.  The result is stored for g's shape and
 * returned in *out (may be NULL).  Allocates its copies; synchronises. */
fia_status fia_ctx_autotune(fia_ctx *ctx, const fia_grid3d *g, const uint8_t *solid, double dt,
                            const fia_advect_opts *opts, int reps, fia_tuning *out);
fia_status fia_ctx_get_stats(fia_ctx *ctx, fia_kernel_stats *out);   /* synchronises */
fia_status fia_ctx_reset_stats(fia_ctx *ctx);

/* ---- Learned warm start of the pressure solve -----------------------------------------------
 * No counterpart in the reference, whose solver is amgcl (src/calc_pressure_3d.cpp:141-166): a
 * 4-level 3-D U-Net (channels 16/32/48/64, 3x3x3 convolutions, BF16 weights and activations on
 * the matrix cores, FP32 accumulation) maps the right-hand side of the pressure system and the
 * obstacle mask to an initial guess for the same PCG (trained against the Jacobi-preconditioned
 * solve; it also runs under multigrid).  Grid sizes must be multiples of 8.
 * Parameters are FP32, fia_unet_param_count() of them (layout in fia_unet.hpp). */
typedef struct fia_unet fia_unet;
size_t fia_unet_param_count(void);
/* network on the context's device, He-normal initial weights drawn from seed */
fia_status fia_unet_create(fia_ctx *ctx, unsigned long long seed, fia_unet **out);
fia_status fia_unet_destroy(fia_unet *net);
fia_status fia_unet_set_params(fia_ctx *ctx, fia_unet *net, const float *params, size_t n);   /* host */
fia_status fia_unet_get_params(fia_ctx *ctx, fia_unet *net, float *params, size_t n);         /* host */
/* y = network output for right-hand side rhs (nx*ny*nz FP64) and obstacle mask solid (optional),
 * in units of *scale = sqrt(mean rhs^2 over fluid cells); pointers on the device when on_device */
fia_status fia_unet_forward(fia_ctx *ctx, fia_unet *net, int nx, int ny, int nz, const double *rhs,
                            const uint8_t *solid, float *y, double *scale, int on_device);
/* Training losses of the network output y against b = rhs/scale on the fluid cells (L: their
 * Neumann Laplacian, no pinned cell): the residual norm |L y - b|^2 / |b|^2, or the energy
 * (y.L y / 2 - b.y) / |b|^2, which differs from half the squared A-norm error of y by a constant. */
enum { FIA_UNET_LOSS_RESIDUAL = 0, FIA_UNET_LOSS_ENERGY = 1 };
/* loss of one sample, its residual norm |L y - b|^2 / |b|^2 and, when grad is not null, the loss
 * gradient with respect to every parameter (host array of fia_unet_param_count() floats); no update */
fia_status fia_unet_loss_grad(fia_ctx *ctx, fia_unet *net, int nx, int ny, int nz, const double *rhs,
                              const uint8_t *solid, int loss_mode, double *loss, double *residual, float *grad,
                              int on_device);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on the mean gradient of `batch` samples stored one
 * after another in rhs (device, batch*nx*ny*nz FP64) and solid (device, same count, optional);
 * the batch means of loss and residual norm before the step are returned */
fia_status fia_unet_train_step(fia_ctx *ctx, fia_unet *net, int nx, int ny, int nz, int batch,
                               const double *rhs, const uint8_t *solid, int loss_mode, float lr, double *mean_loss,
                               double *mean_residual);

/* The right-hand side b = -div(u) dx of the pressure system of grid (0 on obstacle cells and on
 * the pinned cell), as fia_pressure_project assembles it: training data for the network. */
fia_status fia_pressure_rhs(fia_ctx *ctx, fia_grid3d *grid, const uint8_t *solid, double *rhs);

typedef struct fia_learned_info {
    int used_network;            /* 1: continued from the network's guess; 0: fell back to zero guess */
    int probe_iters;             /* iterations run from both starts before choosing                     */
    double guess_scale;          /* energy-optimal factor x0.b / x0.L x0 applied to the network's guess  */
    double guess_rel_residual;   /* |b - L x0| / |b| of the (scaled) guess                                */
    double probe_res_network;    /* relative residual after the probe from the network's guess          */
    double probe_res_zero;       /* ... and from a zero guess                                            */
} fia_learned_info;

/* fia_pressure_project with the learned warm start: the network's guess, scaled by the factor that
 * minimises its energy error, and a zero guess each run probe iterations (5); the solve continues
 * from the network's unless its residual is larger (then from the zero guess).  Same system, tolerance and outputs as fia_pressure_project; iters
 * counts the iterations of the solve that was continued. */
fia_status fia_pressure_project_learned(fia_ctx *ctx, fia_grid3d *grid, const uint8_t *solid, float *pressure,
                                        fia_unet *net, double rel_tol, int max_iters, int *iters,
                                        double *rel_residual, fia_learned_info *info);

/* Diagnostics written to the reference's *.dat series (src/main.cpp:518-597): totals[0] = total
 * mass sum(rho) dx^3 (0 when grid->density is NULL), totals[1..3] = total momenta sum(u) dx^3,
 * sum(v) dx^3, sum(w) dx^3; FP64 sums in a fixed order. */
fia_status fia_grid_totals(fia_ctx *ctx, fia_grid3d *grid, double totals[4]);

/* fia_grid_totals without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid it).  It computes the same sums with the same kernels (bit-identical totals) into a
 * fixed scratch of 12,352 bytes that every GPU the library runs on holds until the process ends,
 * set up when the first context on that GPU is created.  The grid's fields must be on the device
 * (on_device = 1; FIA_ERR_INVALID_ARG otherwise); any grid size is accepted, the scratch holds a
 * fixed number of partial sums.  Calls that overlap on one GPU share that scratch and race: do not
 * run two at once on the same GPU (on any contexts or streams). */
fia_status fia_grid_totals_noalloc(fia_ctx *ctx, const fia_grid3d *grid, double totals[4]);
/* Momentum the walls take out through the pressure of one projection (src/calc_pressure_3d.cpp:
 * 235-266): loss[0] = dx^2 (sum of p over the cells at x = nx-1 - sum over x = 0), likewise y, z.
 * The reference adds it to Grid_3D::total_loss_of_momentum_{x,y,z} after every projection. */
fia_status fia_wall_momentum_loss(fia_ctx *ctx, int nx, int ny, int nz, double dx, const float *pressure,
                                  int on_device, double loss[3]);

/* ---- Many independent scenes ------------------------------------------------------------------
 * No counterpart in the reference, which runs one grid per process.  A scene is a device grid with
 * its own forcing and clock; one step follows the order of the reference's main loop
 * (src/main.cpp:596-740): dt from the CFL number, forcing, velocity self-advection (xyz), pressure
 * projection, density advection. */
typedef struct fia_scene {
    fia_grid3d grid;             /* device fields (on_device = 1), density required                   */
    const uint8_t *solid;        /* device obstacle mask or NULL                                      */
    float *pressure;             /* device nx*ny*nz floats receiving p, or NULL                       */
    double source_x, source_y, source_z, source_radius, source_value;   /* density source sphere   */
    double buoyancy;             /* v += dt * buoyancy * rho (y is up)                                */
    int cfl;                     /* CFL number (an int, as the reference's)                           */
    double movie_fps;            /* frame clipping of dt                                              */
    double time;                 /* in/out: simulated time                                            */
    int frame;                   /* in/out: movie frame                                               */
    /* outputs of the last step */
    double dt;
    int pcg_iters;
    double pcg_residual;
    int used_network;
    int steps;                   /* steps taken                                                       */
    fia_status status;
} fia_scene;

/* The forcing of one scene for a step of length dt: density = max(density, source_value) inside
 * the source sphere; v += dt * buoyancy * (density averaged to the y face) on interior y faces. */
fia_status fia_scene_forcing(fia_ctx *ctx, fia_scene *scene, double dt);
/* One step of every scene.  workers threads, each with its own context (ctxs[w]; one stream and
 * scratch each) and, for the learned warm start, its own network (nets[w], or nets NULL for plain
 * PCG), take scenes largest first.  Scenes are independent, so the result does not depend on
 * workers.  Returns the first failure (every scene's own status is in scene->status). */
fia_status fia_scenes_step(fia_ctx *const *ctxs, int workers, fia_scene *scenes, int n_scenes,
                           const fia_advect_opts *density_opts, const fia_advect_opts *velocity_opts,
                           double rel_tol, int max_iters, fia_unet *const *nets);
/* The same step with the scenes' pressure solves batched: the workers advance every scene to its
 * pressure system and start its solve (as fia_pressure_project, or with nets as
 * fia_pressure_project_learned, probe included); then the CG iterations of all scenes run
 * together on ctxs[0] (one direction and one update launch per iteration for all of them) and each
 * scene is projected; then the workers advect the densities.  Every scene's arithmetic is that of
 * fia_scenes_step, so the results are identical; many small scenes need far fewer launches and
 * polls.  ctxs[0] keeps the solve buffers of all scenes (about 49 bytes per cell, 89 with nets). */
fia_status fia_scenes_step_batched(fia_ctx *const *ctxs, int workers, fia_scene *scenes, int n_scenes,
                                   const fia_advect_opts *density_opts, const fia_advect_opts *velocity_opts,
                                   double rel_tol, int max_iters, fia_unet *const *nets);

#ifdef __cplusplus
}
#endif

#endif /* FIA_H */
