"""MI355X flux-interpolated advection: Python host module over libfia.so (C ABI in include/fia.h).

The functions mirror the reference's entry points and argument names so a caller of the
reference path can switch:

    move_substances_3d(ctx, grid, i_frame, dt, ...)          reference include/move_substances_3d.h:48-67
    advect_velocity_xyz_flux_advection_3d(ctx, grid, dt,...) reference include/advect_velocity_flux_advection_3d.h:110-124
    calc_time_step_length_from_cfl_number_3d(...)            reference src/calc_time_step_length_from_CFL_number.cpp:81-132
    calc_pressure_gradient_term_3d(ctx, grid, ...)           reference src/calc_pressure_3d.cpp:183-269

Arrays are numpy float32 (host: each call copies to the GPU and back, like the reference's
in-place update of Grid_3D) or torch float32 tensors on a ROCm device (device pointers: the call
runs in place on the GPU).  Everything runs in the HIP library; if libfia.so is missing the import
of the first function raises -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfia.so")

FIA_OK = 0
STATUS = {0: "FIA_OK", 1: "FIA_ERR_INVALID_ARG", 2: "FIA_ERR_UNSUPPORTED", 3: "FIA_ERR_CAPACITY",
          4: "FIA_ERR_NONFINITE", 5: "FIA_ERR_DEVICE", 6: "FIA_ERR_NO_DEVICE", 7: "FIA_ERR_NOT_CONVERGED",
          8: "FIA_ERR_CUT"}

# reference strings -> enums (reference src/split_face_3D.cpp:26-43, src/main.cpp:60,79)
INTERP = {"linear": 0, "WENO6": 1, "1Dy_linear": 2, "1Dy_WENO6": 3, "WENO6-optimized": 4}
SPLIT = {"xz-CellFaceAligned": 1, "xz-AxisAligned": 2, "x-CellFaceAligned_z-AxisAligned": 3,
         "x-AxisAligned_z-CellFaceAligned": 4}
PSI = {"const": 0, "linear": 1, "WENO6": 2}
UNET_LOSS = {"residual": 0, "energy": 1}

# exported symbols of libfia.so (the declarations of include/fia.h)
EXPORTS = ["fia_default_opts", "fia_ctx_create", "fia_ctx_create_on_stream", "fia_ctx_stream", "fia_ctx_destroy", "fia_ctx_device_bytes", "fia_last_error",
           "fia_device_count", "fia_move_substances_3d", "fia_advect_flux_pass", "fia_advect_velocity_3d",
           "fia_calc_psi_y", "fia_interpolate", "fia_backtrace_vertices", "fia_split_faces", "fia_cfl_time_step",
           "fia_pressure_project", "fia_ctx_set_timing", "fia_ctx_get_stats", "fia_ctx_reset_stats",
           "fia_unet_param_count", "fia_unet_create", "fia_unet_destroy", "fia_unet_set_params",
           "fia_unet_get_params", "fia_unet_forward", "fia_unet_loss_grad", "fia_unet_train_step", "fia_pressure_rhs",
           "fia_pressure_project_learned", "fia_grid_totals", "fia_grid_totals_noalloc", "fia_wall_momentum_loss", "fia_scene_forcing",
           "fia_scenes_step", "fia_scenes_step_batched", "fia_ctx_set_pressure_preconditioner",
           "fia_ctx_set_graphs", "fia_ctx_set_tuning", "fia_ctx_get_tuning", "fia_ctx_autotune"]


class FiaError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS.get(status, status), message))
        self.status = status


class AdvectOpts(ctypes.Structure):
    _fields_ = [("use_flux_advection", ctypes.c_int), ("use_maccormack", ctypes.c_int),
                ("use_clamping_in_maccormack", ctypes.c_int), ("num_gauss_quad_boundary", ctypes.c_int),
                ("num_gauss_quad_bulk", ctypes.c_int), ("enable_zero_velocity_correction", ctypes.c_int),
                ("split_method", ctypes.c_int), ("num_gauss_quadrature_points", ctypes.c_int),
                ("interpolation_method", ctypes.c_int), ("use_integral", ctypes.c_int),
                ("num_gauss_points_integrate_density", ctypes.c_int),
                ("enable_cell_volume_correction", ctypes.c_int), ("minimal_area", ctypes.c_double)]


class GridC(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_double),
                ("density", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("w", ctypes.c_void_p), ("on_device", ctypes.c_int)]


class LearnedInfo(ctypes.Structure):
    _fields_ = [("used_network", ctypes.c_int), ("probe_iters", ctypes.c_int), ("guess_scale", ctypes.c_double),
                ("guess_rel_residual", ctypes.c_double), ("probe_res_network", ctypes.c_double),
                ("probe_res_zero", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SceneC(ctypes.Structure):
    """fia_scene (include/fia.h)."""
    _fields_ = [("grid", GridC), ("solid", ctypes.c_void_p), ("pressure", ctypes.c_void_p),
                ("source_x", ctypes.c_double), ("source_y", ctypes.c_double), ("source_z", ctypes.c_double),
                ("source_radius", ctypes.c_double), ("source_value", ctypes.c_double), ("buoyancy", ctypes.c_double),
                ("cfl", ctypes.c_int), ("movie_fps", ctypes.c_double), ("time", ctypes.c_double), ("frame", ctypes.c_int),
                ("dt", ctypes.c_double), ("pcg_iters", ctypes.c_int), ("pcg_residual", ctypes.c_double),
                ("used_network", ctypes.c_int), ("steps", ctypes.c_int), ("status", ctypes.c_int)]


class KernelStats(ctypes.Structure):
    _fields_ = [("face_ms", ctypes.c_double), ("face_launches", ctypes.c_longlong), ("pieces", ctypes.c_longlong),
                ("psi_points", ctypes.c_longlong), ("density_points", ctypes.c_longlong),
                ("vertices", ctypes.c_longlong), ("vertex_ms", ctypes.c_double), ("psi_ms", ctypes.c_double),
                ("gather_ms", ctypes.c_double), ("gather_cells", ctypes.c_longlong),
                ("weno_evals", ctypes.c_longlong), ("quad_ms", ctypes.c_double), ("quad_launches", ctypes.c_longlong),
                ("pcg_ms", ctypes.c_double), ("pcg_iters", ctypes.c_longlong), ("pcg_solves", ctypes.c_longlong),
                ("twisted_faces", ctypes.c_longlong), ("general_pieces", ctypes.c_longlong),
                ("graph_captures", ctypes.c_longlong), ("graph_replays", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TuningC(ctypes.Structure):
    """fia_tuning: launch shapes of one grid shape (0 / -1 = the built-in choice)."""
    _fields_ = [("face_block", ctypes.c_int), ("piece_block", ctypes.c_int), ("cell_block", ctypes.c_int),
                ("pcg_chunk_x", ctypes.c_int), ("pcg_recompute", ctypes.c_int), ("pcg_fused", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double


def lib():
    """Loads libfia.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FiaError(6, "libfia.so is not built (run flux-interpolated-advection_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        L.fia_default_opts.argtypes = [ctypes.POINTER(AdvectOpts)]
        L.fia_default_opts.restype = None
        L.fia_ctx_create.argtypes = [_i, ctypes.POINTER(_vp)]
        L.fia_ctx_create_on_stream.argtypes = [_i, _vp, ctypes.POINTER(_vp)]
        L.fia_ctx_stream.argtypes = [_vp]
        L.fia_ctx_stream.restype = _vp
        L.fia_ctx_destroy.argtypes = [_vp]
        L.fia_ctx_device_bytes.argtypes = [_vp]
        L.fia_ctx_device_bytes.restype = ctypes.c_size_t
        L.fia_last_error.restype = ctypes.c_char_p
        L.fia_device_count.restype = _i
        L.fia_move_substances_3d.argtypes = [_vp, ctypes.POINTER(GridC), _i, _d, ctypes.POINTER(AdvectOpts)]
        L.fia_advect_flux_pass.argtypes = [_vp, ctypes.POINTER(GridC), _vp, _d, _i, _i, _i, _i, _i, _i, _d]
        L.fia_advect_velocity_3d.argtypes = [_vp, ctypes.POINTER(GridC), _i, _d, ctypes.POINTER(AdvectOpts)]
        L.fia_calc_psi_y.argtypes = [_vp, _i, _i, _i, _d, _vp, _vp, _i, _i, _i, _i]
        L.fia_interpolate.argtypes = [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _i, _vp, ctypes.c_size_t, _vp, _i]
        L.fia_backtrace_vertices.argtypes = [_vp, ctypes.POINTER(GridC), _i, _d, _i, _vp, _i]
        L.fia_split_faces.argtypes = [_vp, _vp, ctypes.c_size_t, _d, _i, _d, _i, _vp, _vp, _vp, _i]
        L.fia_cfl_time_step.argtypes = [_vp, ctypes.POINTER(GridC), _d, _i, _i, _d, ctypes.POINTER(_d),
                                        ctypes.POINTER(_i)]
        L.fia_pressure_project.argtypes = [_vp, ctypes.POINTER(GridC), _vp, _vp, _i, _d, _i, ctypes.POINTER(_i),
                                           ctypes.POINTER(_d)]
        L.fia_ctx_set_timing.argtypes = [_vp, _i]
        L.fia_ctx_set_pressure_preconditioner.argtypes = [_vp, _i]
        L.fia_ctx_set_graphs.argtypes = [_vp, _i]
        L.fia_ctx_set_tuning.argtypes = [_vp, _i, _i, _i, ctypes.POINTER(TuningC)]
        L.fia_ctx_get_tuning.argtypes = [_vp, _i, _i, _i, ctypes.POINTER(TuningC)]
        L.fia_ctx_autotune.argtypes = [_vp, ctypes.POINTER(GridC), _vp, _d, ctypes.POINTER(AdvectOpts), _i,
                                       ctypes.POINTER(TuningC)]
        L.fia_ctx_get_stats.argtypes = [_vp, ctypes.POINTER(KernelStats)]
        L.fia_ctx_reset_stats.argtypes = [_vp]
        L.fia_unet_param_count.restype = ctypes.c_size_t
        L.fia_unet_create.argtypes = [_vp, ctypes.c_ulonglong, ctypes.POINTER(_vp)]
        L.fia_unet_destroy.argtypes = [_vp]
        L.fia_unet_set_params.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.fia_unet_get_params.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.fia_unet_forward.argtypes = [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, ctypes.POINTER(_d), _i]
        L.fia_unet_loss_grad.argtypes = [_vp, _vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp,
                                         _i]
        L.fia_unet_train_step.argtypes = [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _i, ctypes.c_float, ctypes.POINTER(_d),
                                          ctypes.POINTER(_d)]
        L.fia_pressure_rhs.argtypes = [_vp, ctypes.POINTER(GridC), _vp, _vp]
        L.fia_grid_totals.argtypes = [_vp, ctypes.POINTER(GridC), ctypes.POINTER(_d * 4)]
        L.fia_wall_momentum_loss.argtypes = [_vp, _i, _i, _i, _d, _vp, _i, ctypes.POINTER(_d * 3)]
        L.fia_grid_totals_noalloc.argtypes = [_vp, ctypes.POINTER(GridC), ctypes.POINTER(_d * 4)]
        L.fia_scene_forcing.argtypes = [_vp, ctypes.POINTER(SceneC), _d]
        L.fia_scenes_step.argtypes = [ctypes.POINTER(_vp), _i, ctypes.POINTER(SceneC), _i, ctypes.POINTER(AdvectOpts),
                                      ctypes.POINTER(AdvectOpts), _d, _i, ctypes.POINTER(_vp)]
        L.fia_scenes_step_batched.argtypes = L.fia_scenes_step.argtypes
        L.fia_pressure_project_learned.argtypes = [_vp, ctypes.POINTER(GridC), _vp, _vp, _vp, _d, _i, ctypes.POINTER(_i),
                                                   ctypes.POINTER(_d), ctypes.POINTER(LearnedInfo)]
        _lib = L
    return _lib


def _check(status):
    if status != FIA_OK:
        raise FiaError(status, lib().fia_last_error().decode())


def _torch_first():
    """torch ships its own HIP runtime; when both are loaded, torch's must start first (after
    libfia.so has opened the device, torch reports no GPU).  No-op without torch."""
    try:
        import torch
    except ImportError:
        return
    if torch.version.hip is not None:
        torch.cuda.is_available()


def device_count():
    _torch_first()
    return lib().fia_device_count()


PRECOND = {"jacobi": 0, "mg": 1}


class Context:
    """Owns the device scratch and the HIP stream (fia_ctx)."""

    def __init__(self, device=0, stream=None):
        """stream: a hipStream_t handle (int) of the HIP runtime libfia.so links (ROCm's
        libamdhip64), to enqueue on the caller's stream; torch's streams belong to torch's own
        HIP runtime and are not valid here.  None: the context creates its own stream."""
        _torch_first()
        self._h = _vp()
        if stream is None:
            _check(lib().fia_ctx_create(device, ctypes.byref(self._h)))
        else:
            _check(lib().fia_ctx_create_on_stream(device, _vp(stream), ctypes.byref(self._h)))

    def stream(self):
        return lib().fia_ctx_stream(self._h)

    def set_pressure_preconditioner(self, kind):
        """"mg" (geometric multigrid V-cycle, the default) or "jacobi" for later pressure solves."""
        _check(lib().fia_ctx_set_pressure_preconditioner(self._h, PRECOND[kind]))

    def set_graphs(self, enable):
        """Replay the pressure solves' CG iteration batches from captured HIP graphs (the default)
        or launch every kernel directly; the results are bit-identical."""
        _check(lib().fia_ctx_set_graphs(self._h, 1 if enable else 0))

    def set_tuning(self, shape, **shapes):
        """Launch shapes (fia_tuning fields) for grids of shape (nx, ny, nz); shape None sets the
        default for shapes without their own entry.  Unnamed fields keep their built-in choice."""
        t = TuningC(0, 0, 0, 0, -1, -1)
        for k, v in shapes.items():
            setattr(t, k, int(v))
        nx, ny, nz = shape if shape is not None else (0, 0, 0)
        _check(lib().fia_ctx_set_tuning(self._h, nx, ny, nz, ctypes.byref(t)))

    def get_tuning(self, shape):
        t = TuningC()
        _check(lib().fia_ctx_get_tuning(self._h, shape[0], shape[1], shape[2], ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, grid, opts, dt, solid=None, reps=3):
        """fia_ctx_autotune: times the candidate launch shapes on copies of grid (left unchanged),
        stores the fastest for its shape and returns them as a dict."""
        g = grid._c(with_density=True)
        sp, keep = _solid_ptr(solid, g.on_device)
        t = TuningC()
        _check(lib().fia_ctx_autotune(self._h, ctypes.byref(g), sp, float(dt), ctypes.byref(opts), int(reps),
                                      ctypes.byref(t)))
        return t.as_dict()

    def close(self):
        if self._h:
            lib().fia_ctx_destroy(self._h)
            self._h = _vp()

    def device_bytes(self):
        return lib().fia_ctx_device_bytes(self._h)

    def set_timing(self, enable=True):
        """HIP events around every face/vertex/psi/gather launch on the context's stream."""
        _check(lib().fia_ctx_set_timing(self._h, int(bool(enable))))

    def reset_stats(self):
        _check(lib().fia_ctx_reset_stats(self._h))

    def stats(self):
        s = KernelStats()
        _check(lib().fia_ctx_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(a):
    """(pointer, on_device) of a numpy float array or a torch tensor."""
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data, 0
    # torch tensor
    if not a.is_contiguous():
        raise ValueError("tensors must be contiguous")
    return a.data_ptr(), int(a.is_cuda)


class Grid3D:
    """Grid_3D (reference include/grid_3d.h:14-88): nx, ny, nz, cell length and the MAC fields
    substance_density [nx,ny,nz], velocity u [nx+1,ny,nz], v [nx,ny+1,nz], w [nx,ny,nz+1]."""

    def __init__(self, nx, ny, nz, dx, density=None, u=None, v=None, w=None):
        self.nx, self.ny, self.nz, self.dx = int(nx), int(ny), int(nz), float(dx)
        z = lambda s: np.zeros(s, dtype=np.float32)
        self.density = density if density is not None else z((nx, ny, nz))
        self.u = u if u is not None else z((nx + 1, ny, nz))
        self.v = v if v is not None else z((nx, ny + 1, nz))
        self.w = w if w is not None else z((nx, ny, nz + 1))
        self.pressure = None

    def _c(self, with_density=True):
        d, on = _ptr(self.density) if with_density and self.density is not None else (0, None)
        pu, ou = _ptr(self.u)
        pv, ov = _ptr(self.v)
        pw, ow = _ptr(self.w)
        flags = {ou, ov, ow} | ({on} if on is not None else set())
        if len(flags) != 1:
            raise ValueError("all grid fields must live on the same side (host or device)")
        return GridC(self.nx, self.ny, self.nz, self.dx, d, pu, pv, pw, flags.pop())


def make_opts(use_flux_advection=True, use_MacCormack_scheme=False, use_clamping_in_MacCormack_scheme=True,
              num_gauss_quad_boundary=1, num_gauss_quad_bulk=1, enable_eliminate_zero_velocity_false_diffusion=True,
              split_method="y-CellFaceAligned", integral_method="gauss", num_gauss_quadrature_points=1,
              interpolation_method="WENO6", use_integral=True, num_gauss_quadrature_point_for_integrate_density=1,
              enable_cell_volume_correction=False, minimal_area=1.0):
    """Parses the reference's options once into fia_advect_opts.  Unknown split strings leave faces
    uncut, exactly as the reference does with its default "y-CellFaceAligned"."""
    if interpolation_method not in INTERP:
        raise FiaError(1, "unknown interpolation_method %r" % (interpolation_method,))
    if integral_method not in ("gauss", "analytical", "triangle"):
        raise FiaError(1, "unknown integral_method %r" % (integral_method,))
    o = AdvectOpts()
    lib().fia_default_opts(ctypes.byref(o))
    o.use_flux_advection = int(bool(use_flux_advection))
    o.use_maccormack = int(bool(use_MacCormack_scheme))
    o.use_clamping_in_maccormack = int(bool(use_clamping_in_MacCormack_scheme))
    o.num_gauss_quad_boundary = num_gauss_quad_boundary
    o.num_gauss_quad_bulk = num_gauss_quad_bulk
    o.enable_zero_velocity_correction = int(bool(enable_eliminate_zero_velocity_false_diffusion))
    o.split_method = SPLIT.get(split_method, 0)
    o.num_gauss_quadrature_points = num_gauss_quadrature_points
    o.interpolation_method = INTERP[interpolation_method]
    o.use_integral = int(bool(use_integral))
    o.num_gauss_points_integrate_density = num_gauss_quadrature_point_for_integrate_density
    o.enable_cell_volume_correction = int(bool(enable_cell_volume_correction))
    o.minimal_area = float(minimal_area)
    return o


def move_substances_3d(ctx, grid, i_frame, time_step_length, use_flux_advection=True, use_MacCormack_scheme=False,
                       use_clamping_in_MacCormack_scheme=True, num_gauss_quad_boundary=1, num_gauss_quad_bulk=1,
                       enable_eliminate_zero_velocity_false_diffusion=True, split_method="y-CellFaceAligned",
                       integral_method="gauss", num_gauss_quadrature_points=1, interpolation_method="WENO6",
                       use_integral=True, num_gauss_quadrature_point_for_integrate_density=1,
                       enable_cell_volume_correction=False, minimal_area=1.0):
    """One density step, in place (reference src/move_substances_3d.cpp:38-240)."""
    o = make_opts(use_flux_advection, use_MacCormack_scheme, use_clamping_in_MacCormack_scheme,
                  num_gauss_quad_boundary, num_gauss_quad_bulk, enable_eliminate_zero_velocity_false_diffusion,
                  split_method, integral_method, num_gauss_quadrature_points, interpolation_method, use_integral,
                  num_gauss_quadrature_point_for_integrate_density, enable_cell_volume_correction, minimal_area)
    g = grid._c()
    _check(lib().fia_move_substances_3d(ctx._h, ctypes.byref(g), int(i_frame), float(time_step_length),
                                        ctypes.byref(o)))


def advect_flux_pass(ctx, grid, q, time_step_length, split_method="xz-CellFaceAligned", interpolation_method="WENO6",
                     psi_mode="WENO6", use_zero_velocity_for_backtrace=False, use_integral=True,
                     num_gauss_quadrature_point_for_integrate_density=1, minimal_area=1.0):
    """One flux pass of a cell-centred field q, in place
    (reference src/move_substances_3d.cpp:585-841)."""
    g = grid._c(with_density=False)
    pq, on = _ptr(q)
    if on != g.on_device:
        raise ValueError("q must live where the grid lives")
    _check(lib().fia_advect_flux_pass(ctx._h, ctypes.byref(g), pq, float(time_step_length),
                                      SPLIT.get(split_method, 0), INTERP[interpolation_method], PSI[psi_mode],
                                      int(bool(use_zero_velocity_for_backtrace)), int(bool(use_integral)),
                                      num_gauss_quadrature_point_for_integrate_density, float(minimal_area)))


def advect_velocity_flux_advection_3d(ctx, grid, dim, time_step_length, **kw):
    """Self-advection of velocity component dim (0, 1, 2) or of all three (dim = -1)
    (reference src/advect_velocity_flux_advection_3d.cpp:580-1272)."""
    o = make_opts(**kw)
    g = grid._c(with_density=False)
    _check(lib().fia_advect_velocity_3d(ctx._h, ctypes.byref(g), int(dim), float(time_step_length),
                                        ctypes.byref(o)))


def advect_velocity_xyz_flux_advection_3d(ctx, grid, time_step_length, **kw):
    advect_velocity_flux_advection_3d(ctx, grid, -1, time_step_length, **kw)


def calc_psi_y(ctx, q, dx, mode="WENO6", use_integral=True, num_gauss=1):
    """psi_y of a cell-centred field (reference src/calc_psi_3D.cpp:10-308); returns float64 in
    units of q*dx, shape (nx, ny+1, nz)."""
    q = np.ascontiguousarray(q, dtype=np.float32)
    nx, ny, nz = q.shape
    psi = np.zeros((nx, ny + 1, nz), dtype=np.float64)
    _check(lib().fia_calc_psi_y(ctx._h, nx, ny, nz, float(dx), q.ctypes.data, psi.ctypes.data, PSI[mode],
                                int(bool(use_integral)), num_gauss, 0))
    return psi * dx   # the kernel stores psi / dx


def interpolate(ctx, field, offsets, dx, method, points):
    """Interpolates a staggered field (sample (i,j,k) at ((i,j,k)+offsets)*dx) at physical points."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(len(pts), dtype=np.float32)
    mx, my, mz = field.shape
    _check(lib().fia_interpolate(ctx._h, field.ctypes.data, mx, my, mz, offsets[0], offsets[1], offsets[2],
                                 float(dx), INTERP[method], pts.ctypes.data, len(pts), out.ctypes.data, 0))
    return out


def backtrace_vertices(ctx, grid, time_step_length, interpolation_method="WENO6", dim=-1):
    """Backtraced displacement (in cells) of every vertex of the density (dim=-1) or velocity
    component lattice; shape (lx, ly, lz, 3)."""
    g = grid._c(with_density=False)
    if g.on_device:
        raise ValueError("host grid expected")
    e = [int(dim == a) for a in range(3)]
    shape = (grid.nx + e[0] + 1, grid.ny + e[1] + 1, grid.nz + e[2] + 1)
    out = np.zeros((3,) + shape, dtype=np.float32)
    _check(lib().fia_backtrace_vertices(ctx._h, ctypes.byref(g), int(dim), float(time_step_length),
                                        INTERP[interpolation_method], out.ctypes.data, 0))
    return np.moveaxis(out, 0, -1)


def split_faces(ctx, quads, dx, split_method, minimal_area=1.0, max_pieces=64):
    """Cuts faces given as (n, 4, 3) physical corner positions; returns per face a list of
    (centroid[3], vector_area[3]) pieces (reference src/split_face_3D.cpp:17-93)."""
    q = np.ascontiguousarray(quads, dtype=np.float32).reshape(-1, 12)
    n = len(q)
    cnt = np.zeros(n, dtype=np.int32)
    cen = np.zeros((n, max_pieces, 3)); ar = np.zeros((n, max_pieces, 3))
    _check(lib().fia_split_faces(ctx._h, q.ctypes.data, n, float(dx), SPLIT.get(split_method, 0),
                                 float(minimal_area), max_pieces, cnt.ctypes.data, cen.ctypes.data, ar.ctypes.data,
                                 0))
    return [(cen[i, :cnt[i]].copy(), ar[i, :cnt[i]].copy()) for i in range(n)]


def calc_time_step_length_from_cfl_number_3d(ctx, grid, now_time, i_movie_frame, cfl_number, movie_fps):
    """(dt, write_frame_to_movie_file); cfl_number is truncated to int as in the reference."""
    g = grid._c(with_density=False)
    dt = _d(0); wf = _i(0)
    _check(lib().fia_cfl_time_step(ctx._h, ctypes.byref(g), float(now_time), int(i_movie_frame), int(cfl_number),
                                   float(movie_fps), ctypes.byref(dt), ctypes.byref(wf)))
    return dt.value, bool(wf.value)


def calc_pressure_gradient_term_3d(ctx, grid, solid=None, rel_tol=1e-6, max_iters=20000, warm_start=None):
    """Projection: PCG (FP64; the context's preconditioner, multigrid by default) on the reference's
    Poisson system, then u -= grad p / dx.
    Returns (iterations, relative residual); grid.pressure receives p."""
    g = grid._c(with_density=False)
    n = grid.nx * grid.ny * grid.nz
    if g.on_device:
        import torch
        if grid.pressure is None or not torch.is_tensor(grid.pressure):
            grid.pressure = torch.zeros((grid.nx, grid.ny, grid.nz), dtype=torch.float32, device=grid.u.device)
        pp = grid.pressure.data_ptr()
        sp = solid.data_ptr() if solid is not None else None
    else:
        if grid.pressure is None or not isinstance(grid.pressure, np.ndarray):
            grid.pressure = np.zeros((grid.nx, grid.ny, grid.nz), dtype=np.float32)
        pp = grid.pressure.ctypes.data
        sp = None
        if solid is not None:
            solid = np.ascontiguousarray(solid, dtype=np.uint8)
            sp = solid.ctypes.data
    if warm_start is not None:
        if g.on_device:
            grid.pressure.copy_(warm_start)
        else:
            grid.pressure[...] = warm_start
    it = _i(0); res = _d(0)
    st = lib().fia_pressure_project(ctx._h, ctypes.byref(g), sp, pp, int(warm_start is not None), float(rel_tol),
                                    int(max_iters), ctypes.byref(it), ctypes.byref(res))
    _check(st)
    return it.value, res.value


# ---- learned warm start of the pressure solve (no counterpart in the reference) --------------

def _solid_ptr(solid, on_device):
    if solid is None:
        return None, None
    if on_device:
        return solid.data_ptr(), solid
    keep = np.ascontiguousarray(solid, dtype=np.uint8)
    return keep.ctypes.data, keep


class UNet:
    """The learned pressure warm start (fia_unet): a 4-level 3-D U-Net, 16/32/48/64 channels,
    BF16 weights on the matrix cores.  Parameters are one FP32 vector of param_count() values."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        self._h = _vp()
        _check(lib().fia_unet_create(ctx._h, int(seed), ctypes.byref(self._h)))

    @staticmethod
    def param_count():
        return int(lib().fia_unet_param_count())

    def close(self):
        if self._h:
            lib().fia_unet_destroy(self._h)
            self._h = _vp()

    def get_params(self):
        p = np.empty(self.param_count(), dtype=np.float32)
        _check(lib().fia_unet_get_params(self.ctx._h, self._h, p.ctypes.data, p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().fia_unet_set_params(self.ctx._h, self._h, p.ctypes.data, p.size))

    def save(self, path):
        np.save(path, self.get_params(), allow_pickle=False)

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def forward(self, rhs, solid=None):
        """Network output y (FP32) and its scale for right-hand side rhs [nx,ny,nz] (FP64):
        the pressure guess is scale * y (up to a constant)."""
        nx, ny, nz = rhs.shape
        pr, on = _ptr(rhs)
        sp, keep = _solid_ptr(solid, on)
        if on:
            import torch
            y = torch.empty((nx, ny, nz), dtype=torch.float32, device=rhs.device)
        else:
            y = np.empty((nx, ny, nz), dtype=np.float32)
        sc = _d(0)
        _check(lib().fia_unet_forward(self.ctx._h, self._h, nx, ny, nz, pr, sp, _ptr(y)[0], ctypes.byref(sc), on))
        return y, sc.value

    def loss_grad(self, rhs, solid=None, with_grad=True, loss="residual"):
        """(loss, residual norm |L y - b|^2/|b|^2, FP32 parameter gradient or None) of one sample;
        loss is "residual" (residual norm) or "energy" (y.L y / 2 - b.y, over |b|^2)."""
        nx, ny, nz = rhs.shape
        pr, on = _ptr(rhs)
        sp, keep = _solid_ptr(solid, on)
        g = np.empty(self.param_count(), dtype=np.float32) if with_grad else None
        lv, rv = _d(0), _d(0)
        _check(lib().fia_unet_loss_grad(self.ctx._h, self._h, nx, ny, nz, pr, sp, UNET_LOSS[loss], ctypes.byref(lv),
                                        ctypes.byref(rv), g.ctypes.data if g is not None else None, on))
        return lv.value, rv.value, g

    def train_step(self, rhs_batch, solid_batch=None, lr=1e-3, loss="residual"):
        """One Adam step on a batch [b, nx, ny, nz] of device tensors; returns the batch means of
        the loss and of the residual norm before the step."""
        b, nx, ny, nz = rhs_batch.shape
        pr, on = _ptr(rhs_batch)
        if not on:
            raise ValueError("training batches must be device tensors")
        sp = solid_batch.data_ptr() if solid_batch is not None else None
        lv, rv = _d(0), _d(0)
        _check(lib().fia_unet_train_step(self.ctx._h, self._h, nx, ny, nz, b, pr, sp, UNET_LOSS[loss], float(lr),
                                         ctypes.byref(lv), ctypes.byref(rv)))
        return lv.value, rv.value


def pressure_rhs(ctx, grid, solid=None):
    """Right-hand side b = -div(u) dx of the pressure system (as the projection assembles it)."""
    g = grid._c(with_density=False)
    sp, keep = _solid_ptr(solid, g.on_device)
    if g.on_device:
        import torch
        rhs = torch.empty((grid.nx, grid.ny, grid.nz), dtype=torch.float64, device=grid.u.device)
    else:
        rhs = np.empty((grid.nx, grid.ny, grid.nz), dtype=np.float64)
    _check(lib().fia_pressure_rhs(ctx._h, ctypes.byref(g), sp, _ptr(rhs)[0]))
    return rhs


def calc_pressure_gradient_term_3d_learned(ctx, grid, net, solid=None, rel_tol=1e-6, max_iters=20000):
    """calc_pressure_gradient_term_3d with the U-Net warm start (and its fallback to a zero
    guess).  Returns (iterations, relative residual, info dict)."""
    g = grid._c(with_density=False)
    n = grid.nx * grid.ny * grid.nz
    sp, keep = _solid_ptr(solid, g.on_device)
    if g.on_device:
        import torch
        if grid.pressure is None or not torch.is_tensor(grid.pressure):
            grid.pressure = torch.zeros((grid.nx, grid.ny, grid.nz), dtype=torch.float32, device=grid.u.device)
        pp = grid.pressure.data_ptr()
    else:
        if grid.pressure is None or not isinstance(grid.pressure, np.ndarray):
            grid.pressure = np.zeros((grid.nx, grid.ny, grid.nz), dtype=np.float32)
        pp = grid.pressure.ctypes.data
    it = _i(0); res = _d(0); info = LearnedInfo()
    _check(lib().fia_pressure_project_learned(ctx._h, ctypes.byref(g), sp, pp, net._h, float(rel_tol), int(max_iters),
                                              ctypes.byref(it), ctypes.byref(res), ctypes.byref(info)))
    return it.value, res.value, info.as_dict()


# ---- diagnostics of the reference's output series --------------------------------------------

def grid_totals(ctx, grid):
    """(total mass, momentum x, y, z): sums times the cell volume (reference src/main.cpp:518-597)."""
    g = grid._c(with_density=grid.density is not None)
    out = (_d * 4)()
    _check(lib().fia_grid_totals(ctx._h, ctypes.byref(g), ctypes.byref(out)))
    return tuple(out)


def grid_totals_noalloc(ctx, grid):
    """grid_totals of a device grid without any allocation (fia_grid_totals_noalloc): the same
    sums in a fixed device scratch; calls overlapping on one GPU race."""
    g = grid._c(with_density=grid.density is not None)
    out = (_d * 4)()
    _check(lib().fia_grid_totals_noalloc(ctx._h, ctypes.byref(g), ctypes.byref(out)))
    return tuple(out)


def wall_momentum_loss(ctx, grid, pressure=None):
    """Momentum the walls took out through the pressure of one projection (x, y, z), reference
    src/calc_pressure_3d.cpp:235-266; pressure defaults to grid.pressure."""
    p = grid.pressure if pressure is None else pressure
    pp, on = _ptr(p)
    out = (_d * 3)()
    _check(lib().fia_wall_momentum_loss(ctx._h, grid.nx, grid.ny, grid.nz, grid.dx, pp, on, ctypes.byref(out)))
    return tuple(out)


# ---- many independent scenes (no counterpart in the reference) -------------------------------

def scenes_step(ctxs, scenes, density_opts, velocity_opts, rel_tol=1e-6, max_iters=20000, nets=None,
                batched=False):
    """One step of every scene (a ctypes array of SceneC) on len(ctxs) workers; batched: the
    pressure solves of all scenes run together on ctxs[0] (fia_scenes_step_batched)."""
    w = len(ctxs)
    ca = (_vp * w)(*[c._h for c in ctxs])
    na = (_vp * w)(*[n._h for n in nets]) if nets is not None else None
    fn = lib().fia_scenes_step_batched if batched else lib().fia_scenes_step
    _check(fn(ca, w, scenes, len(scenes), ctypes.byref(density_opts), ctypes.byref(velocity_opts), float(rel_tol),
              int(max_iters), na))
