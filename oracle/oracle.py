"""ctypes wrapper over the CPU oracle (oracle/_build/libfia_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg.  The product (libfia.so and its Python host module) never uses it.

All arrays are numpy float64 in the reference's [x][y][z] layout; `prec` selects the oracle's
arithmetic (0 = FP64, 1 = FP32 like the reference's MY_FLOAT_TYPE) and `cprec` the face-cutter
arithmetic (0 = FP64, 1 = FP32 as in reference src/facecutter.h).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfia_oracle.so")
REF_CUTTER_PATH = os.path.join(HERE, "_ref", "libref_facecutter.so")

INTERP = {"linear": 0, "WENO6": 1, "1Dy_linear": 2, "1Dy_WENO6": 3, "WENO6-optimized": 4}
PSI = {"const": 0, "linear": 1, "WENO6": 2}
SPLIT = {"none": 0, "xz-CellFaceAligned": 1, "xz-AxisAligned": 2,
         "x-CellFaceAligned_z-AxisAligned": 3, "x-AxisAligned_z-CellFaceAligned": 4}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_lib = None


def build():
    """Compiles the oracle (and the reference cutter when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _p(a):
    return a.ctypes.data_as(_dp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_weno1d.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double]
        L.orc_weno1d.restype = ctypes.c_double
        L.orc_weno1d_opt.argtypes = [_dp, _dp, ctypes.c_double]
        L.orc_weno1d_opt.restype = ctypes.c_double
        L.orc_interpolate.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, _dp, ctypes.c_long, _dp]
        L.orc_calc_psi.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_backtrace_vertices.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_cut_face_xieta.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp]
        L.orc_split_face.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.orc_advect_pass.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_move_substances.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_int, _dp]
        L.orc_advect_velocity.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, _dp, _dp, _dp, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, _dp]
        L.orc_face_fluxes.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_cfl_time_step.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, _dp, _dp, _dp, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip]
        L.orc_cfl_time_step.restype = ctypes.c_double
        L.orc_pressure_project.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, _dp, _dp, _dp,
                                           ctypes.POINTER(ctypes.c_uint8), _dp, ctypes.c_double, ctypes.c_int,
                                           _ip, _dp]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gauss_legendre(n):
    x = np.zeros(n); w = np.zeros(n)
    if lib().orc_gauss_legendre(n, _p(x), _p(w)):
        raise ValueError("unsupported Gauss rule %d" % n)
    return x, w


def weno1d(f, d, h, prec=0, opt=False):
    """1D WENO6 of six samples at signed distances d; opt: the "WENO6-optimized" weights
    C / (1e-30 + S^2) (FP64 only)."""
    if opt:
        return lib().orc_weno1d_opt(_p(_f64(f)), _p(_f64(d)), h)
    return lib().orc_weno1d(prec, _p(_f64(f)), _p(_f64(d)), h)


def interpolate(field, offsets, h, method, points, prec=0):
    field = _f64(field)
    pts = _f64(points).reshape(-1, 3)
    out = np.zeros(len(pts))
    mx, my, mz = field.shape
    lib().orc_interpolate(prec, _p(field), mx, my, mz, offsets[0], offsets[1], offsets[2], h,
                          INTERP[method], _p(pts), len(pts), _p(out))
    return out


def calc_psi(q, h, mode="WENO6", use_integral=True, num_gauss=1, comp=-1, n=None, prec=0):
    """psi on the y-faces of the component grid; q has the component grid's shape."""
    q = _f64(q)
    mx, my, mz = q.shape
    if n is None:
        e = [int(comp == a) for a in range(3)]
        n = (mx - e[0], my - e[1], mz - e[2])
    psi = np.zeros((mx, my + 1, mz))
    rc = lib().orc_calc_psi(prec, comp, n[0], n[1], n[2], h, _p(q), PSI[mode], int(use_integral), num_gauss,
                            _p(psi))
    if rc:
        raise ValueError("orc_calc_psi failed %d" % rc)
    return psi


def backtrace_vertices(u, v, w, h, dt, method="WENO6", comp=-1, prec=0):
    u, v, w = _f64(u), _f64(v), _f64(w)
    nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
    e = [int(comp == a) for a in range(3)]
    out = np.zeros((nx + e[0] + 1, ny + e[1] + 1, nz + e[2] + 1, 3))
    lib().orc_backtrace_vertices(prec, comp, nx, ny, nz, h, _p(u), _p(v), _p(w), dt, INTERP[method], _p(out))
    return out


def cut_face_xieta(colours, min_area=1.0, cprec=1, max_pieces=256):
    col = _f64(colours).reshape(8)
    nv = np.zeros(max_pieces, dtype=np.int32)
    verts = np.zeros((max_pieces, 32, 2)); cen = np.zeros((max_pieces, 2)); area = np.zeros(max_pieces)
    n = lib().orc_cut_face_xieta(cprec, _p(col), min_area, max_pieces, nv.ctypes.data_as(_ip), _p(verts),
                                 _p(cen), _p(area))
    if n < 0:
        raise OverflowError("cutter capacity exceeded")
    return [(verts[i, :nv[i]].copy(), cen[i].copy(), area[i]) for i in range(n)]


def split_face(quad, h, split, min_area=1.0, prec=0, cprec=0, max_pieces=256):
    q = _f64(quad).reshape(12)
    cen = np.zeros((max_pieces, 3)); an = np.zeros((max_pieces, 3))
    n = lib().orc_split_face(prec, cprec, _p(q), h, SPLIT[split], min_area, max_pieces, _p(cen), _p(an))
    if n < 0:
        raise OverflowError("cutter capacity exceeded")
    return cen[:n].copy(), an[:n].copy()


def advect_pass(q, u, v, w, h, dt, split="xz-CellFaceAligned", method="WENO6", psi_mode="WENO6",
                zero_velocity=False, use_integral=True, num_gauss=1, min_area=1.0, comp=-1, prec=0, cprec=0,
                nthreads=8, stats=None):
    q = _f64(q).copy()
    u, v, w = _f64(u), _f64(v), _f64(w)
    nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
    st = np.zeros(4)
    rc = lib().orc_advect_pass(prec, cprec, comp, nx, ny, nz, h, _p(u), _p(v), _p(w), _p(q), dt, SPLIT[split],
                               INTERP[method], PSI[psi_mode], int(zero_velocity), int(use_integral), num_gauss,
                               min_area, nthreads, _p(st))
    if rc:
        raise RuntimeError("orc_advect_pass failed %d" % rc)
    if stats is not None:
        stats["pieces"] = stats.get("pieces", 0) + st[0]
    return q


def face_fluxes(q, u, v, w, h, dt, split="xz-CellFaceAligned", method="WENO6", psi_mode="WENO6",
                zero_velocity=False, use_integral=True, num_gauss=1, min_area=1.0, comp=-1, prec=0, cprec=0):
    """Flux through every unique face (x+, y+, z+ orientation of the cell below), in units of q."""
    q = _f64(q)
    u, v, w = _f64(u), _f64(v), _f64(w)
    nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
    mx, my, mz = q.shape
    fx = np.zeros((mx + 1, my, mz)); fy = np.zeros((mx, my + 1, mz)); fz = np.zeros((mx, my, mz + 1))
    rc = lib().orc_face_fluxes(prec, cprec, comp, nx, ny, nz, h, _p(u), _p(v), _p(w), _p(q), dt, SPLIT[split],
                               INTERP[method], PSI[psi_mode], int(zero_velocity), int(use_integral), num_gauss,
                               min_area, _p(fx), _p(fy), _p(fz))
    if rc:
        raise RuntimeError("orc_face_fluxes failed %d" % rc)
    return fx, fy, fz


def gather_fluxes(fx, fy, fz):
    """Cell values from unique-face fluxes: + face minus - face along each axis."""
    return (fx[1:] - fx[:-1]) + (fy[:, 1:] - fy[:, :-1]) + (fz[:, :, 1:] - fz[:, :, :-1])


def move_substances(rho, u, v, w, h, dt, split="xz-CellFaceAligned", method="WENO6", use_integral=True,
                    correction=True, num_gauss=1, min_area=1.0, prec=0, cprec=0, nthreads=8, stats=None):
    rho = _f64(rho).copy()
    u, v, w = _f64(u), _f64(v), _f64(w)
    nx, ny, nz = rho.shape
    st = np.zeros(4)
    rc = lib().orc_move_substances(prec, cprec, nx, ny, nz, h, _p(rho), _p(u), _p(v), _p(w), dt, SPLIT[split],
                                   INTERP[method], int(use_integral), int(correction), num_gauss, min_area,
                                   nthreads, _p(st))
    if rc:
        raise RuntimeError("orc_move_substances failed %d" % rc)
    if stats is not None:
        stats["pieces"] = stats.get("pieces", 0) + st[0]
    return rho


def advect_velocity(u, v, w, h, dt, dim=-1, method="WENO6", use_integral=True, correction=True, num_gauss=1,
                    min_area=1.0, prec=0, cprec=0, nthreads=8):
    u, v, w = _f64(u).copy(), _f64(v).copy(), _f64(w).copy()
    nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
    st = np.zeros(4)
    rc = lib().orc_advect_velocity(prec, cprec, dim, nx, ny, nz, h, _p(u), _p(v), _p(w), dt, INTERP[method],
                                   int(use_integral), int(correction), num_gauss, min_area, nthreads, _p(st))
    if rc:
        raise RuntimeError("orc_advect_velocity failed %d" % rc)
    return u, v, w


def cfl_time_step(u, v, w, h, now_time, i_movie_frame, cfl, fps):
    u, v, w = _f64(u), _f64(v), _f64(w)
    nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
    wf = ctypes.c_int(0)
    dt = lib().orc_cfl_time_step(nx, ny, nz, h, _p(u), _p(v), _p(w), now_time, i_movie_frame, int(cfl), fps,
                                 ctypes.byref(wf))
    return dt, bool(wf.value)


def pressure_project(u, v, w, h, solid=None, tol=1e-8, max_iters=10000, p0=None):
    u, v, w = _f64(u).copy(), _f64(v).copy(), _f64(w).copy()
    nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
    p = np.zeros((nx, ny, nz)) if p0 is None else _f64(p0).copy()
    sp = None
    if solid is not None:
        solid = np.ascontiguousarray(solid, dtype=np.uint8)
        sp = solid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    it = ctypes.c_int(0); res = ctypes.c_double(0)
    rc = lib().orc_pressure_project(nx, ny, nz, h, _p(u), _p(v), _p(w), sp, _p(p), tol, max_iters,
                                    ctypes.byref(it), ctypes.byref(res))
    return u, v, w, p, it.value, res.value, rc


def ref_cutter_available():
    return os.path.exists(REF_CUTTER_PATH)


def ref_cut_face(colours, min_area=1.0, max_pieces=256):
    """The reference's own cutter (oracle/_ref, built from reference src/facecutter.h)."""
    L = ctypes.CDLL(REF_CUTTER_PATH)
    L.ref_cut_face.argtypes = [_fp, ctypes.c_float, ctypes.c_int, _ip, _fp, _fp, _fp]
    col = np.ascontiguousarray(colours, dtype=np.float32).reshape(8)
    nv = np.zeros(max_pieces, dtype=np.int32)
    verts = np.zeros((max_pieces, 32, 2), dtype=np.float32)
    cen = np.zeros((max_pieces, 2), dtype=np.float32); area = np.zeros(max_pieces, dtype=np.float32)
    n = L.ref_cut_face(col.ctypes.data_as(_fp), np.float32(min_area), max_pieces, nv.ctypes.data_as(_ip),
                       verts.ctypes.data_as(_fp), cen.ctypes.data_as(_fp), area.ctypes.data_as(_fp))
    if n < 0:
        raise OverflowError("cutter capacity exceeded")
    return [(verts[i, :nv[i]].astype(np.float64), cen[i].astype(np.float64), float(area[i])) for i in range(n)]
