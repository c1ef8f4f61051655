"""Host replay of the library's flux passes (tests/pass_host.cpp: the kernels' __host__ __device__
bodies in the order fia_api.hip launches them) against the FP64 oracle, on the CPU.

The replay context keeps the GPU context's buffer rules (grow-only, contents not kept, nothing
cleared between calls) and poisons new memory, so these checks also cover reads of entries no
kernel of a call wrote and scratch reused across calls of different grid shapes.  Every call
checks its face lists and slot counts.  Cases:
- non-cubic grids (the reference CLI takes grid_num_x/y/z separately, src/main.cpp:55-57), for
  the density step and for velocity self-advection;
- the 1Dy_linear / const psi / xz-AxisAligned flux pass at CFL 1, whose colours sit on integers at
  the walls (degenerate faces);
- a velocity step followed by a density step through one context, against the oracle's same
  sequence.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_linf

TOL = 1e-5
SPLIT = {"none": 0, "xz-CellFaceAligned": 1, "xz-AxisAligned": 2, "x-CellFaceAligned_z-AxisAligned": 3,
         "x-AxisAligned_z-CellFaceAligned": 4}
INTERP = {"linear": 0, "WENO6": 1, "1Dy_linear": 2}
PSI = {"const": 0, "linear": 1, "WENO6": 2}


class HostReplay:
    def __init__(self, path):
        L = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        fp, dp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.host_ctx_create.restype = vp
        L.host_ctx_destroy.argtypes = [vp]
        L.host_ctx_message.argtypes = [vp]
        L.host_ctx_message.restype = ctypes.c_char_p
        L.host_stat_cols.restype = i
        L.host_set_threads.argtypes = [i]
        L.host_move_substances.argtypes = [vp, i, i, i, d, fp, fp, fp, fp, d, i, i, i, i, i, d, dp, i, ip]
        L.host_advect_velocity.argtypes = [vp, i, i, i, d, fp, fp, fp, i, d, i, i, i, i, d, dp, i, ip]
        L.host_flux_pass.argtypes = [vp, i, i, i, i, d, fp, fp, fp, fp, d, i, i, i, i, i, i, d, dp, i, ip]
        self.L = L
        self.cols = L.host_stat_cols()

    def context(self):
        return self.L.host_ctx_create()

    def close(self, c):
        self.L.host_ctx_destroy(c)

    def _stats(self):
        st = np.zeros((64, self.cols)); rows = ctypes.c_int(0)
        return st, rows

    def _check(self, c, e):
        if e != 0:
            raise RuntimeError("host replay error %d: %s" % (e, self.L.host_ctx_message(c).decode()))

    def move_substances(self, c, rho, u, v, w, h, dt, split="xz-CellFaceAligned", interp="WENO6", integral=True,
                        corr=True, ng=1, min_area=1.0):
        nx, ny, nz = rho.shape
        st, rows = self._stats()
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        e = self.L.host_move_substances(c, nx, ny, nz, h, f(rho), f(u), f(v), f(w), dt, SPLIT[split], INTERP[interp],
                                        int(integral), int(corr), ng, min_area,
                                        st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 64, ctypes.byref(rows))
        self._check(c, e)
        return st[:rows.value]

    def advect_velocity(self, c, u, v, w, h, dt, dim=-1, interp="WENO6", integral=True, corr=True, ng=1):
        nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
        st, rows = self._stats()
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        e = self.L.host_advect_velocity(c, nx, ny, nz, h, f(u), f(v), f(w), dim, dt, INTERP[interp], int(integral),
                                        int(corr), ng, 1.0, st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 64,
                                        ctypes.byref(rows))
        self._check(c, e)
        return st[:rows.value]

    def flux_pass(self, c, q, u, v, w, h, dt, split, interp, psi, zero, integral, ng, comp=-1):
        nx, ny, nz = u.shape[0] - 1, u.shape[1], u.shape[2]
        st, rows = self._stats()
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        e = self.L.host_flux_pass(c, comp, nx, ny, nz, h, f(q), f(u), f(v), f(w), dt, SPLIT[split], INTERP[interp], PSI[psi],
                                  int(zero), int(integral), ng, 1.0, st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  64, ctypes.byref(rows))
        self._check(c, e)
        return st[:rows.value]


def build_replay(out, sanitize=False):
    # the kernels' translation unit is compiled whole (its device code too, which the replay never
    # runs) so that the shared object carries the code object its kernel registrations refer to
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "pass_host.cpp"), "-o", out]
    if sanitize:
        cmd[5:5] = ["-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not found")
    return HostReplay(build_replay(str(tmp_path_factory.mktemp("replay") / "pass_host.so")))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def white_noise_scene(nx, ny, nz, seed, cfl, h):
    rng = np.random.default_rng(seed)
    rho = _f32(rng.uniform(0, 1, size=(nx, ny, nz)))
    u = rng.normal(size=(nx + 1, ny, nz)); u[0] = u[-1] = 0
    v = rng.normal(size=(nx, ny + 1, nz)); v[:, 0] = v[:, -1] = 0
    w = rng.normal(size=(nx, ny, nz + 1)); w[:, :, 0] = w[:, :, -1] = 0
    u, v, w = _f32(u), _f32(v), _f32(w)
    dt = cfl * h / max(abs(u).max(), abs(v).max(), abs(w).max())
    return rho, u, v, w, dt


@pytest.mark.parametrize("shape,seed,cfl", [((12, 20, 9), 3, 1.5), ((10, 16, 24), 5, 2.5), ((17, 8, 13), 7, 1.0)])
def test_replay_move_substances_nonrect_matches_oracle(replay, oracle, shape, seed, cfl):
    nx, ny, nz = shape
    h = 1.0 / 16
    rho, u, v, w, dt = white_noise_scene(nx, ny, nz, seed, cfl, h)
    c = replay.context()
    try:
        q = rho.copy()
        st = replay.move_substances(c, q, u, v, w, h, dt, split="xz-CellFaceAligned")
    finally:
        replay.close(c)
    ref = oracle.move_substances(rho.astype(np.float64), u, v, w, h, dt, split="xz-CellFaceAligned")
    assert np.isfinite(q).all()
    assert rel_linf(q, ref) < TOL
    if cfl > 1.0:
        assert st[0, 0] > 0   # the white-noise field twists faces: the general cutter is exercised


@pytest.mark.parametrize("shape,seed,cfl", [((4, 3, 5), 3, 1.5), ((2, 7, 3), 5, 2.0), ((6, 6, 1), 7, 1.0),
                                             ((1, 5, 6), 9, 1.0), ((3, 3, 3), 4, 3.0)])
def test_replay_tiny_grids_match_oracle(replay, oracle, shape, seed, cfl):
    # grids narrower than the six-sample WENO6 stencil (every stencil clamped) and one cell thick
    nx, ny, nz = shape
    h = 1.0 / 16
    rho, u, v, w, dt = white_noise_scene(nx, ny, nz, seed, cfl, h)
    c = replay.context()
    try:
        q = rho.copy()
        replay.move_substances(c, q, u, v, w, h, dt, split="xz-CellFaceAligned")
    finally:
        replay.close(c)
    ref = oracle.move_substances(rho.astype(np.float64), u, v, w, h, dt, split="xz-CellFaceAligned")
    assert np.isfinite(q).all()
    assert rel_linf(q, ref) < TOL


def test_replay_pass_total_is_its_wall_and_correction_terms(replay):
    # A cell's new value is the sum over its six faces of (upper flux - lower flux) plus the lower
    # face's correction.  Summed over the grid every interior face cancels, so the total is the
    # fluxes through the domain walls plus the corrections of the general cutter's (twisted) faces,
    # which the reference also evaluates in both corner orders: nothing else may create or lose mass.
    nx, ny, nz = 12, 20, 9
    h = 1.0 / 16
    rho, u, v, w, dt = white_noise_scene(nx, ny, nz, 3, 2.0, h)
    c = replay.context()
    try:
        q = rho.copy()
        st = replay.flux_pass(c, q, u, v, w, h, dt, "xz-CellFaceAligned", "WENO6", "WENO6", False, True, 1)
        L = replay.L
        L.host_ctx_fluxes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_long]
        shapes = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
        flux, corr = [], []
        for fam, sh in enumerate(shapes):
            f = np.zeros(sh); k = np.zeros(sh)
            dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            assert L.host_ctx_fluxes(c, fam, dp(f), dp(k), int(f.size)) == 0
            flux.append(f); corr.append(k)
    finally:
        replay.close(c)
    assert st[0, 0] > 0   # twisted faces, so corrections take part
    walls = ((flux[0][-1] - flux[0][0]).sum() + (flux[1][:, -1] - flux[1][:, 0]).sum() +
             (flux[2][:, :, -1] - flux[2][:, :, 0]).sum())
    # corrections enter through the lower face of each cell
    corrections = corr[0][:-1].sum() + corr[1][:, :-1].sum() + corr[2][:, :, :-1].sum()
    assert np.count_nonzero(corr[0]) + np.count_nonzero(corr[1]) + np.count_nonzero(corr[2]) > 0
    total = q.astype(np.float64).sum()
    # to the FP32 rounding of the cells (without the corrections the sum is off by 1e-4 of it)
    assert abs(total - (walls + corrections)) <= 1e-8 * np.abs(q).sum()


def smooth_scene(scenes, shape, seed, cfl, h):
    """Curl noise on a non-cubic box: smooth, divergence-free, no flow through the walls."""
    u, v, w = [_f32(a) for a in scenes.curl_noise(shape, seed=seed)]
    dt = cfl * h / max(abs(u).max(), abs(v).max(), abs(w).max())
    return u, v, w, dt


def rel_linf_either(got, refs):
    """Largest cell difference from the closer of several references (relative to their scale)."""
    got = np.asarray(got, dtype=np.float64)
    err = np.min([np.abs(got - r) for r in refs], axis=0)
    return float(err.max() / max(np.abs(refs[0]).max(), 1e-300))


@pytest.mark.parametrize("shape", [(12, 20, 9), (9, 11, 14)])
@pytest.mark.parametrize("interp", ["WENO6", "linear"])
def test_replay_velocity_nonrect_matches_oracle(replay, oracle, scenes, shape, interp):
    # The wall rows of pinned vertices put face corners exactly on cut lines (axis-aligned colours
    # are integers there); the pieces of such a face depend on rounding in the reference's own
    # FP32 cutter (src/facecutter.h), which the oracle reproduces bit for bit with cprec=1, and its
    # FP64 form (cprec=0) can settle them differently.  Every cell must agree with one of the two.
    h = 1.0 / 16
    u, v, w, dt = smooth_scene(scenes, shape, 11, 2.0, h)
    c = replay.context()
    try:
        gu, gv, gw = u.copy(), v.copy(), w.copy()
        replay.advect_velocity(c, gu, gv, gw, h, dt, interp=interp)
    finally:
        replay.close(c)
    refs = [oracle.advect_velocity(u, v, w, h, dt, method=interp, cprec=cp) for cp in (0, 1)]
    for a, got in enumerate((gu, gv, gw)):
        assert np.isfinite(got).all()
        assert rel_linf_either(got, [r[a] for r in refs]) < TOL


def test_replay_axis_aligned_1dy_pass_at_cfl1(replay, oracle, scenes):
    # the flux pass with 1Dy_linear values, const psi and axis-aligned colours at CFL 1 on the
    # 16^3 curl-noise scene of tests/test_gpu_pass.py: pinned wall vertices put corner colours
    # exactly on integers, so the degenerate-face path runs
    n, h = 16, 1.0 / 16
    rho = _f32(scenes.gaussian_blob(n, centre=(0.5, 0.45, 0.5), sigma=0.12))
    u, v, w = [_f32(a) for a in scenes.curl_noise(n, seed=1)]
    dt = scenes.cfl_dt(u, v, w, h, 1.0)
    c = replay.context()
    try:
        q = rho.copy()
        st = replay.flux_pass(c, q, u, v, w, h, dt, "xz-AxisAligned", "1Dy_linear", "const", False, False, 1)
    finally:
        replay.close(c)
    ref = oracle.advect_pass(rho.astype(np.float64), u, v, w, h, dt, split="xz-AxisAligned", method="1Dy_linear",
                             psi_mode="const", zero_velocity=False, use_integral=False, num_gauss=1)
    assert rel_linf(q, ref) < TOL
    assert st[0, 1] > 0   # degenerate faces were cut by the clipping pipeline


def test_replay_velocity_then_density_in_one_context(replay, oracle, scenes):
    # the step sequence of a simulation loop through one context: velocity self-advection (three
    # component grids), then the density step on the cell grid with the advected velocity; the
    # context's scratch was sized by the larger component grids and keeps their contents
    n, h = 20, 1.0 / 32
    rho = _f32(scenes.gaussian_blob(n, centre=(0.5, 0.45, 0.5), sigma=0.12) + scenes.sphere_source(n))
    u, v, w = [_f32(a) for a in scenes.curl_noise(n, seed=2)]
    dt = scenes.cfl_dt(u, v, w, h, 4.0)
    c = replay.context()
    try:
        gu, gv, gw, q = u.copy(), v.copy(), w.copy(), rho.copy()
        replay.advect_velocity(c, gu, gv, gw, h, dt)
        replay.move_substances(c, q, gu, gv, gw, h, dt)
        # and once more, from the same inputs, through the same (now warm) context: identical bits
        gu2, gv2, gw2, q2 = u.copy(), v.copy(), w.copy(), rho.copy()
        replay.advect_velocity(c, gu2, gv2, gw2, h, dt)
        replay.move_substances(c, q2, gu2, gv2, gw2, h, dt)
    finally:
        replay.close(c)
    ru, rv, rw = oracle.advect_velocity(u, v, w, h, dt)
    ref = oracle.move_substances(rho.astype(np.float64), gu, gv, gw, h, dt)
    for got, want in ((gu, ru), (gv, rv), (gw, rw)):
        assert rel_linf(got, want) < TOL
    assert rel_linf(q, ref) < TOL
    assert np.array_equal(q, q2) and np.array_equal(gu, gu2)
