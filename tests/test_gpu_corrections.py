"""The corrections of the general cutter's faces (twisted faces, evaluated in both corner orders)
are kept sparse: per family a mask bit marks the faces whose FP64 correction the gather adds, and
the correction arrays are zero elsewhere between passes, the next pass on a flux set clearing what
the last one set.  A context whose passes had twisted faces must give, on another grid shape and
then back on the first, and after a velocity self-advection on the staggered component grids,
exactly what fresh contexts give; the twisted faces are counted, so the check is not vacuous.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(scenes, n, cfl, seed):
    h = 1.0 / 32
    rho = scenes.gaussian_blob(n, centre=(0.5, 0.45, 0.5), sigma=0.12)
    u, v, w = scenes.curl_noise(n, seed=seed)
    dt = scenes.cfl_dt(u, v, w, h, cfl)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return h, f32(rho), f32(u), f32(v), f32(w), dt


def _density_step(fia, c, sc):
    h, rho, u, v, w, dt = sc
    n = rho.shape[0]
    grid = fia.Grid3D(n, n, n, h, rho.copy(), u, v, w)
    fia.move_substances_3d(c, grid, 0, dt, split_method="xz-CellFaceAligned", interpolation_method="WENO6")
    return grid.density.copy()


def _velocity_step(fia, c, sc):
    h, rho, u, v, w, dt = sc
    n = rho.shape[0]
    grid = fia.Grid3D(n, n, n, h, None, u.copy(), v.copy(), w.copy())
    fia.advect_velocity_flux_advection_3d(c, grid, -1, dt, interpolation_method="WENO6")
    return grid.u.copy(), grid.v.copy(), grid.w.copy()


def test_corrections_do_not_leak_between_passes(fia, scenes):
    print("case sparse corrections: 24^3 and 20^3 at CFL 4 in one context", flush=True)
    a = _scene(scenes, 24, 4.0, seed=7)
    b = _scene(scenes, 20, 4.0, seed=9)
    c = fia.Context(0)
    try:
        c.set_timing(True)
        c.reset_stats()
        ra1 = _density_step(fia, c, a)
        twisted_a = c.stats()["twisted_faces"]
        rb1 = _density_step(fia, c, b)
        rv1 = _velocity_step(fia, c, b)
        ra2 = _density_step(fia, c, a)
        twisted = c.stats()["twisted_faces"]
    finally:
        c.close()
    print("twisted faces: first step %d, all steps %d" % (twisted_a, twisted), flush=True)
    assert twisted_a > 0
    refs = []
    for fn, sc in ((_density_step, b), (_velocity_step, b)):
        f = fia.Context(0)
        try:
            refs.append(fn(fia, f, sc))
        finally:
            f.close()
    assert np.array_equal(ra2, ra1)
    assert np.array_equal(rb1, refs[0])
    for x, y in zip(rv1, refs[1]):
        assert np.array_equal(x, y)
