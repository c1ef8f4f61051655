"""Seeded synthetic scenes on the reference's MAC layout ([x][y][z], y up, unit cube).

The reference ships only square/circle/Taylor-Green initialisers (src/initialize_grid.cpp:272-640);
these generators build the workloads used by the tests and the benchmark:
  gaussian_blob      density exp(-|x-c|^2 / 2 sigma^2)
  curl_noise         u = discrete curl of a seeded vector potential on MAC edges: the discrete
                     divergence is zero to round-off and the normal velocity vanishes on walls
  sphere_source      density 1 inside a sphere (plume source)
  spherical_obstacles  seeded solid mask of spheres
All arrays are float64 numpy; callers convert to float32.
"""
import numpy as np


def cell_centres(n):
    return (np.arange(n) + 0.5) / n


def gaussian_blob(n, centre=(0.5, 0.3, 0.5), sigma=0.1, peak=1.0):
    x = cell_centres(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    return peak * np.exp(-r2 / (2.0 * sigma * sigma))


def sphere_source(n, centre=(0.5, 0.1, 0.5), radius=0.06, value=1.0):
    x = cell_centres(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    return np.where(r2 <= radius * radius, value, 0.0)


def curl_noise(n, seed=0, kmin=1, kmax=4, amplitude=1.0):
    """Divergence-free MAC velocity: curl of A = sum_k a_k sin(pi k.x)-type modes that vanish on the
    walls (wavenumbers kmin..kmax per axis, amplitude ~ |k|^-2).  n is the cube's edge, or a shape
    (nx, ny, nz) for a box of cells of edge 1 / max(shape).  Returns (u, v, w)."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n) if np.isscalar(n) else tuple(int(a) for a in n)
    nx, ny, nz = shape
    h = 1.0 / max(shape)
    # node and cell-centre coordinates of each axis, scaled to the box edge (x / L in [0, 1]; the
    # longest edge is 1, written as in the cube so that cubic fields keep their bits)
    node = [np.arange(m + 1) * h if m == max(shape) else np.arange(m + 1) / m for m in shape]
    cent = [cell_centres(m) for m in shape]
    # edge-centred potentials: Ax at (i+1/2, j, k), Ay at (i, j+1/2, k), Az at (i, j, k+1/2)
    Ax = np.zeros((nx, ny + 1, nz + 1)); Ay = np.zeros((nx + 1, ny, nz + 1)); Az = np.zeros((nx + 1, ny + 1, nz))
    for kx in range(kmin, kmax + 1):
        for ky in range(kmin, kmax + 1):
            for kz in range(kmin, kmax + 1):
                kk = float(kx * kx + ky * ky + kz * kz)
                a = rng.normal(size=3) / kk
                # sin(pi k x / L) vanishes at the walls x = 0, L for every integer k
                sx_n, sy_n, sz_n = np.sin(np.pi * kx * node[0]), np.sin(np.pi * ky * node[1]), np.sin(np.pi * kz * node[2])
                cx_c, cy_c, cz_c = np.cos(np.pi * kx * cent[0]), np.cos(np.pi * ky * cent[1]), np.cos(np.pi * kz * cent[2])
                Ax += a[0] * cx_c[:, None, None] * sy_n[None, :, None] * sz_n[None, None, :]
                Ay += a[1] * sx_n[:, None, None] * cy_c[None, :, None] * sz_n[None, None, :]
                Az += a[2] * sx_n[:, None, None] * sy_n[None, :, None] * cz_c[None, None, :]
    # u = curl A on faces
    u = (Az[:, 1:, :] - Az[:, :-1, :]) / h - (Ay[:, :, 1:] - Ay[:, :, :-1]) / h      # (nx+1, ny, nz)
    v = (Ax[:, :, 1:] - Ax[:, :, :-1]) / h - (Az[1:, :, :] - Az[:-1, :, :]) / h      # (nx, ny+1, nz)
    w = (Ay[1:, :, :] - Ay[:-1, :, :]) / h - (Ax[:, 1:, :] - Ax[:, :-1, :]) / h      # (nx, ny, nz+1)
    u[0] = u[-1] = 0.0
    v[:, 0] = v[:, -1] = 0.0
    w[:, :, 0] = w[:, :, -1] = 0.0
    s = amplitude / max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    return u * s, v * s, w * s


def divergence(u, v, w, h):
    return ((u[1:] - u[:-1]) + (v[:, 1:] - v[:, :-1]) + (w[:, :, 1:] - w[:, :, :-1])) / h


def spherical_obstacles(n, count=8, rmin=0.03, rmax=0.08, seed=0, keep_clear=None):
    """uint8 mask of `count` seeded spheres (radius in [rmin, rmax], uniform centres)."""
    rng = np.random.default_rng(seed)
    x = cell_centres(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    mask = np.zeros((n, n, n), dtype=np.uint8)
    for _ in range(count):
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(r, 1.0 - r, size=3)
        mask |= ((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r * r).astype(np.uint8)
    if keep_clear is not None:
        mask[keep_clear > 0] = 0
    return mask


def cfl_dt(u, v, w, h, cfl):
    m = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    return cfl * h / m if m > 0 else h
