"""Pins the oracle's interpolation and quadrature by closed forms (CPU only).

WENO6 (reference src/WENO6_interpolation_3d.cpp:63-123): each cubic candidate reproduces cubics
and the weights sum to one, so WENO6 is exact on cubics; with smoothness indicators far below
the 1e-6 epsilon the weights are the ideal ones and WENO6 equals the quintic Lagrange
interpolant of the six samples; at a node it returns the node value.  Gauss-Legendre rules
(reference src/gauss_quadrature_points.cpp:18-299) against numpy's leggauss.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 3, 5, 7, 9, 11, 13, 25, 49])
def test_gauss_legendre_matches_numpy(oracle, n):
    x, w = oracle.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(np.sort(x), np.sort(xr), atol=1e-14)
    assert np.allclose(w[np.argsort(x)], (wr / 2)[np.argsort(xr)], atol=1e-14)
    assert abs(w.sum() - 1.0) < 1e-14
    # exact for degree 2n-1 on [-1,1] (weights normalised to the interval length 2)
    for d in range(2 * n):
        exact = (1.0 - (-1.0) ** (d + 1)) / (d + 1) / 2.0
        assert abs((w * x ** d).sum() - exact) < 1e-12


def test_gauss_rejects_unsupported(oracle):
    with pytest.raises(ValueError):
        oracle.gauss_legendre(2)


def _dist(t, h=1.0):
    # signed distances from the six sample points (at -2..3) to the point t in [0,1)
    return np.array([(t - (i - 2)) * h for i in range(6)])


@pytest.mark.parametrize("prec", [0, 1])
def test_weno1d_constant_and_cubic(oracle, prec):
    rng = np.random.default_rng(0)
    tol = 1e-12 if prec == 0 else 2e-5
    for _ in range(200):
        t = rng.uniform(0, 1)
        c = rng.normal(size=4)
        xs = np.arange(-2, 4, dtype=float)
        f = c[0] + c[1] * xs + c[2] * xs ** 2 + c[3] * xs ** 3
        exact = c[0] + c[1] * t + c[2] * t ** 2 + c[3] * t ** 3
        got = oracle.weno1d(f, _dist(t), 1.0, prec)
        assert abs(got - exact) < tol * max(1, np.abs(f).max())
        k = rng.normal()
        assert abs(oracle.weno1d(np.full(6, k), _dist(t), 1.0, prec) - k) < tol * max(1, abs(k))


def test_weno1d_ideal_weights_give_quintic(oracle):
    rng = np.random.default_rng(1)
    xs = np.arange(-2, 4, dtype=float)
    for _ in range(100):
        t = rng.uniform(0, 1)
        f = 1e-9 * rng.normal(size=6)      # S << 1e-6: ideal weights
        quintic = np.polyval(np.polyfit(xs, f, 5), t)
        assert abs(oracle.weno1d(f, _dist(t), 1.0, 0) - quintic) < 1e-6 * 1e-9


def test_weno1d_node_value(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        f = rng.normal(size=6)
        assert abs(oracle.weno1d(f, _dist(0.0), 1.0, 0) - f[2]) < 1e-13


def test_weno1d_discontinuity_picks_smooth_side(oracle):
    # a step between samples 4 and 5: the interpolant on [x2, x3] stays near the left state
    f = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.0])
    v = oracle.weno1d(f, _dist(0.5), 1.0, 0)
    assert abs(v - 1.0) < 1e-3


OFFSETS = [(0.5, 0.5, 0.5), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)]


@pytest.mark.parametrize("off", OFFSETS)
def test_weno3d_exact_on_tricubic_interior(oracle, off):
    n = 16; h = 1.0 / n
    rng = np.random.default_rng(3)
    shape = tuple(n + (1 if o == 0.0 else 0) for o in off)
    axes = [(np.arange(shape[a]) + off[a]) * h for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    c = rng.normal(size=(4, 4, 4))
    P = lambda x, y, z: sum(c[i, j, k] * x ** i * y ** j * z ** k for i in range(4) for j in range(4) for k in range(4))
    field = P(X, Y, Z)
    pts = rng.uniform(0.3, 0.7, size=(300, 3))    # stencils stay inside the grid
    got = oracle.interpolate(field, off, h, "WENO6", pts)
    assert np.abs(got - P(pts[:, 0], pts[:, 1], pts[:, 2])).max() < 1e-9


def test_weno3d_clamped_constant(oracle):
    n = 8; h = 1.0 / n
    field = np.full((n, n, n), 2.5)
    pts = np.random.default_rng(4).uniform(-0.2, 1.2, size=(200, 3))
    assert np.allclose(oracle.interpolate(field, (0.5, 0.5, 0.5), h, "WENO6", pts), 2.5, atol=1e-13)


@pytest.mark.parametrize("off", OFFSETS)
def test_trilinear_exact_on_trilinear(oracle, off):
    n = 10; h = 0.1
    rng = np.random.default_rng(5)
    shape = tuple(n + (1 if o == 0.0 else 0) for o in off)
    axes = [(np.arange(shape[a]) + off[a]) * h for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    c = rng.normal(size=8)
    P = lambda x, y, z: (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z
                         + c[7] * x * y * z)
    pts = rng.uniform(0.1, 0.9, size=(300, 3))
    got = oracle.interpolate(P(X, Y, Z), off, h, "linear", pts)
    assert np.abs(got - P(pts[:, 0], pts[:, 1], pts[:, 2])).max() < 1e-12


def test_1dy_linear_uses_containing_column(oracle):
    n = 8; h = 1.0 / n
    rng = np.random.default_rng(6)
    field = rng.normal(size=(n, n, n))
    pts = rng.uniform(0.2, 0.8, size=(100, 3))
    got = oracle.interpolate(field, (0.5, 0.5, 0.5), h, "1Dy_linear", pts)
    for p, g in zip(pts, got):
        i, k = int(np.floor(p[0] / h)), int(np.floor(p[2] / h))
        s = p[1] / h - 0.5
        j = int(np.floor(s)); t = s - j
        jl, jh = min(max(j, 0), n - 1), min(max(j + 1, 0), n - 1)
        assert abs(g - ((1 - t) * field[i, jl, k] + t * field[i, jh, k])) < 1e-12
