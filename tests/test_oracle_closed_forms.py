"""Closed-form pins of the oracle's WENO6 weights and of its per-piece quadrature (CPU only).

WENO6 (reference src/WENO6_interpolation_3d.cpp:63-123).  The smoothness indicator of candidate
r is the Jiang-Shu integral  S_r = sum_{l=1..3} int_0^1 (d^l p_r / dx^l)^2 dx  (h = 1) of the
cubic p_r through samples r..r+3 (nodes at -2..3, point t in [0, 1)); the ideal weights C_r are
the unique ones for which sum C_r p_r(t) is the quintic Lagrange interpolant of all six samples;
alpha_r = C_r / (1e-6 + S_r)^2 and the result is sum alpha_r p_r / sum alpha_r.  Everything here
is derived with numpy polynomials, independently of the oracle's integer coefficient table, and
compared with oracle.weno1d on rough samples whose weights are far from ideal (and on samples
whose indicators are comparable to the 1e-6 epsilon, which fixes where epsilon enters).

Quadrature (reference src/polygon_3D.cpp:257-487, src/calc_psi_3D.cpp:177-263).  A uniform
velocity moving every vertex by a fraction of a cell leaves every backtraced face planar and
axis-aligned, so only y faces carry y area and each cell's new value is the difference of the
y-antiderivative Psi between its backtraced top and bottom faces.  For a density that depends on
y alone and is a polynomial the WENO6 samples reproduce (linear with the 1-point Gauss rule,
cubic with 3 points), the psi prefix, the psi point on the face level and the Gauss points of the
residual integral are all exact, and one step must give the exact cell average of the shifted
profile in the interior.  Any change to the Gauss nodes or weights, the interval split at
f = 1/2, the half-cell rule of the psi prefix or its weights breaks this.  A velocity component
linear in y (v = V0 + b y, u and w uniform) compresses the control volumes in y only; the exact
new components are again closed forms.
"""
import numpy as np
import pytest

XS = np.arange(-2, 4, dtype=float)   # sample nodes of one 1D WENO6 evaluation


def _dist(t):
    return t - XS


def _cubic_values(f, t):
    """Candidates p_r(t), r = 0, 1, 2 (cubics through samples r..r+3)."""
    return np.array([np.polyval(np.polyfit(XS[r:r + 4], f[r:r + 4], 3), t) for r in range(3)])


def _jiang_shu(f):
    s = []
    for r in range(3):
        p = np.polyfit(XS[r:r + 4], f[r:r + 4], 3)
        tot = 0.0
        for l in (1, 2, 3):
            d = np.polyder(p, l)
            q = np.polyint(np.polymul(d, d))
            tot += np.polyval(q, 1.0) - np.polyval(q, 0.0)
        s.append(tot)
    return np.array(s)


def _ideal_weights_solved(t):
    # sum_r C_r p_r(t) = quintic(t) for every data vector: solved on the six unit vectors
    # (unique for t strictly between the nodes; at a node every candidate gives the same value)
    A = np.zeros((6, 3)); b = np.zeros(6)
    for i in range(6):
        e = np.zeros(6); e[i] = 1.0
        A[i] = _cubic_values(e, t)
        b[i] = np.polyval(np.polyfit(XS, e, 5), t)
    c, res, rank, _ = np.linalg.lstsq(A, b, rcond=None)
    assert rank == 3 and np.abs(A @ c - b).max() < 1e-12   # consistent: such weights exist
    return c


def _ideal_weights(t):
    # the closed form the reference writes (d4 d5 / 20, -d5 d0 / 10, d0 d1 / 20 with d_i = t - x_i),
    # checked against _ideal_weights_solved below
    d = _dist(t)
    return np.array([d[4] * d[5] / 20, -d[5] * d[0] / 10, d[0] * d[1] / 20])


def weno6_numpy(f, t, eps_form="reference"):
    q = _cubic_values(f, t)
    C = _ideal_weights(t)
    S = _jiang_shu(f)
    if eps_form == "reference":
        a = C / (1e-6 + S) ** 2
    elif eps_form == "optimized":   # the ISPC variant (src/weno.ispc:46-50): 1e-30 added to S^2
        a = C / (1e-30 + S ** 2)
    else:   # a plausible mistake: epsilon added to S^2
        a = C / (1e-6 + S ** 2)
    return float((a * q).sum() / a.sum())


def test_ideal_weights_closed_form():
    for t in np.linspace(0.05, 0.95, 10):
        assert np.abs(_ideal_weights_solved(t) - _ideal_weights(t)).max() < 1e-12
    # and with them the combination of the candidates is the quintic at any t
    rng = np.random.default_rng(3)
    for t in np.concatenate([[0.0], rng.uniform(0, 1, 20)]):
        f = rng.normal(size=6)
        assert abs((_ideal_weights(t) * _cubic_values(f, t)).sum() - np.polyval(np.polyfit(XS, f, 5), t)) < 1e-12


@pytest.mark.parametrize("scale", [3e-4, 3e-3, 3e-2, 1.0, 50.0])
def test_weno1d_matches_jiang_shu(oracle, scale):
    rng = np.random.default_rng(7)
    worst, sep = 0.0, 0.0
    for _ in range(150):
        t = rng.uniform(0, 1)
        f = scale * rng.normal(size=6)                   # rough: weights far from ideal
        f[rng.integers(0, 6)] += scale * 4.0 * rng.choice([-1, 1])   # a jump somewhere
        want = weno6_numpy(f, t)
        got = oracle.weno1d(f, _dist(t), 1.0, 0)
        worst = max(worst, abs(got - want) / np.abs(f).max())
        sep = max(sep, abs(weno6_numpy(f, t, "mistake") - want) / np.abs(f).max())
    assert worst < 1e-11
    # where the indicators are comparable to epsilon the check distinguishes the reference's
    # (eps + S)^2 from eps + S^2 (at larger scales epsilon no longer matters to either)
    if scale <= 3e-3:
        assert sep > 1e-4


def test_weno1d_weights_far_from_ideal(oracle):
    # a step at the stencil's edge: the candidate that straddles it gets a vanishing weight, so
    # the result is far from the quintic; the Jiang-Shu form must still agree
    for f in (np.array([0.0, 1.0, 1.0, 1.0, 1.0, 1.0]), np.array([2.0, 2.0, 2.0, 2.0, 2.0, -1.0])):
        for t in (0.1, 0.5, 0.9):
            want = weno6_numpy(f, t)
            quintic = np.polyval(np.polyfit(XS, f, 5), t)
            assert abs(want - quintic) > 1e-3
            assert abs(oracle.weno1d(f, _dist(t), 1.0, 0) - want) < 1e-12


def test_product_weno6_matches_jiang_shu(cutter_host):
    # the library's FP32 difference form (fia_math.hpp weno6), built for the host
    rng = np.random.default_rng(11)
    worst = 0.0
    for scale in (3e-2, 1.0, 50.0):
        for _ in range(200):
            t = float(np.float32(rng.uniform(0, 1)))
            f = (scale * rng.normal(size=6)).astype(np.float32).astype(np.float64)
            f[rng.integers(0, 6)] += scale * 3.0
            f = f.astype(np.float32).astype(np.float64)
            want = weno6_numpy(f, t)
            got = cutter_host.weno6(f, t, 1.0)
            worst = max(worst, abs(got - want) / np.abs(f).max())
    assert worst < 2e-5


# ---------------------------------------------------------------------------------------------
# "WENO6-optimized" (the reference's ISPC interpolation of quadrature points): the same candidates,
# ideal weights and indicators, alpha = C / (1e-30 + S^2), and coordinates clamped to the samples

@pytest.mark.parametrize("scale", [3e-4, 3e-3, 1.0, 50.0])
def test_weno1d_opt_matches_jiang_shu(oracle, scale):
    rng = np.random.default_rng(13)
    worst, sep = 0.0, 0.0
    for _ in range(150):
        t = rng.uniform(0, 1)
        f = scale * rng.normal(size=6)
        f[rng.integers(0, 6)] += scale * 4.0 * rng.choice([-1, 1])
        want = weno6_numpy(f, t, "optimized")
        got = oracle.weno1d(f, _dist(t), 1.0, opt=True)
        worst = max(worst, abs(got - want) / np.abs(f).max())
        sep = max(sep, abs(weno6_numpy(f, t) - want) / np.abs(f).max())
    assert worst < 1e-11
    if scale <= 3e-3:   # where S is comparable to the reference form's 1e-6 the two rules differ
        assert sep > 1e-4


def test_product_weno6_opt_matches_jiang_shu(cutter_host):
    # the library's FP32 difference form with the optimized weights (weno6<V, true>), on rough
    # data, on data whose indicators vanish (constant samples: the ideal weights), and with the
    # units factor kappa of the psi field (S in units of kappa / 180)
    rng = np.random.default_rng(17)
    worst = 0.0
    for scale in (3e-2, 1.0, 50.0):
        for _ in range(200):
            t = float(np.float32(rng.uniform(0, 1)))
            f = (scale * rng.normal(size=6)).astype(np.float32).astype(np.float64)
            f[rng.integers(0, 6)] += scale * 3.0
            f = f.astype(np.float32).astype(np.float64)
            want = weno6_numpy(f, t, "optimized")
            worst = max(worst, abs(cutter_host.weno6(f, t, 1.0, opt=True) - want) / np.abs(f).max())
            # kappa: samples in cells^2-scaled units, S rescaled back (S_phys = kappa S / 180)
            kap = 1.0 / 4096.0
            g = f / kap
            worst = max(worst, abs(cutter_host.weno6(g, t, kap, dbl=True, opt=True) * kap - want) / np.abs(f).max())
    assert worst < 2e-5
    for t in (0.0, 0.3, 1.0):
        assert cutter_host.weno6(np.full(6, 2.5), t, 1.0, opt=True) == 2.5


def test_split_clamped_holds_the_coordinate_to_the_samples(cutter_host):
    m = 10
    for s, b, t in [(-3.0, 0, 0.0), (-0.25, 0, 0.0), (0.0, 0, 0.0), (3.25, 3, 0.25), (8.5, 8, 0.5), (9.0, 8, 1.0),
                    (9.75, 8, 1.0), (40.0, 8, 1.0)]:
        got_b, got_t = cutter_host.split_clamped(s, m)
        assert (got_b, got_t) == (b, t), s


def test_interpolate_opt_clamps_to_the_samples(oracle):
    # a field cubic along each axis: inside the samples both rules are exact; beyond the first or
    # last sample the optimized rule returns that sample's value exactly (the WENO6 rule, on the
    # clamped samples, comes to the same value only up to its epsilon)
    n = 12
    h = 1.0 / n
    c = lambda s: 0.3 + 0.7 * s - 0.05 * s ** 2 + 0.004 * s ** 3
    idx = np.arange(n, dtype=np.float64)
    f = c(idx)[:, None, None] + 2.0 * c(idx)[None, :, None] - c(idx)[None, None, :]
    off = (0.5, 0.5, 0.5)
    inside = np.array([[3.3, 5.7, 7.1], [4.02, 6.5, 5.5], [5.5, 4.25, 6.75]]) * h
    want = c(inside[:, 0] / h - 0.5) + 2.0 * c(inside[:, 1] / h - 0.5) - c(inside[:, 2] / h - 0.5)
    for method in ("WENO6", "WENO6-optimized"):
        got = oracle.interpolate(f, off, h, method, inside)
        assert np.abs(got - want).max() < 1e-12, method
    beyond = np.array([[n + 0.4, 5.7, 7.1], [-1.3, 6.5, 5.5], [5.5, 4.25, n + 2.0]]) * h
    s = np.clip(beyond / h - 0.5, 0.0, n - 1.0)
    want = c(s[:, 0]) + 2.0 * c(s[:, 1]) - c(s[:, 2])
    got = oracle.interpolate(f, off, h, "WENO6-optimized", beyond)
    assert np.abs(got - want).max() < 1e-12


# ---------------------------------------------------------------------------------------------
# exact fractional translation
# ---------------------------------------------------------------------------------------------
N = 32
H = 1.0 / N
SHIFT = (1.37, 0.61, 2.29)   # cells per step
MARGIN = 8                    # > ceil(shift) + 3 (WENO6 half width) + 1


def _uniform_velocity(n, s):
    u = np.full((n + 1, n, n), s[0]); v = np.full((n, n + 1, n), s[1]); w = np.full((n, n, n + 1), s[2])
    return u, v, w


def _interior(a, m=MARGIN):
    return a[m:-m, m:-m, m:-m]


def _shifted_cell_average(prof, n, h, dy):
    """Average of prof(y) over every cell shifted down by dy (physical), as an (n,n,n) field."""
    yc = (np.arange(n) + 0.5) * h
    lo, hi = yc - dy - 0.5 * h, yc - dy + 0.5 * h
    P = np.polyint(prof)
    avg = (np.polyval(P, hi) - np.polyval(P, lo)) / h
    return np.broadcast_to(avg[None, :, None], (n, n, n))


@pytest.mark.parametrize("prof,ng", [
    (np.array([0.8, 0.3]), 1),                        # linear, one Gauss point per interval
    (np.array([1.7, -2.1, 0.6, 0.5]), 3),             # cubic, three points
    (np.array([1.7, -2.1, 0.6, 0.5]), 5),
])
def test_fractional_translation_is_exact(oracle, prof, ng):
    n, h = N, H
    yc = (np.arange(n) + 0.5) * h
    rho = np.broadcast_to(np.polyval(prof, yc)[None, :, None], (n, n, n)).copy()
    u, v, w = _uniform_velocity(n, SHIFT)
    dt = h   # displacement = velocity * dt = SHIFT cells
    got = oracle.move_substances(rho, u, v, w, h, dt, split="xz-CellFaceAligned", method="WENO6",
                                 use_integral=True, correction=True, num_gauss=ng)
    want = _shifted_cell_average(prof, n, h, SHIFT[1] * h)
    err = np.abs(_interior(got) - _interior(want)).max() / np.abs(want).max()
    assert err < 1e-12, err


def test_fractional_translation_detects_a_wrong_rule(oracle):
    # the same cubic with too few Gauss points for it: the check is sensitive to the rule
    prof = np.array([1.7, -2.1, 0.6, 0.5])
    n, h = N, H
    yc = (np.arange(n) + 0.5) * h
    rho = np.broadcast_to(np.polyval(prof, yc)[None, :, None], (n, n, n)).copy()
    u, v, w = _uniform_velocity(n, SHIFT)
    got = oracle.move_substances(rho, u, v, w, h, h, split="xz-CellFaceAligned", method="WENO6",
                                 use_integral=True, correction=True, num_gauss=1)
    want = _shifted_cell_average(prof, n, h, SHIFT[1] * h)
    assert np.abs(_interior(got) - _interior(want)).max() / np.abs(want).max() > 1e-9


def test_velocity_linear_in_y_is_exact(oracle):
    """v = V0 + b y on the y faces, u and w uniform: every backtraced control volume is the
    original one moved along x and z and stretched in y (vertex y' = y - dt (V0 + b y)), so
    the new components are the exact averages over the stretched volumes."""
    n, h = N, H
    V0, b, U0, W0 = 0.61, 0.5, 1.37, 2.29
    dt = h
    u = np.full((n + 1, n, n), U0); w = np.full((n, n, n + 1), W0)
    yf = np.arange(n + 1) * h
    v = np.broadcast_to((V0 + b * yf)[None, :, None], (n, n + 1, n)).copy()
    nu, nv, nw = oracle.advect_velocity(u, v, w, h, dt, dim=-1, method="WENO6", use_integral=True,
                                        correction=True, num_gauss=1)
    back = lambda y: y - dt * (V0 + b * y)
    # v control volumes: y in [(j - 1/2) h, (j + 1/2) h] around face j
    lo, hi = back((np.arange(n + 1) - 0.5) * h), back((np.arange(n + 1) + 0.5) * h)
    v_want = (V0 * (hi - lo) + 0.5 * b * (hi ** 2 - lo ** 2)) / h
    m = MARGIN
    v_got = nv[m:-m, m:-m, m:-m]
    assert np.abs(v_got - v_want[None, m:-m, None]).max() / np.abs(v_want).max() < 1e-12
    # u and w control volumes: y in [j h, (j + 1) h]; the uniform component is scaled by the stretch
    stretch = (back(h) - back(0.0)) / h
    assert np.abs(nu[m:-m, m:-m, m:-m] - U0 * stretch).max() / U0 < 1e-12
    assert np.abs(nw[m:-m, m:-m, m:-m] - W0 * stretch).max() / W0 < 1e-12
    # and the stretch is visible at this size (the check is not vacuous)
    assert abs(stretch - 1.0) > 1e-2
