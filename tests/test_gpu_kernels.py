"""GPU parity of the individual kernels against the oracle (through the C ABI).

Tolerances: the kernels compute in FP32 (interpolation, quadrature) with FP64 cutting and sums,
the oracle in FP64; FP32 interpolation of O(1) data agrees to ~1e-6 relative.
"""
import numpy as np
import pytest

from conftest import rel_linf

pytestmark = pytest.mark.gpu

FIELD_OFFSETS = {  # staggered sample offsets (cells) and dims deltas
    "centre": ((0.5, 0.5, 0.5), (0, 0, 0)),
    "face_x": ((0.0, 0.5, 0.5), (1, 0, 0)),
    "face_y": ((0.5, 0.0, 0.5), (0, 1, 0)),
    "face_z": ((0.5, 0.5, 0.0), (0, 0, 1)),
}


def smooth_field(shape, seed):
    rng = np.random.default_rng(seed)
    x = [np.linspace(0, 1, s) for s in shape]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        k = rng.uniform(0.5, 3.0, size=3); ph = rng.uniform(0, 6, size=3)
        f += rng.normal() * np.sin(k[0] * 3 * X + ph[0]) * np.cos(k[1] * 3 * Y + ph[1]) * np.sin(k[2] * 3 * Z + ph[2])
    return f


@pytest.mark.parametrize("method", ["WENO6", "linear", "1Dy_linear", "1Dy_WENO6", "WENO6-optimized"])
@pytest.mark.parametrize("kind", list(FIELD_OFFSETS))
def test_interpolate_matches_oracle(fia, oracle, ctx, method, kind):
    print("case interpolate", method, kind, flush=True)
    n = 12; h = 1.0 / n
    off, de = FIELD_OFFSETS[kind]
    shape = (n + de[0], n + de[1], n + de[2])
    f = smooth_field(shape, 3).astype(np.float32)
    rng = np.random.default_rng(5)
    pts = rng.uniform(-0.1, 1.1, size=(2000, 3)).astype(np.float32)   # includes clamped stencils
    got = fia.interpolate(ctx, f, off, h, method, pts)
    ref = oracle.interpolate(f.astype(np.float64), off, h, method, pts.astype(np.float64))
    assert rel_linf(got, ref) < 2e-5


@pytest.mark.parametrize("mode,integral", [("const", True), ("linear", True), ("WENO6", True), ("WENO6", False)])
@pytest.mark.parametrize("ng", [1, 3])
def test_psi_matches_oracle(fia, oracle, ctx, scenes, mode, integral, ng):
    print("case psi", mode, integral, ng, flush=True)
    n = 16; h = 1.0 / n
    q = scenes.gaussian_blob(n, sigma=0.15).astype(np.float32)
    got = fia.calc_psi_y(ctx, q, h, mode=mode, use_integral=integral, num_gauss=ng)
    ref = oracle.calc_psi(q.astype(np.float64), h, mode=mode, use_integral=integral, num_gauss=ng)
    assert rel_linf(got, ref) < 1e-6


# Columns of 250 and 257 cells (eight segments of 32 and 33, the most the held forms take) and 300
# (segments of 38: the sliding-window WENO6 form and the unheld cheap rules), ragged in x and z;
# the WENO6 integral with 3 Gauss points per half cell takes the sliding-window form at any length.
@pytest.mark.parametrize("ny", [250, 257, 300])
@pytest.mark.parametrize("mode,ng", [("WENO6", 1), ("WENO6", 3), ("linear", 1)])
def test_psi_long_columns_match_oracle(fia, oracle, ctx, ny, mode, ng):
    print("case psi long columns", ny, mode, ng, flush=True)
    h = 1.0 / 64
    rng = np.random.default_rng(ny + ng)
    y = (np.arange(ny) + 0.5) / ny
    q = (0.5 + 0.4 * np.sin(7.0 * y)[None, :, None] + 0.05 * rng.standard_normal((5, ny, 70))).astype(np.float32)
    got = fia.calc_psi_y(ctx, q, h, mode=mode, use_integral=True, num_gauss=ng)
    ref = oracle.calc_psi(q.astype(np.float64), h, mode=mode, use_integral=True, num_gauss=ng)
    assert rel_linf(got, ref) < 1e-6


@pytest.mark.parametrize("method", ["WENO6", "linear", "1Dy_linear"])
@pytest.mark.parametrize("dim", [-1, 0, 1, 2])
def test_backtrace_matches_oracle(fia, oracle, ctx, scenes, method, dim):
    print("case backtrace", method, dim, flush=True)
    # h = 2^-k: lattice points then sit exactly on the columns of the piecewise-constant 1Dy
    # rule in both FP32 and FP64 (for other h the reference's floor(x/h) depends on rounding)
    n = 12; h = 1.0 / 16
    u, v, w = [a.astype(np.float32) for a in scenes.curl_noise(n, seed=2)]
    dt = scenes.cfl_dt(u, v, w, h, 2.0)
    grid = fia.Grid3D(n, n, n, h, None, u, v, w)
    disp = fia.backtrace_vertices(ctx, grid, dt, method, dim)
    ref = oracle.backtrace_vertices(u, v, w, h, dt, method, comp=dim)
    e = [int(dim == a) for a in range(3)]
    lat = np.stack(np.meshgrid(*[np.arange(n + e[a] + 1) for a in range(3)], indexing="ij"), axis=-1)
    pos = (lat - 0.5 * np.array(e)) + disp.astype(np.float64)
    assert np.abs(pos * h - ref).max() < 2e-6


def _quads(rng, n, h, stretch):
    base = rng.uniform(2, 6, size=(n, 1, 3))
    corners = np.array([[0, 0, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1]], dtype=float)
    q = base + corners[None] * rng.uniform(0.5, 1.5, size=(n, 1, 1))
    q = q + rng.normal(scale=stretch, size=q.shape)
    return q * h


@pytest.mark.parametrize("split", ["xz-CellFaceAligned", "xz-AxisAligned", "x-CellFaceAligned_z-AxisAligned", "none"])
@pytest.mark.parametrize("min_area", [1.0, 0.1])
def test_split_faces_match_oracle(fia, oracle, ctx, split, min_area):
    print("case split", split, min_area, flush=True)
    rng = np.random.default_rng(11)
    h = 0.125
    quads = _quads(rng, 300, h, 0.6)
    got = fia.split_faces(ctx, quads, h, split, minimal_area=min_area, max_pieces=256)
    for i in range(len(quads)):
        cen, an = oracle.split_face(quads[i].astype(np.float32).astype(np.float64), h, split, min_area=min_area)
        gc, ga = got[i]
        assert len(gc) == len(cen), i
        key = lambda c: np.lexsort(tuple(np.round(c[:, a], 9) for a in (2, 1, 0)))
        oc, oa = cen[key(cen)], an[key(cen)]
        pc, pa = gc[key(gc)], ga[key(gc)]
        assert np.abs(oc - pc).max() < 1e-9 * 8
        assert np.abs(oa - pa).max() < 1e-12 * 8


def test_context_on_callers_stream(fia, ctx, scenes):
    # a context on a stream the caller created with the HIP runtime libfia.so links gives the same
    # result as a context with a stream of its own, on that stream
    import ctypes
    import numpy as np
    print("case context on caller stream", flush=True)
    hip = ctypes.CDLL("libamdhip64.so.7")
    s = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    try:
        c2 = fia.Context(0, stream=s.value)
        try:
            assert c2.stream() == s.value
            n = 16
            h = 1.0 / n
            rho = scenes.gaussian_blob(n, centre=(0.5, 0.45, 0.5), sigma=0.12).astype(np.float32)
            u, v, w = [a.astype(np.float32) for a in scenes.curl_noise(n, seed=2)]
            dt = scenes.cfl_dt(u, v, w, h, 2.0)
            g1 = fia.Grid3D(n, n, n, h, rho.copy(), u, v, w)
            g2 = fia.Grid3D(n, n, n, h, rho.copy(), u, v, w)
            fia.move_substances_3d(ctx, g1, 0, dt, split_method="xz-CellFaceAligned")
            fia.move_substances_3d(c2, g2, 0, dt, split_method="xz-CellFaceAligned")
            assert np.array_equal(g1.density, g2.density)
        finally:
            c2.close()
    finally:
        hip.hipStreamDestroy(s)
