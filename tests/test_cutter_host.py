"""The library's face cutters (csrc/fia_cutter.hpp) built for the host.  The general cutter is
checked against the oracle's recursive cutter (a restatement of reference src/facecutter.h:59-324)
on twisted faces whose corner colours span many cells: the set of pieces must be the same, and no
capacity of the cutter may run out where the reference's recursion would not.  The segment
cutter's one-pass form (all pieces of a face cut at most once per colour) is checked against its
piece-by-piece form and against the oracle's split of the same face (reference
src/split_face_3D.cpp:26-92 with the recursive cutter)."""
import ctypes

import numpy as np
import pytest

@pytest.fixture(scope="module")
def cutter(cutter_host):
    return cutter_host.lib


def _cut(L, col, max_pieces=1024):
    col = np.ascontiguousarray(col, dtype=np.float64).reshape(8)
    cen = np.zeros((max_pieces, 2)); area = np.zeros(max_pieces)
    dp = ctypes.POINTER(ctypes.c_double)
    n = L.host_general_cut(col.ctypes.data_as(dp), 1.0, max_pieces, cen.ctypes.data_as(dp), area.ctypes.data_as(dp))
    return n, cen[:max(n, 0)], area[:max(n, 0)]


@pytest.mark.parametrize("span", [0.8, 3.0, 8.0])
def test_general_cutter_matches_oracle_on_stretched_faces(cutter, oracle, span):
    # random corner colours within +-span cells: most of these quads are twisted (bow ties) and
    # are crossed by up to 2 span integer lines in each direction
    rng = np.random.default_rng(int(span * 10))
    faces = 300 if span > 4 else 600
    for t in range(faces):
        col = np.array([10.3, 20.7]) + span * rng.uniform(-1.0, 1.0, size=(4, 2))
        n, cen, area = _cut(cutter, col)
        assert n >= 0, (t, n)
        ref = oracle.cut_face_xieta(col, min_area=1.0, cprec=0, max_pieces=1024)
        rc = np.array([c for _, c, _ in ref]).reshape(-1, 2)
        ra = np.array([a for _, _, a in ref])
        assert n == len(ref), t
        key = lambda c: np.lexsort((np.round(c[:, 1], 9), np.round(c[:, 0], 9)))
        assert np.abs(cen[key(cen)] - rc[key(rc)]).max(initial=0.0) < 1e-12, t
        assert np.abs(area[key(cen)] - ra[key(rc)]).max(initial=0.0) < 1e-12, t
        assert abs(area.sum() - ra.sum()) < 1e-12


def test_split_coord_keeps_far_points_outside_every_field(cutter):
    # a sliver piece can have a huge but finite centroid; its stencil base must stay far outside
    # every field without overflowing an int (the interior test of the row loads reads b - 2 and
    # b + 3), and ordinary coordinates split exactly as floor
    b, t = ctypes.c_int(), ctypes.c_float()
    for s in (1e12, -1e12, 2.2e9, -2.2e9, 2.0 ** 31, 1e300, -1e300, float("inf"), float("-inf"), float("nan")):
        inside = cutter.host_split_coord(s, 1 << 20, ctypes.byref(b), ctypes.byref(t))
        assert not inside and abs(b.value) <= 1 << 24, s
    for s in (0.0, 2.5, 13.999, 255.25, -0.5, -3.75):
        inside = cutter.host_split_coord(s, 256, ctypes.byref(b), ctypes.byref(t))
        assert b.value == int(np.floor(s)) and abs(t.value - (s - np.floor(s))) < 1e-7
        assert inside == (2 <= b.value <= 252)


def test_fastdiv_matches_integer_division(cutter):
    # face ids decode with the multiply-shift division of CompGrid (ids < 2^31): every divisor up
    # to 600 with every numerator up to 70,000 and the largest multiples below 2^31, then a
    # million random pairs
    assert cutter.host_fastdiv_mismatches(600, 70000, 1000000) == 0


def _backtraced_faces(rng, count, family, strain):
    """Faces of one family as a backtrace leaves them: the lattice face (cells) with every corner
    moved by a smooth shift plus a random displacement of size `strain`."""
    base = {0: [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)],
            1: [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)],
            2: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]}[family]
    for _ in range(count):
        origin = rng.integers(4, 40, size=3).astype(np.float64)
        shift = rng.uniform(-4.0, 4.0, size=3)
        quad = origin + np.array(base, dtype=np.float64) + shift + strain * rng.uniform(-1, 1, size=(4, 3))
        yield quad


@pytest.mark.parametrize("family,variant", [(1, 0), (0, 1), (2, 2)])
def test_one_pass_segment_cutter_matches_per_piece_and_oracle(cutter, oracle, family, variant):
    # xz-CellFaceAligned colours (x - 1/2, z - 1/2); faces with at most one cut per colour, as the
    # face kernel of each family takes them (variant: y faces both splits, x faces z, z faces x)
    rng = np.random.default_rng(40 + family)
    dp = ctypes.POINTER(ctypes.c_double)
    checked = 0
    shapes = {}
    for quad in _backtraced_faces(rng, 6000, family, 0.35):
        col = np.ascontiguousarray(quad[:, [0, 2]] - 0.5)
        v12 = np.ascontiguousarray(quad.reshape(12))
        one = np.zeros((4, 4)); per = np.zeros((4, 4))
        n = cutter.host_face_pieces(v12.ctypes.data_as(dp), col.ctypes.data_as(dp), variant,
                                    one.ctypes.data_as(dp), per.ctypes.data_as(dp))
        if n < 0:
            continue
        checked += 1
        shapes[n] = shapes.get(n, 0) + 1
        one, per = one[:n], per[:n]
        # the two forms: the same segments up to the chord points they share
        assert np.abs(one - per).max() <= 1e-12, (quad, one, per)
        # the oracle's split of the face: the same non-empty pieces (centre, y area)
        cen, an = oracle.split_face(quad, 1.0, "xz-CellFaceAligned")
        ref = [(c, a[1]) for c, a in zip(cen, an) if abs(a[1]) > 1e-13]
        got = [(r[:3], r[3]) for r in one if abs(r[3]) > 1e-13]
        assert len(ref) == len(got), (quad, ref, got)
        for c, a in got:
            d = [np.abs(c - rc).max() for rc, _ in ref]
            k = int(np.argmin(d))
            assert d[k] <= 1e-10 and abs(a - ref[k][1]) <= 1e-12, (quad, c, a, ref)
    # the sample covers the cut configurations the kernels see: one piece (no cut), two (one cut)
    # and, for y faces, four (both cuts)
    assert checked > 1000 and shapes.get(2, 0) > 100 and (family != 1 or shapes.get(4, 0) > 100)


def _cut_counts(lo, hi):
    """Interior integers of the open colour ranges (lo, hi), element-wise."""
    first = np.ceil(lo)
    first = np.where(lo - first == 0, first + 1, first)
    last = np.floor(hi)
    last = np.where(hi - last == 0, last - 1, last)
    return np.maximum(last - first + 1, 0)


def test_gpu_step_scene_covers_both_segment_cutter_paths(oracle, scenes):
    # the 24^3, CFL 4 scene of the GPU step tests (tests/test_gpu_pass.py: h = 1/32, curl noise seed
    # 1), backtraced with the oracle: every face family has faces the one-pass form takes and faces
    # with several cuts along a colour, which the face kernel's block cuts piece by piece, so the
    # GPU parity tests run both paths
    n, h = 24, 1.0 / 32
    u, v, w = scenes.curl_noise(n, seed=1)
    dt = scenes.cfl_dt(u, v, w, h, 4.0)
    P = oracle.backtrace_vertices(u, v, w, h, dt, method="WENO6") / h
    corners = {0: [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)], 1: [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)],
               2: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]}
    one_pass = {0: lambda nx, nz: (nx == 0) & (nz <= 1), 1: lambda nx, nz: (nx <= 1) & (nz <= 1),
                2: lambda nx, nz: (nz == 0) & (nx <= 1)}
    for fam, cs in corners.items():
        ext = [n + (a == fam) for a in range(3)]
        quad = np.stack([P[c[0]:c[0] + ext[0], c[1]:c[1] + ext[1], c[2]:c[2] + ext[2]] for c in cs])
        cx, cz = quad[..., 0] - 0.5, quad[..., 2] - 0.5   # xz-CellFaceAligned colours
        fast = one_pass[fam](_cut_counts(cx.min(0), cx.max(0)), _cut_counts(cz.min(0), cz.max(0)))
        assert fast.sum() > 0 and (~fast).sum() > 0, (fam, fast.sum(), (~fast).sum())


def _nonzero_pieces(pieces):
    out = sorted((round(float(c[0]), 9), round(float(c[1]), 9), round(float(a), 9)) for c, a in pieces if abs(a) > 1e-12)
    return np.array(out).reshape(-1, 3)


def test_pipeline_on_corner_lying_on_a_cut_line(cutter_host, oracle):
    # A face of the v-component grid at a wall row of pinned vertices (axis-aligned colours): two
    # corners sit exactly on the z cut line and one also on the x cut line.  The crossing of the
    # edge that ends on the x line is that corner; computed through the rounded reciprocal its
    # weight missed 1 by an ulp, its z colour rounded below the z line, and the pipeline's loop of
    # piece (x >= 8, z < 2) folded over itself (xi-eta area 0.787 instead of 0.287).
    col = np.array([[8.0, 2.0], [9.0, 2.0], [8.727837562561035, 1.1455326676368713],
                    [7.925276763737202, 2.635922431945801]])
    cen, area = cutter_host.pipeline_pieces(col)
    got = _nonzero_pieces(zip(cen, area))
    assert abs(area.sum() - 1.0) < 1e-12
    for cprec in (0, 1):
        ref = _nonzero_pieces((c, a) for _, c, a in oracle.cut_face_xieta(col, cprec=cprec))
        assert got.shape == ref.shape and np.abs(got - ref).max() < 1e-6


def test_pipeline_matches_recursive_cut_on_degenerate_faces(cutter_host, oracle):
    # corners exactly on integer colours (pinned or clamped vertices) next to moved ones: every
    # face must tile the square, and its pieces must be those of the reference's recursive cut in
    # one of its two arithmetics (the FP32 form of src/facecutter.h, or FP64)
    rng = np.random.default_rng(5)
    bad = checked = 0
    for t in range(3000):
        col = np.floor(rng.uniform(0, 4, size=(4, 2))) + 0.5
        on = rng.choice(4, size=rng.integers(1, 4), replace=False)
        col[on] = np.round(col[on] + rng.uniform(-0.5, 0.5, size=(len(on), 2)))
        off = np.setdiff1d(np.arange(4), on)
        col[off] += rng.uniform(-0.9, 0.9, size=(len(off), 2))
        twisted, degenerate = cutter_host.face_class(col)
        if twisted or not degenerate:
            continue   # the recursive cutter or the segment cutter takes these
        checked += 1
        cen, area = cutter_host.pipeline_pieces(col)   # raises if a stage sees a twisted loop
        assert abs(area.sum() - 1.0) < 1e-9, (t, col.tolist(), area.sum())
        got = _nonzero_pieces(zip(cen, area))
        ok = False
        for cprec in (0, 1):
            ref = _nonzero_pieces((c, a) for _, c, a in oracle.cut_face_xieta(col, cprec=cprec))
            ok = ok or (got.shape == ref.shape and np.abs(got - ref).max() < 1e-5)
        bad += not ok
    assert checked > 1000
    assert bad <= 3, bad
