"""Pins the oracle's face cutter (reference src/facecutter.h, src/split_face_3D.cpp:17-93).

1. Bit-exact against the reference's own cutter: golden pieces in tests/golden/cutter_ref.npz
   (written by tests/golden/make_cutter_golden.py from oracle/_ref, the reference header compiled
   as-is), and live against oracle/_ref when it is present.
2. Invariants the geometry fixes: the pieces tile the xi-eta square, every piece of a non-twisted
   face is convex, known decompositions of undeformed faces, vector areas add up to the face's.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cutter_ref.npz")


def _golden():
    z = np.load(GOLDEN)
    out, vi, pi = [], 0, 0
    for f in range(len(z["colours"])):
        pieces = []
        for _ in range(z["npieces"][f]):
            nv = z["nverts"][pi]
            pieces.append((z["verts"][vi:vi + nv], z["centroid"][pi], z["area"][pi]))
            vi += nv; pi += 1
        out.append((z["colours"][f], float(z["min_area"][f]), pieces))
    return out


def test_fp32_cutter_bit_exact_against_reference_golden(oracle):
    faces = _golden()
    assert len(faces) == 600
    for col, ma, ref in faces:
        mine = oracle.cut_face_xieta(col.astype(np.float64), min_area=ma, cprec=1)
        assert len(mine) == len(ref)
        for (v, c, a), (rv, rc, ra) in zip(mine, ref):
            assert np.array_equal(v.astype(np.float32), rv)
            assert np.array_equal(c.astype(np.float32), rc)
            assert np.float32(a) == ra


def test_fp32_cutter_bit_exact_against_reference_build(oracle):
    if not oracle.ref_cutter_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(99)
    for t in range(1500):
        col = (rng.uniform(-4, 4, size=2) + rng.normal(scale=rng.choice([0.3, 1.0, 3.0]), size=(4, 2)))
        if t % 3 == 0:
            col = np.round(col * 4) / 4
        col = col.astype(np.float32).astype(np.float64)
        ma = 1.0 if t % 2 else float(np.float32(0.15))
        mine = oracle.cut_face_xieta(col, min_area=ma, cprec=1)
        ref = oracle.ref_cut_face(col, min_area=ma)
        assert len(mine) == len(ref)
        for (v, c, a), (rv, rc, ra) in zip(mine, ref):
            assert np.array_equal(v, rv) and np.array_equal(c, rc) and a == ra


def _convex(poly):
    p = np.asarray(poly)
    n = len(p)
    cr = []
    for i in range(n):
        a, b, c = p[i], p[(i + 1) % n], p[(i + 2) % n]
        cr.append((b[0] - a[0]) * (c[1] - b[1]) - (b[1] - a[1]) * (c[0] - b[0]))
    cr = np.array(cr)
    return np.all(cr >= -1e-12) or np.all(cr <= 1e-12)


@pytest.mark.parametrize("cprec", [0, 1])
def test_pieces_tile_the_square(oracle, cprec):
    rng = np.random.default_rng(7)
    for t in range(400):
        col = rng.uniform(-3, 3, size=2) + rng.normal(scale=1.5, size=(4, 2))
        pieces = oracle.cut_face_xieta(col, cprec=cprec)
        total = sum(a for _, _, a in pieces)
        assert abs(total - 1.0) < (1e-12 if cprec == 0 else 1e-5)
        for v, c, a in pieces:
            assert np.all(v >= -1e-6) and np.all(v <= 1 + 1e-6)


def _twisted(col):
    # an iso line crosses the boundary of the colour-interpolated square more than twice
    for d in range(2):
        c = col[:, d]
        for k in range(int(np.floor(c.min())) + 1, int(np.ceil(c.max()))):
            s = (c - k) < 0
            if sum(s[i] != s[(i + 1) % 4] for i in range(4)) > 2:
                return True
    return False


def test_pieces_of_untwisted_faces_are_convex(oracle):
    rng = np.random.default_rng(8)
    seen = 0
    for t in range(600):
        col = rng.uniform(-3, 3, size=2) + rng.normal(scale=1.2, size=(4, 2))
        if _twisted(col):
            continue
        seen += 1
        for v, c, a in oracle.cut_face_xieta(col, cprec=0):
            assert _convex(v)
    assert seen > 300


def test_known_decompositions(oracle):
    h = 0.25
    # undeformed y-face of cell (1,1,1): corners (x,z) in cells, xz-CellFaceAligned -> 4 quarters
    y = 2 * h
    quad = np.array([[1, 2, 1], [1, 2, 2], [2, 2, 2], [2, 2, 1]], dtype=float) * h
    cen, an = oracle.split_face(quad, h, "xz-CellFaceAligned")
    assert len(cen) == 4
    assert np.allclose(sorted(map(tuple, np.round(cen / h, 12))),
                       sorted([(1.25, 2, 1.25), (1.25, 2, 1.75), (1.75, 2, 1.25), (1.75, 2, 1.75)]))
    assert np.allclose(an[:, 1], h * h / 4) and np.allclose(an[:, [0, 2]], 0)
    # the same face with the reference's default split string: one piece
    cen, an = oracle.split_face(quad, h, "none")
    assert len(cen) == 1 and np.allclose(cen[0] / h, (1.5, 2, 1.5)) and np.isclose(an[0, 1], h * h)
    # undeformed x-face: split in z only (x colour constant), 2 halves, zero y area
    quad = np.array([[2, 1, 1], [2, 2, 1], [2, 2, 2], [2, 1, 2]], dtype=float) * h
    cen, an = oracle.split_face(quad, h, "xz-CellFaceAligned")
    assert len(cen) == 2 and np.allclose(an[:, 1], 0) and np.allclose(np.abs(an[:, 0]), h * h / 2)
    # xz-AxisAligned: cell-aligned face not cut
    quad = np.array([[1, 2, 1], [1, 2, 2], [2, 2, 2], [2, 2, 1]], dtype=float) * h
    assert len(oracle.split_face(quad, h, "xz-AxisAligned")[0]) == 1


def test_vector_areas_add_up(oracle):
    rng = np.random.default_rng(9)
    h = 0.1
    for t in range(200):
        quad = (np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0]], dtype=float) + rng.uniform(3, 6, size=3)
                + rng.normal(scale=0.4, size=(4, 3))) * h
        cen, an = oracle.split_face(quad, h, "xz-CellFaceAligned")
        # vector area of the boundary loop of the bilinear face = 1/2 (d1 x d2) of the diagonals
        d1, d2 = quad[2] - quad[0], quad[3] - quad[1]
        assert np.allclose(an.sum(axis=0), 0.5 * np.cross(d1, d2), atol=1e-13)


def test_twisted_face_splits_into_three(oracle):
    # x colours alternate around 0 at the corners: one iso line crosses all four edges
    col = np.array([[0.4, 0.5], [-0.4, 0.5], [0.4, 0.5], [-0.4, 0.5]])
    pieces = oracle.cut_face_xieta(col, cprec=0)
    assert len(pieces) == 3
    assert abs(sum(a for _, _, a in pieces) - 1.0) < 1e-12
