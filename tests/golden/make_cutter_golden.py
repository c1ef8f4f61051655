"""Writes tests/golden/cutter_ref.npz: pieces produced by the reference's own face cutter.

The cutter is the reference's header src/facecutter.h compiled as-is into
oracle/_ref/libref_facecutter.so by oracle/Makefile (run `make -C oracle` with the reference
sources at /root/reference).  Inputs are seeded corner colours covering generic faces, colours
on exact integers and halves, stretched faces and twisted (saddle) colour patterns, with and
without triangle subdivision.  Outputs are stored as flat arrays so the fixture is pure data.

    python tests/golden/make_cutter_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def cases(rng, n):
    out = []
    for t in range(n):
        base = rng.uniform(-3, 3, size=2)
        col = base + rng.normal(scale=rng.choice([0.2, 0.8, 2.0]), size=(4, 2))
        if t % 4 == 0:
            col = np.round(col * 2) / 2                   # exact integers and halves
        if t % 7 == 0:                                      # saddle in x: corners alternate around an integer
            k = np.floor(base[0]) + 0.5
            col[:, 0] = k + np.array([0.3, -0.3, 0.3, -0.3]) * rng.uniform(0.5, 2.0)
        ma = float(np.float32(rng.choice([0.4, 0.1]))) if t % 3 == 0 else 1.0
        out.append((col.astype(np.float32), ma))
    return out


def main():
    if not O.ref_cutter_available():
        raise SystemExit("oracle/_ref/libref_facecutter.so missing: run make -C oracle with the reference present")
    rng = np.random.default_rng(20240601)
    colours, min_area, npieces, nverts, verts, cen, area = [], [], [], [], [], [], []
    for col, ma in cases(rng, 600):
        pieces = O.ref_cut_face(col.astype(np.float64), min_area=ma)
        colours.append(col); min_area.append(ma); npieces.append(len(pieces))
        for v, c, a in pieces:
            nverts.append(len(v)); verts.append(v.astype(np.float32)); cen.append(c.astype(np.float32))
            area.append(np.float32(a))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cutter_ref.npz"),
                        colours=np.array(colours), min_area=np.array(min_area, dtype=np.float32),
                        npieces=np.array(npieces, dtype=np.int32), nverts=np.array(nverts, dtype=np.int32),
                        verts=np.concatenate(verts), centroid=np.array(cen), area=np.array(area))
    print("wrote", len(colours), "faces,", sum(npieces), "pieces")


if __name__ == "__main__":
    main()
