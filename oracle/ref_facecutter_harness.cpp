// ref_facecutter_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Builds the reference's own face cutter (reference src/facecutter.h, header-only, standard
// library only) into oracle/_ref/libref_facecutter.so so the oracle's restatement of it can be
// checked bit for bit.  The glue below does what split_face_optimized does around the cutter
// (reference src/split_face_3D.cpp:45-64): cut the unit xi-eta square with the four corner
// colours, pass polygons with >= 3 vertices to subdivide(), collect the emitted pieces.
// Compiled by oracle/Makefile with -I pointing at the reference sources; nothing of the
// reference is copied here.
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <algorithm>
#include "facecutter.h"

extern "C" int ref_cut_face(const float *colours, float min_area, int max_pieces, int *nverts,
                            float *verts, float *centroid, float *area) {
    point2 col[4];
    for (int i = 0; i < 4; ++i) col[i] = point2(colours[2 * i], colours[2 * i + 1]);
    const vertex2 square[4] = {
        vertex2(point2(0.0, 0.0), col[0]), vertex2(point2(1.0, 0.0), col[1]),
        vertex2(point2(1.0, 1.0), col[2]), vertex2(point2(0.0, 1.0), col[3])};
    int count = 0;
    bool overflow = false;
    cut_face(square, 4, [&](const vertex2 *polygon, int num) {
        if (num < 3) return;
        point2 tmp[MAX_TMP_POINTS];
        for (int i = 0; i < num; ++i) tmp[i] = polygon[i].p;
        subdivide(tmp, num, min_area, [&](const point2 *poly, unsigned char n, const point2 &c, const float a) {
            if (count >= max_pieces) { overflow = true; return; }
            nverts[count] = n;
            for (int k = 0; k < n; ++k) {
                verts[(count * MAX_TMP_POINTS + k) * 2] = poly[k].x[0];
                verts[(count * MAX_TMP_POINTS + k) * 2 + 1] = poly[k].x[1];
            }
            centroid[2 * count] = c.x[0];
            centroid[2 * count + 1] = c.x[1];
            area[count] = a;
            ++count;
        });
    });
    return overflow ? -1 : count;
}
