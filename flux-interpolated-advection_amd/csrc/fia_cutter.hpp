// fia_cutter.hpp - face cutting in xi-eta space, FP64 predicates (gfx950).
//
// The reference cuts a backtraced face recursively (src/facecutter.h:59-224): the unit xi-eta
// square carries a "colour" (cell coordinates of the backtraced corners, src/split_face_3D.cpp:
// 26-43) that is interpolated linearly along polygon edges; each polygon is split at the first
// integer strictly inside its colour range, x colours first, then z.  The pieces are therefore
// the regions  k_s <= c_x < k_{s+1},  m <= c_z < m+1  of the square (sgn(0) = +1: a vertex on an
// iso line belongs to the upper side).
//
// Fast path (no arrays, everything in registers): each piece (s, m) is the square clipped by up
// to four colour half-planes, streamed corner by corner through a four-stage Sutherland-Hodgman
// pipeline whose last stage accumulates the xi-eta area moments and the y component of the 3D
// vector area of the mapped loop.  When every cut meets a polygon in exactly two points this is
// the reference's piece; each stage counts the sign changes it sees and reports a face whose
// colour crosses an iso line more than twice (a twisted face), which is then cut by the exact
// recursive restatement in fia_cutter_general (fallback kernel).
#pragma once
#include "fia_internal.hpp"
#include "fia_math.hpp"

namespace fia {

struct CutVert {
    double xi, eta, cx, cz;
};

// Face geometry: the four backtraced corners in cell units (physical frame) and the bilinear map
// (reference src/cell_face_3D.cpp:68-73), written around corner 0 so that a face whose corners
// share a coordinate maps every point to exactly that coordinate.
struct FaceGeom {
    double p0[3], e1[3], e3[3], e13[3];   // p(xi,eta) = p0 + xi e1 + eta e3 + xi eta e13
    FIA_HD void init(const double v[4][3]) {
        for (int a = 0; a < 3; ++a) {
            p0[a] = v[0][a];
            e1[a] = v[1][a] - v[0][a];
            e3[a] = v[3][a] - v[0][a];
            e13[a] = v[0][a] - v[1][a] + v[2][a] - v[3][a];
        }
    }
    // point relative to corner 0
    FIA_HD void rel(double xi, double eta, double out[3]) const {
        for (int a = 0; a < 3; ++a) out[a] = xi * e1[a] + eta * e3[a] + xi * eta * e13[a];
    }
};

// Last stage: loop moments.  area2 = sum(xi_k eta_{k+1} - xi_{k+1} eta_k) (twice the signed
// area), mxi/meta the centroid moments (reference src/facecutter.h:226-253), a?2 = twice the
// vector area of the mapped loop (reference src/split_face_3D.cpp:78-92; for a
// closed loop the fan around the centre equals the loop sum).
struct Accum {
    int n;
    CutVert first, prev;
    double fr[3], pr[3];
    double area2, mxi, meta, ax2, ay2, az2;
    FIA_HD void reset() { n = 0; area2 = mxi = meta = ax2 = ay2 = az2 = 0.0; }
    FIA_HD void edge(const CutVert &a, const double ra[3], const CutVert &b, const double rb[3]) {
        const double cr = a.xi * b.eta - b.xi * a.eta;
        area2 += cr;
        mxi += cr * (a.xi + b.xi);
        meta += cr * (a.eta + b.eta);
        ax2 += ra[1] * rb[2] - ra[2] * rb[1];
        ay2 += ra[2] * rb[0] - ra[0] * rb[2];
        az2 += ra[0] * rb[1] - ra[1] * rb[0];
    }
    FIA_HD void put(const FaceGeom &g, const CutVert &v) {
        double r[3];
        g.rel(v.xi, v.eta, r);
        if (n == 0) { first = v; fr[0] = r[0]; fr[1] = r[1]; fr[2] = r[2]; }
        else edge(prev, pr, v, r);
        prev = v; pr[0] = r[0]; pr[1] = r[1]; pr[2] = r[2];
        ++n;
    }
    FIA_HD void close() { if (n > 1) edge(prev, pr, first, fr); }
};

// One Sutherland-Hodgman stage: keeps c_dim >= iso (upper = true) or c_dim < iso (upper = false).
// A vertex exactly on the iso line counts as upper (sgn(0) = +1, reference src/facecutter.h:55-57).
struct Stage {
    bool active, upper;
    int dim;
    double iso;
    int n, crossings;
    CutVert first, prev;
    bool prev_in, first_in;
    FIA_HD void reset(bool act, int d, double k, bool up) {
        active = act; dim = d; iso = k; upper = up; n = 0; crossings = 0;
    }
    FIA_HD bool inside(const CutVert &v) const {
        const double c = (dim == 0 ? v.cx : v.cz) - iso;
        return upper ? (c >= 0.0) : (c < 0.0);
    }
    // crossing point on edge a -> b (reference src/facecutter.h:151-161).  An end point on the
    // line is the crossing itself, taken exactly: the weights of the reference's form (t1 = -v0 /
    // (v1 - v0) through a rounded reciprocal) can miss 1 by an ulp there, and the interpolated
    // other colour of that near-copy of the corner can then round to the far side of a later stage's
    // line, which folds the clipped loop over itself (the recursive cut of the reference survives
    // the same near-copy as a zero-area sliver).
    FIA_HD CutVert cross(const CutVert &a, const CutVert &b) const {
        const double v0 = (dim == 0 ? a.cx : a.cz) - iso, v1 = (dim == 0 ? b.cx : b.cz) - iso;
        if (v1 == 0.0) return b;
        if (v0 == 0.0) return a;
        const double inv = 1.0 / (v1 - v0);
        const double t0 = v1 * inv, t1 = -v0 * inv;
        CutVert r;
        r.xi = t0 * a.xi + t1 * b.xi;
        r.eta = t0 * a.eta + t1 * b.eta;
        if (dim == 0) { r.cx = iso; r.cz = t0 * a.cz + t1 * b.cz; }
        else { r.cz = iso; r.cx = t0 * a.cx + t1 * b.cx; }
        return r;
    }
};

// Four stages and the accumulator.  Forwarding is compile-time recursion over the stage index,
// so every stage's state stays in registers (no dynamically indexed arrays).
struct Pipeline {
    Stage st[4];
    Accum acc;

    template <int S> FIA_HD void push(const FaceGeom &g, const CutVert &v) {
        if constexpr (S == 4) {
            acc.put(g, v);
        } else {
            Stage &T = st[S];
            if (!T.active) { push<S + 1>(g, v); return; }
            const bool in = T.inside(v);
            if (T.n == 0) {
                T.first = v; T.first_in = in;
            } else if (T.prev_in != in) {
                ++T.crossings;
                push<S + 1>(g, T.cross(T.prev, v));
            }
            if (in) push<S + 1>(g, v);
            T.prev = v; T.prev_in = in; ++T.n;
        }
    }
    template <int S> FIA_HD void close(const FaceGeom &g) {
        if constexpr (S == 4) {
            acc.close();
        } else {
            Stage &T = st[S];
            if (T.active && T.n > 0 && T.prev_in != T.first_in) {
                ++T.crossings;
                push<S + 1>(g, T.cross(T.prev, T.first));
            }
            close<S + 1>(g);
        }
    }
    FIA_HD bool twisted() const {
        return (st[0].active && st[0].crossings > 2) || (st[1].active && st[1].crossings > 2) ||
               (st[2].active && st[2].crossings > 2) || (st[3].active && st[3].crossings > 2);
    }
};

// Interior integers of the open colour range (min, max) (reference src/facecutter.h:128-131).
FIA_HD void interior_integers(double lo, double hi, int &first, int &last) {
    first = (int)ceil(lo);
    last = (int)floor(hi);
    if (lo - first == 0.0) first += 1;
    if (hi - last == 0.0) last -= 1;
}

// ---------------------------------------------------------------------------------------------
// Piece moments from boundary segments (untwisted faces, no branches on the piece's shape).
//
// On an untwisted face every cut meets the polygon it cuts in two points, so the piece (s, m)
// is the convex region of the square with  xl <= c_x < xh  (x strip s) and  zl <= c_z < zh
// (z band m) whose boundary is: the parts of the four square edges inside strip and band (the
// colour is linear along an edge), the parts of the x chords (cut lines of xl and xh, between
// their crossings on the square's edges, c_z linear along them) inside the band, and the z
// chords of the strip (between the crossings of zl and zh with the strip's boundary).  Area
// moments and the mapped loop's vector area are sums over oriented boundary segments, so each
// piece is the same fixed sequence of segment terms for every face: no vertex lists, and in
// exact arithmetic the same polygon as the recursive cut.  Orientation is counter-clockwise
// (the square's corner order); a chord runs from where the boundary leaves its region to where
// it re-enters.  Inactive bounds are +-kNoBound.
// ---------------------------------------------------------------------------------------------
constexpr double kNoBound = 1e300;

// [lo, hi] &= {u in [0,1]: L <= c0 + u d <= H} for the colour c0 + u d along a segment, with
// r = 1 / d computed once per segment by the caller (not read when d == 0); a constant colour
// keeps everything or nothing, decided with the half-open rule L <= c < H.
FIA_HD void clip_band(double c0, double d, double r, double L, double H, double &lo, double &hi) {
    if (d == 0.0) {
        if (!(c0 >= L && c0 < H)) { lo = 1.0; hi = 0.0; }
        return;
    }
    const double uL = (L - c0) * r, uH = (H - c0) * r;
    lo = fmax(lo, d > 0.0 ? uL : uH);
    hi = fmin(hi, d > 0.0 ? uH : uL);
}

template <bool ALL> struct SegAcc {
    double area2 = 0.0, mxi = 0.0, meta = 0.0, ax2 = 0.0, ay2 = 0.0, az2 = 0.0;
    FIA_HD void add(const FaceGeom &g, double xa, double ea, double xb, double eb) {
        const double cr = xa * eb - xb * ea;
        area2 += cr;
        mxi += cr * (xa + xb);
        meta += cr * (ea + eb);
        double ra[3], rb[3];
        g.rel(xa, ea, ra);
        g.rel(xb, eb, rb);
        ay2 += ra[2] * rb[0] - ra[0] * rb[2];
        if (ALL) {
            ax2 += ra[1] * rb[2] - ra[2] * rb[1];
            az2 += ra[0] * rb[1] - ra[1] * rb[0];
        }
    }
};

// Where the boundary leaves / re-enters one region (last one seen of each).
struct Chord {
    bool has_exit = false, has_entry = false;
    double xe = 0.0, ee = 0.0, ze = 0.0, xn = 0.0, en = 0.0, zn = 0.0;
    // boundary segment a -> b with region membership ia, ib; the crossing is at a + t (b - a)
    FIA_HD void see(bool ia, bool ib, double t, double xa, double ea, double za, double xb, double eb, double zb) {
        if (ia == ib) return;
        const double x = xa + t * (xb - xa), e = ea + t * (eb - ea), z = za + t * (zb - za);
        if (ia) { has_exit = true; xe = x; ee = e; ze = z; }
        else { has_entry = true; xn = x; en = e; zn = z; }
    }
    // the same update written with selects, for chords kept in arrays: a branch per field end
    // lets the compiler merge the two stores into one through a selected address, which puts
    // the array in scratch memory
    FIA_HD void see_sel(bool ia, bool ib, double t, double xa, double ea, double za, double xb, double eb,
                        double zb) {
        const bool ex = ia && !ib, nn = ib && !ia;
        const double x = xa + t * (xb - xa), e = ea + t * (eb - ea), z = za + t * (zb - za);
        xe = ex ? x : xe; ee = ex ? e : ee; ze = ex ? z : ze;
        xn = nn ? x : xn; en = nn ? e : en; zn = nn ? z : zn;
        has_exit = has_exit || ex;
        has_entry = has_entry || nn;
    }
};

// z-chord crossings of a part of the strip boundary segment a -> b (colour z linear along it,
// r = 1 / (zb - za)): membership is decided at the part's end colours z0, z1, and a crossing is
// placed on the whole segment at u = (level - za) r.
FIA_HD void see_z(Chord &lo, Chord &hi, double zl, double zh, double z0, double z1, double xa, double ea, double za,
                  double xb, double eb, double zb, double r) {
    const bool la = z0 >= zl, lb = z1 >= zl, ha = z0 < zh, hb = z1 < zh;
    if (la == lb && ha == hb) return;
    lo.see(la, lb, (zl - za) * r, xa, ea, za, xb, eb, zb);
    hi.see(ha, hb, (zh - za) * r, xa, ea, za, xb, eb, zb);
}

// Each segment's colour differences are inverted once and shared by its crossings and clips:
// one divide per colour of each square edge and one per x chord.
template <bool ALL>
FIA_HD void piece_moments(const double col[4][2], const FaceGeom &g, double xl, double xh, double zl, double zh,
                          SegAcc<ALL> &acc) {
    const double sx[4] = {0.0, 1.0, 1.0, 0.0}, se[4] = {0.0, 0.0, 1.0, 1.0};
    Chord cxl, cxh, czl, czh;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        const double xa = sx[i], ea = se[i], xb = sx[j], eb = se[j];
        const double ca = col[i][0], cb = col[j][0], za = col[i][1], zb = col[j][1];
        const double dc = cb - ca, dz = zb - za;
        const double rc = 1.0 / dc, rz = 1.0 / dz;   // inf on a constant colour, then unused
        // crossings of the x cut lines with this edge (the x chords' ends)
        {
            const bool la = ca >= xl, lb = cb >= xl, ha = ca < xh, hb = cb < xh;
            if (la != lb || ha != hb) {
                cxl.see(la, lb, (xl - ca) * rc, xa, ea, za, xb, eb, zb);
                cxh.see(ha, hb, (xh - ca) * rc, xa, ea, za, xb, eb, zb);
            }
        }
        // the edge's part inside the strip: z chord crossings, then its part inside the band
        double lo = 0.0, hi = 1.0;
        clip_band(ca, dc, rc, xl, xh, lo, hi);
        if (lo < hi) {
            see_z(czl, czh, zl, zh, za + lo * dz, za + hi * dz, xa, ea, za, xb, eb, zb, rz);
            double lo2 = lo, hi2 = hi;
            clip_band(za, dz, rz, zl, zh, lo2, hi2);
            if (lo2 < hi2)
                acc.add(g, xa + lo2 * (xb - xa), ea + lo2 * (eb - ea), xa + hi2 * (xb - xa), ea + hi2 * (eb - ea));
        }
    }
    // x chords (exit -> entry of their regions), their parts inside the band
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const Chord &c = k == 0 ? cxl : cxh;
        if (c.has_exit && c.has_entry) {
            const double dz = c.zn - c.ze, rz = 1.0 / dz;
            see_z(czl, czh, zl, zh, c.ze, c.zn, c.xe, c.ee, c.ze, c.xn, c.en, c.zn, rz);
            double lo = 0.0, hi = 1.0;
            clip_band(c.ze, dz, rz, zl, zh, lo, hi);
            if (lo < hi)
                acc.add(g, c.xe + lo * (c.xn - c.xe), c.ee + lo * (c.en - c.ee), c.xe + hi * (c.xn - c.xe),
                        c.ee + hi * (c.en - c.ee));
        }
    }
    // z chords of the strip
    if (czl.has_exit && czl.has_entry) acc.add(g, czl.xe, czl.ee, czl.xn, czl.en);
    if (czh.has_exit && czh.has_entry) acc.add(g, czh.xe, czh.ee, czh.xn, czh.en);
}

// Moment terms of one oriented boundary segment a -> b (those SegAcc<false>::add sums), so that a
// segment bounding two pieces is evaluated once and added to one, subtracted from the other.
struct SegTerm {
    double area2, mxi, meta, ay2;
};
FIA_HD SegTerm seg_term(const FaceGeom &g, double xa, double ea, double xb, double eb) {
    const double cr = xa * eb - xb * ea;
    double ra[3], rb[3];
    g.rel(xa, ea, ra);
    g.rel(xb, eb, rb);
    return SegTerm{cr, cr * (xa + xb), cr * (ea + eb), ra[2] * rb[0] - ra[0] * rb[2]};
}
struct PieceAcc {
    double area2 = 0.0, mxi = 0.0, meta = 0.0, ay2 = 0.0;
    FIA_HD void add(const SegTerm &t) { area2 += t.area2; mxi += t.mxi; meta += t.meta; ay2 += t.ay2; }
    FIA_HD void sub(const SegTerm &t) { area2 -= t.area2; mxi -= t.mxi; meta -= t.meta; ay2 -= t.ay2; }
};

// Every piece of an untwisted, non-degenerate face cut at most once per colour, in one pass: the
// x level X0 (when xcut; strips s = 0: c_x < X0 and s = 1: c_x >= X0) and the z level Z0 (when
// zcut; bands m likewise).  NX / NZ (0 or 1) compile the strip / band split in or out; a level
// that is not cut leaves one strip / band.  The boundary is piece_moments' for each piece, with
// each segment's terms evaluated once:
//  - a square edge's part inside strip s and band m belongs to piece (s, m) alone, with the same
//    end points as piece_moments computes for that piece;
//  - the x chord (exit -> entry of c_x >= X0) inside band m bounds (1, m), and reversed (0, m);
//  - the z chord of strip s (exit -> entry of c_z >= Z0) bounds (s, 1), and reversed (s, 0).
// Strip 0 reuses strip 1's x chord and its z crossing, so the pieces on both sides of a chord
// share its points (piece_moments recomputes them from the other end, equal up to rounding).
template <int NX, int NZ>
FIA_HD void face_pieces(const double col[4][2], const FaceGeom &g, bool xcut, double X0, bool zcut, double Z0,
                        PieceAcc (&acc)[2][2]) {
    const double sx[4] = {0.0, 1.0, 1.0, 0.0}, se[4] = {0.0, 0.0, 1.0, 1.0};
    // strip / band bounds; a split that is not cut leaves strip / band 1 empty
    const double X = (NX && xcut) ? X0 : kNoBound, Z = (NZ && zcut) ? Z0 : kNoBound;
    const double xl[2] = {-kNoBound, X}, xh[2] = {X, kNoBound};
    const double zl[2] = {-kNoBound, Z}, zh[2] = {Z, kNoBound};
    Chord cx;        // x chord: boundary leaving / entering c_x >= X0
    Chord cz[2];     // z chord of each strip: its boundary leaving / entering c_z >= Z0
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        const double xa = sx[i], ea = se[i], xb = sx[j], eb = se[j];
        const double ca = col[i][0], cb = col[j][0], za = col[i][1], zb = col[j][1];
        const double dc = cb - ca, dz = zb - za;
        const double rc = 1.0 / dc, rz = 1.0 / dz;   // inf on a constant colour, then unused
        if (NX) cx.see(ca >= X, cb >= X, (X - ca) * rc, xa, ea, za, xb, eb, zb);
#pragma unroll
        for (int s = 0; s <= NX; ++s) {
            double lo = 0.0, hi = 1.0;
            clip_band(ca, dc, rc, xl[s], xh[s], lo, hi);
            if (lo < hi) {
                if (NZ) cz[s].see_sel(za + lo * dz >= Z, za + hi * dz >= Z, (Z - za) * rz, xa, ea, za, xb, eb, zb);
#pragma unroll
                for (int m = 0; m <= NZ; ++m) {
                    double lo2 = lo, hi2 = hi;
                    clip_band(za, dz, rz, zl[m], zh[m], lo2, hi2);
                    if (lo2 < hi2)
                        acc[s][m].add(seg_term(g, xa + lo2 * (xb - xa), ea + lo2 * (eb - ea), xa + hi2 * (xb - xa),
                                               ea + hi2 * (eb - ea)));
                }
            }
        }
    }
    if (NX && cx.has_exit && cx.has_entry) {
        const double dz = cx.zn - cx.ze, rz = 1.0 / dz;
        if (NZ) {
            // strip 1 runs the chord exit -> entry, strip 0 entry -> exit
            const bool l0 = cx.ze >= Z, l1 = cx.zn >= Z;
            const double t = (Z - cx.ze) * rz;
            cz[1].see_sel(l0, l1, t, cx.xe, cx.ee, cx.ze, cx.xn, cx.en, cx.zn);
            cz[0].see_sel(l1, l0, t, cx.xe, cx.ee, cx.ze, cx.xn, cx.en, cx.zn);
        }
#pragma unroll
        for (int m = 0; m <= NZ; ++m) {
            double lo = 0.0, hi = 1.0;
            clip_band(cx.ze, dz, rz, zl[m], zh[m], lo, hi);
            if (lo < hi) {
                const SegTerm t = seg_term(g, cx.xe + lo * (cx.xn - cx.xe), cx.ee + lo * (cx.en - cx.ee),
                                           cx.xe + hi * (cx.xn - cx.xe), cx.ee + hi * (cx.en - cx.ee));
                acc[1][m].add(t);
                acc[0][m].sub(t);
            }
        }
    }
    if (NZ) {
#pragma unroll
        for (int s = 0; s <= NX; ++s) {
            if (cz[s].has_exit && cz[s].has_entry) {
                const SegTerm t = seg_term(g, cz[s].xe, cz[s].ee, cz[s].xn, cz[s].en);
                acc[s][1].add(t);
                acc[s][0].sub(t);
            }
        }
    }
}

// True if an integer strictly inside the range of one colour has the corners on alternating
// sides (c0, c2 >= k > c1, c3 or the reverse): that iso line crosses the square's boundary four
// times and the face is twisted.  Without such a line for either colour no cut of any piece
// meets more than two crossings (cutting a loop replaces a run of vertices on one side of a
// later line by two points on the same side, so it never adds sign changes) and the pieces are
// those of piece_moments.
FIA_HD bool alternating_iso(double c0, double c1, double c2, double c3) {
    auto side = [](double a0, double a2, double b1, double b3) {
        const double hi = fmax(a0, a2);
        double k = floor(fmin(a0, a2));
        if (k == hi) k -= 1.0;   // a line through the colour maximum cuts nothing
        return k > fmax(b1, b3);
    };
    return side(c0, c2, c1, c3) || side(c1, c3, c0, c2);
}

FIA_HD bool face_twisted(const double col[4][2]) {
    return alternating_iso(col[0][0], col[1][0], col[2][0], col[3][0]) ||
           alternating_iso(col[0][1], col[1][1], col[2][1], col[3][1]);
}

// True if a corner colour is exactly an integer strictly inside its colour's range.  The cut
// line of that integer then runs through the corner, and where an edge of constant colour lies on
// it the recursive cut decides which side gets the edge from the rounding of interpolated
// colours; such faces take the recursive cutter, as twisted ones do.
FIA_HD bool face_degenerate(const double col[4][2]) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const double lo = fmin(fmin(col[0][d], col[1][d]), fmin(col[2][d], col[3][d]));
        const double hi = fmax(fmax(col[0][d], col[1][d]), fmax(col[2][d], col[3][d]));
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double v = col[c][d];
            if (v == floor(v) && v > lo && v < hi) return true;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------------------------
// General cutter: the reference's recursion (src/facecutter.h:59-224) on an explicit stack, and
// subdivide (src/facecutter.h:255-324), in FP64.  Used for twisted faces and minimal_area < 1.
// Pieces come out in a different order than the reference's depth-first order; the set is the
// same.  Capacities: kGenVerts vertices per polygon (the reference's MAX_TMP_POINTS), kGenStack
// pending polygons, kGenTris pending triangles; exceeding one returns its KE_CAP_* code.  The
// pieces below a cut line (which hold no more interior integers along that axis) are cut to the
// end before the remainder above it, so the stack holds a few remainders however many lines
// cross the face.
// ---------------------------------------------------------------------------------------------
constexpr int kGenVerts = 32;
constexpr int kGenStack = 8;
constexpr int kGenTris = 40;

struct GenPoly {
    int n, dim;
    CutVert v[kGenVerts];
};
// copies a polygon's n live vertices only (the arrays live in private scratch memory, where a
// whole-struct copy would move all kGenVerts entries every time the stack is pushed or popped)
FIA_HD void gen_copy(GenPoly &d, const GenPoly &s) {
    d.n = s.n;
    d.dim = s.dim;
    for (int i = 0; i < s.n; ++i) d.v[i] = s.v[i];
}

// centroid and |area| of a xi-eta polygon (reference src/facecutter.h:226-253)
FIA_HD void gen_center(const double (*p)[2], int num, double &cx, double &cy, double &area) {
    double a2 = 0.0, mx = 0.0, my = 0.0;
    for (int n = 0; n < num; ++n) {
        const int m = (n + 1 == num) ? 0 : n + 1;
        const double a = p[n][0] * p[m][1] - p[m][0] * p[n][1];
        a2 += a;
        mx += a * (p[n][0] + p[m][0]);
        my += a * (p[n][1] + p[m][1]);
    }
    area = fabs(0.5 * a2);
    if (area != 0.0) {
        const double s = 1.0 / (6.0 * area);
        cx = mx * s; cy = my * s;
    } else {
        cx = cy = 0.0;
        for (int n = 0; n < num; ++n) { cx += p[n][0]; cy += p[n][1]; }
        cx /= num; cy /= num;
    }
}

// emit(pts, num, centre_xi, centre_eta) for every piece.  Returns KE_NONE, KE_CAP_VERTS,
// KE_CAP_STACK, KE_CAP_TRIS (a polygon, the stack or the triangle list overflowed) or KE_CUT (a
// cut line crossed a piece's boundary an odd number of times; the reference stops the program
// there, src/facecutter.h:190-203).
template <class Emit>
FIA_HD int general_cut(const double col[4][2], double min_area, Emit &emit) {
    GenPoly stack[kGenStack];
    int sp = 0;
    {
        GenPoly &q = stack[sp++];
        q.n = 4; q.dim = 0;
        const double sq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int i = 0; i < 4; ++i) { q.v[i].xi = sq[i][0]; q.v[i].eta = sq[i][1]; q.v[i].cx = col[i][0]; q.v[i].cz = col[i][1]; }
    }
    GenPoly cur, arr0, arr1;
    while (sp > 0) {
        gen_copy(cur, stack[--sp]);
        const int dim = cur.dim, num = cur.n;
        double lo = 1e300, hi = -1e300;
        for (int n = 0; n < num; ++n) {
            const double c = dim == 0 ? cur.v[n].cx : cur.v[n].cz;
            lo = fmin(lo, c); hi = fmax(hi, c);
        }
        int first, last;
        interior_integers(lo, hi, first, last);
        if (first <= last) {
            const double iso = first;
            // arr0 is the piece on v[0]'s side, closed last; when that side is above the cut its
            // stack slot is taken now, under the pieces below the cut
            int reserved = -1;
            if (!((dim == 0 ? cur.v[0].cx : cur.v[0].cz) - iso < 0.0)) {
                if (sp >= kGenStack) return KE_CAP_STACK;
                reserved = sp++;
            }
            arr0.n = 0; arr1.n = 0; arr0.dim = dim; arr1.dim = dim;
            int slot = 0;
            arr0.v[arr0.n++] = cur.v[0];
            for (int n = 0; n < num; ++n) {
                const int m = (n + 1 == num) ? 0 : n + 1;
                const double v0 = (dim == 0 ? cur.v[n].cx : cur.v[n].cz) - iso;
                const double v1 = (dim == 0 ? cur.v[m].cx : cur.v[m].cz) - iso;
                const bool neg0 = v0 < 0.0, neg1 = v1 < 0.0;
                if (neg0 != neg1) {
                    const double inv = 1.0 / (v1 - v0);
                    const double t0 = v1 * inv, t1 = -v0 * inv;
                    CutVert add;
                    add.xi = t0 * cur.v[n].xi + t1 * cur.v[m].xi;
                    add.eta = t0 * cur.v[n].eta + t1 * cur.v[m].eta;
                    if (dim == 0) { add.cx = iso; add.cz = t0 * cur.v[n].cz + t1 * cur.v[m].cz; }
                    else { add.cz = iso; add.cx = t0 * cur.v[n].cx + t1 * cur.v[m].cx; }
                    if (slot == 0) {
                        arr0.v[arr0.n++] = add;
                        slot = 1; arr1.n = 0;
                        if (t1 < 1.0) arr1.v[arr1.n++] = add;
                    } else {
                        arr1.v[arr1.n++] = add;
                        if (sp >= kGenStack) return KE_CAP_STACK;
                        gen_copy(stack[sp++], arr1);
                        slot = 0;
                        if (t1 < 1.0) arr0.v[arr0.n++] = add;
                    }
                }
                if (n < num - 1) {
                    GenPoly &dst = slot == 0 ? arr0 : arr1;
                    dst.v[dst.n++] = cur.v[m];
                }
                if (arr0.n >= kGenVerts - 2 || arr1.n >= kGenVerts - 2) return KE_CAP_VERTS;
            }
            if (slot != 0) return KE_CUT;   // odd number of sign changes: cannot happen on a loop
            if (reserved >= 0) {
                gen_copy(stack[reserved], arr0);
            } else {
                if (sp >= kGenStack) return KE_CAP_STACK;
                gen_copy(stack[sp++], arr0);
            }
        } else if (dim == 0) {
            cur.dim = 1;
            if (sp >= kGenStack) return KE_CAP_STACK;
            gen_copy(stack[sp++], cur);
        } else if (num >= 3) {
            // subdivide (reference src/facecutter.h:255-324)
            double pts[kGenVerts][2];
            for (int n = 0; n < num; ++n) { pts[n][0] = cur.v[n].xi; pts[n][1] = cur.v[n].eta; }
            double tri[kGenTris][3][2];
            int tp = 0;
            if (num == 3) {
                for (int k = 0; k < 3; ++k) { tri[0][k][0] = pts[k][0]; tri[0][k][1] = pts[k][1]; }
                tp = 1;
            } else {
                double cx, cy, area;
                gen_center(pts, num, cx, cy, area);
                if (area > min_area) {
                    for (int n = num - 1; n >= 0; --n) {   // pushed in reverse: popped in polygon order
                        const int m = (n + 1 == num) ? 0 : n + 1;
                        if (tp >= kGenTris) return KE_CAP_TRIS;
                        tri[tp][0][0] = cx; tri[tp][0][1] = cy;
                        tri[tp][1][0] = pts[n][0]; tri[tp][1][1] = pts[n][1];
                        tri[tp][2][0] = pts[m][0]; tri[tp][2][1] = pts[m][1];
                        ++tp;
                    }
                } else {
                    if (area != 0.0) emit(pts, num, cx, cy);
                    continue;
                }
            }
            while (tp > 0) {
                double t[3][2];
                --tp;
                for (int k = 0; k < 3; ++k) { t[k][0] = tri[tp][k][0]; t[k][1] = tri[tp][k][1]; }
                double cx, cy, area;
                gen_center(t, 3, cx, cy, area);
                if (area > min_area) {
                    int imax = 0;
                    double maxlen = -1.0;
                    for (int n = 0; n < 3; ++n) {
                        const int m = (n + 1) % 3;
                        const double dx = t[n][0] - t[m][0], dy = t[n][1] - t[m][1];
                        const double len = dx * dx + dy * dy;
                        if (maxlen < len) { maxlen = len; imax = n; }
                    }
                    const int i1 = (imax + 1) % 3, i2 = (imax + 2) % 3;
                    const double mx = 0.5 * (t[imax][0] + t[i1][0]), my = 0.5 * (t[imax][1] + t[i1][1]);
                    if (tp + 2 > kGenTris) return KE_CAP_TRIS;
                    // tri1 = {mid, p1, p2} pushed first so tri0 = {p0, mid, p2} is processed first
                    tri[tp][0][0] = mx; tri[tp][0][1] = my;
                    tri[tp][1][0] = t[i1][0]; tri[tp][1][1] = t[i1][1];
                    tri[tp][2][0] = t[i2][0]; tri[tp][2][1] = t[i2][1];
                    ++tp;
                    tri[tp][0][0] = t[imax][0]; tri[tp][0][1] = t[imax][1];
                    tri[tp][1][0] = mx; tri[tp][1][1] = my;
                    tri[tp][2][0] = t[i2][0]; tri[tp][2][1] = t[i2][1];
                    ++tp;
                } else if (area != 0.0) {
                    emit(t, 3, cx, cy);
                }
            }
        }
    }
    return KE_NONE;
}

}  // namespace fia
