/*
 * fia_oracle.cpp -- CPU restatement of the reference's 3D flux-interpolated advection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP path is checked against and
 * the CPU baseline that bench.py times.  Nothing in the product (libfia.so, the C++ shim, the
 * Python host module) links, loads or calls it.
 *
 * It keeps the reference's structure on purpose: every cell evaluates its own six faces
 * (interior faces are computed twice, once per neighbouring cell, with opposite orientation)
 * and every face backtraces its own four vertices, as in
 * src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:45-313.  The
 * reference's per-thread quadrature-point lists are replaced by direct accumulation into the
 * owning cell (identical arithmetic per point, no lists), parallelised over x-slabs of cells.
 *
 * Precision: template parameter R is the field/geometry type (double = FP64 oracle, float = the
 * reference's MY_FLOAT_TYPE), C is the face-cutter type (the reference cutter is always float,
 * src/facecutter.h:33-42).  Build with -ffp-contract=off so no FMA changes the rounding.
 *
 * Every function cites the reference file:line it restates.  Functions whose result no
 * reference test, golden vector or closed form pins are marked "parity unpinned".
 */
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------------------
// Index helpers: [x][y][z], z fastest (reference include/utils.h:92,116,140,164).
// ---------------------------------------------------------------------------------------------
static inline size_t idx3(int ix, int iy, int iz, int my, int mz) {
    return ((size_t)ix * my + iy) * mz + iz;
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------------------------
// Gauss-Legendre rules with weights normalised to sum 1 on [-1,1]
// (reference src/gauss_quadrature_points.cpp:18-299, getters :306-431).  Nodes are computed by
// Newton iteration on P_n; the reference stores the same rules as 16-digit literals.
// ---------------------------------------------------------------------------------------------
struct GaussRule { std::vector<double> x, w; };
static GaussRule gauss_legendre(int n) {
    GaussRule g; g.x.resize(n); g.w.resize(n);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; ++k) { double p2 = p1; p1 = p0; p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k; }
            double dp = n * (z * p0 - p1) / (z * z - 1.0);
            double dz = p0 / dp; z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = 0.0;
        for (int k = 1; k <= n; ++k) { double p2 = p1; p1 = p0; p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k; }
        double dp = n * (z * p0 - p1) / (z * z - 1.0);
        g.x[n - 1 - i] = z;                                   // ascending
        g.w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp) / 2.0; // sum 1
    }
    if (n == 1) { g.x[0] = 0.0; g.w[0] = 1.0; }
    return g;
}
static bool gauss_n_supported(int n) {
    return n == 1 || n == 3 || n == 5 || n == 7 || n == 9 || n == 11 || n == 13 || n == 25 || n == 49;
}

// ---------------------------------------------------------------------------------------------
// Staggered field: sample (i,j,k) sits at ((i+ox)h, (j+oy)h, (k+oz)h); dims (mx,my,mz).
// Cell centres have o = 1/2, the face axis of a face array has o = 0.
// ---------------------------------------------------------------------------------------------
template <class R> struct Field {
    const R *d; int mx, my, mz; double ox, oy, oz;
    R at(int i, int j, int k) const { return d[idx3(i, j, k, my, mz)]; }
};

// 1D WENO6 interpolation from six samples f[0..5] with signed distances d[i] = p - x_i
// (reference src/WENO6_interpolation_3d.cpp:63-123): Newton-form cubic candidates, ideal weights
// C0=d4 d5/20h^2, C1=-d5 d0/10h^2, C2=d0 d1/20h^2, smoothness indicators as integer quadratic
// forms /180, alpha = C/(1e-6+S)^2.
template <class R> static R weno1d(const R f[6], const R d[6], R h, bool opt = false) {
    R q[3];
    for (int i = 0; i < 3; ++i) {
        q[i] = f[i] + (f[i + 1] - f[i]) * d[i] / h
             + (f[i + 2] - 2 * f[i + 1] + f[i]) * d[i] * d[i + 1] / (2 * h * h)
             + (f[i + 3] - 3 * f[i + 2] + 3 * f[i + 1] - f[i]) * d[i] * d[i + 1] * d[i + 2] / (6 * h * h * h);
    }
    R c[3];
    c[0] = d[4] * d[5] / (20 * h * h);
    c[1] = -d[5] * d[0] / (10 * h * h);
    c[2] = d[0] * d[1] / (20 * h * h);
    R s[3];
    s[0] = (-3579 * f[3] * f[2] + 2634 * f[3] * f[1] - 683 * f[3] * f[0]
            - 6927 * f[2] * f[1] + 1854 * f[2] * f[0] - 1659 * f[1] * f[0]
            + 814 * f[3] * f[3] + 4326 * f[2] * f[2] + 2976 * f[1] * f[1]
            + 244 * f[0] * f[0]) / R(180.0);
    s[1] = (-3777 * f[3] * f[2] + 1074 * f[3] * f[1] - 1269 * f[2] * f[1]
            + 1986 * f[3] * f[3] + 1986 * f[2] * f[2] + 244 * f[1] * f[1]
            + 244 * f[4] * f[4] - 1269 * f[4] * f[3] + 1074 * f[4] * f[2]
            - 293 * f[4] * f[1]) / R(180.0);
    s[2] = (-3579 * f[3] * f[2] + 4326 * f[3] * f[3] + 814 * f[2] * f[2]
            + 2976 * f[4] * f[4] + 244 * f[5] * f[5] - 683 * f[5] * f[2]
            - 6927 * f[4] * f[3] + 2634 * f[4] * f[2] - 1659 * f[5] * f[4]
            + 1854 * f[5] * f[3]) / R(180.0);
    R a[3];
    const R eps = R(0.000001);
    // "WENO6-optimized" (reference src/weno.ispc:46-50, its rcp taken as exact):
    // alpha = C / (1e-30 + S^2)
    if (opt) for (int i = 0; i < 3; ++i) a[i] = c[i] / (R(1e-30) + s[i] * s[i]);
    else for (int i = 0; i < 3; ++i) a[i] = c[i] / ((eps + s[i]) * (eps + s[i]));
    R sum = a[0] + a[1] + a[2];
    R w0 = a[0] / sum, w1 = a[1] / sum, w2 = a[2] / sum;
    return w0 * q[0] + w1 * q[1] + w2 * q[2];
}

// 3D WENO6 on a staggered field, passes z then y then x, clamp-to-edge indexing
// (reference src/WENO6_interpolation_3d.cpp:7-258 cell centre; :260-1020 face x/y/z variants,
// which differ only in the per-axis offset o and the clamp bound).
template <class R> static R weno3d(const Field<R> &F, R px, R py, R pz, R h) {
    const int bx = (int)std::floor((px - R(F.ox) * h) / h);
    const int by = (int)std::floor((py - R(F.oy) * h) / h);
    const int bz = (int)std::floor((pz - R(F.oz) * h) / h);
    R dz[6], dy[6], dxv[6];
    for (int i = 0; i < 6; ++i) {
        dz[i] = pz - R(bz - 2 + i + F.oz) * h;
        dy[i] = py - R(by - 2 + i + F.oy) * h;
        dxv[i] = px - R(bx - 2 + i + F.ox) * h;
    }
    R zv[6][6];
    for (int ix = 0; ix < 6; ++ix) {
        const int jx = clampi(bx - 2 + ix, 0, F.mx - 1);
        for (int iy = 0; iy < 6; ++iy) {
            const int jy = clampi(by - 2 + iy, 0, F.my - 1);
            R v[6];
            for (int iz = 0; iz < 6; ++iz) v[iz] = F.at(jx, jy, clampi(bz - 2 + iz, 0, F.mz - 1));
            zv[ix][iy] = weno1d<R>(v, dz, h);
        }
    }
    R yv[6];
    for (int ix = 0; ix < 6; ++ix) yv[ix] = weno1d<R>(zv[ix], dy, h);
    return weno1d<R>(yv, dxv, h);
}

// "WENO6-optimized": the reference's ISPC interpolation of quadrature points
// (src/sginterp3.cpp:203-383 with presort, src/sginterp3.ispc:79-115).  Per axis the coordinate in
// sample units is held to [0, m - 1], the base cell to [0, m - 2] (src/sginterp3.cpp:254-258) and
// the fraction to [0, 1] (:299-301); samples base-2 .. base+3 clamp to the field; passes run along
// the fastest axis first (z, then y, then x, as the reference hands it z as its x); the 1D rule
// is weno1d with alpha = C / (1e-30 + S^2).
template <class R> static void clamped_split(R s, int m, int &b, R &t) {
    const R c = std::min(std::max(s, R(0)), R(m - 1));
    const int f = (int)std::floor(c);
    b = f < m - 2 ? f : (m >= 2 ? m - 2 : 0);
    t = c - R(b);
}
template <class R> static R weno3d_opt(const Field<R> &F, R px, R py, R pz, R h) {
    int bx, by, bz;
    R tx, ty, tz;
    clamped_split<R>(px / h - R(F.ox), F.mx, bx, tx);
    clamped_split<R>(py / h - R(F.oy), F.my, by, ty);
    clamped_split<R>(pz / h - R(F.oz), F.mz, bz, tz);
    R dz[6], dy[6], dxv[6];
    for (int i = 0; i < 6; ++i) {
        dz[i] = (tz - R(i - 2)) * h;
        dy[i] = (ty - R(i - 2)) * h;
        dxv[i] = (tx - R(i - 2)) * h;
    }
    R zv[6][6];
    for (int ix = 0; ix < 6; ++ix) {
        const int jx = clampi(bx - 2 + ix, 0, F.mx - 1);
        for (int iy = 0; iy < 6; ++iy) {
            const int jy = clampi(by - 2 + iy, 0, F.my - 1);
            R v[6];
            for (int iz = 0; iz < 6; ++iz) v[iz] = F.at(jx, jy, clampi(bz - 2 + iz, 0, F.mz - 1));
            zv[ix][iy] = weno1d<R>(v, dz, h, true);
        }
    }
    R yv[6];
    for (int ix = 0; ix < 6; ++ix) yv[ix] = weno1d<R>(zv[ix], dy, h, true);
    return weno1d<R>(yv, dxv, h, true);
}

// WENO6 along y inside the column floor(x/h), floor(z/h)
// (reference src/WENO6_interpolation_1d.cpp:315-779).
template <class R> static R weno1dy(const Field<R> &F, R px, R py, R pz, R h) {
    const int cx = clampi((int)std::floor(px / h), 0, F.mx - 1);
    const int cz = clampi((int)std::floor(pz / h), 0, F.mz - 1);
    const int by = (int)std::floor((py - R(F.oy) * h) / h);
    R v[6], d[6];
    for (int i = 0; i < 6; ++i) {
        v[i] = F.at(cx, clampi(by - 2 + i, 0, F.my - 1), cz);
        d[i] = py - R(by - 2 + i + F.oy) * h;
    }
    return weno1d<R>(v, d, h);
}

// Trilinear, clamp-to-edge, passes z, y, x (reference src/linear_interpolation_3d.cpp:7-95).
template <class R> static R linear3d(const Field<R> &F, R px, R py, R pz, R h) {
    const R sx = (px - R(F.ox) * h) / h, sy = (py - R(F.oy) * h) / h, sz = (pz - R(F.oz) * h) / h;
    const int bx = (int)std::floor(sx), by = (int)std::floor(sy), bz = (int)std::floor(sz);
    const R c0 = sz - bz, c1 = R(1) - c0;
    const R b0 = sy - by, b1 = R(1) - b0;
    const R a0 = sx - bx, a1 = R(1) - a0;
    R yz[2];
    for (int ix = 0; ix < 2; ++ix) {
        const int jx = clampi(bx + ix, 0, F.mx - 1);
        R zz[2];
        for (int iy = 0; iy < 2; ++iy) {
            const int jy = clampi(by + iy, 0, F.my - 1);
            const R v0 = F.at(jx, jy, clampi(bz, 0, F.mz - 1));
            const R v1 = F.at(jx, jy, clampi(bz + 1, 0, F.mz - 1));
            zz[iy] = c1 * v0 + c0 * v1;
        }
        yz[ix] = b1 * zz[0] + b0 * zz[1];
    }
    return a1 * yz[0] + a0 * yz[1];
}

// Linear in y inside the column that contains the point
// (reference src/linear_interpolation_1d.cpp:314-374: x index floor((x-(ox-1/2)h)/h)).
template <class R> static R linear1dy(const Field<R> &F, R px, R py, R pz, R h) {
    const int cx = clampi((int)std::floor((px - (R(F.ox) - R(0.5)) * h) / h), 0, F.mx - 1);
    const int cz = clampi((int)std::floor((pz - (R(F.oz) - R(0.5)) * h) / h), 0, F.mz - 1);
    const R sy = (py - R(F.oy) * h) / h;
    const int by = (int)std::floor(sy);
    const R b0 = sy - by, b1 = R(1) - b0;
    const R v0 = F.at(cx, clampi(by, 0, F.my - 1), cz);
    const R v1 = F.at(cx, clampi(by + 1, 0, F.my - 1), cz);
    return b1 * v0 + b0 * v1;
}

enum Interp { I_LINEAR = 0, I_WENO6 = 1, I_1DY_LINEAR = 2, I_1DY_WENO6 = 3, I_WENO6_OPT = 4 };
enum PsiMode { P_CONST = 0, P_LINEAR = 1, P_WENO6 = 2 };
enum Split { S_NONE = 0, S_XZ_CELL = 1, S_XZ_AXIS = 2, S_XCELL_ZAXIS = 3, S_XAXIS_ZCELL = 4 };

template <class R> static R interp(const Field<R> &F, int method, R px, R py, R pz, R h) {
    switch (method) {
        case I_LINEAR: return linear3d(F, px, py, pz, h);
        case I_WENO6: return weno3d(F, px, py, pz, h);
        case I_1DY_LINEAR: return linear1dy(F, px, py, pz, h);
        case I_1DY_WENO6: return weno1dy(F, px, py, pz, h);
        case I_WENO6_OPT: return weno3d_opt(F, px, py, pz, h);
    }
    return R(0);
}

// MAC velocity (reference src/grid_3d.cpp:545-573).
template <class R> struct Velocity {
    Field<R> u, v, w;
    void at(int method, R px, R py, R pz, R h, R out[3]) const {
        out[0] = interp(u, method, px, py, pz, h);
        out[1] = interp(v, method, px, py, pz, h);
        out[2] = interp(w, method, px, py, pz, h);
    }
};

// ---------------------------------------------------------------------------------------------
// Component grid G: density (comp = -1) or one MAC velocity component (comp = 0,1,2).
// G has dims M = N + e_comp; its vertex lattice is shifted by -h/2 along comp
// (reference src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:30-43,
// 343-350, 391-405).
// ---------------------------------------------------------------------------------------------
struct Comp {
    int comp; int nx, ny, nz; int e[3]; int m[3];
    Comp(int c, int nx_, int ny_, int nz_) : comp(c), nx(nx_), ny(ny_), nz(nz_) {
        for (int a = 0; a < 3; ++a) e[a] = (a == c);
        m[0] = nx + e[0]; m[1] = ny + e[1]; m[2] = nz + e[2];
    }
    double origin(int a) const { return -0.5 * e[a]; }   // in cells
};

// ---------------------------------------------------------------------------------------------
// psi on the y-faces of G (reference src/calc_psi_3D.cpp:10-308 for density,
// :311-537 for velocity).  WENO6 for a velocity component is not present in the reference (the
// branch is commented out at calc_psi_3D.cpp:436-518); here it is the density rule applied on G
// with the column index taken directly (parity unpinned for comp >= 0 with WENO6).
// ---------------------------------------------------------------------------------------------
template <class R>
static void calc_psi(const Comp &G, const R *q, R h, int mode, int use_integral, int ng, R *psi) {
    const int mx = G.m[0], my = G.m[1], mz = G.m[2];
    const Field<R> Q{q, mx, my, mz, 0.5 - 0.5 * G.e[0], 0.5 - 0.5 * G.e[1], 0.5 - 0.5 * G.e[2]};
    GaussRule gr = gauss_legendre(ng);
    for (int ix = 0; ix < mx; ++ix)
        for (int iz = 0; iz < mz; ++iz) psi[idx3(ix, 0, iz, my + 1, mz)] = R(0);
    const bool integ = use_integral != 0;
    for (int iy = 1; iy < my + 1; ++iy) {
        for (int ix = 0; ix < mx; ++ix) {
            for (int iz = 0; iz < mz; ++iz) {
                const R prev = psi[idx3(ix, iy - 1, iz, my + 1, mz)];
                R val;
                if (!integ || mode == P_CONST) {
                    val = prev + q[idx3(ix, iy - 1, iz, my, mz)] * h;
                } else if (mode == P_LINEAR) {
                    // reference calc_psi_3D.cpp:111-130 / :410-431
                    const R xc = R(ix + 0.5 - 0.5 * G.e[0]) * h, zc = R(iz + 0.5 - 0.5 * G.e[2]) * h;
                    const R y0 = R((iy - 1) + 0.25 - 0.5 * G.e[1]) * h;
                    const R y1 = R((iy - 1) + 0.75 - 0.5 * G.e[1]) * h;
                    const R i0 = linear3d(Q, xc, y0, zc, h), i1 = linear3d(Q, xc, y1, zc, h);
                    val = prev + i0 * (h / R(2)) + i1 * (h / R(2));
                } else {
                    // reference calc_psi_3D.cpp:192-262
                    val = prev;
                    const R oy = R(-0.5 * G.e[1]) * h;
                    const R xc = R(ix + 0.5 - 0.5 * G.e[0]) * h, zc = R(iz + 0.5 - 0.5 * G.e[2]) * h;
                    for (int half = 0; half < 2; ++half) {
                        const R a = R(iy - 1.0 + 0.5 * half) * h + oy;
                        const R b = R(iy - 0.5 + 0.5 * half) * h + oy;
                        const R len = b - a;
                        for (int g = 0; g < ng; ++g) {
                            const R yg = (len / R(2)) * R(gr.x[g]) + (len / R(2)) + a;
                            R rho;
                            if (G.comp < 0) {
                                rho = weno1dy(Q, xc, yg, zc, h);
                            } else {
                                // column taken directly (see header comment)
                                const int by = (int)std::floor((yg - R(Q.oy) * h) / h);
                                R v[6], d[6];
                                for (int i = 0; i < 6; ++i) {
                                    v[i] = Q.at(ix, clampi(by - 2 + i, 0, my - 1), iz);
                                    d[i] = yg - R(by - 2 + i + Q.oy) * h;
                                }
                                rho = weno1d<R>(v, d, h);
                            }
                            val += rho * (h / R(2)) * R(gr.w[g]);
                        }
                    }
                }
                psi[idx3(ix, iy, iz, my + 1, mz)] = val;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Face cutter (reference src/facecutter.h:33-324), templated on its arithmetic type C.
// ---------------------------------------------------------------------------------------------
template <class C> struct P2 { C x[2]; };
template <class C> struct V2 { P2<C> p, c; };
static const int kMaxTmp = 32;   // MAX_TMP_POINTS (reference src/facecutter.h:29)

template <class C> struct Cutter {
    bool overflow = false;
    std::function<void(const V2<C> *, int)> out;

    static C sgn(C v) { return v < C(0) ? C(-1) : C(1); }

    // reference src/facecutter.h:59-224 (the recursive, non-STACKFREE variant)
    void cut(const V2<C> *poly, int num, int dim) {
        C lo = std::numeric_limits<C>::max(), hi = std::numeric_limits<C>::lowest();
        for (int n = 0; n < num; ++n) { lo = std::min(lo, poly[n].c.x[dim]); hi = std::max(hi, poly[n].c.x[dim]); }
        int first = (int)std::ceil(lo), last = (int)std::floor(hi);
        if (lo - first == C(0)) first += 1;
        if (hi - last == C(0)) last -= 1;
        bool produced = false;
        if (first <= last) {
            const C iso = C(first);
            V2<C> arr[2][kMaxTmp + 4];
            int head[2] = {0, 0};
            int slot = 0;
            arr[0][head[0]++] = poly[0];
            for (int n = 0; n < num; ++n) {
                const int m = (n + 1) % num;
                const C v0 = poly[n].c.x[dim] - iso, v1 = poly[m].c.x[dim] - iso;
                if ((v0 != C(0) || v1 != C(0)) && sgn(v0) * sgn(v1) < C(0)) {
                    // 1.0 is a double literal in the reference: divide in double, round to C
                    const C inv = C(1.0 / (double)(v1 - v0));
                    const C t0 = v1 * inv, t1 = -v0 * inv;
                    produced = true;
                    V2<C> add;
                    add.p.x[0] = t0 * poly[n].p.x[0] + t1 * poly[m].p.x[0];
                    add.p.x[1] = t0 * poly[n].p.x[1] + t1 * poly[m].p.x[1];
                    add.c.x[dim] = iso;
                    add.c.x[1 - dim] = t0 * poly[n].c.x[1 - dim] + t1 * poly[m].c.x[1 - dim];
                    if (slot == 0) {
                        arr[0][head[0]++] = add;
                        slot = 1; head[1] = 0;
                        if (t1 < C(1)) arr[1][head[1]++] = add;
                    } else {
                        arr[1][head[1]++] = add;
                        if (!check(head)) return;
                        cut(arr[1], head[1], dim);
                        slot = 0;
                        if (t1 < C(1)) arr[0][head[0]++] = add;
                    }
                }
                if (n < num - 1) arr[slot][head[slot]++] = poly[m];
                if (!check(head)) return;
            }
            if (produced) {
                if (slot != 0) { overflow = true; return; }   // reference exits (facecutter.h:191-204)
                cut(arr[0], head[0], dim);
            }
        }
        if (!produced) {
            if (dim == 0) cut(poly, num, 1);
            else out(poly, num);
        }
    }
    bool check(const int head[2]) {
        if (head[0] >= kMaxTmp || head[1] >= kMaxTmp) { overflow = true; return false; }
        return true;
    }
};

// reference src/facecutter.h:226-253
template <class C> static void center_of_gravity(const P2<C> *p, int num, P2<C> &c, C &area) {
    c.x[0] = c.x[1] = C(0); area = C(0);
    for (int n = 0; n < num; ++n) {
        const P2<C> &p0 = p[n], &p1 = p[(n + 1) % num];
        const C a = p0.x[0] * p1.x[1] - p1.x[0] * p0.x[1];
        area += a;
        c.x[0] += a * (p0.x[0] + p1.x[0]);
        c.x[1] += a * (p0.x[1] + p1.x[1]);
    }
    area = std::abs(C(0.5) * area);
    if (area != C(0)) {
        const C s = C(1.0 / (6.0 * (double)area));   // double literals as in the reference
        c.x[0] = c.x[0] * s; c.x[1] = c.x[1] * s;
    } else {
        c.x[0] = c.x[1] = C(0);
        for (int n = 0; n < num; ++n) { c.x[0] = c.x[0] + p[n].x[0]; c.x[1] = c.x[1] + p[n].x[1]; }
        const C s = C(1.0 / (double)num);
        c.x[0] = c.x[0] * s; c.x[1] = c.x[1] * s;
    }
}

// reference src/facecutter.h:255-324
template <class C>
static void subdivide(const P2<C> *poly, int num, C min_area,
                      const std::function<void(const P2<C> *, int, const P2<C> &, C)> &emit) {
    std::function<void(const P2<C> *)> tri_split = [&](const P2<C> *pt) {
        P2<C> c; C area;
        center_of_gravity(pt, 3, c, area);
        if (area > min_area) {
            int imax = 0;
            C maxlen = std::numeric_limits<C>::lowest();
            for (int n = 0; n < 3; ++n) {
                const P2<C> &a = pt[n], &b = pt[(n + 1) % 3];
                const C len = (a.x[0] - b.x[0]) * (a.x[0] - b.x[0]) + (a.x[1] - b.x[1]) * (a.x[1] - b.x[1]);
                if (maxlen < len) { maxlen = len; imax = n; }
            }
            const P2<C> &p0 = pt[imax], &p1 = pt[(imax + 1) % 3], &p2 = pt[(imax + 2) % 3];
            P2<C> mid; mid.x[0] = C(0.5) * (p0.x[0] + p1.x[0]); mid.x[1] = C(0.5) * (p0.x[1] + p1.x[1]);
            const P2<C> t0[3] = {p0, mid, p2};
            const P2<C> t1[3] = {mid, p1, p2};
            tri_split(t0);
            tri_split(t1);
        } else if (area != C(0)) {
            emit(pt, 3, c, area);
        }
    };
    if (num == 3) {
        tri_split(poly);
    } else {
        P2<C> c; C area;
        center_of_gravity(poly, num, c, area);
        if (area > min_area) {
            for (int n = 0; n < num; ++n) {
                const P2<C> t[3] = {c, poly[n], poly[(n + 1) % num]};
                tri_split(t);
            }
        } else if (area != C(0)) {
            emit(poly, num, c, area);
        }
    }
}

// Cut-and-subdivide of one face given its four corner colours; emits xi-eta pieces.
// (reference src/split_face_3D.cpp:17-65)
template <class C>
static bool split_xieta(const C colour[4][2], C min_area,
                        const std::function<void(const P2<C> *, int, const P2<C> &, C)> &emit) {
    V2<C> poly[4];
    const C sq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int i = 0; i < 4; ++i) {
        poly[i].p.x[0] = sq[i][0]; poly[i].p.x[1] = sq[i][1];
        poly[i].c.x[0] = colour[i][0]; poly[i].c.x[1] = colour[i][1];
    }
    Cutter<C> cu;
    bool bad = false;
    cu.out = [&](const V2<C> *p, int num) {
        if (num >= 3) {
            if (num > kMaxTmp) { bad = true; return; }
            P2<C> tmp[kMaxTmp];
            for (int i = 0; i < num; ++i) tmp[i] = p[i].p;
            subdivide<C>(tmp, num, min_area, emit);
        }
    };
    cu.cut(poly, 4, 0);
    return !(cu.overflow || bad);
}

// Corner colours of a face for a split method (reference src/split_face_3D.cpp:26-43).
template <class C, class R>
static void face_colours(const R vert[4][3], R h, int split, C col[4][2]) {
    for (int i = 0; i < 4; ++i) {
        col[i][0] = col[i][1] = C(0);
        const double x = (double)(vert[i][0] / h), z = (double)(vert[i][2] / h);
        switch (split) {
            case S_XZ_CELL: col[i][0] = C(x - 0.5); col[i][1] = C(z - 0.5); break;
            case S_XZ_AXIS: col[i][0] = C(x); col[i][1] = C(z); break;
            case S_XCELL_ZAXIS: col[i][0] = C(x - 0.5); col[i][1] = C(z); break;
            case S_XAXIS_ZCELL: col[i][0] = C(x); col[i][1] = C(z - 0.5); break;
            default: break;
        }
    }
}

// Bilinear xi-eta -> real-space map (reference src/cell_face_3D.cpp:68-73).  The reference
// writes it as (1-xi)(1-eta) v0 + xi(1-eta) v1 + xi eta v2 + (1-xi) eta v3; here it is the same
// polynomial written around corner 0, v0 + xi (v1-v0) + eta (v3-v0) + xi eta (v0-v1+v2-v3), so
// that a coordinate shared by all four corners comes out exactly.  That matters on faces that
// stay on a y-level (every face on the top and bottom walls): the integral-psi rule switches
// branch at floor(c_y) with a jump of the order of the interpolation error, and the reference's
// form can round c_y to either side of the level.
template <class R> static void bilinear(const R v[4][3], R xi, R eta, R out[3]) {
    for (int k = 0; k < 3; ++k) {
        const R e1 = v[1][k] - v[0][k], e3 = v[3][k] - v[0][k], e13 = v[0][k] - v[1][k] + v[2][k] - v[3][k];
        out[k] = v[0][k] + (xi * e1 + eta * e3 + xi * eta * e13);
    }
}

// Piece centre and area-scaled normal (reference src/split_face_3D.cpp:67-93).
template <class R, class C>
static void piece_geometry(const R v[4][3], const P2<C> *poly, int num, const P2<C> &cen, R pos[3], R an[3]) {
    bilinear<R>(v, R(cen.x[0]), R(cen.x[1]), pos);
    R tp[kMaxTmp][3];
    for (int n = 0; n < num; ++n) bilinear<R>(v, R(poly[n].x[0]), R(poly[n].x[1]), tp[n]);
    an[0] = an[1] = an[2] = R(0);
    auto cross_add = [&](const R *o, const R *p1, const R *p2, R s) {
        const R a0 = p1[0] - o[0], a1 = p1[1] - o[1], a2 = p1[2] - o[2];
        const R b0 = p2[0] - o[0], b1 = p2[1] - o[1], b2 = p2[2] - o[2];
        an[0] += s * (a1 * b2 - a2 * b1);
        an[1] += s * (a2 * b0 - a0 * b2);
        an[2] += s * (a0 * b1 - a1 * b0);
    };
    if (num == 3) {
        cross_add(tp[0], tp[1], tp[2], R(0.5));
    } else if (num > 3) {
        R acc[3] = {0, 0, 0};
        for (int n = 0; n < num; ++n) {
            const R *p1 = tp[n], *p2 = tp[(n + 1) % num];
            const R a0 = p1[0] - pos[0], a1 = p1[1] - pos[1], a2 = p1[2] - pos[2];
            const R b0 = p2[0] - pos[0], b1 = p2[1] - pos[1], b2 = p2[2] - pos[2];
            acc[0] += a1 * b2 - a2 * b1; acc[1] += a2 * b0 - a0 * b2; acc[2] += a0 * b1 - a1 * b0;
        }
        an[0] = acc[0] * R(0.5); an[1] = acc[1] * R(0.5); an[2] = acc[2] * R(0.5);
    }
}

// Vertex order of the six faces of a cell (reference
// src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:77-112):
// 0 = z-, 1 = x+, 2 = y+, 3 = x-, 4 = y-, 5 = z+.
static const int kFaceVerts[6][4][3] = {
    {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}},
    {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}},
    {{0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}},
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},
    {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}},
    {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},
};

// Backtrace of one lattice vertex of G and clamp to the domain
// (reference src/calc_backtraced_face_3D.cpp:9-128 forward direction, :131-159 clamp).
template <class R>
static void backtrace_vertex(const Comp &G, const Velocity<R> &vel, int method, R h, R dt,
                             int vi, int vj, int vk, bool zero_velocity, R out[3]) {
    const int vid[3] = {vi, vj, vk};
    for (int a = 0; a < 3; ++a) out[a] = R(vid[a]) * h;
    for (int a = 0; a < 3; ++a) if (G.e[a]) out[a] -= R(0.5) * h;
    if (zero_velocity) return;
    bool pinned = false;
    if (G.comp >= 0) {
        const int n = (G.comp == 0 ? G.nx : (G.comp == 1 ? G.ny : G.nz));
        pinned = vid[G.comp] <= 1 || vid[G.comp] >= n;
    }
    R u[3];
    vel.at(method, out[0], out[1], out[2], h, u);
    if (!pinned) for (int a = 0; a < 3; ++a) out[a] -= u[a] * dt;
    const int n3[3] = {G.nx, G.ny, G.nz};
    for (int a = 0; a < 3; ++a) {
        const R lo = R(0) - R(G.e[a]) * R(0.5) * h;
        const R hi = R(0) + h * R(n3[a]) + R(G.e[a]) * R(0.5) * h;
        if (out[a] < lo) out[a] = lo;
        if (out[a] > hi) out[a] = hi;
    }
}

struct PassOpts {
    int comp; int split; int interp; int psi_mode; int zero_velocity; int use_integral;
    int ng; double min_area;
};

// One flux pass over G: psi, all faces of all cells, accumulation (reference
// src/move_substances_3d.cpp:585-841 and src/advect_velocity_flux_advection_3d.cpp:252-576,
// producers src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:7-597,
// src/backtrace_and_calc_all_quadrature_point_list.cpp:7-280, quadrature
// src/polygon_3D.cpp:57-77,257-487, consumers
// src/calc_density_in_cell_from_quadrature_point_list_for_integrate_density.cpp:119-201,
// src/calc_density_in_cell_from_quadrature_point_list.cpp:98-132,
// src/calc_velocity_from_quadrature_point_list_use_integral.cpp:12-181,
// src/calc_velocity_from_quadrature_point_list.cpp:8-129).
// The velocity field used for backtracing is vel; q (G-sized) is advected in place.
// Returns 0 on success, 3 on cutter capacity overflow.
// Everything one pass needs to integrate one face: the fields, psi and the options.
template <class R, class C> struct FaceIntegrator {
    const Comp &G;
    const Velocity<R> &vel;
    R h, dt;
    const PassOpts &o;
    Field<R> Q, PS;
    GaussRule gr;
    R oy, h3;

    FaceIntegrator(const Comp &g, const Velocity<R> &v, R h_, R dt_, const R *q, const R *psi, const PassOpts &opt)
        : G(g), vel(v), h(h_), dt(dt_), o(opt),
          Q{q, g.m[0], g.m[1], g.m[2], 0.5 - 0.5 * g.e[0], 0.5 - 0.5 * g.e[1], 0.5 - 0.5 * g.e[2]},
          PS{psi, g.m[0], g.m[1] + 1, g.m[2], 0.5 - 0.5 * g.e[0], -0.5 * g.e[1], 0.5 - 0.5 * g.e[2]},
          gr(gauss_legendre(opt.ng)), oy(R(g.origin(1)) * h_), h3(h_ * h_ * h_) {}

    // Contribution of the face with corner lattice indices lat[4] (reference vertex order) to the
    // cell that owns it, split into the density-point and psi-point parts.  Returns false on
    // cutter overflow.
    bool face(const int lat[4][3], R &acc_density, R &acc_psi, long &pieces) const {
        R vert[4][3];
        for (int k = 0; k < 4; ++k)
            // "WENO6-optimized" backtraces with WENO6 (reference src/grid_3d.cpp:557)
            backtrace_vertex<R>(G, vel, o.interp == I_WENO6_OPT ? I_WENO6 : o.interp, h, dt, lat[k][0], lat[k][1],
                                lat[k][2], o.zero_velocity != 0,
                                vert[k]);
        C col[4][2];
        face_colours<C, R>(vert, h, o.split, col);
        auto emit = [&](const P2<C> *poly, int num, const P2<C> &cen, C) {
            ++pieces;
            R pos[3], an[3];
            piece_geometry<R, C>(vert, poly, num, cen, pos, an);
            if (o.use_integral) {
                // reference src/polygon_3D.cpp:275-478
                const int j = (int)std::floor((pos[1] - oy) / h);
                const R over = ((pos[1] - oy) / h) - R(j);
                if (over < R(0.5)) {
                    const R a = R(j) * h + oy, b = pos[1], len = b - a;
                    for (int g = 0; g < o.ng; ++g) {
                        const R yg = (len / R(2)) * R(gr.x[g]) + (len / R(2)) + a;
                        const R wy = an[1] * over * h * R(gr.w[g]);
                        acc_density += interp(Q, o.interp, pos[0], yg, pos[2], h) * wy / h3;
                    }
                } else {
                    for (int half = 0; half < 2; ++half) {
                        const R a = half == 0 ? R(j) * h + oy : (R(j) + R(0.5)) * h + oy;
                        const R b = half == 0 ? (R(j) + R(0.5)) * h + oy : (R(j) + over) * h + oy;
                        const R frac = half == 0 ? R(0.5) : over - R(0.5);
                        const R len = b - a;
                        for (int g = 0; g < o.ng; ++g) {
                            const R yg = (len / R(2)) * R(gr.x[g]) + (len / R(2)) + a;
                            const R wy = an[1] * frac * h * R(gr.w[g]);
                            acc_density += interp(Q, o.interp, pos[0], yg, pos[2], h) * wy / h3;
                        }
                    }
                }
                const R py = R(j) * h + oy;
                acc_psi += interp(PS, o.interp, pos[0], py, pos[2], h) * an[1] / h3;
            } else {
                acc_psi += interp(PS, o.interp, pos[0], pos[1], pos[2], h) * an[1] / h3;
            }
        };
        return split_xieta<C>(col, C(o.min_area), emit);
    }
};

// One flux pass over G: psi, all faces of all cells, accumulation (reference
// src/move_substances_3d.cpp:585-841 and src/advect_velocity_flux_advection_3d.cpp:252-576,
// producers src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:7-597,
// src/backtrace_and_calc_all_quadrature_point_list.cpp:7-280, quadrature
// src/polygon_3D.cpp:57-77,257-487, consumers
// src/calc_density_in_cell_from_quadrature_point_list_for_integrate_density.cpp:119-201,
// src/calc_density_in_cell_from_quadrature_point_list.cpp:98-132,
// src/calc_velocity_from_quadrature_point_list_use_integral.cpp:12-181,
// src/calc_velocity_from_quadrature_point_list.cpp:8-129).
// The velocity field used for backtracing is vel; q (G-sized) is advected in place.
// Returns 0 on success, 3 on cutter capacity overflow.
template <class R, class C>
static int advect_pass(const Comp &G, const Velocity<R> &vel, R h, R dt, R *q, const PassOpts &o,
                       int nthreads, double *stats) {
    const int mx = G.m[0], my = G.m[1], mz = G.m[2];
    std::vector<R> psi((size_t)mx * (my + 1) * mz);
    calc_psi<R>(G, q, h, o.psi_mode, o.use_integral, o.ng, psi.data());
    const FaceIntegrator<R, C> fi(G, vel, h, dt, q, psi.data(), o);
    std::vector<R> result((size_t)mx * my * mz, R(0));
    std::atomic<int> fail{0};
    std::atomic<long> npieces{0};

    auto work = [&](int x0, int x1) {
        long local_pieces = 0;
        for (int ix = x0; ix < x1; ++ix)
        for (int iy = 0; iy < my; ++iy)
        for (int iz = 0; iz < mz; ++iz) {
            R acc_density = R(0), acc_psi = R(0);
            for (int f = 0; f < 6; ++f) {
                int lat[4][3];
                for (int k = 0; k < 4; ++k) {
                    lat[k][0] = ix + kFaceVerts[f][k][0]; lat[k][1] = iy + kFaceVerts[f][k][1];
                    lat[k][2] = iz + kFaceVerts[f][k][2];
                }
                if (!fi.face(lat, acc_density, acc_psi, local_pieces)) { fail = 3; return; }
            }
            result[idx3(ix, iy, iz, my, mz)] = acc_density + acc_psi;
        }
        npieces += local_pieces;
    };
    int nt = std::max(1, std::min(nthreads, mx));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const int x0 = (int)((long)mx * t / nt), x1 = (int)((long)mx * (t + 1) / nt);
        th.emplace_back(work, x0, x1);
    }
    for (auto &t : th) t.join();
    if (fail) return fail;
    std::copy(result.begin(), result.end(), q);
    if (stats) stats[0] += (double)npieces.load();
    return 0;
}

// Flux through every unique face of G in the orientation of the reference's faces 1 (x+),
// 2 (y+) and 5 (z+) of the cell below it, normalised to units of q (the contribution to that
// cell times h^3 / h^3): fx ((mx+1) my mz), fy (mx (my+1) mz), fz (mx my (mz+1)).
template <class R, class C>
static int face_fluxes(const Comp &G, const Velocity<R> &vel, R h, R dt, const R *q, const PassOpts &o,
                       double *fx, double *fy, double *fz) {
    const int mx = G.m[0], my = G.m[1], mz = G.m[2];
    std::vector<R> psi((size_t)mx * (my + 1) * mz);
    calc_psi<R>(G, q, h, o.psi_mode, o.use_integral, o.ng, psi.data());
    const FaceIntegrator<R, C> fi(G, vel, h, dt, q, psi.data(), o);
    double *out[3] = {fx, fy, fz};
    const int face_type[3] = {1, 2, 4 + 1};
    long pieces = 0;
    for (int fam = 0; fam < 3; ++fam) {
        const int dx = mx + (fam == 0), dy = my + (fam == 1), dz = mz + (fam == 2);
        for (int i = 0; i < dx; ++i) for (int j = 0; j < dy; ++j) for (int k = 0; k < dz; ++k) {
            // the '+' face of cell (i,j,k) - e_fam: shift the reference table back by one along fam
            int lat[4][3];
            const int ft = face_type[fam];
            for (int c = 0; c < 4; ++c) {
                lat[c][0] = i + kFaceVerts[ft][c][0] - (fam == 0);
                lat[c][1] = j + kFaceVerts[ft][c][1] - (fam == 1);
                lat[c][2] = k + kFaceVerts[ft][c][2] - (fam == 2);
            }
            R a = R(0), b = R(0);
            if (!fi.face(lat, a, b, pieces)) return 3;
            out[fam][idx3(i, j, k, dy, dz)] = (double)((a + b) * h * h * h / (h * h * h));
        }
    }
    return 0;
}

// Split used by each velocity component (reference
// src/advect_velocity_flux_advection_3d.cpp:622-628, :804-810, :997-1003) and by its error pass
// (:721, :913, :1096).  Density error pass uses xz-AxisAligned (src/move_substances_3d.cpp:189).
static int velocity_split(int comp, bool one_dy, bool error_pass) {
    const bool flip = one_dy != error_pass;
    if (comp == 0) return flip ? S_XCELL_ZAXIS : S_XAXIS_ZCELL;
    if (comp == 1) return flip ? S_XZ_AXIS : S_XZ_CELL;
    return flip ? S_XAXIS_ZCELL : S_XCELL_ZAXIS;
}

struct StepOpts {
    int split; int interp; int use_integral; int correction; int ng; double min_area;
};

// psi mode that the reference derives from interpolation_method
// (reference src/move_substances_3d.cpp:61-83, src/advect_velocity_flux_advection_3d.cpp:601-620).
static int psi_mode_of(int interp, int &use_integral) {
    if (interp == I_LINEAR) return P_LINEAR;
    if (interp == I_1DY_LINEAR) { use_integral = 0; return P_CONST; }
    return P_WENO6;
}

// Density step with the zero-velocity correction (reference src/move_substances_3d.cpp:38-240).
template <class R, class C>
static int move_substances(int nx, int ny, int nz, R h, R dt, R *rho, const Velocity<R> &vel,
                           const StepOpts &s, int nthreads, double *stats) {
    Comp G(-1, nx, ny, nz);
    const size_t n = (size_t)nx * ny * nz;
    std::vector<R> before(rho, rho + n);
    int integ = s.use_integral;
    const int pm = psi_mode_of(s.interp, integ);
    PassOpts p1{-1, s.split, s.interp, pm, 0, integ, s.ng, s.min_area};
    int rc = advect_pass<R, C>(G, vel, h, dt, rho, p1, nthreads, stats);
    if (rc) return rc;
    if (s.correction) {
        std::vector<R> e(before);
        PassOpts p2{-1, s.split, s.interp, pm, 1, 0, s.ng, s.min_area};
        rc = advect_pass<R, C>(G, vel, h, dt, e.data(), p2, nthreads, stats);
        if (rc) return rc;
        for (size_t i = 0; i < n; ++i) e[i] = -e[i] + before[i];
        PassOpts p3{-1, S_XZ_AXIS, I_1DY_LINEAR, P_CONST, 0, 0, s.ng, s.min_area};
        rc = advect_pass<R, C>(G, vel, h, dt, e.data(), p3, nthreads, stats);
        if (rc) return rc;
        for (size_t i = 0; i < n; ++i) rho[i] += e[i];
    }
    return 0;
}

// Self-advection of one velocity component (reference
// src/advect_velocity_flux_advection_3d.cpp:580-759 for x, :763-952 for y, :956-1134 for z).
// comp_in/comp_out are G-sized; vel is the velocity used to backtrace (pre-step velocity).
template <class R, class C>
static int advect_component(int comp, int nx, int ny, int nz, R h, R dt, const R *comp_in, R *comp_out,
                            const Velocity<R> &vel, const StepOpts &s, int nthreads, double *stats) {
    Comp G(comp, nx, ny, nz);
    const size_t n = (size_t)G.m[0] * G.m[1] * G.m[2];
    std::copy(comp_in, comp_in + n, comp_out);
    int integ = s.use_integral;
    const int pm = psi_mode_of(s.interp, integ);
    const bool one_dy = (s.interp == I_1DY_LINEAR || s.interp == I_1DY_WENO6);
    PassOpts p1{comp, velocity_split(comp, one_dy, false), s.interp, pm, 0, integ, s.ng, s.min_area};
    int rc = advect_pass<R, C>(G, vel, h, dt, comp_out, p1, nthreads, stats);
    if (rc) return rc;
    if (s.correction) {
        std::vector<R> e(comp_in, comp_in + n);
        PassOpts p2{comp, velocity_split(comp, one_dy, false), s.interp, pm, 1, 0, s.ng, s.min_area};
        rc = advect_pass<R, C>(G, vel, h, dt, e.data(), p2, nthreads, stats);
        if (rc) return rc;
        for (size_t i = 0; i < n; ++i) e[i] = -e[i] + comp_in[i];
        PassOpts p3{comp, velocity_split(comp, false, true), I_1DY_LINEAR, P_CONST, 0, 0, s.ng, s.min_area};
        rc = advect_pass<R, C>(G, vel, h, dt, e.data(), p3, nthreads, stats);
        if (rc) return rc;
        for (size_t i = 0; i < n; ++i) comp_out[i] += e[i];
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Pressure projection (reference src/calc_pressure_3d.cpp:18-269).  The reference solves with
// amgcl; the solver here is Jacobi-PCG in FP64 on the same matrix (the last cell pinned to 0),
// which the residual check ties down independently of the solver.  Solid cells (optional) are
// extra Neumann walls and their faces get zero velocity (obstacles are not in the reference).
// ---------------------------------------------------------------------------------------------
static int pressure_project(int nx, int ny, int nz, double h, double *u, double *v, double *w,
                            const uint8_t *solid, double *p_out, double tol, int max_iters,
                            int *iters_out, double *res_out) {
    const size_t n = (size_t)nx * ny * nz;
    auto fluid = [&](int i, int j, int k) {
        return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz && !(solid && solid[idx3(i, j, k, ny, nz)]);
    };
    size_t pin = n;   // last fluid cell
    for (size_t c = n; c-- > 0;) { int k = c % nz, j = (c / nz) % ny, i = c / ((size_t)ny * nz); if (fluid(i, j, k)) { pin = c; break; } }
    if (solid) {   // zero the faces of obstacle cells
        for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
            if (fluid(i, j, k)) continue;
            u[idx3(i, j, k, ny, nz)] = 0; u[idx3(i + 1, j, k, ny, nz)] = 0;
            v[idx3(i, j, k, ny + 1, nz)] = 0; v[idx3(i, j + 1, k, ny + 1, nz)] = 0;
            w[idx3(i, j, k, ny, nz + 1)] = 0; w[idx3(i, j, k + 1, ny, nz + 1)] = 0;
        }
    }
    std::vector<double> b(n, 0.0), diag(n, 0.0);
    for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
        const size_t c = idx3(i, j, k, ny, nz);
        if (!fluid(i, j, k) || c == pin) continue;
        const double div = (u[idx3(i + 1, j, k, ny, nz)] - u[idx3(i, j, k, ny, nz)])
                         + (v[idx3(i, j + 1, k, ny + 1, nz)] - v[idx3(i, j, k, ny + 1, nz)])
                         + (w[idx3(i, j, k + 1, ny, nz + 1)] - w[idx3(i, j, k, ny, nz + 1)]);
        b[c] = -div * h;   // system L p = -b_ref with L = -A (positive semidefinite)
        int nb = 0;
        nb += fluid(i - 1, j, k); nb += fluid(i + 1, j, k); nb += fluid(i, j - 1, k);
        nb += fluid(i, j + 1, k); nb += fluid(i, j, k - 1); nb += fluid(i, j, k + 1);
        diag[c] = nb;
    }
    auto apply = [&](const std::vector<double> &x, std::vector<double> &y) {
        for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
            const size_t c = idx3(i, j, k, ny, nz);
            if (!fluid(i, j, k) || c == pin) { y[c] = 0; continue; }
            double s = diag[c] * x[c];
            auto sub = [&](int a, int bb, int cc) {
                if (!fluid(a, bb, cc)) return;
                const size_t d = idx3(a, bb, cc, ny, nz);
                if (d == pin) return;
                s -= x[d];
            };
            sub(i - 1, j, k); sub(i + 1, j, k); sub(i, j - 1, k); sub(i, j + 1, k); sub(i, j, k - 1); sub(i, j, k + 1);
            y[c] = s;
        }
    };
    std::vector<double> x(n, 0.0), r(b), z(n), pd(n), q(n);
    if (p_out) for (size_t c = 0; c < n; ++c) x[c] = (fluid(c / ((size_t)ny * nz), (c / nz) % ny, c % nz) && c != pin) ? p_out[c] : 0.0;
    apply(x, q);
    for (size_t c = 0; c < n; ++c) r[c] = b[c] - q[c];
    double bn = 0; for (double t : b) bn += t * t; bn = std::sqrt(bn);
    auto norm = [&](const std::vector<double> &a) { double s = 0; for (double t : a) s += t * t; return std::sqrt(s); };
    int it = 0;
    double rn = norm(r);
    if (bn == 0.0) bn = 1.0;
    double rz = 0;
    for (size_t c = 0; c < n; ++c) { z[c] = diag[c] > 0 ? r[c] / diag[c] : 0.0; rz += r[c] * z[c]; }
    pd = z;
    while (rn > tol * bn && it < max_iters) {
        apply(pd, q);
        double pq = 0; for (size_t c = 0; c < n; ++c) pq += pd[c] * q[c];
        const double alpha = rz / pq;
        for (size_t c = 0; c < n; ++c) { x[c] += alpha * pd[c]; r[c] -= alpha * q[c]; }
        double rz2 = 0;
        for (size_t c = 0; c < n; ++c) { z[c] = diag[c] > 0 ? r[c] / diag[c] : 0.0; rz2 += r[c] * z[c]; }
        const double beta = rz2 / rz; rz = rz2;
        for (size_t c = 0; c < n; ++c) pd[c] = z[c] + beta * pd[c];
        rn = norm(r);
        ++it;
    }
    // u -= grad p / h on interior faces between two fluid cells (reference :196-234)
    for (int i = 1; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k)
        if (fluid(i, j, k) && fluid(i - 1, j, k)) u[idx3(i, j, k, ny, nz)] -= (x[idx3(i, j, k, ny, nz)] - x[idx3(i - 1, j, k, ny, nz)]) / h;
    for (int i = 0; i < nx; ++i) for (int j = 1; j < ny; ++j) for (int k = 0; k < nz; ++k)
        if (fluid(i, j, k) && fluid(i, j - 1, k)) v[idx3(i, j, k, ny + 1, nz)] -= (x[idx3(i, j, k, ny, nz)] - x[idx3(i, j - 1, k, ny, nz)]) / h;
    for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 1; k < nz; ++k)
        if (fluid(i, j, k) && fluid(i, j, k - 1)) w[idx3(i, j, k, ny, nz + 1)] -= (x[idx3(i, j, k, ny, nz)] - x[idx3(i, j, k - 1, ny, nz)]) / h;
    if (p_out) for (size_t c = 0; c < n; ++c) p_out[c] = x[c];
    if (iters_out) *iters_out = it;
    if (res_out) *res_out = rn / bn;
    return rn <= tol * bn ? 0 : 7;
}

}  // namespace orc

// =============================================================================================
// C API (double arrays in and out; prec selects R = double (0) or float (1), cprec selects the
// cutter type C = double (0) or float (1)).
// =============================================================================================
using namespace orc;

template <class R> static std::vector<R> to_r(const double *p, size_t n) { return std::vector<R>(p, p + n); }

template <class R>
static Velocity<R> make_vel(int nx, int ny, int nz, const std::vector<R> &u, const std::vector<R> &v, const std::vector<R> &w) {
    Velocity<R> vel;
    vel.u = Field<R>{u.data(), nx + 1, ny, nz, 0.0, 0.5, 0.5};
    vel.v = Field<R>{v.data(), nx, ny + 1, nz, 0.5, 0.0, 0.5};
    vel.w = Field<R>{w.data(), nx, ny, nz + 1, 0.5, 0.5, 0.0};
    return vel;
}

extern "C" int orc_gauss_legendre(int n, double *x, double *w) {
    if (!gauss_n_supported(n)) return 1;
    GaussRule g = gauss_legendre(n);
    for (int i = 0; i < n; ++i) { x[i] = g.x[i]; w[i] = g.w[i]; }
    return 0;
}

extern "C" double orc_weno1d(int prec, const double *f, const double *d, double h) {
    if (prec) { float ff[6], dd[6]; for (int i = 0; i < 6; ++i) { ff[i] = (float)f[i]; dd[i] = (float)d[i]; } return weno1d<float>(ff, dd, (float)h); }
    return weno1d<double>(f, d, h);
}
extern "C" double orc_weno1d_opt(const double *f, const double *d, double h) { return weno1d<double>(f, d, h, true); }

template <class R>
static void interp_t(const double *field, int mx, int my, int mz, double ox, double oy, double oz, double h,
                     int method, const double *pts, long n, double *out) {
    std::vector<R> f = to_r<R>(field, (size_t)mx * my * mz);
    Field<R> F{f.data(), mx, my, mz, ox, oy, oz};
    for (long i = 0; i < n; ++i)
        out[i] = (double)interp<R>(F, method, R(pts[3 * i]), R(pts[3 * i + 1]), R(pts[3 * i + 2]), R(h));
}
extern "C" void orc_interpolate(int prec, const double *field, int mx, int my, int mz, double ox, double oy, double oz,
                     double h, int method, const double *pts, long n, double *out) {
    if (prec) interp_t<float>(field, mx, my, mz, ox, oy, oz, h, method, pts, n, out);
    else interp_t<double>(field, mx, my, mz, ox, oy, oz, h, method, pts, n, out);
}

template <class R>
static void psi_t(int comp, int nx, int ny, int nz, double h, const double *q, int mode, int integ, int ng, double *psi) {
    Comp G(comp, nx, ny, nz);
    const size_t nq = (size_t)G.m[0] * G.m[1] * G.m[2], np = (size_t)G.m[0] * (G.m[1] + 1) * G.m[2];
    std::vector<R> qq = to_r<R>(q, nq), pp(np);
    calc_psi<R>(G, qq.data(), R(h), mode, integ, ng, pp.data());
    for (size_t i = 0; i < np; ++i) psi[i] = pp[i];
}
extern "C" int orc_calc_psi(int prec, int comp, int nx, int ny, int nz, double h, const double *q, int mode, int integ,
                 int ng, double *psi) {
    if (!gauss_n_supported(ng)) return 1;
    if (prec) psi_t<float>(comp, nx, ny, nz, h, q, mode, integ, ng, psi);
    else psi_t<double>(comp, nx, ny, nz, h, q, mode, integ, ng, psi);
    return 0;
}

// Backtraced positions (physical units) of every vertex of the lattice of component comp.
template <class R>
static void bt_t(int comp, int nx, int ny, int nz, double h, const double *u, const double *v, const double *w,
                 double dt, int method, double *out) {
    std::vector<R> uu = to_r<R>(u, (size_t)(nx + 1) * ny * nz), vv = to_r<R>(v, (size_t)nx * (ny + 1) * nz),
                   ww = to_r<R>(w, (size_t)nx * ny * (nz + 1));
    Velocity<R> vel = make_vel<R>(nx, ny, nz, uu, vv, ww);
    Comp G(comp, nx, ny, nz);
    const int lx = G.m[0] + 1, ly = G.m[1] + 1, lz = G.m[2] + 1;
    for (int i = 0; i < lx; ++i) for (int j = 0; j < ly; ++j) for (int k = 0; k < lz; ++k) {
        R p[3];
        backtrace_vertex<R>(G, vel, method, R(h), R(dt), i, j, k, false, p);
        const size_t c = idx3(i, j, k, ly, lz);
        for (int a = 0; a < 3; ++a) out[3 * c + a] = p[a];
    }
}
extern "C" void orc_backtrace_vertices(int prec, int comp, int nx, int ny, int nz, double h, const double *u, const double *v,
                            const double *w, double dt, int method, double *out) {
    if (prec) bt_t<float>(comp, nx, ny, nz, h, u, v, w, dt, method, out);
    else bt_t<double>(comp, nx, ny, nz, h, u, v, w, dt, method, out);
}

// xi-eta pieces of one face from its corner colours (for checking the cutter against the
// reference-built facecutter).  Writes up to max_pieces pieces: vertex count, up to 32 xi-eta
// vertices, centroid and area.  Returns the piece count, or -1 on overflow.
template <class C>
static int cut_t(const double *colours, double min_area, int max_pieces, int *nverts, double *verts,
                 double *centroid, double *area) {
    C col[4][2];
    for (int i = 0; i < 4; ++i) { col[i][0] = (C)colours[2 * i]; col[i][1] = (C)colours[2 * i + 1]; }
    int cnt = 0; bool over = false;
    auto emit = [&](const P2<C> *poly, int num, const P2<C> &c, C a) {
        if (cnt >= max_pieces) { over = true; return; }
        nverts[cnt] = num;
        for (int k = 0; k < num; ++k) { verts[(cnt * kMaxTmp + k) * 2] = poly[k].x[0]; verts[(cnt * kMaxTmp + k) * 2 + 1] = poly[k].x[1]; }
        centroid[2 * cnt] = c.x[0]; centroid[2 * cnt + 1] = c.x[1]; area[cnt] = a;
        ++cnt;
    };
    if (!split_xieta<C>(col, (C)min_area, emit) || over) return -1;
    return cnt;
}
extern "C" int orc_cut_face_xieta(int cprec, const double *colours, double min_area, int max_pieces, int *nverts,
                       double *verts, double *centroid, double *area) {
    if (cprec) return cut_t<float>(colours, min_area, max_pieces, nverts, verts, centroid, area);
    return cut_t<double>(colours, min_area, max_pieces, nverts, verts, centroid, area);
}

// Pieces (3D centre, vector area) of one face given by its 4 vertices (physical units).
template <class R, class C>
static int split3d_t(const double *quad, double h, int split, double min_area, int max_pieces, double *cen, double *an) {
    R v[4][3];
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 3; ++k) v[i][k] = (R)quad[3 * i + k];
    C col[4][2];
    face_colours<C, R>(v, (R)h, split, col);
    int cnt = 0; bool over = false;
    auto emit = [&](const P2<C> *poly, int num, const P2<C> &c, C) {
        if (cnt >= max_pieces) { over = true; return; }
        R pos[3], a[3];
        piece_geometry<R, C>(v, poly, num, c, pos, a);
        for (int k = 0; k < 3; ++k) { cen[3 * cnt + k] = pos[k]; an[3 * cnt + k] = a[k]; }
        ++cnt;
    };
    if (!split_xieta<C>(col, (C)min_area, emit) || over) return -1;
    return cnt;
}
extern "C" int orc_split_face(int prec, int cprec, const double *quad, double h, int split, double min_area, int max_pieces,
                   double *cen, double *an) {
    if (prec) return cprec ? split3d_t<float, float>(quad, h, split, min_area, max_pieces, cen, an)
                           : split3d_t<float, double>(quad, h, split, min_area, max_pieces, cen, an);
    return cprec ? split3d_t<double, float>(quad, h, split, min_area, max_pieces, cen, an)
                 : split3d_t<double, double>(quad, h, split, min_area, max_pieces, cen, an);
}

template <class R, class C>
static int pass_t(int comp, int nx, int ny, int nz, double h, const double *u, const double *v, const double *w,
                  double *q, double dt, int split, int method, int psi_mode, int zero_vel, int integ, int ng,
                  double min_area, int nthreads, double *stats) {
    std::vector<R> uu = to_r<R>(u, (size_t)(nx + 1) * ny * nz), vv = to_r<R>(v, (size_t)nx * (ny + 1) * nz),
                   ww = to_r<R>(w, (size_t)nx * ny * (nz + 1));
    Velocity<R> vel = make_vel<R>(nx, ny, nz, uu, vv, ww);
    Comp G(comp, nx, ny, nz);
    const size_t nq = (size_t)G.m[0] * G.m[1] * G.m[2];
    std::vector<R> qq = to_r<R>(q, nq);
    PassOpts o{comp, split, method, psi_mode, zero_vel, integ, ng, min_area};
    int rc = advect_pass<R, C>(G, vel, R(h), R(dt), qq.data(), o, nthreads, stats);
    for (size_t i = 0; i < nq; ++i) q[i] = qq[i];
    return rc;
}
extern "C" int orc_advect_pass(int prec, int cprec, int comp, int nx, int ny, int nz, double h, const double *u, const double *v,
                    const double *w, double *q, double dt, int split, int method, int psi_mode, int zero_vel,
                    int integ, int ng, double min_area, int nthreads, double *stats) {
    if (!gauss_n_supported(ng)) return 1;
    if (prec) return cprec ? pass_t<float, float>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, nthreads, stats)
                           : pass_t<float, double>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, nthreads, stats);
    return cprec ? pass_t<double, float>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, nthreads, stats)
                 : pass_t<double, double>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, nthreads, stats);
}

template <class R, class C>
static int move_t(int nx, int ny, int nz, double h, double *rho, const double *u, const double *v, const double *w,
                  double dt, int split, int method, int integ, int corr, int ng, double min_area, int nthreads,
                  double *stats) {
    std::vector<R> uu = to_r<R>(u, (size_t)(nx + 1) * ny * nz), vv = to_r<R>(v, (size_t)nx * (ny + 1) * nz),
                   ww = to_r<R>(w, (size_t)nx * ny * (nz + 1));
    Velocity<R> vel = make_vel<R>(nx, ny, nz, uu, vv, ww);
    const size_t n = (size_t)nx * ny * nz;
    std::vector<R> rr = to_r<R>(rho, n);
    StepOpts s{split, method, integ, corr, ng, min_area};
    int rc = move_substances<R, C>(nx, ny, nz, R(h), R(dt), rr.data(), vel, s, nthreads, stats);
    for (size_t i = 0; i < n; ++i) rho[i] = rr[i];
    return rc;
}
extern "C" int orc_move_substances(int prec, int cprec, int nx, int ny, int nz, double h, double *rho, const double *u,
                        const double *v, const double *w, double dt, int split, int method, int integ, int corr,
                        int ng, double min_area, int nthreads, double *stats) {
    if (!gauss_n_supported(ng)) return 1;
    if (method == I_1DY_WENO6) return 2;
    if (prec) return cprec ? move_t<float, float>(nx, ny, nz, h, rho, u, v, w, dt, split, method, integ, corr, ng, min_area, nthreads, stats)
                           : move_t<float, double>(nx, ny, nz, h, rho, u, v, w, dt, split, method, integ, corr, ng, min_area, nthreads, stats);
    return cprec ? move_t<double, float>(nx, ny, nz, h, rho, u, v, w, dt, split, method, integ, corr, ng, min_area, nthreads, stats)
                 : move_t<double, double>(nx, ny, nz, h, rho, u, v, w, dt, split, method, integ, corr, ng, min_area, nthreads, stats);
}

template <class R, class C>
static int ff_t(int comp, int nx, int ny, int nz, double h, const double *u, const double *v, const double *w,
                const double *q, double dt, int split, int method, int psi_mode, int zero_vel, int integ, int ng,
                double min_area, double *fx, double *fy, double *fz) {
    std::vector<R> uu = to_r<R>(u, (size_t)(nx + 1) * ny * nz), vv = to_r<R>(v, (size_t)nx * (ny + 1) * nz),
                   ww = to_r<R>(w, (size_t)nx * ny * (nz + 1));
    Velocity<R> vel = make_vel<R>(nx, ny, nz, uu, vv, ww);
    Comp G(comp, nx, ny, nz);
    std::vector<R> qq = to_r<R>(q, (size_t)G.m[0] * G.m[1] * G.m[2]);
    PassOpts o{comp, split, method, psi_mode, zero_vel, integ, ng, min_area};
    return face_fluxes<R, C>(G, vel, R(h), R(dt), qq.data(), o, fx, fy, fz);
}
extern "C" int orc_face_fluxes(int prec, int cprec, int comp, int nx, int ny, int nz, double h, const double *u,
                               const double *v, const double *w, const double *q, double dt, int split, int method,
                               int psi_mode, int zero_vel, int integ, int ng, double min_area, double *fx,
                               double *fy, double *fz) {
    if (!gauss_n_supported(ng)) return 1;
    if (prec) return cprec ? ff_t<float, float>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, fx, fy, fz)
                           : ff_t<float, double>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, fx, fy, fz);
    return cprec ? ff_t<double, float>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, fx, fy, fz)
                 : ff_t<double, double>(comp, nx, ny, nz, h, u, v, w, q, dt, split, method, psi_mode, zero_vel, integ, ng, min_area, fx, fy, fz);
}

// Velocity self-advection; dim = 0,1,2 advects one component, -1 all three through the pre-step
// velocity (reference src/advect_velocity_flux_advection_3d.cpp:1138-1272).
template <class R, class C>
static int vel_t(int dim, int nx, int ny, int nz, double h, double *u, double *v, double *w, double dt, int method,
                 int integ, int corr, int ng, double min_area, int nthreads, double *stats) {
    const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz, nw = (size_t)nx * ny * (nz + 1);
    std::vector<R> uu = to_r<R>(u, nu), vv = to_r<R>(v, nv), ww = to_r<R>(w, nw);
    Velocity<R> vel = make_vel<R>(nx, ny, nz, uu, vv, ww);
    StepOpts s{0, method, integ, corr, ng, min_area};
    std::vector<R> ou(uu), ov(vv), ow(ww);
    int rc = 0;
    if (dim == 0 || dim == -1) { rc = advect_component<R, C>(0, nx, ny, nz, R(h), R(dt), uu.data(), ou.data(), vel, s, nthreads, stats); if (rc) return rc; }
    if (dim == 1 || dim == -1) { rc = advect_component<R, C>(1, nx, ny, nz, R(h), R(dt), vv.data(), ov.data(), vel, s, nthreads, stats); if (rc) return rc; }
    if (dim == 2 || dim == -1) { rc = advect_component<R, C>(2, nx, ny, nz, R(h), R(dt), ww.data(), ow.data(), vel, s, nthreads, stats); if (rc) return rc; }
    for (size_t i = 0; i < nu; ++i) u[i] = ou[i];
    for (size_t i = 0; i < nv; ++i) v[i] = ov[i];
    for (size_t i = 0; i < nw; ++i) w[i] = ow[i];
    return 0;
}
extern "C" int orc_advect_velocity(int prec, int cprec, int dim, int nx, int ny, int nz, double h, double *u, double *v,
                        double *w, double dt, int method, int integ, int corr, int ng, double min_area, int nthreads,
                        double *stats) {
    if (!gauss_n_supported(ng)) return 1;
    if (method == I_1DY_WENO6) return 2;
    if (method == I_WENO6_OPT) return 2;   // the reference's velocity path has no branch for it
    if (prec) return cprec ? vel_t<float, float>(dim, nx, ny, nz, h, u, v, w, dt, method, integ, corr, ng, min_area, nthreads, stats)
                           : vel_t<float, double>(dim, nx, ny, nz, h, u, v, w, dt, method, integ, corr, ng, min_area, nthreads, stats);
    return cprec ? vel_t<double, float>(dim, nx, ny, nz, h, u, v, w, dt, method, integ, corr, ng, min_area, nthreads, stats)
                 : vel_t<double, double>(dim, nx, ny, nz, h, u, v, w, dt, method, integ, corr, ng, min_area, nthreads, stats);
}

// CFL step (reference src/calc_time_step_length_from_CFL_number.cpp:81-132).
extern "C" double orc_cfl_time_step(int nx, int ny, int nz, double h, const double *u, const double *v, const double *w,
                         double now_time, int i_movie_frame, int cfl, double fps, int *write_frame) {
    double m = 0;
    for (size_t i = 0; i < (size_t)(nx + 1) * ny * nz; ++i) m = std::max(m, std::fabs(u[i]));
    for (size_t i = 0; i < (size_t)nx * (ny + 1) * nz; ++i) m = std::max(m, std::fabs(v[i]));
    for (size_t i = 0; i < (size_t)nx * ny * (nz + 1); ++i) m = std::max(m, std::fabs(w[i]));
    const double dt = cfl * h / m;
    const double next = (i_movie_frame + 1) / fps;
    if (now_time + dt > next) { *write_frame = 1; return next - now_time; }
    *write_frame = 0;
    return dt;
}

extern "C" int orc_pressure_project(int nx, int ny, int nz, double h, double *u, double *v, double *w, const uint8_t *solid,
                         double *p, double tol, int max_iters, int *iters, double *res) {
    return pressure_project(nx, ny, nz, h, u, v, w, solid, p, tol, max_iters, iters, res);
}

