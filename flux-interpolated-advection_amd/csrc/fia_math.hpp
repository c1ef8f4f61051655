// fia_math.hpp - interpolation kernels of the flux-interpolated advection path (gfx950).
//
// Coordinates are in cell units.  A staggered field has dims (mx,my,mz) and its sample (i,j,k)
// sits at (i,j,k) + o; callers pass the sample coordinate s = p - o split into an integer base
// floor(s) and a fraction t in [0,1).  Index clamping to [0, m-1] is clamp-to-edge as in the
// reference (src/WENO6_interpolation_3d.cpp:28-60).
//
// WENO6 is evaluated in difference form: the six samples enter through their first differences
// g_k = f_{k+1} - f_k and the result is f_2 + delta.  In exact arithmetic this equals the
// reference's form (src/WENO6_interpolation_3d.cpp:63-123: Newton-form cubic candidates,
// ideal weights C0 = d4 d5/20, C1 = -d5 d0/10, C2 = d0 d1/20, smoothness indicators /180,
// alpha = C/(1e-6+S)^2); the smoothness indicators annihilate constants, so they are rewritten
// as quadratic forms of the differences.  Evaluated that way nothing cancels when the samples
// are large and close together (the y-prefix psi reaches n*q at the top of a column), so psi
// can be interpolated with FP32 arithmetic around an FP64 base.  kappa rescales S to the
// reference's physical units (S_phys = kappa * S) so the 1e-6 epsilon keeps its meaning.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#define FIA_HD __host__ __device__ __forceinline__

namespace fia {

FIA_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
FIA_HD size_t idx3(int i, int j, int k, int my, int mz) { return ((size_t)i * my + j) * mz + k; }

FIA_HD float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// Coefficients of one WENO6 evaluation that depend only on the fraction t (shared by every
// evaluation of a pass).  With d_i = t + 2 - i: candidate r is f_2 + g-linear Newton terms with
// the products a_r = d_r d_{r+1} / 2 and b_r = d_r d_{r+1} d_{r+2} / 6; the ideal weights are
// C0 = d4 d5 / 20, C1 = -d5 d0 / 10, C2 = d0 d1 / 20 (all positive on [0,1]).
struct WenoCoef {
    float t, t1, a0, b0, a1, b1, a2, b2, c0, c1, c2, k;
    FIA_HD explicit WenoCoef(float tt, float kappa) {
        const float d0 = tt + 2.0f, d1 = tt + 1.0f, d2 = tt, d3 = tt - 1.0f, d4 = tt - 2.0f, d5 = tt - 3.0f;
        t = tt; t1 = d1;
        a0 = d0 * d1 * 0.5f; b0 = d0 * d1 * d2 * (1.0f / 6.0f);
        a1 = d1 * d2 * 0.5f; b1 = d1 * d2 * d3 * (1.0f / 6.0f);
        a2 = d2 * d3 * 0.5f; b2 = d2 * d3 * d4 * (1.0f / 6.0f);
        c0 = d4 * d5 * (1.0f / 20.0f);
        c1 = -d5 * d0 * (1.0f / 10.0f);
        c2 = d0 * d1 * (1.0f / 20.0f);
        k = kappa * (1.0f / 180.0f);
    }
};

// delta = WENO6(f)(t) - f_2 from the five first differences g_k = f_{k+1} - f_k.
FIA_HD float weno6_delta(float g0, float g1, float g2, float g3, float g4, const WenoCoef &w) {
    // Newton-form cubic candidates on samples {0..3}, {1..4}, {2..5}, relative to f_2
    const float h0 = g1 - g0, h1 = g2 - g1, h2 = g3 - g2, h3 = g4 - g3;
    const float q0 = (g0 * w.t1 - g1) + h0 * w.a0 + (h1 - h0) * w.b0;
    const float q1 = g1 * w.t + h1 * w.a1 + (h2 - h1) * w.b1;
    const float q2 = g2 * w.t + h2 * w.a2 + (h3 - h2) * w.b2;
    // smoothness indicators (x180): quadratic forms over shared products
    const float p00 = g0 * g0, p11 = g1 * g1, p22 = g2 * g2, p33 = g3 * g3, p44 = g4 * g4;
    const float p01 = g0 * g1, p02 = g0 * g2, p12 = g1 * g2, p13 = g1 * g3, p23 = g2 * g3;
    const float p24 = g2 * g4, p34 = g3 * g4;
    const float s0 = 244.0f * p00 + 1561.0f * p11 + 814.0f * p22 - 1171.0f * p01 + 683.0f * p02 - 1951.0f * p12;
    const float s1 = 244.0f * p11 + 961.0f * p22 + 244.0f * p33 - 781.0f * p12 + 293.0f * p13 - 781.0f * p23;
    const float s2 = 814.0f * p22 + 1561.0f * p33 + 244.0f * p44 - 1951.0f * p23 + 683.0f * p24 - 1171.0f * p34;
    const float e0 = w.k * s0 + 1e-6f, e1 = w.k * s1 + 1e-6f, e2 = w.k * s2 + 1e-6f;
    const float E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
    if (fmaxf(e0, fmaxf(e1, e2)) < 1e9f) {
        // weights C_r / e_r^2 scaled by E0 E1 E2 (one reciprocal): below e = 1e9 every product
        // stays under 1e36 and above 1e-36, and the weights are normalised before they meet q
        const float a0 = w.c0 * (E1 * E2), a1 = w.c1 * (E0 * E2), a2 = w.c2 * (E0 * E1);
        const float r = fast_rcp(a0 + a1 + a2);
        return (a0 * r) * q0 + (a1 * r) * q1 + (a2 * r) * q2;
    }
    const float a0 = w.c0 * fast_rcp(E0), a1 = w.c1 * fast_rcp(E1), a2 = w.c2 * fast_rcp(E2);
    return (a0 * q0 + a1 * q1 + a2 * q2) * fast_rcp(a0 + a1 + a2);
}

// WENO6 of six samples of type V (float, or double for the y-prefix psi).
template <class V> FIA_HD V weno6(const V f[6], const WenoCoef &w) {
    const float g0 = (float)(f[1] - f[0]), g1 = (float)(f[2] - f[1]), g2 = (float)(f[3] - f[2]);
    const float g3 = (float)(f[4] - f[3]), g4 = (float)(f[5] - f[4]);
    return f[2] + (V)weno6_delta(g0, g1, g2, g3, g4, w);
}
template <class V> FIA_HD V weno6(const V f[6], float t, float kappa) { return weno6<V>(f, WenoCoef(t, kappa)); }

// Clamped loader of a staggered field.
template <class V> struct Loader {
    const V *__restrict__ d;
    int mx, my, mz;
    FIA_HD V operator()(int i, int j, int k) const {
        return d[idx3(clampi(i, 0, mx - 1), clampi(j, 0, my - 1), clampi(k, 0, mz - 1), my, mz)];
    }
};

// Six consecutive samples of row (x, y) along z from z0.  When the whole row lies inside the field
// (interior) they are one contiguous run of memory and come in as wide loads; otherwise every
// sample is clamped on its own.
template <class V> FIA_HD void load_z6(const Loader<V> &L, int x, int y, int z0, bool interior, V f[6]) {
    if (interior) {
        const V *p = L.d + idx3(clampi(x, 0, L.mx - 1), clampi(y, 0, L.my - 1), z0, L.my, L.mz);
        struct Row { V v[6]; } r;
        __builtin_memcpy(&r, p, sizeof(r));
#pragma unroll
        for (int i = 0; i < 6; ++i) f[i] = r.v[i];
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) f[i] = L(x, y, z0 + i);
    }
}

// Six consecutive samples at a 32-bit element offset from the field's base (a row that lies
// inside the field: the caller has checked the bounds, and that the byte offset fits in 32 bits).
// The address is the uniform base plus a 32-bit byte offset, the form of a global load with a
// scalar base.
template <class V> FIA_HD void load_row6(const V *__restrict__ base, unsigned off, V f[6]) {
    struct Row { V v[6]; } r;
    const unsigned bytes = off * (unsigned)sizeof(V);
    __builtin_memcpy(&r, reinterpret_cast<const char *>(base) + bytes, sizeof(r));
#pragma unroll
    for (int i = 0; i < 6; ++i) f[i] = r.v[i];
}

// 3D WENO6, passes z then y then x (reference src/WENO6_interpolation_3d.cpp:7-258).  A pass
// whose sample coordinate is on a grid node (SX/SY/SZ) is skipped: WENO6 at t = 0 returns f_2
// exactly (every candidate interpolates node 2 and the weights sum to 1).
template <bool SX, bool SY, bool SZ, class V, class L>
FIA_HD V weno3d_t(const L &ld, int bx, int by, int bz, float tx, float ty, float tz, float kappa) {
    V yv[6];
#pragma unroll
    for (int ix = 0; ix < 6; ++ix) {
        if (SX && ix != 2) continue;
        V zr[6];
#pragma unroll
        for (int iy = 0; iy < 6; ++iy) {
            if (SY && iy != 2) continue;
            if (SZ) {
                zr[iy] = ld(bx - 2 + ix, by - 2 + iy, bz);
            } else {
                V f[6];
#pragma unroll
                for (int iz = 0; iz < 6; ++iz) f[iz] = ld(bx - 2 + ix, by - 2 + iy, bz - 2 + iz);
                zr[iy] = weno6<V>(f, tz, kappa);
            }
        }
        yv[ix] = SY ? zr[2] : weno6<V>(zr, ty, kappa);
    }
    return SX ? yv[2] : weno6<V>(yv, tx, kappa);
}

// Runtime dispatch on which passes sit on a node (uniform inside a wave for lattice vertices).
template <class V, class L>
FIA_HD V weno3d(const L &ld, int bx, int by, int bz, float tx, float ty, float tz, float kappa) {
    const int m = (tx == 0.0f ? 4 : 0) | (ty == 0.0f ? 2 : 0) | (tz == 0.0f ? 1 : 0);
    switch (m) {
        case 0: return weno3d_t<false, false, false, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 1: return weno3d_t<false, false, true, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 2: return weno3d_t<false, true, false, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 3: return weno3d_t<false, true, true, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 4: return weno3d_t<true, false, false, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 5: return weno3d_t<true, false, true, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 6: return weno3d_t<true, true, false, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        default: return ld(bx, by, bz);
    }
}

// Trilinear, passes z, y, x (reference src/linear_interpolation_3d.cpp:7-95).
template <class V, class L>
FIA_HD V linear3d(const L &ld, int bx, int by, int bz, float tx, float ty, float tz) {
    V yz[2];
#pragma unroll
    for (int ix = 0; ix < 2; ++ix) {
        V zz[2];
#pragma unroll
        for (int iy = 0; iy < 2; ++iy) {
            const V v0 = ld(bx + ix, by + iy, bz), v1 = ld(bx + ix, by + iy, bz + 1);
            zz[iy] = (V)(1.0f - tz) * v0 + (V)tz * v1;
        }
        yz[ix] = (V)(1.0f - ty) * zz[0] + (V)ty * zz[1];
    }
    return (V)(1.0f - tx) * yz[0] + (V)tx * yz[1];
}

// Linear in y inside column (cx, cz) (reference src/linear_interpolation_1d.cpp:314-374).
template <class V, class L> FIA_HD V linear1dy(const L &ld, int cx, int by, int cz, float ty) {
    return (V)(1.0f - ty) * ld(cx, by, cz) + (V)ty * ld(cx, by + 1, cz);
}

// WENO6 in y inside column (cx, cz) (reference src/WENO6_interpolation_1d.cpp:315-430).
template <class V, class L> FIA_HD V weno1dy(const L &ld, int cx, int by, int cz, float ty, float kappa) {
    V f[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) f[i] = ld(cx, by - 2 + i, cz);
    return weno6<V>(f, ty, kappa);
}

// floor split of a double sample coordinate into base and fraction
FIA_HD void split_coord(double s, int &b, float &t) {
    const double f = floor(s);
    b = (int)f;
    t = (float)(s - f);
}

enum InterpMethod { IM_LINEAR = 0, IM_WENO6 = 1, IM_1DY_LINEAR = 2, IM_1DY_WENO6 = 3 };

// Interpolation of a staggered field at p (cell units, in the same frame as the field's
// offsets o, which are in cells).  o_col is the physical offset used by the 1Dy-WENO6 column
// rule (floor(p_x/h), reference src/WENO6_interpolation_1d.cpp:320-322).
template <class V, class L>
FIA_HD V interp_point(const L &ld, int method, double px, double py, double pz, double ox, double oy, double oz,
                      float kappa) {
    int bx, by, bz;
    float tx, ty, tz;
    split_coord(py - oy, by, ty);
    if (method == IM_1DY_LINEAR) {
        // column of the containing cell: floor(p - (o - 1/2))
        const int cx = (int)floor(px - ox + 0.5), cz = (int)floor(pz - oz + 0.5);
        return linear1dy<V>(ld, cx, by, cz, ty);
    }
    if (method == IM_1DY_WENO6) {
        const int cx = (int)floor(px), cz = (int)floor(pz);
        return weno1dy<V>(ld, cx, by, cz, ty, kappa);
    }
    split_coord(px - ox, bx, tx);
    split_coord(pz - oz, bz, tz);
    if (method == IM_LINEAR) return linear3d<V>(ld, bx, by, bz, tx, ty, tz);
    return weno3d<V>(ld, bx, by, bz, tx, ty, tz, kappa);
}

}  // namespace fia
