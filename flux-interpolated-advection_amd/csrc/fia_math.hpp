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
// in the differences (a common squared first difference plus forms of second differences).
// Evaluated that way nothing cancels when the samples are large and close together (the
// y-prefix psi reaches n*q at the top of a column), so psi can be interpolated with FP32
// arithmetic around an FP64 base.  kappa rescales S to the reference's physical units
// (S_phys = kappa * S) so the 1e-6 epsilon keeps its meaning.
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

// p for every lane of the wave (the host replays run one lane at a time: p itself)
FIA_HD bool wave_all(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __all(p);
#else
    return p;
#endif
}

// Coefficients of one WENO6 evaluation that depend only on the fraction t (shared by every
// evaluation of a pass).  With d_i = t + 2 - i the ideal weights are C0 = d4 d5 / 20,
// C1 = -d5 d0 / 10, C2 = d0 d1 / 20 (all positive on [0,1]).  Candidate r (the Newton-form cubic
// on samples r..r+3) equals f_2 + t g_2 + A h_r + B_r D_r with the second and third differences
// h_k = g_{k+1} - g_k, D_k = h_{k+1} - h_k, A = t (t-1) / 2 and B_0 = t (t-1) (t+4) / 6,
// B_1 = t (t-1) (t+1) / 6, B_2 = t (t-1) (t-2) / 6.  eps is the reference's 1e-6 in the units of
// the indicators below (S_phys = kappa S / 180); kappa in (0, 1e3] keeps eps^4 a normal float.
// eps2 is the 1e-30 that the "WENO6-optimized" variant adds to S_phys^2 (reference
// src/weno.ispc:10-51), in the squared units of the same indicators.
struct WenoCoef {
    float t, A, B0, B1, B2, c0, c1, c2, eps, eps2;
    FIA_HD explicit WenoCoef(float tt, float kappa) {
        const float p = tt * (tt - 1.0f);
        t = tt;
        A = 0.5f * p;
        B0 = p * (tt + 4.0f) * (1.0f / 6.0f);
        B1 = p * (tt + 1.0f) * (1.0f / 6.0f);
        B2 = p * (tt - 2.0f) * (1.0f / 6.0f);
        c0 = (tt - 2.0f) * (tt - 3.0f) * (1.0f / 20.0f);
        c1 = (3.0f - tt) * (tt + 2.0f) * (1.0f / 10.0f);
        c2 = (tt + 2.0f) * (tt + 1.0f) * (1.0f / 20.0f);
        eps = 1.8e-4f * fast_rcp(kappa);
        const float s = 180.0f * fast_rcp(kappa);
        eps2 = 1e-30f * (s * s);
    }
};

// delta = WENO6(f)(t) - f_2 from the five first differences g_k = f_{k+1} - f_k.
//
// The reference's smoothness indicators (x180), quadratic forms of the four samples of each
// candidate, are in these differences the common 180 g_2^2 plus a form of two second differences:
//   S_0 = 180 g_2^2 + 244 h_0^2 - 683 h_0 h_1 + 634 h_1^2
//   S_1 = 180 g_2^2 + 244 h_1^2 - 293 h_1 h_2 + 244 h_2^2
//   S_2 = 180 g_2^2 + 634 h_2^2 - 683 h_2 h_3 + 244 h_3^2
// The weights C_r / (1e-6 + S_phys,r)^2 are normalised, so the factor kappa / 180 of S_phys
// cancels and e_r = S_r + eps stands in for 1e-6 + S_phys,r.  OPT: the "WENO6-optimized" weights
// C_r / (1e-30 + S_phys,r^2) (reference src/weno.ispc:46-50, whose rcp is taken as exact here),
// with three reciprocals: e_r^2 + eps2 can be as small as eps2.
template <bool OPT = false>
FIA_HD float weno6_delta(float g0, float g1, float g2, float g3, float g4, const WenoCoef &w) {
    const float h0 = g1 - g0, h1 = g2 - g1, h2 = g3 - g2, h3 = g4 - g3;
    const float D0 = h1 - h0, D1 = h2 - h1, D2 = h3 - h2;
    const float u0 = w.A * h0 + w.B0 * D0, u1 = w.A * h1 + w.B1 * D1, u2 = w.A * h2 + w.B2 * D2;
    const float p00 = h0 * h0, p11 = h1 * h1, p22 = h2 * h2, p33 = h3 * h3;
    const float p01 = h0 * h1, p12 = h1 * h2, p23 = h2 * h3;
    const float base = 180.0f * (g2 * g2) + (OPT ? 0.0f : w.eps);
    const float e0 = base + 244.0f * p00 - 683.0f * p01 + 634.0f * p11;
    const float e1 = base + 244.0f * p11 - 293.0f * p12 + 244.0f * p22;
    const float e2 = base + 634.0f * p22 - 683.0f * p23 + 244.0f * p33;
    const float lin = w.t * g2;
    if constexpr (OPT) {
        const float a0 = w.c0 * fast_rcp(e0 * e0 + w.eps2), a1 = w.c1 * fast_rcp(e1 * e1 + w.eps2),
                    a2 = w.c2 * fast_rcp(e2 * e2 + w.eps2);
        return lin + (a0 * u0 + a1 * u1 + a2 * u2) * fast_rcp(a0 + a1 + a2);
    }
    const float E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
    if (fmaxf(e0, fmaxf(e1, e2)) < 1e9f) {
        // weights C_r / e_r^2 scaled by E0 E1 E2 (one reciprocal): for eps <= e < 1e9 every
        // product stays between eps^4 and 1e36
        const float a0 = w.c0 * (E1 * E2), a1 = w.c1 * (E0 * E2), a2 = w.c2 * (E0 * E1);
        return lin + (a0 * u0 + a1 * u1 + a2 * u2) * fast_rcp(a0 + a1 + a2);
    }
    const float a0 = w.c0 * fast_rcp(E0), a1 = w.c1 * fast_rcp(E1), a2 = w.c2 * fast_rcp(E2);
    return lin + (a0 * u0 + a1 * u1 + a2 * u2) * fast_rcp(a0 + a1 + a2);
}

// WENO6 of six samples of type V (float, or double for the y-prefix psi).
template <class V, bool OPT = false> FIA_HD V weno6(const V f[6], const WenoCoef &w) {
    const float g0 = (float)(f[1] - f[0]), g1 = (float)(f[2] - f[1]), g2 = (float)(f[3] - f[2]);
    const float g3 = (float)(f[4] - f[3]), g4 = (float)(f[5] - f[4]);
    return f[2] + (V)weno6_delta<OPT>(g0, g1, g2, g3, g4, w);
}
template <class V, bool OPT = false> FIA_HD V weno6(const V f[6], float t, float kappa) {
    return weno6<V, OPT>(f, WenoCoef(t, kappa));
}

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
template <bool SX, bool SY, bool SZ, bool OPT, class V, class L>
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
                zr[iy] = weno6<V, OPT>(f, tz, kappa);
            }
        }
        yv[ix] = SY ? zr[2] : weno6<V, OPT>(zr, ty, kappa);
    }
    return SX ? yv[2] : weno6<V, OPT>(yv, tx, kappa);
}

// Runtime dispatch on which passes sit on a node (uniform inside a wave for lattice vertices).
template <class V, bool OPT = false, class L>
FIA_HD V weno3d(const L &ld, int bx, int by, int bz, float tx, float ty, float tz, float kappa) {
    const int m = (tx == 0.0f ? 4 : 0) | (ty == 0.0f ? 2 : 0) | (tz == 0.0f ? 1 : 0);
    switch (m) {
        case 0: return weno3d_t<false, false, false, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 1: return weno3d_t<false, false, true, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 2: return weno3d_t<false, true, false, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 3: return weno3d_t<false, true, true, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 4: return weno3d_t<true, false, false, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 5: return weno3d_t<true, false, true, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
        case 6: return weno3d_t<true, true, false, OPT, V>(ld, bx, by, bz, tx, ty, tz, kappa);
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

// floor(s) as an index, held to |b| <= 2^24 so that b plus a stencil offset never overflows an
// int (a huge but finite s, such as the centroid of a sliver piece, would otherwise convert to
// INT_MAX and wrap in the interior tests of the stencil loads).  A stencil that far outside a
// field clamps every sample to the same edge sample, so the value is that of the exact floor.
FIA_HD int floor_index(double s) { return (int)fmin(fmax(floor(s), -16777216.0), 16777216.0); }

// floor split of a double sample coordinate into base and fraction
FIA_HD void split_coord(double s, int &b, float &t) {
    const double f = floor(s);
    b = floor_index(s);
    t = (float)(s - f);
}

// The "WENO6-optimized" rule's split along one axis of m samples (reference
// src/sginterp3.cpp:254-258 and 299-301): the coordinate held to the samples' range [0, m - 1],
// the base to [0, m - 2] and the fraction to [0, 1].  Beyond the last sample the value is that
// sample's (the candidates all pass through it at t = 1), where the WENO6 rule extrapolates from
// the clamped samples.
FIA_HD void split_coord_clamped(double s, int m, int &b, float &t) {
    const double c = fmin(fmax(s, 0.0), (double)(m - 1));
    const int f = (int)floor(c);
    b = f < m - 2 ? f : (m >= 2 ? m - 2 : 0);
    t = (float)(c - (double)b);
}

// IM_WENO6_OPT: "WENO6-optimized", the quadrature-point interpolation of the reference's ISPC path
// (src/sginterp3.cpp:203-383, src/sginterp3.ispc:79-115): passes z, y, x as WENO6, with the
// clamped split above and the weights of weno6_delta<true>.
enum InterpMethod { IM_LINEAR = 0, IM_WENO6 = 1, IM_1DY_LINEAR = 2, IM_1DY_WENO6 = 3, IM_WENO6_OPT = 4 };

// Interpolation of a staggered field at p (cell units, in the same frame as the field's
// offsets o, which are in cells).  o_col is the physical offset used by the 1Dy-WENO6 column
// rule (floor(p_x/h), reference src/WENO6_interpolation_1d.cpp:320-322).
template <class V, class L>
FIA_HD V interp_point(const L &ld, int method, double px, double py, double pz, double ox, double oy, double oz,
                      float kappa) {
    int bx, by, bz;
    float tx, ty, tz;
    if (method == IM_WENO6_OPT) {
        split_coord_clamped(px - ox, ld.mx, bx, tx);
        split_coord_clamped(py - oy, ld.my, by, ty);
        split_coord_clamped(pz - oz, ld.mz, bz, tz);
        return weno3d<V, true>(ld, bx, by, bz, tx, ty, tz, kappa);
    }
    split_coord(py - oy, by, ty);
    if (method == IM_1DY_LINEAR) {
        // column of the containing cell: floor(p - (o - 1/2))
        const int cx = floor_index(px - ox + 0.5), cz = floor_index(pz - oz + 0.5);
        return linear1dy<V>(ld, cx, by, cz, ty);
    }
    if (method == IM_1DY_WENO6) {
        const int cx = floor_index(px), cz = floor_index(pz);
        return weno1dy<V>(ld, cx, by, cz, ty, kappa);
    }
    split_coord(px - ox, bx, tx);
    split_coord(pz - oz, bz, tz);
    if (method == IM_LINEAR) return linear3d<V>(ld, bx, by, bz, tx, ty, tz);
    return weno3d<V>(ld, bx, by, bz, tx, ty, tz, kappa);
}

}  // namespace fia
