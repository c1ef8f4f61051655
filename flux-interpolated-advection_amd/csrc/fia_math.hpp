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

// delta = WENO6(f)(t) - f_2 from the five first differences of the six samples.
FIA_HD float weno6_delta(float g0, float g1, float g2, float g3, float g4, float t, float kappa) {
    const float d0 = t + 2.0f, d1 = t + 1.0f, d2 = t, d3 = t - 1.0f, d4 = t - 2.0f, d5 = t - 3.0f;
    // Newton-form cubic candidates on samples {0..3}, {1..4}, {2..5}, relative to f_2
    const float q0 = -(g0 + g1) + g0 * d0 + (g1 - g0) * (d0 * d1 * 0.5f)
                   + (g2 - 2.0f * g1 + g0) * (d0 * d1 * d2 * (1.0f / 6.0f));
    const float q1 = -g1 + g1 * d1 + (g2 - g1) * (d1 * d2 * 0.5f)
                   + (g3 - 2.0f * g2 + g1) * (d1 * d2 * d3 * (1.0f / 6.0f));
    const float q2 = g2 * d2 + (g3 - g2) * (d2 * d3 * 0.5f)
                   + (g4 - 2.0f * g3 + g2) * (d2 * d3 * d4 * (1.0f / 6.0f));
    const float c0 = d4 * d5 * (1.0f / 20.0f);
    const float c1 = -d5 * d0 * (1.0f / 10.0f);
    const float c2 = d0 * d1 * (1.0f / 20.0f);
    const float k = kappa * (1.0f / 180.0f);
    const float s0 = k * (244.0f * g0 * g0 + 1561.0f * g1 * g1 + 814.0f * g2 * g2
                          - 1171.0f * g0 * g1 + 683.0f * g0 * g2 - 1951.0f * g1 * g2);
    const float s1 = k * (244.0f * g1 * g1 + 961.0f * g2 * g2 + 244.0f * g3 * g3
                          - 781.0f * g1 * g2 + 293.0f * g1 * g3 - 781.0f * g2 * g3);
    const float s2 = k * (814.0f * g2 * g2 + 1561.0f * g3 * g3 + 244.0f * g4 * g4
                          - 1951.0f * g2 * g3 + 683.0f * g2 * g4 - 1171.0f * g3 * g4);
    const float e0 = 1e-6f + s0, e1 = 1e-6f + s1, e2 = 1e-6f + s2;
    const float a0 = c0 * fast_rcp(e0 * e0);
    const float a1 = c1 * fast_rcp(e1 * e1);
    const float a2 = c2 * fast_rcp(e2 * e2);
    return (a0 * q0 + a1 * q1 + a2 * q2) * fast_rcp(a0 + a1 + a2);
}

// WENO6 of six samples of type V (float, or double for the y-prefix psi).
template <class V> FIA_HD V weno6(const V f[6], float t, float kappa) {
    const float g0 = (float)(f[1] - f[0]), g1 = (float)(f[2] - f[1]), g2 = (float)(f[3] - f[2]);
    const float g3 = (float)(f[4] - f[3]), g4 = (float)(f[5] - f[4]);
    return f[2] + (V)weno6_delta(g0, g1, g2, g3, g4, t, kappa);
}

// Clamped loader of a staggered field.
template <class V> struct Loader {
    const V *__restrict__ d;
    int mx, my, mz;
    FIA_HD V operator()(int i, int j, int k) const {
        return d[idx3(clampi(i, 0, mx - 1), clampi(j, 0, my - 1), clampi(k, 0, mz - 1), my, mz)];
    }
};

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
