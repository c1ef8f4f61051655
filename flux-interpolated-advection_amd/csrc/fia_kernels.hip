// fia_kernels.hip - gfx950 kernels of the flux-interpolated advection step.
//
// One flux pass over a component grid G (density, or one staggered velocity component):
//   psi_kernel        y-prefix psi on the y-faces of G: 64 z-columns of one x plane per block, the
//                     column's y range in eight segments (one wave each), FP64 sums
//                     (psi_weno_held_kernel: the WENO6 integral rule with the segment held)
//   vertex_kernel     backtraced displacement of every lattice vertex, one thread per vertex
//   face_count_kernel slots of every face (x strips x z bands); twisted and degenerate faces listed
//   scan_*            exclusive scan of the slot counts (three kernels over tiles)
//   seg_face_kernel   boundary-segment cutter, one thread per face (multi-cut faces piece by piece)
//   degenerate_kernel clipping pipeline for faces with a corner colour on an interior integer
//   zero_velocity_face_kernel  the zero-velocity correction pass: the undeformed y faces' pieces
//                     in closed form (at most 2 x 2 axis-aligned rectangles per face)
//   piece_kernel      quadrature of every piece slot (WENO6 integral rule: piece_kernel<1>)
//   face_sum_kernel   face flux = ordered sum of its slots
//   general_*         the exact recursive cutter for the few twisted faces: pieces counted and
//                     stored per face, integrated one thread per piece, summed per face; their
//                     corrections are marked in a bit mask per family (corr_reset_kernel clears
//                     them before the flux set's next pass)
//   gather_kernel     per cell: sum of its six signed face fluxes (deterministic, no atomics)
// Each interior face is computed once; its flux enters the two neighbouring cells with opposite
// signs, so nothing is lost or created inside the domain.
#include "fia_cutter.hpp"
#include "fia_internal.hpp"
#include "fia_math.hpp"

namespace fia {

constexpr int kBlock = 256;

static inline unsigned grid_for(size_t n, int block) { return (unsigned)((n + block - 1) / block); }
// a tunable block size: 64, 128 or 256 threads (the kernels' launch bound), anything else 256
static inline int block_or_default(int b) { return (b == 64 || b == 128) ? b : kBlock; }

// Every kernel is an index computation around a __host__ __device__ body, so the bodies can also
// run over host arrays (a CPU emulation of the kernels used to check their logic, on host threads).
// Their atomics are the compiler's host atomics in that build.
FIA_HD int atomic_add_i32(int *p, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}
FIA_HD void atomic_or_u32(unsigned *p, unsigned v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    __atomic_fetch_or(p, v, __ATOMIC_RELAXED);
#endif
}
FIA_HD void atomic_cas_i32(int *p, int cmp, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicCAS(p, cmp, v);
#else
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
#endif
}

// ---------------------------------------------------------------------------------------------
// psi (reference src/calc_psi_3D.cpp:10-308 density, :311-537 velocity)
// ---------------------------------------------------------------------------------------------
// Increment of psi across cell jj of column (i, k): the integral of q over the cell in y.
FIA_HD float psi_increment(const PsiParams &P, const Loader<float> &Q, int i, int jj, int k) {
    if (!P.integral || P.mode == 0) return Q(i, jj, k);
    if (P.mode == 1) {
        // trilinear at (x_c, jj + 1/4, z_c) and (x_c, jj + 3/4, z_c): x, z on nodes
        const float a = 0.25f * Q(i, jj - 1, k) + 0.75f * Q(i, jj, k);   // sample jj - 1/4
        const float b = 0.75f * Q(i, jj, k) + 0.25f * Q(i, jj + 1, k);   // sample jj + 1/4
        return 0.5f * a + 0.5f * b;
    }
    // Gauss-Legendre on [jj, jj+1/2] and [jj+1/2, jj+1] over WENO6 in y
    float inc = 0.0f;
    for (int half = 0; half < 2; ++half) {
        const double a = jj + 0.5 * half;
        for (int gq = 0; gq < P.gauss.n; ++gq) {
            const double y = 0.25 * P.gauss.x[gq] + 0.25 + a;   // G coords
            int by; float ty;
            split_coord(y - 0.5, by, ty);
            inc += 0.5f * (float)P.gauss.w[gq] * weno1dy<float>(Q, i, by, k, ty, P.kappa);
        }
    }
    return inc;
}

// The y-prefix of a column is summed in kPsiSegs segments of its cells: segment w covers cells
// [w S, min(my, (w + 1) S)), S = ceil(my / kPsiSegs); psi at the top of a cell of segment w is the
// sum of the totals of segments 0..w-1 (in that order) plus the running sum inside segment w.
constexpr int kPsiSegs = 8;
FIA_HD void psi_segment(int my, int w, int &j0, int &j1) {
    const int seg = (my + kPsiSegs - 1) / kPsiSegs;
    j0 = min(my, w * seg);
    j1 = min(my, j0 + seg);
}

// The WENO6 integral increments of cells [j0, j1) with the seven rows they read held in registers
// and slid up the column: every Gauss point of the lower half of cell jj has its stencil base at
// row jj - 1 and every one of the upper half at row jj (the points lie strictly inside the halves),
// so the two halves read rows jj-3..jj+2 and jj-2..jj+3, and each cell loads one new row.  The
// samples, fractions and sums are psi_increment's, in its order.  WRITE: psi above each cell from
// `off`; returns the running sum (the segment total from off = 0).
template <bool WRITE>
FIA_HD double psi_weno_segment(const PsiParams &P, const Loader<float> &Q, int i, int k, int j0, int j1, double off) {
    const int my = P.g.m[1], mz = P.g.m[2];
    double *col = P.psi + idx3(i, 0, k, my + 1, mz);
    float win[7];
#pragma unroll
    for (int r = 0; r < 7; ++r) win[r] = Q(i, j0 - 3 + r, k);
    double acc = off;
    for (int jj = j0; jj < j1; ++jj) {
        float inc = 0.0f;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const double a = jj + 0.5 * half;
            for (int gq = 0; gq < P.gauss.n; ++gq) {
                const double y = 0.25 * P.gauss.x[gq] + 0.25 + a;   // G coords
                int by; float ty;
                split_coord(y - 0.5, by, ty);
                inc += 0.5f * (float)P.gauss.w[gq] * weno6<float>(&win[half], WenoCoef(ty, P.kappa));
            }
        }
        acc += (double)inc;
        if (WRITE) col[(size_t)(jj + 1) * mz] = acc;
#pragma unroll
        for (int r = 0; r < 6; ++r) win[r] = win[r + 1];
        win[6] = Q(i, jj + 4, k);
    }
    return acc;
}

FIA_HD double psi_segment_total(const PsiParams &P, const Loader<float> &Q, int i, int k, int j0, int j1) {
    if (P.integral && P.mode == 2) return psi_weno_segment<false>(P, Q, i, k, j0, j1, 0.0);
    double s = 0.0;
#pragma unroll 8
    for (int jj = j0; jj < j1; ++jj) s += (double)psi_increment(P, Q, i, jj, k);
    return s;
}
// writes psi on the faces above the cells of [j0, j1) from `off`, the psi of face j0
FIA_HD void psi_segment_write(const PsiParams &P, const Loader<float> &Q, int i, int k, int j0, int j1, double off) {
    const int my = P.g.m[1], mz = P.g.m[2];
    double *col = P.psi + idx3(i, 0, k, my + 1, mz);
    if (j0 == 0) col[0] = 0.0;
    if (P.integral && P.mode == 2) {
        psi_weno_segment<true>(P, Q, i, k, j0, j1, off);
        return;
    }
    double acc = off;
#pragma unroll 8
    for (int jj = j0; jj < j1; ++jj) {
        acc += (double)psi_increment(P, Q, i, jj, k);
        col[(size_t)(jj + 1) * mz] = acc;
    }
}

// psi of column (i, k) with the segment order of psi_kernel (host replay of the kernel)
FIA_HD void psi_body(const PsiParams &P, long colid) {
    const int my = P.g.m[1], mz = P.g.m[2];
    const int i = (int)(colid / mz), k = (int)(colid % mz);
    const Loader<float> Q{P.q, P.g.m[0], my, mz};
    double tot[kPsiSegs];
    for (int w = 0; w < kPsiSegs; ++w) {
        int j0, j1;
        psi_segment(my, w, j0, j1);
        tot[w] = psi_segment_total(P, Q, i, k, j0, j1);
    }
    double off = 0.0;
    for (int w = 0; w < kPsiSegs; ++w) {
        int j0, j1;
        psi_segment(my, w, j0, j1);
        psi_segment_write(P, Q, i, k, j0, j1, off);
        off += tot[w];
    }
}

// One block per 64 z-columns of one x plane: lanes are consecutive z (coalesced rows of q and
// psi), the block's kPsiSegs waves take the y segments, so a 256^3 grid runs 8,192 waves instead
// of one thread per column.  Each wave sums its segment's increments, the totals meet in LDS, and
// each wave writes its segment's prefix from the sum of the segments below it.  Segments of at
// most kPsiHeld cells under the cheap rules (const, linear) keep their increments in registers
// between the two steps (q is read once), and so does the WENO6 integral rule with one Gauss point
// per half cell (psi_weno_held_kernel, below); otherwise (longer segments, more Gauss points) they
// are computed again.
constexpr int kPsiHeld = 33;   // 33 cells: segments of 257-row columns (a y-velocity grid of 256 cells)
// HELD: the increments stay in registers (cheap rules, segments of at most kPsiHeld cells); a
// template parameter, so the other form keeps its own register budget
FIA_HD bool psi_held(const PsiParams &P) {
    return !(P.integral && P.mode == 2) && (P.g.m[1] + kPsiSegs - 1) / kPsiSegs <= kPsiHeld;
}
template <bool HELD> __global__ void __launch_bounds__(64 * kPsiSegs) psi_kernel(PsiParams P) {
    const int my = P.g.m[1], mz = P.g.m[2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.y, k = blockIdx.x * 64 + lane;
    const bool own = k < mz;
    const Loader<float> Q{P.q, P.g.m[0], my, mz};
    int j0, j1;
    psi_segment(my, w, j0, j1);
    __shared__ double tot[kPsiSegs][64];
    if constexpr (HELD) {
        float inc[kPsiHeld];
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < kPsiHeld; ++r) {
            inc[r] = (own && j0 + r < j1) ? psi_increment(P, Q, i, j0 + r, k) : 0.0f;
            s += (double)inc[r];
        }
        tot[w][lane] = s;
        __syncthreads();
        if (!own) return;
        double acc = 0.0;
        for (int b = 0; b < w; ++b) acc += tot[b][lane];
        double *col = P.psi + idx3(i, 0, k, my + 1, mz);
        if (j0 == 0) col[0] = 0.0;
#pragma unroll
        for (int r = 0; r < kPsiHeld; ++r) {
            if (j0 + r < j1) {
                acc += (double)inc[r];
                col[(size_t)(j0 + r + 1) * mz] = acc;
            }
        }
    } else {
        tot[w][lane] = own ? psi_segment_total(P, Q, i, k, j0, j1) : 0.0;
        __syncthreads();
        if (!own) return;
        double off = 0.0;
        for (int b = 0; b < w; ++b) off += tot[b][lane];
        psi_segment_write(P, Q, i, k, j0, j1, off);
    }
}

// The WENO6 integral rule in held form (segments of at most kPsiHeld cells, kPsiWenoGauss Gauss
// point per half cell, the default rule): the rows of the segment and its six neighbours are
// loaded at once, the increments are computed once (not again for the prefix) and kept in
// registers.  Each increment is psi_weno_segment's, with the same operations in the same order.
constexpr int kPsiWenoGauss = 1;
FIA_HD bool psi_weno_held(const PsiParams &P) {
    return P.integral && P.mode == 2 && (P.g.m[1] + kPsiSegs - 1) / kPsiSegs <= kPsiHeld &&
           P.gauss.n <= kPsiWenoGauss;
}
// bound to 128 VGPRs (three values spilled): two 8-wave blocks per CU instead of one (0.127 ->
// 0.106 ms per launch on the 256^3 velocity grids)
__global__ void __launch_bounds__(64 * kPsiSegs) __attribute__((amdgpu_waves_per_eu(4))) psi_weno_held_kernel(PsiParams P) {
    const int my = P.g.m[1], mz = P.g.m[2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.y, k = blockIdx.x * 64 + lane;
    const bool own = k < mz;
    const Loader<float> Q{P.q, P.g.m[0], my, mz};
    int j0, j1;
    psi_segment(my, w, j0, j1);
    __shared__ double tot[kPsiSegs][64];
    // rows j0 - 3 .. j0 + kPsiHeld + 2 of the column (clamped), all loads issued before any use
    float rows[kPsiHeld + 6];
#pragma unroll
    for (int r = 0; r < kPsiHeld + 6; ++r) rows[r] = Q(i, j0 - 3 + r, k);
    float inc[kPsiHeld];
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < kPsiHeld; ++r) {
        const int jj = j0 + r;
        float v = 0.0f;
        if (own && jj < j1) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const double a = jj + 0.5 * half;
                for (int gq = 0; gq < P.gauss.n; ++gq) {
                    const double y = 0.25 * P.gauss.x[gq] + 0.25 + a;   // G coords
                    int by; float ty;
                    split_coord(y - 0.5, by, ty);
                    v += 0.5f * (float)P.gauss.w[gq] * weno6<float>(&rows[r + half], WenoCoef(ty, P.kappa));
                }
            }
        }
        inc[r] = v;
        s += (double)v;
    }
    tot[w][lane] = s;
    __syncthreads();
    if (!own) return;
    double acc = 0.0;
    for (int b = 0; b < w; ++b) acc += tot[b][lane];
    double *col = P.psi + idx3(i, 0, k, my + 1, mz);
    if (j0 == 0) col[0] = 0.0;
#pragma unroll
    for (int r = 0; r < kPsiHeld; ++r) {
        if (j0 + r < j1) {
            acc += (double)inc[r];
            col[(size_t)(j0 + r + 1) * mz] = acc;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// vertex backtrace (reference src/calc_backtraced_face_3D.cpp:9-159)
// ---------------------------------------------------------------------------------------------
FIA_HD void vertex_body(const VertexParams &P, long id) {
    const int ly = P.g.m[1] + 1, lz = P.g.m[2] + 1;
    const int k = (int)(id % lz), j = (int)((id / lz) % ly), i = (int)(id / ((long)ly * lz));
    const int lat[3] = {i, j, k};
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = lat[a] - 0.5 * P.g.e[a];
    bool pinned = false;
    if (P.g.comp >= 0) {
        const int li = lat[P.g.comp], n = P.g.n[P.g.comp];
        pinned = li <= 1 || li >= n;
    }
    float d[3] = {0.0f, 0.0f, 0.0f};
    if (!pinned) {
        const int nx = P.g.n[0], ny = P.g.n[1], nz = P.g.n[2];
        const Loader<float> U{P.u, nx + 1, ny, nz}, V{P.v, nx, ny + 1, nz}, W{P.w, nx, ny, nz + 1};
        const float vu = interp_point<float>(U, P.interp, p[0], p[1], p[2], 0.0, 0.5, 0.5, 1.0f);
        const float vv = interp_point<float>(V, P.interp, p[0], p[1], p[2], 0.5, 0.0, 0.5, 1.0f);
        const float vw = interp_point<float>(W, P.interp, p[0], p[1], p[2], 0.5, 0.5, 0.0, 1.0f);
        d[0] = -P.dt_over_h * vu;
        d[1] = -P.dt_over_h * vv;
        d[2] = -P.dt_over_h * vw;
    }
    for (int a = 0; a < 3; ++a) {
        // clamp the position to [-e_a/2, n_a + e_a/2] (reference :131-159) as a bound on d
        const float lo = (float)(-0.5 * P.g.e[a] - p[a]);
        const float hi = (float)(P.g.n[a] + 0.5 * P.g.e[a] - p[a]);
        float v = d[a];
        if (!(v >= lo)) v = (v != v) ? v : lo;   // NaN passes through and is caught by the face kernel
        if (v > hi) v = hi;
        P.disp[a][id] = v;
    }
}
__global__ void __launch_bounds__(kBlock) vertex_kernel(VertexParams P) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)(P.g.m[0] + 1) * (P.g.m[1] + 1) * (P.g.m[2] + 1)) return;
    vertex_body(P, id);
}

// ---------------------------------------------------------------------------------------------
// Face setup: corners of face (fam, i, j, k) of G in "+" orientation, backtraced, in the
// physical frame (cell units), plus corner colours.  Orientation of the reference's faces 1, 2, 5
// (src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:83-112).
// ---------------------------------------------------------------------------------------------
FIA_HD void face_corner_offsets(int fam, int c, int o[3]) {
    // x-face: (j,k),(j+1,k),(j+1,k+1),(j,k+1); y-face: (i,k),(i,k+1),(i+1,k+1),(i+1,k);
    // z-face: (i,j),(i+1,j),(i+1,j+1),(i,j+1)
    const int a = (c == 1 || c == 2), b = (c == 2 || c == 3);
    o[0] = o[1] = o[2] = 0;
    if (fam == 0) { o[1] = a; o[2] = b; }
    else if (fam == 1) { o[2] = a; o[0] = b; }
    else { o[0] = a; o[1] = b; }
}

FIA_HD bool load_face(const FaceParams &P, int fam, int i, int j, int k, double v[4][3],
                                          double col[4][2]) {
    const int ly = P.g.m[1] + 1, lz = P.g.m[2] + 1;
    // the twelve displacements first, in one block (loaded corner by corner under the test of
    // the pointer, each corner's loads were waited for before the next corner's were issued)
    const bool moving = P.disp[0] != nullptr;
    float d[4][3];
    if (moving) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            int o[3];
            face_corner_offsets(fam, c, o);
            const size_t vi = idx3(i + o[0], j + o[1], k + o[2], ly, lz);
#pragma unroll
            for (int a = 0; a < 3; ++a) d[c][a] = P.disp[a][vi];
        }
    }
    bool finite = true;
    for (int c = 0; c < 4; ++c) {
        int o[3];
        face_corner_offsets(fam, c, o);
        const int li = i + o[0], lj = j + o[1], lk = k + o[2];
        v[c][0] = li - 0.5 * P.g.e[0];
        v[c][1] = lj - 0.5 * P.g.e[1];
        v[c][2] = lk - 0.5 * P.g.e[2];
        if (moving) {
            v[c][0] += (double)d[c][0];
            v[c][1] += (double)d[c][1];
            v[c][2] += (double)d[c][2];
        }
        finite = finite & std::isfinite(v[c][0]) & std::isfinite(v[c][1]) & std::isfinite(v[c][2]);
        // colours (reference src/split_face_3D.cpp:26-43)
        const double x = v[c][0], z = v[c][2];
        double cx = 0.0, cz = 0.0;
        switch (P.split) {
            case 1: cx = x - 0.5; cz = z - 0.5; break;
            case 2: cx = x; cz = z; break;
            case 3: cx = x - 0.5; cz = z; break;
            case 4: cx = x; cz = z - 0.5; break;
            default: break;
        }
        col[c][0] = cx; col[c][1] = cz;
    }
    return finite;
}

// Flux contribution of one piece with centre c (physical frame, cells) and y vector area ay
// (cells^2), in units of q (reference quadrature src/polygon_3D.cpp:57-77 and 257-487,
// consumers src/calc_density_in_cell_from_quadrature_point_list_for_integrate_density.cpp:119-201
// and src/calc_density_in_cell_from_quadrature_point_list.cpp:98-132).
FIA_HD double piece_flux(const FaceParams &P, const double c[3], double ay, int &npts) {
    const int mx = P.g.m[0], my = P.g.m[1], mz = P.g.m[2];
    const double ex = 0.5 * P.g.e[0], ey = 0.5 * P.g.e[1], ez = 0.5 * P.g.e[2];
    // field offsets in the physical frame: q at G cell centres, psi on G y-faces
    const double oqx = 0.5 - ex, oqy = 0.5 - ey, oqz = 0.5 - ez;
    const double opx = 0.5 - ex, opy = -ey, opz = 0.5 - ez;
    const Loader<double> PS{P.psi, mx, my + 1, mz};
    // the quadrature points' interpolation ("WENO6-optimized": the ISPC variant)
    const int im = (P.interp == IM_WENO6 && P.weno_opt) ? IM_WENO6_OPT : P.interp;
    if (!P.integral) {
        return ay * interp_point<double>(PS, im, c[0], c[1], c[2], opx, opy, opz, P.kappa_psi);
    }
    const Loader<float> Q{P.q, mx, my, mz};
    const double gy = c[1] + ey;             // G coordinate along y
    const double jf = floor(gy);
    const double f = gy - jf;
    const double level = jf - ey;            // physical-frame y of G face level j
    const double psi = interp_point<double>(PS, im, c[0], level, c[2], opx, opy, opz, P.kappa_psi);
    float dens = 0.0f;
    const int nh = f < 0.5 ? 1 : 2;
    for (int half = 0; half < nh; ++half) {
        double a, len, frac;
        if (nh == 1) { a = level; len = f; frac = f; }
        else if (half == 0) { a = level; len = 0.5; frac = 0.5; }
        else { a = level + 0.5; len = f - 0.5; frac = f - 0.5; }
        for (int gq = 0; gq < P.gauss.n; ++gq) {
            const double y = (0.5 * len) * P.gauss.x[gq] + 0.5 * len + a;
            const float val = interp_point<float>(Q, im, c[0], y, c[2], oqx, oqy, oqz, P.kappa_q);
            dens += val * (float)(frac * P.gauss.w[gq]);
        }
    }
    npts += nh * P.gauss.n;
    return ay * (psi + (double)dens);
}

// WENO6 integral rule (the same quadrature as piece_flux with IM_WENO6 and integral psi) with
// the work shared between the points of a piece.  Every point has the piece's x and z, so the
// sample bases and fractions along x and z are common to psi and q.  psi is evaluated on the G
// face level J = floor(c_y + e_y) where its y sample coordinate is exactly J: only the z and x
// passes run.  The density points of the lower half (sample coordinate in [J - 1/2, J)) use
// sample rows J-3..J+2, those of the upper half rows J-2..J+3, so the z pass runs once over the
// seven rows J-3..J+3 and each point adds its own y and x passes.  Pass order z, y, x and every
// 1D evaluation are those of weno3d, so each value equals the per-point rule's (up to the
// t == 0 skips, which return sample f_2 where this evaluates WENO6 at t = 0).
// OPT: the "WENO6-optimized" interpolation (IM_WENO6_OPT): clamped x and z splits and weights;
// a density point beyond the first or last sample row takes that row's value (the clamped rule's
// result there), read from the z pass's window.
template <int NG, bool OPT>
FIA_HD double piece_flux_weno_integral(const FaceParams &P, const double c[3], double ay, int &npts,
                                                           int &nevals) {
    const int mx = P.g.m[0], my = P.g.m[1], mz = P.g.m[2];
    const double ex = 0.5 * P.g.e[0], ey = 0.5 * P.g.e[1], ez = 0.5 * P.g.e[2];
    int bx, bz;
    float tx, tz;
    if constexpr (OPT) {
        split_coord_clamped(c[0] - (0.5 - ex), mx, bx, tx);
        split_coord_clamped(c[2] - (0.5 - ez), mz, bz, tz);
    } else {
        split_coord(c[0] - (0.5 - ex), bx, tx);
        split_coord(c[2] - (0.5 - ez), bz, tz);
    }
    const double gy = c[1] + ey;
    const double jf = floor(gy);
    const double f = gy - jf;
    const int J = floor_index(jf);
    const bool zin = bz - 2 >= 0 && bz + 3 <= mz - 1;   // the z rows need no clamping
    // Stencils wholly inside the field (the common case) take their rows at 32-bit offsets from
    // the field's base, one add per row; the others are clamped row by row.
    const bool small = (size_t)mx * (size_t)(my + 1) * (size_t)mz * sizeof(double) < ((size_t)1 << 32);
    const bool xin = small && zin && bx - 2 >= 0 && bx + 3 <= mx - 1;
    // psi on level J
    double psi;
    {
        const Loader<double> PS{P.psi, mx, my + 1, mz};
        const WenoCoef wz(tz, P.kappa_psi), wx(tx, P.kappa_psi);
        const bool in = xin && J >= 0 && J <= my;
        const unsigned sx = (unsigned)(my + 1) * (unsigned)mz;
        const unsigned o0 = in ? ((unsigned)(bx - 2) * (unsigned)(my + 1) + (unsigned)J) * (unsigned)mz + (unsigned)(bz - 2)
                               : 0u;
        double xr[6];
        if (wave_all(in)) {
            // every lane's stencil inside the field (the common case): the rows' loads carry no
            // test, so they need not each be waited for before the next row's are issued
#pragma unroll
            for (int ix = 0; ix < 6; ++ix) {
                double fz[6];
                load_row6<double>(P.psi, o0 + (unsigned)ix * sx, fz);
                xr[ix] = weno6<double, OPT>(fz, wz);
            }
        } else {
#pragma unroll
            for (int ix = 0; ix < 6; ++ix) {
                double fz[6];
                if (in) load_row6<double>(P.psi, o0 + (unsigned)ix * sx, fz);
                else load_z6<double>(PS, bx - 2 + ix, J, bz - 2, zin, fz);
                xr[ix] = weno6<double, OPT>(fz, wz);
            }
        }
        psi = weno6<double, OPT>(xr, wx);
    }
    // density points: half h, Gauss node g -> point p = h * NG + g
    const Loader<float> Q{P.q, mx, my, mz};
    const int nh = f < 0.5 ? 1 : 2;
    const double level = jf - ey, oqy = 0.5 - ey;
    const int B = J - 1;   // sample row base of the lower-half points
    float ty[2 * NG], wt[2 * NG];
    int edge[2 * NG];   // OPT: window row of a point beyond the sample rows, else -1
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            double a, len, frac;
            if (nh == 1) { a = level; len = f; frac = f; }
            else if (h == 0) { a = level; len = 0.5; frac = 0.5; }
            else { a = level + 0.5; len = f - 0.5; frac = f - 0.5; }
            const double y = (0.5 * len) * P.gauss.x[g] + 0.5 * len + a;
            const double s = y - oqy;
            ty[h * NG + g] = (float)(s - (double)(B + h));
            wt[h * NG + g] = (float)(frac * P.gauss.w[g]);
            edge[h * NG + g] = -1;
            if constexpr (OPT) {
                if (s < 0.0 || s > (double)(my - 1)) {
                    const int row = s < 0.0 ? 0 : my - 1;
                    edge[h * NG + g] = clampi(row - (B - 2), 0, 6);
                }
            }
        }
    }
    const WenoCoef wz(tz, P.kappa_q), wx(tx, P.kappa_q);
    const bool qin = xin && B - 2 >= 0 && B + 4 <= my - 1;
    const unsigned qsx = (unsigned)my * (unsigned)mz;
    const unsigned q0 = qin ? ((unsigned)(bx - 2) * (unsigned)my + (unsigned)(B - 2)) * (unsigned)mz + (unsigned)(bz - 2)
                            : 0u;
    float yv[2 * NG][6];
#pragma unroll
    for (int ix = 0; ix < 6; ++ix) {
        float zr[7];
#pragma unroll
        for (int r = 0; r < 7; ++r) {
            if (r == 6 && nh == 1) { zr[6] = 0.0f; continue; }
            float fz[6];
            if (qin) load_row6<float>(P.q, q0 + (unsigned)ix * qsx + (unsigned)r * (unsigned)mz, fz);
            else load_z6<float>(Q, bx - 2 + ix, B - 2 + r, bz - 2, zin, fz);
            zr[r] = weno6<float, OPT>(fz, wz);
        }
#pragma unroll
        for (int p = 0; p < 2 * NG; ++p) {
            const int h = p / NG;
            if (h < nh) {
                if (OPT && edge[p] >= 0) yv[p][ix] = zr[edge[p]];
                else yv[p][ix] = weno6<float, OPT>(&zr[h], WenoCoef(ty[p], P.kappa_q));
            }
        }
    }
    float dens = 0.0f;
#pragma unroll
    for (int p = 0; p < 2 * NG; ++p)
        if (p / NG < nh) dens += weno6<float, OPT>(yv[p], wx) * wt[p];
    npts += nh * NG;
    nevals += 7 + 6 * (nh == 2 ? 7 : 6) + 7 * nh * NG;
    return ay * (psi + (double)dens);
}

// face id (< 2^30, checked on the host) -> (i, j, k), id = (i fy + j) fz + k
FIA_HD void decode_face(const CompGrid &g, int fam, long id, int &i, int &j, int &k) {
    const FastDiv &dz = g.fz[fam], &dy = g.fy[fam];
    const uint32_t u = (uint32_t)id, q = dz.div(u), r = dy.div(q);
    k = (int)(u - q * dz.d);
    j = (int)(q - r * dy.d);
    i = (int)r;
}

FIA_HD void report(int *err, int code) { atomic_cas_i32(err, 0, code); }

// Corner colours -> number of piece slots (the product of the x and z strip counts).
FIA_HD void colour_ranges(const double col[4][2], int &x0, int &nxc, int &z0, int &nzc) {
    double lx = col[0][0], hx = col[0][0], lz = col[0][1], hz = col[0][1];
    for (int c = 1; c < 4; ++c) {
        lx = fmin(lx, col[c][0]); hx = fmax(hx, col[c][0]);
        lz = fmin(lz, col[c][1]); hz = fmax(hz, col[c][1]);
    }
    int x1, z1;
    interior_integers(lx, hx, x0, x1);
    interior_integers(lz, hz, z0, z1);
    nxc = x1 >= x0 ? x1 - x0 + 1 : 0;
    nzc = z1 >= z0 ? z1 - z0 + 1 : 0;
}

FIA_HD void queue_general(const FaceParams &P, int fam, long id) {
    const int slot = atomic_add_i32(P.twisted_count, 1);
    P.twisted[slot] = ((unsigned)fam << 30) | (unsigned)id;
}
FIA_HD void queue_degenerate(const FaceParams &P, int fam, long id) {
    const int slot = atomic_add_i32(P.degen_count, 1);
    P.degen[slot] = ((unsigned)fam << 30) | (unsigned)id;
}
// Faces of family FAM whose pieces the one-pass form cuts: at most one cut along each colour a
// face of that family is commonly cut along (y faces: x and z; x faces: z; z faces: x).  Measured
// on a 256^3 curl-noise step at CFL 4 these are 96-98 % of the faces of each family; the others
// are cut piece by piece.
template <int FAM> FIA_HD bool seg_one_pass(int nxc, int nzc) {
    if (FAM == 0) return nxc == 0 && nzc <= 1;
    if (FAM == 2) return nzc == 0 && nxc <= 1;
    return nxc <= 1 && nzc <= 1;
}

// Phase 1: slots each face needs, (x strips) x (z strips).  Faces for the general cutter (twisted
// ones, or all when minimal_area < 1) are queued and get no slots; a non-finite face reports an
// error and gets none (zero flux).  Degenerate faces keep their slots and are queued for the
// clipping pipeline (degenerate_kernel).
template <int FAM> FIA_HD void face_count_body(const FaceParams &P, long id) {
    unsigned n = 0;
    if (P.fam_mask & (1 << FAM)) {
        int i, j, k;
        decode_face(P.g, FAM, id, i, j, k);
        double v[4][3], col[4][2];
        if (!load_face(P, FAM, i, j, k, v, col)) {
            report(P.error, KE_NONFINITE);
        } else if (P.general_all || face_twisted(col)) {
            queue_general(P, FAM, id);
        } else {
            int x0, nxc, z0, nzc;
            colour_ranges(col, x0, nxc, z0, nzc);
            n = (unsigned)((nxc + 1) * (nzc + 1));
            if (face_degenerate(col)) queue_degenerate(P, FAM, id);
        }
    }
    P.face_n[P.fam_base[FAM] + id] = n;
}
template <int FAM> __global__ void __launch_bounds__(kBlock) face_count_kernel(FaceParams P) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)P.g.faces(FAM)) return;
    face_count_body<FAM>(P, id);
}


// Value of a piece under the non-integral rule with linear or 1Dy_linear psi: piece_flux's
// non-integral branch for those two methods, each with its method fixed at compile time.
FIA_HD double piece_value_linear(const FaceParams &P, const double c[3], double ay) {
    const double ex = 0.5 * P.g.e[0], ey = 0.5 * P.g.e[1], ez = 0.5 * P.g.e[2];
    const Loader<double> PS{P.psi, P.g.m[0], P.g.m[1] + 1, P.g.m[2]};
    const double psi = P.interp == IM_LINEAR
                           ? interp_point<double>(PS, IM_LINEAR, c[0], c[1], c[2], 0.5 - ex, -ey, 0.5 - ez, P.kappa_psi)
                           : interp_point<double>(PS, IM_1DY_LINEAR, c[0], c[1], c[2], 0.5 - ex, -ey, 0.5 - ez,
                                                  P.kappa_psi);
    return ay * psi;
}

// Piece record of slot q from its moments: centre (cells, physical frame) and y vector area; an
// empty piece stores a zero area.  With fuse_value the piece's value goes to piece_val instead.
FIA_HD void store_piece(const FaceParams &P, const FaceGeom &geom, double area2, double mxi, double meta, double ay2,
                        unsigned long long q) {
    double ay = 0.0, c3[3] = {0.0, 0.0, 0.0};
    if (area2 != 0.0) {
        const double inv3a = 1.0 / (3.0 * area2);
        double r[3];
        geom.rel(mxi * inv3a, meta * inv3a, r);
        c3[0] = geom.p0[0] + r[0]; c3[1] = geom.p0[1] + r[1]; c3[2] = geom.p0[2] + r[2];
        ay = 0.5 * ay2;
        if (!std::isfinite(c3[0]) || !std::isfinite(c3[1]) || !std::isfinite(c3[2])) ay = 0.0;   // degenerate sliver
    }
    if (P.fuse_value) {
        P.piece_val[q] = ay != 0.0 ? piece_value_linear(P, c3, ay) : 0.0;
        return;
    }
    P.piece_c[0][q] = c3[0];
    P.piece_c[1][q] = c3[1];
    P.piece_c[2][q] = c3[2];
    P.piece_ay[q] = ay;   // a piece with no y area contributes exactly zero
}


// Segment cutter (boundary segments; degenerate faces take the clipping pipeline in
// degenerate_kernel).  One block of kBlock threads per kBlock faces of family FAM:
//  1. each thread cuts all pieces of its face in one pass (face_pieces) when seg_one_pass admits
//     it, into its slots s * (z bands) + m;
//  2. the block's other faces (a few percent) are listed in LDS with their slot counts, and the
//     block's threads share out their pieces, one piece per thread with piece_moments.
// Faces with no slots (general cutter, non-finite, family not computed) are left alone.
template <int FAM> FIA_HD unsigned seg_face_body(const FaceParams &P, long id) {
    constexpr int NX = FAM == 0 ? 0 : 1, NZ = FAM == 2 ? 0 : 1;
    const size_t gid = P.fam_base[FAM] + id;
    const unsigned n = P.face_n[gid];
    if (n == 0) return 0;
    int i, j, k;
    decode_face(P.g, FAM, id, i, j, k);
    double v[4][3], col[4][2];
    load_face(P, FAM, i, j, k, v, col);
    if (face_degenerate(col)) return 0;
    int x0, nxc, z0, nzc;
    colour_ranges(col, x0, nxc, z0, nzc);
    if (!seg_one_pass<FAM>(nxc, nzc)) return n;   // cut piece by piece
    FaceGeom geom;
    geom.init(v);
    PieceAcc acc[2][2];
    face_pieces<NX, NZ>(col, geom, nxc == 1, (double)x0, nzc == 1, (double)z0, acc);
    const unsigned long long s0 = P.face_start[gid];
#pragma unroll
    for (int s = 0; s <= NX; ++s) {
#pragma unroll
        for (int m = 0; m <= NZ; ++m) {
            if (s > nxc || m > nzc) continue;
            const PieceAcc &a = acc[s][m];
            store_piece(P, geom, a.area2, a.mxi, a.meta, a.ay2, s0 + (unsigned long long)(s * (nzc + 1) + m));
        }
    }
    return 0;
}
// piece `local` (slot s * (z bands) + m) of face id of family FAM, on its own
template <int FAM> FIA_HD void seg_piece_body(const FaceParams &P, long id, unsigned local) {
    const size_t gid = P.fam_base[FAM] + id;
    int i, j, k;
    decode_face(P.g, FAM, id, i, j, k);
    double v[4][3], col[4][2];
    load_face(P, FAM, i, j, k, v, col);
    FaceGeom geom;
    geom.init(v);
    int x0, nxc, z0, nzc;
    colour_ranges(col, x0, nxc, z0, nzc);
    const int s = (int)local / (nzc + 1), m = (int)local - s * (nzc + 1);
    SegAcc<false> acc;
    piece_moments(col, geom, s > 0 ? (double)(x0 + s - 1) : -kNoBound, s < nxc ? (double)(x0 + s) : kNoBound,
                  m > 0 ? (double)(z0 + m - 1) : -kNoBound, m < nzc ? (double)(z0 + m) : kNoBound, acc);
    store_piece(P, geom, acc.area2, acc.mxi, acc.meta, acc.ay2, P.face_start[gid] + local);
}
// x and z faces (one cut direction each) fit 128 VGPRs, four waves per SIMD, with one spilled value
// in the piece-by-piece phase (1.03 -> 0.86 ms per launch at 256^3); the y faces (two cut
// directions) need 160 and keep three waves (bound to 128 they spill 126 values)
template <int FAM>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(FAM == 1 ? 1 : 4))) seg_face_kernel(FaceParams P) {
    __shared__ unsigned s_face[kBlock];   // the block's faces cut piece by piece (index in the family)
    __shared__ unsigned s_end[kBlock];    // their slot counts, then the inclusive prefix of them
    __shared__ int s_count;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < (long)P.g.faces(FAM)) {
        const unsigned n = seg_face_body<FAM>(P, id);
        if (n != 0) {
            const int k = atomicAdd(&s_count, 1);
            s_face[k] = (unsigned)id;
            s_end[k] = n;
        }
    }
    __syncthreads();
    const int nm = s_count;
    if (nm == 0) return;   // the same for every thread of the block
    if (threadIdx.x == 0) {
        unsigned total = 0;
        for (int k = 0; k < nm; ++k) { total += s_end[k]; s_end[k] = total; }
    }
    __syncthreads();
    const unsigned total = s_end[nm - 1];
    for (unsigned p = threadIdx.x; p < total; p += blockDim.x) {
        int lo = 0, hi = nm - 1;   // the face whose slot range holds p: s_end[lo - 1] <= p < s_end[lo]
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_end[mid] > p) hi = mid;
            else lo = mid + 1;
        }
        seg_piece_body<FAM>(P, (long)s_face[lo], p - (lo > 0 ? s_end[lo - 1] : 0u));
    }
}

// ---------------------------------------------------------------------------------------------
// Zero-velocity pass (reference src/move_substances_3d.cpp:133-153: backtrace with zero velocity).
// Every face is its own undeformed lattice face and only y faces carry y area.  A y face spans one
// cell in x and z, so each colour range holds at most one integer strictly inside it: the cutter's
// pieces are up to 2 x 2 axis-aligned rectangles, whose xi-eta centroids map to their centres and
// whose y vector areas are their sizes.  One thread per y face sums them in the cutter's order
// (x strips outer, z bands inner).  Not used when minimal_area < 1 (pieces are subdivided then).
// ---------------------------------------------------------------------------------------------
// colour offset of the split along x (d = 0) or z (d = 1): c = position - offset
// (reference src/split_face_3D.cpp:26-43); false: no cut along that axis
FIA_HD bool split_offset(int split, int d, double &off) {
    switch (split) {
        case 1: off = 0.5; return true;
        case 2: off = 0.0; return true;
        case 3: off = d == 0 ? 0.5 : 0.0; return true;
        case 4: off = d == 0 ? 0.0 : 0.5; return true;
        default: off = 0.0; return false;
    }
}

// pieces of the unit interval [a, a+1] cut at the integer colour strictly inside (c = x - off)
FIA_HD int unit_interval_pieces(double a, bool cut, double off, double lo[2], double hi[2]) {
    if (cut) {
        const double c0 = a - off;
        const double k = floor(c0) + 1.0;   // smallest integer above c0; equals c0 + 1 when c0 is one
        if (k < c0 + 1.0) {
            const double x = k + off;
            lo[0] = a; hi[0] = x; lo[1] = x; hi[1] = a + 1.0;
            return 2;
        }
    }
    lo[0] = a; hi[0] = a + 1.0;
    return 1;
}

// WENO6 psi of the 2 x 2 pieces at once (non-integral rule): the four centres lie on psi's sample
// level j, so each value is a z pass then an x pass over the 6 x 6 samples of that level around it.
// The two centres along an axis are half a cell apart, so their sample bases differ by 0 or 1 and
// the four stencils fit in 7 x 7 samples.  Columns are visited in x order: each gets one z pass
// per z centre (14 evaluations in all) into a 6-column window, and a piece's x pass runs when the
// window holds its six columns (4 more), the same arithmetic as weno3d.
FIA_HD double zero_velocity_weno4(const FaceParams &P, int j, const double xl[2], const double xh[2],
                                  const double zl[2], const double zh[2]) {
    const double opx = 0.5 - 0.5 * P.g.e[0], opz = 0.5 - 0.5 * P.g.e[2];
    int bx[2], bz[2];
    float tx[2], tz[2];
    for (int a = 0; a < 2; ++a) {
        split_coord(0.5 * (xl[a] + xh[a]) - opx, bx[a], tx[a]);
        split_coord(0.5 * (zl[a] + zh[a]) - opz, bz[a], tz[a]);
    }
    const int dx = bx[1] - bx[0], dz = bz[1] - bz[0];   // 0 or 1
    const Loader<double> PS{P.psi, P.g.m[0], P.g.m[1] + 1, P.g.m[2]};
    const WenoCoef wz0(tz[0], P.kappa_psi), wz1(tz[1], P.kappa_psi);
    double w0[6] = {0, 0, 0, 0, 0, 0}, w1[6] = {0, 0, 0, 0, 0, 0};   // z passes of the last 6 columns
    double flux = 0.0;
#pragma unroll 1
    for (int c = 0; c < 7; ++c) {
        double f[7];
#pragma unroll
        for (int r = 0; r < 7; ++r) f[r] = PS(bx[0] - 2 + c, j, bz[0] - 2 + r);
        double g[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) g[r] = dz ? f[r + 1] : f[r];
        const double v0 = weno6<double>(f, wz0), v1 = weno6<double>(g, wz1);
#pragma unroll
        for (int r = 0; r < 5; ++r) { w0[r] = w0[r + 1]; w1[r] = w1[r + 1]; }
        w0[5] = v0; w1[5] = v1;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (c != 5 + (a == 1 ? dx : 0)) continue;
            const WenoCoef wx(tx[a], P.kappa_psi);
            flux += ((xh[a] - xl[a]) * (zh[0] - zl[0])) * weno6<double>(w0, wx);
            flux += ((xh[a] - xl[a]) * (zh[1] - zl[1])) * weno6<double>(w1, wx);
        }
    }
    return flux;
}

// FAST: WENO6 non-integral rule with both axes cut in two (the host decides from the split and the
// grid's staggering); otherwise the general quadrature per piece.
template <bool FAST> FIA_HD void zero_velocity_face_body(const FaceParams &P, long id) {
    int i, j, k;
    decode_face(P.g, 1, id, i, j, k);
    const double x0 = i - 0.5 * P.g.e[0], y = j - 0.5 * P.g.e[1], z0 = k - 0.5 * P.g.e[2];
    double ox, oz;
    const bool cx = split_offset(P.split, 0, ox), cz = split_offset(P.split, 1, oz);
    double xl[2], xh[2], zl[2], zh[2];
    const int nx = unit_interval_pieces(x0, cx, ox, xl, xh), nz = unit_interval_pieces(z0, cz, oz, zl, zh);
    if constexpr (FAST) {
        P.flux[1][id] = (nx == 2 && nz == 2) ? zero_velocity_weno4(P, j, xl, xh, zl, zh) : 0.0;
        return;
    } else {
        double f = 0.0;
        int npts = 0;
#pragma unroll 1
        for (int q = 0; q < nx * nz; ++q) {
            const int a = q / nz, b = q - a * nz;   // x strips outer, z bands inner
            const double c3[3] = {0.5 * (xl[a] + xh[a]), y, 0.5 * (zl[b] + zh[b])};
            f += piece_flux(P, c3, (xh[a] - xl[a]) * (zh[b] - zl[b]), npts);
        }
        P.flux[1][id] = f;
    }
}
template <bool FAST> __global__ void __launch_bounds__(kBlock) zero_velocity_face_kernel(FaceParams P) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)P.g.faces(1)) return;
    zero_velocity_face_body<FAST>(P, id);
}

// Work counters: block sum in LDS, then one atomic per counter and block on one of kStatSlots
// counter triples (spread so the atomics do not serialise on one address).
__device__ __forceinline__ void add_stats(unsigned long long *stats, int a, int b, int c) {
    __shared__ unsigned long long sh[kStatWords][kBlock / 64];
    unsigned long long x[kStatWords] = {(unsigned long long)a, (unsigned long long)b, (unsigned long long)c};
#pragma unroll
    for (int w = 0; w < kStatWords; ++w)
        for (int off = 32; off > 0; off >>= 1) x[w] += __shfl_down(x[w], off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int w = 0; w < kStatWords; ++w) sh[w][wave] = x[w];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int slot = (int)(blockIdx.x % kStatSlots);
        for (int w = 0; w < kStatWords; ++w) {
            unsigned long long t = 0;
            for (int v = 0; v < (int)(blockDim.x >> 6); ++v) t += sh[w][v];
            atomicAdd(stats + kStatWords * slot + w, t);
        }
    }
}

// Phase 3: quadrature of every piece slot, one thread per slot (the expensive WENO work runs
// converged regardless of how many pieces each face has).  QM selects the quadrature: the
// WENO6 integral rule with NG = QM Gauss points per half (1 or 2), or any mode (QM = 0).
template <int QM, bool OPT = false>
FIA_HD void piece_body(const FaceParams &P, unsigned long long q, int &npieces, int &npts, int &nevals) {
    const double ay = P.piece_ay[q];
    double val = 0.0;
    if (ay != 0.0) {
        const double c3[3] = {P.piece_c[0][q], P.piece_c[1][q], P.piece_c[2][q]};
        if constexpr (QM == 0) val = piece_flux(P, c3, ay, npts);
        else val = piece_flux_weno_integral<QM, OPT>(P, c3, ay, npts, nevals);
        npieces = 1;
    }
    P.piece_val[q] = val;
}
// The one-point WENO6 integral rule (QM = 1) is held to 128 VGPRs, four waves per SIMD, which its
// VALU-bound quadrature needs to hide the row loads.  OPT: the "WENO6-optimized" interpolation
// (its three reciprocals per evaluation do not fit 128 VGPRs; it is left to the compiler).
template <int QM, bool OPT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(QM == 1 && !OPT ? 4 : 1)))
piece_kernel(FaceParams P, unsigned long long nslots) {
    const unsigned long long q = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    int npieces = 0, npts = 0, nevals = 0;
    if (q < nslots) piece_body<QM, OPT>(P, q, npieces, npts, nevals);
    if (QM != 0 && P.stats) add_stats(P.stats, npieces, npts, nevals);   // counted for the WENO6 rule only
}

// Phase 4: face flux = ordered sum of its slots (deterministic).
template <int FAM> FIA_HD void face_sum_body(const FaceParams &P, long id) {
    const size_t gid = P.fam_base[FAM] + id;
    const unsigned n = P.face_n[gid];   // 0 for faces of the general cutter (it runs later)
    const unsigned long long s0 = P.face_start[gid];
    double f = 0.0;
    for (unsigned q = 0; q < n; ++q) f += P.piece_val[s0 + q];
    P.flux[FAM][id] = f;
}
template <int FAM> __global__ void __launch_bounds__(kBlock) face_sum_kernel(FaceParams P) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)P.g.faces(FAM)) return;
    face_sum_body<FAM>(P, id);
}

// Degenerate faces of the segment-cutter mode (a corner colour exactly on an interior integer):
// the clipping pipeline cuts every (strip, band) piece of the face into its slot, with the
// interpolated colours of the recursive cut, so the side an edge on a cut line goes to is decided
// the same way.  A grid-stride loop over the queued faces (few, near walls or pinned vertices).
FIA_HD void degenerate_body(const FaceParams &P, int t) {
    const unsigned code = P.degen[t];
    const int fam = (int)(code >> 30);
    const long id = (long)(code & ((1u << 30) - 1u));
    int i, j, k;
    decode_face(P.g, fam, id, i, j, k);
    double v[4][3], col[4][2];
    load_face(P, fam, i, j, k, v, col);
    FaceGeom geom;
    geom.init(v);
    int x0, nxc, z0, nzc;
    colour_ranges(col, x0, nxc, z0, nzc);
    unsigned long long slot = P.face_start[P.fam_base[fam] + id];
    for (int s = 0; s <= nxc; ++s) {
        for (int m = 0; m <= nzc; ++m, ++slot) {
            Pipeline pp;
            pp.st[0].reset(s > 0, 0, (double)(x0 + s - 1), true);
            pp.st[1].reset(s < nxc, 0, (double)(x0 + s), false);
            pp.st[2].reset(m > 0, 1, (double)(z0 + m - 1), true);
            pp.st[3].reset(m < nzc, 1, (double)(z0 + m), false);
            pp.acc.reset();
            const double sq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int c = 0; c < 4; ++c) {
                CutVert cv;
                cv.xi = sq[c][0]; cv.eta = sq[c][1]; cv.cx = col[c][0]; cv.cz = col[c][1];
                pp.push<0>(geom, cv);
            }
            pp.close<0>(geom);
            double ay = 0.0, c3[3] = {0.0, 0.0, 0.0};
            if (pp.acc.n >= 3 && pp.acc.area2 != 0.0) {
                const double inv3a = 1.0 / (3.0 * pp.acc.area2);
                double r[3];
                geom.rel(pp.acc.mxi * inv3a, pp.acc.meta * inv3a, r);
                c3[0] = geom.p0[0] + r[0]; c3[1] = geom.p0[1] + r[1]; c3[2] = geom.p0[2] + r[2];
                ay = 0.5 * pp.acc.ay2;
            }
            P.piece_c[0][slot] = c3[0];
            P.piece_c[1][slot] = c3[1];
            P.piece_c[2][slot] = c3[2];
            P.piece_ay[slot] = ay;
        }
    }
}
__global__ void __launch_bounds__(64) degenerate_kernel(FaceParams P) {
    const int count = *P.degen_count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) degenerate_body(P, t);
}
// With fuse_value (no quadrature launch): the values of the degenerate faces' pieces from the
// records degenerate_kernel stored.
FIA_HD void degenerate_value_body(const FaceParams &P, int t) {
    const unsigned code = P.degen[t];
    const size_t gid = P.fam_base[code >> 30] + (code & ((1u << 30) - 1u));
    const unsigned n = P.face_n[gid];
    const unsigned long long s0 = P.face_start[gid];
    for (unsigned k = 0; k < n; ++k) {
        const unsigned long long q = s0 + k;
        const double ay = P.piece_ay[q];
        const double c3[3] = {P.piece_c[0][q], P.piece_c[1][q], P.piece_c[2][q]};
        P.piece_val[q] = ay != 0.0 ? piece_value_linear(P, c3, ay) : 0.0;
    }
}
__global__ void __launch_bounds__(64) degenerate_value_kernel(FaceParams P) {
    const int count = *P.degen_count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x)
        degenerate_value_body(P, t);
}

// ---------------------------------------------------------------------------------------------
// Twisted faces: the exact recursive cutter (fia_cutter.hpp general_cut).  The reference evaluates
// every face twice, once per neighbouring cell with the corner order reversed
// (src/backtrace_and_calc_all_quadrature_point_list_for_integrate_density.cpp:77-112); for a
// twisted face its cutter's pieces depend on that order, so both evaluations are done: the lower
// cell's ('+' order v0 v1 v2 v3, orientation 0) goes to flux, the upper cell's (v0 v3 v2 v1,
// orientation 1) enters through corr = upper + flux.  In steps, so that no thread integrates a
// whole face alone:
//   general_count_kernel  pieces of each orientation of each listed face (t: 2t, 2t + 1)
//   general_scan_kernel   their exclusive scan (one block) and the total, read back to size
//                         the piece records
//   general_emit_kernel   the pieces again, each stored as a record (centre, y area) in its range
//   general_piece_kernel  the quadrature of every piece, one thread per piece
//   general_sum_kernel    each orientation's pieces summed in cut order into flux and corr
// Ranges and order depend only on the faces, so the sums are deterministic.
// ---------------------------------------------------------------------------------------------

// corners and colours of listed face t in orientation o; returns its family and index
FIA_HD void general_face(const FaceParams &P, int t, int o, int &fam, long &id, double v[4][3], double col[4][2]) {
    const unsigned code = P.twisted[t];
    fam = (int)(code >> 30);   // 0, 1 or 2
    id = (long)(code & ((1u << 30) - 1u));
    int i, j, k;
    decode_face(P.g, fam, id, i, j, k);
    double v0[4][3], c0[4][2];
    load_face(P, fam, i, j, k, v0, c0);
    const int order[2][4] = {{0, 1, 2, 3}, {0, 3, 2, 1}};
    for (int c = 0; c < 4; ++c) {
        for (int a = 0; a < 3; ++a) v[c][a] = v0[order[o][c]][a];
        col[c][0] = c0[order[o][c]][0]; col[c][1] = c0[order[o][c]][1];
    }
}

// centre (cells, physical frame) and y vector area of a piece the recursive cutter emits
FIA_HD void general_piece_record(const FaceGeom &geom, const double (*pts)[2], int num, double cxi, double ceta,
                                 double c3[3], double &ay) {
    double ay2 = 0.0;
    double r0[3], rp[3];
    geom.rel(pts[0][0], pts[0][1], r0);
    rp[0] = r0[0]; rp[1] = r0[1]; rp[2] = r0[2];
    for (int n = 1; n <= num; ++n) {
        double rn[3];
        if (n == num) { rn[0] = r0[0]; rn[1] = r0[1]; rn[2] = r0[2]; }
        else geom.rel(pts[n][0], pts[n][1], rn);
        ay2 += rp[2] * rn[0] - rp[0] * rn[2];
        rp[0] = rn[0]; rp[1] = rn[1]; rp[2] = rn[2];
    }
    double r[3];
    geom.rel(cxi, ceta, r);
    c3[0] = geom.p0[0] + r[0]; c3[1] = geom.p0[1] + r[1]; c3[2] = geom.p0[2] + r[2];
    ay = 0.5 * ay2;
}

FIA_HD void general_count_body(const FaceParams &P, int t) {
    for (int o = 0; o < 2; ++o) {
        int fam;
        long id;
        double v[4][3], col[4][2];
        general_face(P, t, o, fam, id, v, col);
        unsigned n = 0;
        auto count = [&](const double (*)[2], int, double, double) { ++n; };
        const int code = general_cut(col, P.min_area, count);
        if (code != KE_NONE) { report(P.error, code); n = 0; }
        P.gen_n[2 * t + o] = n;
    }
}
__global__ void __launch_bounds__(64) general_count_kernel(FaceParams P) {
    const int count = *P.twisted_count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) general_count_body(P, t);
}

FIA_HD void general_emit_body(const FaceParams &P, int t) {
    for (int o = 0; o < 2; ++o) {
        int fam;
        long id;
        double v[4][3], col[4][2];
        general_face(P, t, o, fam, id, v, col);
        FaceGeom geom;
        geom.init(v);
        const unsigned long long s0 = P.gen_start[2 * t + o];
        const unsigned cap = P.gen_n[2 * t + o];
        unsigned k = 0;
        auto emit = [&](const double (*pts)[2], int num, double cxi, double ceta) {
            if (k >= cap) return;   // the same cut as the count: never taken
            double c3[3], ay;
            general_piece_record(geom, pts, num, cxi, ceta, c3, ay);
            const unsigned long long q = s0 + k++;
            P.gen_c[0][q] = c3[0]; P.gen_c[1][q] = c3[1]; P.gen_c[2][q] = c3[2];
            P.gen_ay[q] = ay;
        };
        if (cap > 0) general_cut(col, P.min_area, emit);
    }
}
__global__ void __launch_bounds__(64) general_emit_kernel(FaceParams P) {
    const int count = *P.twisted_count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) general_emit_body(P, t);
}

FIA_HD void general_piece_body(const FaceParams &P, unsigned long long q) {
    const double ay = P.gen_ay[q];
    double val = 0.0;
    if (ay != 0.0) {
        const double c3[3] = {P.gen_c[0][q], P.gen_c[1][q], P.gen_c[2][q]};
        int npts = 0;
        val = piece_flux(P, c3, ay, npts);
    }
    P.gen_val[q] = val;
}
__global__ void __launch_bounds__(kBlock) general_piece_kernel(FaceParams P, unsigned long long npieces) {
    const unsigned long long q = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < npieces) general_piece_body(P, q);
}

FIA_HD void general_sum_body(const FaceParams &P, int t) {
    double f[2];
    for (int o = 0; o < 2; ++o) {
        const unsigned long long s0 = P.gen_start[2 * t + o];
        const unsigned n = P.gen_n[2 * t + o];
        double s = 0.0;
        for (unsigned k = 0; k < n; ++k) s += P.gen_val[s0 + k];
        f[o] = s;
    }
    const unsigned code = P.twisted[t];
    const int fam = (int)(code >> 30);
    const long id = (long)(code & ((1u << 30) - 1u));
    P.flux[fam][id] = f[0];
    P.corr[fam][id] = f[1] + f[0];
    atomic_or_u32(&P.corr_mask[fam][id >> 5], 1u << (id & 31));
}
__global__ void __launch_bounds__(kBlock) general_sum_kernel(FaceParams P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < *P.twisted_count) general_sum_body(P, t);
}

// ---------------------------------------------------------------------------------------------
// gather: q(i,j,k) = sum over the six faces of +-flux (the '+' face of a cell is its own
// orientation, the '-' face is the neighbour's '+' face reversed).
// ---------------------------------------------------------------------------------------------
// The correction of the lower face is read only where its mask bit is set (faces of the general
// cutter); elsewhere it is zero and the sum adds +0.0, as it would from a stored zero.
FIA_HD double corr_at(const FluxView &v, int fam, size_t lo) {
    return ((v.m[fam][lo >> 5] >> (lo & 31)) & 1u) ? v.c[fam][lo] : 0.0;
}

struct GatherArgs {
    CompGrid g;
    FluxView s1, s2;
    const float *base;
    float *out;
    int mode, fam_mask;
};

FIA_HD double cell_sum(const CompGrid &g, const FluxView &v, int i, int j, int k, int fm) {
    const int my = g.m[1], mz = g.m[2];
    double s = 0.0;
    if (fm & 1) {
        const size_t lo = idx3(i, j, k, my, mz);
        s += v.f[0][idx3(i + 1, j, k, my, mz)] - v.f[0][lo] + corr_at(v, 0, lo);
    }
    if (fm & 2) {
        const size_t lo = idx3(i, j, k, my + 1, mz);
        s += v.f[1][idx3(i, j + 1, k, my + 1, mz)] - v.f[1][lo] + corr_at(v, 1, lo);
    }
    if (fm & 4) {
        const size_t lo = idx3(i, j, k, my, mz + 1);
        s += v.f[2][idx3(i, j, k + 1, my, mz + 1)] - v.f[2][lo] + corr_at(v, 2, lo);
    }
    return s;
}

FIA_HD void gather_body(const GatherArgs &A, long id) {
    const int mz = A.g.m[2], my = A.g.m[1];
    const int k = (int)(id % mz), j = (int)((id / mz) % my), i = (int)(id / ((long)my * mz));
    double s = cell_sum(A.g, A.s1, i, j, k, A.fam_mask);
    if (A.mode == 1) s += cell_sum(A.g, A.s2, i, j, k, 7);
    if (A.mode == 2) s = (double)A.base[id] - s;
    A.out[id] = (float)s;
}
__global__ void __launch_bounds__(kBlock) gather_kernel(GatherArgs A) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)A.g.cells()) return;
    gather_body(A, id);
}

// Back to the invariant of a flux set between passes: every correction zero, every mask bit clear.
// Word w of a family's mask: each set bit's correction to zero, then the word.
FIA_HD void corr_reset_word(unsigned *mask, double *corr, size_t w) {
    unsigned b = mask[w];
    if (b == 0u) return;
    while (b) {
        const int t = __builtin_ctz(b);
        corr[w * 32 + (size_t)t] = 0.0;
        b &= b - 1u;
    }
    mask[w] = 0u;
}
struct CorrResetArgs {
    FluxView v;
    size_t words[3];
};
__global__ void __launch_bounds__(kBlock) corr_reset_kernel(CorrResetArgs A) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t w = id;
    for (int a = 0; a < 3; ++a) {
        if (w < A.words[a]) { corr_reset_word(A.v.m[a], A.v.c[a], w); return; }
        w -= A.words[a];
    }
}

// ---------------------------------------------------------------------------------------------
// utility kernels
// ---------------------------------------------------------------------------------------------
__global__ void interpolate_kernel(const float *field, int mx, int my, int mz, double ox, double oy, double oz,
                                   int method, float kappa, const float *pts, size_t n, float *out) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const Loader<float> L{field, mx, my, mz};
    out[id] = interp_point<float>(L, method, pts[3 * id], pts[3 * id + 1], pts[3 * id + 2], ox, oy, oz, kappa);
}

__global__ void __launch_bounds__(64) split_faces_kernel(const double *quads, size_t n, int split, double min_area,
                                                         int max_pieces, int *count, double *cen, double *area,
                                                         int *error) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    double v[4][3], col[4][2];
    for (int c = 0; c < 4; ++c) {
        for (int a = 0; a < 3; ++a) v[c][a] = quads[12 * id + 3 * c + a];
        const double x = v[c][0], z = v[c][2];
        double cx = 0.0, cz = 0.0;
        switch (split) {
            case 1: cx = x - 0.5; cz = z - 0.5; break;
            case 2: cx = x; cz = z; break;
            case 3: cx = x - 0.5; cz = z; break;
            case 4: cx = x; cz = z - 0.5; break;
            default: break;
        }
        col[c][0] = cx; col[c][1] = cz;
    }
    FaceGeom geom;
    geom.init(v);
    int cnt = 0;
    int over = KE_NONE;
    auto put = [&](const double r[3], double ax2, double ay2, double az2) {
        if (cnt >= max_pieces) { over = KE_CAPACITY; return; }
        const size_t o = (id * max_pieces + cnt) * 3;
        cen[o] = geom.p0[0] + r[0]; cen[o + 1] = geom.p0[1] + r[1]; cen[o + 2] = geom.p0[2] + r[2];
        area[o] = 0.5 * ax2; area[o + 1] = 0.5 * ay2; area[o + 2] = 0.5 * az2;
        ++cnt;
    };
    // same cutter choice as the flux passes: twisted faces (and every face when minimal_area < 1)
    // take the recursive cutter, degenerate ones the streaming pipeline, the rest boundary segments
    bool twisted = min_area < 1.0 || face_twisted(col);
    const bool seg = !face_degenerate(col);
    if (!twisted) {
        int x0, nxc, z0, nzc;
        colour_ranges(col, x0, nxc, z0, nzc);
        for (int s = 0; s <= nxc && !twisted; ++s) {
            for (int m = 0; m <= nzc; ++m) {
                const double xl = s > 0 ? (double)(x0 + s - 1) : -kNoBound, xh = s < nxc ? (double)(x0 + s) : kNoBound;
                const double zl = m > 0 ? (double)(z0 + m - 1) : -kNoBound, zh = m < nzc ? (double)(z0 + m) : kNoBound;
                double area2, mxi, meta, ax2, ay2, az2;
                if (seg) {
                    SegAcc<true> acc;
                    piece_moments(col, geom, xl, xh, zl, zh, acc);
                    area2 = acc.area2; mxi = acc.mxi; meta = acc.meta; ax2 = acc.ax2; ay2 = acc.ay2; az2 = acc.az2;
                } else {
                    Pipeline pp;
                    pp.st[0].reset(s > 0, 0, xl, true);
                    pp.st[1].reset(s < nxc, 0, xh, false);
                    pp.st[2].reset(m > 0, 1, zl, true);
                    pp.st[3].reset(m < nzc, 1, zh, false);
                    pp.acc.reset();
                    const double sq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
                    for (int c = 0; c < 4; ++c) {
                        CutVert cv;
                        cv.xi = sq[c][0]; cv.eta = sq[c][1]; cv.cx = col[c][0]; cv.cz = col[c][1];
                        pp.push<0>(geom, cv);
                    }
                    pp.close<0>(geom);
                    if (pp.twisted()) { twisted = true; break; }
                    area2 = pp.acc.n >= 3 ? pp.acc.area2 : 0.0;
                    mxi = pp.acc.mxi; meta = pp.acc.meta; ax2 = pp.acc.ax2; ay2 = pp.acc.ay2; az2 = pp.acc.az2;
                }
                if (area2 != 0.0) {
                    const double inv3a = 1.0 / (3.0 * area2);
                    double r[3];
                    geom.rel(mxi * inv3a, meta * inv3a, r);
                    put(r, ax2, ay2, az2);
                }
            }
        }
    }
    if (twisted) {
        cnt = 0;
        over = false;
        auto emit = [&](const double (*pts)[2], int num, double cxi, double ceta) {
            double a2[3] = {0, 0, 0};
            double r0[3], rp[3];
            geom.rel(pts[0][0], pts[0][1], r0);
            rp[0] = r0[0]; rp[1] = r0[1]; rp[2] = r0[2];
            for (int q = 1; q <= num; ++q) {
                double rn[3];
                if (q == num) { rn[0] = r0[0]; rn[1] = r0[1]; rn[2] = r0[2]; }
                else geom.rel(pts[q][0], pts[q][1], rn);
                a2[0] += rp[1] * rn[2] - rp[2] * rn[1];
                a2[1] += rp[2] * rn[0] - rp[0] * rn[2];
                a2[2] += rp[0] * rn[1] - rp[1] * rn[0];
                rp[0] = rn[0]; rp[1] = rn[1]; rp[2] = rn[2];
            }
            double r[3];
            geom.rel(cxi, ceta, r);
            put(r, a2[0], a2[1], a2[2]);
        };
        const int code = general_cut(col, min_area, emit);
        if (code != KE_NONE) over = code;
    }
    count[id] = cnt;
    if (over) atomicCAS(error, 0, over);
}

__global__ void absmax_kernel(const float *a, size_t n, unsigned int *out_bits) {
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = fabsf(a[i]);
        m = (v > m || v != v) ? v : m;
    }
    // wave reduction then one atomic per wave; |x| bit patterns order like the floats
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_down(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, __float_as_uint(m));
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
hipError_t launch_psi(const PsiParams &p, hipStream_t s) {
    if (p.g.m[0] > 65535) return hipErrorInvalidValue;   // one grid row of blocks per x plane
    const dim3 grid((unsigned)((p.g.m[2] + 63) / 64), (unsigned)p.g.m[0]);
    if (psi_held(p)) hipLaunchKernelGGL(psi_kernel<true>, grid, dim3(64 * kPsiSegs), 0, s, p);
    else if (psi_weno_held(p)) hipLaunchKernelGGL(psi_weno_held_kernel, grid, dim3(64 * kPsiSegs), 0, s, p);
    else hipLaunchKernelGGL(psi_kernel<false>, grid, dim3(64 * kPsiSegs), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_vertices(const VertexParams &p, hipStream_t s) {
    hipLaunchKernelGGL(vertex_kernel, dim3(grid_for(p.g.lattice(), kBlock)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_face_count(const FaceParams &p, hipStream_t s) {
    // every family writes its entries so the scan sees no stale counts
    const int b = block_or_default(p.face_block);
    hipLaunchKernelGGL(face_count_kernel<0>, dim3(grid_for(p.g.faces(0), b)), dim3(b), 0, s, p);
    hipLaunchKernelGGL(face_count_kernel<1>, dim3(grid_for(p.g.faces(1), b)), dim3(b), 0, s, p);
    hipLaunchKernelGGL(face_count_kernel<2>, dim3(grid_for(p.g.faces(2), b)), dim3(b), 0, s, p);
    return hipGetLastError();
}

int piece_quadrature_mode(const FaceParams &p) {
    if (p.interp == IM_WENO6 && p.integral && (p.gauss.n == 1 || p.gauss.n == 2)) return p.gauss.n;
    return 0;
}

hipError_t launch_pieces(const FaceParams &p, unsigned long long nslots, hipStream_t s) {
    if (nslots == 0) return hipSuccess;
    const int b = block_or_default(p.piece_block);
    const dim3 grid(grid_for(nslots, b));
    const int qm = piece_quadrature_mode(p);
    if (p.weno_opt) {
        if (qm == 1) hipLaunchKernelGGL((piece_kernel<1, true>), grid, dim3(b), 0, s, p, nslots);
        else if (qm == 2) hipLaunchKernelGGL((piece_kernel<2, true>), grid, dim3(b), 0, s, p, nslots);
        else hipLaunchKernelGGL((piece_kernel<0, true>), grid, dim3(b), 0, s, p, nslots);
    } else if (qm == 1) {
        hipLaunchKernelGGL((piece_kernel<1, false>), grid, dim3(b), 0, s, p, nslots);
    } else if (qm == 2) {
        hipLaunchKernelGGL((piece_kernel<2, false>), grid, dim3(b), 0, s, p, nslots);
    } else {
        hipLaunchKernelGGL((piece_kernel<0, false>), grid, dim3(b), 0, s, p, nslots);
    }
    return hipGetLastError();
}

hipError_t launch_face_sum(const FaceParams &p, hipStream_t s) {
    const int b = block_or_default(p.face_block);
    if (p.fam_mask & 1) hipLaunchKernelGGL(face_sum_kernel<0>, dim3(grid_for(p.g.faces(0), b)), dim3(b), 0, s, p);
    if (p.fam_mask & 2) hipLaunchKernelGGL(face_sum_kernel<1>, dim3(grid_for(p.g.faces(1), b)), dim3(b), 0, s, p);
    if (p.fam_mask & 4) hipLaunchKernelGGL(face_sum_kernel<2>, dim3(grid_for(p.g.faces(2), b)), dim3(b), 0, s, p);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Exclusive scan of the per-face slot counts into 64-bit starts.  Three plain kernels over tiles
// of kScanTile counts: tile sums, one block scanning the tile sums, then each tile scanned in
// LDS and offset.  No block waits on another block's progress.  The last kernel also writes the
// slot bounds of the three families (bounds[f] = start of family f, bounds[3] = total).
// ---------------------------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

// block-wide exclusive scan of one value per thread (kBlock threads, 64-wide waves); total out
__device__ __forceinline__ unsigned long long block_exclusive_scan(unsigned long long x, unsigned long long &total) {
    __shared__ unsigned long long wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    unsigned long long before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();   // wsum may be reused by the caller's next scan
    total = all;
    return before + incl - x;
}

__global__ void __launch_bounds__(kBlock) scan_tile_sums_kernel(const unsigned *in, size_t n,
                                                                unsigned long long *tile_sum) {
    // the tile read striped over the block (coalesced); the sum does not depend on the order
    const size_t tile0 = (size_t)blockIdx.x * kScanTile;
    unsigned long long s = 0;
    if (tile0 + kScanTile <= n) {
        // a full tile (all but the last): no bounds tests, so the loads are issued together (a
        // load under a test is compiled into a branch and waited for at once)
        unsigned v[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) v[i] = in[tile0 + (size_t)i * kBlock + threadIdx.x];
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) s += v[i];
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            const size_t k = tile0 + (size_t)i * kBlock + threadIdx.x;
            if (k < n) s += in[k];
        }
    }
    unsigned long long total;
    (void)block_exclusive_scan(s, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// one block: tile_sum[0..ntiles) -> exclusive prefix in place
__global__ void __launch_bounds__(kBlock) scan_tile_offsets_kernel(unsigned long long *tile_sum, int ntiles) {
    const int per = (ntiles + kBlock - 1) / kBlock;
    const int b0 = threadIdx.x * per;
    unsigned long long s = 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < ntiles) s += tile_sum[b0 + i];
    unsigned long long total;
    unsigned long long run = block_exclusive_scan(s, total);
    for (int i = 0; i < per; ++i) {
        if (b0 + i < ntiles) {
            const unsigned long long v = tile_sum[b0 + i];
            tile_sum[b0 + i] = run;
            run += v;
        }
    }
}

// one block: gen_start[0..m) = exclusive scan of gen_n, gen_start[m] = total
__global__ void __launch_bounds__(kBlock) general_scan_kernel(const unsigned *n, unsigned long long *start, int m) {
    const int per = (m + kBlock - 1) / kBlock;
    const int b0 = threadIdx.x * per;
    unsigned long long s = 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < m) s += n[b0 + i];
    unsigned long long total;
    unsigned long long run = block_exclusive_scan(s, total);
    for (int i = 0; i < per; ++i) {
        if (b0 + i < m) {
            start[b0 + i] = run;
            run += n[b0 + i];
        }
    }
    if (threadIdx.x == 0) start[m] = total;
}

// LDS index of tile element e, padded by one word per kScanItems elements so that both the striped
// (e = i kBlock + t) and the blocked (e = t kScanItems + i) accesses of a wave spread over the banks
__device__ __forceinline__ int scan_pad(int e) { return e + e / kScanItems; }

// Each tile is loaded striped over the block (coalesced), transposed in LDS so that every thread
// scans kScanItems consecutive counts, and its starts stored striped again.
__global__ void __launch_bounds__(kBlock) scan_tiles_kernel(const unsigned *in, size_t n,
                                                            const unsigned long long *tile_off,
                                                            unsigned long long *out, FaceParams P,
                                                            unsigned long long *bounds) {
    __shared__ unsigned sv[kScanTile + kScanTile / kScanItems];
    __shared__ unsigned long long so[kScanTile + kScanTile / kScanItems];
    const size_t tile0 = (size_t)blockIdx.x * kScanTile;
    const bool full = tile0 + kScanTile <= n;   // all tiles but the last: loads without tests
    if (full) {
        unsigned w[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) w[i] = in[tile0 + (size_t)i * kBlock + threadIdx.x];
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) sv[scan_pad(i * kBlock + threadIdx.x)] = w[i];
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            const int e = i * kBlock + threadIdx.x;
            const size_t k = tile0 + e;
            sv[scan_pad(e)] = k < n ? in[k] : 0u;
        }
    }
    __syncthreads();
    unsigned v[kScanItems];
    unsigned long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = sv[scan_pad(threadIdx.x * kScanItems + i)];
        s += v[i];
    }
    unsigned long long total;
    unsigned long long run = tile_off[blockIdx.x] + block_exclusive_scan(s, total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        so[scan_pad(threadIdx.x * kScanItems + i)] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int e = i * kBlock + threadIdx.x;
        const size_t k = tile0 + e;
        if (k < n) {
            const unsigned long long r = so[scan_pad(e)];
            out[k] = r;
            if (k == P.fam_base[1]) bounds[1] = r;
            if (k == P.fam_base[2]) bounds[2] = r;
            if (k == n - 1) bounds[3] = r + sv[scan_pad(e)];
        }
    }
}

size_t scan_temp_words(size_t n) { return (n + kScanTile - 1) / kScanTile; }

hipError_t launch_scan(unsigned long long *tile_tmp, const unsigned *in, unsigned long long *out, size_t n,
                       const FaceParams &p, unsigned long long *bounds, hipStream_t s) {
    const size_t ntiles = scan_temp_words(n);
    if (ntiles == 0 || ntiles > (size_t)INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3((unsigned)ntiles), dim3(kBlock), 0, s, in, n, tile_tmp);
    hipLaunchKernelGGL(scan_tile_offsets_kernel, dim3(1), dim3(kBlock), 0, s, tile_tmp, (int)ntiles);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)ntiles), dim3(kBlock), 0, s, in, n,
                       (const unsigned long long *)tile_tmp, out, p, bounds);
    return hipGetLastError();
}

// Both axes of every undeformed y face are cut in two when the colour at the face's low corner is
// not an integer: position i - e/2 minus the split offset.
static bool zero_velocity_cuts_both(const FaceParams &p) {
    double ox, oz;
    const bool cx = split_offset(p.split, 0, ox), cz = split_offset(p.split, 1, oz);
    const double fx = 0.5 * p.g.e[0] + ox, fz = 0.5 * p.g.e[2] + oz;   // fractional part of -colour
    return cx && cz && fx != floor(fx) && fz != floor(fz);
}

hipError_t launch_zero_velocity_faces(const FaceParams &p, hipStream_t s) {
    const int b = block_or_default(p.face_block);
    const dim3 grid(grid_for(p.g.faces(1), b));
    // (the closed form's shared stencils are the WENO6 rule's; "WENO6-optimized" takes the
    // general quadrature per piece)
    if (p.interp == IM_WENO6 && !p.weno_opt && !p.integral && zero_velocity_cuts_both(p))
        hipLaunchKernelGGL(zero_velocity_face_kernel<true>, grid, dim3(b), 0, s, p);
    else
        hipLaunchKernelGGL(zero_velocity_face_kernel<false>, grid, dim3(b), 0, s, p);
    return hipGetLastError();
}

// grid-stride kernels over a face list of `count` faces: one 64-lane block per 64 faces, at most
// 1024 blocks (one wave per SIMD at these kernels' register counts)
static unsigned list_blocks(int count) { return (unsigned)std::min(1024, (count + 63) / 64); }

hipError_t launch_seg_faces(const FaceParams &p, hipStream_t s) {
    const int b = block_or_default(p.face_block);
    if (p.fam_mask & 1) hipLaunchKernelGGL(seg_face_kernel<0>, dim3(grid_for(p.g.faces(0), b)), dim3(b), 0, s, p);
    if (p.fam_mask & 2) hipLaunchKernelGGL(seg_face_kernel<1>, dim3(grid_for(p.g.faces(1), b)), dim3(b), 0, s, p);
    if (p.fam_mask & 4) hipLaunchKernelGGL(seg_face_kernel<2>, dim3(grid_for(p.g.faces(2), b)), dim3(b), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_faces_degenerate(const FaceParams &p, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(degenerate_kernel, dim3(list_blocks(count)), dim3(64), 0, s, p);
    if (p.fuse_value) hipLaunchKernelGGL(degenerate_value_kernel, dim3(list_blocks(count)), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_general_count(const FaceParams &p, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(general_count_kernel, dim3(list_blocks(count)), dim3(64), 0, s, p);
    hipLaunchKernelGGL(general_scan_kernel, dim3(1), dim3(kBlock), 0, s, (const unsigned *)p.gen_n, p.gen_start,
                       2 * count);
    return hipGetLastError();
}

hipError_t launch_general_pieces(const FaceParams &p, int count, unsigned long long npieces, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(general_emit_kernel, dim3(list_blocks(count)), dim3(64), 0, s, p);
    if (npieces > 0)
        hipLaunchKernelGGL(general_piece_kernel, dim3(grid_for(npieces, kBlock)), dim3(kBlock), 0, s, p, npieces);
    hipLaunchKernelGGL(general_sum_kernel, dim3(grid_for((size_t)count, kBlock)), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_gather(const CompGrid &g, const FluxView &s1, const FluxView *s2, const float *base, float *out,
                         int mode, int fam_mask, hipStream_t s, int block) {
    GatherArgs A;
    A.g = g;
    A.s1 = s1;
    A.s2 = s2 ? *s2 : FluxView{};
    A.base = base; A.out = out; A.mode = mode; A.fam_mask = fam_mask;
    const int b = block_or_default(block);
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(g.cells(), b)), dim3(b), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_corr_reset(const FluxView &v, const size_t words[3], hipStream_t s) {
    CorrResetArgs A;
    A.v = v;
    for (int a = 0; a < 3; ++a) A.words[a] = words[a];
    const size_t n = words[0] + words[1] + words[2];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(corr_reset_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_interpolate(const float *field, int mx, int my, int mz, double ox, double oy, double oz, int method,
                              float kappa, const float *pts, size_t n, float *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(interpolate_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, field, mx, my, mz, ox, oy,
                       oz, method, kappa, pts, n, out);
    return hipGetLastError();
}

hipError_t launch_split_faces(const double *quads, size_t n, int split, double min_area, int max_pieces, int *count,
                              double *cen, double *area, int *error, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(split_faces_kernel, dim3(grid_for(n, 64)), dim3(64), 0, s, quads, n, split, min_area,
                       max_pieces, count, cen, area, error);
    return hipGetLastError();
}

hipError_t launch_absmax(const float *a, size_t n, unsigned int *out_bits, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>(grid_for(n, kBlock), 2048);
    hipLaunchKernelGGL(absmax_kernel, dim3(blocks), dim3(kBlock), 0, s, a, n, out_bits);
    return hipGetLastError();
}

}  // namespace fia
